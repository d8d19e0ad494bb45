"""Measure the tile edge of the host-operand runtime on this GPU and write
the built-in table (csrc/tune_table.inc) that libblasx consults by default.

    python -m blasx_amd.tune [--out PATH] [--sizes 4096,8192,16384,32768]

For each element type (S, D, Z) and cube size n, blasx_autotune_tile_size
times one C = A B + C call on pinned n x n operands per candidate edge
(256, 512, 1024, 2048 within a 32 MiB cache slot) after a warm-up call and
keeps the fastest.  Rebuild the library afterwards (make -C blasx_amd) to
compile the table in; $BLASX_TUNE_FILE entries and run-time measurements
still take precedence.
"""
import argparse
import ctypes as C
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", default=os.path.join(HERE, "csrc", "tune_table.inc"))
    ap.add_argument("--sizes", default="4096,8192,16384,32768")
    ap.add_argument("--reps", type=int, default=1)
    args = ap.parse_args(argv)
    sys.path.insert(0, os.path.dirname(HERE))
    import blasx_amd as bx

    L = bx.lib()
    if not bx.gpu_available():
        print("tune: libblasx reports no usable GPU", file=sys.stderr)
        return 1
    L.blasx_autotune_tile_size.argtypes = [C.c_int] * 5 + [C.POINTER(C.c_int),
                                                           C.POINTER(C.c_double)]
    sizes = [int(x) for x in args.sizes.split(",") if x]
    rows = []
    for dt, name in ((bx.S, "S"), (bx.D, "D"), (bx.Z, "Z")):
        for n in sizes:
            best = C.c_int()
            times = (C.c_double * 4)()
            t0 = time.perf_counter()
            rc = L.blasx_autotune_tile_size(dt, n, n, n, args.reps, C.byref(best), times)
            if rc:
                print(f"tune: {name} {n}^3 failed ({rc})", file=sys.stderr, flush=True)
                return 1
            ts = ", ".join(f"{256 << i}: {times[i]:.3f}s" for i in range(4) if times[i] > 0)
            print(f"tune: {name} {n}^3 -> T = {best.value} ({ts}; "
                  f"{time.perf_counter() - t0:.0f}s)", file=sys.stderr, flush=True)
            rows.append((dt, n, best.value, ts))
    with open(args.out, "w") as f:
        f.write("/* tune_table.inc -- measured tile edges of the host-operand runtime, one\n"
                " * {dtype, M, N, K, T} entry per line (dtype: 0 S, 1 D, 2 Z), written by\n"
                " * blasx_amd/tune.py from blasx_autotune_tile_size runs on an MI355X\n"
                " * (seconds per call of each candidate edge in the comments). */\n")
        for dt, n, T, ts in rows:
            f.write(f"{{{dt}, {n}, {n}, {n}, {T}}}, /* {ts} */\n")
    print(f"tune: wrote {args.out}", file=sys.stderr)
    return 0


if __name__ == "__main__":
    sys.exit(main())
