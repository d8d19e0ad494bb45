"""Measure the tile edge of the host-operand runtime on this GPU and write
the built-in table (csrc/tune_table.inc) that libblasx consults by default;
with --grouped, measure the block tile shapes of the grouped launches and
write csrc/tune_grouped.inc.

    python -m blasx_amd.tune [--out PATH] [--sizes 4096,8192,16384,32768]
    python -m blasx_amd.tune --grouped [--out PATH]

Tile edge: for each element type (S, D, Z) and cube size n,
blasx_autotune_tile_size times one C = A B + C call on pinned n x n operands
per candidate edge (256, 512, 1024, 2048 within a 32 MiB cache slot) after a
warm-up call and keeps the fastest.

Grouped shapes: for S and D and each workload (count problems of one M x N x
K, distinct operands in HBM), one blasx_gemm_grouped_device call per block
shape forced with blasx_set_grouped_shape (128 x 128, 128 x 64, 64 x 64),
timed with HIP events over repeated calls after a warm-up; the three
median times are kept with the workload's load (its tile count at 128 x 128).

Rebuild the library afterwards (make -C blasx_amd) to compile a table in;
$BLASX_TUNE_FILE entries and run-time measurements still take precedence.
"""
import argparse
import ctypes as C
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))


# grouped workloads: (M, N, K) and the problem counts measured for it
GROUPED_WORKLOADS = [((s, s, s), (1, 4, 16, 64, 256)) for s in (64, 128, 256, 512, 1024)]
GROUPED_WORKLOADS += [((2048, 2048, 2048), (1, 4, 16, 64)),
                      ((4096, 64, 4096), (1, 8, 64)), ((64, 4096, 4096), (1, 8, 64)),
                      ((2048, 128, 512), (1, 8, 64)), ((512, 512, 8192), (1, 8, 64)),
                      ((256, 256, 4096), (1, 16, 128))]


def _time_grouped(bx, descs, min_ms=30.0, max_reps=50):
    bx.gemm_grouped_device(descs)  # warm-up
    bx.synchronize()
    e0, e1 = bx.Event(), bx.Event()
    ms, reps = 0.0, 0
    times = []
    while reps < max_reps and (ms < min_ms or reps < 3):
        e0.record()
        bx.gemm_grouped_device(descs)
        e1.record()
        e1.synchronize()
        t = e0.elapsed_ms(e1)
        times.append(t)
        ms += t
        reps += 1
    times.sort()
    return times[len(times) // 2]


def tune_grouped(bx, out):
    import numpy as np

    L = bx.lib()
    rows = []
    for dt, name, np_dt in ((bx.S, "S", np.float32), (bx.D, "D", np.float64)):
        for (M, N, K), counts in GROUPED_WORKLOADS:
            for cnt in counts:
                A = bx.DeviceArray(M, K * cnt, np_dt)
                B = bx.DeviceArray(K, N * cnt, np_dt)
                Cm = bx.DeviceArray(M, N * cnt, np_dt)
                for i, x in enumerate((A, B, Cm)):
                    x.fill_uniform(5, i)
                es = np.dtype(np_dt).itemsize
                descs = [bx.make_desc(np_dt, "N", "N", M, N, K, 1.0, A.ptr + i * M * K * es, M,
                                      B.ptr + i * K * N * es, K, 1.0, Cm.ptr + i * M * N * es, M)
                         for i in range(cnt)]
                load = cnt * ((M + 127) // 128) * ((N + 127) // 128)
                ts = []
                try:
                    for shape in range(3):
                        assert L.blasx_set_grouped_shape(dt, shape) == 0
                        ts.append(_time_grouped(bx, descs))
                finally:
                    L.blasx_set_grouped_shape(dt, -1)
                for x in (A, B, Cm):
                    x.free()
                best = min(range(3), key=lambda i: ts[i])
                tf = [2.0 * M * N * K * cnt / (t * 1e-3) / 1e12 for t in ts]
                note = ", ".join(f"{lbl}: {t:.3f} ms {f:.1f} TF" for lbl, t, f in
                                 zip(("128x128", "128x64", "64x64"), ts, tf))
                print(f"tune: grouped {name} {cnt} x {M}x{N}x{K} (load {load}) -> shape {best} "
                      f"({note})", file=sys.stderr, flush=True)
                rows.append((dt, M, N, K, load, [max(1, round(t * 1000)) for t in ts]))
    with open(out, "w") as f:
        f.write("/* tune_grouped.inc -- measured block tile shapes of the grouped launches, one\n"
                " * {dtype, M, N, K, load, t0, t1, t2} entry per line: a launch of `count` equal\n"
                " * NN problems (dtype: 0 S, 1 D; load: the launch's tile count at 128 x 128)\n"
                " * took t0, t1, t2 microseconds (median) on 128 x 128, 128 x 64 and 64 x 64 block\n"
                " * tiles.  Written by blasx_amd/tune.py --grouped on an MI355X. */\n")
        for dt, M, N, K, load, us in rows:
            f.write(f"{{{dt}, {M}, {N}, {K}, {load}, {us[0]}, {us[1]}, {us[2]}}},\n")
    print(f"tune: wrote {out}", file=sys.stderr)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", default=None)
    ap.add_argument("--sizes", default="4096,8192,16384,32768")
    ap.add_argument("--reps", type=int, default=1)
    ap.add_argument("--grouped", action="store_true",
                    help="measure the grouped launches' block tile shapes instead")
    args = ap.parse_args(argv)
    sys.path.insert(0, os.path.dirname(HERE))
    import blasx_amd as bx

    L = bx.lib()
    if not bx.gpu_available():
        print("tune: libblasx reports no usable GPU", file=sys.stderr)
        return 1
    if args.grouped:
        return tune_grouped(bx, args.out or os.path.join(HERE, "csrc", "tune_grouped.inc"))
    args.out = args.out or os.path.join(HERE, "csrc", "tune_table.inc")
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
            ts = ", ".join(f"{256 << i}: {times[i] * 1e3:.3f} ms" for i in range(4) if times[i] > 0)
            print(f"tune: {name} {n}^3 -> T = {best.value} ({ts}; "
                  f"{time.perf_counter() - t0:.0f}s)", file=sys.stderr, flush=True)
            rows.append((dt, n, best.value, ts))
    with open(args.out, "w") as f:
        f.write("/* tune_table.inc -- measured tile edges of the host-operand runtime, one\n"
                " * {dtype, M, N, K, T} entry per line (dtype: 0 S, 1 D, 2 Z), written by\n"
                " * blasx_amd/tune.py from blasx_autotune_tile_size runs on an MI355X\n"
                " * (time per call of each candidate edge in the comments). */\n")
        for dt, n, T, ts in rows:
            f.write(f"{{{dt}, {n}, {n}, {n}, {T}}}, /* {ts} */\n")
    print(f"tune: wrote {args.out}", file=sys.stderr)
    return 0


if __name__ == "__main__":
    sys.exit(main())
