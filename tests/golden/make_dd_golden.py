"""Writes dd_golden.npz: exact GEMM results (mpmath, 120 digits) for small
inputs with a wide dynamic range and heavy cancellation.  They pin the
oracle's double-double GEMM (oracle/oracle.c orc_gemm_dd_d / orc_gemm_dd_z),
which in turn is the accuracy reference for the split-precision GEMM.

Inputs come from numpy's PCG64 with a fixed seed, so this script does not
depend on anything under oracle/.  Run: python tests/golden/make_dd_golden.py
"""
import os

import mpmath
import numpy as np

mpmath.mp.dps = 120
HERE = os.path.dirname(os.path.abspath(__file__))


def wide(rng, shape, phi):
    return (rng.random(shape) - 0.5) * np.exp(phi * rng.standard_normal(shape))


def exact_real(ta, tb, alpha, A, B, beta, C0):
    opA = A if ta == "N" else A.T
    opB = B if tb == "N" else B.T
    M, K = opA.shape
    N = opB.shape[1]
    hi = np.zeros((M, N))
    lo = np.zeros((M, N))
    for i in range(M):
        for j in range(N):
            s = mpmath.mpf(0)
            for p in range(K):
                s += mpmath.mpf(opA[i, p]) * mpmath.mpf(opB[p, j])
            v = mpmath.mpf(alpha) * s + mpmath.mpf(beta) * mpmath.mpf(C0[i, j])
            h = float(v)
            hi[i, j] = h
            lo[i, j] = float(v - mpmath.mpf(h))
    return hi, lo


def exact_cplx(ta, tb, alpha, A, B, beta, C0):
    def op(X, t):
        return X if t == "N" else (X.T if t == "T" else X.conj().T)
    opA, opB = op(A, ta), op(B, tb)
    M, K = opA.shape
    N = opB.shape[1]
    hi = np.zeros((M, N), dtype=np.complex128)
    lo = np.zeros((M, N), dtype=np.complex128)
    al, be = mpmath.mpc(alpha), mpmath.mpc(beta)
    for i in range(M):
        for j in range(N):
            s = mpmath.mpc(0)
            for p in range(K):
                s += mpmath.mpc(opA[i, p]) * mpmath.mpc(opB[p, j])
            v = al * s + be * mpmath.mpc(C0[i, j])
            hr, hi_ = float(v.real), float(v.imag)
            hi[i, j] = complex(hr, hi_)
            lo[i, j] = complex(float(v.real - mpmath.mpf(hr)), float(v.imag - mpmath.mpf(hi_)))
    return hi, lo


def main():
    rng = np.random.Generator(np.random.PCG64(20260101))
    out = {}
    # real: M=5, N=4, K=9 with phi=4 dynamic range, plus a cancellation-heavy case
    cases = [("N", "N", 5, 4, 9, 1.5, 0.5, 4.0), ("T", "N", 6, 3, 11, -0.75, 2.0, 2.0),
             ("N", "T", 4, 6, 13, 1.0, 0.0, 6.0)]
    for ci, (ta, tb, M, N, K, al, be, phi) in enumerate(cases):
        A = wide(rng, (M, K) if ta == "N" else (K, M), phi)
        B = wide(rng, (K, N) if tb == "N" else (N, K), phi)
        if ci == 0:
            # exact cancellation: column 0 of B is orthogonal-ish to row 0 of A
            B[:, 0] = 0.0
            B[0, 0], B[1, 0] = A[0, 1], -A[0, 0]
        C0 = wide(rng, (M, N), phi)
        hi, lo = exact_real(ta, tb, al, A, B, be, C0)
        out.update({f"r{ci}_A": A, f"r{ci}_B": B, f"r{ci}_C": C0, f"r{ci}_hi": hi,
                    f"r{ci}_lo": lo,
                    f"r{ci}_meta": np.array([M, N, K, al, be, ord(ta), ord(tb)])})
    ccases = [("N", "C", 4, 3, 7, 1.0 - 0.5j, 0.25 + 1j, 3.0), ("C", "T", 3, 5, 8, 2j, 0.0, 2.0)]
    for ci, (ta, tb, M, N, K, al, be, phi) in enumerate(ccases):
        A = wide(rng, (M, K) if ta == "N" else (K, M), phi) + 1j * wide(
            rng, (M, K) if ta == "N" else (K, M), phi)
        B = wide(rng, (K, N) if tb == "N" else (N, K), phi) + 1j * wide(
            rng, (K, N) if tb == "N" else (N, K), phi)
        C0 = wide(rng, (M, N), phi) + 1j * wide(rng, (M, N), phi)
        hi, lo = exact_cplx(ta, tb, al, A, B, be, C0)
        out.update({f"z{ci}_A": A, f"z{ci}_B": B, f"z{ci}_C": C0, f"z{ci}_hi": hi,
                    f"z{ci}_lo": lo,
                    f"z{ci}_meta": np.array([M, N, K, al, be, ord(ta), ord(tb)],
                                            dtype=np.complex128)})
    np.savez(os.path.join(HERE, "dd_golden.npz"), **out)
    print("wrote", os.path.join(HERE, "dd_golden.npz"))


if __name__ == "__main__":
    main()
