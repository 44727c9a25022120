"""Writes the golden vectors that pin the oracle's third-party arithmetic.

The reference calls LAPACK ?gttrf/?gttrs (SolveImpEqnUpdate_QT.F90:63,74;
SolveImpEqnUpdate_QR.F90:44,61; SolveImpEqnUpdate_QZ.F90:58,64;
SolvePressureCorrection.F90:139-141) and FFTW 3.3.4 unnormalised r2c / c2c /
c2r transforms (SolvePressureCorrection.F90:106-113,156-160).  Neither ships
with the reference, so the vectors come from scipy.linalg.lapack (the
reference LAPACK algorithm) and numpy.fft (the same unnormalised DFT) in this
container.

Run: python tests/golden/make_golden.py   (writes lapack.npz and fft.npz here)
"""
import os

import numpy as np
import scipy
from scipy.linalg import lapack

HERE = os.path.dirname(os.path.abspath(__file__))


def tridiag_cases(rng):
    cases = []
    # 1. general real system where partial pivoting happens
    n = 9
    cases.append(("real_pivot", rng.normal(size=n - 1), rng.normal(size=n) * 0.3,
                  rng.normal(size=n - 1), rng.normal(size=n)))
    # 2. row-normalised Helmholtz-like system (unit diagonal, small off-diagonals)
    n = 33
    cases.append(("real_helmholtz", -0.007 * rng.random(n - 1), np.ones(n),
                  -0.007 * rng.random(n - 1), rng.normal(size=n)))
    # 3. identity wall rows (the q_r system)
    n = 17
    dl, du = -0.01 * rng.random(n - 1), -0.01 * rng.random(n - 1)
    dl[-1] = 0.0
    du[0] = 0.0
    b = rng.normal(size=n)
    b[0] = b[-1] = 0.0
    cases.append(("real_walls", dl, np.ones(n), du, b))
    return cases


def complex_cases(rng):
    cases = []
    n = 12
    cases.append(("complex_pivot", rng.normal(size=n - 1) + 1j * rng.normal(size=n - 1),
                  0.3 * (rng.normal(size=n) + 1j * rng.normal(size=n)),
                  rng.normal(size=n - 1) + 1j * rng.normal(size=n - 1),
                  rng.normal(size=n) + 1j * rng.normal(size=n)))
    # Poisson-like: real coefficients (diag = -(am+ap) - sigma), complex rhs
    n = 20
    am = np.concatenate([[0.0], 50 + 50 * rng.random(n - 1)])
    ap = np.concatenate([50 + 50 * rng.random(n - 1), [0.0]])
    for name, sigma in (("poisson_low", 1e-3), ("poisson_high", 400.0)):
        d = -(am + ap) - sigma
        cases.append((name, am[1:].astype(complex), d.astype(complex), ap[:-1].astype(complex),
                      rng.normal(size=n) + 1j * rng.normal(size=n)))
    return cases


def main():
    rng = np.random.default_rng(20240601)
    out = {}
    for name, dl, d, du, b in tridiag_cases(rng):
        dlf, df, duf, du2, ipiv, info = lapack.dgttrf(dl, d, du)
        x, info2 = lapack.dgttrs(dlf, df, duf, du2, ipiv, b)[:2]
        for key, val in (("dl", dl), ("d", d), ("du", du), ("b", b), ("dlf", dlf), ("df", df),
                         ("duf", duf), ("du2", du2), ("ipiv", ipiv - 1), ("x", x)):
            out[f"{name}__{key}"] = np.asarray(val)
        out[f"{name}__info"] = np.array([info, info2])
    for name, dl, d, du, b in complex_cases(rng):
        dlf, df, duf, du2, ipiv, info = lapack.zgttrf(dl, d, du)
        x, info2 = lapack.zgttrs(dlf, df, duf, du2, ipiv, b)[:2]
        for key, val in (("dl", dl), ("d", d), ("du", du), ("b", b), ("dlf", dlf), ("df", df),
                         ("duf", duf), ("du2", du2), ("ipiv", ipiv - 1), ("x", x)):
            out[f"{name}__{key}"] = np.asarray(val)
        out[f"{name}__info"] = np.array([info, info2])
    out["versions"] = np.array([scipy.__version__, np.__version__])
    np.savez(os.path.join(HERE, "lapack.npz"), **out)

    fft = {}
    for n in (2, 6, 15, 30, 64, 96, 100, 384, 512):
        x = rng.normal(size=n)
        fft[f"r2c_{n}__x"] = x
        fft[f"r2c_{n}__X"] = np.fft.rfft(x)
        X = np.fft.rfft(rng.normal(size=n))
        # c2r ignores the imaginary parts of the zero and (even n) Nyquist bins
        X[0] += 0.37j
        if n % 2 == 0:
            X[-1] -= 0.21j
        fft[f"c2r_{n}__X"] = X
        fft[f"c2r_{n}__x"] = np.fft.irfft(X, n) * n
    for n in (3, 8, 12, 45, 64, 128, 250):
        z = rng.normal(size=n) + 1j * rng.normal(size=n)
        fft[f"c2c_{n}__z"] = z
        fft[f"c2c_{n}__fwd"] = np.fft.fft(z)
        fft[f"c2c_{n}__bwd"] = np.fft.ifft(z) * n
    np.savez(os.path.join(HERE, "fft.npz"), **fft)
    print("wrote", os.path.join(HERE, "lapack.npz"), os.path.join(HERE, "fft.npz"))


if __name__ == "__main__":
    main()
