"""The oracle's third-party arithmetic against golden vectors.

LAPACK ?gttrf/?gttrs (reference algorithm, partial pivoting) from
scipy.linalg.lapack and the unnormalised DFT from numpy.fft, both written by
tests/golden/make_golden.py.  The factorisations must agree to the last few
ulps (same elimination order), the solutions and transforms to roundoff.
"""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
LAPACK = np.load(os.path.join(HERE, "golden", "lapack.npz"))
FFT = np.load(os.path.join(HERE, "golden", "fft.npz"))


def _cases(prefix_set):
    names = sorted({k.split("__")[0] for k in LAPACK.files if "__" in k})
    return [n for n in names if n.split("_")[0] in prefix_set]


@pytest.mark.parametrize("case", _cases({"real"}))
def test_dgttrf_dgttrs(oracle, case):
    g = {k.split("__")[1]: LAPACK[k] for k in LAPACK.files if k.startswith(case + "__")}
    dl, d, du, du2, ipiv, info = oracle.dgttrf(g["dl"], g["d"], g["du"])
    assert info == 0
    np.testing.assert_array_equal(ipiv, g["ipiv"])
    for got, ref in ((dl, g["dlf"]), (d, g["df"]), (du, g["duf"]), (du2, g["du2"])):
        np.testing.assert_allclose(got, ref, rtol=4e-16, atol=1e-300)
    x = oracle.dgttrs(dl, d, du, du2, ipiv, g["b"])
    np.testing.assert_allclose(x, g["x"], rtol=1e-14, atol=1e-15 * np.abs(g["x"]).max())


@pytest.mark.parametrize("case", _cases({"complex", "poisson"}))
def test_zgttrf_zgttrs(oracle, case):
    g = {k.split("__")[1]: LAPACK[k] for k in LAPACK.files if k.startswith(case + "__")}
    dl, d, du, du2, ipiv, info = oracle.zgttrf(g["dl"], g["d"], g["du"])
    assert info == 0
    np.testing.assert_array_equal(ipiv, g["ipiv"])
    scale = max(np.abs(g["df"]).max(), 1.0)
    for got, ref in ((dl, g["dlf"]), (d, g["df"]), (du, g["duf"]), (du2, g["du2"])):
        np.testing.assert_allclose(got, ref, rtol=1e-14, atol=1e-15 * scale)
    x = oracle.zgttrs(dl, d, du, du2, ipiv, g["b"])
    np.testing.assert_allclose(x, g["x"], rtol=1e-12, atol=1e-13 * np.abs(g["x"]).max())


def test_pivoting_is_exercised():
    # the golden set holds systems with and without row interchanges
    piv = [LAPACK[k] for k in LAPACK.files if k.endswith("__ipiv")]
    assert any((p != np.arange(len(p))).any() for p in piv)
    assert any((p == np.arange(len(p))).all() for p in piv)


def _fft_sizes(kind):
    return sorted({int(k.split("__")[0].split("_")[1]) for k in FFT.files if k.startswith(kind + "_")})


@pytest.mark.parametrize("n", _fft_sizes("r2c"))
def test_r2c(oracle, n):
    x, X = FFT[f"r2c_{n}__x"], FFT[f"r2c_{n}__X"]
    got = oracle.fft_r2c(x)
    np.testing.assert_allclose(got, X, rtol=0, atol=1e-13 * np.abs(X).max())


@pytest.mark.parametrize("n", _fft_sizes("c2r"))
def test_c2r(oracle, n):
    X, x = FFT[f"c2r_{n}__X"], FFT[f"c2r_{n}__x"]
    got = oracle.fft_c2r(X, n)
    np.testing.assert_allclose(got, x, rtol=0, atol=1e-13 * np.abs(x).max())


@pytest.mark.parametrize("n", _fft_sizes("c2c"))
def test_c2c(oracle, n):
    z = FFT[f"c2c_{n}__z"]
    for sign, key in ((-1, "fwd"), (1, "bwd")):
        ref = FFT[f"c2c_{n}__{key}"]
        got = oracle.fft_c2c(z, sign)
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-13 * np.abs(ref).max())


@pytest.mark.parametrize("n", [8, 12, 30])
def test_single_mode(oracle, n):
    # a pure exponential transforms to a single bin of height n
    for m in range(n):
        z = np.exp(2j * np.pi * m * np.arange(n) / n)
        X = oracle.fft_c2c(z, -1)
        ref = np.zeros(n, complex)
        ref[m] = n
        np.testing.assert_allclose(X, ref, atol=1e-12 * n)
