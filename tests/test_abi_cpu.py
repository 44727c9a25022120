"""C-ABI checks that need no GPU: the library loads and exports every entry
point include/tc_abi.h declares; the host-side grid, wall-mask and starting
field code agrees with the oracle; bad bou.in values are rejected with a
message; the Fortran host binds the same ABI."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tc_abi.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(tc_[a-z_0-9]+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol(tc):
    lib = tc.lib()
    syms = declared_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    assert set(syms) == set(tc.ABI_SYMBOLS)


def test_params_struct_matches_header(tc):
    text = open(HEADER).read()
    body = text[text.index("typedef struct tc_params {"):text.index("} tc_params;")]
    names = re.findall(r"([a-z_0-9]+)\s*[,;]", body.split("{", 1)[1])
    assert names == [n for n, _ in tc.PARAM_FIELDS]


CONFIGS = [
    dict(nthm=64, nzm=64, nrm=65),
    dict(nthm=48, nzm=40, nrm=33, istr=0),
    dict(nthm=30, nzm=18, nrm=21, istr=6, str=4.0),
    dict(nthm=512, nzm=512, nrm=385, lamb=20.0),
]


@pytest.mark.parametrize("cfg", CONFIGS)
def test_grid_vectors_match_oracle_bitwise(tc, oracle, cfg):
    p = tc.configs.small_64()
    p.update(cfg)
    o = oracle.Oracle(p)
    for v in tc.VECS:
        a, b = tc.grid_vector(p, v), o.vec(v)
        assert a.shape == b.shape and np.array_equal(a, b), v


def test_perturbed_start_is_seeded_and_respects_walls(tc):
    p = tc.configs.scaled(tc.configs.small_64(), 16, 12, 17)
    a = tc.init_perturbed_couette(p, 42, 1e-2)
    b = tc.init_perturbed_couette(p, 42, 1e-2)
    c = tc.init_perturbed_couette(p, 43, 1e-2)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert not np.array_equal(a[0], c[0])
    qt, qr, qz = a
    assert np.all(qr[0] == 0) and np.all(qr[-1] == 0) and np.all(qt[-1] == 0) and np.all(qz[-1] == 0)
    rm = tc.grid_vector(p, "rm")[:17]
    eta = p["rint"] / p["rext"]
    A = (p["virotext"] - eta ** 2 * p["virotint"]) / (1 - eta ** 2)
    B = (p["virotint"] - p["virotext"]) * p["rint"] ** 2 / (1 - eta ** 2)
    dev = qt[:17] - (A * rm + B / rm)[:, None, None]
    assert 0.9e-2 < np.abs(dev).max() <= 1e-2 and abs(dev.mean()) < 1e-3
    assert 0.9e-2 < np.abs(qz[:17]).max() <= 1e-2


@pytest.mark.parametrize("bad,msg", [
    (dict(istr=5), "ISTR"),
    (dict(minstrt=2, multmin=1, minstrz=0), "MINSTRZ"),
    (dict(nsst=2), "NSST"),
    (dict(rext=0.5), "RINT"),
])
def test_bad_bou_values_rejected(tc, bad, msg):
    p = tc.configs.scaled(tc.configs.small_64(), 16, 12, 17)
    p.update(bad)
    with pytest.raises(tc.TCError, match=msg):
        tc.grid_vector(p, "rc")


def _has_gpu(tc):
    try:
        s = tc.Solver(tc.configs.scaled(tc.configs.small_64(), 8, 8, 9))
        s.close()
        return True
    except tc.TCError:
        return False


def test_restart_info_reads_the_documented_header(tc, tmp_path):
    """restart_info parses the header layout include/tc_abi.h documents for
    tc_write_restart files and rejects other files."""
    import struct
    path = tmp_path / "x.tcr"
    hdr = struct.pack("<8s8i8d", b"TCAFIDR1", 64, 48, 33, 3, 120, 1, 4, 0,
                      0.75, 2.5e-3, 0.4, 1.0, 2.0, 6.2831853, 1.0, 1.2)
    path.write_bytes(hdr + np.zeros(8).tobytes())
    assert tc.restart_info(str(path)) == dict(nthm=64, nzm=48, nrm=33, nsst=3, ntime=120,
                                              time=0.75, dt=2.5e-3)
    path.write_bytes(b"TCAFIDR0" + hdr[8:])
    with pytest.raises(tc.TCError):
        tc.restart_info(str(path))
    path.write_bytes(hdr[:50])
    with pytest.raises(tc.TCError):
        tc.restart_info(str(path))


def test_no_silent_fallback_without_device(tc):
    if _has_gpu(tc):
        pytest.skip("a GPU is visible")
    with pytest.raises(tc.TCError, match="no HIP device"):
        tc.Solver(tc.configs.small_64())


def test_fortran_host_binds_the_abi(tc, tmp_path):
    host = os.path.join(ROOT, "taylor-couette-afid_amd", "tc_host")
    if not os.path.exists(host):
        pytest.skip("tc_host not built")
    out = subprocess.run(["nm", "-u", host], capture_output=True, text=True).stdout
    used = set(re.findall(r"\b(tc_[a-z_0-9]+)\b", out))
    assert used <= set(declared_symbols()), used - set(declared_symbols())
    assert {"tc_create", "tc_explicit_terms_qt", "tc_solve_pressure_correction",
            "tc_time_marcher", "tc_calc_torque"} <= used
    if _has_gpu(tc):
        return
    # without a device the host stops at tc_create with the library's message
    from helpers import write_bou
    write_bou(tmp_path / "bou.in", tc.configs.scaled(tc.configs.small_64(), 8, 8, 9))
    r = subprocess.run([host], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "no HIP device" in r.stdout
