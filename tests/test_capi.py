"""The C ABI library loads, exports every symbol include/owl_amd.h declares,
and its host-side preparation (grid indices, Kaiser-sinc weights, STF, CFS-PML
profiles, CFL bound) is bit-identical to the oracle's.  No GPU needed."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from owl_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "owl_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(owl_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    declared = _declared()
    assert declared, "no declarations parsed"
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    assert set(declared) == set(_lib.exported_symbols())
    assert L.owl_version().decode().startswith("owl_amd")


def _locate(dims, x, y, z, ox, oy, oz, dx, dy, dz):
    idx = np.zeros(6, np.int32)
    w = np.zeros(54, np.float32)
    st = _lib.lib().owl_locate_point(dims, x, y, z, ox, oy, oz, dx, dy, dz,
                                     idx.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p))
    _lib.check(st)
    return idx, w.reshape(6, 9)


@pytest.mark.parametrize("x,z", [(2000.0, 20.0), (2005.0, 15.0), (2003.7, 21.9), (0.0, 0.0),
                                 (1234.5, 987.25), (44.99, 5.0001), (-35.0, 65.0)])
def test_locate_matches_oracle(oracle32, x, z):
    idx, w = _locate(2, x, 0.0, z, 0.0, 0.0, 0.0, 10.0, 1.0, 10.0)
    pt = oracle32.locate(x, z, 0.0, 0.0, 10.0, 10.0)
    assert (idx[0], idx[2], idx[3], idx[5]) == (pt.gx, pt.gz, pt.hx, pt.hz)
    for k, name in [(0, "ix"), (2, "iz"), (3, "hxw"), (5, "hzw")]:
        assert np.array_equal(w[k], np.array(getattr(pt, name)[:], np.float32)), name


def test_nint_half_away_from_zero():
    # x = 2005 m, dx = 10 m: (x - ox)/dx = 200.5 -> nint = 201, gx = 202
    idx, _ = _locate(2, 2005.0, 0.0, -5.0, 0.0, 0.0, 0.0, 10.0, 1.0, 10.0)
    assert idx[0] == 202
    # z = -5: -0.5 -> nint = -1, gz = 0
    assert idx[2] == 0


@pytest.mark.parametrize("f0,dt", [(10.0, 0.001), (8.0, 0.001), (10.0, 0.0008), (10.0, 0.0004),
                                   (15.0, 0.001), (25.0, 0.0005)])
def test_stf_matches_oracle(oracle32, f0, dt):
    hnt = C.c_int()
    n = _lib.lib().owl_source_time_function(0, f0, dt, 1.0, None, 0, C.byref(hnt))
    buf = np.zeros(n, np.float32)
    _lib.lib().owl_source_time_function(0, f0, dt, 1.0, buf.ctypes.data_as(C.c_void_p), n, C.byref(hnt))
    ref, href = oracle32.stf(f0, dt)
    assert n == len(ref) and hnt.value == href
    assert np.array_equal(buf, ref)


def test_stf_lengths_from_float32_arithmetic():
    # ceiling(2/f0/dt) evaluated left to right in float32
    got = []
    for f0, dt in [(10.0, 0.001), (8.0, 0.001), (10.0, 0.0008), (10.0, 0.0004)]:
        hnt = C.c_int()
        got.append((_lib.lib().owl_source_time_function(0, f0, dt, 1.0, None, 0, C.byref(hnt)), hnt.value))
    exp = []
    for f0, dt in [(10.0, 0.001), (8.0, 0.001), (10.0, 0.0008), (10.0, 0.0004)]:
        n = int(np.ceil(np.float32(np.float32(2.0) / np.float32(f0)) / np.float32(dt)))
        exp.append((n, int(np.floor(n / 2.0 + 0.5))))
    assert got == exp


@pytest.mark.parametrize("n,pml,d,dt,vp,f0", [(401, 40, 10.0, 0.001, 4100.0, 10.0),
                                              (256, 20, 10.0, 0.0009, 1500.0, 8.0),
                                              (64, 7, 5.0, 0.0004, 2500.0, 25.0)])
def test_cfspml_matches_oracle(oracle32, n, pml, d, dt, vp, f0):
    arrs = [np.zeros(n + 2 * pml, np.float32) for _ in range(4)]
    _lib.check(_lib.lib().owl_cfspml_axis(n, pml, d, dt, vp, f0,
                                          *[a.ctypes.data_as(C.c_void_p) for a in arrs]))
    ref = oracle32.cfspml_axis(n, pml, d, dt, vp, f0)
    for a, k in zip(arrs, ["ai", "bi", "ah", "bh"]):
        assert np.array_equal(a, ref[k]), k
    # masks: nonzero damping exactly inside the slabs
    i = np.arange(-pml + 1, n + pml + 1)
    assert np.array_equal(arrs[0] != 0, (i <= 0) | (i >= n + 1))
    assert np.array_equal(arrs[2] != 0, (i <= 1) | (i >= n + 1))


def test_dt_stable_matches_oracle(oracle32):
    for R in range(2, 9):
        for dims in (2, 3):
            a = _lib.lib().owl_dt_stable(R, dims, 4000.0, 10.0, 10.0, 10.0)
            b = oracle32.dt_stable(R, dims, 4000.0, 10.0, 10.0, 10.0)
            assert a == b


def test_errors_are_reported_not_fatal():
    st = _lib.lib().owl_cfspml_axis(0, 0, 0.0, 0.0, 0.0, 0.0, None, None, None, None)
    assert st == _lib.OWL_ERR_ARG
    assert b"invalid" in _lib.lib().owl_last_error()
