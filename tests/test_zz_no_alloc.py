"""owl_acoustic2d_forward_static, the forward that allocates no device memory,
against the allocating entry points (owl_acoustic2d_create +
owl_acoustic2d_forward) on the same inputs: traces bit-identical, on two calls
in a row (the second on a smaller problem, so stale contents of the shared
static scratch would show), and a run that needs more than the static scratch
is refused before anything is launched.
"""
import ctypes as C

import numpy as np
import pytest

from owl_amd import Acoustic2DBatch, Geometry, Receiver, Source, Window, _lib
from owl_amd._lib import lib

pytestmark = pytest.mark.gpu


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _model(nz, nx, seed):
    rng = np.random.default_rng(seed)
    z = np.arange(nz)[:, None] / max(nz - 1, 1)
    vp = (1800.0 + 1500.0 * z + 40.0 * rng.standard_normal((nz, nx))).astype(np.float32)
    rho = (1000.0 + 0.31 * vp).astype(np.float32)
    return np.broadcast_to(vp, (nz, nx)).copy(), rho


def _static(b):
    flat = np.full(sum(b.nr) * b.nt, np.nan, np.float32)
    rc = lib().owl_acoustic2d_forward_static(C.byref(b.cfg), _ptr(b._vp), _ptr(b._rho), b.nshot,
                                             C.cast(b._shots, C.c_void_p), _ptr(flat))
    return rc, b._split(flat)


def _case_large():
    nx, nz = 150, 110
    vp, rho = _model(nz, nx, 1)
    shots = [(Geometry(id=1, srcr=[Source(x=420.0, z=35.0, f0=18.0)],
                       recr=[Receiver(x=12.0 + 9.0 * k, z=20.0) for k in range(150)]), None),
             (Geometry(id=2, srcr=[Source(x=900.0, z=300.0, f0=12.0, mechanism="force", polar=0.4)],
                       recr=[Receiver(x=610.0 + 7.0 * k, z=55.0 + 2.0 * k) for k in range(60)]),
              Window(51, 140, 11, 100, 0.0, 0.0))]
    return Acoustic2DBatch(nx, nz, 10.0, 10.0, vp, rho, shots, 1e-3, 0.35, pml=12, fdhalf=4)


def _case_small():
    nx, nz = 64, 48
    vp, rho = _model(nz, nx, 2)
    shots = [(Geometry(id=7, srcr=[Source(x=315.0, z=205.0, f0=20.0)],
                       recr=[Receiver(x=20.0 + 10.0 * k, z=30.0) for k in range(55)]), None)]
    return Acoustic2DBatch(nx, nz, 10.0, 10.0, vp, rho, shots, 1e-3, 0.2, pml=10, fdhalf=3,
                           free_surface=True)


def test_static_forward_matches_the_allocating_forward(gpu):
    for name, make in (("two shots, uneven windows, 150x110", _case_large),
                       ("one shot, free surface, 64x48", _case_small)):
        print(f"case: static forward, {name}", flush=True)
        b = make()
        ref = b.forward()
        rc, got = _static(b)
        b.close()
        assert rc == _lib.OWL_OK, lib().owl_last_error()
        for r, g in zip(ref, got):
            assert np.abs(r).max() > 0
            assert np.array_equal(r, g)


def test_static_forward_refuses_a_run_larger_than_its_scratch(gpu):
    # 8 full-model shots of 600 x 600 need far more than 32 MiB of fields
    nx = nz = 600
    vp, rho = _model(nz, nx, 3)
    shots = [(Geometry(id=k + 1, srcr=[Source(x=100.0 + 500.0 * k, z=50.0, f0=10.0)],
                       recr=[Receiver(x=10.0 * q, z=20.0) for q in range(1, 500)]), None) for k in range(8)]
    b = Acoustic2DBatch(nx, nz, 10.0, 10.0, vp, rho, shots, 1e-3, 0.01, pml=20, fdhalf=4)
    print("case: static forward over its scratch limit", flush=True)
    rc, _ = _static(b)
    b.close()
    assert rc == _lib.OWL_ERR_NOMEM
    assert b"static device scratch" in lib().owl_last_error()
    # the library still works after the refusal
    s = _case_small()
    ref = s.forward()
    rc, got = _static(s)
    s.close()
    assert rc == _lib.OWL_OK
    assert np.array_equal(ref[0], got[0])
