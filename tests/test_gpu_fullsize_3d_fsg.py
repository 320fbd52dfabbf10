"""Full-size parity of the 3D acoustic, 3D elastic and 2D FSG workloads on
the GPU.  Each runs the benchmark batch itself, all shots and all steps:

* 3D acoustic: 256^3, 10 m, R = 4, dt 0.9 ms, 1500 steps, 20-cell PML,
  5-layer model, 16 shots with uneven windows (OWL's set_adaptive_range),
  100 x 100 receivers each, autotuned launch shape;
* 3D elastic: 200 x 200 x 150, 10 m, R = 4, dt 0.8 ms, 2500 steps, flat free
  surface, 8 vertical-force shots, 50 x 50 three-component receivers;
* 2D tilted-TI FSG: 801 x 401, 5 m, R = 4, dt 0.4 ms, 5000 steps, 64 shots
  with uneven windows (200-799 columns), without and with the flat free
  surface on OWL's depth-refined mesh (refinement 4, the benchmark's).

Checks: a trace sample at step t depends only on steps up to t, so the
float32 oracle run for the first n steps of a shot must equal the first n
samples of the GPU's full-length traces bit for bit (n sized so that the
oracle takes seconds; the smallest FSG window without the surface is checked
over all 5000 steps); and every shot modelled alone must give the same
traces, bit for bit, as inside the full batch.  The same prefixes from the
float64 oracle state the FP32 scheme's error against an FP64 build of the
same scheme at each workload (printed, and bounded by 1e-4 relative L2).
"""
import numpy as np
import pytest

from helpers import rel_l2
from owl_amd import Acoustic3DBatch, Elastic3DBatch, FSG2DBatch
from owl_amd.synthetic import (acoustic3d_layered_model, acoustic3d_shots, elastic3d_model, elastic3d_shots,
                               fsg2d_shots, fsg2d_tti_model)
from test_gpu_acoustic3d import _oracle as _oracle3
from test_gpu_elastic2d import _window_grid
from test_gpu_elastic3d import _oracle as _oracle_e3
from test_gpu_fsg2d import _oracle_args, _sub

pytestmark = pytest.mark.gpu


def _area(win, axes="xy"):
    a = 1
    for ax in axes:
        a *= win[f"n{ax}end"] - win[f"n{ax}beg"] + 1
    return a


def test_c2_acoustic3d_full_size(oracle32, oracle64, gpu):
    vp, rho = acoustic3d_layered_model()
    shots = acoustic3d_shots()
    dt, nt = 0.0009, 1500
    print("case: 3D acoustic 256^3, 16 shots, 1500 steps", flush=True)
    b = Acoustic3DBatch(256, 256, 256, 10.0, 10.0, 10.0, vp, rho, shots, dt, (nt - 1) * dt, pml=20, fdhalf=4)
    assert b.nt == nt
    shape = b.autotune(3)
    print(f"launch shape {shape}", flush=True)
    tr = b.forward()
    b.close()
    k = int(np.argmin([_area(w) for _, w in shots]))
    print(f"case: shot {k} alone, window {shots[k][1]}", flush=True)
    bs = Acoustic3DBatch(256, 256, 256, 10.0, 10.0, 10.0, vp, rho, [shots[k]], dt, (nt - 1) * dt, pml=20,
                         fdhalf=4)
    solo = bs.forward()[0]
    bs.close()
    assert np.abs(solo).max() > 0
    assert np.array_equal(solo, tr[k])
    n = 300
    print(f"case: oracle, shot {k}, first {n} steps", flush=True)
    ref = _oracle3(oracle32, vp, rho, shots[k][1], shots[k][0], 10.0, dt, n, 20, 4)
    assert np.abs(ref).max() > 0
    assert np.array_equal(tr[k][:, :n], ref), f"max diff {np.abs(tr[k][:, :n] - ref).max()}"
    ref64 = _oracle3(oracle64, vp, rho, shots[k][1], shots[k][0], 10.0, dt, n, 20, 4)
    e64 = rel_l2(tr[k][:, :n], ref64)
    print(f"C2 shot {k}, first {n} steps: FP32 GPU vs FP64 oracle rel-L2 {e64:.3e}", flush=True)
    assert e64 < 1e-4


def test_c3_elastic3d_full_size(oracle32, oracle64, gpu):
    vp, vs, rho = elastic3d_model()
    shots = elastic3d_shots()
    dt, nt = 0.0008, 2500
    print("case: 3D elastic 200x200x150, 8 shots, 2500 steps", flush=True)
    b = Elastic3DBatch(200, 200, 150, 10.0, 10.0, 10.0, vp, vs, rho, shots, dt, (nt - 1) * dt, pml=20, fdhalf=4,
                       free_surface=True)
    assert b.nt == nt
    tr = b.forward()
    b.close()
    k = 5
    print(f"case: shot {k} alone", flush=True)
    bs = Elastic3DBatch(200, 200, 150, 10.0, 10.0, 10.0, vp, vs, rho, [shots[k]], dt, (nt - 1) * dt, pml=20,
                        fdhalf=4, free_surface=True)
    solo = bs.forward()[0]
    bs.close()
    for a, c in zip(solo, tr[k]):
        assert np.abs(a).max() > 0
        assert np.array_equal(a, c)
    n = 100
    print(f"case: oracle, shot 0, first {n} steps", flush=True)
    ref = _oracle_e3(oracle32, vp, vs, rho, None, shots[0][0], 10.0, dt, n, 20, 4, True)
    for comp, (a, c) in enumerate(zip(tr[0], ref)):
        assert np.abs(c).max() > 0
        assert np.array_equal(a[:, :n], c), f"component {comp}: max diff {np.abs(a[:, :n] - c).max()}"
    ref64 = _oracle_e3(oracle64, vp, vs, rho, None, shots[0][0], 10.0, dt, n, 20, 4, True)
    e64 = rel_l2(np.concatenate([a[:, :n] for a in tr[0]]), np.concatenate(ref64))
    print(f"C3 shot 0, first {n} steps: FP32 GPU vs FP64 oracle rel-L2 {e64:.3e}", flush=True)
    assert e64 < 1e-4


def test_c4_fsg2d_full_size(oracle32, oracle64, gpu):
    cij, rho = fsg2d_tti_model()
    shots = fsg2d_shots()
    dt, nt = 0.0004, 5000
    print("case: FSG 801x401, 64 shots, 5000 steps", flush=True)
    b = FSG2DBatch(801, 401, 5.0, 5.0, cij, rho, shots, dt, (nt - 1) * dt, pml=20, fdhalf=4)
    assert b.nt == nt
    tr = b.forward()
    b.close()
    widths = [w["nxend"] - w["nxbeg"] + 1 for _, w in shots]
    small, large = int(np.argmin(widths)), int(np.argmax(widths))
    print(f"case: shot {small} alone", flush=True)
    bs = FSG2DBatch(801, 401, 5.0, 5.0, cij, rho, [shots[small]], dt, (nt - 1) * dt, pml=20, fdhalf=4)
    solo = bs.forward()[0]
    bs.close()
    for a, c in zip(solo, tr[small]):
        assert np.array_equal(a, c)
    for k, n in ((small, nt), (large, 1000)):
        geom, win = shots[k]
        print(f"case: oracle, shot {k} (width {widths[k]}), first {n} steps", flush=True)
        g, sl = _window_grid(rho, win, 5.0, dt, n, 20, 4)
        g["aniso"] = 1
        srcs, recs = _oracle_args(geom, g, 5.0)
        ref = oracle32.fsg2d_forward(g, _sub(cij, sl), np.ascontiguousarray(rho[sl]), srcs, recs)
        for comp, (a, c) in enumerate(zip(tr[k], ref)):
            assert np.abs(c).max() > 0
            assert np.array_equal(a[:, :n], c), f"shot {k} comp {comp}: max diff {np.abs(a[:, :n] - c).max()}"
        if k == small:
            ref64 = oracle64.fsg2d_forward(g, _sub(cij, sl), np.ascontiguousarray(rho[sl]), srcs, recs)
            e64 = rel_l2(np.concatenate([a[:, :n] for a in tr[k]]), np.concatenate(ref64))
            print(f"C4 shot {k}, {n} steps: FP32 GPU vs FP64 oracle rel-L2 {e64:.3e}", flush=True)
            assert e64 < 1e-4


def test_c4_fsg2d_free_surface_full_size(oracle32, oracle64, gpu):
    """The configs' FSG workload with its flat free surface on the refined
    mesh (the benchmark's batch): the smallest window's first 1500 steps
    bit-identical to the float32 oracle and stated against the float64 one;
    the widest window's first 300 steps bit-identical."""
    cij, rho = fsg2d_tti_model()
    shots = fsg2d_shots()
    dt, nt = 0.0004, 5000
    surf = dict(refine=4.0)
    print("case: FSG 801x401 with free surface, 64 shots, 5000 steps", flush=True)
    b = FSG2DBatch(801, 401, 5.0, 5.0, cij, rho, shots, dt, (nt - 1) * dt, pml=20, fdhalf=4, surface=surf)
    tr = b.forward()
    b.close()
    widths = [w["nxend"] - w["nxbeg"] + 1 for _, w in shots]
    small, large = int(np.argmin(widths)), int(np.argmax(widths))
    for k, n in ((small, 1500), (large, 300)):
        geom, win = shots[k]
        print(f"case: oracle with free surface, shot {k} (width {widths[k]}), first {n} steps", flush=True)
        g, sl = _window_grid(rho, win, 5.0, dt, n, 20, 4)
        g["aniso"] = 1
        srcs, recs = _oracle_args(geom, g, 5.0)
        ref = oracle32.fsg2d_forward_fs(g, surf, _sub(cij, sl), np.ascontiguousarray(rho[sl]), srcs, recs)
        for comp, (a, c) in enumerate(zip(tr[k], ref)):
            assert np.abs(c).max() > 0
            assert np.array_equal(a[:, :n], c), f"shot {k} comp {comp}: max diff {np.abs(a[:, :n] - c).max()}"
        if k == small:
            ref64 = oracle64.fsg2d_forward_fs(g, surf, _sub(cij, sl), np.ascontiguousarray(rho[sl]), srcs, recs)
            e64 = rel_l2(np.concatenate([a[:, :n] for a in tr[k]]), np.concatenate(ref64))
            print(f"C4 free surface, shot {k}, {n} steps: FP32 GPU vs FP64 oracle rel-L2 {e64:.3e}", flush=True)
            assert e64 < 1e-4
