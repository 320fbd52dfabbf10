"""Full-size parity of the 2D acoustic workloads on the GPU.

* 2D forward, 401 x 401, dx = dz = 10 m, R = 4, dt = 1 ms, 2000 steps,
  40-cell CFS-PML, one 10 Hz Ricker explosion at (x 2000, z 20), 391
  receivers at z = 20 every 10 m: traces bit-identical to the float32 oracle
  (target rel-L2 <= 1e-5); the float32 scheme's error against the float64
  oracle is printed and bounded.
* 2D FWI gradient, the benchmark batch itself: 32 shots in one batch, the
  autotuned fused kernels, one shared model, binomial checkpoints with 40
  snapshots, graph replays.  Traces of three shots are bit-identical to the
  oracle; OWL's gradient algorithm agrees per shot within 1e-4 (target 1e-3);
  the exact gradient agrees with the oracle's exact adjoint within 1e-5.  For
  the exact gradient, 29 shots get observed data equal to their own synthetic
  traces, so their residual is exactly zero and the batch gradient is the sum
  of the three remaining shots' gradients, which the oracle computes.
"""
import numpy as np
import pytest

from helpers import oracle_grid, oracle_receivers, oracle_sources, rel_l2
from owl_amd import (OWL_CKPT_BINOMIAL, OWL_GRAD_EXACT, OWL_STORE_F32, Acoustic2DBatch, Geometry, Receiver,
                     Source, Window)
from owl_amd.synthetic import acoustic2d_initial_model, acoustic2d_shots, acoustic2d_true_model

pytestmark = pytest.mark.gpu

NX = NZ = 401
DX = 10.0
DT = 0.001
TMAX = 1.999
PML = 40
R = 4
CHECK = (0, 15, 31)   # shots compared with the oracle
W = Window(1, NX, 1, NZ, 0.0, 0.0)


def _grid(nt):
    return oracle_grid(W, DX, DX, DT, nt, PML, R)


@pytest.fixture(scope="module")
def models():
    vp, rho = acoustic2d_true_model(NX, NZ, DX, seed=0)
    vp0, rho0 = acoustic2d_initial_model(vp, rho, 10.0)
    return vp, rho, vp0, rho0


def test_c0_forward_full_size(oracle32, oracle64, gpu, models):
    vp, rho, _, _ = models
    geom = Geometry(id=1, srcr=[Source(x=2000.0, z=20.0, f0=10.0)],
                    recr=[Receiver(x=50.0 + 10.0 * k, z=20.0) for k in range(391)])
    print("case: 2D forward 401x401, 2000 steps, 1 shot", flush=True)
    b = Acoustic2DBatch(NX, NZ, DX, DX, vp, rho, [(geom, None)], DT, TMAX, pml=PML, fdhalf=R)
    assert b.nt == 2000
    tr = b.forward()[0]
    b.close()
    g = _grid(2000)
    src, rec = oracle_sources(geom), oracle_receivers(geom, W, DX, DX)
    ref = oracle32.acoustic2d_forward(g, vp, rho, src, rec)
    assert np.abs(ref).max() > 0
    assert rel_l2(tr, ref) <= 1e-5
    assert np.array_equal(tr, ref), f"max |diff| {np.abs(tr - ref).max()}"
    ref64 = oracle64.acoustic2d_forward(g, vp, rho, src, rec)
    e64 = rel_l2(tr, ref64)
    print(f"C0 traces: FP32 GPU vs FP64 oracle rel-L2 {e64:.3e}", flush=True)
    assert e64 < 1e-4


@pytest.fixture(scope="module")
def c1(gpu, models, oracle32):
    """The benchmark batch: observe from the true model, switch to the
    initial model, autotune, and keep traces, OWL gradients and the exact
    gradient of the three checked shots."""
    vp, rho, vp0, rho0 = models
    shots = acoustic2d_shots(32)
    print("case: 2D FWI gradient batch, 32 shots 401x401, 2000 steps", flush=True)
    b = Acoustic2DBatch(NX, NZ, DX, DX, vp, rho, [(g, None) for g in shots], DT, TMAX, pml=PML, fdhalf=R)
    obs = b.forward()
    b.set_model(vp0, rho0)
    syn = b.forward(save_boundary=True)
    out = dict(shots=shots, obs=obs, syn=syn, nt=b.nt)
    # OWL's algorithm, per shot, from the boundary records of this forward
    adj = []
    for k in range(32):
        if k in CHECK:
            _, a = _l2_adjoint(oracle32, shots[k], syn[k], obs[k])
            adj.append(a)
        else:
            adj.append(np.zeros_like(syn[k]))
    grads = b.adjoint(adj)
    out["owl"] = {k: grads[k] for k in CHECK}
    out["adj"] = {k: adj[k] for k in CHECK}
    # exact gradient of the three shots through the full batch
    mixed = [obs[k] if k in CHECK else syn[k] for k in range(32)]
    b.set_observed(mixed)
    ex = dict(mode=OWL_GRAD_EXACT, ckpt=OWL_CKPT_BINOMIAL, ckpt_param=40, store=OWL_STORE_F32)
    runs = [b.gradient(**ex) for _ in range(3)]   # eager (autotunes), captured, replayed
    out["exact"] = runs
    out["rows"] = b.fused_rows()
    b.close()
    return out


def _l2_adjoint(o, geom, syn, obs):
    """OWL's L2 adjoint source syn - obs with the discard rule (float32
    oracle)."""
    rec = oracle_receivers(geom, W, DX, DX)
    return o.adjoint_source_l2([r["weight"] for r in rec], syn, obs)


def test_c1_traces_full_size(oracle32, c1, models):
    _, _, vp0, rho0 = models
    print(f"fused rows (forward, adjoint): {c1['rows']}", flush=True)
    for k in CHECK:
        geom = c1["shots"][k]
        ref = oracle32.acoustic2d_forward(_grid(c1["nt"]), vp0, rho0, oracle_sources(geom),
                                          oracle_receivers(geom, W, DX, DX))
        assert np.array_equal(c1["syn"][k], ref), f"shot {k}: rel {rel_l2(c1['syn'][k], ref)}"


def test_c1_owl_gradient_full_size(oracle32, c1, models):
    _, _, vp0, rho0 = models
    for k in CHECK:
        geom = c1["shots"][k]
        rv, rr, _ = oracle32.acoustic2d_gradient_owl(_grid(c1["nt"]), vp0, rho0, oracle_sources(geom),
                                                     oracle_receivers(geom, W, DX, DX), c1["adj"][k])
        gv, gr = c1["owl"][k]
        ev, er = rel_l2(gv, rv), rel_l2(gr, rr)
        print(f"shot {k}: OWL gradient rel-L2 vp {ev:.2e} rho {er:.2e}", flush=True)
        assert np.abs(rv).max() > 0
        assert ev <= 1e-4 and er <= 1e-4


def test_c1_exact_gradient_full_size(oracle32, c1, models):
    vp, rho, vp0, rho0 = models
    runs = c1["exact"]
    m0, gv0, gr0, syn0 = runs[0]
    for m, gv, gr, _ in runs[1:]:           # graph capture and replay equal eager
        assert np.array_equal(m, m0) and np.array_equal(gv, gv0) and np.array_equal(gr, gr0)
    for k in range(32):
        if k not in CHECK:
            assert m0[k] == 0.0
    sv = np.zeros_like(gv0)
    sr = np.zeros_like(gr0)
    for k in CHECK:
        geom = c1["shots"][k]
        rv, rr, mis, tr = oracle32.acoustic2d_gradient_exact(_grid(c1["nt"]), vp0, rho0, oracle_sources(geom),
                                                            oracle_receivers(geom, W, DX, DX), c1["obs"][k])
        assert np.array_equal(tr, syn0[k])
        assert abs(m0[k] - mis) <= 1e-6 * mis
        sv += rv
        sr += rr
    ev, er = rel_l2(gv0, sv), rel_l2(gr0, sr)
    print(f"exact gradient (3 shots in the 32-shot batch): rel-L2 vp {ev:.2e} rho {er:.2e}", flush=True)
    assert ev <= 1e-5 and er <= 1e-5


def test_c1_fp64_error_statement(oracle64, c1, models):
    """Error of the FP32 GPU scheme against the FP64 oracle at full size, one
    shot: traces, OWL's gradient and the exact gradient."""
    _, _, vp0, rho0 = models
    k = CHECK[1]
    geom = c1["shots"][k]
    g = _grid(c1["nt"])
    src, rec = oracle_sources(geom), oracle_receivers(geom, W, DX, DX)
    tr64 = oracle64.acoustic2d_forward(g, vp0, rho0, src, rec)
    e_tr = rel_l2(c1["syn"][k], tr64)
    rv, rr, _ = oracle64.acoustic2d_gradient_owl(g, vp0, rho0, src, rec, c1["adj"][k].astype(np.float64))
    gv, gr = c1["owl"][k]
    e_owl = (rel_l2(gv, rv), rel_l2(gr, rr))
    print(f"C1 shot {k} vs FP64 oracle: traces {e_tr:.3e}; OWL gradient vp {e_owl[0]:.3e} rho {e_owl[1]:.3e}",
          flush=True)
    assert e_tr < 1e-4
    assert max(e_owl) < 1e-3
