"""GPU exact adjoint (the differentiable path) against the oracle's exact
adjoint, against itself under every storage policy, against finite
differences, and through torch.autograd.  Tolerances: FP32 states agree
with the oracle to rel-L2 1e-5; checkpoint recomputation is bit-identical
to full storage; FP16 / BF16 snapshots bound the gradient error at 5e-3 /
3e-2 (stated in the test); finite differences on 10 seeded perturbations
agree to 1e-2 in FP32."""
import numpy as np
import pytest

from helpers import oracle_grid, oracle_receivers, oracle_sources, rel_l2, simple_geometry
from owl_amd import (OWL_CKPT_BINOMIAL, OWL_CKPT_FULL, OWL_CKPT_UNIFORM, OWL_GRAD_EXACT,
                     OWL_STORE_BF16, OWL_STORE_F16, OWL_STORE_F32, Acoustic2DBatch, Window)
from owl_amd.synthetic import gardner, random_field

pytestmark = pytest.mark.gpu

NZ, NX, DX, DT, TMAX, PML = 36, 48, 10.0, 0.001, 0.24, 10


def _case(seed=2):
    vp = (2000.0 * (1 + 0.05 * random_field((NZ, NX), 4.0, seed))).astype(np.float32)
    rho = gardner(vp)
    rx = 15.0 + 10.0 * np.arange(45)
    geoms = [simple_geometry(1, 200.3, 150.7, rx, np.full(45, 40.5), f0=20.0),
             simple_geometry(2, 301.0, 120.0, rx, np.full(45, 40.5), f0=20.0)]
    return vp, rho, geoms


def _batch(vp, rho, geoms):
    return Acoustic2DBatch(NX, NZ, DX, DX, vp, rho, [(g, None) for g in geoms], DT, TMAX, pml=PML)


def _obs(oracle32, vp, rho, geoms, nt):
    w = Window(1, NX, 1, NZ, 0.0, 0.0)
    grid = oracle_grid(w, DX, DX, DT, nt, PML, 4)
    return [oracle32.acoustic2d_forward(grid, (vp * 1.02).astype(np.float32), rho, oracle_sources(g),
                                        oracle_receivers(g, w, DX, DX)) for g in geoms]


def test_exact_gradient_matches_oracle(oracle32, gpu):
    vp, rho, geoms = _case()
    b = _batch(vp, rho, geoms)
    obs = _obs(oracle32, vp, rho, geoms, b.nt)
    print("case: exact gradient, full FP32 storage", flush=True)
    mis, gv, gr, syn = b.gradient(obs, mode=OWL_GRAD_EXACT, ckpt=OWL_CKPT_FULL)
    w = Window(1, NX, 1, NZ, 0.0, 0.0)
    sv = np.zeros_like(gv)
    sr = np.zeros_like(gr)
    for k, g in enumerate(geoms):
        grid = oracle_grid(w, DX, DX, DT, b.nt, PML, 4)
        rv, rr, m, tr = oracle32.acoustic2d_gradient_exact(grid, vp, rho, oracle_sources(g),
                                                           oracle_receivers(g, w, DX, DX), obs[k])
        assert np.array_equal(tr, syn[k])
        assert abs(mis[k] - m) <= 1e-6 * m
        sv += rv
        sr += rr
    assert rel_l2(gv, sv) < 1e-5, rel_l2(gv, sv)
    assert rel_l2(gr, sr) < 1e-5, rel_l2(gr, sr)
    b.close()


@pytest.mark.parametrize("ckpt,param", [(OWL_CKPT_UNIFORM, 7), (OWL_CKPT_UNIFORM, 1),
                                        (OWL_CKPT_BINOMIAL, 4), (OWL_CKPT_BINOMIAL, 0),
                                        (OWL_CKPT_BINOMIAL, 1), (OWL_CKPT_BINOMIAL, 2),
                                        (OWL_CKPT_BINOMIAL, 40)])
def test_checkpointing_is_bit_identical_to_full_storage(oracle32, gpu, ckpt, param):
    vp, rho, geoms = _case(3)
    b = _batch(vp, rho, geoms)
    obs = _obs(oracle32, vp, rho, geoms, b.nt)
    print(f"case: checkpoint policy {ckpt} param {param}", flush=True)
    m0, gv0, gr0, _ = b.gradient(obs, mode=OWL_GRAD_EXACT, ckpt=OWL_CKPT_FULL)
    m1, gv1, gr1, _ = b.gradient(obs, mode=OWL_GRAD_EXACT, ckpt=ckpt, ckpt_param=param)
    assert np.array_equal(m0, m1)
    assert np.array_equal(gv0, gv1)
    assert np.array_equal(gr0, gr1)
    b.close()


@pytest.mark.parametrize("ckpt,param,store,tol", [
    (OWL_CKPT_FULL, 0, OWL_STORE_F16, 5e-3), (OWL_CKPT_FULL, 0, OWL_STORE_BF16, 3e-2),
    (OWL_CKPT_BINOMIAL, 5, OWL_STORE_F16, 5e-3), (OWL_CKPT_UNIFORM, 10, OWL_STORE_BF16, 3e-2)])
def test_reduced_precision_storage(oracle32, gpu, ckpt, param, store, tol):
    vp, rho, geoms = _case(4)
    b = _batch(vp, rho, geoms)
    obs = _obs(oracle32, vp, rho, geoms, b.nt)
    print(f"case: {ckpt}/{param} storage {store}", flush=True)
    _, gv0, gr0, _ = b.gradient(obs, mode=OWL_GRAD_EXACT, ckpt=OWL_CKPT_FULL)
    _, gv1, gr1, _ = b.gradient(obs, mode=OWL_GRAD_EXACT, ckpt=ckpt, ckpt_param=param, store=store)
    print(f"rel-L2 vp {rel_l2(gv1, gv0):.2e} rho {rel_l2(gr1, gr0):.2e}")
    assert rel_l2(gv1, gv0) < tol
    assert rel_l2(gr1, gr0) < tol
    b.close()


def test_exact_gradient_finite_differences(oracle32, gpu):
    """10 seeded smooth perturbations (zero on the edge lines, which set the
    PML strength): central differences of 1/2|d - obs|^2 from the GPU forward
    agree with <g, dm>."""
    vp, rho, geoms = _case(5)
    b = _batch(vp, rho, geoms)
    obs = _obs(oracle32, vp, rho, geoms, b.nt)
    print("case: finite differences, 10 perturbations", flush=True)
    _, gv, gr, _ = b.gradient(obs, mode=OWL_GRAD_EXACT, ckpt=OWL_CKPT_BINOMIAL, ckpt_param=8)
    taper = np.zeros((NZ, NX))
    taper[1:-1, 1:-1] = 1.0

    def phi(v, r):
        b.set_model(v, r)
        tr = b.forward()
        return 0.5 * sum(float(np.sum((t.astype(np.float64) - o) ** 2)) for t, o in zip(tr, obs))

    for seed in range(10):
        dm = random_field((NZ, NX), 3.0, 200 + seed) * taper
        for which in ("vp", "rho"):
            base = vp if which == "vp" else rho
            step = 2e-3 * float(np.abs(base).max()) * dm
            plus = (base + step).astype(np.float32)
            minus = (base - step).astype(np.float32)
            dact = (plus.astype(np.float64) - minus) / 2
            if which == "vp":
                fd = (phi(plus, rho) - phi(minus, rho)) / 2
                an = float(np.sum(gv * dact))
            else:
                fd = (phi(vp, plus) - phi(vp, minus)) / 2
                an = float(np.sum(gr * dact))
            assert abs(fd / an - 1) < 1e-2, (seed, which, fd, an)
    b.close()


@pytest.mark.parametrize("ckpt,param", [(OWL_CKPT_UNIFORM, 9), (OWL_CKPT_BINOMIAL, 5)])
def test_autograd_backward_equals_gradient(oracle32, gpu, ckpt, param):
    import torch
    from owl_amd.autograd import propagate
    vp, rho, geoms = _case(6)
    b = _batch(vp, rho, geoms)
    obs = _obs(oracle32, vp, rho, geoms, b.nt)
    print(f"case: torch.autograd through the exact adjoint, policy {ckpt} param {param}", flush=True)
    _, gv0, gr0, _ = b.gradient(obs, mode=OWL_GRAD_EXACT, ckpt=OWL_CKPT_FULL)
    tv = torch.tensor(vp, dtype=torch.float64, device="cuda:0", requires_grad=True)
    tr_ = torch.tensor(rho, dtype=torch.float64, device="cuda:0", requires_grad=True)
    d = propagate(tv, tr_, b, ckpt, param, OWL_STORE_F32)
    o = torch.tensor(np.concatenate(obs, 0), device="cuda:0")
    loss = 0.5 * ((d.double() - o.double()) ** 2).sum()
    loss.backward()
    assert rel_l2(tv.grad.cpu().numpy(), gv0) < 1e-6
    assert rel_l2(tr_.grad.cpu().numpy(), gr0) < 1e-6
    b.close()


@pytest.mark.parametrize("R,zrows,arows", [(4, 32, 8), (2, 7, 3), (8, 16, 16), (5, 1, 5)])
def test_fused_step_is_bit_identical_to_sweeps(gpu, R, zrows, arows):
    """The fused z-marching forward and adjoint steps against the two sweeps
    and the two-pass adjoint: five shots on uneven windows, traces, misfits
    and the vp gradient bit-identical for full FP32 and FP16 storage, uniform
    and binomial checkpoints, at several z-row counts of either kernel.  The
    fused adjoint sums the four density terms of a cell per step before its
    FP64 accumulation (the two-pass adjoint accumulates them apart), so the
    rho gradient agrees to FP64 rounding: rel-L2 <= 1e-12."""
    from owl_amd.synthetic import gardner, random_field
    nz, nx, dx = 60, 140, 10.0
    vp = (2000.0 * (1 + 0.05 * random_field((nz, nx), 4.0, 11))).astype(np.float32)
    rho = gardner(vp)
    wins = [Window(1, 140, 1, 60, 0.0, 0.0), Window(11, 70, 1, 45, 100.0, 0.0),
            Window(50, 131, 5, 60, 490.0, 40.0), Window(30, 62, 1, 33, 290.0, 0.0),
            Window(3, 96, 2, 58, 20.0, 10.0)]
    geoms = []
    for k, w in enumerate(wins):
        x0 = (w.nxbeg - 1) * dx
        rx = x0 + 5.0 + 10.0 * np.arange(w.nx - 1)
        sx = x0 + 0.5 * (w.nx - 1) * dx + 3.3
        sz = (w.nzbeg - 1) * dx + 80.0
        geoms.append(simple_geometry(k + 1, sx, sz, rx, np.full(len(rx), (w.nzbeg - 1) * dx + 30.0), f0=20.0))
    pml = max(10, R)
    b = Acoustic2DBatch(nx, nz, dx, dx, vp, rho, list(zip(geoms, wins)), 0.001, 0.15, pml=pml, fdhalf=R)
    obs = b.forward()
    b.set_model((vp * 0.98).astype(np.float32), rho)
    cases = [(OWL_CKPT_FULL, 0, OWL_STORE_F32), (OWL_CKPT_FULL, 0, OWL_STORE_F16),
             (OWL_CKPT_UNIFORM, 9, OWL_STORE_F32), (OWL_CKPT_BINOMIAL, 6, OWL_STORE_F32),
             (OWL_CKPT_BINOMIAL, 5, OWL_STORE_F16)]
    for ckpt, param, store in cases:
        print(f"case: fused vs sweeps R={R} zrows={zrows} arows={arows} ckpt={ckpt} param={param} store={store}",
              flush=True)
        out = []
        for mode in ("sweeps", "fused"):
            b.set_step_kernel(mode, zrows)
            b.set_adjoint_rows(arows)
            b.set_profiling(True)
            out.append(b.gradient(obs, mode=OWL_GRAD_EXACT, ckpt=ckpt, ckpt_param=param, store=store))
            st = b.kernel_stats()
            b.set_profiling(False)
            assert (st["step_fused"]["launches"] > 0) == (mode != "sweeps")
        m0, gv0, gr0, s0 = out[0]
        assert np.abs(gv0).max() > 0
        for m1, gv1, gr1, s1 in out[1:]:
            assert np.array_equal(m0, m1)
            for x, y in zip(s0, s1):
                assert np.array_equal(x, y)
            assert np.array_equal(gv0, gv1)
            assert rel_l2(gr1, gr0) <= 1e-12, rel_l2(gr1, gr0)
    b.close()


def test_graph_replays_are_bit_identical_to_eager(gpu):
    """The exact-adjoint gradient captured into a HIP graph: four replays,
    with eager forward modelling and synchronisations between them, each
    bit-identical to the eager result, for binomial FP32 and uniform FP16
    storage on a five-shot batch (fused steps)."""
    from owl_amd.synthetic import gardner, random_field
    nz, nx, dx = 50, 90, 10.0
    vp = (2000.0 * (1 + 0.05 * random_field((nz, nx), 4.0, 12))).astype(np.float32)
    rho = gardner(vp)
    rx = 15.0 + 10.0 * np.arange(85)
    geoms = [simple_geometry(k + 1, 100.0 + 150.0 * k, 60.0, rx, np.full(85, 30.0), f0=20.0) for k in range(5)]
    b = Acoustic2DBatch(nx, nz, dx, dx, vp, rho, [(g, None) for g in geoms], 0.001, 0.15, pml=10)
    obs = b.forward()
    b.set_model((vp * 0.98).astype(np.float32), rho)
    for ckpt, param, store in [(OWL_CKPT_BINOMIAL, 7, OWL_STORE_F32), (OWL_CKPT_UNIFORM, 11, OWL_STORE_F16)]:
        print(f"case: graph replays ckpt={ckpt} param={param} store={store}", flush=True)
        b.set_graphs(False)
        ref = b.gradient(obs, mode=OWL_GRAD_EXACT, ckpt=ckpt, ckpt_param=param, store=store)
        b.set_graphs(True)
        for k in range(6):   # eager (signature), capture + launch, four replays
            got = b.gradient(None, mode=OWL_GRAD_EXACT, ckpt=ckpt, ckpt_param=param, store=store)
            assert np.array_equal(got[0], ref[0]), k
            assert np.array_equal(got[1], ref[1]), k
            assert np.array_equal(got[2], ref[2]), k
            for x, y in zip(got[3], ref[3]):
                assert np.array_equal(x, y), k
            if k % 2:
                b.forward()   # eager work between replays
    b.set_graphs(False)
    b.close()


def test_autotuned_rows_give_the_same_gradient(gpu):
    """The z-row autotuner picks rows from its candidate sets, and the
    autotuned batch's exact gradient equals one with fixed rows: the misfit
    and the vp gradient bit for bit, the rho gradient to FP64 rounding."""
    from owl_amd.synthetic import gardner, random_field
    nz, nx, dx = 60, 140, 10.0
    vp = (2000.0 * (1 + 0.05 * random_field((nz, nx), 4.0, 12))).astype(np.float32)
    rho = gardner(vp)
    wins = [Window(1, 140, 1, 60, 0.0, 0.0), Window(11, 70, 1, 45, 100.0, 0.0),
            Window(50, 131, 5, 60, 490.0, 40.0), Window(30, 62, 1, 33, 290.0, 0.0)]
    geoms = []
    for k, w in enumerate(wins):
        x0 = (w.nxbeg - 1) * dx
        rx = x0 + 5.0 + 10.0 * np.arange(w.nx - 1)
        geoms.append(simple_geometry(k + 1, x0 + 0.5 * (w.nx - 1) * dx + 3.3, (w.nzbeg - 1) * dx + 80.0, rx,
                                     np.full(len(rx), (w.nzbeg - 1) * dx + 30.0), f0=20.0))
    print("case: autotuned fused rows", flush=True)
    b = Acoustic2DBatch(nx, nz, dx, dx, vp, rho, list(zip(geoms, wins)), 0.001, 0.15, pml=10, fdhalf=4)
    obs = b.forward()
    b.set_model((vp * 0.98).astype(np.float32), rho)
    b.set_graphs(False)
    kw = dict(mode=OWL_GRAD_EXACT, ckpt=OWL_CKPT_BINOMIAL, ckpt_param=6, store=OWL_STORE_F32)
    m0, gv0, gr0, _ = b.gradient(obs, **kw)
    z, a = b.fused_rows()
    assert z in (8, 12, 16, 24, 32) and a in (4, 6, 8, 12, 16)
    b.set_step_kernel("fused", 16)
    b.set_adjoint_rows(8)
    m1, gv1, gr1, _ = b.gradient(obs, **kw)
    assert np.array_equal(m0, m1)
    assert np.abs(gv0).max() > 0
    # the density terms are split between accumulators by chunk, so the rho
    # gradient depends on the rows at FP64 rounding only
    assert np.array_equal(gv0, gv1) and rel_l2(gr1, gr0) <= 1e-12, rel_l2(gr1, gr0)
    b.close()


def test_device_path_equals_host_path(gpu):
    """The device entry points (model, traces and gradients in device memory)
    give the host path's traces and float gradients bit for bit, for both
    model layouts, and a second model whose edge velocities move the PML
    profiles equals a batch planned on that model."""
    import torch
    vp, rho, geoms = _case(8)
    vp2 = (vp * 1.07).astype(np.float32)  # moves every window's edge velocity
    dtr = None
    for k, (v, r) in enumerate(((vp, rho), (vp2, rho))):
        print(f"case: device vs host differentiable path, model {k}", flush=True)
        host = _batch(v, r, geoms)
        dev = _batch(vp, rho, geoms)  # planned on the first model, updated on the device
        tr_h = np.concatenate(host.forward_store(OWL_CKPT_BINOMIAL, 5, OWL_STORE_F32), 0)
        if dtr is None:
            dtr = np.random.default_rng(3).standard_normal(tr_h.shape).astype(np.float32)
        gv_h, gr_h = host.backward(dtr)
        for xf in (True, False):
            tv = torch.tensor(v if xf else np.ascontiguousarray(v.T), device="cuda:0")
            tr_ = torch.tensor(r if xf else np.ascontiguousarray(r.T), device="cuda:0")
            out = torch.empty(tr_h.shape, dtype=torch.float32, device="cuda:0")
            torch.cuda.synchronize()
            dev.set_model_device(tv.data_ptr(), tr_.data_ptr(), x_fastest=xf)
            dev.forward_store_device(out.data_ptr(), OWL_CKPT_BINOMIAL, 5, OWL_STORE_F32)
            assert np.array_equal(out.cpu().numpy(), tr_h)
            g = torch.tensor(dtr, device="cuda:0")
            shape = (NZ, NX) if xf else (NX, NZ)
            gv = torch.empty(shape, dtype=torch.float32, device="cuda:0")
            gr = torch.empty(shape, dtype=torch.float32, device="cuda:0")
            torch.cuda.synchronize()
            dev.backward_device(g.data_ptr(), gv.data_ptr(), gr.data_ptr(), x_fastest=xf)
            gvn, grn = gv.cpu().numpy(), gr.cpu().numpy()
            if not xf:
                gvn, grn = gvn.T, grn.T
            assert np.abs(gv_h).max() > 0
            assert np.array_equal(gvn, gv_h.astype(np.float32))
            assert np.array_equal(grn, gr_h.astype(np.float32))
        host.close()
        dev.close()


@pytest.mark.parametrize("ckpt,param,store", [(OWL_CKPT_BINOMIAL, 5, OWL_STORE_F32),
                                              (OWL_CKPT_BINOMIAL, 7, OWL_STORE_F16),
                                              (OWL_CKPT_UNIFORM, 9, OWL_STORE_F32)])
def test_split_calls_equal_the_joint_gradient(oracle32, gpu, ckpt, param, store):
    """forward_store then backward(syn - obs) gives the joint exact gradient
    call's vp and rho gradients bit for bit: the split binomial schedule
    resumes where its forward call stopped instead of starting again."""
    vp, rho, geoms = _case(9)
    b = _batch(vp, rho, geoms)
    obs = _obs(oracle32, vp, rho, geoms, b.nt)
    print(f"case: split vs joint exact gradient, ckpt {ckpt} param {param} store {store}", flush=True)
    _, gv0, gr0, syn0 = b.gradient(obs, mode=OWL_GRAD_EXACT, ckpt=ckpt, ckpt_param=param, store=store)
    syn = b.forward_store(ckpt, param, store)
    for x, y in zip(syn, syn0):
        assert np.array_equal(x, y)
    gv, gr = b.backward([s - o for s, o in zip(syn, obs)])
    assert np.abs(gv0).max() > 0
    assert np.array_equal(gv, gv0)
    assert np.array_equal(gr, gr0)
    b.close()
