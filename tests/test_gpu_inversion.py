"""FWI outer loop on the GPU (owl_amd.inversion.FWI2D over the C ABI):
the misfit-only entry equals the gradient call's misfit, and a few L-BFGS,
CG and steepest-descent iterations with OWL's line search, quadratic and
linear step sizes lower the misfit at every iteration and move the model
towards the true one."""
import numpy as np
import pytest

from owl_amd import OWL_GRAD_EXACT, OWL_GRAD_OWL, Acoustic2DBatch, Geometry, Receiver, Source
from owl_amd.inversion import FWI2D

pytestmark = pytest.mark.gpu

NX, NZ, DX = 90, 50, 10.0


def _models():
    z = np.arange(NZ)[:, None] * DX
    x = np.arange(NX)[None, :] * DX
    base = (1800.0 + 1.2 * z + 0.0 * x).astype(np.float32)
    bump = 150.0 * np.exp(-((z - 260.0) ** 2 + (x - 450.0) ** 2) / (2 * 60.0 ** 2))
    return (base + bump).astype(np.float32), base, np.full((NZ, NX), 2000.0, np.float32)


def _batch(vp, rho):
    shots = [Geometry(id=k + 1, srcr=[Source(x=100.0 + 230.0 * k, z=20.0, f0=12.0)],
                      recr=[Receiver(x=10.0 * r, z=20.0) for r in range(1, NX - 1)]) for k in range(4)]
    return Acoustic2DBatch(NX, NZ, DX, DX, vp, rho, [(g, None) for g in shots], 1e-3, 0.7, pml=15)


def test_misfit_entry_matches_gradient_misfit(gpu):
    vp, vp0, rho = _models()
    b = _batch(vp, rho)
    obs = b.forward()
    b.set_model(vp0, rho)
    print("case: misfit-only entry vs gradient misfit", flush=True)
    m1, syn = b.misfit(obs, want_traces=True)
    m2, _, _, syn2 = b.gradient(obs, mode=OWL_GRAD_OWL)
    b.close()
    assert np.all(m1 > 0)
    assert np.array_equal(m1, m2)
    for a, c in zip(syn, syn2):
        assert np.array_equal(a, c)


@pytest.mark.parametrize("search,step,mode", [("l-bfgs", "line-search", OWL_GRAD_EXACT),
                                              ("cg", "quadratic", OWL_GRAD_OWL),
                                              ("sd", "linear", OWL_GRAD_EXACT)])
def test_fwi_iterations_reduce_the_misfit(gpu, search, step, mode):
    vp, vp0, rho = _models()
    b = _batch(vp, rho)
    obs = b.forward()
    print(f"case: FWI {search} {step} mode {mode}", flush=True)
    # the near-surface part of the gradient is tapered off (OWL's 'taper'
    # processing) and steps stay in the range where the misfit is close to
    # quadratic (a few m/s for this 12 Hz problem)
    f = FWI2D(b, obs, vp0, rho, params={"vp": {"min": 1500.0, "max": 3000.0, "step_max": 8.0}},
              search=search, step_method=step, mode=mode, processing=("taper", "smooth"),
              process_kw=dict(taper_z=(100.0, 0.0), smooth=(20.0, 20.0)))
    err0 = np.linalg.norm(vp0 - vp)
    prev = None
    for it in range(3):
        m = f.iterate()
        m0, m1, s = f.misfits[-1]
        print(f"  iteration {it + 1}: misfit {m0:.4e} -> {m1:.4e}, step {s:.3g}", flush=True)
        assert m1 < m0
        if prev is not None:
            assert m0 == pytest.approx(prev, rel=1e-6)
        prev = m
    b.close()
    assert np.linalg.norm(f.p["vp"].m - vp) < err0
