"""FWI outer loop (owl_amd.inversion; OWL's module_inversion_search_direction
.f90, module_inversion_step_size.f90, module_gradient.f90 processing), host
only.  Pinned by what the mathematics fixes:
* Polak-Ribiere CG with exact line searches minimises an n-dimensional convex
  quadratic in n iterations;
* OWL's L-BFGS two-loop recursion equals the dense BFGS inverse-Hessian
  update started from gamma I (gamma = s'y / y'y of the last pair) applied to
  the gradient, while the history is within the 10-pair memory, and keeps the
  last 10 pairs after that;
* the quadratic step is the vertex of the parabola through its three points;
  the line search finds a point below the starting misfit on convex
  functions and returns step 0 when no trial improves;
* the range check halves the scaling factor after each test (OWL's loop).
"""
import numpy as np
import pytest

from owl_amd import inversion as inv


def _quadratic(n, seed):
    rng = np.random.default_rng(seed)
    q = rng.standard_normal((n, n))
    A = q @ q.T + n * np.eye(n)
    b = rng.standard_normal(n)
    return A, b


def test_cg_minimises_a_quadratic_in_n_steps():
    n = 6
    A, b = _quadratic(n, 1)
    x = np.zeros(n)
    pg = ps = None
    for _ in range(n):
        g = (A @ x - b).astype(np.float32)
        d = inv.search_direction_cg(g, pg, ps).astype(np.float64)
        alpha = -(g @ d) / (d @ A @ d)
        x = x + alpha * d
        pg, ps = g, d.astype(np.float32)
    assert np.linalg.norm(A @ x - b) < 1e-3 * np.linalg.norm(b)


def test_cg_first_iteration_and_beta_clip():
    g = np.array([1.0, -2.0, 0.5], np.float32)
    assert np.array_equal(inv.search_direction_cg(g), -g)
    # <g, g - gp> < 0 -> beta = 0 -> steepest descent
    gp = np.array([3.0, -4.0, 1.0], np.float32)
    assert np.allclose(inv.search_direction_cg(g, gp, np.ones(3, np.float32)), -g)
    # zero previous gradient -> beta = 0
    assert np.allclose(inv.search_direction_cg(g, np.zeros(3, np.float32), np.ones(3, np.float32)), -g)


def _dense_bfgs(g, xs, gs):
    s = [xs[i + 1] - xs[i] for i in range(len(xs) - 1)]
    y = [gs[i + 1] - gs[i] for i in range(len(gs) - 1)]
    gamma = (s[-1] @ y[-1]) / (y[-1] @ y[-1])
    H = gamma * np.eye(len(g))
    for si, yi in zip(s, y):
        rho = 1.0 / (yi @ si)
        V = np.eye(len(g)) - rho * np.outer(yi, si)
        H = V.T @ H @ V + rho * np.outer(si, si)
    return -H @ g


@pytest.mark.parametrize("k", [2, 4, 10])
def test_lbfgs_equals_dense_bfgs(k):
    n = 12
    A, b = _quadratic(n, k)
    rng = np.random.default_rng(k + 10)
    xs = [rng.standard_normal(n) for _ in range(k)]
    gs = [A @ x - b for x in xs]
    # OWL's indexing: models[0..k-1] are the models of iterations 1..k, the
    # current gradient gs[k-1] is taken at models[k-1]
    d = inv.search_direction_lbfgs(gs[-1].astype(np.float32), [x.astype(np.float32) for x in xs],
                                   [g.astype(np.float32) for g in gs[:-1]])
    ref = _dense_bfgs(gs[-1], xs, gs)
    assert np.linalg.norm(d - ref) < 2e-4 * np.linalg.norm(ref)


def test_lbfgs_keeps_ten_pairs():
    n = 8
    A, b = _quadratic(n, 3)
    rng = np.random.default_rng(5)
    xs = [rng.standard_normal(n) for _ in range(14)]
    gs = [A @ x - b for x in xs]
    f = lambda a: [v.astype(np.float32) for v in a]  # noqa: E731
    full = inv.search_direction_lbfgs(gs[-1].astype(np.float32), f(xs), f(gs[:-1]))
    last = inv.search_direction_lbfgs(gs[-1].astype(np.float32), f(xs[-11:]), f(gs[-11:-1]))
    assert np.allclose(full, last, rtol=1e-5, atol=1e-6)
    ref = _dense_bfgs(gs[-1], xs[-11:], gs[-11:])
    assert np.linalg.norm(full - ref) < 2e-4 * np.linalg.norm(ref)


def _param(srch, step_max=100.0):
    p = inv.Parameter("vp", np.full(srch.shape, 2000.0, np.float32), 1000.0, 5000.0, step_max)
    p.srch = srch.astype(np.float32)
    return p


def test_initial_step_and_range_check():
    p = _param(np.array([0.5, -2.0, 1.0]))
    assert p.initial_step() == np.float32(50.0)
    # 0.2 is already suitable; OWL's loop halves once more after the test
    assert inv.suitable_scaling([p], 0.2) == pytest.approx(0.1)
    assert inv.suitable_scaling([p], 1.0) == pytest.approx(0.25)
    q = _param(np.zeros(3))
    assert q.initial_step() == 0.0
    # the update clips to [min, max]
    p.m[:] = 4990.0
    assert p.updated(-1.0).max() <= 5000.0 and p.updated(1.0).max() == 5000.0


def test_quadratic_step_is_the_parabola_vertex():
    p = _param(np.array([1.0, -1.0, 0.25]))
    f = lambda s: 3.0 * (s - 0.21) ** 2 + 1.5  # noqa: E731
    step, mis = inv.step_quadratic([p], f(0.0), f)
    assert step == pytest.approx(0.21, rel=1e-9)
    assert mis == pytest.approx(1.5)


def test_line_search_on_convex_functions():
    p = _param(np.array([1.0, -1.0, 0.25]))
    calls = []

    def f(s):
        calls.append(s)
        return 2.0 * (s - 0.4) ** 2 + 0.3
    step, mis = inv.step_linesearch([p], f(0.0), f)
    assert mis < f(0.0) and mis == pytest.approx(f(step))
    assert len(calls) <= inv.NSEARCH_MAX + 1

    # the first trial (the suitable factor: 1.0 fails the range check, 0.5
    # passes, then one more halving) already lowers the misfit: taken at once
    step2, _ = inv.step_linesearch([p], 10.0, lambda s: 10.0 - s)
    assert step2 == pytest.approx(0.25)

    # nothing below the starting misfit: step 0
    step3, mis3 = inv.step_linesearch([p], 1.0, lambda s: 1.0 + s)
    assert step3 == 0.0 and mis3 == 1.0


def test_linear_step_coefficient():
    # one shot, two traces; the trial data move by s * u and the residual is 3 u
    p = _param(np.array([1.0, -1.0, 0.25]))
    u = np.array([[1.0, 2.0, 0.0, -1.0], [0.5, 0.0, 0.5, 1.0]], np.float32)
    d0 = np.ones_like(u)
    obs = [d0 + 3.0 * u]
    traces = lambda s: [d0 + np.float32(s) * u]  # noqa: E731
    misfit = lambda s: float(np.sum((obs[0] - traces(s)[0]) ** 2))  # noqa: E731
    step, mis = inv.step_linear([p], misfit(0.0), traces, obs, misfit)
    # trial at s0 = 0.1: err23 = s0^2 |u_r|^2, err13 = 9 |u_r|^2 per trace r,
    # so sum1 / sum2 = 9 / s0^2 = 900; the search starts from twice that and
    # halves, trials outside the step range not counted, until the
    # perturbation is below step_max (s < 1): 1800 / 2^11
    assert mis < misfit(0.0)
    assert step == pytest.approx(1800.0 / 2 ** 11, rel=1e-5)


def test_gradient_processing():
    rng = np.random.default_rng(0)
    w = rng.standard_normal((20, 30)).astype(np.float32)
    assert np.array_equal(inv.process_gradient(w, ["scale"], 10.0, 10.0, scale=2.0), 2 * w)
    mb = inv.process_gradient(w, ["max_balance"], 10.0, 10.0)
    assert mb.max() == pytest.approx(1.0)
    m = (rng.random((20, 30)) > 0.5).astype(np.float32)
    assert np.array_equal(inv.process_gradient(w, ["mask"], 10.0, 10.0, mask=m), w * m)
    c = np.full((20, 30), 3.0, np.float32)
    assert np.allclose(inv.process_gradient(c, ["smooth"], 10.0, 10.0, smooth=(30.0, 20.0)), 3.0)
    t = inv.process_gradient(c, ["taper"], 10.0, 10.0, taper_x=(50.0, 0.0), taper_z=(0.0, 40.0))
    assert t[10, 0] == pytest.approx(0.0, abs=1e-6) and t[10, -1] == 3.0 and t[0, 15] == 3.0
    assert t[-1, 15] == pytest.approx(0.0, abs=1e-6)
    assert np.all(np.diff(t[10, :6]) > 0)
    assert not inv.process_gradient(w, ["scale"], 10.0, 10.0, it=5, update_iter=(1, 3)).any()
    rb = inv.process_gradient(w, ["rms_balance"], 10.0, 10.0)
    assert np.sqrt(np.mean(rb.astype(np.float64) ** 2)) == pytest.approx(1.0, rel=1e-5)
    with pytest.raises(ValueError):
        inv.process_gradient(w, ["nonsense"], 10.0, 10.0)
