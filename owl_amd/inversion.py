"""FWI outer loop around the GPU gradient: search directions, step-size
searches, gradient processing and the iteration driver.

Mirrors OWL's src/inverse modules, array for array:
* search directions (module_inversion_search_direction.f90:39-342):
  steepest descent, Polak-Ribiere conjugate gradient with beta >= 0, and
  L-BFGS with the two-loop recursion over the last 10 (model, gradient)
  pairs;
* step sizes (module_inversion_step_size.f90:51-1171): the initial step
  max_perturbation / max|srch| per parameter, the range check that halves the
  scaling factor until every perturbation stays below step_max, the
  'linear', 'quadratic' and 'line-search' searches, and the model update with
  clipping to [min, max];
* gradient processing (src/inverse/2d/module_gradient.f90:738-900): scale,
  max_balance, rms_balance, smooth, taper and mask, in the order given, with
  the per-parameter update-iteration window.

The forward and gradient evaluations are the caller's (for 2D acoustic FWI,
FWI2D below runs them through Acoustic2DBatch on the GPU); this module holds
the host logic between them.  OWL keeps every array on disk between stages
(dir_iter_model); here they stay in memory.  Sums are float32 as in OWL's
Fortran (numpy's pairwise order, not Fortran's sequential one).
"""
import math

import numpy as np

NSTEP_LBFGS = 10       # nstep_lbfgs, module_inversion_search_direction.f90:31
NSEARCH_MAX = 6        # nsearch_max, module_inversion_step_size.f90:41
FLOAT_TINY = np.finfo(np.float32).tiny
FLOAT_HUGE = np.finfo(np.float32).max

# default step_max per parameter family (compute_initial_step_size,
# module_inversion_step_size.f90:231-258)
STEP_MAX_DEFAULT = {"vp": 100.0, "vs": 100.0, "rho": 100.0, "epsilon": 0.1, "delta": 0.1, "gamma": 0.1,
                    "eta": 0.1, "theta": 0.1 * math.pi / 2, "phi": 0.1 * math.pi / 2, "c": 1.0e9, "mt": 1.0}


def _f32(a):
    return np.asarray(a, np.float32)


def _dot(a, b):
    return np.float32(np.sum(_f32(a) * _f32(b), dtype=np.float32))


# ---------------------------------------------------------------------------
# Search directions
# ---------------------------------------------------------------------------
def search_direction_sd(grad):
    """Steepest descent: -grad."""
    return -_f32(grad)


def search_direction_cg(grad, prev_grad=None, prev_srch=None):
    """Polak-Ribiere conjugate gradient (compute_search_direction_cg_single_
    parameter): beta = max(0, <g, g - g_prev> / <g_prev, g_prev>), 0 when the
    denominator is 0; -grad on the first iteration."""
    g = _f32(grad)
    if prev_grad is None:
        return -g
    pg, ps = _f32(prev_grad), _f32(prev_srch)
    s1 = _dot(pg, pg)
    s2 = _dot(g, g - pg)
    beta = np.float32(0.0) if s1 == 0 else np.float32(max(np.float32(0.0), s2 / s1))
    return -g + beta * ps


def search_direction_lbfgs(grad, models, grads, nstep=NSTEP_LBFGS):
    """L-BFGS two-loop recursion (compute_search_direction_lbfgs_single_
    parameter).  models[i], grads[i] for i = 0 .. it-1 are the model the
    gradient of iteration i+1 was taken at (models[0] = the starting model)
    and that gradient; grad is the current one (iteration it+1, at
    models[it]).  models needs one more entry than grads: len(models) ==
    len(grads) + 1.  With one gradient only (first iteration): -grad."""
    q = _f32(grad).copy()
    it = len(grads) + 1          # OWL's iter
    if it == 1:
        return -q
    if len(models) != len(grads) + 1:
        raise ValueError("L-BFGS needs the model of every gradient plus the current model")
    G = list(grads) + [q.copy()]          # G[i-1] = gradient of iteration i
    X = list(models)                      # X[i-1] = model of iteration i

    def pair(i):
        # s_i, y_i between iterations i and i+1 (1-based as in OWL)
        return _f32(X[i]) - _f32(X[i - 1]), _f32(G[i]) - _f32(G[i - 1])

    lo = max(it - nstep, 1)
    alpha = []
    for i in range(it - 1, lo - 1, -1):
        si, yi = pair(i)
        s1, s2 = _dot(si, q), _dot(yi, si)
        a = np.float32(0.0) if s2 == 0 else np.float32(s1 / s2)
        alpha.append(a)
        q = q - a * yi
    si, yi = pair(it - 1)
    s1, s2 = _dot(si, yi), _dot(yi, yi)
    srch = np.float32(0.0) * q if s2 == 0 else np.float32(s1 / s2) * q
    for l, i in zip(range(len(alpha) - 1, -1, -1), range(lo, it)):
        si, yi = pair(i)
        s1, s2 = _dot(yi, srch), _dot(yi, si)
        beta = np.float32(0.0) if s2 == 0 else np.float32(s1 / s2)
        srch = srch + si * (alpha[l] - beta)
    return -srch


# ---------------------------------------------------------------------------
# Gradient processing
# ---------------------------------------------------------------------------
def _blackman_ramp(n):
    """Rising Blackman half-window of n samples (0 .. 1); parity unpinned:
    FLIT's taper shape is not in the reference tree."""
    if n <= 0:
        return np.ones(0, np.float32)
    k = np.arange(n, dtype=np.float64)
    x = k / (2.0 * n)
    return (0.42 - 0.5 * np.cos(2 * np.pi * x) + 0.08 * np.cos(4 * np.pi * x)).astype(np.float32)


def taper(w, nz=(0, 0), nx=(0, 0)):
    """Blackman tapers of nz[0], nz[1] cells at the top and bottom and nx[0],
    nx[1] at the left and right of a (nz, nx) array."""
    w = _f32(w).copy()
    for axis, (a, b) in ((0, nz), (1, nx)):
        n = w.shape[axis]
        ramp = np.ones(n, np.float32)
        if a:
            ramp[:a] *= _blackman_ramp(a)
        if b:
            ramp[n - b:] *= _blackman_ramp(b)[::-1]
        w *= ramp[:, None] if axis == 0 else ramp[None, :]
    return w


def gauss_smooth(w, sigma_z, sigma_x):
    """Gaussian smoothing with standard deviations in cells (OWL's
    gauss_filt with [smooth_z/dz, smooth_x/dx]); edges replicated.  Parity
    unpinned: FLIT's gauss_filt is not in the reference tree."""
    from scipy.ndimage import gaussian_filter
    return gaussian_filter(_f32(w), sigma=(float(sigma_z), float(sigma_x)), mode="nearest").astype(np.float32)


def process_gradient(w, steps, dx, dz, it=1, update_iter=None, scale=1.0, smooth=None, taper_x=(0.0, 0.0),
                     taper_z=(0.0, 0.0), mask=None, rms_balance_x=None, rms_balance_z=None):
    """process_model_single_parameter (module_gradient.f90:738-900) for one
    (nz, nx) gradient: the operations in `steps` applied in order, each one
    of 'scale', 'max_balance', 'rms_balance', 'rms_balance_x',
    'rms_balance_z', 'smooth', 'taper', 'mask'.  update_iter = (first,
    last): outside that iteration window the gradient is zero."""
    w = _f32(w).copy()
    if update_iter is not None and not (update_iter[0] <= it <= update_iter[1]):
        return np.zeros_like(w)
    for op in steps:
        if op in ("", None):
            continue
        if op == "scale":
            w = w * np.float32(scale)
        elif op == "max_balance":
            m = w.max()
            if m != 0:
                w = w / m
        elif op == "rms_balance":
            # FLIT's mean(w, 2): the root mean square (parity unpinned)
            m = np.float32(np.sqrt(np.mean(w.astype(np.float64) ** 2)))
            if m != 0:
                w = w / m
        elif op in ("rms_balance_x", "rms_balance_z"):
            axis = 1 if op == "rms_balance_x" else 0
            length = rms_balance_x if axis == 1 else rms_balance_z
            d = dx if axis == 1 else dz
            length = (1.0 * d) if length is None else length
            n = int(np.floor(length / d + 0.5))
            if n % 2 == 0:
                n -= 1
            h = (n - 1) // 2
            pad = [(0, 0), (0, 0)]
            pad[axis] = (n, n)
            wt = np.pad(w, pad)          # zero padding (pad_array default)
            out = np.empty_like(w)
            for j in range(w.shape[axis]):
                sl = [slice(None), slice(None)]
                sl[axis] = slice(j + n - h, j + n + h + 1)
                nrm = np.float32(np.sqrt(np.sum(wt[tuple(sl)].astype(np.float64) ** 2)))
                tgt = [slice(None), slice(None)]
                tgt[axis] = j
                with np.errstate(divide="ignore", invalid="ignore"):
                    out[tuple(tgt)] = w[tuple(tgt)] / nrm
            w = np.where(np.isfinite(out), out, 0).astype(np.float32)   # return_normal
        elif op == "smooth":
            sx, sz = smooth if smooth is not None else (3.0 * dx, 3.0 * dz)
            w = gauss_smooth(w, sz / dz, sx / dx)
        elif op == "taper":
            tz = tuple(int(math.floor(t / dz + 0.5)) for t in taper_z)
            tx = tuple(int(math.floor(t / dx + 0.5)) for t in taper_x)
            w = taper(w, tz, tx)
        elif op == "mask":
            if mask is None:
                raise ValueError("mask step without a mask array")
            w = w * _f32(mask)
        else:
            raise ValueError(f"unknown gradient processing step {op!r}")
    return w


# ---------------------------------------------------------------------------
# Step size
# ---------------------------------------------------------------------------
class Parameter:
    """One inverted model parameter: its array, search direction, bounds and
    step_max (model_m, model_srch, model_min/max, model_step_max)."""

    def __init__(self, name, model, vmin=-FLOAT_HUGE, vmax=FLOAT_HUGE, step_max=None):
        self.name = name
        self.m = _f32(model).copy()
        self.srch = np.zeros_like(self.m)
        self.vmin, self.vmax = np.float32(vmin), np.float32(vmax)
        family = name if name in STEP_MAX_DEFAULT else ("c" if name[:1] in "cC" else name)
        self.step_max = np.float32(STEP_MAX_DEFAULT.get(family, 0.0) if step_max is None else step_max)
        self.step = np.float32(0.0)

    def initial_step(self, scale_factor=1.0):
        """initial_step_size: step = step_max / max|srch| (0 for a zero
        direction)."""
        step_max = np.float32(self.step_max * scale_factor)
        ms = np.abs(self.srch).max() if self.srch.size else 0.0
        self.step = np.float32(0.0) if ms == 0 else np.float32(step_max / ms)
        return self.step

    def updated(self, s):
        """m + step * srch * s, clipped to [vmin, vmax] (update_model)."""
        return np.clip(self.m + self.step * self.srch * np.float32(s), self.vmin, self.vmax).astype(np.float32)

    def in_range(self, s, scale_factor=1.0):
        step_max = np.float32(self.step_max * scale_factor)
        if step_max == 0:
            return True
        return bool(np.abs(self.step * self.srch * np.float32(s)).max() < step_max)


def suitable_scaling(params, s, tries=20):
    """Halve the scaling factor until every parameter's perturbation is below
    its step_max (the 'check if the initial step scalar is suitable' loop):
    returns the factor after the last halving, or raises."""
    cnt, ok = 1, False
    while not ok and cnt < tries:
        ok = all(p.in_range(s) for p in params)
        s = s / 2.0
        cnt += 1
    if not ok:
        raise RuntimeError("cannot find a suitable initial step size")
    return s


def step_quadratic(params, misfit0, misfit_at):
    """compute_step_size_quadratic: misfits at 0.1 s0 and s0 and the vertex
    of the parabola through (0, misfit0) and those two points.  Returns
    (step, misfit at step)."""
    for p in params:
        p.initial_step()
    s = suitable_scaling(params, 1.0)
    step1, step2 = 0.1 * s, s
    f1, f2 = misfit_at(step1), misfit_at(step2)
    num = 0.5 * ((misfit0 - f2) * step1 ** 2 + (-misfit0 + f1) * step2 ** 2)
    den = -(f2 * step1) + misfit0 * (step1 - step2) + f1 * step2
    step = num / den
    return step, misfit_at(step)


def step_linesearch(params, misfit0, misfit_at, jumpout_factor=1.0, nsearch_max=NSEARCH_MAX):
    """compute_step_size_linesearch: quadratic fit with interval bounding and
    a bisection switch-over, starting from the suitable scaling factor; a
    trial that lowers the misfit at once ends the search.  Returns (step,
    misfit); step 0 when no trial beats jumpout_factor * misfit0."""
    for p in params:
        p.initial_step()
    a0 = suitable_scaling(params, 1.0)
    al, fl = 0.0, misfit0
    ar = a0
    fr = misfit_at(ar)
    step1, misfit1 = ar, fr
    step2, misfit2 = None, None
    if fr < fl:
        ac, fc = ar, fr
    else:
        ac = 0.5 * (al + ar)
        fc = misfit_at(ac)
        step2, misfit2 = ac, fc
        if not fc < fl:
            cnt = 3
            while cnt < nsearch_max:
                if fc < fl and fc < fr:
                    acr, bcr = ac - ar, ac ** 2 - ar ** 2
                    arl, brl = ar - al, ar ** 2 - al ** 2
                    alc, blc = al - ac, al ** 2 - ac ** 2
                    num = bcr * fl + brl * fc + blc * fr
                    den = acr * fl + arl * fc + alc * fr
                    if den == 0.0:
                        break
                    an = 0.5 * num / den
                    fn = misfit_at(an)
                    cnt += 1
                    if an > ac:
                        if fn >= fc:
                            ar, fr = an, fn
                        else:
                            al, fl, ac, fc = ac, fc, an, fn
                    else:
                        if fn >= fc:
                            al, fl = an, fn
                        else:
                            ar, fr, ac, fc = ac, fc, an, fn
                else:
                    a2 = ac + 1.0e-1 * (ar - al)
                    f2 = misfit_at(a2)
                    cnt += 1
                    if fc < f2:
                        ar, fr = a2, f2
                    else:
                        al, fl = ac, fc
                    ac = 0.5 * (al + ar)
                    fc = misfit_at(ac)
                    cnt += 1
                if abs(ar - al) <= 0.05:
                    if fc >= misfit1:
                        ac, fc = step1, misfit1
                    if step2 is not None and fc >= misfit2:
                        ac, fc = step2, misfit2
                    break
    if fc >= jumpout_factor * misfit0:
        return 0.0, misfit0
    return ac, fc


def step_linear(params, misfit0, traces_at, obs, misfit_at, jumpout_factor=1.0, nsearch_max=NSEARCH_MAX,
                enforce_update=False):
    """compute_step_size_linear with the L2 waveform misfit: one trial at the
    suitable scaling factor (starting from 0.2), the step coefficient
    sum1 / sum2 from the per-trace misfits err23 = |d(s) - d(0)|^2 and
    err13 = |obs - d(0)|^2 (compute_step_coef), then halvings from twice that
    coefficient until the misfit is below jumpout_factor * misfit0 (or one
    trial when enforce_update).  traces_at(s) -> list of (nr, nt) traces per
    shot; obs likewise.  Returns (step, misfit)."""
    for p in params:
        p.initial_step()
    s = suitable_scaling(params, 0.2)
    d0 = traces_at(0.0)
    d1 = traces_at(s)
    sum1 = np.float32(0.0)
    sum2 = np.float32(0.0)
    for a0, a1, ob in zip(d0, d1, obs):
        e23 = np.sum((_f32(a1) - _f32(a0)) ** 2, axis=-1, dtype=np.float32)
        e13 = np.sum((_f32(ob) - _f32(a0)) ** 2, axis=-1, dtype=np.float32)
        keep = (np.linalg.norm(a0, axis=-1) > 0) & (np.linalg.norm(a1, axis=-1) > 0) & \
               (np.linalg.norm(ob, axis=-1) > 0)
        sum1 += np.float32(np.sum(e23[keep] * e13[keep], dtype=np.float32))
        sum2 += np.float32(np.sum(e23[keep] ** 2, dtype=np.float32))
    if not (np.isfinite(sum1) and np.isfinite(sum2)):
        raise RuntimeError("step coefficient is not finite")
    step = float(sum1 / (sum2 + FLOAT_TINY)) * 2.0
    if not step <= FLOAT_HUGE:
        raise RuntimeError("step size NaN")
    trial = 2.0 * misfit0
    cnt = 1
    if enforce_update:
        while cnt < 2:
            step /= 2.0
            if not all(p.in_range(step) for p in params):
                continue
            trial = misfit_at(step)
            cnt += 1
        return step, trial
    while trial > jumpout_factor * misfit0 and cnt < nsearch_max:
        step /= 2.0
        if not all(p.in_range(step) for p in params):
            continue
        trial = misfit_at(step)
        cnt += 1
    if trial > jumpout_factor * misfit0:
        return 0.0, misfit0
    return step, trial


# ---------------------------------------------------------------------------
# Driver for 2D acoustic FWI on the GPU
# ---------------------------------------------------------------------------
class FWI2D:
    """Iterative 2D acoustic FWI of vp (and optionally rho) with the GPU
    gradient of an Acoustic2DBatch: gradient, processing, search direction,
    step size, update (OWL's main_fwi iteration, src/inverse/main_fwi.f90).

    batch: an Acoustic2DBatch holding the shots; obs: observed traces per
    shot.  params: dict name -> dict(min=, max=, step_max=) for 'vp' and
    optionally 'rho'.  mode: OWL_GRAD_OWL or OWL_GRAD_EXACT (with the given
    checkpointing)."""

    def __init__(self, batch, obs, vp, rho, params=None, search="l-bfgs", step_method="line-search",
                 processing=(), process_kw=None, mode=None, ckpt=None, ckpt_param=0, store=None,
                 jumpout_factor=1.0):
        from . import _lib
        self.b = batch
        self.obs = [_f32(o) for o in obs]
        self.b.set_observed(self.obs)
        params = params or {"vp": {}}
        self.p = {}
        for name, kw in params.items():
            arr = vp if name == "vp" else rho
            self.p[name] = Parameter(name, arr, kw.get("min", -FLOAT_HUGE), kw.get("max", FLOAT_HUGE),
                                     kw.get("step_max"))
        self.fixed = {"vp": _f32(vp), "rho": _f32(rho)}
        self.search = search.lower()
        self.step_method = step_method
        self.processing = tuple(processing)
        self.process_kw = dict(process_kw or {})
        self.mode = _lib.OWL_GRAD_OWL if mode is None else mode
        self.ckpt = _lib.OWL_CKPT_FULL if ckpt is None else ckpt
        self.ckpt_param = ckpt_param
        self.store = _lib.OWL_STORE_F32 if store is None else store
        self.jumpout_factor = jumpout_factor
        self.it = 0
        self.history = {n: dict(models=[p.m.copy()], grads=[], srch=None) for n, p in self.p.items()}
        self.misfits = []

    def _model(self, overrides=None):
        m = {k: v for k, v in self.fixed.items()}
        for n, p in self.p.items():
            m[n] = p.m
        if overrides:
            m.update(overrides)
        return m["vp"], m["rho"]

    def _misfit_at(self, s):
        vp, rho = self._model({n: p.updated(s) for n, p in self.p.items()})
        self.b.set_model(vp, rho)
        mis, _ = self.b.misfit()
        return float(np.sum(mis))

    def _traces_at(self, s):
        vp, rho = self._model({n: p.updated(s) for n, p in self.p.items()})
        self.b.set_model(vp, rho)
        return self.b.forward()

    def iterate(self):
        """One FWI iteration; returns the misfit after the update."""
        self.it += 1
        vp, rho = self._model()
        self.b.set_model(vp, rho)
        self.b.set_observed(self.obs)
        mis, gv, gr, _ = self.b.gradient(want_traces=False, mode=self.mode, ckpt=self.ckpt,
                                         ckpt_param=self.ckpt_param, store=self.store)
        misfit0 = float(np.sum(mis))
        grads = {"vp": _f32(gv), "rho": _f32(gr)}
        dx, dz = self.b.dx, self.b.dz
        for n, p in self.p.items():
            g = process_gradient(grads[n], self.processing, dx, dz, it=self.it, **self.process_kw)
            h = self.history[n]
            if self.search in ("sd", "steepest-descent"):
                p.srch = search_direction_sd(g)
            elif self.search in ("cg", "conjugate-gradient"):
                prev = h["grads"][-1] if h["grads"] else None
                p.srch = search_direction_cg(g, prev, h["srch"])
            elif self.search == "l-bfgs":
                p.srch = search_direction_lbfgs(g, h["models"], h["grads"])
            else:
                raise ValueError(f"unknown search method {self.search!r}")
            h["grads"].append(g)
            h["srch"] = p.srch.copy()
        if self.step_method == "line-search":
            step, misfit = step_linesearch(list(self.p.values()), misfit0, self._misfit_at, self.jumpout_factor)
        elif self.step_method == "quadratic":
            step, misfit = step_quadratic(list(self.p.values()), misfit0, self._misfit_at)
        elif self.step_method == "linear":
            step, misfit = step_linear(list(self.p.values()), misfit0, self._traces_at, self.obs,
                                       self._misfit_at, self.jumpout_factor)
        else:
            raise ValueError(f"unknown step-size method {self.step_method!r}")
        for n, p in self.p.items():
            p.m = p.updated(step)
            self.history[n]["models"].append(p.m.copy())
            # L-BFGS keeps the last nstep pairs
            hm, hg = self.history[n]["models"], self.history[n]["grads"]
            if len(hg) > NSTEP_LBFGS:
                del hm[0]
                del hg[0]
        self.misfits.append((misfit0, misfit, step))
        return misfit
