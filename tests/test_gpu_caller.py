"""OWL's 2D drivers over the drop-in modules (owl_amd/fortran/owl_modeling2
and owl_gradient2: parameter file, geometry files, models on disk, per-shot
windows) give the library batch's results bit for bit:

* acoustic modelling of three shots on adaptive windows with data_dt = 2 dt:
  the SU files equal the batch's traces resampled by owl_resample_traces;
* elastic-tti modelling with 'thomsen' parameters, a free surface from
  file_topo and depths measured from the surface: the two SU components
  equal the FSG batch's with the same surface;
* one acoustic FWI gradient: the driver's misfit is the sum of the batch's
  per-trace waveform misfits and grad_vp.bin, grad_rho.bin the shots'
  window gradients (adjoint sources from owl_adjoint_source) summed on the
  model grid.
"""
import os
import subprocess

import numpy as np
import pytest

from owl_amd import Acoustic2DBatch, set_adaptive_range
from owl_amd.formats import read_su, resample, write_su
from owl_amd.geometry import Geometry, Receiver, Source
from owl_amd.misfit import adjoint_source
from owl_amd.synthetic import gardner, random_field
from test_caller import GRADIENT, MODELING, write_project

pytestmark = pytest.mark.gpu

NX, NZ, DX, DT, TMAX, PML = 120, 50, 10.0, 1.0e-3, 0.3, 15


def _run(exe, param, what):
    if not os.path.exists(exe):
        pytest.fail(f"{os.path.basename(exe)} is not built (__graft_entry__.build() builds it)")
    print(f"case: {what}", flush=True)
    r = subprocess.run([exe, param], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def _acoustic_project(d, vp, rho, extra):
    shots = [dict(id=sid, sources=[(sx, 20.0, "explosion", f"ricker {f0} 1.0 0.0")],
                  receivers=[(x0 + 10.0 * k, 15.0, 1.0) for k in range(nr)])
             for sid, sx, f0, x0, nr in [(4, 255.0, 14.0, 30.0, 70), (8, 640.0, 12.0, 420.0, 60),
                                         (9, 1010.0, 15.0, 700.0, 45)]]
    params = {"nx": NX, "nz": NZ, "dx": DX, "dz": DX, "dt": DT, "tmax": TMAX, "ns": 3, "npml": PML,
              "yn_adpx": "y", "adp_extrax": 120.0, "which_medium": "acoustic-iso",
              "dir_synthetic": os.path.join(d, "syn"), "dir_working": os.path.join(d, "work"),
              "dir_record": os.path.join(d, "obs")}
    params.update(extra)
    p = write_project(d, params, shots, {"vp": vp, "rho": rho})
    geoms, wins = [], []
    for s in shots:
        (sx, sz, _, wav), = s["sources"]
        g = Geometry(id=s["id"], srcr=[Source(x=sx, z=sz, f0=float(wav.split()[1]))],
                     recr=[Receiver(x=x, z=z, weight=w) for x, z, w in s["receivers"]])
        wins.append(set_adaptive_range(g, NX, NZ, DX, DX, yn_adpx=True, adp_extrax=120.0))
        geoms.append(g)
    return p, shots, geoms, wins


def _model(seed):
    zz = np.arange(NZ)[:, None] * DX
    vp = ((1900.0 + 1.5 * zz) * (1.0 + 0.04 * random_field((NZ, NX), 4.0, seed))).astype(np.float32)
    return vp, gardner(vp)


def test_owl_modeling2_acoustic_with_data_dt(gpu, tmp_path):
    vp, rho = _model(11)
    data_dt = 2.0 * DT
    p, shots, geoms, wins = _acoustic_project(str(tmp_path), vp, rho,
                                              {"model_name": "vp, rho", "data_dt": data_dt})
    assert any(w.nx < NX for w in wins)
    b = Acoustic2DBatch(NX, NZ, DX, DX, vp, rho, list(zip(geoms, wins)), DT, TMAX, pml=PML)
    ref = b.forward()
    b.close()
    _run(MODELING, p, "owl_modeling2 acoustic-iso, 3 shots, adaptive windows, data_dt = 2 dt")
    dnt = int(round(TMAX / data_dt)) + 1
    for s, tr in zip(shots, ref):
        syn, sdt = read_su(str(tmp_path / "syn" / f"shot_{s['id']}_seismogram_p.su"))
        assert sdt == pytest.approx(data_dt) and syn.shape == (tr.shape[0], dnt)
        assert np.abs(tr).max() > 0
        assert np.array_equal(syn, resample(tr, DT, dnt, data_dt))


def test_owl_modeling2_elastic_tti_free_surface(gpu, tmp_path):
    from owl_amd import FSG2DBatch
    from owl_amd.fsg2d import tti_cij
    nz, nx, d, dt, tmax, pml = 44, 56, 10.0, 0.4e-3, 0.2, 10
    vp = (2600.0 * (1 + 0.05 * random_field((nz, nx), 3.0, 3))).astype(np.float32)
    vs = (vp / 1.8).astype(np.float32)
    rho = gardner(vp)
    eps = np.full((nz, nx), 0.2, np.float32)
    dlt = np.full((nz, nx), 0.08, np.float32)
    the = (0.4 + 0.3 * random_field((nz, nx), 5.0, 4)).astype(np.float32)
    xt = np.linspace(-300.0, 900.0, 25).astype(np.float32)
    zt = (12.0 * np.sin(xt / 150.0) - 0.06 * xt).astype(np.float32)
    rec = [(20.0 + 9.0 * k, 5.0 + 0.5 * k) for k in range(50)]
    shots = [dict(id=2, sources=[(213.0, 31.0, "force 20.0 0.0", "ricker 20.0 1.0 0.0")],
                  receivers=[(x, z, 1.0) for x, z in rec])]
    params = {"nx": nx, "nz": nz, "dx": d, "dz": d, "dt": dt, "tmax": tmax, "ns": 1, "npml": pml,
              "which_medium": "elastic-tti", "anisotropy_type": "thomsen",
              "model_name": "vp, vs, rho, epsilon, delta, theta", "yn_free_surface": "y",
              "free_surface_dz_refine": 3.0, "dz_max": 14.0, "topo_interp": "cubic",
              "measure_source_depth_from_surface": "y", "measure_receiver_depth_from_surface": "y",
              "dir_synthetic": str(tmp_path / "syn")}
    p = write_project(str(tmp_path), params, shots,
                      {"vp": vp, "vs": vs, "rho": rho, "epsilon": eps, "delta": dlt, "theta": the}, topo=(xt, zt))
    geom = Geometry(id=2, srcr=[Source(x=213.0, z=31.0, f0=20.0, mechanism="force",
                                          polar=float(np.float32(20.0) * np.float32(0.017453292519943295)))],
                    recr=[Receiver(x=x, z=z) for x, z in rec])
    cij = tti_cij("thomsen", vp, vs, rho, eps, dlt, the)
    b = FSG2DBatch(nx, nz, d, d, cij, rho, [(geom, None)], dt, tmax, pml=pml, fdhalf=4, aniso=True,
                   surface=dict(topo=(xt, zt), interp="cubic", refine=3.0, dz_max=14.0,
                                src_from_surface=True, rec_from_surface=True))
    tx, tz = b.forward()[0]
    b.close()
    _run(MODELING, p, "owl_modeling2 elastic-tti thomsen, free surface from file_topo")
    sx, _ = read_su(str(tmp_path / "syn" / "shot_2_seismogram_x.su"))
    sz, _ = read_su(str(tmp_path / "syn" / "shot_2_seismogram_z.su"))
    assert np.abs(tz).max() > 0
    assert np.array_equal(sx, tx) and np.array_equal(sz, tz)


MISFIT_CASES = {"waveform": ({"misfit_type": "waveform"}, {}),
                "adaptive": ({"misfit_type": "adaptive", "tlag_max": 0.04, "penalty_method": "power",
                              "penalty_power": 2.0, "deconv_eps": 0.05},
                             dict(dt=DT, tlag_max=0.04, penalty="power", penalty_power=2.0, deconv_eps=0.05)),
                "local-adaptive": ({"misfit_type": "local-adaptive", "tlag_max": 0.04, "lawi_sigma": 0.03},
                                   dict(dt=DT, tlag_max=0.04, lawi_sigma=0.03)),
                "adaptive-spacetime": ({"misfit_type": "adaptive-spacetime", "tlag_max": 0.04,
                                        "adaptive_half_window": 2},
                                       dict(dt=DT, tlag_max=0.04, adaptive_hw=2))}


@pytest.mark.parametrize("kind", list(MISFIT_CASES))
def test_owl_gradient2_acoustic(gpu, tmp_path, kind):
    vp, rho = _model(5)
    vt = (vp * (1.0 + 0.03 * random_field((NZ, NX), 6.0, 9))).astype(np.float32)
    fpar, kw = MISFIT_CASES[kind]
    p, shots, geoms, wins = _acoustic_project(str(tmp_path), vp, rho, dict({"model_update": "vp, rho"}, **fpar))
    bt = Acoustic2DBatch(NX, NZ, DX, DX, vt, rho, list(zip(geoms, wins)), DT, TMAX, pml=PML)
    obs = bt.forward()
    bt.close()
    os.makedirs(tmp_path / "obs")
    for s, o in zip(shots, obs):
        write_su(str(tmp_path / "obs" / f"shot_{s['id']}_seismogram_p.su"), o, DT)
    b = Acoustic2DBatch(NX, NZ, DX, DX, vp, rho, list(zip(geoms, wins)), DT, TMAX, pml=PML)
    syn = b.forward(save_boundary=True)
    adjs, misfit = [], 0.0
    for g, o, s in zip(geoms, obs, syn):
        a, m = adjoint_source(kind, o, s, np.array([r.weight for r in g.recr], np.float32), **kw)
        adjs.append(a)
        misfit += m.sum()
    grads = b.adjoint(adjs)
    b.close()
    full = [np.zeros((NZ, NX), np.float32), np.zeros((NZ, NX), np.float32)]
    for (gv, gr), w in zip(grads, wins):
        sl = (slice(w.nzbeg - 1, w.nzend), slice(w.nxbeg - 1, w.nxend))
        full[0][sl] += gv
        full[1][sl] += gr
    _run(GRADIENT, p, f"owl_gradient2 acoustic-iso, 3 shots, {kind} misfit")
    # The FFT-based misfits run hipFFT: this python process may bind the
    # rocFFT build that torch bundles, the Fortran program binds the system's,
    # and the two may round some transform lengths differently (1-2 ulp).
    # Without FFTs the two processes agree bit for bit.
    exact = kind == "waveform"
    for s, a in zip(shots, adjs):
        got, _ = read_su(str(tmp_path / "work" / "iteration_1" / "adjoint_source" / f"shot_{s['id']}_seismogram_p.su"))
        if exact:
            assert np.array_equal(got, a), s["id"]
        else:
            assert np.abs(got - a).max() <= 1e-6 * np.abs(a).max(), (s["id"], float(np.abs(got - a).max()))
    for k, name in enumerate(("vp", "rho")):
        got = np.fromfile(str(tmp_path / "work" / "iteration_1" / "model" / f"grad_{name}.bin"),
                          np.float32).reshape(NX, NZ).T
        assert np.abs(full[k]).max() > 0
        if exact:
            assert np.array_equal(got, full[k]), name
        else:
            assert np.linalg.norm(got - full[k]) <= 1e-5 * np.linalg.norm(full[k]), name
    with open(tmp_path / "work" / "data_misfit.txt") as f:
        it, value = f.read().split()
    assert int(it) == 1 and float(value) == pytest.approx(misfit, rel=1e-12 if exact else 1e-6)
