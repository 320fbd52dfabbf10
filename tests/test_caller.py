"""OWL's driver set-up on the host (owl_amd/fortran/owl_caller.f90 through
owl_modeling2 / owl_gradient2 --dry-run, which stop before any GPU call) and
the trace resampling of su%resamp (owl_resample_traces), CPU only:

* the parameter file: "key = value" lines, comments, the last of repeated
  keys, lists and logicals, defaults, a required key that is missing;
* load_geometry (module_source_receiver.f90:183-551): shot files listed in
  geometry.txt, force angles in degrees, the source and receiver validity
  rules (model extent, sx/rx limits, offsets, rec_index, zero weights,
  positions above the topography of file_topo), src_index, sid_select,
  src_exclude, sid_exclude and duplicate ids;
* the per-shot windows equal owl_amd.set_adaptive_range's on the same
  geometry, and a model read on a coarser grid is resampled linearly;
* resampling equals the natural cubic spline of scipy, keeps all-zero
  traces zero and is the identity at the same sampling; the oracle's
  free-surface Lamb traces at dt = 0.2 ms resampled to data_dt = 1 ms match
  the Lamb solution at 1 ms as OWL's own Lamb test samples them
  (test/test_elastic_lamb.py:47-48).
"""
import os
import subprocess

import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from owl_amd.formats import resample

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MODELING = os.path.join(ROOT, "owl_amd", "fortran", "owl_modeling2")
GRADIENT = os.path.join(ROOT, "owl_amd", "fortran", "owl_gradient2")


def _need(exe):
    if not os.path.exists(exe):
        pytest.fail(f"{os.path.basename(exe)} is not built (__graft_entry__.build() builds it)")


def write_project(d, params, shots, models, topo=None):
    """A project directory as OWL reads it: param.rb, geometry/geometry.txt
    with one file per shot, float32 (nz, nx) models.  shots: list of dicts
    id, sources [(x, z, mech_line, wavelet_line)], receivers [(x, z, w)]."""
    os.makedirs(os.path.join(d, "geometry"), exist_ok=True)
    os.makedirs(os.path.join(d, "model"), exist_ok=True)
    with open(os.path.join(d, "geometry", "geometry.txt"), "w") as f:
        for s in shots:
            f.write(f"shot_{s['id']}_geometry.txt\n")
    for s in shots:
        with open(os.path.join(d, "geometry", f"shot_{s['id']}_geometry.txt"), "w") as f:
            f.write(f"{s['id']}\n\n{len(s['sources'])}\n")
            for x, z, mech, wav in s["sources"]:
                f.write(f"{float(x)!r} 0.0 {float(z)!r}\n{mech}\n{wav}\n0.0 0.0\n")
            f.write(f"\n{len(s['receivers'])}\n")
            for x, z, w in s["receivers"]:
                f.write(f"{float(x)!r} 0.0 {float(z)!r} {float(w)!r}\n")
    for name, a in models.items():
        np.asfortranarray(a, np.float32).T.tofile(os.path.join(d, "model", f"{name}.bin"))
    lines = [f"{k} = {v}" for k, v in params.items()]
    lines.append("file_geometry = " + os.path.join(d, "geometry", "geometry.txt"))
    lines += [f"file_{k} = " + os.path.join(d, "model", f"{k}.bin") for k in models]
    if topo is not None:
        tp = os.path.join(d, "topo.txt")
        with open(tp, "w") as f:
            for x, z in zip(*topo):
                f.write(f"{float(x)!r} {float(z)!r}\n")
        lines.append(f"file_topo = {tp}")
    path = os.path.join(d, "param.rb")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path


def dry_run(exe, param):
    _need(exe)
    r = subprocess.run([exe, param, "--dry-run"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    out = dict(shots=[], model_stats={})
    for ln in r.stdout.splitlines():
        t = ln.split()
        if not t:
            continue
        if t[0] == "shot":
            out["shots"].append(dict(id=int(t[1]), ns=int(t[2]), nr=int(t[3]), live=int(t[4]),
                                     window=tuple(int(v) for v in t[5:9]), src=[], rec=[]))
        elif t[0] == "src":
            out["shots"][-1]["src"].append(dict(mech=t[1], wavelet=t[2], x=float(t[3]), z=float(t[4]),
                                                f0=float(t[5]), amp=float(t[6]), polar=float(t[7]), t0=float(t[8])))
        elif t[0] == "rec":
            out["shots"][-1]["rec"].append(tuple(float(v) for v in t[1:5]))
        elif t[0] == "model":
            out["model_stats"][t[1]] = tuple(float(v) for v in t[2:5])
        elif t[0] == "shots":
            out["nshot"] = int(t[1])
        else:
            out[t[0]] = t[1:]
    return out


def _line(n, x0, dx, z, w=1.0):
    return [(x0 + k * dx, z, w) for k in range(n)]


NX, NZ, DX = 80, 40, 10.0


def _vp():
    return (2000.0 + 10.0 * np.arange(NZ)[:, None] + 0.0 * np.arange(NX)[None, :]).astype(np.float32)


def test_parameter_file_rules(tmp_path):
    shots = [dict(id=1, sources=[(200.0, 20.0, "explosion", "ricker 12.5 2.0e3 0.01")],
                  receivers=_line(10, 100.0, 20.0, 10.0))]
    p = write_project(str(tmp_path), {"nx": NX, "nz": NZ, "dx": DX, "dz": DX, "dt": "0.5e-3", "tmax": 0.4,
                                      "ns": 1, "model_name": "vp,   rho", "which_medium": "acoustic-iso",
                                      "verbose": "yes", "npml": 12},
                      shots, {"vp": _vp(), "rho": np.full((NZ, NX), 1000.0, np.float32)})
    with open(p, "a") as f:
        # a comment, a repeated key (the last line wins), blanks around '='
        f.write("# tmax = 9\n   tmax    =   0.6   # the last one\ndata_dt = 1.0e-3\n")
    o = dry_run(MODELING, p)
    assert o["which_medium"] == ["acoustic-iso"]
    assert o["grid"] == ["80", "40", "12"]
    dt, tmax, ddt, dtmax = (float(v) for v in o["time"])
    assert dt == pytest.approx(0.5e-3) and tmax == pytest.approx(0.6)
    assert ddt == pytest.approx(1e-3) and dtmax == pytest.approx(0.6)  # data_tmax defaults to tmax
    assert o["models"] == ["vp", "rho"]
    assert o["data"] == ["p"]
    (s,) = o["shots"]
    assert s["src"][0]["f0"] == pytest.approx(12.5) and s["src"][0]["t0"] == pytest.approx(0.01)
    # a missing required key stops with OWL's message
    with open(p) as f:
        text = f.read()
    with open(p, "w") as f:
        f.write("\n".join(ln for ln in text.splitlines() if not ln.startswith("nz ")) + "\n")
    r = subprocess.run([MODELING, p, "--dry-run"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "required parameter nz" in r.stdout
    r = subprocess.run([MODELING, str(tmp_path / "absent.rb")], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "not found" in r.stdout


def test_load_geometry_rules_and_windows(tmp_path):
    from owl_amd import set_adaptive_range
    from owl_amd.geometry import Geometry, Receiver, Source
    shots = [
        # receivers 4 (weight 0) and beyond offset 260 m dropped
        dict(id=11, sources=[(250.0, 20.0, "force 30.0 0.0", "ricker 10.0 1.0 0.0")],
             receivers=[(x, 10.0, 0.0 if k == 4 else 1.0) for k, x in enumerate(np.arange(0.0, 790.0, 30.0))]),
        # a source outside the model: the shot has no valid source and is dropped
        dict(id=12, sources=[(900.0, 20.0, "explosion", "ricker 10.0 1.0 0.0")],
             receivers=_line(5, 100.0, 10.0, 10.0)),
        dict(id=13, sources=[(500.0, 30.0, "explosion", "ricker 10.0 1.0 0.0")],
             receivers=_line(30, 300.0, 15.0, 12.0)),
        dict(id=14, sources=[(650.0, 30.0, "explosion", "ricker 10.0 1.0 0.0")],
             receivers=_line(10, 600.0, 10.0, 12.0)),
        dict(id=15, sources=[(700.0, 30.0, "explosion", "ricker 10.0 1.0 0.0")],
             receivers=_line(10, 600.0, 10.0, 12.0)),
    ]
    params = {"nx": NX, "nz": NZ, "dx": DX, "dz": DX, "dt": 1e-3, "tmax": 0.3, "ns": 5,
              "model_name": "vp", "offset_max": 260.0, "rec_index": "1, 1, 1000", "yn_adpx": "y",
              "adp_extrax": 40.0, "src_exclude": 5}
    p = write_project(str(tmp_path), params, shots, {"vp": _vp()})
    o = dry_run(MODELING, p)
    assert [s["id"] for s in o["shots"]] == [11, 13, 14]
    s11 = o["shots"][0]
    assert s11["src"][0]["mech"] == "force"
    assert s11["src"][0]["polar"] == pytest.approx(np.radians(30.0), rel=1e-6)
    for (x, z, w, aoff), (rx, _, w0) in zip(s11["rec"], shots[0]["receivers"]):
        assert aoff == pytest.approx(np.hypot(rx - 250.0, 10.0 - 20.0), rel=1e-6)
        expect = w0 != 0 and aoff <= 260.0
        assert (w != 0) == expect, (rx, w, aoff)
    # windows: the library's rule on the receivers load_geometry kept
    for s, shot in zip(o["shots"], [shots[0], shots[2], shots[3]]):
        g = Geometry(id=s["id"], srcr=[Source(x=shot["sources"][0][0], z=shot["sources"][0][1])],
                     recr=[Receiver(x=x, z=z, weight=w) for x, z, w, _ in s["rec"]])
        w = set_adaptive_range(g, NX, NZ, DX, DX, yn_adpx=True, adp_extrax=40.0, offset_max=260.0)
        assert s["window"] == (w.nxbeg, w.nxend, w.nzbeg, w.nzend)
        assert s["live"] == sum(r.weight != 0 for r in g.recr)
    # rec_index 2, 3, 9: every third receiver from the second, up to the ninth
    params.update({"rec_index": "2, 3, 9", "src_exclude": 0, "sid_select": "13, 15"})
    p = write_project(str(tmp_path / "b"), params, shots, {"vp": _vp()})
    o = dry_run(MODELING, p)
    assert [s["id"] for s in o["shots"]] == [13, 15]
    live = [k + 1 for k, r in enumerate(o["shots"][0]["rec"]) if r[2] != 0]
    assert live and all(j in (2, 5, 8) for j in live)
    # duplicate ids stop
    shots[3]["id"] = 13
    p = write_project(str(tmp_path / "c"), dict(params, sid_select=0, rec_index="1, 1, 100"), shots,
                      {"vp": _vp()})
    r = subprocess.run([MODELING, p, "--dry-run"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "Duplicate source ID" in r.stdout


def test_topography_drops_sources_and_receivers_above_the_surface(tmp_path):
    xt = np.array([-300.0, 0.0, 400.0, 1100.0])
    zt = np.array([0.0, 0.0, 60.0, 60.0])  # elevation up: the surface rises to z = -60
    shots = [dict(id=1, sources=[(100.0, 20.0, "explosion", "ricker 10.0 1.0 0.0")],
                  receivers=[(x, -30.0, 1.0) for x in np.arange(0.0, 700.0, 50.0)]),
             dict(id=2, sources=[(600.0, -70.0, "explosion", "ricker 10.0 1.0 0.0")],
                  receivers=_line(5, 500.0, 10.0, 10.0))]
    params = {"nx": NX, "nz": NZ, "dx": DX, "dz": DX, "dt": 1e-3, "tmax": 0.3, "ns": 2, "model_name": "vp, vs",
              "which_medium": "elastic-tti", "yn_free_surface": "y", "topo_interp": "linear", "npml": 10}
    p = write_project(str(tmp_path), params, shots, {"vp": _vp(), "vs": _vp() / 1.8}, topo=(xt, zt))
    o = dry_run(MODELING, p)
    # shot 2's source sits above the surface (z = -70 < -60): dropped
    assert [s["id"] for s in o["shots"]] == [1]
    assert o["data"] == ["x", "z"]
    assert float(o["free_surface"][2]) == pytest.approx(15.0)  # dz_max defaults to 1.5 dz
    surf = -np.interp(np.arange(0.0, 700.0, 50.0), xt, zt)  # depth of the surface
    for (x, z, w, _), zs in zip(o["shots"][0]["rec"], surf):
        assert (w != 0) == (z >= zs - 1e-4), (x, z, zs)
    # measured from the surface: every receiver counts
    p = write_project(str(tmp_path / "b"), dict(params, measure_receiver_depth_from_surface="y"), shots,
                      {"vp": _vp(), "vs": _vp() / 1.8}, topo=(xt, zt))
    o = dry_run(MODELING, p)
    assert all(r[2] != 0 for r in o["shots"][0]["rec"])


def test_model_resampled_to_the_target_grid(tmp_path):
    nx0, nz0 = 41, 21
    x0 = np.arange(nx0) * 20.0
    z0 = np.arange(nz0) * 20.0
    m0 = (1500.0 + 3.0 * z0[:, None] + 0.5 * x0[None, :] + 40.0 * np.sin(x0[None, :] / 90.0)).astype(np.float32)
    shots = [dict(id=1, sources=[(200.0, 20.0, "explosion", "ricker 10.0 1.0 0.0")],
                  receivers=_line(10, 100.0, 20.0, 10.0))]
    # nnz and ddz set the depth extent (zmax defaults to (nnz - 1) ddz); xmax
    # rounds up to the grid
    params = {"nx": nx0, "nz": nz0, "dx": 20.0, "dz": 20.0, "ddx": 10.0, "ddz": 10.0, "nnz": 41, "xmax": 795.0,
              "dt": 1e-3, "tmax": 0.3, "ns": 1, "model_name": "vp"}
    p = write_project(str(tmp_path), params, shots, {"vp": m0})
    o = dry_run(MODELING, p)
    nx, nz = int(o["grid"][0]), int(o["grid"][1])
    assert (nx, nz) == (81, 41)  # ceiling(795/10) + 1, ceiling(400/10) + 1
    xs = np.arange(nx) * 10.0
    zs = np.arange(nz) * 10.0
    t = np.stack([np.interp(zs, z0, m0[:, j].astype(np.float64)) for j in range(nx0)], axis=1)
    ref = np.stack([np.interp(xs, x0, t[i]) for i in range(nz)], axis=0)
    lo, hi, total = o["model_stats"]["vp"]
    assert lo == pytest.approx(ref.min(), rel=1e-6) and hi == pytest.approx(ref.max(), rel=1e-6)
    assert total == pytest.approx(ref.sum(), rel=1e-6)


def test_gradient_driver_reads_fwi_keys(tmp_path):
    shots = [dict(id=3, sources=[(200.0, 20.0, "explosion", "ricker 10.0 1.0 0.0")],
                  receivers=_line(10, 100.0, 20.0, 10.0))]
    params = {"nx": NX, "nz": NZ, "dx": DX, "dz": DX, "dt": 1e-3, "tmax": 0.3, "ns": 1,
              "model_update": "vp", "model_aux": "rho", "misfit_type": "envelope"}
    p = write_project(str(tmp_path), params, shots, {"vp": _vp(), "rho": np.full((NZ, NX), 1900.0, np.float32)})
    o = dry_run(GRADIENT, p)
    assert o["models"] == ["vp"]  # model_update
    # the model_aux model is read too, for the solvers' get_model
    assert o["model_stats"]["rho"][:2] == (1900.0, 1900.0)
    assert o["model_stats"]["vp"][1] == pytest.approx(float(_vp().max()))


def test_resample_matches_natural_cubic_spline():
    rng = np.random.default_rng(5)
    dt, nt, ddt, nnt = 2.0e-4, 2501, 1.0e-3, 501
    tr = rng.standard_normal((4, nt)).astype(np.float32)
    tr[2] = 0.0
    out = resample(tr, dt, nnt, ddt)
    t = (np.arange(nt, dtype=np.float32) * np.float32(dt)).astype(np.float64)
    tq = (np.arange(nnt, dtype=np.float32) * np.float32(ddt)).astype(np.float64)
    for k in (0, 1, 3):
        ref = CubicSpline(t, tr[k].astype(np.float64), bc_type="natural")(tq)
        assert np.abs(out[k] - ref).max() < 2e-6 * np.abs(ref).max()
    assert not out[2].any()
    assert np.array_equal(resample(tr, dt, nt, dt), tr)
    # a band-limited signal well inside both Nyquists comes back to the sampled values
    s = np.sin(2 * np.pi * 12.0 * t)[None].astype(np.float32)
    assert np.abs(resample(s, dt, nnt, ddt)[0] - np.sin(2 * np.pi * 12.0 * tq)).max() < 1e-5
    with pytest.raises(Exception):
        resample(tr, dt, nnt, 0.0)


def test_lamb_traces_resampled_to_data_dt(oracle64):
    """OWL's Lamb test records at data_dt = 1 ms a run at dt = 0.2 ms; the
    oracle's flat free-surface traces, resampled so, match the Lamb solution
    at 1 ms with OWL's tolerance (shape and energy below 0.20)."""
    from lamb_halfspace import lamb_surface_velocity
    from test_oracle_pinning_fsg_surface import DT, F0, OFFSETS, RHO, TMAX, VP, VS, XC, _iso, _shape_error, \
        _surface, NX as LNX, NZ as LNZ, DX as LDX
    o = oracle64
    data_dt = 1.0e-3
    nt = int(round(TMAX / DT)) + 1
    cij, rho = _iso()
    g = dict(nx=LNX, nz=LNZ, dx=LDX, dz=LDX, dt=DT, nt=nt, pml=15, R=4, aniso=0)
    stf, _ = o.stf(F0, DT, 1.0, "ricker")
    s = np.zeros(nt)
    s[1:1 + len(stf)] = stf[:nt - 1]
    dnt = int(round(TMAX / data_dt)) + 1
    src = [dict(x=XC, z=0.0, f0=F0, amp=1.0e6, polar=0.0, mechanism="force")]
    rec = [dict(x=XC + r, z=0.0, weight=1.0) for r in OFFSETS]
    tx, tz = o.fsg2d_forward_fs(g, _surface(0.0), cij, rho, src, rec)
    vx = resample(tx.astype(np.float32), DT, dnt, data_dt).T.astype(np.float64)
    vz = resample(tz.astype(np.float32), DT, dnt, data_dt).T.astype(np.float64)
    # the Lamb solution from the source time function sampled at data_dt
    ts = np.arange(dnt) * data_dt
    sd = np.interp(ts, np.arange(nt) * DT, s)
    rt, rn = lamb_surface_velocity(OFFSETS, sd, data_dt, VP, VS, RHO)
    a = np.concatenate([vx.ravel(), vz.ravel()])
    b = np.concatenate([rt.ravel(), rn.ravel()])
    K = a @ b / (a @ a)
    shape = 0.5 * (_shape_error(vx, rt) + _shape_error(vz, rn))
    energy = abs(np.sqrt(((K * vx) ** 2 + (K * vz) ** 2).sum()) / np.sqrt((rt ** 2 + rn ** 2).sum()) - 1.0)
    assert shape < 0.20 and energy < 0.20, (shape, energy)
