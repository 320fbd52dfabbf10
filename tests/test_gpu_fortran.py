"""The Fortran host side (owl_amd/fortran: iso_c_binding module, the drop-in
acoustic_iso_2d module with OWL's wave_solver_acoustic_iso_2d, the owl_fwd2d
driver) runs the same C ABI as the Python host and writes OWL's files: its SU
traces must equal the Python path's and the float32 oracle's bit for bit, and
its .grd gradients the Python adjoint's.  OWL's own two callers, restated
line for line in tests/fortran/owl_caller_2d.f90 (module_modeling.f90:74-95,
module_gradient.f90:116-135 and 342-346), drive the drop-in module over
several shots with adaptive windows and give the Python batch's traces and
window gradients bit for bit.
"""
import os
import subprocess

import numpy as np
import pytest

from helpers import oracle_grid, oracle_receivers, oracle_sources, simple_geometry, window_model
from owl_amd import Acoustic2DBatch, Window
from owl_amd.formats import read_grd, read_su, write_su
from owl_amd.synthetic import gardner, random_field

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "owl_amd", "fortran", "owl_fwd2d")
CALLER = os.path.join(ROOT, "tests", "fortran", "owl_caller_2d")


def _job(path, nx, nz, dx, dt, tmax, pml, R, recon, f_vp, f_rho, src, rx, rz, syn, adj=None, work=None):
    lines = [f"{nx} {nz} {dx} {dx} {dt} {tmax} {pml} {R} {int(recon)}", f_vp, f_rho,
             f"{src[0]} {src[1]} {src[2]} 1.0", f"{len(rx)}"]
    lines += [f"{x!r} {z!r}" for x, z in zip(rx, rz)]
    lines.append(syn)
    if recon:
        lines += [adj, work]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def test_fortran_driver_forward_adjoint(oracle32, gpu, tmp_path):
    if not os.path.exists(DRIVER):
        pytest.fail("owl_fwd2d is not built (__graft_entry__.build() builds it)")
    nz, nx, dx, dt, tmax, pml, R = 61, 81, 10.0, 0.001, 0.3, 20, 4
    zz = np.arange(nz)[:, None] * dx
    vp = ((2000.0 + zz) * (1.0 + 0.05 * random_field((nz, nx), 4.0, 7))).astype(np.float32)
    rho = gardner(vp)
    f_vp, f_rho = str(tmp_path / "vp.bin"), str(tmp_path / "rho.bin")
    np.asfortranarray(vp).T.tofile(f_vp)          # (nz, nx) column-major = x-major rows
    np.asfortranarray(rho).T.tofile(f_rho)
    rx = [float(np.float32(37.5 + 9.0 * k)) for k in range(60)]
    rz = [100.0] * 60
    src = (403.5, 212.5, 15.0)
    for d in ("syn", "adj", "work"):
        os.makedirs(tmp_path / d)

    geom = simple_geometry(1, src[0], src[1], rx, rz, f0=src[2])
    b = Acoustic2DBatch(nx, nz, dx, dx, vp, rho, [(geom, None)], dt, tmax, pml=pml, fdhalf=R)
    ref = b.forward(save_boundary=True)[0]
    adj = (np.random.default_rng(3).standard_normal(ref.shape).astype(np.float32) * 1e-3
           + 0.5 * ref).astype(np.float32)
    gv_py, gr_py = b.adjoint([adj])[0]
    b.close()
    w = Window(1, nx, 1, nz, 0.0, 0.0)
    orc = oracle32.acoustic2d_forward(oracle_grid(w, dx, dx, dt, ref.shape[1], pml, R),
                                      window_model(vp, w), window_model(rho, w),
                                      oracle_sources(geom), oracle_receivers(geom, w, dx, dx))
    write_su(str(tmp_path / "adj" / "shot_1_seismogram_p.su"), adj, dt)

    job = str(tmp_path / "job.txt")
    _job(job, nx, nz, dx, dt, tmax, pml, R, True, f_vp, f_rho, src, rx, rz,
         str(tmp_path / "syn"), str(tmp_path / "adj"), str(tmp_path / "work"))
    print("case: owl_fwd2d forward+adjoint 61x81 R=4", flush=True)
    r = subprocess.run([DRIVER, job], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    syn, sdt = read_su(str(tmp_path / "syn" / "shot_1_seismogram_p.su"))
    assert sdt == pytest.approx(dt)
    assert np.array_equal(syn, ref)
    assert np.array_equal(syn, orc)
    gv = read_grd(str(tmp_path / "work" / "shot_1_grad_vp.grd"))
    gr = read_grd(str(tmp_path / "work" / "shot_1_grad_rho.grd"))
    assert np.abs(gv_py).max() > 0
    assert np.array_equal(gv, gv_py)
    assert np.array_equal(gr, gr_py)


def test_owl_callers_restated_over_the_drop_in_module(gpu, tmp_path):
    from owl_amd import set_adaptive_range
    from owl_amd.geometry import Geometry, Receiver, Source
    if not os.path.exists(CALLER):
        pytest.fail("owl_caller_2d is not built (__graft_entry__.build() builds it)")
    nz, nx, dx, dt, tmax, pml = 50, 120, 10.0, 0.001, 0.3, 15
    zz = np.arange(nz)[:, None] * dx
    vp = ((1900.0 + 1.5 * zz) * (1.0 + 0.04 * random_field((nz, nx), 4.0, 11))).astype(np.float32)
    rho = gardner(vp)
    f_vp, f_rho = str(tmp_path / "vp.bin"), str(tmp_path / "rho.bin")
    np.asfortranarray(vp).T.tofile(f_vp)
    np.asfortranarray(rho).T.tofile(f_rho)
    shots = [(3, 255.0, 20.0, 14.0), (7, 612.5, 35.0, 12.0), (9, 1010.0, 20.0, 15.0)]
    recs = [[(30.0 + 10.0 * k, 20.0) for k in range(110)], [(500.0 + 7.5 * k, 40.0) for k in range(60)],
            [(600.0 + 9.0 * k, 25.0) for k in range(60)]]
    dirs = {d: str(tmp_path / d) for d in ("syn", "scratch", "adj")}
    for d in dirs.values():
        os.makedirs(d)
    lines = [f"{nx} {nz} {dx} {dx} {dt} {tmax} {pml}", f_vp, f_rho, "1 150.0", f"{len(shots)}"]
    for (sid, sx, sz, f0), rr in zip(shots, recs):
        lines.append(f"{sid} {sx} {sz} {f0} {len(rr)}")
        lines += [f"{x!r} {z!r}" for x, z in rr]
    lines += [dirs["syn"], dirs["scratch"], dirs["adj"]]
    job = str(tmp_path / "job.txt")
    with open(job, "w") as f:
        f.write("\n".join(lines) + "\n")

    # the Python batch on the same windows (OWL's rule through the C ABI)
    geoms, wins = [], []
    for (sid, sx, sz, f0), rr in zip(shots, recs):
        g = Geometry(id=sid, srcr=[Source(x=sx, z=sz, f0=f0)], recr=[Receiver(x=x, z=z) for x, z in rr])
        wins.append(set_adaptive_range(g, nx, nz, dx, dx, yn_adpx=True, adp_extrax=150.0))
        geoms.append(g)
    assert any(w.nx < nx for w in wins)
    b = Acoustic2DBatch(nx, nz, dx, dx, vp, rho, list(zip(geoms, wins)), dt, tmax, pml=pml)
    ref = b.forward(save_boundary=True)

    print("case: restated module_modeling loop, 3 shots with adaptive windows", flush=True)
    r = subprocess.run([CALLER, "forward", job], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    adjs = []
    for (sid, _, _, _), tr in zip(shots, ref):
        syn, _ = read_su(os.path.join(dirs["syn"], f"shot_{sid}_seismogram_p.su"))
        assert np.abs(tr).max() > 0
        assert np.array_equal(syn, tr)
        adj = (0.5 * tr + 1e-4 * np.random.default_rng(sid).standard_normal(tr.shape)).astype(np.float32)
        write_su(os.path.join(dirs["adj"], f"shot_{sid}_seismogram_p.su"), adj, dt)
        adjs.append(adj)
    grads = b.adjoint(adjs)
    b.close()

    print("case: restated module_gradient loop (forward with reconstruct, adjoint)", flush=True)
    r = subprocess.run([CALLER, "gradient", job], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    for (sid, _, _, _), (gv, gr), w in zip(shots, grads, wins):
        fv = read_grd(os.path.join(dirs["scratch"], f"shot_{sid}_grad_vp.grd"))
        fr = read_grd(os.path.join(dirs["scratch"], f"shot_{sid}_grad_rho.grd"))
        assert fv.shape == (w.nz, w.nx)
        assert np.abs(gv).max() > 0
        assert np.array_equal(fv, gv)
        assert np.array_equal(fr, gr)


SOLVERS = os.path.join(ROOT, "tests", "fortran", "owl_solvers_3d_tti")


def _read_grd3(path):
    raw = open(path, "rb").read()
    head = raw[:4000].decode("ascii", errors="replace")
    kv = {k.strip(): v.strip() for k, v in (ln.split("=", 1) for ln in head.split("\r\n") if "=" in ln)}
    n = [int(kv[f"n{a}"]) for a in (1, 2, 3)]
    return np.frombuffer(raw[4000:], "<f4")[: n[0] * n[1] * n[2]].reshape(n[2], n[1], n[0]).transpose(2, 1, 0)


def test_acoustic_iso_3d_drop_in(gpu, tmp_path):
    from owl_amd import Acoustic3DBatch
    from owl_amd.geometry import Geometry, Receiver, Source
    if not os.path.exists(SOLVERS):
        pytest.fail("owl_solvers_3d_tti is not built (__graft_entry__.build() builds it)")
    nz, ny, nx, d, dt, tmax, pml = 22, 26, 30, 10.0, 0.001, 0.15, 8
    vp = (2200.0 * (1 + 0.05 * random_field((nz, ny, nx), 3.0, 5))).astype(np.float32)
    rho = gardner(vp)
    files = {k: str(tmp_path / f"{k}.bin") for k in ("vp", "rho")}
    np.asfortranarray(vp).T.tofile(files["vp"])
    np.asfortranarray(rho).T.tofile(files["rho"])
    rec = [(20.0 + 9.0 * k, 60.5, 30.0) for k in range(28)]
    dirs = {k: str(tmp_path / k) for k in ("syn", "adj", "work")}
    for v in dirs.values():
        os.makedirs(v)
    geom = Geometry(id=4, srcr=[Source(x=143.0, y=121.7, z=95.0, f0=20.0)],
                    recr=[Receiver(x=x, y=y, z=z) for x, y, z in rec])
    b = Acoustic3DBatch(nx, ny, nz, d, d, d, vp, rho, [(geom, None)], dt, tmax, pml=pml, fdhalf=4)
    ref = b.forward(save_boundary=True)[0]
    adj = (0.5 * ref).astype(np.float32)
    gv, gr = b.adjoint([adj])[0]
    b.close()
    write_su(os.path.join(dirs["adj"], "shot_4_seismogram_p.su"), adj, dt)
    job = str(tmp_path / "job3d.txt")
    with open(job, "w") as f:
        f.write("\n".join([f"{nx} {ny} {nz} {d} {dt} {tmax} {pml}", files["vp"], files["rho"],
                           "4 143.0 121.7 95.0 20.0", f"{len(rec)}"] + [f"{x!r} {y!r} {z!r}" for x, y, z in rec]
                          + [dirs["syn"], dirs["adj"], dirs["work"]]) + "\n")
    print("case: acoustic_iso_3d drop-in forward + adjoint", flush=True)
    r = subprocess.run([SOLVERS, "3d", job], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    syn, _ = read_su(os.path.join(dirs["syn"], "shot_4_seismogram_p.su"))
    assert np.abs(ref).max() > 0 and np.array_equal(syn, ref)
    fv = _read_grd3(os.path.join(dirs["work"], "shot_4_grad_vp.grd"))
    fr = _read_grd3(os.path.join(dirs["work"], "shot_4_grad_rho.grd"))
    assert fv.shape == (nz, ny, nx) and np.abs(gv).max() > 0
    assert np.array_equal(fv, gv) and np.array_equal(fr, gr)


def test_elastic_tti_2d_drop_in_thomsen(gpu, tmp_path):
    _tti_drop_in(tmp_path, surface=None)


def test_elastic_tti_2d_drop_in_free_surface(gpu, tmp_path):
    """The drop-in with free_surface, a file_topo and depths measured from
    the surface equals the library batch with the same surface."""
    xt = np.linspace(-300.0, 900.0, 25).astype(np.float32)
    zt = (12.0 * np.sin(xt / 150.0) - 0.06 * xt).astype(np.float32)
    _tti_drop_in(tmp_path, surface=dict(topo=(xt, zt), interp="cubic", refine=3.0, dz_max=14.0,
                                        src_from_surface=True, rec_from_surface=True))


def _tti_drop_in(tmp_path, surface):
    from owl_amd import FSG2DBatch
    from owl_amd.fsg2d import tti_cij, tti_param_gradient
    from owl_amd.geometry import Geometry, Receiver, Source
    if not os.path.exists(SOLVERS):
        pytest.fail("owl_solvers_3d_tti is not built (__graft_entry__.build() builds it)")
    nz, nx, d, dt, tmax, pml = 44, 56, 10.0, 0.8e-3, 0.25, 10
    vp = (2600.0 * (1 + 0.05 * random_field((nz, nx), 3.0, 3))).astype(np.float32)
    vs = (vp / 1.8).astype(np.float32)
    rho = gardner(vp)
    eps = np.full((nz, nx), 0.2, np.float32)
    dlt = np.full((nz, nx), 0.08, np.float32)
    the = (0.4 + 0.3 * random_field((nz, nx), 5.0, 4)).astype(np.float32)
    names = ("vp", "vs", "rho", "eps", "del", "the")
    files = {k: str(tmp_path / f"{k}.bin") for k in names}
    for k, a in zip(names, (vp, vs, rho, eps, dlt, the)):
        np.asfortranarray(a).T.tofile(files[k])
    rec = [(20.0 + 9.0 * k, 35.0 + k) for k in range(50)]
    dirs = {k: str(tmp_path / k) for k in ("syn", "adj", "work")}
    for v in dirs.values():
        os.makedirs(v)
    geom = Geometry(id=2, srcr=[Source(x=213.0, z=181.0, f0=20.0, mechanism="force")],
                    recr=[Receiver(x=x, z=z) for x, z in rec])
    cij = tti_cij("thomsen", vp, vs, rho, eps, dlt, the)
    if surface is not None:
        dt = 0.4e-3  # the refined rows near the surface
    b = FSG2DBatch(nx, nz, d, d, cij, rho, [(geom, None)], dt, tmax, pml=pml, fdhalf=4, aniso=True,
                   surface=surface)
    tx, tz = b.forward(save_boundary=True)[0]
    ax, az = (0.5 * tx).astype(np.float32), (0.5 * tz).astype(np.float32)
    g = b.adjoint([(ax, az)])[0]
    b.close()
    pg = tti_param_gradient("thomsen", g, vp, vs, rho, eps, dlt, the)
    write_su(os.path.join(dirs["adj"], "shot_2_seismogram_x.su"), ax, dt)
    write_su(os.path.join(dirs["adj"], "shot_2_seismogram_z.su"), az, dt)
    job = str(tmp_path / "jobtti.txt")
    extra = []
    if surface is not None:
        topo = str(tmp_path / "topo.txt")
        with open(topo, "w") as f:
            for x, z in zip(*surface["topo"]):
                f.write(f"{float(x)!r} {float(z)!r}\n")
        extra = [f"{topo} {surface['refine']!r} {surface['dz_max']!r} 1 1"]
    with open(job, "w") as f:
        f.write("\n".join([f"{nx} {nz} {d} {dt!r} {tmax} {pml}", "thomsen"] + [files[k] for k in names]
                          + ["2 213.0 181.0 20.0", f"{len(rec)}"] + [f"{x!r} {z!r}" for x, z in rec]
                          + [dirs["syn"], dirs["adj"], dirs["work"]] + extra) + "\n")
    print(f"case: elastic_tti_2d drop-in forward + adjoint, Thomsen, surface {surface is not None}", flush=True)
    r = subprocess.run([SOLVERS, "tti", job], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    sx, _ = read_su(os.path.join(dirs["syn"], "shot_2_seismogram_x.su"))
    sz, _ = read_su(os.path.join(dirs["syn"], "shot_2_seismogram_z.su"))
    assert np.abs(tz).max() > 0 and np.array_equal(sx, tx) and np.array_equal(sz, tz)
    for k in ("vp", "vs", "epsilon", "delta"):
        a = read_grd(os.path.join(dirs["work"], f"shot_2_grad_{k}.grd"))
        assert np.abs(pg["grad_" + k]).max() > 0
        assert np.array_equal(a, pg["grad_" + k]), k
    assert np.array_equal(read_grd(os.path.join(dirs["work"], "shot_2_grad_rho.grd")), g["grad_rho"])
