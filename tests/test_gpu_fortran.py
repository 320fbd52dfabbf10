"""The Fortran host side (owl_amd/fortran: iso_c_binding module, the
wave_solver_acoustic_iso_2d mirror and the owl_fwd2d driver) runs the same
C ABI as the Python host and writes OWL's files: its SU traces must equal the
Python path's and the float32 oracle's bit for bit, and its .grd gradients
the Python adjoint's.
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

DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "owl_amd", "fortran", "owl_fwd2d")


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
