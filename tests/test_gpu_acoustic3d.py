"""GPU parity of the 3D acoustic forward (C ABI -> HIP) against the oracle:
traces bit-identical to the float32 oracle (same operation order, no FMA
contraction), per shot, for a batch with uneven windows; independent of
the z-chunk launch shape."""
import numpy as np
import pytest

from helpers import rel_l2
from owl_amd import Acoustic3DBatch, Geometry, Receiver, Source
from owl_amd.synthetic import gardner, random_field

pytestmark = pytest.mark.gpu


def _model(nz, ny, nx, seed=0):
    vp = (2200.0 * (1 + 0.05 * random_field((nz, ny, nx), 3.0, seed))).astype(np.float32)
    return vp, gardner(vp)


def _oracle(o, vp, rho, win, geom, dx, dt, nt, pml, R):
    nz, ny, nx = vp.shape
    w = dict(nxbeg=1, nxend=nx, nybeg=1, nyend=ny, nzbeg=1, nzend=nz)
    w.update(win or {})
    sub_vp = vp[w["nzbeg"] - 1:w["nzend"], w["nybeg"] - 1:w["nyend"], w["nxbeg"] - 1:w["nxend"]]
    sub_rho = rho[w["nzbeg"] - 1:w["nzend"], w["nybeg"] - 1:w["nyend"], w["nxbeg"] - 1:w["nxend"]]
    f = np.float32
    g = dict(nx=sub_vp.shape[2], ny=sub_vp.shape[1], nz=sub_vp.shape[0], dx=dx, dy=dx, dz=dx,
             ox=float(f(w["nxbeg"] - 1) * f(dx)), oy=float(f(w["nybeg"] - 1) * f(dx)),
             oz=float(f(w["nzbeg"] - 1) * f(dx)), dt=dt, nt=nt, pml=pml, R=R)
    x1 = f(g["nx"] - 1) * f(dx) + f(g["ox"])
    y1 = f(g["ny"] - 1) * f(dx) + f(g["oy"])
    z1 = f(g["nz"] - 1) * f(dx) + f(g["oz"])
    recs = []
    for r in geom.recr:
        ok = g["ox"] <= f(r.x) <= x1 and g["oy"] <= f(r.y) <= y1 and g["oz"] <= f(r.z) <= z1
        recs.append(dict(x=r.x, y=r.y, z=r.z, weight=r.weight if ok else 0.0))
    srcs = [dict(x=s.x, y=s.y, z=s.z, f0=s.f0) for s in geom.srcr]
    return o.acoustic3d_forward(g, sub_vp, sub_rho, srcs, recs)


@pytest.mark.parametrize("R", [4, 2, 8])
def test_forward3d_matches_oracle(oracle32, gpu, R):
    nz, ny, nx, dx, dt, tmax, pml = 22, 26, 30, 10.0, 0.001, 0.12, max(8, R)
    vp, rho = _model(nz, ny, nx, seed=R)
    geoms = [
        (Geometry(id=1, srcr=[Source(x=143.0, y=121.7, z=95.0, f0=20.0)],
                  recr=[Receiver(x=20.0 + 9.0 * k, y=60.5, z=30.0) for k in range(28)]), None),
        (Geometry(id=2, srcr=[Source(x=210.0, y=150.0, z=80.0, f0=20.0)],
                  recr=[Receiver(x=130.0 + 10.0 * k, y=140.0, z=40.0) for k in range(15)]),
         dict(nxbeg=12, nxend=30, nybeg=5, nyend=26, nzbeg=1, nzend=20)),
    ]
    print(f"case: 3D forward R={R}", flush=True)
    b = Acoustic3DBatch(nx, ny, nz, dx, dx, dx, vp, rho, geoms, dt, tmax, pml=pml, fdhalf=R)
    outs = [b.forward()]
    for kc, var in ((16, 0), (64, 8), (5, 24)):
        b.set_launch(kc, var)
        outs.append(b.forward())
    for o in outs[1:]:
        for a, c in zip(o, outs[0]):
            assert np.array_equal(a, c)
    out = outs[0]
    for (geom, win), g in zip(geoms, out):
        ref = _oracle(oracle32, vp, rho, win, geom, dx, dt, b.nt, pml, R)
        assert np.abs(ref).max() > 0
        assert rel_l2(g, ref) <= 1e-6, rel_l2(g, ref)
        assert np.array_equal(g, ref), f"shot {geom.id}: max diff {np.abs(g - ref).max()}"
    b.close()


def test_forward3d_zchunk_invariance(gpu):
    nz, ny, nx, dx = 40, 20, 24, 10.0
    vp, rho = _model(nz, ny, nx, seed=9)
    geom = Geometry(id=1, srcr=[Source(x=115.0, y=95.0, z=200.0, f0=20.0)],
                    recr=[Receiver(x=20.0 + 10.0 * k, y=90.0, z=50.0) for k in range(20)])
    print("case: 3D z-chunk invariance", flush=True)
    b = Acoustic3DBatch(nx, ny, nz, dx, dx, dx, vp, rho, [(geom, None)], 0.001, 0.1, pml=8)
    ref = b.forward()[0]
    for kc, var in ((1, 1), (7, 0), (13, 1), (200, 0), (1, 8), (13, 24), (64, 8), (64, 1)):
        b.set_launch(kc, var)
        assert np.array_equal(b.forward()[0], ref), (kc, var)
    b.close()
