"""OWL's on-disk formats at the solver boundary.

* SU traces (module_su.f90:1454-1607): 240-byte little-endian trace header
  (ns int16 at byte 114, dt in microseconds int16 at 116, d1 float at 180,
  ntr int32 at 204) followed by nt float32 samples, trace after trace.
* 2D grid files (module_grid.f90:317-477): 50 lines of 80 ASCII bytes
  (key = value, CR LF at columns 79-80), then float32 values of the
  (n1, n2) array in column-major order starting at byte 4001.
"""
import os

import numpy as np

SU_HEADER = 240


def write_su(path, traces, dt):
    """Write traces (nr, nt) as an SU file (init_su sets ns, dt and d1)."""
    tr = np.ascontiguousarray(traces, dtype="<f4")
    nr, nt = tr.shape
    hdr = np.zeros((nr, SU_HEADER), np.uint8)
    h16 = np.zeros(1, "<i2")
    for ir in range(nr):
        h16[0] = nt
        hdr[ir, 114:116] = h16.view(np.uint8)
        h16[0] = int(np.float32(dt) * np.float32(1.0e6))
        hdr[ir, 116:118] = h16.view(np.uint8)
        hdr[ir, 180:184] = np.array([dt], "<f4").view(np.uint8)
        hdr[ir, 204:208] = np.array([nr], "<i4").view(np.uint8)
    with open(path, "wb") as f:
        for ir in range(nr):
            f.write(hdr[ir].tobytes())
            f.write(tr[ir].tobytes())


def read_su(path):
    """Read an SU file: returns (traces (nr, nt) float32, dt)."""
    raw = np.fromfile(path, np.uint8)
    if raw.size < SU_HEADER:
        raise ValueError(f"{path}: not an SU file")
    nt = int(raw[114:116].view("<i2")[0])
    if nt <= 0:
        raise ValueError(f"{path}: bad ns")
    rec = SU_HEADER + 4 * nt
    if raw.size % rec:
        raise ValueError(f"{path}: size is not a multiple of the trace length")
    nr = raw.size // rec
    d1 = float(raw[180:184].view("<f4")[0])
    dt = d1 if d1 > 0 else int(raw[116:118].view("<i2")[0]) * 1.0e-6
    body = raw.reshape(nr, rec)[:, SU_HEADER:]
    return np.ascontiguousarray(body).view("<f4").reshape(nr, nt).astype(np.float32), dt


def _header_lines(n, d, o):
    lines = ["# Start 4000-byte ASCII header "]
    for ax in (1, 2):
        k = ax - 1
        lines += [f"n{ax} = {n[k]}", f"d{ax} = {d[k]:.7E}", f"o{ax} = {o[k]:.7E}",
                  f"l{ax} = {o[k] + (n[k] - 1) * d[k]:.7E}", f"label{ax} = Axis {ax}",
                  f"unit{ax} = unitless"]
    lines += [""] * (49 - len(lines))
    lines.append("# End 4000-byte ASCII header")
    out = bytearray()
    for ln in lines:
        b = ln.encode("ascii")[:78].ljust(78, b" ")
        out += b + b"\r\n"
    assert len(out) == 4000
    return bytes(out)


def write_grd(path, array, d=(1.0, 1.0), o=(0.0, 0.0)):
    """Write a 2D grid file; array is (n1, n2) = (nz, nx)."""
    a = np.asarray(array, dtype="<f4")
    with open(path, "wb") as f:
        f.write(_header_lines(a.shape, d, o))
        f.write(np.asfortranarray(a).tobytes(order="F"))


def read_grd(path):
    """Read a 2D grid file: returns the (n1, n2) float32 array."""
    raw = open(path, "rb").read()
    head = raw[:4000].decode("ascii", errors="replace")
    kv = {}
    for ln in head.split("\r\n"):
        if "=" in ln:
            k, v = ln.split("=", 1)
            kv[k.strip()] = v.strip()
    body = np.frombuffer(raw[4000:], "<f4")
    n1 = int(kv.get("n1", body.size))
    n2 = int(kv.get("n2", body.size // max(n1, 1)))
    return body[: n1 * n2].reshape(n2, n1).T.astype(np.float32)


def shot_file(directory, shot_id, kind, name):
    """OWL's per-shot file names: shot_<id>_seismogram_<c>.su, shot_<id>_grad_<m>.grd."""
    if kind == "seismogram":
        return os.path.join(directory, f"shot_{shot_id}_seismogram_{name}.su")
    return os.path.join(directory, f"shot_{shot_id}_grad_{name}.grd")


def resample(traces, dt, nnt, ddt, oot=0.0):
    """su%resamp (module_su.f90:425-468) through owl_resample_traces:
    traces (nr, nt) sampled at k dt to (nr, nnt) at oot + k ddt, FLIT's
    interp 'cubic' per trace, all-zero traces kept zero.  Returns the input
    unchanged when nnt, ddt and oot already match, as OWL does."""
    from . import _lib
    tr = np.ascontiguousarray(traces, dtype=np.float32)
    if tr.ndim != 2:
        raise ValueError("traces must be (nr, nt)")
    nr, nt = tr.shape
    if nt == nnt and np.float32(dt) == np.float32(ddt) and oot == 0:
        return tr.copy()
    out = np.empty((nr, int(nnt)), np.float32)
    _lib.check(_lib.lib().owl_resample_traces(nr, nt, dt, tr.ctypes.data, int(nnt), ddt, oot, out.ctypes.data))
    return out
