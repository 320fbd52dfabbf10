"""Host side of the FSG path without a GPU: the library's M-PML
damping-ratio scan (host_prep.cpp, used when a batch is planned) is
bit-identical to the oracle's restatement of OWL's damp_ratio for
isotropic, VTI and tilted-TI media and the Becache model-III medium, in
both directions."""
import numpy as np

from owl_amd.fsg2d import CIJ, damp_ratio
from test_oracle_pinning_fsg2d import tilted_ti


def _media():
    out = []
    for (vp, vs, rho, eps, dl, th) in [(3000, 1700, 2000, 0, 0, 0), (3000, 1500, 2000, 0.3, 0.05, 0.0),
                                        (3000, 1500, 2000, 0.3, 0.05, 0.4), (2500, 1400, 2100, 0.15, 0.1, 0.8),
                                        (4500, 2500, 2500, 0.25, 0.0, 1.3), (1800, 1000, 1900, 0.1, 0.15, 0.2)]:
        m = tilted_ti(vp, vs, rho, eps, dl, th)
        out.append(([float(m[k]) for k in CIJ], float(rho)))
    c, s = np.cos(0.5), np.sin(0.5)
    M = np.array([[c * c, s * s, 2 * c * s], [s * s, c * c, -2 * c * s], [-c * s, c * s, c * c - s * s]])
    Cr = M @ np.array([[4e9, 7.5e9, 0], [7.5e9, 20e9, 0], [0, 0, 2e9]]) @ M.T
    out.append(([Cr[0, 0], Cr[0, 1], Cr[0, 2], Cr[1, 1], Cr[1, 2], Cr[2, 2]], 1000.0))
    return out


def test_damp_ratio_matches_oracle(oracle32):
    seen = set()
    for c, rho in _media():
        c32 = [float(np.float32(v)) for v in c]
        for d in ("x", "z"):
            mine = damp_ratio(c32, rho, d)
            ref = oracle32.fsg_damp_ratio(c32, rho, d)
            assert mine == ref, (c32, rho, d, mine, ref)
            seen.add(round(mine, 3))
    assert len(seen) > 3  # the media exercise different ratios
