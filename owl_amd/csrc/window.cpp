// Per-shot aperture window: OWL's set_adaptive_range
// (src/lib/module_utility.f90:410-510) with the absolute-offset rule of
// prepare_geometry (src/lib/module_source_receiver.f90:421-436).
//
// All arithmetic is float32 in the order the Fortran writes it, so the
// window's node ranges and the receivers that keep a non-zero weight are
// OWL's: floor/ceiling snap the window ends to the grid, int() truncates
// toward zero, and the absolute offset is sqrt of the three squared float32
// differences summed left to right.
#include <cmath>
#include <limits>

#include "common.hpp"
#include "owl_amd.h"

using namespace owl;

namespace {

struct AxisRange {
    float beg, end;
    int nbeg, nend;
};

// Mean of the source coordinates (FLIT's mean: float32 sum over the array
// in index order divided by its size).
float source_mean(int ns, const owl_source* src, int axis) {
    float s = 0.0f;
    for (int k = 0; k < ns; ++k) s += axis == 0 ? src[k].x : axis == 1 ? src[k].y : src[k].z;
    return s / float(ns);
}

float coord(const owl_receiver& r, int axis) { return axis == 0 ? r.x : axis == 1 ? r.y : r.z; }
float coord(const owl_source& s, int axis) { return axis == 0 ? s.x : axis == 1 ? s.y : s.z; }

int clip(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// One axis of the window.  When adaptive, the range spans the live
// receivers and every source, widened by `extra`, clipped to the model
// extent [vmin, vmax] and snapped outward to the grid; receivers outside it
// lose their weight.  Otherwise it is the whole axis.
AxisRange adaptive_axis(int axis, bool adaptive, float extra, float vmin, float vmax, float o, float d, int n,
                        int ns, const owl_source* src, int nr, owl_receiver* rec) {
    AxisRange A{};
    if (adaptive) {
        // minval/maxval with an all-false mask give +huge/-huge
        float rlo = std::numeric_limits<float>::max(), rhi = -std::numeric_limits<float>::max();
        for (int k = 0; k < nr; ++k) {
            if (rec[k].weight == 0.0f) continue;
            const float c = coord(rec[k], axis);
            if (c < rlo) rlo = c;
            if (c > rhi) rhi = c;
        }
        float slo = coord(src[0], axis), shi = slo;
        for (int k = 1; k < ns; ++k) {
            const float c = coord(src[k], axis);
            if (c < slo) slo = c;
            if (c > shi) shi = c;
        }
        const float lo = (rlo < slo ? rlo : slo) - extra;
        const float hi = (rhi > shi ? rhi : shi) + extra;
        float beg = vmin > lo ? vmin : lo;
        float end = vmax < hi ? vmax : hi;
        const float kb = float(int(std::floor((beg - o) / d)));
        const float ke = float(int(std::ceil((end - o) / d)));
        beg = kb * d + o;
        end = ke * d + o;
        for (int k = 0; k < nr; ++k) {
            const float c = coord(rec[k], axis);
            if (c < beg || c > end) rec[k].weight = 0.0f;
        }
        A.beg = beg;
        A.end = end;
    } else {
        A.beg = o;
        A.end = o + float(n - 1) * d;
    }
    A.nbeg = clip(int((A.beg - o) / d) + 1, 1, n);
    A.nend = clip(int((A.end - o) / d) + 1, 1, n);
    return A;
}

}  // namespace

extern "C" owl_status owl_set_adaptive_range(const owl_window_params* p, int ns, const owl_source* src, int nr,
                                             owl_receiver* rec, int* window, float* range) {
    return guarded([&] {
        if (!p || !window) fail(OWL_ERR_ARG, "null argument");
        if (ns < 1 || !src) fail(OWL_ERR_ARG, "a shot needs at least one source");
        if (nr < 0 || (nr > 0 && !rec)) fail(OWL_ERR_ARG, "bad receiver array");
        if (p->nx < 1 || p->ny < 1 || p->nz < 1 || !(p->dx > 0) || !(p->dy > 0) || !(p->dz > 0))
            fail(OWL_ERR_ARG, "invalid model grid");
        // receivers outside [offset_min, offset_max] first
        const float mx = source_mean(ns, src, 0), my = source_mean(ns, src, 1), mz = source_mean(ns, src, 2);
        for (int k = 0; k < nr; ++k) {
            const float ax = rec[k].x - mx, ay = rec[k].y - my, az = rec[k].z - mz;
            const float aoff = std::sqrt(ax * ax + ay * ay + az * az);
            if (aoff < p->offset_min || aoff > p->offset_max) rec[k].weight = 0.0f;
        }
        const AxisRange X = adaptive_axis(0, p->yn_adpx != 0, p->adp_extrax, p->xmin, p->xmax, p->ox, p->dx, p->nx,
                                          ns, src, nr, rec);
        const AxisRange Y = adaptive_axis(1, p->yn_adpy != 0, p->adp_extray, p->ymin, p->ymax, p->oy, p->dy, p->ny,
                                          ns, src, nr, rec);
        const AxisRange Z = adaptive_axis(2, p->yn_adpz != 0, p->adp_extraz, p->zmin, p->zmax, p->oz, p->dz, p->nz,
                                          ns, src, nr, rec);
        window[0] = X.nbeg;
        window[1] = X.nend;
        window[2] = Y.nbeg;
        window[3] = Y.nend;
        window[4] = Z.nbeg;
        window[5] = Z.nend;
        if (range) {
            range[0] = X.beg;
            range[1] = X.end;
            range[2] = Y.beg;
            range[3] = Y.end;
            range[4] = Z.beg;
            range[5] = Z.end;
        }
    });
}
