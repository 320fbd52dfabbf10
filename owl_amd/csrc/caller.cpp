// Caller-side helpers of OWL's drivers that sit on the solver boundary:
//
// * owl_resample_traces: su%resamp (src/lib/module_su.f90:425-468): every
//   trace that is not all zero becomes FLIT's
//   interp(w, nt, dt, 0, nnt, ddt, oot, 'cubic'); all-zero traces stay zero.
//   FLIT's interp is not part of the OWL tree; its 'cubic' is taken as the
//   natural cubic spline through the samples evaluated in double, the end
//   segments' cubics extended beyond the first and last sample (the same
//   restatement as the topography's ginterp of the free-surface mesh).
// * owl_ginterp: FLIT's ginterp(x, y, xi, method) with method 'linear' or
//   'cubic' as load_geometry and the free-surface mesh call it
//   (module_source_receiver.f90:302-316).
#include <vector>

#include "common.hpp"
#include "fs_mesh.hpp"
#include "owl_amd.h"

using namespace owl;

extern "C" owl_status owl_resample_traces(int nr, int nt, float dt, const float* in, int nnt, float ddt, float oot,
                                   float* out) {
    return guarded([&] {
        if (nr < 0 || nt < 1 || nnt < 1) fail(OWL_ERR_ARG, "resample: nr >= 0, nt >= 1 and nnt >= 1 are required");
        if (!(dt > 0) || !(ddt > 0)) fail(OWL_ERR_ARG, "resample: dt and ddt must be positive");
        if (nr > 0 && (!in || !out)) fail(OWL_ERR_ARG, "resample: null trace array");
        std::vector<float> t(static_cast<size_t>(nt)), tq(static_cast<size_t>(nnt));
        for (int k = 0; k < nt; ++k) t[size_t(k)] = 0.0f + float(k) * dt;
        for (int k = 0; k < nnt; ++k) tq[size_t(k)] = oot + float(k) * ddt;
        for (int r = 0; r < nr; ++r) {
            const float* w = in + size_t(r) * size_t(nt);
            float* o = out + size_t(r) * size_t(nnt);
            bool live = false;
            for (int k = 0; k < nt && !live; ++k) live = w[k] != 0.0f;
            if (!live) {
                for (int k = 0; k < nnt; ++k) o[k] = 0.0f;
                continue;
            }
            fsm::ginterp(t.data(), w, nt, tq.data(), nnt, o, true);
        }
    });
}

extern "C" owl_status owl_ginterp(int n, const float* x, const float* y, int m, const float* xi, float* yi, int cubic) {
    return guarded([&] {
        if (n < 1 || m < 0) fail(OWL_ERR_ARG, "ginterp: n >= 1 and m >= 0 are required");
        if (!x || !y || (m > 0 && (!xi || !yi))) fail(OWL_ERR_ARG, "ginterp: null array");
        for (int k = 1; k < n; ++k)
            if (!(x[k] > x[k - 1])) fail(OWL_ERR_ARG, "ginterp: x must increase");
        fsm::ginterp(x, y, n, xi, m, yi, cubic != 0);
    });
}
