// Stateless entry points of the C ABI and the per-thread error message.
#include <algorithm>
#include <cstring>
#include <string>

#include "common.hpp"
#include "host_prep.hpp"
#include "owl_amd.h"

namespace owl {
namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string& msg) { g_last_error = msg; }
void clear_last_error() { g_last_error.clear(); }
const char* last_error_cstr() { return g_last_error.c_str(); }
}  // namespace owl

using namespace owl;

extern "C" {

const char* owl_version(void) { return "owl_amd 0.1.0 (gfx950)"; }

const char* owl_last_error(void) { return last_error_cstr(); }

float owl_dt_stable(int fdhalf, int dims, float vmax, float dx, float dy, float dz) {
    if (fdhalf < 2 || fdhalf > 8) return 0.0f;
    return dt_stable(fdhalf, dims, vmax, dx, dy, dz);
}

owl_status owl_locate_point(int dims, float x, float y, float z, float ox, float oy, float oz, float dx,
                            float dy, float dz, int* idx, float* w) {
    return guarded([&] {
        if (!idx || !w) fail(OWL_ERR_ARG, "null output");
        if (dims != 2 && dims != 3) fail(OWL_ERR_ARG, "dims must be 2 or 3");
        const Located L = locate(dims, x, y, z, ox, oy, oz, dx, dy, dz);
        idx[0] = L.gx; idx[1] = L.gy; idx[2] = L.gz;
        idx[3] = L.hx; idx[4] = L.hy; idx[5] = L.hz;
        std::memcpy(w + 0, L.ix, sizeof(L.ix));
        std::memcpy(w + 9, L.iy, sizeof(L.iy));
        std::memcpy(w + 18, L.iz, sizeof(L.iz));
        std::memcpy(w + 27, L.hxw, sizeof(L.hxw));
        std::memcpy(w + 36, L.hyw, sizeof(L.hyw));
        std::memcpy(w + 45, L.hzw, sizeof(L.hzw));
    });
}

int owl_source_time_function(int wavelet, float f0, float dt, float f0factor, float* stf, int cap, int* hnt) {
    if (!(f0 > 0) || !(dt > 0)) return 0;
    int h = 0;
    const std::vector<float> w = source_time_function(wavelet, f0, dt, f0factor > 0 ? f0factor : 1.0f, &h);
    if (hnt) *hnt = h;
    if (stf) std::copy(w.begin(), w.begin() + std::min<size_t>(w.size(), size_t(std::max(cap, 0))), stf);
    return int(w.size());
}

owl_status owl_cfspml_axis(int n, int pml, float d, float dt, float pmlvp, float f0max, float* ai, float* bi,
                           float* ah, float* bh) {
    return guarded([&] {
        if (n < 1 || pml < 1 || !(d > 0)) fail(OWL_ERR_ARG, "invalid axis");
        const PmlAxis P = cfspml_axis(n, pml, d, dt, pmlvp, f0max);
        const size_t len = size_t(n + 2 * pml);
        if (ai) std::copy(P.ai.begin(), P.ai.begin() + len, ai);
        if (bi) std::copy(P.bi.begin(), P.bi.begin() + len, bi);
        if (ah) std::copy(P.ah.begin(), P.ah.begin() + len, ah);
        if (bh) std::copy(P.bh.begin(), P.bh.begin() + len, bh);
    });
}

}  // extern "C"
