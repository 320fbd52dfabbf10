// L2 waveform misfit and adjoint source of two-component (vx, vz) traces,
// shared by the 2D elastic solvers.  ShotT provides nr, rec0 and trace0;
// traces hold [2][nt][nr] per shot.
#pragma once
#include <hip/hip_runtime.h>

namespace owl {

// L2 waveform misfit and adjoint source per shot and component (as in 3D):
// discard rule on each component's own largest trace norms, FP64 Neumaier
// sums, traces [nt][nr] per component.
template <class ShotT>
__global__ __launch_bounds__(256) void k_adjoint_source_2c(const ShotT* __restrict__ shots,
                                                           const float* __restrict__ syn, const float* __restrict__ obs,
                                                           const float* __restrict__ rec_w, float* adj,
                                                           double* misfit, int nt, float threshold, double* scratch) {
    const ShotT S = shots[blockIdx.x];
    const int comp = blockIdx.y;
    const size_t base = S.trace0 + size_t(comp) * nt * S.nr;
    __shared__ double red[256];
    __shared__ double mo_s, ms_s;
    double* nobs = scratch + 2 * (2 * size_t(S.rec0) + size_t(comp) * S.nr);
    double* nsyn = nobs + S.nr;
    double lmo = 0, lms = 0;
    for (int r = threadIdx.x; r < S.nr; r += blockDim.x) {
        double a = 0, b = 0;
        for (int t = 0; t < nt; ++t) {
            const double o = obs[base + size_t(t) * S.nr + r];
            const double sv = syn[base + size_t(t) * S.nr + r];
            a += o * o;
            b += sv * sv;
        }
        nobs[r] = sqrt(a);
        nsyn[r] = sqrt(b);
        lmo = fmax(lmo, nobs[r]);
        lms = fmax(lms, nsyn[r]);
    }
    red[threadIdx.x] = lmo;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) mo_s = red[0];
    __syncthreads();
    red[threadIdx.x] = lms;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) ms_s = red[0];
    __syncthreads();
    const double mo = mo_s, ms = ms_s;
    double sum = 0, cmp = 0;
    for (int r = threadIdx.x; r < S.nr; r += blockDim.x) {
        const bool keep = !(rec_w[S.rec0 + r] == 0.0f || nobs[r] < threshold * mo || nsyn[r] < threshold * ms ||
                            nobs[r] * nsyn[r] == 0.0);
        for (int t = 0; t < nt; ++t) {
            const size_t q = base + size_t(t) * S.nr + r;
            if (keep) {
                const float d = obs[q] - syn[q];
                const double d2 = double(d) * double(d);
                const double y = sum + d2;
                cmp += (fabs(sum) >= fabs(d2)) ? (sum - y) + d2 : (d2 - y) + sum;
                sum = y;
                adj[q] = syn[q] - obs[q];
            } else {
                adj[q] = 0.0f;
            }
        }
    }
    red[threadIdx.x] = sum + cmp;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) misfit[2 * blockIdx.x + comp] = red[0];
}

}  // namespace owl
