// Misfits and adjoint sources of OWL's FWI beyond the L2 waveform misfit,
// batched over the traces of a shot on the GPU
// (src/inverse/module_inversion_adjoint_source.f90: compute_adjsrc_waveform
// :371-383, compute_adjsrc_corr :391-413, compute_adjsrc_envelope :424-452,
// compute_adjsrc_phase :460-481, compute_adjsrc_adaptive :490-537 with
// xcorr_penalty :321-363; the per-shot driver and its trace-discard rule
// compute_adjoint_source :985-1160).
//
// Layout: traces receiver-major [ntr][nt], as the SU traces hold them.  One
// block per trace for the reductions (FP64, Neumaier-compensated like the
// library's L2 misfit); the Hilbert transforms are batched single-precision
// FFTs (hipFFT) of all traces at once: analytic signal z = x + i H(x) =
// IFFT(h . FFT(x)), h = 1 at DC (and Nyquist for even nt), 2 on the positive
// frequencies, 0 on the negative ones.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "common.hpp"
#include "owl_amd.h"

namespace owl {
namespace mf {

constexpr int kThreads = 256;

struct Cplx {
    float x, y;
};

// Neumaier-compensated block sum of one double per thread; every thread
// returns the total.
__device__ double block_sum(double v, double* sh) {
    double s = v, c = 0.0;
    sh[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        s = 0.0;
        for (int k = 0; k < int(blockDim.x); ++k) {
            const double a = sh[k], t = s + a;
            c += fabs(s) >= fabs(a) ? (s - t) + a : (a - t) + s;
            s = t;
        }
        sh[0] = s + c;
    }
    __syncthreads();
    const double r = sh[0];
    __syncthreads();
    return r;
}

__device__ float block_max(float v, float* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int w = int(blockDim.x) / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    const float r = sh[0];
    __syncthreads();
    return r;
}

// Per-trace L2 norms (FP64) of obs and syn.
__global__ __launch_bounds__(kThreads) void k_norms(const float* __restrict__ obs, const float* __restrict__ syn,
                                                    int nt, double* no, double* ns) {
    __shared__ double sh[kThreads];
    const size_t b = size_t(blockIdx.x) * nt;
    double a = 0, c = 0;
    for (int t = threadIdx.x; t < nt; t += blockDim.x) {
        const double o = obs[b + t], s = syn[b + t];
        a += o * o;
        c += s * s;
    }
    const double A = block_sum(a, sh), Cc = block_sum(c, sh);
    if (threadIdx.x == 0) {
        no[blockIdx.x] = sqrt(A);
        ns[blockIdx.x] = sqrt(Cc);
    }
}

// L2 waveform: misfit sum (obs - syn)^2, adjoint syn - obs.
__global__ __launch_bounds__(kThreads) void k_waveform(const float* __restrict__ obs, const float* __restrict__ syn,
                                                       const int* __restrict__ keep, int nt, float* adj,
                                                       double* mis) {
    __shared__ double sh[kThreads];
    const size_t b = size_t(blockIdx.x) * nt;
    const bool on = keep[blockIdx.x] != 0;
    double e = 0;
    for (int t = threadIdx.x; t < nt; t += blockDim.x) {
        const float o = obs[b + t], s = syn[b + t];
        const float d = o - s;
        e += double(d) * double(d);
        adj[b + t] = on ? s - o : 0.0f;
    }
    const double E = block_sum(e, sh);
    if (threadIdx.x == 0) mis[blockIdx.x] = on ? E : 0.0;
}

// Zero-lag normalised correlation (compute_adjsrc_corr).
__global__ __launch_bounds__(kThreads) void k_corr(const float* __restrict__ obs, const float* __restrict__ syn,
                                                   const int* __restrict__ keep, int nt, float* adj, double* mis) {
    __shared__ double sh[kThreads];
    __shared__ float shf[kThreads];
    const size_t b = size_t(blockIdx.x) * nt;
    if (!keep[blockIdx.x]) {
        for (int t = threadIdx.x; t < nt; t += blockDim.x) adj[b + t] = 0.0f;
        if (threadIdx.x == 0) mis[blockIdx.x] = 0.0;
        return;
    }
    double oo = 0, ss = 0, so = 0;
    float smax = 0.0f, omax = 0.0f;
    for (int t = threadIdx.x; t < nt; t += blockDim.x) {
        const double o = obs[b + t], s = syn[b + t];
        oo += o * o;
        ss += s * s;
        so += s * o;
        smax = fmaxf(smax, fabsf(syn[b + t]));
        omax = fmaxf(omax, fabsf(obs[b + t]));
    }
    const double OO = block_sum(oo, sh), SS = block_sum(ss, sh), SO = block_sum(so, sh);
    const float SM = block_max(smax, shf), OM = block_max(omax, shf);
    const double no = sqrt(OO), ns = sqrt(SS);
    double e = 0;
    if (SM == 0.0f && OM != 0.0f) {
        for (int t = threadIdx.x; t < nt; t += blockDim.x) {
            const double o = obs[b + t] / no;
            e += o * o;
            adj[b + t] = float(-o);
        }
    } else {
        for (int t = threadIdx.x; t < nt; t += blockDim.x) {
            const double s = syn[b + t], o = obs[b + t];
            const double d = s / ns - o / no;
            e += d * d;
            adj[b + t] = float((SO / (ns * ns) * s - o) / (no * ns));
        }
    }
    const double E = block_sum(e, sh);
    if (threadIdx.x == 0) mis[blockIdx.x] = E;
}

// Real traces -> complex buffer [ntr][nt] (imaginary part zero).
__global__ void k_to_complex(const float* __restrict__ x, Cplx* z, size_t n) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        z[i] = Cplx{x[i], 0.0f};
}

// Spectrum -> analytic-signal spectrum, with the 1/nt of the inverse FFT.
__global__ void k_analytic(Cplx* Z, int nt, size_t n) {
    const float inv = 1.0f / float(nt);
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
        const int f = int(i % size_t(nt));
        float h;
        if (f == 0 || (nt % 2 == 0 && f == nt / 2))
            h = 1.0f;
        else if (f < (nt + 1) / 2)
            h = 2.0f;
        else
            h = 0.0f;
        Z[i].x *= h * inv;
        Z[i].y *= h * inv;
    }
}

// Envelope misfit (compute_adjsrc_envelope): from the analytic signals of
// obs (zo) and syn (zs), the misfit and the two real signals the adjoint
// needs: r1 = syn * c and r2 = H(syn) * c, c = (Es - Eo)/(Es + 0.01 max Es).
__global__ __launch_bounds__(kThreads) void k_envelope(const Cplx* __restrict__ zo, const Cplx* __restrict__ zs,
                                                       const int* __restrict__ keep, int nt, float* r1, float* r2,
                                                       double* mis) {
    __shared__ double sh[kThreads];
    __shared__ float shf[kThreads];
    const size_t b = size_t(blockIdx.x) * nt;
    const bool on = keep[blockIdx.x] != 0;
    double e = 0;
    float emax = 0.0f;
    for (int t = threadIdx.x; t < nt; t += blockDim.x) {
        const Cplx o = zo[b + t], s = zs[b + t];
        const float eo = sqrtf(o.x * o.x + o.y * o.y), es = sqrtf(s.x * s.x + s.y * s.y);
        const float d = eo - es;
        e += double(d) * double(d);
        emax = fmaxf(emax, es);
    }
    const double E = block_sum(e, sh);
    const float EM = block_max(emax, shf);
    for (int t = threadIdx.x; t < nt; t += blockDim.x) {
        const Cplx o = zo[b + t], s = zs[b + t];
        const float eo = sqrtf(o.x * o.x + o.y * o.y), es = sqrtf(s.x * s.x + s.y * s.y);
        const float c = (on && EM != 0.0f) ? (es - eo) / (es + 1.0e-2f * EM) : 0.0f;
        r1[b + t] = s.x * c;
        r2[b + t] = s.y * c;
    }
    if (threadIdx.x == 0) mis[blockIdx.x] = on ? E : 0.0;
}

// Instantaneous-phase misfit (compute_adjsrc_phase): from the analytic
// signals, the misfit sum (phi_d - phi_u)^2 and the real signals
// r1 = -(phi_u - phi_d) H(syn) / env and r2 = (phi_u - phi_d) syn / env,
// env = Es^2 + 0.05 max Es^2.
__global__ __launch_bounds__(kThreads) void k_phase(const Cplx* __restrict__ zo, const Cplx* __restrict__ zs,
                                                    const int* __restrict__ keep, int nt, float* r1, float* r2,
                                                    double* mis) {
    __shared__ double sh[kThreads];
    __shared__ float shf[kThreads];
    const size_t b = size_t(blockIdx.x) * nt;
    const bool on = keep[blockIdx.x] != 0;
    double e = 0;
    float emax = 0.0f;
    for (int t = threadIdx.x; t < nt; t += blockDim.x) {
        const Cplx o = zo[b + t], s = zs[b + t];
        const float d = atan2f(o.y, o.x) - atan2f(s.y, s.x);
        e += double(d) * double(d);
        emax = fmaxf(emax, s.x * s.x + s.y * s.y);
    }
    const double E = block_sum(e, sh);
    const float EM = block_max(emax, shf);
    for (int t = threadIdx.x; t < nt; t += blockDim.x) {
        const Cplx o = zo[b + t], s = zs[b + t];
        const float dphi = atan2f(s.y, s.x) - atan2f(o.y, o.x);
        const float env = (s.x * s.x + s.y * s.y) + 5.0e-2f * EM;
        r1[b + t] = on ? -dphi * s.y / env : 0.0f;
        r2[b + t] = on ? dphi * s.x / env : 0.0f;
    }
    if (threadIdx.x == 0) mis[blockIdx.x] = on ? E : 0.0;
}

// adj = r1 + sign * H(r2), from the analytic signal z2 of r2.
__global__ void k_combine(const float* __restrict__ r1, const Cplx* __restrict__ z2, float sign, float* adj,
                          size_t n) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        adj[i] = r1[i] + sign * z2[i].y;
}

// Adaptive misfit (compute_adjsrc_adaptive).  Per trace: syn and obs
// tapered (ramp over tl samples at each end) and zero-padded to nnt, as
// complex rows U, D.
__global__ void k_awi_load(const float* __restrict__ obs, const float* __restrict__ syn,
                           const float* __restrict__ ramp, Cplx* U, Cplx* D, int nt, int nnt, int tl) {
    const size_t tr = blockIdx.y;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nnt; k += gridDim.x * blockDim.x) {
        float u = 0.0f, d = 0.0f;
        if (k < nt) {
            float r = 1.0f;
            if (k < tl) r *= ramp[k];
            if (k >= nt - tl) r *= ramp[nt - 1 - k];
            u = syn[tr * nt + k] * r;
            d = obs[tr * nt + k] * r;
        }
        U[tr * nnt + k] = Cplx{u, 0.0f};
        D[tr * nnt + k] = Cplx{d, 0.0f};
    }
}

// Spectra -> U = U conj(D) / den / nnt (the matching filter's spectrum with
// the inverse FFT's 1/nnt) and D = D / den / nnt (kept for the adjoint),
// den = |D|^2 + eps max |D|^2.  One block per trace.
__global__ __launch_bounds__(kThreads) void k_awi_filter(Cplx* U, Cplx* D, int nnt, float eps, int* valid) {
    __shared__ float sh[kThreads];
    Cplx* u = U + size_t(blockIdx.x) * nnt;
    Cplx* d = D + size_t(blockIdx.x) * nnt;
    float m = 0.0f;
    for (int f = threadIdx.x; f < nnt; f += blockDim.x) m = fmaxf(m, d[f].x * d[f].x + d[f].y * d[f].y);
    const float e = eps * block_max(m, sh);
    const float inv = 1.0f / float(nnt);
    // an all-zero obs spectrum (e = 0) gives zero rows and no filter
    for (int f = threadIdx.x; f < nnt; f += blockDim.x) {
        const Cplx a = u[f], b = d[f];
        const float den = (b.x * b.x + b.y * b.y) + e;
        u[f] = e > 0.0f ? Cplx{(a.x * b.x + a.y * b.y) / den * inv, (a.y * b.x - a.x * b.y) / den * inv}
                        : Cplx{0.0f, 0.0f};
        d[f] = e > 0.0f ? Cplx{b.x / den * inv, b.y / den * inv} : Cplx{0.0f, 0.0f};
    }
    if (valid && threadIdx.x == 0) valid[blockIdx.x] = e > 0.0f;
}

// Space-time adaptive misfit: trace i's filter is the mean of the valid
// filters of traces i - hw .. i + hw (rows of X), in trace order; keep[i] =
// weight != 0, trace i valid and a non-empty group.
__global__ void k_stawi_group(const Cplx* __restrict__ X, const int* __restrict__ valid,
                              const float* __restrict__ weight, Cplx* G, int* keep, int ntr, int nnt, int hw) {
    const int i = blockIdx.y;
    int l = 0;
    for (int j = max(0, i - hw); j <= min(ntr - 1, i + hw); ++j) l += valid[j];
    const float inv = l > 0 ? 1.0f / float(l) : 0.0f;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nnt; k += gridDim.x * blockDim.x) {
        float s = 0.0f;
        for (int j = max(0, i - hw); j <= min(ntr - 1, i + hw); ++j)
            if (valid[j]) s += X[size_t(j) * nnt + k].x;
        G[size_t(i) * nnt + k] = Cplx{s * inv, 0.0f};
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) keep[i] = weight[i] != 0.0f && valid[i] && l > 0;
}

// From the filter xc (real part of U): misfit 0.5 sum (w xc)^2 / sum xc^2
// (FP64) and U <- (w^2 - 2 misfit) xc / sum xc^2, the misfit's derivative
// by xc.  One block per trace.
__global__ __launch_bounds__(kThreads) void k_awi_misfit(Cplx* U, const float* __restrict__ w,
                                                         const int* __restrict__ keep, int nnt, double* mis) {
    __shared__ double sh[kThreads];
    Cplx* u = U + size_t(blockIdx.x) * nnt;
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < nnt; k += blockDim.x) {
        const double x = u[k].x, wx = double(w[k]) * x;
        a += x * x;
        b += wx * wx;
    }
    const double sxc = block_sum(a, sh), swx = block_sum(b, sh);
    const bool ok = keep[blockIdx.x] && sxc > 0.0;
    const double f = ok ? 0.5 * swx / sxc : 0.0;
    for (int k = threadIdx.x; k < nnt; k += blockDim.x) {
        const double wk = w[k];
        u[k] = Cplx{ok ? float((wk * wk - 2.0 * f) * double(u[k].x) / sxc) : 0.0f, 0.0f};
    }
    if (threadIdx.x == 0) mis[blockIdx.x] = f;
}

// Local adaptive misfit (compute_adjsrc_local_adaptive).  eps and eta of
// each trace from the whole-trace filter: eps = deconv_eps max |D0|^2 is
// taken by k_lawi_eps from the spectra, eta = deconv_eps max xc0^2 by
// k_lawi_eta from the filter (FLT_EPSILON for a zero).
__global__ __launch_bounds__(kThreads) void k_lawi_eps(const Cplx* __restrict__ D, int nnt, float rel, float* e) {
    __shared__ float sh[kThreads];
    const Cplx* d = D + size_t(blockIdx.x) * nnt;
    float m = 0.0f;
    for (int f = threadIdx.x; f < nnt; f += blockDim.x) m = fmaxf(m, d[f].x * d[f].x + d[f].y * d[f].y);
    const float v = rel * block_max(m, sh);
    if (threadIdx.x == 0) e[blockIdx.x] = v != 0.0f ? v : 1.1920929e-7f;
}

__global__ __launch_bounds__(kThreads) void k_lawi_eta(const Cplx* __restrict__ X, int nnt, float rel, float* eta) {
    __shared__ float sh[kThreads];
    const Cplx* x = X + size_t(blockIdx.x) * nnt;
    float m = 0.0f;
    for (int k = threadIdx.x; k < nnt; k += blockDim.x) m = fmaxf(m, x[k].x * x[k].x);
    const float v = rel * block_max(m, sh);
    if (threadIdx.x == 0) eta[blockIdx.x] = v != 0.0f ? v : 1.1920929e-7f;
}

// Windowed copies of the tapered traces (row j = trace * nc + window c,
// window centre i = i0 + c + 1 on the 1-based sample axis); windows past the
// last sample are zero rows.
__global__ void k_lawi_load(const Cplx* __restrict__ U0, const Cplx* __restrict__ D0, Cplx* U, Cplx* D, int nt,
                            int nnt, int nc, int i0, float inv2s2) {
    const int row = blockIdx.y, tr = row / nc, c = row % nc, i = i0 + c + 1;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nnt; k += gridDim.x * blockDim.x) {
        const float dk = float(k + 1 - i);
        const float g = i <= nt ? expf(-dk * dk * inv2s2) : 0.0f;
        U[size_t(row) * nnt + k] = Cplx{U0[size_t(tr) * nnt + k].x * g, 0.0f};
        D[size_t(row) * nnt + k] = Cplx{D0[size_t(tr) * nnt + k].x * g, 0.0f};
    }
}

// As k_awi_filter with the trace's eps.
__global__ __launch_bounds__(kThreads) void k_lawi_filter(Cplx* U, Cplx* D, int nnt, int nc,
                                                          const float* __restrict__ eps) {
    Cplx* u = U + size_t(blockIdx.x) * nnt;
    Cplx* d = D + size_t(blockIdx.x) * nnt;
    const float e = eps[blockIdx.x / nc], inv = 1.0f / float(nnt);
    for (int f = threadIdx.x; f < nnt; f += blockDim.x) {
        const Cplx a = u[f], b = d[f];
        const float den = (b.x * b.x + b.y * b.y) + e;
        u[f] = Cplx{(a.x * b.x + a.y * b.y) / den * inv, (a.y * b.x - a.x * b.y) / den * inv};
        d[f] = Cplx{b.x / den * inv, b.y / den * inv};
    }
}

// One window: v = sum (w xc)^2 / (sum xc^2 + eta), v^2 into vt2[row], and
// U <- v (w^2 - v) xc / (sum xc^2 + eta).
__global__ __launch_bounds__(kThreads) void k_lawi_misfit(Cplx* U, const float* __restrict__ w, int nnt, int nc,
                                                          const float* __restrict__ eta, double* vt2) {
    __shared__ double sh[kThreads];
    Cplx* u = U + size_t(blockIdx.x) * nnt;
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < nnt; k += blockDim.x) {
        const double x = u[k].x, wx = double(w[k]) * x;
        a += x * x;
        b += wx * wx;
    }
    const double sx = block_sum(a, sh) + double(eta[blockIdx.x / nc]), v = block_sum(b, sh) / sx;
    for (int k = threadIdx.x; k < nnt; k += blockDim.x) {
        const double wk = w[k];
        u[k] = Cplx{float(v * (wk * wk - v) * double(u[k].x) / sx), 0.0f};
    }
    if (threadIdx.x == 0) vt2[blockIdx.x] = v * v;
}

// acc[tr][k] += sum over the chunk's windows of a_c(k) g_c(k), in window
// order; mis[tr] += the windows' v^2 (thread 0 of the first block).
__global__ void k_lawi_accum(const Cplx* __restrict__ A, const double* __restrict__ vt2, double* acc, double* mis,
                             int nt, int nnt, int nc, int i0, float inv2s2) {
    const int tr = blockIdx.y;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nt; k += gridDim.x * blockDim.x) {
        double s = acc[size_t(tr) * nt + k];
        for (int c = 0; c < nc && i0 + c < nt; ++c) {
            const float dk = float(k + 1 - (i0 + c + 1));
            s += double(A[(size_t(tr) * nc + c) * nnt + k].x) * double(expf(-dk * dk * inv2s2));
        }
        acc[size_t(tr) * nt + k] = s;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double m = mis[tr];
        for (int c = 0; c < nc && i0 + c < nt; ++c) m += vt2[size_t(tr) * nc + c];
        mis[tr] = m;
    }
}

// adj = acc / sum_i g_i^2, tapered at the end; the misfit of kept traces.
__global__ void k_lawi_out(const double* __restrict__ acc, const double* __restrict__ gsum,
                           const double* __restrict__ mis_acc, const float* __restrict__ ramp,
                           const int* __restrict__ keep, float* adj, double* mis, int nt, int tl) {
    const int tr = blockIdx.y;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nt; k += gridDim.x * blockDim.x) {
        float a = keep[tr] ? float(acc[size_t(tr) * nt + k] / gsum[k]) : 0.0f;
        if (k >= nt - tl) a *= ramp[nt - 1 - k];
        adj[size_t(tr) * nt + k] = a;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) mis[tr] = keep[tr] ? mis_acc[tr] : 0.0;
}

// U <- U . D (D holds D / den / nnt).
__global__ void k_awi_back(Cplx* U, const Cplx* __restrict__ D, size_t n) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
        const Cplx a = U[i], b = D[i];
        U[i] = Cplx{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
    }
}

// adj = the first nt samples of Re(U), tapered at the end.
__global__ void k_awi_out(const Cplx* __restrict__ U, const float* __restrict__ ramp, const int* __restrict__ keep,
                          float* adj, int nt, int nnt, int tl) {
    const size_t tr = blockIdx.y;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nt; k += gridDim.x * blockDim.x) {
        float a = keep[tr] ? U[tr * nnt + k].x : 0.0f;
        if (k >= nt - tl) a *= ramp[nt - 1 - k];
        adj[tr * nt + k] = a;
    }
}

void fft_check(hipfftResult r, const char* what) {
    if (r != HIPFFT_SUCCESS) fail(OWL_ERR_DEVICE, std::string(what) + " failed (hipFFT " + std::to_string(int(r)) + ")");
}

// Batched analytic signals of ntr real traces of nt samples.
struct Analytic {
    hipfftHandle plan{};
    bool have = false;
    Analytic(int nt, int ntr, hipStream_t st) {
        fft_check(hipfftPlan1d(&plan, nt, HIPFFT_C2C, ntr), "hipfftPlan1d");
        have = true;
        fft_check(hipfftSetStream(plan, st), "hipfftSetStream");
    }
    ~Analytic() {
        if (have) (void)hipfftDestroy(plan);
    }
    void run(const float* x, Cplx* z, int nt, size_t n, hipStream_t st) {
        const unsigned g = unsigned(std::min<size_t>((n + 255) / 256, 65535));
        hipLaunchKernelGGL(k_to_complex, dim3(g), dim3(256), 0, st, x, z, n);
        auto* zc = reinterpret_cast<hipfftComplex*>(z);
        fft_check(hipfftExecC2C(plan, zc, zc, HIPFFT_FORWARD), "forward FFT");
        hipLaunchKernelGGL(k_analytic, dim3(g), dim3(256), 0, st, z, nt, n);
        fft_check(hipfftExecC2C(plan, zc, zc, HIPFFT_BACKWARD), "inverse FFT");
    }
};

// Host set-up of the adaptive misfit: padded length, taper, lag penalty.
struct AwiShape {
    int nnt = 0, tl = 0, maxlag = 0;
    float eps = 0.0f;
    std::vector<float> ramp, w;  // rising Hann half-window of tl samples; penalty per padded lag
};

int next_power_235(int n) {
    for (int m = std::max(n, 1);; ++m) {
        int k = m;
        for (int p : {2, 3, 5})
            while (k % p == 0) k /= p;
        if (k == 1) return m;
    }
}

AwiShape awi_shape(const owl_misfit_params* p, int nt, int pad = 2) {
    if (!p) fail(OWL_ERR_ARG, "the adaptive misfit needs its parameters");
    if (!(p->dt > 0) || !(p->tlag_max > 0) || !(p->deconv_eps >= 0))
        fail(OWL_ERR_ARG, "adaptive misfit: dt and tlag_max must be positive, deconv_eps non-negative");
    if (p->penalty < OWL_PENALTY_LINEAR || p->penalty > OWL_PENALTY_EXP) fail(OWL_ERR_ARG, "unknown lag penalty");
    AwiShape a;
    a.maxlag = int(std::ceil(p->tlag_max / p->dt));  // ceiling(tlag_max/dt) in float32
    a.nnt = next_power_235(pad * nt);
    a.tl = int(std::floor(0.025f * float(nt) + 0.5f));
    a.eps = p->deconv_eps;
    if (a.maxlag + 1 > a.nnt) fail(OWL_ERR_ARG, "tlag_max exceeds the padded trace length");
    a.ramp.resize(size_t(a.tl));
    for (int k = 0; k < a.tl; ++k) a.ramp[size_t(k)] = float(0.5 - 0.5 * std::cos(M_PI * k / a.tl));
    // xcorr_penalty over lags -maxlag .. maxlag, laid onto the padded lags
    const int L = 2 * a.maxlag + 1;
    std::vector<double> pen(static_cast<size_t>(L));
    const double dt = p->dt, T = p->tlag_max, pw = std::fabs(double(p->penalty_power));
    const double scaling = 1.0 / std::sqrt(2.0 * std::log(100.0));
    for (int i = -a.maxlag; i <= a.maxlag; ++i) {
        const double x = std::fabs(i * dt / T);
        double v = x;
        if (p->penalty == OWL_PENALTY_POWER) v = std::pow(x, pw);
        if (p->penalty == OWL_PENALTY_GAUSSIAN) v = std::exp(-(i * dt) * (i * dt) / (scaling * T * T));
        if (p->penalty == OWL_PENALTY_EXP) v = std::exp(-std::fabs(i * dt) / (scaling * T));
        pen[size_t(i + a.maxlag)] = v;
    }
    if (p->penalty == OWL_PENALTY_GAUSSIAN || p->penalty == OWL_PENALTY_EXP) {
        const double lo = *std::min_element(pen.begin(), pen.end()), hi = *std::max_element(pen.begin(), pen.end());
        for (double& v : pen) v = 1.0 - std::pow(hi > lo ? (v - lo) / (hi - lo) : 0.0, pw);
    }
    a.w.assign(size_t(a.nnt), 0.0f);
    for (int k = 0; k <= a.maxlag; ++k) a.w[size_t(k)] = float(pen[size_t(k + a.maxlag)]);
    for (int k = 1; k <= a.maxlag; ++k) a.w[size_t(a.nnt - k)] = float(pen[size_t(a.maxlag - k)]);
    return a;
}

// A batched C2C plan of n-point transforms, destroyed with the object.
struct FftPlan {
    hipfftHandle p{};
    FftPlan(int n, int batch, hipStream_t st) {
        fft_check(hipfftPlan1d(&p, n, HIPFFT_C2C, batch), "hipfftPlan1d");
        fft_check(hipfftSetStream(p, st), "hipfftSetStream");
    }
    ~FftPlan() { (void)hipfftDestroy(p); }
    FftPlan(const FftPlan&) = delete;
    FftPlan& operator=(const FftPlan&) = delete;
    void run(Cplx* z, int dir) const {
        auto* zc = reinterpret_cast<hipfftComplex*>(z);
        fft_check(hipfftExecC2C(p, zc, zc, dir), dir == HIPFFT_FORWARD ? "forward FFT" : "inverse FFT");
    }
};

// compute_adjsrc_adaptive_spacetime for every trace of the gather, with the
// driver's weight rule (module_inversion_adjoint_source.f90:1171-1215).
void adaptive_spacetime(const AwiShape& aw, const owl_misfit_params& p, int ntr, int nt, const float* weight,
                        const float* obs, const float* syn, float* adj, double* mis, hipStream_t st) {
    const int nnt = aw.nnt;
    const size_t m = size_t(ntr) * size_t(nnt);
    DevBuf<Cplx> U(m), D(m), G(m);
    DevBuf<float> d_ramp(std::max<size_t>(aw.ramp.size(), 1)), d_w(static_cast<size_t>(nnt)),
        d_wt(static_cast<size_t>(ntr));
    DevBuf<int> valid(static_cast<size_t>(ntr)), keep(static_cast<size_t>(ntr));
    if (!aw.ramp.empty()) d_ramp.upload(aw.ramp.data(), aw.ramp.size(), st);
    d_w.upload(aw.w.data(), aw.w.size(), st);
    std::vector<float> wt(size_t(ntr), 1.0f);
    if (weight) wt.assign(weight, weight + ntr);
    d_wt.upload(wt.data(), wt.size(), st);
    FftPlan plan(nnt, ntr, st);
    const dim3 rows(unsigned(std::min(64, (nnt + 255) / 256)), unsigned(ntr));
    hipLaunchKernelGGL(k_awi_load, rows, dim3(256), 0, st, obs, syn, d_ramp.get(), U.get(), D.get(), nt, nnt, aw.tl);
    plan.run(U.get(), HIPFFT_FORWARD);
    plan.run(D.get(), HIPFFT_FORWARD);
    hipLaunchKernelGGL(k_awi_filter, dim3(ntr), dim3(kThreads), 0, st, U.get(), D.get(), nnt, aw.eps, valid.get());
    plan.run(U.get(), HIPFFT_BACKWARD);
    hipLaunchKernelGGL(k_stawi_group, rows, dim3(256), 0, st, U.get(), valid.get(), d_wt.get(), G.get(), keep.get(),
                       ntr, nnt, p.adaptive_hw);
    hipLaunchKernelGGL(k_awi_misfit, dim3(ntr), dim3(kThreads), 0, st, G.get(), d_w.get(), keep.get(), nnt, mis);
    plan.run(G.get(), HIPFFT_FORWARD);
    const unsigned g = unsigned(std::min<size_t>((m + 255) / 256, 65535));
    hipLaunchKernelGGL(k_awi_back, dim3(g), dim3(256), 0, st, G.get(), D.get(), m);
    plan.run(G.get(), HIPFFT_BACKWARD);
    hipLaunchKernelGGL(k_awi_out, dim3(unsigned(std::min(64, (nt + 255) / 256)), unsigned(ntr)), dim3(256), 0, st,
                       G.get(), d_ramp.get(), keep.get(), adj, nt, nnt, aw.tl);
}

// compute_adjsrc_local_adaptive for every trace: the whole-trace filter for
// eps and eta, then the per-sample windows in chunks of nc windows of every
// trace (about 1 GiB of spectra per chunk).
void local_adaptive(const AwiShape& aw, const owl_misfit_params& p, int ntr, int nt, const float* obs,
                    const float* syn, const int* keep, float* adj, double* mis, hipStream_t st) {
    const int nnt = aw.nnt;
    const size_t whole = size_t(ntr) * size_t(nnt);
    DevBuf<Cplx> U0(whole), D0(whole);
    DevBuf<float> d_ramp(std::max<size_t>(aw.ramp.size(), 1)), d_w(static_cast<size_t>(nnt)),
        d_eps(static_cast<size_t>(ntr)), d_eta(static_cast<size_t>(ntr));
    if (!aw.ramp.empty()) d_ramp.upload(aw.ramp.data(), aw.ramp.size(), st);
    d_w.upload(aw.w.data(), aw.w.size(), st);
    const float sg = p.lawi_sigma / p.dt, inv2s2 = 1.0f / (2.0f * sg * sg);
    {
        // the tapered traces (kept real in U0, D0) and the whole-trace filter
        DevBuf<Cplx> Uf(whole), Df(whole);
        const dim3 rows(unsigned(std::min(64, (nnt + 255) / 256)), unsigned(ntr));
        hipLaunchKernelGGL(k_awi_load, rows, dim3(256), 0, st, obs, syn, d_ramp.get(), U0.get(), D0.get(), nt, nnt,
                           aw.tl);
        OWL_HIP(hipMemcpyAsync(Uf.get(), U0.get(), whole * sizeof(Cplx), hipMemcpyDeviceToDevice, st));
        OWL_HIP(hipMemcpyAsync(Df.get(), D0.get(), whole * sizeof(Cplx), hipMemcpyDeviceToDevice, st));
        FftPlan plan(nnt, ntr, st);
        plan.run(Uf.get(), HIPFFT_FORWARD);
        plan.run(Df.get(), HIPFFT_FORWARD);
        hipLaunchKernelGGL(k_lawi_eps, dim3(ntr), dim3(kThreads), 0, st, Df.get(), nnt, aw.eps, d_eps.get());
        hipLaunchKernelGGL(k_lawi_filter, dim3(ntr), dim3(kThreads), 0, st, Uf.get(), Df.get(), nnt, 1,
                           d_eps.get());
        plan.run(Uf.get(), HIPFFT_BACKWARD);
        hipLaunchKernelGGL(k_lawi_eta, dim3(ntr), dim3(kThreads), 0, st, Uf.get(), nnt, aw.eps, d_eta.get());
    }
    // sum_i g_i(k)^2 over the windows i = 1 .. nt (host, double)
    std::vector<double> gsum(size_t(nt), 0.0);
    for (int k = 0; k < nt; ++k)
        for (int i = 1; i <= nt; ++i) {
            const float dk = float(k + 1 - i);
            const double g = std::exp(double(-dk * dk * inv2s2));
            gsum[size_t(k)] += g * g;
        }
    DevBuf<double> d_gsum(static_cast<size_t>(nt)), acc(static_cast<size_t>(ntr) * nt),
        macc(static_cast<size_t>(ntr));
    d_gsum.upload(gsum.data(), gsum.size(), st);
    acc.zero(st);
    macc.zero(st);
    const size_t budget = size_t(1) << 30;
    const int nc = int(std::max<size_t>(1, std::min<size_t>(size_t(nt), budget / (whole * 2 * sizeof(Cplx)))));
    const size_t m = whole * size_t(nc);
    DevBuf<Cplx> U(m), D(m);
    DevBuf<double> vt2(size_t(ntr) * nc);
    FftPlan plan(nnt, ntr * nc, st);
    const unsigned gx = unsigned(std::min(64, (nnt + 255) / 256)), gb = unsigned(std::min<size_t>((m + 255) / 256, 65535));
    const dim3 trows(unsigned(std::min(64, (nt + 255) / 256)), unsigned(ntr));
    for (int i0 = 0; i0 < nt; i0 += nc) {
        hipLaunchKernelGGL(k_lawi_load, dim3(gx, unsigned(ntr * nc)), dim3(256), 0, st, U0.get(), D0.get(), U.get(),
                           D.get(), nt, nnt, nc, i0, inv2s2);
        plan.run(U.get(), HIPFFT_FORWARD);
        plan.run(D.get(), HIPFFT_FORWARD);
        hipLaunchKernelGGL(k_lawi_filter, dim3(unsigned(ntr * nc)), dim3(kThreads), 0, st, U.get(), D.get(), nnt, nc,
                           d_eps.get());
        plan.run(U.get(), HIPFFT_BACKWARD);
        hipLaunchKernelGGL(k_lawi_misfit, dim3(unsigned(ntr * nc)), dim3(kThreads), 0, st, U.get(), d_w.get(), nnt,
                           nc, d_eta.get(), vt2.get());
        plan.run(U.get(), HIPFFT_FORWARD);
        hipLaunchKernelGGL(k_awi_back, dim3(gb), dim3(256), 0, st, U.get(), D.get(), m);
        plan.run(U.get(), HIPFFT_BACKWARD);
        hipLaunchKernelGGL(k_lawi_accum, trows, dim3(256), 0, st, U.get(), vt2.get(), acc.get(), macc.get(), nt, nnt,
                           nc, i0, inv2s2);
    }
    hipLaunchKernelGGL(k_lawi_out, trows, dim3(256), 0, st, acc.get(), d_gsum.get(), macc.get(), d_ramp.get(), keep,
                       adj, mis, nt, aw.tl);
}

}  // namespace mf
}  // namespace owl

using namespace owl;
using namespace owl::mf;

extern "C" {

owl_status owl_adjoint_source(int type, int ntr, int nt, const float* obs, const float* syn, const float* weight,
                              float threshold, float* adj, double* misfit) {
    return owl_adjoint_source_ex(type, nullptr, ntr, nt, obs, syn, weight, threshold, adj, misfit);
}

owl_status owl_adjoint_source_ex(int type, const owl_misfit_params* prm, int ntr, int nt, const float* obs,
                                 const float* syn, const float* weight, float threshold, float* adj,
                                 double* misfit) {
    return guarded([&] {
        if (ntr < 0 || nt < 1 || (ntr > 0 && (!obs || !syn || !adj || !misfit))) fail(OWL_ERR_ARG, "null argument");
        if (type != OWL_MISFIT_WAVEFORM && type != OWL_MISFIT_CORR && type != OWL_MISFIT_ENVELOPE &&
            type != OWL_MISFIT_PHASE && type != OWL_MISFIT_ADAPTIVE && type != OWL_MISFIT_LOCAL_ADAPTIVE &&
            type != OWL_MISFIT_ADAPTIVE_SPACETIME)
            fail(OWL_ERR_ARG, "unknown misfit type");
        AwiShape aw{};
        if (type == OWL_MISFIT_ADAPTIVE || type == OWL_MISFIT_LOCAL_ADAPTIVE) aw = awi_shape(prm, nt);
        if (type == OWL_MISFIT_ADAPTIVE_SPACETIME) {
            aw = awi_shape(prm, nt, 4);
            if (prm->adaptive_hw < 0) fail(OWL_ERR_ARG, "space-time adaptive misfit: adaptive_hw must be >= 0");
        }
        if (type == OWL_MISFIT_LOCAL_ADAPTIVE && !(prm->lawi_sigma > 0))
            fail(OWL_ERR_ARG, "local adaptive misfit: lawi_sigma must be positive");
        if (ntr == 0) return;
        const size_t n = size_t(ntr) * nt;
        hipStream_t st = nullptr;
        OWL_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { (void)hipStreamDestroy(s); }
        } guard{st};
        DevBuf<float> d_obs(n), d_syn(n), d_adj(n);
        DevBuf<double> d_no(ntr), d_ns(ntr), d_mis(ntr);
        DevBuf<int> d_keep(ntr);
        d_obs.upload(obs, n, st);
        d_syn.upload(syn, n, st);
        hipLaunchKernelGGL(k_norms, dim3(ntr), dim3(kThreads), 0, st, d_obs.get(), d_syn.get(), nt, d_no.get(),
                           d_ns.get());
        // the trace-discard rule of compute_adjoint_source (:1042-1054):
        // zero weight, a norm below threshold x the largest one of the shot,
        // or a zero product of the two norms
        std::vector<double> no(ntr), ns(ntr);
        OWL_HIP(hipMemcpyAsync(no.data(), d_no.get(), ntr * sizeof(double), hipMemcpyDeviceToHost, st));
        OWL_HIP(hipMemcpyAsync(ns.data(), d_ns.get(), ntr * sizeof(double), hipMemcpyDeviceToHost, st));
        OWL_HIP(hipStreamSynchronize(st));
        double mo = 0, ms = 0;
        for (int r = 0; r < ntr; ++r) {
            mo = std::max(mo, no[r]);
            ms = std::max(ms, ns[r]);
        }
        std::vector<int> keep(ntr);
        for (int r = 0; r < ntr; ++r) {
            const bool w0 = weight && weight[r] == 0.0f;
            keep[r] = !(w0 || no[r] < double(threshold) * mo || ns[r] < double(threshold) * ms || no[r] * ns[r] == 0.0);
        }
        d_keep.upload(keep.data(), size_t(ntr), st);
        if (type == OWL_MISFIT_WAVEFORM) {
            hipLaunchKernelGGL(k_waveform, dim3(ntr), dim3(kThreads), 0, st, d_obs.get(), d_syn.get(), d_keep.get(),
                               nt, d_adj.get(), d_mis.get());
        } else if (type == OWL_MISFIT_CORR) {
            hipLaunchKernelGGL(k_corr, dim3(ntr), dim3(kThreads), 0, st, d_obs.get(), d_syn.get(), d_keep.get(), nt,
                               d_adj.get(), d_mis.get());
        } else if (type == OWL_MISFIT_ADAPTIVE) {
            const size_t m = size_t(ntr) * size_t(aw.nnt);
            DevBuf<Cplx> U(m), D(m);
            DevBuf<float> d_ramp(std::max<size_t>(aw.ramp.size(), 1)), d_w(size_t(aw.nnt));
            if (!aw.ramp.empty()) d_ramp.upload(aw.ramp.data(), aw.ramp.size(), st);
            d_w.upload(aw.w.data(), aw.w.size(), st);
            hipfftHandle plan{};
            fft_check(hipfftPlan1d(&plan, aw.nnt, HIPFFT_C2C, ntr), "hipfftPlan1d");
            struct PlanGuard {
                hipfftHandle p;
                ~PlanGuard() { (void)hipfftDestroy(p); }
            } pg{plan};
            fft_check(hipfftSetStream(plan, st), "hipfftSetStream");
            auto* uc = reinterpret_cast<hipfftComplex*>(U.get());
            auto* dc = reinterpret_cast<hipfftComplex*>(D.get());
            const dim3 rows(unsigned(std::min(64, (aw.nnt + 255) / 256)), unsigned(ntr));
            hipLaunchKernelGGL(k_awi_load, rows, dim3(256), 0, st, d_obs.get(), d_syn.get(), d_ramp.get(), U.get(),
                               D.get(), nt, aw.nnt, aw.tl);
            fft_check(hipfftExecC2C(plan, uc, uc, HIPFFT_FORWARD), "forward FFT");
            fft_check(hipfftExecC2C(plan, dc, dc, HIPFFT_FORWARD), "forward FFT");
            hipLaunchKernelGGL(k_awi_filter, dim3(ntr), dim3(kThreads), 0, st, U.get(), D.get(), aw.nnt, aw.eps,
                               static_cast<int*>(nullptr));
            fft_check(hipfftExecC2C(plan, uc, uc, HIPFFT_BACKWARD), "inverse FFT");
            hipLaunchKernelGGL(k_awi_misfit, dim3(ntr), dim3(kThreads), 0, st, U.get(), d_w.get(), d_keep.get(),
                               aw.nnt, d_mis.get());
            fft_check(hipfftExecC2C(plan, uc, uc, HIPFFT_FORWARD), "forward FFT");
            const unsigned g = unsigned(std::min<size_t>((m + 255) / 256, 65535));
            hipLaunchKernelGGL(k_awi_back, dim3(g), dim3(256), 0, st, U.get(), D.get(), m);
            fft_check(hipfftExecC2C(plan, uc, uc, HIPFFT_BACKWARD), "inverse FFT");
            hipLaunchKernelGGL(k_awi_out, dim3(unsigned(std::min(64, (nt + 255) / 256)), unsigned(ntr)), dim3(256), 0,
                               st, U.get(), d_ramp.get(), d_keep.get(), d_adj.get(), nt, aw.nnt, aw.tl);
        } else if (type == OWL_MISFIT_ADAPTIVE_SPACETIME) {
            adaptive_spacetime(aw, *prm, ntr, nt, weight, d_obs.get(), d_syn.get(), d_adj.get(), d_mis.get(), st);
        } else if (type == OWL_MISFIT_LOCAL_ADAPTIVE) {
            local_adaptive(aw, *prm, ntr, nt, d_obs.get(), d_syn.get(), d_keep.get(), d_adj.get(), d_mis.get(), st);
        } else {
            DevBuf<Cplx> zo(n), zs(n), z2(n);
            DevBuf<float> r1(n), r2(n);
            Analytic A(nt, ntr, st);
            A.run(d_obs.get(), zo.get(), nt, n, st);
            A.run(d_syn.get(), zs.get(), nt, n, st);
            if (type == OWL_MISFIT_ENVELOPE)
                hipLaunchKernelGGL(k_envelope, dim3(ntr), dim3(kThreads), 0, st, zo.get(), zs.get(), d_keep.get(),
                                   nt, r1.get(), r2.get(), d_mis.get());
            else
                hipLaunchKernelGGL(k_phase, dim3(ntr), dim3(kThreads), 0, st, zo.get(), zs.get(), d_keep.get(), nt,
                                   r1.get(), r2.get(), d_mis.get());
            A.run(r2.get(), z2.get(), nt, n, st);
            const unsigned g = unsigned(std::min<size_t>((n + 255) / 256, 65535));
            // envelope: syn c - H(H(syn) c); phase: -dphi H(syn)/env + H(dphi syn/env)
            hipLaunchKernelGGL(k_combine, dim3(g), dim3(256), 0, st, r1.get(), z2.get(),
                               type == OWL_MISFIT_ENVELOPE ? -1.0f : 1.0f, d_adj.get(), n);
        }
        OWL_HIP(hipGetLastError());
        OWL_HIP(hipMemcpyAsync(adj, d_adj.get(), n * sizeof(float), hipMemcpyDeviceToHost, st));
        OWL_HIP(hipMemcpyAsync(misfit, d_mis.get(), ntr * sizeof(double), hipMemcpyDeviceToHost, st));
        OWL_HIP(hipStreamSynchronize(st));
    });
}

}  // extern "C"
