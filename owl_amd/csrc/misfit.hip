// Misfits and adjoint sources of OWL's FWI beyond the L2 waveform misfit,
// batched over the traces of a shot on the GPU
// (src/inverse/module_inversion_adjoint_source.f90: compute_adjsrc_waveform
// :371-383, compute_adjsrc_corr :391-413, compute_adjsrc_envelope :424-452,
// compute_adjsrc_phase :460-481; the per-shot driver and its trace-discard
// rule compute_adjoint_source :985-1160).
//
// Layout: traces receiver-major [ntr][nt], as the SU traces hold them.  One
// block per trace for the reductions (FP64, Neumaier-compensated like the
// library's L2 misfit); the Hilbert transforms are batched single-precision
// FFTs (hipFFT) of all traces at once: analytic signal z = x + i H(x) =
// IFFT(h . FFT(x)), h = 1 at DC (and Nyquist for even nt), 2 on the positive
// frequencies, 0 on the negative ones.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

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

}  // namespace mf
}  // namespace owl

using namespace owl;
using namespace owl::mf;

extern "C" {

owl_status owl_adjoint_source(int type, int ntr, int nt, const float* obs, const float* syn, const float* weight,
                              float threshold, float* adj, double* misfit) {
    return guarded([&] {
        if (ntr < 0 || nt < 1 || (ntr > 0 && (!obs || !syn || !adj || !misfit))) fail(OWL_ERR_ARG, "null argument");
        if (type != OWL_MISFIT_WAVEFORM && type != OWL_MISFIT_CORR && type != OWL_MISFIT_ENVELOPE &&
            type != OWL_MISFIT_PHASE)
            fail(OWL_ERR_ARG, "unknown misfit type");
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
