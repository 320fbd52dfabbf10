// 3D isotropic elastic SSG-FD velocity-stress time stepping for MI355X
// (gfx950): OWL's elastic_vhtiort_3d forward solver for anisotropy_type =
// 'iso' (module_wavefield.f90:52-1090, module_cfspml.f90:428-625,
// submodule_forward_modeling.f90:31-140) for a batch of shots, each on its
// own window, in every launch.  Force sources, three-component velocity
// receivers, CFS-MPML boundaries (zero damping ratios for isotropic media)
// and an optional flat free surface on a uniform mesh.
//
// Layout: per shot a padded grid, x fastest, then y, then z; rows padded to
// a multiple of 64 floats.  Each sweep tiles x-y with 64 x cy thread blocks
// and marches along z with per-thread register queues for the z stencils.
// FP32 in the reference's operation order (no FMA contraction).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "common.hpp"
#include "host_prep.hpp"
#include "owl_amd.h"
#include "profiler.hpp"

namespace owl {
namespace e3 {

constexpr int kPitchAlign = 64;
constexpr int kCY = 4;  // thread rows per block

struct Coef {
    float c[8];
};

struct ShotE {
    int nx, ny, nz;
    int px;                // row pitch (floats)
    int nyp, nzp;          // padded rows per plane, padded planes
    long long pxy;         // plane pitch
    long long off;         // element offset of node (lo, lo, lo)
    int tile0, ntx, nty, nzc;
    int prof;              // M-PML profiles: x (di dh ai ah), y, z
    int nr, rec0;
    int ninj, inj0;
    long long trace0;      // three [nt][nr] blocks (vx, vy, vz)
};

// Wavefields and memory variables (named by derivative and field as in OWL).
struct EF {
    float *sxx, *syy, *szz, *syz, *sxz, *sxy, *vx, *vy, *vz;
    float *m_dxvx, *m_dyvy, *m_dzvz, *m_dyvz, *m_dzvy, *m_dxvz, *m_dzvx, *m_dxvy, *m_dyvx;
    float *m_dxxx, *m_dyxy, *m_dzxz, *m_dxxy, *m_dyyy, *m_dzyz, *m_dxxz, *m_dyyz, *m_dzzz;
};

// Medium: c11 (= c22 = c33), c12 (= c13 = c23), the harmonic 4-point shear
// moduli at the three shear-stress positions, and the 2-point densities at
// the three velocity positions.
struct EM {
    const float *c11, *c12, *mu_yz, *mu_xz, *mu_xy, *rx, *ry, *rz;
};

struct Tap {
    long long cell;
    float w1, w2, w3;
    int row;        // amplitude row (source) or receiver index
    int tbeg, tend;
    int field;      // 0 vx, 1 vy, 2 vz
};

__device__ __forceinline__ size_t ate(const ShotE& S, int lo, int i, int j, int k) {
    return size_t(S.off) + size_t(i - lo) + size_t(j - lo) * size_t(S.px) + size_t(k - lo) * size_t(S.pxy);
}

// damp_coef + damp_wavefield (module_wavefield.f90:40-47) with kappa = 1.
// A node whose own-axis profile is zero keeps a zero memory variable and an
// unchanged derivative, so the update is skipped there.
__device__ __forceinline__ float damp(float* w, size_t q, float v, float d, float alpha, float adt) {
    if (d == 0.0f) return v;
    const float a_ = d / 1.0f;
    const float b_ = d / 1.0f + alpha;
    const float ca = -(2.0f * adt * a_ / (2.0f + adt * b_));
    const float cb = (2.0f - adt * b_) / (2.0f + adt * b_);
    const float m = ca * v + cb * w[q];
    w[q] = m;
    return (v + m) / 1.0f;
}

struct Prof {
    const float *dix, *dhx, *aix, *ahx, *diy, *dhy, *aiy, *ahy, *diz, *dhz, *aiz, *ahz;
    int pml;
};

__device__ __forceinline__ Prof prof_of(const float* prof, const ShotE& S, int pml) {
    const int lx = S.nx + 2 * pml, ly = S.ny + 2 * pml, lz = S.nz + 2 * pml;
    const float* p = prof + S.prof;
    Prof P;
    P.dix = p; P.dhx = p + lx; P.aix = p + 2 * lx; P.ahx = p + 3 * lx;
    p += 4 * lx;
    P.diy = p; P.dhy = p + ly; P.aiy = p + 2 * ly; P.ahy = p + 3 * ly;
    p += 4 * ly;
    P.diz = p; P.dhz = p + lz; P.aiz = p + 2 * lz; P.ahz = p + 3 * lz;
    P.pml = pml;
    return P;
}

// ---------------------------------------------------------------------------
// Stress sweep: normal stresses at integer nodes, syz at (i, Y, Z), sxz at
// (X, j, Z), sxy at (X, Y, k), all stored at the same index (i, j, k).
// module_wavefield.f90:58-306 (free surface :592-821).  The z stencils come
// from register queues over planes k-R .. k+R of vx, vy, vz.
// ---------------------------------------------------------------------------
template <int R>
__global__ __launch_bounds__(256) void ke_stress(const ShotE* __restrict__ shots, const int* __restrict__ tile_shot,
                                                 EF F, EM M, const float* __restrict__ prof, Coef c, float dt,
                                                 float idx, float idy, float idz, int pml, int kc, int fs) {
    const int tile = blockIdx.x;
    const ShotE S = shots[tile_shot[tile]];
    int lt = tile - S.tile0;
    const int tx = lt % S.ntx;
    lt /= S.ntx;
    const int ty = lt % S.nty, tzc = lt / S.nty;
    const int nx = S.nx, ny = S.ny, nz = S.nz;
    const int lo = 1 - pml - R;
    const int i = -pml + 1 + tx * 64 + int(threadIdx.x);
    const int j = -pml + 1 + ty * int(blockDim.y) + int(threadIdx.y);
    if (i > nx + pml || j > ny + pml) return;
    const int kt = -pml + 1 + tzc * kc, k1 = min(kt + kc - 1, nz + pml);
    const int k0 = fs ? max(kt, 1) : kt;
    if (k0 > k1) return;
    const Prof P = prof_of(prof, S, pml);
    const float adt = fabsf(dt);
    const float* __restrict__ vx = F.vx;
    const float* __restrict__ vy = F.vy;
    const float* __restrict__ vz = F.vz;
    const size_t sx = 1, sy = S.px, sz = S.pxy;
    const int ii = i + pml - 1, jj = j + pml - 1;
    size_t q = ate(S, lo, i, j, k0);
    // queues: plane k + m - R at slot m, m = 0 .. 2R
    float qx[2 * R + 1], qy[2 * R + 1], qz[2 * R + 1];
#pragma unroll
    for (int m = 0; m < 2 * R; ++m) {
        qx[m] = vx[q + (m - R) * sz];
        qy[m] = vy[q + (m - R) * sz];
        qz[m] = vz[q + (m - R) * sz];
    }
    const bool xin = !(i <= 0 || i >= nx + 1), yin = !(j <= 0 || j >= ny + 1);
    const bool Xin = !(i <= 1 || i >= nx + 1), Yin = !(j <= 1 || j >= ny + 1);
    for (int k = k0; k <= k1; ++k, q += sz) {
        qx[2 * R] = vx[q + R * sz];
        qy[2 * R] = vy[q + R * sz];
        qz[2 * R] = vz[q + R * sz];
        const int kk = k + pml - 1;
        // ---- normal stresses
        {
            float a = c.c[0] * (vx[q + sx] - vx[q]);
            float b = c.c[0] * (vy[q + sy] - vy[q]);
#pragma unroll
            for (int m = 2; m <= R; ++m) {
                a = a + c.c[m - 1] * (vx[q + m * sx] - vx[q - (m - 1) * sx]);
                b = b + c.c[m - 1] * (vy[q + m * sy] - vy[q - (m - 1) * sy]);
            }
            float pdxvx = idx * a, pdyvy = idy * b, pdzvz;
            const float c11 = M.c11[q], c12 = M.c12[q];
            if (fs && k == 1) {
                if (!xin || !yin) {
                    pdxvx = damp(F.m_dxvx, q, pdxvx, P.dix[ii], P.aix[ii], adt);
                    pdyvy = damp(F.m_dyvy, q, pdyvy, P.diy[jj], P.aiy[jj], adt);
                }
                pdzvz = -(c12 / c11 * pdxvx) - c12 / c11 * pdyvy;
            } else {
                float s = c.c[0] * (qz[R + 1] - qz[R]);
#pragma unroll
                for (int m = 2; m <= R; ++m) s = s + c.c[m - 1] * (qz[R + m] - qz[R - m + 1]);
                pdzvz = idz * s;
                if (!xin || !yin || (!fs && k <= 0) || k >= nz + 1) {
                    pdxvx = damp(F.m_dxvx, q, pdxvx, P.dix[ii], P.aix[ii], adt);
                    pdyvy = damp(F.m_dyvy, q, pdyvy, P.diy[jj], P.aiy[jj], adt);
                    pdzvz = damp(F.m_dzvz, q, pdzvz, P.diz[kk], P.aiz[kk], adt);
                }
            }
            F.sxx[q] = F.sxx[q] + dt * (c11 * pdxvx + c12 * pdyvy + c12 * pdzvz);
            F.syy[q] = F.syy[q] + dt * (c12 * pdxvx + c11 * pdyvy + c12 * pdzvz);
            F.szz[q] = F.szz[q] + dt * (c12 * pdxvx + c12 * pdyvy + c11 * pdzvz);
        }
        // ---- shear stresses with a z derivative: syz, sxz (Z = k)
        if (!fs || k >= 2) {
            float a = c.c[0] * (qy[R] - qy[R - 1]);   // d-z vy
            float b = c.c[0] * (vz[q] - vz[q - sy]);  // d-y vz
            float e = c.c[0] * (qx[R] - qx[R - 1]);   // d-z vx
            float f = c.c[0] * (vz[q] - vz[q - sx]);  // d-x vz
#pragma unroll
            for (int m = 2; m <= R; ++m) {
                a = a + c.c[m - 1] * (qy[R + m - 1] - qy[R - m]);
                b = b + c.c[m - 1] * (vz[q + (m - 1) * sy] - vz[q - m * sy]);
                e = e + c.c[m - 1] * (qx[R + m - 1] - qx[R - m]);
                f = f + c.c[m - 1] * (vz[q + (m - 1) * sx] - vz[q - m * sx]);
            }
            float pdzvy = idz * a, pdyvz = idy * b, pdzvx = idz * e, pdxvz = idx * f;
            const bool zpml = (!fs && k <= 1) || k >= nz + 1;
            if (!xin || !Yin || zpml) {
                pdyvz = damp(F.m_dyvz, q, pdyvz, P.dhy[jj], P.ahy[jj], adt);
                pdzvy = damp(F.m_dzvy, q, pdzvy, P.dhz[kk], P.ahz[kk], adt);
            }
            F.syz[q] = F.syz[q] + dt * M.mu_yz[q] * (pdzvy + pdyvz);
            if (!Xin || !yin || zpml) {
                pdxvz = damp(F.m_dxvz, q, pdxvz, P.dhx[ii], P.ahx[ii], adt);
                pdzvx = damp(F.m_dzvx, q, pdzvx, P.dhz[kk], P.ahz[kk], adt);
            }
            F.sxz[q] = F.sxz[q] + dt * M.mu_xz[q] * (pdzvx + pdxvz);
        }
        // ---- sxy (X, Y, k)
        {
            float a = c.c[0] * (vx[q] - vx[q - sy]);  // d-y vx
            float b = c.c[0] * (vy[q] - vy[q - sx]);  // d-x vy
#pragma unroll
            for (int m = 2; m <= R; ++m) {
                a = a + c.c[m - 1] * (vx[q + (m - 1) * sy] - vx[q - m * sy]);
                b = b + c.c[m - 1] * (vy[q + (m - 1) * sx] - vy[q - m * sx]);
            }
            float pdyvx = idy * a, pdxvy = idx * b;
            if (!Xin || !Yin || (!fs && (k <= 0 || k >= nz + 1))) {
                pdxvy = damp(F.m_dxvy, q, pdxvy, P.dhx[ii], P.ahx[ii], adt);
                pdyvx = damp(F.m_dyvx, q, pdyvx, P.dhy[jj], P.ahy[jj], adt);
            }
            F.sxy[q] = F.sxy[q] + dt * M.mu_xy[q] * (pdyvx + pdxvy);
        }
#pragma unroll
        for (int m = 0; m < 2 * R; ++m) {
            qx[m] = qx[m + 1];
            qy[m] = qy[m + 1];
            qz[m] = qz[m + 1];
        }
    }
}

// ---------------------------------------------------------------------------
// Velocity sweep: vx at (X, j, k), vy at (i, Y, k), vz at (i, j, Z).
// module_wavefield.f90:316-503 (free surface :852-1090).  z stencils from
// queues over planes k-R .. k+R of sxz, syz, szz.
// ---------------------------------------------------------------------------
template <int R>
__global__ __launch_bounds__(256) void ke_velocity(const ShotE* __restrict__ shots,
                                                   const int* __restrict__ tile_shot, EF F, EM M,
                                                   const float* __restrict__ prof, Coef c, float dt, float idx,
                                                   float idy, float idz, int pml, int kc, int fs) {
    const int tile = blockIdx.x;
    const ShotE S = shots[tile_shot[tile]];
    int lt = tile - S.tile0;
    const int tx = lt % S.ntx;
    lt /= S.ntx;
    const int ty = lt % S.nty, tzc = lt / S.nty;
    const int nx = S.nx, ny = S.ny, nz = S.nz;
    const int lo = 1 - pml - R;
    const int i = -pml + 1 + tx * 64 + int(threadIdx.x);
    const int j = -pml + 1 + ty * int(blockDim.y) + int(threadIdx.y);
    if (i > nx + pml || j > ny + pml) return;
    const int kt = -pml + 1 + tzc * kc, k1 = min(kt + kc - 1, nz + pml);
    const int k0 = fs ? max(kt, 1) : kt;
    if (k0 > k1) return;
    const Prof P = prof_of(prof, S, pml);
    const float adt = fabsf(dt);
    const float* __restrict__ sxx = F.sxx;
    const float* __restrict__ syy = F.syy;
    const float* __restrict__ szz = F.szz;
    const float* __restrict__ syz = F.syz;
    const float* __restrict__ sxz = F.sxz;
    const float* __restrict__ sxy = F.sxy;
    const size_t sx = 1, sy = S.px, sz = S.pxy;
    const int ii = i + pml - 1, jj = j + pml - 1;
    size_t q = ate(S, lo, i, j, k0);
    float qa[2 * R + 1], qb[2 * R + 1], qc[2 * R + 1];  // sxz, syz, szz
#pragma unroll
    for (int m = 0; m < 2 * R; ++m) {
        qa[m] = sxz[q + (m - R) * sz];
        qb[m] = syz[q + (m - R) * sz];
        qc[m] = szz[q + (m - R) * sz];
    }
    const bool xin = !(i <= 0 || i >= nx + 1), yin = !(j <= 0 || j >= ny + 1);
    const bool Xin = !(i <= 1 || i >= nx + 1), Yin = !(j <= 1 || j >= ny + 1);
    for (int k = k0; k <= k1; ++k, q += sz) {
        qa[2 * R] = sxz[q + R * sz];
        qb[2 * R] = syz[q + R * sz];
        qc[2 * R] = szz[q + R * sz];
        const int kk = k + pml - 1;
        const bool zin = !((!fs && k <= 0) || k >= nz + 1);
        // ---- vx (X, j, k)
        {
            float a = c.c[0] * (sxx[q] - sxx[q - sx]);       // d-x sxx
            float b = c.c[0] * (sxy[q + sy] - sxy[q]);       // d+y sxy
            float e = c.c[0] * (qa[R + 1] - qa[R]);          // d+z sxz
#pragma unroll
            for (int m = 2; m <= R; ++m) {
                a = a + c.c[m - 1] * (sxx[q + (m - 1) * sx] - sxx[q - m * sx]);
                b = b + c.c[m - 1] * (sxy[q + m * sy] - sxy[q - (m - 1) * sy]);
                e = e + c.c[m - 1] * (qa[R + m] - qa[R - m + 1]);
            }
            float pdxxx = idx * a, pdyxy = idy * b, pdzxz = idz * e;
            if (!Xin || !yin || !zin) {
                pdxxx = damp(F.m_dxxx, q, pdxxx, P.dhx[ii], P.ahx[ii], adt);
                pdyxy = damp(F.m_dyxy, q, pdyxy, P.diy[jj], P.aiy[jj], adt);
                pdzxz = damp(F.m_dzxz, q, pdzxz, P.diz[kk], P.aiz[kk], adt);
            }
            F.vx[q] = F.vx[q] + dt / M.rx[q] * (pdxxx + pdyxy + pdzxz);
        }
        // ---- vy (i, Y, k)
        {
            float a = c.c[0] * (sxy[q + sx] - sxy[q]);       // d+x sxy
            float b = c.c[0] * (syy[q] - syy[q - sy]);       // d-y syy
            float e = c.c[0] * (qb[R + 1] - qb[R]);          // d+z syz
#pragma unroll
            for (int m = 2; m <= R; ++m) {
                a = a + c.c[m - 1] * (sxy[q + m * sx] - sxy[q - (m - 1) * sx]);
                b = b + c.c[m - 1] * (syy[q + (m - 1) * sy] - syy[q - m * sy]);
                e = e + c.c[m - 1] * (qb[R + m] - qb[R - m + 1]);
            }
            float pdxxy = idx * a, pdyyy = idy * b, pdzyz = idz * e;
            if (!xin || !Yin || !zin) {
                pdxxy = damp(F.m_dxxy, q, pdxxy, P.dix[ii], P.aix[ii], adt);
                pdyyy = damp(F.m_dyyy, q, pdyyy, P.dhy[jj], P.ahy[jj], adt);
                pdzyz = damp(F.m_dzyz, q, pdzyz, P.diz[kk], P.aiz[kk], adt);
            }
            F.vy[q] = F.vy[q] + dt / M.ry[q] * (pdxxy + pdyyy + pdzyz);
        }
        // ---- vz (i, j, Z = k)
        if (!fs || k >= 2) {
            float a = c.c[0] * (sxz[q + sx] - sxz[q]);       // d+x sxz
            float b = c.c[0] * (syz[q + sy] - syz[q]);       // d+y syz
            float e = c.c[0] * (qc[R] - qc[R - 1]);          // d-z szz
#pragma unroll
            for (int m = 2; m <= R; ++m) {
                a = a + c.c[m - 1] * (sxz[q + m * sx] - sxz[q - (m - 1) * sx]);
                b = b + c.c[m - 1] * (syz[q + m * sy] - syz[q - (m - 1) * sy]);
                e = e + c.c[m - 1] * (qc[R + m - 1] - qc[R - m]);
            }
            float pdxxz = idx * a, pdyyz = idy * b, pdzzz = idz * e;
            const bool Zin = !((!fs && k <= 1) || k >= nz + 1);
            if (!xin || !yin || !Zin) {
                pdxxz = damp(F.m_dxxz, q, pdxxz, P.dix[ii], P.aix[ii], adt);
                pdyyz = damp(F.m_dyyz, q, pdyyz, P.diy[jj], P.aiy[jj], adt);
                pdzzz = damp(F.m_dzzz, q, pdzzz, P.dhz[kk], P.ahz[kk], adt);
            }
            F.vz[q] = F.vz[q] + dt / M.rz[q] * (pdxxz + pdyyz + pdzzz);
        }
#pragma unroll
        for (int m = 0; m < 2 * R; ++m) {
            qa[m] = qa[m + 1];
            qb[m] = qb[m + 1];
            qc[m] = qc[m + 1];
        }
    }
}

// Free surface, whole padded x-y planes (module_wavefield.f90:580-590,
// 823-845; submodule_forward_modeling.f90:64-73).  mode 0: zero the
// velocities above the surface; 1: mirror the stresses; 2: mirror the
// velocities for the receivers (NKW planes).
template <int R>
__global__ void ke_surface(const ShotE* __restrict__ shots, EF F, int pml, int mode) {
    const ShotE S = shots[blockIdx.y];
    const int lo = 1 - pml - R;
    const long long plane = (long long)S.px * S.nyp;
    const int nk = mode == 2 ? kNkw : R;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < plane; e += (long long)gridDim.x * blockDim.x) {
        const size_t b = size_t(S.off) + size_t(e);
        auto at = [&](int k) { return b + size_t(k - lo) * size_t(S.pxy); };
        if (mode == 0) {
            for (int m = 1; m <= R; ++m) {
                F.vx[at(1 - m)] = 0.0f;
                F.vy[at(1 - m)] = 0.0f;
                F.vz[at(2 - m)] = 0.0f;
            }
        } else if (mode == 1) {
            for (int m = 1; m <= R; ++m) {
                F.szz[at(1 - m)] = -F.szz[at(1 + m)];
                F.szz[at(1)] = 0.0f;
                F.sxx[at(1 - m)] = -F.sxx[at(1 + m)];
                F.syy[at(1 - m)] = -F.syy[at(1 + m)];
                F.sxy[at(1 - m)] = -F.sxy[at(1 + m)];
                F.syz[at(2 - m)] = -F.syz[at(1 + m)];
                F.sxz[at(2 - m)] = -F.sxz[at(1 + m)];
            }
        } else {
            for (int m = 1; m <= nk; ++m) {
                F.vx[at(1 - m)] = F.vx[at(1 + m)];
                F.vy[at(1 - m)] = F.vy[at(1 + m)];
                F.vz[at(2 - m)] = F.vz[at(1 + m)];
            }
        }
    }
}

// Force injection of step t (add_source, module_wavefield.f90:1115-1175);
// one block per shot, taps grouped by (field, cell) in source order.
__global__ __launch_bounds__(1024) void ke_inject(const ShotE* __restrict__ shots, EF F,
                                                  const int* __restrict__ inj_beg, const Tap* __restrict__ inj,
                                                  const float* __restrict__ amp_tab, int nt, int t) {
    const ShotE S = shots[blockIdx.x];
    for (int c = threadIdx.x; c < S.ninj; c += blockDim.x) {
        const int b = inj_beg[S.inj0 + c], e = inj_beg[S.inj0 + c + 1];
        const long long cell = inj[b].cell;
        const int fld = inj[b].field;
        float* f = fld == 0 ? F.vx : (fld == 1 ? F.vy : F.vz);
        float acc = f[cell];
        for (int k = b; k < e; ++k) {
            const Tap tp = inj[k];
            if (t < tp.tbeg || t > tp.tend) continue;
            acc = acc + amp_tab[size_t(tp.row) * nt + (t - 1)] * tp.w1 * tp.w2 * tp.w3;
        }
        f[cell] = acc;
    }
}

// Receivers of step t: vx at (hx, gy, gz), vy at (gx, hy, gz), vz at
// (gx, gy, hz) (submodule_forward_modeling.f90:75-140).
__global__ __launch_bounds__(256) void ke_record(const ShotE* __restrict__ shots, EF F,
                                                 const int* __restrict__ rec_beg, const Tap* __restrict__ rec,
                                                 const float* __restrict__ rec_w, float* traces, int nt, int t) {
    const ShotE S = shots[blockIdx.y];
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < 3 * S.nr; r += gridDim.x * blockDim.x) {
        const int comp = r / S.nr, ir = r % S.nr;
        const int g = 3 * S.rec0 + comp * S.nr + ir;
        const float w = rec_w[S.rec0 + ir];
        const float* f = comp == 0 ? F.vx : (comp == 1 ? F.vy : F.vz);
        float acc = 0.0f;
        for (int k = rec_beg[g]; k < rec_beg[g + 1]; ++k) {
            const Tap tp = rec[k];
            acc = acc + f[tp.cell] * tp.w1 * tp.w2 * tp.w3 * w;
        }
        traces[S.trace0 + size_t(comp) * nt * S.nr + size_t(t - 1) * S.nr + ir] = acc;
    }
}

struct ShotHostE {
    int id = 0, nx = 0, ny = 0, nz = 0, nxbeg = 1, nybeg = 1, nzbeg = 1;
    float ox = 0, oy = 0, oz = 0;
    std::vector<owl_source> src;
    std::vector<owl_receiver> rec;
};

}  // namespace e3
}  // namespace owl

using namespace owl;
using namespace owl::e3;

struct owl_elastic3d {
    owl_acoustic_config cfg{};
    int R = 4, pml = 15, nt = 0, kc = 64, cy = kCY, fs = 0;
    float dt = 0, idx = 0, idy = 0, idz = 0;
    e3::Coef coef{};
    std::vector<ShotHostE> hs;
    std::vector<ShotE> sd;
    std::vector<int> tile_shot;
    size_t field_elems = 0, trace_elems = 0;
    int nrec_total = 0, max_nr = 0;
    double pts_full = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0, last_updates = 0;
    Profiler prof;
    std::vector<float> vp_full, vs_full, rho_full;
    DevBuf<ShotE> d_shots;
    DevBuf<int> d_tile_shot, d_inj_beg, d_rec_beg;
    DevBuf<float> d_prof, d_amp, d_rec_w, d_traces;
    DevBuf<float> med[8];
    DevBuf<float> fld[27];
    DevBuf<Tap> d_inj, d_rec;
    ~owl_elastic3d() {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

inline size_t cell_e(const owl_elastic3d* b, const ShotE& S, int i, int j, int k) {
    const int lo = 1 - b->pml - b->R, h = b->pml + b->R;
    if (i < lo || i > S.nx + h || j < lo || j > S.ny + h || k < lo || k > S.nz + h)
        fail(OWL_ERR_ARG, "a source/receiver stencil reaches outside the padded grid; increase pml");
    return size_t(S.off) + size_t(i - lo) + size_t(j - lo) * size_t(S.px) + size_t(k - lo) * size_t(S.pxy);
}

inline size_t model_index(const owl_elastic3d* b, const ShotHostE& H, int i, int j, int k) {
    const int ic = std::min(std::max(i, 1), H.nx), jc = std::min(std::max(j, 1), H.ny),
              kc = std::min(std::max(k, 1), H.nz);
    const size_t nz = size_t(b->cfg.nz), ny = size_t(b->cfg.ny);
    return size_t(H.nzbeg - 1 + kc - 1) + size_t(H.nybeg - 1 + jc - 1) * nz + size_t(H.nxbeg - 1 + ic - 1) * nz * ny;
}

// 1D CFS-MPML profile of one axis (module_cfspml.f90:498-620): for i <= 1
// (unless the top is a free surface) and i >= n, nd = |i - edge| / pml,
// d0 = ln(1/R0) (2 + 1) pmlvp / (2 pml d) nd^2, alpha = max(alphamax (1 -
// nd), alphamax / 4); zero elsewhere.  Arrays di, dh, ai, ah of n + 2 pml.
void mpml_axis(int n, int pml, float d, float pmlvp, float alphamax, bool top_off, std::vector<float>& out) {
    const int len = n + 2 * pml;
    std::vector<float> di(len, 0.f), dh(len, 0.f), ai(len, 0.f), ah(len, 0.f);
    const float R0 = 1.0e-5f, extra = 0.25f * alphamax, lg = std::log(1.0f / R0);
    for (int i = -pml + 1; i <= n + pml; ++i) {
        int side = 0;
        if (!top_off && i <= 1) side = -1;
        else if (i >= n) side = 1;
        if (!side) continue;
        const int ref = side < 0 ? 1 : n, q = i + pml - 1;
        const float nd = float(std::abs(i - ref)) * 1.0f / float(pml);
        const float ndh = std::fabs(float(i - ref) - 0.5f) * 1.0f / float(pml);
        di[q] = lg * (2 + 1.0f) * pmlvp / (2.0f * float(pml) * d) * (nd * nd);
        dh[q] = lg * (2 + 1.0f) * pmlvp / (2.0f * float(pml) * d) * (ndh * ndh);
        ai[q] = std::max(alphamax * (1.0f - nd), extra);
        ah[q] = std::max(alphamax * (1.0f - ndh), extra);
    }
    for (auto* v : {&di, &dh, &ai, &ah}) out.insert(out.end(), v->begin(), v->end());
}

void plan_e(owl_elastic3d* b, const owl_acoustic_config* cfg, const float* vp, const float* vs, const float* rho,
            int nshot, const owl_shot* shots) {
    if (!cfg || !vp || !vs || !rho || nshot < 1 || !shots) fail(OWL_ERR_ARG, "null argument or no shots");
    if (!fd_coefficients(cfg->fdhalf, b->coef.c)) fail(OWL_ERR_ARG, "fdhalf must be 2..8");
    if (cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 1 || !(cfg->dx > 0) || !(cfg->dy > 0) || !(cfg->dz > 0) ||
        !(cfg->dt > 0))
        fail(OWL_ERR_ARG, "invalid grid or dt");
    if (cfg->pml < cfg->fdhalf || cfg->pml + cfg->fdhalf < kNkw) fail(OWL_ERR_ARG, "pml must be >= fdhalf");
    if (cfg->data_dt != cfg->dt || cfg->data_tmax != cfg->tmax)
        fail(OWL_ERR_ARG, "data_dt/data_tmax must equal dt/tmax (trace resampling is not provided)");
    b->cfg = *cfg;
    b->R = cfg->fdhalf;
    b->pml = cfg->pml;
    b->fs = cfg->free_surface ? 1 : 0;
    b->dt = cfg->dt;
    b->nt = time_steps(cfg->tmax, cfg->dt);
    if (b->nt < 2) fail(OWL_ERR_ARG, "tmax/dt gives fewer than 2 steps");
    b->idx = float(1.0 / double(cfg->dx));
    b->idy = float(1.0 / double(cfg->dy));
    b->idz = float(1.0 / double(cfg->dz));
    const size_t nfull = size_t(cfg->nx) * cfg->ny * cfg->nz;
    b->vp_full.assign(vp, vp + nfull);
    b->vs_full.assign(vs, vs + nfull);
    b->rho_full.assign(rho, rho + nfull);
    const float f0factor = cfg->f0factor > 0 ? cfg->f0factor : 1.0f;
    const int R = b->R, pml = b->pml, nt = b->nt, lo = 1 - pml - R;
    b->hs.assign(nshot, ShotHostE{});
    b->sd.assign(nshot, ShotE{});
    b->tile_shot.clear();
    b->nrec_total = b->max_nr = 0;
    b->pts_full = 0;
    long long off = 0, trace = 0;
    int prof = 0;
    std::vector<float> profs, amp, rec_w;
    std::vector<int> inj_beg{0};
    std::vector<Tap> inj;
    std::vector<std::vector<Tap>> rec_taps;  // per (shot, component, receiver)
    for (int s = 0; s < nshot; ++s) {
        const owl_shot& in = shots[s];
        ShotHostE& H = b->hs[s];
        ShotE& S = b->sd[s];
        H.id = in.id;
        H.nxbeg = in.nxbeg > 0 ? in.nxbeg : 1;
        H.nybeg = in.nybeg > 0 ? in.nybeg : 1;
        H.nzbeg = in.nzbeg > 0 ? in.nzbeg : 1;
        const int nxend = in.nxend > 0 ? in.nxend : cfg->nx, nyend = in.nyend > 0 ? in.nyend : cfg->ny,
                  nzend = in.nzend > 0 ? in.nzend : cfg->nz;
        if (H.nxbeg > nxend || nxend > cfg->nx || H.nybeg > nyend || nyend > cfg->ny || H.nzbeg > nzend ||
            nzend > cfg->nz)
            fail(OWL_ERR_ARG, "shot window outside the model");
        if (b->fs && H.nzbeg != 1) fail(OWL_ERR_ARG, "with a free surface every window starts at the top");
        H.nx = nxend - H.nxbeg + 1;
        H.ny = nyend - H.nybeg + 1;
        H.nz = nzend - H.nzbeg + 1;
        H.ox = float(H.nxbeg - 1) * cfg->dx + cfg->ox;
        H.oy = float(H.nybeg - 1) * cfg->dy + cfg->oy;
        H.oz = float(H.nzbeg - 1) * cfg->dz + cfg->oz;
        if (in.ns < 1 || !in.sources || in.nr < 1 || !in.receivers)
            fail(OWL_ERR_ARG, "shot without sources or receivers");
        H.src.assign(in.sources, in.sources + in.ns);
        H.rec.assign(in.receivers, in.receivers + in.nr);
        // pmlvp = sqrt(max c11/rho over the window's faces); vmax for the CFL
        float pm = 0.f, vmax = 0.f, f0max = 0.f;
        for (int i = 1; i <= H.nx; ++i)
            for (int j = 1; j <= H.ny; ++j)
                for (int k = 1; k <= H.nz; ++k) {
                    const size_t m = model_index(b, H, i, j, k);
                    const float r = b->rho_full[m], p = b->vp_full[m], v = b->vs_full[m];
                    if (!(r > 0.f) || !(p > 0.f) || !(v > 0.f)) fail(OWL_ERR_ARG, "vp, vs and rho must be positive");
                    const float c11 = r * (p * p);
                    vmax = std::max(vmax, std::sqrt(c11 / r));
                    if (i == 1 || i == H.nx || j == 1 || j == H.ny || k == 1 || k == H.nz) pm = std::max(pm, c11 / r);
                }
        const float pmlvp = std::sqrt(pm);
        const float dts = dt_stable(R, 3, vmax, cfg->dx, cfg->dy, cfg->dz);
        if (cfg->dt > dts)
            fail(OWL_ERR_CFL, "dt = " + std::to_string(cfg->dt) + " exceeds the stable dt = " + std::to_string(dts));
        for (auto& sc : H.src) f0max = std::max(f0max, sc.f0);
        S.nx = H.nx;
        S.ny = H.ny;
        S.nz = H.nz;
        S.px = (H.nx + 2 * (pml + R) + kPitchAlign - 1) / kPitchAlign * kPitchAlign;
        S.nyp = H.ny + 2 * (pml + R);
        S.nzp = H.nz + 2 * (pml + R);
        S.pxy = (long long)S.px * S.nyp;
        S.off = off;
        off += S.pxy * S.nzp;
        off = (off + 63) / 64 * 64;
        S.nty = (H.ny + 2 * pml + b->cy - 1) / b->cy;
        S.ntx = (H.nx + 2 * pml + 63) / 64;
        S.nzc = (H.nz + 2 * pml + b->kc - 1) / b->kc;
        S.tile0 = int(b->tile_shot.size());
        b->tile_shot.insert(b->tile_shot.end(), size_t(S.ntx) * S.nty * S.nzc, s);
        b->pts_full += double(H.nx + 2 * pml) * double(H.ny + 2 * pml) * double(H.nz + 2 * pml);
        S.prof = prof;
        const float amax = float(double(1.0f * f0max) * M_PI);
        mpml_axis(H.nx, pml, cfg->dx, pmlvp, amax, false, profs);
        mpml_axis(H.ny, pml, cfg->dy, pmlvp, amax, false, profs);
        mpml_axis(H.nz, pml, cfg->dz, pmlvp, amax, b->fs != 0, profs);
        prof += 4 * (H.nx + 2 * pml) + 4 * (H.ny + 2 * pml) + 4 * (H.nz + 2 * pml);
        S.nr = int(H.rec.size());
        S.trace0 = trace;
        trace += 3LL * nt * S.nr;
        // force sources; above the first node below a free surface they move
        // to z = dz with amplitude x 1.2 (module_vars.f90:650-655)
        const float xmax = float(H.nx - 1) * cfg->dx + H.ox, ymax = float(H.ny - 1) * cfg->dy + H.oy,
                    zmax = float(H.nz - 1) * cfg->dz + H.oz;
        std::vector<Tap> st;
        for (const owl_source& sc0 : H.src) {
            owl_source sc = sc0;
            if (sc.mechanism != OWL_FORCE) fail(OWL_ERR_ARG, "the elastic solver here takes force sources");
            if (b->fs && sc.z < cfg->dz) {
                sc.z = cfg->dz;
                sc.amp = sc.amp * 1.2f;
            }
            if (!(sc.x >= H.ox && sc.x <= xmax && sc.y >= H.oy && sc.y <= ymax && sc.z >= H.oz && sc.z <= zmax))
                fail(OWL_ERR_ARG, "source outside its shot window");
            const Located L = locate(3, sc.x, sc.y, sc.z, H.ox, H.oy, H.oz, cfg->dx, cfg->dy, cfg->dz);
            int hnt = 0;
            const std::vector<float> stf = source_time_function(sc.wavelet, sc.f0, cfg->dt, f0factor, &hnt);
            const int nbeg = fortran_nint(sc.t0 / cfg->dt) + 1, nend = nbeg + int(stf.size()) - 1;
            const float rs = b->rho_full[model_index(b, H, L.gx, L.gy, L.gz)];
            const float sp = std::sin(sc.polar), cp = std::cos(sc.polar);
            const float sa = std::sin(sc.azimuth), ca = std::cos(sc.azimuth);
            const float fac[3] = {sp * ca, sp * sa, cp};
            const int gx[3] = {L.hx, L.gx, L.gx}, gy[3] = {L.gy, L.hy, L.gy}, gz[3] = {L.gz, L.gz, L.hz};
            const float* wx[3] = {L.hxw, L.ix, L.ix};
            const float* wy[3] = {L.iy, L.hyw, L.iy};
            const float* wz[3] = {L.iz, L.iz, L.hzw};
            for (int c = 0; c < 3; ++c) {
                const int row = int(amp.size() / nt);
                for (int t = 1; t <= nt; ++t) {
                    const float a = (t >= nbeg && t <= nend) ? stf[t - nbeg] * sc.amp * cfg->dt : 0.0f;
                    amp.push_back(fac[c] * (a / rs));
                }
                for (int irz = -kNkw; irz <= kNkw; ++irz)
                    for (int iry = -kNkw; iry <= kNkw; ++iry)
                        for (int irx = -kNkw; irx <= kNkw; ++irx) {
                            const float w1 = wx[c][irx + kNkw], w2 = wy[c][iry + kNkw], w3 = wz[c][irz + kNkw];
                            if (w1 == 0.f || w2 == 0.f || w3 == 0.f) continue;
                            if (b->fs && gz[c] + irz < 2) continue;
                            st.push_back(Tap{(long long)cell_e(b, S, gx[c] + irx, gy[c] + iry, gz[c] + irz), w1, w2,
                                             w3, row, nbeg, nend, c});
                        }
            }
        }
        std::stable_sort(st.begin(), st.end(), [](const Tap& x, const Tap& y) {
            return x.field != y.field ? x.field < y.field : x.cell < y.cell;
        });
        S.inj0 = int(inj_beg.size()) - 1;
        S.ninj = 0;
        for (size_t k = 0; k < st.size(); ++k) {
            if (k == 0 || st[k].cell != st[k - 1].cell || st[k].field != st[k - 1].field) {
                if (k) inj_beg.push_back(int(inj.size()));
                ++S.ninj;
            }
            inj.push_back(st[k]);
        }
        if (!st.empty()) inj_beg.push_back(int(inj.size()));
        // receivers: three components (outside the window: weight 0)
        S.rec0 = b->nrec_total;
        std::vector<std::vector<Tap>> comp(3 * S.nr);
        for (int r = 0; r < S.nr; ++r) {
            const owl_receiver& rc = H.rec[r];
            const bool valid = rc.x >= H.ox && rc.x <= xmax && rc.y >= H.oy && rc.y <= ymax && rc.z >= H.oz &&
                               rc.z <= zmax;
            const float w = (valid && rc.weight != 0.f) ? rc.weight : 0.f;
            rec_w.push_back(w);
            if (w == 0.f) continue;
            const Located L = locate(3, rc.x, rc.y, rc.z, H.ox, H.oy, H.oz, cfg->dx, cfg->dy, cfg->dz);
            const int gx[3] = {L.hx, L.gx, L.gx}, gy[3] = {L.gy, L.hy, L.gy}, gz[3] = {L.gz, L.gz, L.hz};
            const float* wx[3] = {L.hxw, L.ix, L.ix};
            const float* wy[3] = {L.iy, L.hyw, L.iy};
            const float* wz[3] = {L.iz, L.iz, L.hzw};
            for (int c = 0; c < 3; ++c)
                for (int irz = -kNkw; irz <= kNkw; ++irz)
                    for (int iry = -kNkw; iry <= kNkw; ++iry)
                        for (int irx = -kNkw; irx <= kNkw; ++irx) {
                            const float w1 = wx[c][irx + kNkw], w2 = wy[c][iry + kNkw], w3 = wz[c][irz + kNkw];
                            if (w1 == 0.f || w2 == 0.f || w3 == 0.f) continue;
                            comp[size_t(c) * S.nr + r].push_back(
                                Tap{(long long)cell_e(b, S, gx[c] + irx, gy[c] + iry, gz[c] + irz), w1, w2, w3, r, 0, 0,
                                    c});
                        }
        }
        for (auto& v : comp) rec_taps.push_back(std::move(v));
        b->nrec_total += S.nr;
        b->max_nr = std::max(b->max_nr, S.nr);
    }
    b->field_elems = size_t(off);
    b->trace_elems = size_t(trace);
    // receiver CSR in (shot, component, receiver) order: entry 3*rec0 + c*nr + r
    std::vector<int> rec_beg{0};
    std::vector<Tap> rec;
    for (auto& v : rec_taps) {
        rec.insert(rec.end(), v.begin(), v.end());
        rec_beg.push_back(int(rec.size()));
    }
    if (!b->stream) {
        OWL_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        OWL_HIP(hipEventCreate(&b->ev0));
        OWL_HIP(hipEventCreate(&b->ev1));
    }
    hipStream_t stm = b->stream;
    b->d_shots.upload(b->sd, stm);
    b->d_tile_shot.upload(b->tile_shot, stm);
    b->d_prof.upload(profs, stm);
    if (amp.empty()) amp.push_back(0.f);
    b->d_amp.upload(amp, stm);
    b->d_inj_beg.upload(inj_beg, stm);
    if (inj.empty()) inj.push_back(Tap{});
    b->d_inj.upload(inj, stm);
    b->d_rec_beg.upload(rec_beg, stm);
    if (rec.empty()) rec.push_back(Tap{});
    b->d_rec.upload(rec, stm);
    b->d_rec_w.upload(rec_w, stm);
    // padded media (edge replication, permute(...,321)), then the effective
    // values at the staggered positions in OWL's operation order
    const size_t n = b->field_elems;
    std::vector<float> c11(n, 0.f), c12(n, 0.f), c44(n, 1.f), rh(n, 1.f);
    for (size_t s = 0; s < b->hs.size(); ++s) {
        const ShotHostE& H = b->hs[s];
        const ShotE& S = b->sd[s];
        for (int k = lo; k < lo + S.nzp; ++k)
            for (int j = lo; j < lo + S.nyp; ++j)
                for (int i = lo; i < lo + S.px; ++i) {
                    const size_t m = model_index(b, H, i, j, k);
                    const float r = b->rho_full[m], p = b->vp_full[m], v = b->vs_full[m];
                    const float a = r * (p * p), mu = r * (v * v);
                    const size_t q = size_t(S.off) + size_t(i - lo) + size_t(j - lo) * S.px + size_t(k - lo) * S.pxy;
                    c11[q] = a;
                    c12[q] = a - 2.0f * mu;
                    c44[q] = mu;
                    rh[q] = r;
                }
    }
    std::vector<float> myz(n, 0.f), mxz(n, 0.f), mxy(n, 0.f), rx(n, 1.f), ry(n, 1.f), rz(n, 1.f);
    for (size_t s = 0; s < b->sd.size(); ++s) {
        const ShotE& S = b->sd[s];
        const size_t sx = 1, sy = S.px, sz = S.pxy;
        for (int k = lo + 1; k < lo + S.nzp; ++k)
            for (int j = lo + 1; j < lo + S.nyp; ++j)
                for (int i = lo + 1; i < lo + S.nx + 2 * (pml + R); ++i) {
                    const size_t q = size_t(S.off) + size_t(i - lo) + size_t(j - lo) * S.px + size_t(k - lo) * S.pxy;
                    myz[q] = 4.0f / (1.0f / c44[q - sy - sz] + 1.0f / c44[q - sz] + 1.0f / c44[q - sy] + 1.0f / c44[q]);
                    mxz[q] = 4.0f / (1.0f / c44[q - sx - sz] + 1.0f / c44[q - sz] + 1.0f / c44[q - sx] + 1.0f / c44[q]);
                    mxy[q] = 4.0f / (1.0f / c44[q - sx - sy] + 1.0f / c44[q - sy] + 1.0f / c44[q - sx] + 1.0f / c44[q]);
                    rx[q] = 0.5f * (rh[q - sx] + rh[q]);
                    ry[q] = 0.5f * (rh[q - sy] + rh[q]);
                    rz[q] = 0.5f * (rh[q - sz] + rh[q]);
                }
    }
    const std::vector<float>* media[8] = {&c11, &c12, &myz, &mxz, &mxy, &rx, &ry, &rz};
    for (int k = 0; k < 8; ++k) b->med[k].upload(*media[k], stm);
    for (auto& f : b->fld) f.alloc(b->field_elems);
    b->d_traces.alloc(std::max<size_t>(b->trace_elems, 1));
    OWL_HIP(hipStreamSynchronize(stm));
}

EF fields(owl_elastic3d* b) {
    EF F;
    float** p = &F.sxx;
    for (int k = 0; k < 27; ++k) p[k] = b->fld[k].get();
    return F;
}

EM media(owl_elastic3d* b) {
    return EM{b->med[0].get(), b->med[1].get(), b->med[2].get(), b->med[3].get(),
              b->med[4].get(), b->med[5].get(), b->med[6].get(), b->med[7].get()};
}

// Algorithmic bytes per sweep: stress sweep reads vx, vy, vz, the six
// stresses and c11, c12 and the three shear moduli and writes six stresses
// (80 B per point); velocity sweep reads six stresses, three velocities and
// three densities and writes three velocities (60 B); memory variables in
// the slabs add 8 B per damped direction and derivative.
void sweep_bytes(const owl_elastic3d* b, double& bs, double& bv) {
    bs = bv = 0;
    for (auto& S : b->sd) {
        const double lx = S.nx + 2.0 * b->pml, ly = S.ny + 2.0 * b->pml, lz = S.nz + 2.0 * b->pml;
        const double pts = lx * ly * lz;
        const double slab = 2.0 * b->pml * (ly * lz + lx * lz + lx * ly);
        bs += 80.0 * pts + 8.0 * 3.0 * slab;
        bv += 60.0 * pts + 8.0 * 3.0 * slab;
    }
}

template <int R>
void run_forward_e(owl_elastic3d* b) {
    hipStream_t st = b->stream;
    for (auto& f : b->fld) f.zero(st);
    b->d_traces.zero(st);
    const EF F = fields(b);
    const EM M = media(b);
    const unsigned nshot = unsigned(b->sd.size());
    const dim3 blk(64, unsigned(b->cy)), grd(unsigned(b->tile_shot.size()));
    double bs, bv;
    sweep_bytes(b, bs, bv);
    const unsigned rg = unsigned(std::max(1, (3 * b->max_nr + 255) / 256));
    OWL_HIP(hipEventRecord(b->ev0, st));
    // step t: [free surface: mirror v above it] -> record(t) -> update -> source
    for (int t = 1; t <= b->nt; ++t) {
        if (b->fs) hipLaunchKernelGGL(ke_surface<R>, dim3(256, nshot), dim3(256), 0, st, b->d_shots.get(), F, b->pml, 2);
        int ev = b->prof.begin(st, 2, 0);
        hipLaunchKernelGGL(ke_record, dim3(rg, nshot), dim3(256), 0, st, b->d_shots.get(), F, b->d_rec_beg.get(),
                           b->d_rec.get(), b->d_rec_w.get(), b->d_traces.get(), b->nt, t);
        b->prof.end(st, ev);
        if (t == b->nt) break;  // the last update changes nothing that is recorded
        if (b->fs) hipLaunchKernelGGL(ke_surface<R>, dim3(256, nshot), dim3(256), 0, st, b->d_shots.get(), F, b->pml, 0);
        ev = b->prof.begin(st, 0, bs);
        hipLaunchKernelGGL(ke_stress<R>, grd, blk, 0, st, b->d_shots.get(), b->d_tile_shot.get(), F, M, b->d_prof.get(),
                           b->coef, b->dt, b->idx, b->idy, b->idz, b->pml, b->kc, b->fs);
        b->prof.end(st, ev);
        if (b->fs) hipLaunchKernelGGL(ke_surface<R>, dim3(256, nshot), dim3(256), 0, st, b->d_shots.get(), F, b->pml, 1);
        ev = b->prof.begin(st, 1, bv);
        hipLaunchKernelGGL(ke_velocity<R>, grd, blk, 0, st, b->d_shots.get(), b->d_tile_shot.get(), F, M,
                           b->d_prof.get(), b->coef, b->dt, b->idx, b->idy, b->idz, b->pml, b->kc, b->fs);
        b->prof.end(st, ev);
        ev = b->prof.begin(st, 2, 0);
        hipLaunchKernelGGL(ke_inject, dim3(nshot), dim3(1024), 0, st, b->d_shots.get(), F, b->d_inj_beg.get(),
                           b->d_inj.get(), b->d_amp.get(), b->nt, t);
        b->prof.end(st, ev);
        if (b->prof.pending() > 20000) {
            OWL_HIP(hipStreamSynchronize(st));
            b->prof.collect();
        }
    }
    OWL_HIP(hipGetLastError());
    OWL_HIP(hipEventRecord(b->ev1, st));
    OWL_HIP(hipStreamSynchronize(st));
    b->prof.collect();
    float ms = 0.f;
    OWL_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_ms = ms;
    b->last_updates = double(b->nt - 1) * b->pts_full;
}

template <class Fn>
void dispatch_Re(int R, Fn&& f) {
    switch (R) {
        case 2: f(std::integral_constant<int, 2>{}); break;
        case 3: f(std::integral_constant<int, 3>{}); break;
        case 4: f(std::integral_constant<int, 4>{}); break;
        case 5: f(std::integral_constant<int, 5>{}); break;
        case 6: f(std::integral_constant<int, 6>{}); break;
        case 7: f(std::integral_constant<int, 7>{}); break;
        case 8: f(std::integral_constant<int, 8>{}); break;
        default: fail(OWL_ERR_ARG, "fdhalf must be 2..8");
    }
}

}  // namespace

extern "C" {

owl_status owl_elastic3d_create(const owl_acoustic_config* cfg, const float* vp, const float* vs, const float* rho,
                                int nshot, const owl_shot* shots, owl_elastic3d** out) {
    return guarded([&] {
        if (!out) fail(OWL_ERR_ARG, "out is null");
        *out = nullptr;
        auto* b = new owl_elastic3d();
        try {
            plan_e(b, cfg, vp, vs, rho, nshot, shots);
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

void owl_elastic3d_destroy(owl_elastic3d* b) { delete b; }

int owl_elastic3d_nt(const owl_elastic3d* b) { return b ? b->nt : 0; }

owl_status owl_elastic3d_forward(owl_elastic3d* b, float* traces) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        dispatch_Re(b->R, [&](auto r) { run_forward_e<decltype(r)::value>(b); });
        if (traces) {
            std::vector<float> h(b->trace_elems);
            OWL_HIP(hipMemcpy(h.data(), b->d_traces.get(), h.size() * sizeof(float), hipMemcpyDeviceToHost));
            // device [3][nt][nr] per shot -> host [3][nr][nt] per shot
            for (auto& S : b->sd)
                for (int c = 0; c < 3; ++c)
                    for (int r = 0; r < S.nr; ++r)
                        for (int t = 0; t < b->nt; ++t)
                            traces[size_t(S.trace0) + (size_t(c) * S.nr + r) * b->nt + t] =
                                h[size_t(S.trace0) + size_t(c) * b->nt * S.nr + size_t(t) * S.nr + r];
        }
    });
}

owl_status owl_elastic3d_set_launch(owl_elastic3d* b, int planes, int rows) {
    return guarded([&] {
        if (!b || planes < 1 || (rows != 1 && rows != 2 && rows != 4 && rows != 8 && rows != 16))
            fail(OWL_ERR_ARG, "bad launch shape");
        b->kc = planes;
        b->cy = rows;
        int lt = 0;
        b->tile_shot.clear();
        for (size_t s = 0; s < b->sd.size(); ++s) {
            ShotE& S = b->sd[s];
            S.nty = (S.ny + 2 * b->pml + b->cy - 1) / b->cy;
            S.ntx = (S.nx + 2 * b->pml + 63) / 64;
            S.nzc = (S.nz + 2 * b->pml + b->kc - 1) / b->kc;
            S.tile0 = lt;
            lt += S.ntx * S.nty * S.nzc;
            b->tile_shot.insert(b->tile_shot.end(), size_t(S.ntx) * S.nty * S.nzc, int(s));
        }
        b->d_shots.upload(b->sd, b->stream);
        b->d_tile_shot.upload(b->tile_shot, b->stream);
        OWL_HIP(hipStreamSynchronize(b->stream));
    });
}

owl_status owl_elastic3d_set_profiling(owl_elastic3d* b, int on) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        b->prof.enable(on != 0);
    });
}

owl_status owl_elastic3d_kernel_stats(const owl_elastic3d* b, int kind, double* total_ms, long long* launches,
                                      double* bytes_total) {
    return guarded([&] {
        if (!b || kind < 0 || kind >= Profiler::kKinds) fail(OWL_ERR_ARG, "bad kernel kind");
        if (total_ms) *total_ms = b->prof.ms(kind);
        if (launches) *launches = b->prof.launches(kind);
        if (bytes_total) *bytes_total = b->prof.bytes(kind);
    });
}

owl_status owl_elastic3d_last_stats(const owl_elastic3d* b, double* ms, double* updates) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        if (ms) *ms = b->last_ms;
        if (updates) *updates = b->last_updates;
    });
}

}  // extern "C"
