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
#include "record_store.hpp"

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
    long long strip0;      // boundary records: nt blocks of strip_len
    long long g0;          // window gradient accumulators (nx*ny*nz, x fastest)
    int nadj, adj0;        // adjoint-source cells (CSR)
    int hnt;               // last reverse step OWL images for this shot (its first source's hnt)
};

// Wavefields and memory variables (named by derivative and field as in OWL).
struct EF {
    float *sxx, *syy, *szz, *syz, *sxz, *sxy, *vx, *vy, *vz;
    float *m_dxvx, *m_dyvy, *m_dzvz, *m_dyvz, *m_dzvy, *m_dxvz, *m_dzvx, *m_dxvy, *m_dyvx;
    float *m_dxxx, *m_dyxy, *m_dzxz, *m_dxxy, *m_dyyy, *m_dzyz, *m_dxxz, *m_dyyz, *m_dzzz;
};

// The nine wavefields a sweep writes.  The stress sweep reads the old
// stresses and the velocities from EF and writes the new stresses here; the
// velocity sweep reads the new stresses from here and the old velocities
// from EF.  In-place stepping passes EF's own pointers.
struct EV {
    float *sxx, *syy, *szz, *syz, *sxz, *sxy, *vx, *vy, *vz;
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
    int field;      // 0 vx, 1 vy, 2 vz, 3 sxx, 4 syy, 5 szz
};

// One receiver component's Kaiser-sinc box: the separable weights along x,
// y, z (0 = tap left out) and the cell of offsets (-kNkw, -kNkw, -kNkw),
// which may lie outside the grid where those weights are 0.
struct RecBox {
    long long cell0;
    float w[3][2 * kNkw + 1];
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

// Load and store at a 32-bit byte offset from a kernel-argument base (the
// sweeps' fields stay below 2 GiB, checked on the host): every array of a
// sweep is addressed by a few shared offset registers against scalar bases.
__device__ __forceinline__ float ldo(const float* base, unsigned off) {
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + off);
}
__device__ __forceinline__ void sto(float* base, unsigned off, float v) {
    *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + off) = v;
}
__device__ __forceinline__ float damp_o(float* w, unsigned qb, float v, float d, float alpha, float adt) {
    if (d == 0.0f) return v;
    const float a_ = d / 1.0f;
    const float b_ = d / 1.0f + alpha;
    const float ca = -(2.0f * adt * a_ / (2.0f + adt * b_));
    const float cb = (2.0f - adt * b_) / (2.0f + adt * b_);
    const float m = ca * v + cb * ldo(w, qb);
    sto(w, qb, m);
    return (v + m) / 1.0f;
}

// x stencil sum_m c_m (w(i+off+m) - w(i+off-m+1)) from the lanes' values of
// the current plane: lane l holds column i and reads columns
// i - R + off + 1 .. i + R + off from lanes l - R + off + 1 .. l + R + off.
template <int R>
__device__ __forceinline__ float sten_lanes(float v, int l, int off, const Coef& c) {
    float xv[2 * R];
#pragma unroll
    for (int k = 0; k < 2 * R; ++k) xv[k] = __shfl(v, l - R + off + 1 + k, 64);
    float s = c.c[0] * (xv[R] - xv[R - 1]);
#pragma unroll
    for (int m = 2; m <= R; ++m) s = s + c.c[m - 1] * (xv[R - 1 + m] - xv[R - m]);
    return s;
}

// y stencil sum_m c_m (w(j+off+m) - w(j+off-m+1)) by loads at byte offsets
// yo[k] = q + (k - R) rows, k = 0 .. 2R.
template <int R>
__device__ __forceinline__ float sten_y(const float* w, const unsigned* yo, int off, const Coef& c) {
    float s = c.c[0] * (ldo(w, yo[R + off + 1]) - ldo(w, yo[R + off]));
#pragma unroll
    for (int m = 2; m <= R; ++m) s = s + c.c[m - 1] * (ldo(w, yo[R + off + m]) - ldo(w, yo[R + off + 1 - m]));
    return s;
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
                                                 EF F, EV O, EM M, const float* __restrict__ prof, Coef c, float dt,
                                                 float idx, float idy, float idz, int pml, int kc, int fs,
                                                 int ntile) {
    constexpr int W = 64 - 2 * R;  // columns a wavefront updates; the other lanes feed its x stencils
    // XCD-contiguous order: blocks are dealt to the 8 XCDs in turn, so block b
    // takes tile (b % 8) * per + b / 8 and each XCD walks one contiguous range
    // of tiles, whose y neighbours then share its L2
    const int per = (ntile + 7) / 8;
    const int tile = int(blockIdx.x % 8u) * per + int(blockIdx.x / 8u);
    if (tile >= ntile) return;
    const ShotE S = shots[tile_shot[tile]];
    int lt = tile - S.tile0;
    const int tx = lt % S.ntx;
    lt /= S.ntx;
    const int ty = lt % S.nty, tzc = lt / S.nty;
    const int nx = S.nx, ny = S.ny, nz = S.nz;
    const int lo = 1 - pml - R;
    const int l = int(threadIdx.x);
    const int i = -pml + 1 + tx * W - R + l;
    // blockDim.x = 64: the row is wavefront-uniform (scalar registers)
    const int j = -pml + 1 + ty * int(blockDim.y) + __builtin_amdgcn_readfirstlane(int(threadIdx.y));
    if (j > ny + pml) return;  // whole wavefront
    const int kt = -pml + 1 + tzc * kc, k1 = min(kt + kc - 1, nz + pml);
    const int k0 = fs ? max(kt, 1) : kt;
    if (k0 > k1) return;
    const bool mcol = i <= nx + pml + R;
    const bool own = l >= R && l <= 63 - R && i <= nx + pml;
    const Prof P = prof_of(prof, S, pml);
    const float adt = fabsf(dt);
    const float* __restrict__ vx = F.vx;
    const float* __restrict__ vy = F.vy;
    const float* __restrict__ vz = F.vz;
    const unsigned syb = unsigned(S.px) * 4u, szb = unsigned(S.pxy) * 4u;
    const int ii = i + pml - 1, jj = j + pml - 1;
    unsigned q = unsigned(ate(S, lo, mcol ? i : lo, j, k0)) * 4u;
    // queues: plane k + m - R at slot m, m = 0 .. 2R
    float qx[2 * R + 1], qy[2 * R + 1], qz[2 * R + 1];
#pragma unroll
    for (int m = 0; m < 2 * R; ++m) {
        const unsigned r = q + unsigned(m - R) * szb;
        qx[m] = mcol ? ldo(vx, r) : 0.0f;
        qy[m] = mcol ? ldo(vy, r) : 0.0f;
        qz[m] = mcol ? ldo(vz, r) : 0.0f;
    }
    const bool xin = !(i <= 0 || i >= nx + 1), yin = !(j <= 0 || j >= ny + 1);
    const bool Xin = !(i <= 1 || i >= nx + 1), Yin = !(j <= 1 || j >= ny + 1);
    // The inputs of one plane: the queues' fronts (plane k + R) and, on the
    // owned lanes, the moduli and old stresses.  Plane k + 1's are loaded
    // while plane k is computed (the y stencils stay direct loads).
    struct Plane {
        float fx, fy, fz;
        float c11, c12, myz, mxz, mxy;
        float sxx, syy, szz, syz, sxz, sxy;
    };
    auto load = [&](unsigned qp, Plane& w) {
        const unsigned r = qp + unsigned(R) * szb;
        w.fx = mcol ? ldo(vx, r) : 0.0f;
        w.fy = mcol ? ldo(vy, r) : 0.0f;
        w.fz = mcol ? ldo(vz, r) : 0.0f;
        if (own) {
            w.c11 = ldo(M.c11, qp);
            w.c12 = ldo(M.c12, qp);
            w.myz = ldo(M.mu_yz, qp);
            w.mxz = ldo(M.mu_xz, qp);
            w.mxy = ldo(M.mu_xy, qp);
            w.sxx = ldo(F.sxx, qp);
            w.syy = ldo(F.syy, qp);
            w.szz = ldo(F.szz, qp);
            w.syz = ldo(F.syz, qp);
            w.sxz = ldo(F.sxz, qp);
            w.sxy = ldo(F.sxy, qp);
        }
    };
    Plane cur{}, nxt{};
    load(q, cur);
    for (int k = k0; k <= k1; ++k, q += szb) {
        if (k < k1) load(q + szb, nxt);
        qx[2 * R] = cur.fx;
        qy[2 * R] = cur.fy;
        qz[2 * R] = cur.fz;
        // x stencils of this plane from the lanes (all lanes take part)
        const float sx_vx = sten_lanes<R>(qx[R], l, 0, c);    // d+x vx
        const float sx_vz = sten_lanes<R>(qz[R], l, -1, c);   // d-x vz
        const float sx_vy = sten_lanes<R>(qy[R], l, -1, c);   // d-x vy
        if (own) {
            const int kk = k + pml - 1;
            unsigned yo[2 * R + 1];  // rows j - R .. j + R of this plane
#pragma unroll
            for (int m = 0; m <= 2 * R; ++m) yo[m] = q + unsigned(m - R) * syb;
            // ---- normal stresses
            {
                float pdxvx = idx * sx_vx, pdyvy = idy * sten_y<R>(vy, yo, 0, c), pdzvz;
                const float c11 = cur.c11, c12 = cur.c12;
                if (fs && k == 1) {
                    if (!xin || !yin) {
                        pdxvx = damp_o(F.m_dxvx, q, pdxvx, P.dix[ii], P.aix[ii], adt);
                        pdyvy = damp_o(F.m_dyvy, q, pdyvy, P.diy[jj], P.aiy[jj], adt);
                    }
                    pdzvz = -(c12 / c11 * pdxvx) - c12 / c11 * pdyvy;
                } else {
                    float s = c.c[0] * (qz[R + 1] - qz[R]);
#pragma unroll
                    for (int m = 2; m <= R; ++m) s = s + c.c[m - 1] * (qz[R + m] - qz[R - m + 1]);
                    pdzvz = idz * s;
                    if (!xin || !yin || (!fs && k <= 0) || k >= nz + 1) {
                        pdxvx = damp_o(F.m_dxvx, q, pdxvx, P.dix[ii], P.aix[ii], adt);
                        pdyvy = damp_o(F.m_dyvy, q, pdyvy, P.diy[jj], P.aiy[jj], adt);
                        pdzvz = damp_o(F.m_dzvz, q, pdzvz, P.diz[kk], P.aiz[kk], adt);
                    }
                }
                sto(O.sxx, q, cur.sxx + dt * (c11 * pdxvx + c12 * pdyvy + c12 * pdzvz));
                sto(O.syy, q, cur.syy + dt * (c12 * pdxvx + c11 * pdyvy + c12 * pdzvz));
                sto(O.szz, q, cur.szz + dt * (c12 * pdxvx + c12 * pdyvy + c11 * pdzvz));
            }
            // ---- shear stresses with a z derivative: syz, sxz (Z = k)
            if (!fs || k >= 2) {
                float a = c.c[0] * (qy[R] - qy[R - 1]);   // d-z vy
                float e = c.c[0] * (qx[R] - qx[R - 1]);   // d-z vx
#pragma unroll
                for (int m = 2; m <= R; ++m) {
                    a = a + c.c[m - 1] * (qy[R + m - 1] - qy[R - m]);
                    e = e + c.c[m - 1] * (qx[R + m - 1] - qx[R - m]);
                }
                float pdzvy = idz * a, pdyvz = idy * sten_y<R>(vz, yo, -1, c), pdzvx = idz * e, pdxvz = idx * sx_vz;
                const bool zpml = (!fs && k <= 1) || k >= nz + 1;
                if (!xin || !Yin || zpml) {
                    pdyvz = damp_o(F.m_dyvz, q, pdyvz, P.dhy[jj], P.ahy[jj], adt);
                    pdzvy = damp_o(F.m_dzvy, q, pdzvy, P.dhz[kk], P.ahz[kk], adt);
                }
                sto(O.syz, q, cur.syz + dt * cur.myz * (pdzvy + pdyvz));
                if (!Xin || !yin || zpml) {
                    pdxvz = damp_o(F.m_dxvz, q, pdxvz, P.dhx[ii], P.ahx[ii], adt);
                    pdzvx = damp_o(F.m_dzvx, q, pdzvx, P.dhz[kk], P.ahz[kk], adt);
                }
                sto(O.sxz, q, cur.sxz + dt * cur.mxz * (pdzvx + pdxvz));
            }
            // ---- sxy (X, Y, k)
            {
                float pdyvx = idy * sten_y<R>(vx, yo, -1, c), pdxvy = idx * sx_vy;
                if (!Xin || !Yin || (!fs && (k <= 0 || k >= nz + 1))) {
                    pdxvy = damp_o(F.m_dxvy, q, pdxvy, P.dhx[ii], P.ahx[ii], adt);
                    pdyvx = damp_o(F.m_dyvx, q, pdyvx, P.dhy[jj], P.ahy[jj], adt);
                }
                sto(O.sxy, q, cur.sxy + dt * cur.mxy * (pdyvx + pdxvy));
            }
        }
#pragma unroll
        for (int m = 0; m < 2 * R; ++m) {
            qx[m] = qx[m + 1];
            qy[m] = qy[m + 1];
            qz[m] = qz[m + 1];
        }
        cur = nxt;
    }
}

// ---------------------------------------------------------------------------
// Velocity sweep: vx at (X, j, k), vy at (i, Y, k), vz at (i, j, Z).
// module_wavefield.f90:316-503 (free surface :852-1090).  z stencils from
// queues over planes k-R .. k+R of sxz, syz, szz.
// ---------------------------------------------------------------------------
template <int R>
__global__ __launch_bounds__(256) void ke_velocity(const ShotE* __restrict__ shots,
                                                   const int* __restrict__ tile_shot, EF F, EV O, EM M,
                                                   const float* __restrict__ prof, Coef c, float dt, float idx,
                                                   float idy, float idz, int pml, int kc, int fs, int ntile) {
    constexpr int W = 64 - 2 * R;
    // XCD-contiguous order: blocks are dealt to the 8 XCDs in turn, so block b
    // takes tile (b % 8) * per + b / 8 and each XCD walks one contiguous range
    // of tiles, whose y neighbours then share its L2
    const int per = (ntile + 7) / 8;
    const int tile = int(blockIdx.x % 8u) * per + int(blockIdx.x / 8u);
    if (tile >= ntile) return;
    const ShotE S = shots[tile_shot[tile]];
    int lt = tile - S.tile0;
    const int tx = lt % S.ntx;
    lt /= S.ntx;
    const int ty = lt % S.nty, tzc = lt / S.nty;
    const int nx = S.nx, ny = S.ny, nz = S.nz;
    const int lo = 1 - pml - R;
    const int l = int(threadIdx.x);
    const int i = -pml + 1 + tx * W - R + l;
    // blockDim.x = 64: the row is wavefront-uniform (scalar registers)
    const int j = -pml + 1 + ty * int(blockDim.y) + __builtin_amdgcn_readfirstlane(int(threadIdx.y));
    if (j > ny + pml) return;  // whole wavefront
    const int kt = -pml + 1 + tzc * kc, k1 = min(kt + kc - 1, nz + pml);
    const int k0 = fs ? max(kt, 1) : kt;
    if (k0 > k1) return;
    const bool mcol = i <= nx + pml + R;
    const bool own = l >= R && l <= 63 - R && i <= nx + pml;
    const Prof P = prof_of(prof, S, pml);
    const float adt = fabsf(dt);
    const float* __restrict__ sxx = O.sxx;
    const float* __restrict__ syy = O.syy;
    const float* __restrict__ szz = O.szz;
    const float* __restrict__ syz = O.syz;
    const float* __restrict__ sxz = O.sxz;
    const float* __restrict__ sxy = O.sxy;
    const unsigned syb = unsigned(S.px) * 4u, szb = unsigned(S.pxy) * 4u;
    const int ii = i + pml - 1, jj = j + pml - 1;
    unsigned q = unsigned(ate(S, lo, mcol ? i : lo, j, k0)) * 4u;
    float qa[2 * R + 1], qb[2 * R + 1], qc[2 * R + 1];  // sxz, syz, szz
#pragma unroll
    for (int m = 0; m < 2 * R; ++m) {
        const unsigned r = q + unsigned(m - R) * szb;
        qa[m] = mcol ? ldo(sxz, r) : 0.0f;
        qb[m] = mcol ? ldo(syz, r) : 0.0f;
        qc[m] = mcol ? ldo(szz, r) : 0.0f;
    }
    const bool xin = !(i <= 0 || i >= nx + 1), yin = !(j <= 0 || j >= ny + 1);
    const bool Xin = !(i <= 1 || i >= nx + 1), Yin = !(j <= 1 || j >= ny + 1);
    // The inputs of one plane: the queues' fronts (plane k + R), sxx and sxy
    // and, on the owned lanes, the densities and old velocities.  Plane
    // k + 1's are loaded while plane k is computed (the y stencils stay
    // direct loads).
    struct Plane {
        float fa, fb, fc, cxx, cxy;
        float rx, ry, rz, vx, vy, vz;
    };
    auto load = [&](unsigned qp, Plane& w) {
        const unsigned r = qp + unsigned(R) * szb;
        w.fa = mcol ? ldo(sxz, r) : 0.0f;
        w.fb = mcol ? ldo(syz, r) : 0.0f;
        w.fc = mcol ? ldo(szz, r) : 0.0f;
        w.cxx = mcol ? ldo(sxx, qp) : 0.0f;
        w.cxy = mcol ? ldo(sxy, qp) : 0.0f;
        if (own) {
            w.rx = ldo(M.rx, qp);
            w.ry = ldo(M.ry, qp);
            w.rz = ldo(M.rz, qp);
            w.vx = ldo(F.vx, qp);
            w.vy = ldo(F.vy, qp);
            w.vz = ldo(F.vz, qp);
        }
    };
    Plane cur{}, nxt{};
    load(q, cur);
    for (int k = k0; k <= k1; ++k, q += szb) {
        if (k < k1) load(q + szb, nxt);
        qa[2 * R] = cur.fa;
        qb[2 * R] = cur.fb;
        qc[2 * R] = cur.fc;
        // x stencils of this plane from the lanes (all lanes take part)
        const float cxx = cur.cxx, cxy = cur.cxy;
        const float sx_xx = sten_lanes<R>(cxx, l, -1, c);    // d-x sxx
        const float sx_xy = sten_lanes<R>(cxy, l, 0, c);     // d+x sxy
        const float sx_xz = sten_lanes<R>(qa[R], l, 0, c);   // d+x sxz
        if (own) {
            const int kk = k + pml - 1;
            const bool zin = !((!fs && k <= 0) || k >= nz + 1);
            unsigned yo[2 * R + 1];  // rows j - R .. j + R of this plane
#pragma unroll
            for (int m = 0; m <= 2 * R; ++m) yo[m] = q + unsigned(m - R) * syb;
            // ---- vx (X, j, k)
            {
                float e = c.c[0] * (qa[R + 1] - qa[R]);          // d+z sxz
#pragma unroll
                for (int m = 2; m <= R; ++m) e = e + c.c[m - 1] * (qa[R + m] - qa[R - m + 1]);
                float pdxxx = idx * sx_xx, pdyxy = idy * sten_y<R>(sxy, yo, 0, c), pdzxz = idz * e;
                if (!Xin || !yin || !zin) {
                    pdxxx = damp_o(F.m_dxxx, q, pdxxx, P.dhx[ii], P.ahx[ii], adt);
                    pdyxy = damp_o(F.m_dyxy, q, pdyxy, P.diy[jj], P.aiy[jj], adt);
                    pdzxz = damp_o(F.m_dzxz, q, pdzxz, P.diz[kk], P.aiz[kk], adt);
                }
                sto(O.vx, q, cur.vx + dt / cur.rx * (pdxxx + pdyxy + pdzxz));
            }
            // ---- vy (i, Y, k)
            {
                float e = c.c[0] * (qb[R + 1] - qb[R]);          // d+z syz
#pragma unroll
                for (int m = 2; m <= R; ++m) e = e + c.c[m - 1] * (qb[R + m] - qb[R - m + 1]);
                float pdxxy = idx * sx_xy, pdyyy = idy * sten_y<R>(syy, yo, -1, c), pdzyz = idz * e;
                if (!xin || !Yin || !zin) {
                    pdxxy = damp_o(F.m_dxxy, q, pdxxy, P.dix[ii], P.aix[ii], adt);
                    pdyyy = damp_o(F.m_dyyy, q, pdyyy, P.dhy[jj], P.ahy[jj], adt);
                    pdzyz = damp_o(F.m_dzyz, q, pdzyz, P.diz[kk], P.aiz[kk], adt);
                }
                sto(O.vy, q, cur.vy + dt / cur.ry * (pdxxy + pdyyy + pdzyz));
            }
            // ---- vz (i, j, Z = k)
            if (!fs || k >= 2) {
                float e = c.c[0] * (qc[R] - qc[R - 1]);          // d-z szz
#pragma unroll
                for (int m = 2; m <= R; ++m) e = e + c.c[m - 1] * (qc[R + m - 1] - qc[R - m]);
                float pdxxz = idx * sx_xz, pdyyz = idy * sten_y<R>(syz, yo, 0, c), pdzzz = idz * e;
                const bool Zin = !((!fs && k <= 1) || k >= nz + 1);
                if (!xin || !yin || !Zin) {
                    pdxxz = damp_o(F.m_dxxz, q, pdxxz, P.dix[ii], P.aix[ii], adt);
                    pdyyz = damp_o(F.m_dyyz, q, pdyyz, P.diy[jj], P.aiy[jj], adt);
                    pdzzz = damp_o(F.m_dzzz, q, pdzzz, P.dhz[kk], P.ahz[kk], adt);
                }
                sto(O.vz, q, cur.vz + dt / cur.rz * (pdxxz + pdyyz + pdzzz));
            }
        }
#pragma unroll
        for (int m = 0; m < 2 * R; ++m) {
            qa[m] = qa[m + 1];
            qb[m] = qb[m + 1];
            qc[m] = qc[m + 1];
        }
        cur = nxt;
    }
}

// Free surface, whole padded x-y planes (module_wavefield.f90:580-590,
// 823-845; submodule_forward_modeling.f90:64-73).  mode 0: zero the
// velocities above the surface; 1: mirror the stresses; 2: mirror the
// velocities for the receivers (NKW planes); 3: zero the velocities above
// it over the planes modes 0 and 2 touch.
template <int R>
__global__ void ke_surface(const ShotE* __restrict__ shots, EF F, int pml, int mode) {
    const ShotE S = shots[blockIdx.y];
    const int lo = 1 - pml - R;
    const long long plane = (long long)S.px * S.nyp;
    const int nk = mode == 2 ? kNkw : (mode == 3 ? (R > kNkw ? R : kNkw) : R);
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < plane; e += (long long)gridDim.x * blockDim.x) {
        const size_t b = size_t(S.off) + size_t(e);
        auto at = [&](int k) { return b + size_t(k - lo) * size_t(S.pxy); };
        if (mode == 0 || mode == 3) {
            for (int m = 1; m <= nk; ++m) {
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

// Source injection of step t (add_source, module_wavefield.f90:1115-1228):
// forces into vx, vy, vz, explosions into sxx, syy, szz; one block per shot,
// taps grouped by (field, cell) in source order.
__global__ __launch_bounds__(1024) void ke_inject(const ShotE* __restrict__ shots, EF F,
                                                  const int* __restrict__ inj_beg, const Tap* __restrict__ inj,
                                                  const float* __restrict__ amp_tab, int nt, int t) {
    const ShotE S = shots[blockIdx.x];
    for (int c = threadIdx.x; c < S.ninj; c += blockDim.x) {
        const int b = inj_beg[S.inj0 + c], e = inj_beg[S.inj0 + c + 1];
        const long long cell = inj[b].cell;
        const int fld = inj[b].field;
        float* f = fld == 0 ? F.vx : fld == 1 ? F.vy : fld == 2 ? F.vz : fld == 3 ? F.sxx : fld == 4 ? F.syy : F.szz;
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
// (gx, gy, hz) (submodule_forward_modeling.f90:75-140).  The taps of a box
// are summed z, y, x outermost to innermost, skipping zero weights, as OWL
// lists them.
__global__ __launch_bounds__(256) void ke_record(const ShotE* __restrict__ shots, EF F,
                                                 const RecBox* __restrict__ boxes,
                                                 const float* __restrict__ rec_w, float* traces, int nt, int t) {
    constexpr int K = 2 * kNkw + 1;
    const ShotE S = shots[blockIdx.y];
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < 3 * S.nr; r += gridDim.x * blockDim.x) {
        const int comp = r / S.nr, ir = r % S.nr;
        const RecBox& B = boxes[3 * S.rec0 + comp * S.nr + ir];
        const float w = rec_w[S.rec0 + ir];
        const float* f = comp == 0 ? F.vx : (comp == 1 ? F.vy : F.vz);
        float wx[K];
#pragma unroll
        for (int k = 0; k < K; ++k) wx[k] = B.w[0][k];
        float acc = 0.0f;
        for (int kz = 0; kz < K; ++kz) {
            const float w3 = B.w[2][kz];
            if (w3 == 0.f) continue;
            for (int ky = 0; ky < K; ++ky) {
                const float w2 = B.w[1][ky];
                if (w2 == 0.f) continue;
                const float* row = f + (B.cell0 + (long long)ky * S.px + (long long)kz * S.pxy);
#pragma unroll
                for (int kx = 0; kx < K; ++kx) {
                    const float w1 = wx[kx];
                    if (w1 == 0.f) continue;
                    acc = acc + row[kx] * w1 * w2 * w3 * w;
                }
            }
        }
        traces[S.trace0 + size_t(comp) * nt * S.nr + size_t(t - 1) * S.nr + ir] = acc;
    }
}

// ---------------------------------------------------------------------------
// OWL's elastic gradient (elastic_vhtiort_3d/submodule_fwi_gradient.f90).
// ---------------------------------------------------------------------------

// Boundary records (module_boundary_saving.f90:83-236): per velocity
// component the x faces i in [2-R, 1] and [nx+1, nx+R], the y faces, the z
// faces, each over [1-R, n+R] on the other two axes, first index fastest;
// vx's six sections, then vy's, then vz's.
struct EStrip {
    long long a, b, c;  // section sizes: R*ey*ez, ex*R*ez, ex*ey*R
    int ex, ey;
};

__host__ __device__ inline EStrip estrip_geo(int nx, int ny, int nz, int R) {
    EStrip g;
    g.ex = nx + 2 * R;
    g.ey = ny + 2 * R;
    const int ez = nz + 2 * R;
    g.a = (long long)R * g.ey * ez;
    g.b = (long long)g.ex * R * ez;
    g.c = (long long)g.ex * g.ey * R;
    return g;
}

// save != 0: record of step t <- (vx, vy, vz); else the reverse.
template <int R>
__global__ void ke_strips(const ShotE* __restrict__ shots, float* vx, float* vy, float* vz, RecStore st, int t,
                          int pml, int mode) {
    const ShotE S = shots[blockIdx.y];
    const EStrip g = estrip_geo(S.nx, S.ny, S.nz, R);
    const long long one = 2 * (g.a + g.b + g.c), len = 3 * one;
    const int lo = 1 - pml - R, nx = S.nx, ny = S.ny, nz = S.nz;
    const long long rec0 = S.strip0 + (long long)(t - 1) * len;
    const int rid = int(blockIdx.y) * st.nt + (t - 1);
    const float sc = mode == kRecAmax ? 1.0f : rec_scale(st, rid), inv = 1.0f / sc;
    float m = 0.0f;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < len;
         e += (long long)gridDim.x * blockDim.x) {
        const int comp = int(e / one);
        long long r = e - comp * one;
        int i, j, k;
        if (r < 2 * g.a) {
            const bool hi = r >= g.a;
            if (hi) r -= g.a;
            i = (hi ? nx + 1 : 2 - R) + int(r % R);
            j = 1 - R + int((r / R) % g.ey);
            k = 1 - R + int(r / (R * (long long)g.ey));
        } else if ((r -= 2 * g.a) < 2 * g.b) {
            const bool hi = r >= g.b;
            if (hi) r -= g.b;
            i = 1 - R + int(r % g.ex);
            j = (hi ? ny + 1 : 2 - R) + int((r / g.ex) % R);
            k = 1 - R + int(r / (g.ex * (long long)R));
        } else {
            r -= 2 * g.b;
            const bool hi = r >= g.c;
            if (hi) r -= g.c;
            i = 1 - R + int(r % g.ex);
            j = 1 - R + int((r / g.ex) % g.ey);
            k = (hi ? nz + 1 : 2 - R) + int(r / (g.ex * (long long)g.ey));
        }
        float* f = comp == 0 ? vx : (comp == 1 ? vy : vz);
        const size_t q = ate(S, lo, i, j, k);
        if (mode == kRecAmax)
            m = fmaxf(m, fabsf(f[q]));
        else if (mode == kRecSave)
            rec_put(st, rec0 + e, f[q], inv);
        else
            f[q] = rec_get(st, rec0 + e, sc);
    }
    if (mode == kRecAmax) rec_amax(st, rid, m);
}

// Adjoint sources into the adjoint velocities (submodule_fwi_gradient.f90:
// 256-330): cells grouped by (component, cell), contributions of one cell
// in receiver order; adj holds three [nt][nr] blocks per shot.
__global__ void ke_inject_adjoint(const ShotE* __restrict__ shots, float* vx, float* vy, float* vz,
                                  const int* __restrict__ cell_beg, const Tap* __restrict__ taps,
                                  const float* __restrict__ rec_w, const float* __restrict__ adj, int nt, int t) {
    const ShotE S = shots[blockIdx.y];
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < S.nadj; c += gridDim.x * blockDim.x) {
        const int b = cell_beg[S.adj0 + c], e = cell_beg[S.adj0 + c + 1];
        const long long cell = taps[b].cell;
        const int fld = taps[b].field;
        float* f = fld == 0 ? vx : (fld == 1 ? vy : vz);
        const float* a_t = adj + S.trace0 + size_t(fld) * nt * S.nr + size_t(t - 1) * S.nr;
        float acc = f[cell];
        for (int k = b; k < e; ++k) {
            const Tap tp = taps[k];
            acc = acc + a_t[tp.row] * tp.w1 * tp.w2 * tp.w3 * rec_w[S.rec0 + tp.row];
        }
        f[cell] = acc;
    }
}

// Isotropic compliances (submodule_fwi_gradient.f90:73-95) with c22 = c33 =
// c11 and c13 = c23 = c12, in the reference's expression order: s22 = s33 =
// s11 and s23 = s12 bit for bit, s13 has the other denominator.
__device__ __forceinline__ void compliances(float c11, float c12, float c44, float& s11, float& s12, float& s13,
                                            float& s44) {
    const float c13 = c12, c22 = c11, c23 = c12, c33 = c11;
    const float den1 = c13 * c13 * c22 - 2.0f * c12 * c13 * c23 + c11 * (c23 * c23) + c12 * c12 * c33 - c11 * c22 * c33;
    const float den2 = c13 * c13 * c22 - 2.0f * c12 * c13 * c23 + c12 * c12 * c33 + c11 * (c23 * c23 - c22 * c33);
    s11 = (c23 * c23 - c22 * c33) / den1;
    s12 = (-c13 * c23 + c12 * c33) / den2;
    s13 = (c13 * c22 - c12 * c23) / den1;
    s44 = c44 != 0.0f ? 1.0f / c44 : 0.0f;
}

// Imaging condition (compute_gradient, submodule_fwi_gradient.f90:702-827)
// over the window interior.  Accumulates, in FP64, the three combinations
// the chain rule (:486-495) needs:
//   A = g11 + g12 + g13 + g22 + g23 + g33      (vp)
//   B = g44 + g55 + g66 - 2 (g12 + g13 + g23)  (vs)
//   C = sum over x, y, z of (v(i)+v(i+1) - prev)(vr(i)+vr(i+1))  (rho)
// with g_ij = -(strain of the forward increment) . (adjoint strain).
__global__ __launch_bounds__(256) void ke_image(const ShotE* __restrict__ shots, EV cur, EV prev, EV adj,
                                                const float* __restrict__ c11, const float* __restrict__ c12,
                                                const float* __restrict__ c44, double* gA, double* gB, double* gC,
                                                int prev_zero, int pml, int R, int t) {
    const ShotE S = shots[blockIdx.y];
    if (t < S.hnt) return;  // this shot's reverse loop has ended (whole block)
    const int nx = S.nx, ny = S.ny, nz = S.nz, lo = 1 - pml - R;
    const long long n = (long long)nx * ny * nz;
    const size_t sx = 1, sy = S.px, sz = S.pxy;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        const int i = 1 + int(e % nx), j = 1 + int((e / nx) % ny), k = 1 + int(e / ((long long)nx * ny));
        const size_t q = ate(S, lo, i, j, k);
        float s11, s12, s13, s44;
        compliances(c11[q], c12[q], c44[q], s11, s12, s13, s44);
        const float s22 = s11, s33 = s11, s23 = s12;
        auto avg4 = [](const float* f, size_t b, size_t u, size_t w) {
            return 0.25f * (f[b] + f[b + u] + f[b + w] + f[b + u + w]);
        };
        float txx = cur.sxx[q], tyy = cur.syy[q], tzz = cur.szz[q];
        float tyz = avg4(cur.syz, q, sy, sz), txz = avg4(cur.sxz, q, sx, sz), txy = avg4(cur.sxy, q, sx, sy);
        float svx = cur.vx[q] + cur.vx[q + sx], svy = cur.vy[q] + cur.vy[q + sy], svz = cur.vz[q] + cur.vz[q + sz];
        if (!prev_zero) {
            txx = txx - prev.sxx[q];
            tyy = tyy - prev.syy[q];
            tzz = tzz - prev.szz[q];
            tyz = tyz - avg4(prev.syz, q, sy, sz);
            txz = txz - avg4(prev.sxz, q, sx, sz);
            txy = txy - avg4(prev.sxy, q, sx, sy);
            svx = svx - (prev.vx[q] + prev.vx[q + sx]);
            svy = svy - (prev.vy[q] + prev.vy[q + sy]);
            svz = svz - (prev.vz[q] + prev.vz[q + sz]);
        }
        const float rxx = adj.sxx[q], ryy = adj.syy[q], rzz = adj.szz[q];
        const float ryz = avg4(adj.syz, q, sy, sz), rxz = avg4(adj.sxz, q, sx, sz), rxy = avg4(adj.sxy, q, sx, sy);
        const float exx = s11 * txx + s12 * tyy + s13 * tzz;
        const float eyy = s12 * txx + s22 * tyy + s23 * tzz;
        const float ezz = s13 * txx + s23 * tyy + s33 * tzz;
        const float fxx = s11 * rxx + s12 * ryy + s13 * rzz;
        const float fyy = s12 * rxx + s22 * ryy + s23 * rzz;
        const float fzz = s13 * rxx + s23 * ryy + s33 * rzz;
        const double i11 = double(exx * fxx), i22 = double(eyy * fyy), i33 = double(ezz * fzz);
        const double i12 = double(exx * fyy) + double(eyy * fxx);
        const double i13 = double(exx * fzz) + double(ezz * fxx);
        const double i23 = double(eyy * fzz) + double(ezz * fyy);
        const double i44 = double((s44 * tyz) * (s44 * ryz)), i55 = double((s44 * txz) * (s44 * rxz)),
                     i66 = double((s44 * txy) * (s44 * rxy));
        const float rvx = adj.vx[q] + adj.vx[q + sx], rvy = adj.vy[q] + adj.vy[q + sy], rvz = adj.vz[q] + adj.vz[q + sz];
        const size_t o = size_t(S.g0 + e);
        gA[o] = gA[o] - (i11 + i12 + i13 + i22 + i23 + i33);
        gB[o] = gB[o] - (i44 + i55 + i66) + 2.0 * (i12 + i13 + i23);
        gC[o] = gC[o] + double(svx * rvx) + double(svy * rvy) + double(svz * rvz);
    }
}

// L2 waveform misfit and adjoint source per shot and component
// (compute_adjsrc_waveform, module_inversion_adjoint_source.f90:1020-1125,
// 371-383): the trace discard rule uses each component's own maximum trace
// norms; traces [nt][nr] per component; FP64 Neumaier sums.
__global__ __launch_bounds__(256) void ke_adjoint_source(const ShotE* __restrict__ shots,
                                                         const float* __restrict__ syn, const float* __restrict__ obs,
                                                         const float* __restrict__ rec_w, float* adj, double* misfit,
                                                         int nt, float threshold, double* scratch) {
    const ShotE S = shots[blockIdx.x];
    const int comp = blockIdx.y;
    const size_t base = S.trace0 + size_t(comp) * nt * S.nr;
    __shared__ double red[256];
    __shared__ double mo_s, ms_s;
    double* nobs = scratch + 2 * (3 * size_t(S.rec0) + size_t(comp) * S.nr);
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
    if (threadIdx.x == 0) misfit[3 * blockIdx.x + comp] = red[0];
}

// Sum of the per-shot window accumulators on the full model (nz, ny, nx),
// z fastest (module_gradient.f90:366-388): the three arrays in one pass.
__global__ void ke_merge(const ShotE* __restrict__ shots, const int* __restrict__ win, int nshot,
                         const double* __restrict__ gA, const double* __restrict__ gB, const double* __restrict__ gC,
                         double* fA, double* fB, double* fC, int nx, int ny, int nz) {
    const long long n = (long long)nx * ny * nz;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        const int kz = int(e % nz), jy = int((e / nz) % ny), ix = int(e / ((long long)nz * ny));
        double a = 0.0, b = 0.0, c = 0.0;
        for (int s = 0; s < nshot; ++s) {
            const ShotE& S = shots[s];
            const int i = ix - (win[3 * s] - 1), j = jy - (win[3 * s + 1] - 1), k = kz - (win[3 * s + 2] - 1);
            if (i < 0 || i >= S.nx || j < 0 || j >= S.ny || k < 0 || k >= S.nz) continue;
            const size_t o = size_t(S.g0) + i + size_t(j) * S.nx + size_t(k) * S.nx * S.ny;
            a += gA[o];
            b += gB[o];
            c += gC[o];
        }
        fA[e] = a;
        fB[e] = b;
        fC[e] = c;
    }
}

struct ShotHostE {
    int id = 0, nx = 0, ny = 0, nz = 0, nxbeg = 1, nybeg = 1, nzbeg = 1, hnt = 1;
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
    size_t field_elems = 0, trace_elems = 0, strip_elems = 0, grad_elems = 0;
    int nrec_total = 0, max_nr = 0, max_nadj = 0;
    double pts_full = 0;
    bool have_boundary = false, fwd_state_valid = false, have_obs = false;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0, last_updates = 0;
    Profiler prof;
    std::vector<float> vp_full, vs_full, rho_full;
    DevBuf<ShotE> d_shots;
    DevBuf<int> d_tile_shot, d_inj_beg;
    DevBuf<float> d_prof, d_amp, d_rec_w, d_traces;
    DevBuf<float> med[9];   // c11, c12, mu_yz, mu_xz, mu_xy, rho_x, rho_y, rho_z, c44 at the nodes
    DevBuf<float> fld[27];  // nine wavefields, eighteen memory variables
    DevBuf<Tap> d_inj;
    DevBuf<RecBox> d_rec_box;
    // gradient: boundary records, the other half of the reverse ping-pong,
    // the adjoint field, adjoint-source CSR, accumulators
    DevBuf<float> d_strips, d_adj, d_obs;
    DevBuf<unsigned> d_rscale;  // per (shot, step) boundary record: largest magnitude (16-bit storage)
    int rec_store = OWL_STORE_F32;
    DevBuf<float> fB[9];
    DevBuf<float> aF[27];
    DevBuf<int> d_adj_beg, d_win;
    DevBuf<Tap> d_adj_taps;
    DevBuf<double> d_gA, d_gB, d_gC, d_fA, d_fB, d_fC, d_misfit, d_scratch;
    // the forward time loop as one HIP graph (modelling without profiling
    // or boundary records), re-captured when a launch argument changes
    LoopGraph graph;
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
        fail(OWL_ERR_ARG, "data_dt/data_tmax must equal dt/tmax (resample the traces with owl_resample_traces)");
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
    long long off = 0, trace = 0, strip = 0, grad = 0;
    int prof = 0;
    std::vector<float> profs, amp, rec_w;
    std::vector<int> inj_beg{0};
    std::vector<Tap> inj;
    std::vector<RecBox> boxes;  // per (shot, component, receiver)
    std::vector<int> adj_beg{0};
    std::vector<Tap> adj_taps;
    b->max_nadj = 0;
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
        S.ntx = (H.nx + 2 * pml + (64 - 2 * R) - 1) / (64 - 2 * R);  // sweeps update 64 - 2R columns per wavefront
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
        {
            const EStrip g = estrip_geo(H.nx, H.ny, H.nz, R);
            S.strip0 = strip;
            strip += 3LL * 2 * (g.a + g.b + g.c) * nt;
        }
        S.g0 = grad;
        grad += (long long)H.nx * H.ny * H.nz;
        // sources; above the first node below a free surface they move to
        // z = dz with amplitude x 1.2 (module_vars.f90:650-655).  A force adds
        // to vx, vy, vz with amp / rho(gx, gy, gz); an explosion subtracts
        // amp / (dx dy dz) w from sxx, syy, szz at (gx, gy, gz), its wavelet
        // the time derivative (time_integration = -1, :664); amp = stf amp dt
        const float xmax = float(H.nx - 1) * cfg->dx + H.ox, ymax = float(H.ny - 1) * cfg->dy + H.oy,
                    zmax = float(H.nz - 1) * cfg->dz + H.oz;
        std::vector<Tap> st;
        for (const owl_source& sc0 : H.src) {
            owl_source sc = sc0;
            if (sc.mechanism != OWL_FORCE && sc.mechanism != OWL_EXPLOSION)
                fail(OWL_ERR_ARG, "the elastic solver takes force and explosion sources");
            if (b->fs && sc.z < cfg->dz) {
                sc.z = cfg->dz;
                sc.amp = sc.amp * 1.2f;
            }
            if (!(sc.x >= H.ox && sc.x <= xmax && sc.y >= H.oy && sc.y <= ymax && sc.z >= H.oz && sc.z <= zmax))
                fail(OWL_ERR_ARG, "source outside its shot window");
            const Located L = locate(3, sc.x, sc.y, sc.z, H.ox, H.oy, H.oz, cfg->dx, cfg->dy, cfg->dz);
            int hnt = 0;
            std::vector<float> stf = source_time_function(sc.wavelet, sc.f0, cfg->dt, f0factor, &hnt);
            if (sc.mechanism == OWL_EXPLOSION) stf = stf_time_derivative(stf, cfg->dt, nt);
            if (&sc0 == &H.src[0]) {
                H.hnt = hnt;
                S.hnt = std::max(hnt, 1);
            }
            const int nbeg = fortran_nint(sc.t0 / cfg->dt) + 1, nend = nbeg + int(stf.size()) - 1;
            if (sc.mechanism == OWL_EXPLOSION) {
                const int row = int(amp.size() / nt);
                for (int t = 1; t <= nt; ++t) {
                    const float a = (t >= nbeg && t <= nend) ? stf[t - nbeg] * sc.amp * cfg->dt : 0.0f;
                    amp.push_back(-(a / (cfg->dx * cfg->dy * cfg->dz)));
                }
                for (int c = 3; c < 6; ++c)
                    for (int irz = -kNkw; irz <= kNkw; ++irz)
                        for (int iry = -kNkw; iry <= kNkw; ++iry)
                            for (int irx = -kNkw; irx <= kNkw; ++irx) {
                                const float w1 = L.ix[irx + kNkw], w2 = L.iy[iry + kNkw], w3 = L.iz[irz + kNkw];
                                if (w1 == 0.f || w2 == 0.f || w3 == 0.f) continue;
                                if (b->fs && L.gz + irz < 2) continue;
                                st.push_back(Tap{(long long)cell_e(b, S, L.gx + irx, L.gy + iry, L.gz + irz), w1, w2,
                                                 w3, row, nbeg, nend, c});
                            }
                continue;
            }
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
        const size_t box0 = boxes.size();
        boxes.resize(box0 + 3 * size_t(S.nr), RecBox{});
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
            for (int c = 0; c < 3; ++c) {
                for (int irz = -kNkw; irz <= kNkw; ++irz)
                    for (int iry = -kNkw; iry <= kNkw; ++iry)
                        for (int irx = -kNkw; irx <= kNkw; ++irx) {
                            const float w1 = wx[c][irx + kNkw], w2 = wy[c][iry + kNkw], w3 = wz[c][irz + kNkw];
                            if (w1 == 0.f || w2 == 0.f || w3 == 0.f) continue;
                            comp[size_t(c) * S.nr + r].push_back(
                                Tap{(long long)cell_e(b, S, gx[c] + irx, gy[c] + iry, gz[c] + irz), w1, w2, w3, r, 0, 0,
                                    c});
                        }
                RecBox& B = boxes[box0 + size_t(c) * S.nr + r];
                const long long lo = 1 - b->pml - b->R;
                B.cell0 = S.off + (gx[c] - kNkw - lo) + (gy[c] - kNkw - lo) * (long long)S.px +
                          (gz[c] - kNkw - lo) * S.pxy;
                for (int k = 0; k < 2 * kNkw + 1; ++k) {
                    B.w[0][k] = wx[c][k];
                    B.w[1][k] = wy[c][k];
                    B.w[2][k] = wz[c][k];
                }
            }
        }
        // adjoint sources: the receivers' taps, below the free surface only
        // (submodule_fwi_gradient.f90:256-330), grouped by (component, cell)
        // in receiver order
        {
            std::vector<Tap> at;
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < S.nr; ++r)
                    for (const Tap& tp : comp[size_t(c) * S.nr + r]) {
                        const long long rel = tp.cell - S.off;
                        const int kk = int(rel / S.pxy) + lo;
                        if (b->fs && kk < 2) continue;
                        at.push_back(tp);
                    }
            std::stable_sort(at.begin(), at.end(), [](const Tap& x, const Tap& y) {
                return x.field != y.field ? x.field < y.field : x.cell < y.cell;
            });
            S.adj0 = int(adj_beg.size()) - 1;
            S.nadj = 0;
            for (size_t k = 0; k < at.size(); ++k) {
                if (k == 0 || at[k].cell != at[k - 1].cell || at[k].field != at[k - 1].field) {
                    if (k) adj_beg.push_back(int(adj_taps.size()));
                    ++S.nadj;
                }
                adj_taps.push_back(at[k]);
            }
            if (!at.empty()) adj_beg.push_back(int(adj_taps.size()));
            b->max_nadj = std::max(b->max_nadj, S.nadj);
        }
        b->nrec_total += S.nr;
        b->max_nr = std::max(b->max_nr, S.nr);
    }
    b->field_elems = size_t(off);
    if (b->field_elems >= (size_t(1) << 29))
        fail(OWL_ERR_ARG, "the shots' padded grids exceed 2 GiB per field; split the batch");
    b->trace_elems = size_t(trace);
    b->strip_elems = size_t(strip);
    b->grad_elems = size_t(grad);
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
    // receiver boxes in (shot, component, receiver) order: entry 3*rec0 + c*nr + r
    if (boxes.empty()) boxes.push_back(RecBox{});
    b->d_rec_box.upload(boxes, stm);
    b->d_rec_w.upload(rec_w, stm);
    b->d_adj_beg.upload(adj_beg, stm);
    if (adj_taps.empty()) adj_taps.push_back(Tap{});
    b->d_adj_taps.upload(adj_taps, stm);
    {
        std::vector<int> win;
        for (auto& H : b->hs) win.insert(win.end(), {H.nxbeg, H.nybeg, H.nzbeg});
        b->d_win.upload(win, stm);
    }
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
    const std::vector<float>* media[9] = {&c11, &c12, &myz, &mxz, &mxy, &rx, &ry, &rz, &c44};
    for (int k = 0; k < 9; ++k) b->med[k].upload(*media[k], stm);
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

EV out_of(const EF& F) { return EV{F.sxx, F.syy, F.szz, F.syz, F.sxz, F.sxy, F.vx, F.vy, F.vz}; }

EV ev_of(DevBuf<float>* f) {
    return EV{f[0].get(), f[1].get(), f[2].get(), f[3].get(), f[4].get(), f[5].get(), f[6].get(), f[7].get(), f[8].get()};
}

EF with_fields(EF F, const EV& V) {
    F.sxx = V.sxx; F.syy = V.syy; F.szz = V.szz; F.syz = V.syz; F.sxz = V.sxz; F.sxy = V.sxy;
    F.vx = V.vx; F.vy = V.vy; F.vz = V.vz;
    return F;
}

void maybe_collect(owl_elastic3d* b) {
    if (b->prof.pending() > 20000) {
        OWL_HIP(hipStreamSynchronize(b->stream));
        b->prof.collect();
    }
}

// One time step of the wavefields F (memory variables in F) into O (O may
// be F's own fields): [free surface: zero v above it] -> stresses ->
// [mirror stresses] -> velocities (module_wavefield.f90:52-1090).
template <int R>
void step_e(owl_elastic3d* b, const EF& F, const EV& O, float dt, double bs, double bv) {
    hipStream_t st = b->stream;
    const unsigned nshot = unsigned(b->sd.size());
    if (64 * b->cy > 256) fail(OWL_ERR_ARG, "launch shape exceeds the sweeps' 256-thread blocks");
    const int ntile = int(b->tile_shot.size());
    const dim3 blk(64, unsigned(b->cy)), grd(unsigned((ntile + 7) / 8 * 8));
    const EM M = media(b);
    if (b->fs) hipLaunchKernelGGL(ke_surface<R>, dim3(256, nshot), dim3(256), 0, st, b->d_shots.get(), F, b->pml, 0);
    int ev = b->prof.begin(st, 0, bs);
    hipLaunchKernelGGL(ke_stress<R>, grd, blk, 0, st, b->d_shots.get(), b->d_tile_shot.get(), F, O, M, b->d_prof.get(),
                       b->coef, dt, b->idx, b->idy, b->idz, b->pml, b->kc, b->fs, ntile);
    b->prof.end(st, ev);
    if (b->fs)
        hipLaunchKernelGGL(ke_surface<R>, dim3(256, nshot), dim3(256), 0, st, b->d_shots.get(), with_fields(F, O),
                           b->pml, 1);
    ev = b->prof.begin(st, 1, bv);
    hipLaunchKernelGGL(ke_velocity<R>, grd, blk, 0, st, b->d_shots.get(), b->d_tile_shot.get(), F, O, M,
                       b->d_prof.get(), b->coef, dt, b->idx, b->idy, b->idz, b->pml, b->kc, b->fs, ntile);
    b->prof.end(st, ev);
}

void strips_e(owl_elastic3d* b, int R, float* vx, float* vy, float* vz, int t, int save);

// Every device address and launch parameter the forward loop passes to its
// kernels: a captured loop is replayed only while these are unchanged.
std::vector<uintptr_t> forward_signature_e(const owl_elastic3d* b) {
    std::vector<uintptr_t> v{uintptr_t(b->R), uintptr_t(b->nt), uintptr_t(b->kc), uintptr_t(b->cy),
                             uintptr_t(b->tile_shot.size()), uintptr_t(b->max_nr), uintptr_t(b->fs),
                             uintptr_t(float_bits(b->dt)), uintptr_t(float_bits(b->idx)), uintptr_t(float_bits(b->idy)),
                             uintptr_t(float_bits(b->idz)), uintptr_t(b->d_shots.get()),
                             uintptr_t(b->d_tile_shot.get()), uintptr_t(b->d_rec_box.get()),
                             uintptr_t(b->d_rec_w.get()), uintptr_t(b->d_traces.get()),
                             uintptr_t(b->d_inj_beg.get()), uintptr_t(b->d_inj.get()), uintptr_t(b->d_amp.get()),
                             uintptr_t(b->d_prof.get())};
    for (auto& f : b->fld) v.push_back(uintptr_t(f.get()));
    for (auto& f : b->med) v.push_back(uintptr_t(f.get()));
    return v;
}

template <int R>
void run_forward_e(owl_elastic3d* b, bool save_boundary) {
    hipStream_t st = b->stream;
    for (auto& f : b->fld) f.zero(st);
    b->d_traces.zero(st);
    if (save_boundary) alloc_records(b->d_strips, b->d_rscale, b->strip_elems, b->rec_store, b->sd.size(), b->nt, st);
    const EF F = fields(b);
    const unsigned nshot = unsigned(b->sd.size());
    double bs, bv;
    sweep_bytes(b, bs, bv);
    const unsigned rg = unsigned(std::max(1, (3 * b->max_nr + 255) / 256));
    // step t: [boundary record of the state at the start of the step] ->
    // [free surface: mirror v above it] -> record(t) -> update -> source
    auto loop = [&] {
        for (int t = 1; t <= b->nt; ++t) {
            if (save_boundary) strips_e(b, R, F.vx, F.vy, F.vz, t, 1);
            if (b->fs)
                hipLaunchKernelGGL(ke_surface<R>, dim3(256, nshot), dim3(256), 0, st, b->d_shots.get(), F, b->pml, 2);
            int ev = b->prof.begin(st, 2, 0);
            hipLaunchKernelGGL(ke_record, dim3(rg, nshot), dim3(256), 0, st, b->d_shots.get(), F, b->d_rec_box.get(),
                               b->d_rec_w.get(), b->d_traces.get(), b->nt, t);
            b->prof.end(st, ev);
            if (t == b->nt) {
                // the last update changes nothing that is recorded; the state
                // at the start of step nt is the reverse pass's final state,
                // whose velocities above a free surface are zero (the mirror
                // undone)
                if (b->fs)
                    hipLaunchKernelGGL(ke_surface<R>, dim3(256, nshot), dim3(256), 0, st, b->d_shots.get(), F, b->pml,
                                       3);
                break;
            }
            step_e<R>(b, F, out_of(F), b->dt, bs, bv);
            ev = b->prof.begin(st, 2, 0);
            hipLaunchKernelGGL(ke_inject, dim3(nshot), dim3(1024), 0, st, b->d_shots.get(), F, b->d_inj_beg.get(),
                               b->d_inj.get(), b->d_amp.get(), b->nt, t);
            b->prof.end(st, ev);
            if (!b->graph.capturing) maybe_collect(b);
        }
    };
    const bool graph_ok = b->graph.enabled && !b->prof.on() && !save_boundary;
    OWL_HIP(hipEventRecord(b->ev0, st));
    if (graph_ok)
        b->graph.run(st, forward_signature_e(b), loop);
    else
        loop();
    OWL_HIP(hipGetLastError());
    OWL_HIP(hipEventRecord(b->ev1, st));
    OWL_HIP(hipStreamSynchronize(st));
    b->prof.collect();
    float ms = 0.f;
    OWL_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_ms = ms;
    b->last_updates = double(b->nt - 1) * b->pts_full;
    b->have_boundary = b->fwd_state_valid = false;  // until the records are known to be usable
    if (save_boundary) check_records_finite(b->d_rscale, b->rec_store, b->sd.size() * size_t(b->nt));
    b->have_boundary = save_boundary;
    b->fwd_state_valid = save_boundary;
}

void strips_e(owl_elastic3d* b, int R, float* vx, float* vy, float* vz, int t, int save) {
    size_t mx = 1;
    for (auto& S : b->sd) {
        const EStrip g = estrip_geo(S.nx, S.ny, S.nz, R);
        mx = std::max<size_t>(mx, size_t(6 * (g.a + g.b + g.c)));
    }
    const dim3 grd(unsigned(std::min<size_t>(4096, (mx + 255) / 256)), unsigned(b->sd.size()));
    const RecStore st{b->d_strips.get(), b->d_rscale.get(), b->rec_store, b->nt};
    const int ev = b->prof.begin(b->stream, 4, 0);
    dispatch_Re(R, [&](auto r) {
        if (save && b->rec_store != OWL_STORE_F32)
            hipLaunchKernelGGL(ke_strips<decltype(r)::value>, grd, dim3(256), 0, b->stream, b->d_shots.get(), vx, vy,
                               vz, st, t, b->pml, int(kRecAmax));
        hipLaunchKernelGGL(ke_strips<decltype(r)::value>, grd, dim3(256), 0, b->stream, b->d_shots.get(), vx, vy, vz,
                           st, t, b->pml, save ? int(kRecSave) : int(kRecRestore));
    });
    b->prof.end(b->stream, ev);
}

// OWL's reverse pass (compute_adjoint, elastic_vhtiort_3d/
// submodule_fwi_gradient.f90): t = nt .. hnt, prev = state; state =
// update(state, -dt); at t = nt the final state; boundary record t into the
// state; adjoint = update(adjoint, -dt); adjoint sources; imaging.  The
// forward field steps out of place (the previous state stays in the other
// buffer for the imaging); at t = nt the update of the zero field is the
// zero field, so the final state is used as is with prev = 0.
template <int R>
void run_reverse_e(owl_elastic3d* b) {
    hipStream_t st = b->stream;
    if (!b->have_boundary || !b->fwd_state_valid)
        fail(OWL_ERR_STATE, "the adjoint needs a preceding forward run with boundary saving");
    if (!b->d_adj.get() || b->d_adj.size() < b->trace_elems) fail(OWL_ERR_STATE, "no adjoint source");
    // OWL's loop ends at the first source's hnt
    for (int k = 9; k < 27; ++k) b->fld[k].zero(st);  // the reverse run's memory variables start at zero
    for (auto& f : b->fB) {
        f.alloc(b->field_elems);
        f.zero(st);
    }
    for (auto& f : b->aF) {
        f.alloc(b->field_elems);
        f.zero(st);
    }
    const size_t ng = std::max<size_t>(b->grad_elems, 1);
    for (auto* g : {&b->d_gA, &b->d_gB, &b->d_gC}) {
        g->alloc(ng);
        g->zero(st);
    }
    const unsigned nshot = unsigned(b->sd.size());
    // each shot is imaged only while t >= its own hnt; the batch runs down to
    // the smallest one
    int hnt = b->nt;
    for (auto& H : b->hs) hnt = std::min(hnt, std::max(H.hnt, 1));
    const int cc = b->cfg.cc_step_interval > 0 ? b->cfg.cc_step_interval : 1;
    double bs, bv;
    sweep_bytes(b, bs, bv);
    EF X = fields(b);              // current state (with memory variables)
    EV Y = ev_of(b->fB);           // the other half of the ping-pong
    EF A;
    {
        float** p = &A.sxx;
        for (int k = 0; k < 27; ++k) p[k] = b->aF[k].get();
    }
    const EV adj = out_of(A);
    // imaging: 27 wavefields (three with up to 3 extra neighbours, mostly
    // cached), c11, c12, c44, three FP64 accumulators read and written
    const double image_bytes = (27.0 * 4 + 12.0 + 48.0) * double(b->grad_elems);
    const unsigned ig = unsigned(std::max<size_t>(1, std::min<size_t>(2048, (b->grad_elems / nshot + 255) / 256)));
    OWL_HIP(hipEventRecord(b->ev0, st));
    for (int t = b->nt; t >= hnt; --t) {
        EV cur = out_of(X), prev = out_of(X);
        if (t != b->nt) {
            step_e<R>(b, X, Y, -b->dt, bs, bv);
            cur = Y;
        }
        strips_e(b, R, cur.vx, cur.vy, cur.vz, t, 0);
        step_e<R>(b, A, adj, -b->dt, bs, bv);
        int ev = b->prof.begin(st, 2, 0);
        if (b->max_nadj > 0)
            hipLaunchKernelGGL(ke_inject_adjoint, dim3(unsigned((b->max_nadj + 255) / 256), nshot), dim3(256), 0, st,
                               b->d_shots.get(), adj.vx, adj.vy, adj.vz, b->d_adj_beg.get(), b->d_adj_taps.get(),
                               b->d_rec_w.get(), b->d_adj.get(), b->nt, t);
        b->prof.end(st, ev);
        if (t % cc == 0) {
            ev = b->prof.begin(st, 3, image_bytes);
            hipLaunchKernelGGL(ke_image, dim3(ig, nshot), dim3(256), 0, st, b->d_shots.get(), cur, prev, adj,
                               b->med[0].get(), b->med[1].get(), b->med[8].get(), b->d_gA.get(), b->d_gB.get(),
                               b->d_gC.get(), t == b->nt ? 1 : 0, b->pml, R, t);
            b->prof.end(st, ev);
        }
        if (t != b->nt) {
            // the new state becomes current; its old buffers take the next step
            const EV old = out_of(X);
            X = with_fields(X, Y);
            Y = old;
        }
        maybe_collect(b);
    }
    OWL_HIP(hipGetLastError());
    OWL_HIP(hipEventRecord(b->ev1, st));
    OWL_HIP(hipStreamSynchronize(st));
    b->prof.collect();
    float ms = 0.f;
    OWL_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_ms = ms;
    b->last_updates = 2.0 * double(b->nt - hnt) * b->pts_full;
    b->fwd_state_valid = false;
}

void traces_to_host_e(const owl_elastic3d* b, const float* dev, float* out) {
    std::vector<float> h(b->trace_elems);
    OWL_HIP(hipMemcpy(h.data(), dev, h.size() * sizeof(float), hipMemcpyDeviceToHost));
    // device [3][nt][nr] per shot -> host [3][nr][nt] per shot
    for (auto& S : b->sd)
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < S.nr; ++r)
                for (int t = 0; t < b->nt; ++t)
                    out[size_t(S.trace0) + (size_t(c) * S.nr + r) * b->nt + t] =
                        h[size_t(S.trace0) + size_t(c) * b->nt * S.nr + size_t(t) * S.nr + r];
}

void traces_to_device_e(owl_elastic3d* b, const float* in, DevBuf<float>& dev) {
    std::vector<float> h(b->trace_elems);
    for (auto& S : b->sd)
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < S.nr; ++r)
                for (int t = 0; t < b->nt; ++t)
                    h[size_t(S.trace0) + size_t(c) * b->nt * S.nr + size_t(t) * S.nr + r] =
                        in[size_t(S.trace0) + (size_t(c) * S.nr + r) * b->nt + t];
    dev.alloc(std::max<size_t>(h.size(), 1));
    dev.upload(h.data(), h.size(), b->stream);
    OWL_HIP(hipStreamSynchronize(b->stream));
}

// Chain rule (submodule_fwi_gradient.f90:486-495) with vp, vs in km/s and
// rho in g/cm^3: grad_vp = 2 rho vp A, grad_vs = 2 rho vs B, grad_rho = C.
inline void chain_rule(float vp, float vs, float rho, double A, double B, double C, double& gv, double& gs,
                       double& gr) {
    const double r = double(rho * 1.0e-3f), p = double(vp * 1.0e-3f), v = double(vs * 1.0e-3f);
    gv = 2.0 * r * p * A;
    gs = 2.0 * r * v * B;
    gr = C;
}

// Tiles of the current launch shape (kc z planes, cy y rows per tile) and
// their shots, uploaded on the batch's stream.
void retile_e(owl_elastic3d* b) {
    int lt = 0;
    b->tile_shot.clear();
    for (size_t s = 0; s < b->sd.size(); ++s) {
        ShotE& S = b->sd[s];
        S.nty = (S.ny + 2 * b->pml + b->cy - 1) / b->cy;
        S.ntx = (S.nx + 2 * b->pml + (64 - 2 * b->R) - 1) / (64 - 2 * b->R);
        S.nzc = (S.nz + 2 * b->pml + b->kc - 1) / b->kc;
        S.tile0 = lt;
        lt += S.ntx * S.nty * S.nzc;
        b->tile_shot.insert(b->tile_shot.end(), size_t(S.ntx) * S.nty * S.nzc, int(s));
    }
    b->d_shots.upload(b->sd, b->stream);
    b->d_tile_shot.upload(b->tile_shot, b->stream);
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

owl_status owl_elastic3d_forward(owl_elastic3d* b, float* traces, int save_boundary) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        dispatch_Re(b->R, [&](auto r) { run_forward_e<decltype(r)::value>(b, save_boundary != 0); });
        if (traces) traces_to_host_e(b, b->d_traces.get(), traces);
    });
}

owl_status owl_elastic3d_adjoint(owl_elastic3d* b, const float* adj, float* grad_vp, float* grad_vs,
                                 float* grad_rho) {
    return guarded([&] {
        if (!b || !adj) fail(OWL_ERR_ARG, "null argument");
        traces_to_device_e(b, adj, b->d_adj);
        dispatch_Re(b->R, [&](auto r) { run_reverse_e<decltype(r)::value>(b); });
        std::vector<double> gA(b->grad_elems), gB(b->grad_elems), gC(b->grad_elems);
        OWL_HIP(hipMemcpy(gA.data(), b->d_gA.get(), gA.size() * sizeof(double), hipMemcpyDeviceToHost));
        OWL_HIP(hipMemcpy(gB.data(), b->d_gB.get(), gB.size() * sizeof(double), hipMemcpyDeviceToHost));
        OWL_HIP(hipMemcpy(gC.data(), b->d_gC.get(), gC.size() * sizeof(double), hipMemcpyDeviceToHost));
        // per-shot window gradients (nz_s, ny_s, nx_s), z fastest, NaN/Inf -> 0
        for (size_t s = 0; s < b->sd.size(); ++s) {
            const ShotE& S = b->sd[s];
            const ShotHostE& H = b->hs[s];
            for (int k = 0; k < S.nz; ++k)
                for (int j = 0; j < S.ny; ++j)
                    for (int i = 0; i < S.nx; ++i) {
                        const size_t src = size_t(S.g0) + i + size_t(j) * S.nx + size_t(k) * S.nx * S.ny;
                        const size_t dst = size_t(S.g0) + k + size_t(j) * S.nz + size_t(i) * S.nz * S.ny;
                        const size_t m = model_index(b, H, i + 1, j + 1, k + 1);
                        double gv, gs, gr;
                        chain_rule(b->vp_full[m], b->vs_full[m], b->rho_full[m], gA[src], gB[src], gC[src], gv, gs,
                                   gr);
                        const float a = float(gv), c = float(gs), e = float(gr);
                        if (grad_vp) grad_vp[dst] = std::isfinite(a) ? a : 0.f;
                        if (grad_vs) grad_vs[dst] = std::isfinite(c) ? c : 0.f;
                        if (grad_rho) grad_rho[dst] = std::isfinite(e) ? e : 0.f;
                    }
        }
    });
}

owl_status owl_elastic3d_gradient(owl_elastic3d* b, const float* obs, double* misfit, double* grad_vp,
                                  double* grad_vs, double* grad_rho, float* traces_out) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        if (obs) {
            traces_to_device_e(b, obs, b->d_obs);
            b->have_obs = true;
        }
        if (!b->have_obs) fail(OWL_ERR_STATE, "no observed data");
        dispatch_Re(b->R, [&](auto r) { run_forward_e<decltype(r)::value>(b, true); });
        const double fwd_ms = b->last_ms, fwd_up = b->last_updates;
        hipStream_t st = b->stream;
        const unsigned nshot = unsigned(b->sd.size());
        b->d_adj.alloc(std::max<size_t>(b->trace_elems, 1));
        b->d_misfit.alloc(3 * size_t(nshot));
        b->d_scratch.alloc(std::max<size_t>(6 * size_t(b->nrec_total), 1));
        hipLaunchKernelGGL(ke_adjoint_source, dim3(nshot, 3), dim3(256), 0, st, b->d_shots.get(), b->d_traces.get(),
                           b->d_obs.get(), b->d_rec_w.get(), b->d_adj.get(), b->d_misfit.get(), b->nt, b->cfg.trace_discard_threshold,
                           b->d_scratch.get());
        dispatch_Re(b->R, [&](auto r) { run_reverse_e<decltype(r)::value>(b); });
        const size_t nfull = size_t(b->cfg.nx) * b->cfg.ny * b->cfg.nz;
        for (auto* g : {&b->d_fA, &b->d_fB, &b->d_fC}) g->alloc(nfull);
        const unsigned mg = unsigned(std::min<size_t>(8192, (nfull + 255) / 256));
        hipLaunchKernelGGL(ke_merge, dim3(mg), dim3(256), 0, st, b->d_shots.get(), b->d_win.get(), int(nshot),
                           b->d_gA.get(), b->d_gB.get(), b->d_gC.get(), b->d_fA.get(), b->d_fB.get(), b->d_fC.get(),
                           b->cfg.nx, b->cfg.ny, b->cfg.nz);
        OWL_HIP(hipGetLastError());
        OWL_HIP(hipStreamSynchronize(st));
        if (misfit) {
            std::vector<double> m3(3 * size_t(nshot));
            OWL_HIP(hipMemcpy(m3.data(), b->d_misfit.get(), m3.size() * sizeof(double), hipMemcpyDeviceToHost));
            for (unsigned s = 0; s < nshot; ++s) misfit[s] = m3[3 * s] + m3[3 * s + 1] + m3[3 * s + 2];
        }
        if (grad_vp || grad_vs || grad_rho) {
            std::vector<double> fA(nfull), fB(nfull), fC(nfull);
            OWL_HIP(hipMemcpy(fA.data(), b->d_fA.get(), nfull * sizeof(double), hipMemcpyDeviceToHost));
            OWL_HIP(hipMemcpy(fB.data(), b->d_fB.get(), nfull * sizeof(double), hipMemcpyDeviceToHost));
            OWL_HIP(hipMemcpy(fC.data(), b->d_fC.get(), nfull * sizeof(double), hipMemcpyDeviceToHost));
            for (size_t e = 0; e < nfull; ++e) {
                double gv, gs, gr;
                chain_rule(b->vp_full[e], b->vs_full[e], b->rho_full[e], fA[e], fB[e], fC[e], gv, gs, gr);
                if (grad_vp) grad_vp[e] = std::isfinite(gv) ? gv : 0.0;
                if (grad_vs) grad_vs[e] = std::isfinite(gs) ? gs : 0.0;
                if (grad_rho) grad_rho[e] = std::isfinite(gr) ? gr : 0.0;
            }
        }
        if (traces_out) traces_to_host_e(b, b->d_traces.get(), traces_out);
        b->last_ms += fwd_ms;
        b->last_updates += fwd_up;
    });
}

owl_status owl_elastic3d_set_launch(owl_elastic3d* b, int planes, int rows) {
    return guarded([&] {
        // blocks of 64 x rows threads; the sweeps are compiled for at most 256
        if (!b || planes < 1 || (rows != 1 && rows != 2 && rows != 4)) fail(OWL_ERR_ARG, "bad launch shape");
        b->kc = planes;
        b->cy = rows;
        retile_e(b);
        OWL_HIP(hipStreamSynchronize(b->stream));
    });
}

// Launch-shape autotuner: times `steps` steps of the sweeps on the batch
// under each candidate (z planes per tile, y rows per tile) and keeps the
// fastest; the fields are zeroed afterwards (every forward starts from
// zero).  Stencil timing does not depend on the field values.
owl_status owl_elastic3d_autotune(owl_elastic3d* b, int steps, int* planes, int* rows) {
    return guarded([&] {
        if (!b || steps < 1) fail(OWL_ERR_ARG, "null batch or no steps");
        static const int kPlanes[] = {8, 12, 16, 32, 64, 128, 256};
        static const int kRows[] = {1, 2, 4};
        hipStream_t st = b->stream;
        const bool prof_on = b->prof.on();
        b->prof.enable(false);
        for (auto& f : b->fld) f.zero(st);
        const EF F = fields(b);
        double bs, bv;
        sweep_bytes(b, bs, bv);
        float best = -1.f;
        int bk = b->kc, by = b->cy;
        for (int kc : kPlanes)
            for (int cy : kRows) {
                b->kc = kc;
                b->cy = cy;
                retile_e(b);
                dispatch_Re(b->R, [&](auto r) {
                    step_e<decltype(r)::value>(b, F, out_of(F), b->dt, bs, bv);  // warm-up
                    OWL_HIP(hipEventRecord(b->ev0, st));
                    for (int t = 0; t < steps; ++t) step_e<decltype(r)::value>(b, F, out_of(F), b->dt, bs, bv);
                    OWL_HIP(hipEventRecord(b->ev1, st));
                });
                OWL_HIP(hipStreamSynchronize(st));
                float ms = 0.f;
                OWL_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
                if (best < 0 || ms < best) {
                    best = ms;
                    bk = kc;
                    by = cy;
                }
            }
        b->kc = bk;
        b->cy = by;
        retile_e(b);
        for (auto& f : b->fld) f.zero(st);
        OWL_HIP(hipStreamSynchronize(st));
        b->prof.enable(prof_on);
        b->fwd_state_valid = false;
        b->have_boundary = false;
        if (planes) *planes = bk;
        if (rows) *rows = by;
    });
}

owl_status owl_elastic3d_set_record_storage(owl_elastic3d* b, int kind) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        check_record_kind(kind);
        if (kind != b->rec_store) {
            b->rec_store = kind;
            b->d_strips.release();
            b->have_boundary = false;
            b->fwd_state_valid = false;
        }
    });
}

owl_status owl_elastic3d_set_graphs(owl_elastic3d* b, int on) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        b->graph.enabled = on != 0;
        b->graph.reset();
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
