// 2D isotropic elastic SSG-FD velocity-stress time stepping and OWL's
// gradient for MI355X (gfx950): OWL's elastic_vhtiort_2d for
// anisotropy_type = 'iso' (module_wavefield.f90:42-134,
// module_cfspml.f90:285-578, submodule_forward_modeling.f90:31-180,
// module_boundary_saving.f90:53-160, submodule_fwi_gradient.f90) for a batch
// of shots, each on its own window, in every launch.  Force and explosion
// sources, two-component velocity receivers, CFS-MPML boundaries (zero
// damping ratios for isotropic media, so per-axis profiles), and OWL's free
// surface on its depth-refined mesh (update_wavefield_free_surface,
// module_wavefield.f90:143-289; module_vars.f90:310-371) as an
// instantiation of the two sweeps with per-row dz scalings.
//
// Layout: per shot a padded grid, x fastest, rows padded to a multiple of 64
// floats.  Each sweep tiles x with 64-wide thread rows and marches along z
// with per-thread register queues for the z stencils.  FP32 in the
// reference's operation order (no FMA contraction); gradients in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "common.hpp"
#include "fs_mesh.hpp"
#include "host_prep.hpp"
#include "owl_amd.h"
#include "profiler.hpp"
#include "record_store.hpp"
#include "trace_misfit.hpp"

namespace owl {
namespace e2 {

constexpr int kPitchAlign = 64;

struct Coef {
    float c[8];
};

struct ShotE2 {
    int nx, nz;
    int px, nzp;           // row pitch (floats), padded rows
    long long off;         // element offset of node (lo, lo)
    int tile0, ntx, nzc;
    int prof;              // profiles: x (ai bi ah bh), z (ai bi ah bh)
    int nr, rec0;
    int ninj, inj0;
    int nadj, adj0;
    long long trace0;      // two [nt][nr] blocks (vx, vz)
    long long strip0;      // boundary records: nt blocks
    long long g0;          // window gradient accumulators (nx*nz, x fastest)
    int hnt;               // last reverse step OWL images for this shot (its first source's hnt)
    int sc0;               // free surface: dz/dz_i, dz/dz_h per row j at [sc0 + j + pml], [sc0 + nrj + j + pml]
};

// Wavefields and memory variables (named by derivative and field as in OWL).
struct EF2 {
    float *sxx, *szz, *sxz, *vx, *vz;
    float *m_dxvx, *m_dzvz, *m_dxvz, *m_dzvx, *m_dxxx, *m_dzxz, *m_dxxz, *m_dzzz;
};

// The five wavefields a sweep writes (the stress sweep reads the old
// stresses and the velocities from EF2 and writes the new stresses here;
// the velocity sweep reads the new stresses from here).
struct EV2 {
    float *sxx, *szz, *sxz, *vx, *vz;
};

// c11 (= c33), c13, the 4-point harmonic c55 at the shear-stress position,
// the 2-point densities at the velocity positions.
struct EM2 {
    const float *c11, *c13, *mu, *rx, *rz;
};

struct Tap {
    long long cell;
    float w1, w2;
    int row;
    int tbeg, tend;
    int field;  // 0 vx, 1 vz, 2 sxx, 3 szz
};

struct Prof2 {
    const float *axi, *bxi, *axh, *bxh, *azi, *bzi, *azh, *bzh;
};

__device__ __forceinline__ Prof2 prof_of(const float* prof, const ShotE2& S, int pml) {
    const int lx = S.nx + 2 * pml, lz = S.nz + 2 * pml;
    const float* p = prof + S.prof;
    return Prof2{p, p + lx, p + 2 * lx, p + 3 * lx, p + 4 * lx, p + 4 * lx + lz, p + 4 * lx + 2 * lz,
                 p + 4 * lx + 3 * lz};
}

__device__ __forceinline__ size_t at2(const ShotE2& S, int lo, int i, int j) {
    return size_t(S.off) + size_t(i - lo) + size_t(j - lo) * size_t(S.px);
}

// Load and store at a 32-bit byte offset from a kernel-argument base (the
// batch's fields stay below 2 GiB, checked on the host): a sweep addresses
// every array with a few shared offset registers against scalar bases.
__device__ __forceinline__ float ldo(const float* base, unsigned off) {
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + off);
}
__device__ __forceinline__ void sto(float* base, unsigned off, float v) {
    *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + off) = v;
}

// Memory-variable update m = a d + b m, d = (d + m) / kappa with kappa = 1
// (module_wavefield.f90:60-66).
__device__ __forceinline__ float mpml(float* m, unsigned qb, float d, float a, float b) {
    const float v = a * d + b * ldo(m, qb);
    sto(m, qb, v);
    return d + v;
}

// x stencil sum_m c_m (w(i+off+m) - w(i+off-m+1)) from the lanes' values of
// the current row: lane l holds column i and reads columns
// i - R + off + 1 .. i + R + off from lanes l - R + off + 1 .. l + R + off.
template <int R>
__device__ __forceinline__ float sten_lanes(float v, int l, int off, const float* c) {
    float xv[2 * R];
#pragma unroll
    for (int k = 0; k < 2 * R; ++k) xv[k] = __shfl(v, l - R + off + 1 + k, 64);
    float s = c[0] * (xv[R] - xv[R - 1]);
#pragma unroll
    for (int m = 2; m <= R; ++m) s = s + c[m - 1] * (xv[R - 1 + m] - xv[R - m]);
    return s;
}

// ---------------------------------------------------------------------------
// Stress sweep: sxx, szz at (i, j), sxz at storage (i, j) = position
// (i - 1/2, j - 1/2) (module_wavefield.f90:52-93).  z stencils from queues
// over rows j-R .. j+R of vx and vz; x stencils from the other lanes' current
// rows by shuffles, so a wavefront of 64 lanes updates the middle 64 - 2R
// columns and the others only feed them.
// ---------------------------------------------------------------------------
template <int R, bool FS>
__global__ __launch_bounds__(256) void k2e_stress(const ShotE2* __restrict__ shots, const int* __restrict__ tile_shot,
                                                  EF2 F, EV2 O, EM2 M, const float* __restrict__ prof,
                                                  const float* __restrict__ scal, Coef c, float dt, float idx,
                                                  float idz, int pml, int kc) {
    constexpr int W = 64 - 2 * R;
    // one wavefront per tile (blockDim.x = 64): tile, shot and row indices
    // are wavefront-uniform and live in scalar registers
    const int tile = __builtin_amdgcn_readfirstlane(int(blockIdx.x) * int(blockDim.y) + int(threadIdx.y));
    const ShotE2 S = shots[tile_shot[tile]];
    int lt = tile - S.tile0;
    const int tx = lt % S.ntx, tz = lt / S.ntx;
    const int nx = S.nx, nz = S.nz, lo = 1 - pml - R;
    const int l = int(threadIdx.x);
    const int i = -pml + 1 + tx * W - R + l;
    // with the free surface the rows start at the surface row 1
    const int j0 = (FS ? 1 : -pml + 1) + tz * kc, j1 = min(j0 + kc - 1, nz + pml);
    if (j0 > j1) return;  // padding tile beyond the shot's rows (whole wavefront)
    const bool mcol = i <= nx + pml + R;
    const bool own = l >= R && l <= 63 - R && i <= nx + pml;
    const Prof2 P = prof_of(prof, S, pml);
    const float* __restrict__ sci = scal + S.sc0 + pml;
    const float* __restrict__ sch = sci + (nz + 2 * pml + 2);
    const float* __restrict__ vx = F.vx;
    const float* __restrict__ vz = F.vz;
    const unsigned szb = unsigned(S.px) * 4u;
    const int ii = i + pml - 1;
    unsigned q = unsigned(at2(S, lo, mcol ? i : lo, j0)) * 4u;
    float qx[2 * R + 1], qz[2 * R + 1];  // rows j + m - R at slot m
#pragma unroll
    for (int m = 0; m < 2 * R; ++m) {
        const unsigned r = q + unsigned(m - R) * szb;
        qx[m] = mcol ? ldo(vx, r) : 0.0f;
        qz[m] = mcol ? ldo(vz, r) : 0.0f;
    }
    const bool xin = !(i <= 0 || i >= nx + 1), Xin = !(i <= 1 || i >= nx + 1);
    for (int j = j0; j <= j1; ++j, q += szb) {
        {
            const unsigned r = q + unsigned(R) * szb;
            qx[2 * R] = mcol ? ldo(vx, r) : 0.0f;
            qz[2 * R] = mcol ? ldo(vz, r) : 0.0f;
        }
        const float sxvx = sten_lanes<R>(qx[R], l, 0, c.c);   // d+x vx
        const float sxvz = sten_lanes<R>(qz[R], l, -1, c.c);  // d-x vz
        if (own) {
            const int jj = j + pml - 1;
            {
                const float c11 = ldo(M.c11, q), c13 = ldo(M.c13, q);
                float pdxvx = idx * sxvx, pdzvz;
                if (FS && j == 1) {
                    // the surface row: dz vz from the traction condition, x damping only
                    if (!xin) pdxvx = mpml(F.m_dxvx, q, pdxvx, P.axi[ii], P.bxi[ii]);
                    pdzvz = -(c13 / c11 * pdxvx);
                } else {
                    float b = c.c[0] * (qz[R + 1] - qz[R]);
#pragma unroll
                    for (int m = 2; m <= R; ++m) b = b + c.c[m - 1] * (qz[R + m] - qz[R - m + 1]);
                    pdzvz = FS ? idz * b * sci[j] : idz * b;
                    if (!xin || (!FS && j <= 0) || j >= nz + 1) {
                        pdxvx = mpml(F.m_dxvx, q, pdxvx, P.axi[ii], P.bxi[ii]);
                        pdzvz = mpml(F.m_dzvz, q, pdzvz, P.azi[jj], P.bzi[jj]);
                    }
                }
                sto(O.sxx, q, ldo(F.sxx, q) + dt * (c11 * pdxvx + c13 * pdzvz));
                sto(O.szz, q, ldo(F.szz, q) + dt * (c13 * pdxvx + c11 * pdzvz));
            }
            if (!FS || j >= 2) {
                float a = c.c[0] * (qx[R] - qx[R - 1]);        // d-z vx
#pragma unroll
                for (int m = 2; m <= R; ++m) a = a + c.c[m - 1] * (qx[R + m - 1] - qx[R - m]);
                float pdzvx = FS ? idz * a * sch[j] : idz * a, pdxvz = idx * sxvz;
                if (!Xin || (!FS && j <= 1) || j >= nz + 1) {
                    pdxvz = mpml(F.m_dxvz, q, pdxvz, P.axh[ii], P.bxh[ii]);
                    pdzvx = mpml(F.m_dzvx, q, pdzvx, P.azh[jj], P.bzh[jj]);
                }
                sto(O.sxz, q, ldo(F.sxz, q) + dt * ldo(M.mu, q) * (pdzvx + pdxvz));
            }
        }
#pragma unroll
        for (int m = 0; m < 2 * R; ++m) {
            qx[m] = qx[m + 1];
            qz[m] = qz[m + 1];
        }
    }
}

// ---------------------------------------------------------------------------
// Velocity sweep: vx at storage (i, j) = (i - 1/2, j), vz at (i, j - 1/2)
// (module_wavefield.f90:95-134).  z stencils from queues of sxz and szz, x
// stencils of sxx and sxz by shuffles as in the stress sweep.
// ---------------------------------------------------------------------------
template <int R, bool FS>
__global__ __launch_bounds__(256) void k2e_velocity(const ShotE2* __restrict__ shots,
                                                    const int* __restrict__ tile_shot, EF2 F, EV2 O, EM2 M,
                                                    const float* __restrict__ prof, const float* __restrict__ scal,
                                                    Coef c, float dt, float idx, float idz, int pml, int kc) {
    constexpr int W = 64 - 2 * R;
    // one wavefront per tile (blockDim.x = 64): tile, shot and row indices
    // are wavefront-uniform and live in scalar registers
    const int tile = __builtin_amdgcn_readfirstlane(int(blockIdx.x) * int(blockDim.y) + int(threadIdx.y));
    const ShotE2 S = shots[tile_shot[tile]];
    int lt = tile - S.tile0;
    const int tx = lt % S.ntx, tz = lt / S.ntx;
    const int nx = S.nx, nz = S.nz, lo = 1 - pml - R;
    const int l = int(threadIdx.x);
    const int i = -pml + 1 + tx * W - R + l;
    const int j0 = (FS ? 1 : -pml + 1) + tz * kc, j1 = min(j0 + kc - 1, nz + pml);
    if (j0 > j1) return;
    const bool mcol = i <= nx + pml + R;
    const bool own = l >= R && l <= 63 - R && i <= nx + pml;
    const Prof2 P = prof_of(prof, S, pml);
    const float* __restrict__ sci = scal + S.sc0 + pml;
    const float* __restrict__ sch = sci + (nz + 2 * pml + 2);
    const float* __restrict__ sxx = O.sxx;
    const float* __restrict__ szz = O.szz;
    const float* __restrict__ sxz = O.sxz;
    const unsigned szb = unsigned(S.px) * 4u;
    const int ii = i + pml - 1;
    unsigned q = unsigned(at2(S, lo, mcol ? i : lo, j0)) * 4u;
    float qa[2 * R + 1], qc[2 * R + 1];  // sxz, szz
#pragma unroll
    for (int m = 0; m < 2 * R; ++m) {
        const unsigned r = q + unsigned(m - R) * szb;
        qa[m] = mcol ? ldo(sxz, r) : 0.0f;
        qc[m] = mcol ? ldo(szz, r) : 0.0f;
    }
    const bool xin = !(i <= 0 || i >= nx + 1), Xin = !(i <= 1 || i >= nx + 1);
    for (int j = j0; j <= j1; ++j, q += szb) {
        {
            const unsigned r = q + unsigned(R) * szb;
            qa[2 * R] = mcol ? ldo(sxz, r) : 0.0f;
            qc[2 * R] = mcol ? ldo(szz, r) : 0.0f;
        }
        const float cxx = mcol ? ldo(sxx, q) : 0.0f;
        const float sxxx = sten_lanes<R>(cxx, l, -1, c.c);    // d-x sxx
        const float sxxz = sten_lanes<R>(qa[R], l, 0, c.c);   // d+x sxz
        if (own) {
            const int jj = j + pml - 1;
            {
                float b = c.c[0] * (qa[R + 1] - qa[R]);       // d+z sxz
#pragma unroll
                for (int m = 2; m <= R; ++m) b = b + c.c[m - 1] * (qa[R + m] - qa[R - m + 1]);
                float pdxxx = idx * sxxx, pdzxz = FS ? idz * b * sci[j] : idz * b;
                if (!Xin || (!FS && j <= 0) || j >= nz + 1) {
                    pdxxx = mpml(F.m_dxxx, q, pdxxx, P.axh[ii], P.bxh[ii]);
                    pdzxz = mpml(F.m_dzxz, q, pdzxz, P.azi[jj], P.bzi[jj]);
                }
                sto(O.vx, q, ldo(F.vx, q) + dt / ldo(M.rx, q) * (pdxxx + pdzxz));
            }
            if (!FS || j >= 2) {
                float b = c.c[0] * (qc[R] - qc[R - 1]);       // d-z szz
#pragma unroll
                for (int m = 2; m <= R; ++m) b = b + c.c[m - 1] * (qc[R + m - 1] - qc[R - m]);
                float pdxxz = idx * sxxz, pdzzz = FS ? idz * b * sch[j] : idz * b;
                if (!xin || (!FS && j <= 1) || j >= nz + 1) {
                    pdxxz = mpml(F.m_dxxz, q, pdxxz, P.axi[ii], P.bxi[ii]);
                    pdzzz = mpml(F.m_dzzz, q, pdzzz, P.azh[jj], P.bzh[jj]);
                }
                sto(O.vz, q, ldo(F.vz, q) + dt / ldo(M.rz, q) * (pdxxz + pdzzz));
            }
        }
#pragma unroll
        for (int m = 0; m < 2 * R; ++m) {
            qa[m] = qa[m + 1];
            qc[m] = qc[m + 1];
        }
    }
}

// Free surface, after the stress sweep (module_wavefield.f90:223-245): the
// odd stress mirrors szz(1-j) = -szz(1+j), sxx(1-j) = -sxx(1+j),
// sxz(2-j) = -sxz(1+j), j = 1 .. R, and szz(1) = 0, over every stored
// column.  The velocities above the surface stay zero throughout (no sweep,
// source or record writes them), which is what OWL's update zeroes them to.
__global__ void k2e_fs_mirror(const ShotE2* __restrict__ shots, float* sxx, float* szz, float* sxz, int pml, int R) {
    const ShotE2 S = shots[blockIdx.y];
    const int lo = 1 - pml - R, w = S.nx + 2 * (pml + R);
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < w * R; e += gridDim.x * blockDim.x) {
        const int i = lo + e % w, j = 1 + e / w;
        const size_t up = at2(S, lo, i, 1 + j);
        szz[at2(S, lo, i, 1 - j)] = -szz[up];
        sxx[at2(S, lo, i, 1 - j)] = -sxx[up];
        sxz[at2(S, lo, i, 2 - j)] = -sxz[up];
        if (j == 1) szz[at2(S, lo, i, 1)] = 0.0f;
    }
}

// Source injection of step t (add_source): forces into vx, vz, explosions
// into sxx, szz; one block per shot, taps grouped by (field, cell) in
// source order.
__global__ __launch_bounds__(1024) void k2e_inject(const ShotE2* __restrict__ shots, float* vx, float* vz,
                                                   float* sxx, float* szz, const int* __restrict__ inj_beg,
                                                   const Tap* __restrict__ inj, const float* __restrict__ amp_tab,
                                                   int nt, int t) {
    const ShotE2 S = shots[blockIdx.x];
    for (int c = threadIdx.x; c < S.ninj; c += blockDim.x) {
        const int b = inj_beg[S.inj0 + c], e = inj_beg[S.inj0 + c + 1];
        const long long cell = inj[b].cell;
        const int fd = inj[b].field;
        float* f = fd == 0 ? vx : fd == 1 ? vz : fd == 2 ? sxx : szz;
        float acc = f[cell];
        for (int k = b; k < e; ++k) {
            const Tap tp = inj[k];
            if (t < tp.tbeg || t > tp.tend) continue;
            acc = acc + amp_tab[size_t(tp.row) * nt + (t - 1)] * tp.w1 * tp.w2;
        }
        f[cell] = acc;
    }
}

// Receivers of step t: vx at (hx, gz), vz at (gx, hz).
__global__ __launch_bounds__(256) void k2e_record(const ShotE2* __restrict__ shots, const float* __restrict__ vx,
                                                  const float* __restrict__ vz, const int* __restrict__ rec_beg,
                                                  const Tap* __restrict__ rec, const float* __restrict__ rec_w,
                                                  float* traces, int nt, int t) {
    const ShotE2 S = shots[blockIdx.y];
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < 2 * S.nr; r += gridDim.x * blockDim.x) {
        const int comp = r / S.nr, ir = r % S.nr;
        const int g = 2 * S.rec0 + comp * S.nr + ir;
        const float w = rec_w[S.rec0 + ir];
        const float* f = comp == 0 ? vx : vz;
        float acc = 0.0f;
        for (int k = rec_beg[g]; k < rec_beg[g + 1]; ++k) {
            const Tap tp = rec[k];
            acc = acc + f[tp.cell] * tp.w1 * tp.w2 * w;
        }
        traces[S.trace0 + size_t(comp) * nt * S.nr + size_t(t - 1) * S.nr + ir] = acc;
    }
}

// Boundary records (module_boundary_saving.f90:89-160): for vx, then vz:
// left [2-R, 1] x [1-R, nz+R], right [nx+1, nx+R] x [1-R, nz+R], top
// [1-R, nx+R] x [2-R, 1], bottom [1-R, nx+R] x [nz+1, nz+R], x fastest.
// save != 0: record <- fields; else fields <- record.
__host__ __device__ inline long long strip2_len(int nx, int nz, int R) {
    return 2LL * (2LL * R * (nz + 2 * R) + 2LL * R * (nx + 2 * R));
}

template <int R>
__global__ void k2e_strips(const ShotE2* __restrict__ shots, float* vx, float* vz, RecStore st, int t, int pml,
                           int mode) {
    const ShotE2 S = shots[blockIdx.y];
    const int nx = S.nx, nz = S.nz, lo = 1 - pml - R;
    const long long ex = nx + 2 * R, ez = nz + 2 * R;
    const long long a = R * ez, b = ex * R, one = 2 * a + 2 * b, len = 2 * one;
    const long long rec0 = S.strip0 + (long long)(t - 1) * len;
    const int rid = int(blockIdx.y) * st.nt + (t - 1);
    const float sc = mode == kRecAmax ? 1.0f : rec_scale(st, rid), inv = 1.0f / sc;
    float m = 0.0f;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < len;
         e += (long long)gridDim.x * blockDim.x) {
        const int comp = int(e / one);
        long long r = e - comp * one;
        int i, j;
        if (r < 2 * a) {
            const bool hi = r >= a;
            if (hi) r -= a;
            i = (hi ? nx + 1 : 2 - R) + int(r % R);
            j = 1 - R + int(r / R);
        } else {
            r -= 2 * a;
            const bool hi = r >= b;
            if (hi) r -= b;
            i = 1 - R + int(r % ex);
            j = (hi ? nz + 1 : 2 - R) + int(r / ex);
        }
        float* f = comp == 0 ? vx : vz;
        const size_t q = at2(S, lo, i, j);
        if (mode == kRecAmax)
            m = fmaxf(m, fabsf(f[q]));
        else if (mode == kRecSave)
            rec_put(st, rec0 + e, f[q], inv);
        else
            f[q] = rec_get(st, rec0 + e, sc);
    }
    if (mode == kRecAmax) rec_amax(st, rid, m);
}

// Adjoint sources into the adjoint velocities: the receivers' taps,
// grouped by (component, cell) in receiver order; adj holds two [nt][nr]
// blocks per shot.
__global__ void k2e_inject_adjoint(const ShotE2* __restrict__ shots, float* vx, float* vz,
                                   const int* __restrict__ cell_beg, const Tap* __restrict__ taps,
                                   const float* __restrict__ rec_w, const float* __restrict__ adj, int nt, int t) {
    const ShotE2 S = shots[blockIdx.y];
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < S.nadj; c += gridDim.x * blockDim.x) {
        const int b = cell_beg[S.adj0 + c], e = cell_beg[S.adj0 + c + 1];
        const long long cell = taps[b].cell;
        const int fld = taps[b].field;
        float* f = fld == 0 ? vx : vz;
        const float* a_t = adj + S.trace0 + size_t(fld) * nt * S.nr + size_t(t - 1) * S.nr;
        float acc = f[cell];
        for (int k = b; k < e; ++k) {
            const Tap tp = taps[k];
            acc = acc + a_t[tp.row] * tp.w1 * tp.w2 * rec_w[S.rec0 + tp.row];
        }
        f[cell] = acc;
    }
}

// Imaging condition (compute_gradient, submodule_fwi_gradient.f90:440-520)
// over the window interior, with the compliances of :60-80.  FP64
// accumulators of the chain rule's combinations (:330-360):
//   A = g11 + g13 + g33 (vp), B = g55 - 2 g13 (vs), C = rho term,
// g_ij = -(adjoint strain) . (strain of the forward increment).
__global__ __launch_bounds__(256) void k2e_image(const ShotE2* __restrict__ shots, EV2 cur, EV2 prev, EV2 adj,
                                                 const float* __restrict__ c11, const float* __restrict__ c13,
                                                 const float* __restrict__ c55, double* gA, double* gB, double* gC,
                                                 int prev_zero, int pml, int R, int t) {
    const ShotE2 S = shots[blockIdx.y];
    if (t < S.hnt) return;  // this shot's reverse loop has ended (whole block)
    const int nx = S.nx, nz = S.nz, lo = 1 - pml - R;
    const long long n = (long long)nx * nz;
    const size_t sx = 1, sz = S.px;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        const int i = 1 + int(e % nx), j = 1 + int(e / nx);
        const size_t q = at2(S, lo, i, j);
        const float a11 = c11[q], a13 = c13[q], a33 = c11[q], a55 = c55[q];
        float s11 = 0.f, s13 = 0.f, s33 = 0.f, s55 = 0.f;
        if (a13 * a13 - a11 * a33 != 0.0f) {
            s11 = a33 / (-(a13 * a13) + a11 * a33);
            s13 = a13 / (a13 * a13 - a11 * a33);
            s33 = a11 / (-(a13 * a13) + a11 * a33);
        }
        if (a55 != 0.0f)
            s55 = 1.0f / a55;
        else
            s11 = s13 = s33 = s55 = 0.0f;
        float txx = cur.sxx[q], tzz = cur.szz[q];
        float txz = 0.25f * (cur.sxz[q] + cur.sxz[q + sx] + cur.sxz[q + sz] + cur.sxz[q + sx + sz]);
        float svx = cur.vx[q] + cur.vx[q + sx], svz = cur.vz[q] + cur.vz[q + sz];
        if (!prev_zero) {
            txx = txx - prev.sxx[q];
            tzz = tzz - prev.szz[q];
            txz = txz - 0.25f * (prev.sxz[q] + prev.sxz[q + sx] + prev.sxz[q + sz] + prev.sxz[q + sx + sz]);
            svx = svx - (prev.vx[q] + prev.vx[q + sx]);
            svz = svz - (prev.vz[q] + prev.vz[q + sz]);
        }
        const float rxx = adj.sxx[q], rzz = adj.szz[q];
        const float rxz = 0.25f * (adj.sxz[q] + adj.sxz[q + sx] + adj.sxz[q + sz] + adj.sxz[q + sx + sz]);
        const float exx = s11 * txx + s13 * tzz, ezz = s13 * txx + s33 * tzz, exz = s55 * txz;
        const float fxx = s11 * rxx + s13 * rzz, fzz = s13 * rxx + s33 * rzz, fxz = s55 * rxz;
        const double i11 = double(fxx * exx), i33 = double(fzz * ezz), i55 = double(fxz * exz);
        const double i13 = double(fxx * ezz) + double(fzz * exx);
        const float rvx = adj.vx[q] + adj.vx[q + sx], rvz = adj.vz[q] + adj.vz[q + sz];
        const size_t o = size_t(S.g0 + e);
        gA[o] = gA[o] - (i11 + i13 + i33);
        gB[o] = gB[o] - i55 + 2.0 * i13;
        gC[o] = gC[o] + double(svx * rvx) + double(svz * rvz);
    }
}

// Sum of the per-shot window accumulators on the full model (nz, nx), z
// fastest (module_gradient.f90:366-388).
__global__ void k2e_merge(const ShotE2* __restrict__ shots, const int* __restrict__ win, int nshot,
                          const double* __restrict__ gA, const double* __restrict__ gB,
                          const double* __restrict__ gC, double* fA, double* fB, double* fC, int nx, int nz) {
    const long long n = (long long)nx * nz;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        const int kz = int(e % nz), ix = int(e / nz);
        double a = 0.0, b = 0.0, c = 0.0;
        for (int s = 0; s < nshot; ++s) {
            const ShotE2& S = shots[s];
            const int i = ix - (win[2 * s] - 1), k = kz - (win[2 * s + 1] - 1);
            if (i < 0 || i >= S.nx || k < 0 || k >= S.nz) continue;
            const size_t o = size_t(S.g0) + i + size_t(k) * S.nx;
            a += gA[o];
            b += gB[o];
            c += gC[o];
        }
        fA[e] = a;
        fB[e] = b;
        fC[e] = c;
    }
}

struct ShotHostE2 {
    int id = 0, nx = 0, nz = 0, nxbeg = 1, nzbeg = 1, hnt = 1;
    float ox = 0, oz = 0;
    std::vector<owl_source> src;
    std::vector<owl_receiver> rec;
};

}  // namespace e2
}  // namespace owl

using namespace owl;
using namespace owl::e2;

struct owl_elastic2d {
    owl_acoustic_config cfg{};
    bool fs = false;                       // OWL's free surface on the depth-refined mesh
    float fs_refine = 4.0f, fs_dzmax = 0.0f;
    std::vector<owl::fsm::SsgMesh> mesh;   // per shot (free surface)
    std::vector<std::vector<float>> rho_map;  // per shot: rho on the refined rows, column i (1..nx) at [(i-1) nrj + j + pml]
    int R = 4, pml = 15, nt = 0, kc = 32, cy = 4;
    float dt = 0, idx = 0, idz = 0;
    e2::Coef coef{};
    std::vector<ShotHostE2> hs;
    std::vector<ShotE2> sd;
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
    DevBuf<ShotE2> d_shots;
    DevBuf<int> d_tile_shot, d_inj_beg, d_rec_beg, d_adj_beg, d_win;
    DevBuf<float> d_prof, d_amp, d_rec_w, d_traces, d_strips, d_adj, d_obs, d_sc;
    DevBuf<unsigned> d_rscale;  // per (shot, step) boundary record: largest magnitude (16-bit storage)
    int rec_store = OWL_STORE_F32;
    DevBuf<float> med[6];   // c11, c13, mu (at sxz), rho_x, rho_z, c55 at the nodes
    DevBuf<float> fld[13];  // five wavefields, eight memory variables
    DevBuf<float> fB[5];    // the other half of the reverse ping-pong
    DevBuf<float> aF[13];   // adjoint field
    DevBuf<Tap> d_inj, d_rec, d_adj_taps;
    DevBuf<double> d_gA, d_gB, d_gC, d_fA, d_fB, d_fC, d_misfit, d_scratch;
    // the forward time loop as one HIP graph (modelling without profiling
    // or boundary records), re-captured when a launch argument changes
    LoopGraph graph;
    ~owl_elastic2d() {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

inline size_t cell2(const owl_elastic2d* b, const ShotE2& S, int i, int j) {
    const int lo = 1 - b->pml - b->R, h = b->pml + b->R;
    if (i < lo || i > S.nx + h || j < lo || j > S.nz + h)
        fail(OWL_ERR_ARG, "a source/receiver stencil reaches outside the padded grid; increase pml");
    return size_t(S.off) + size_t(i - lo) + size_t(j - lo) * size_t(S.px);
}

inline size_t model_index2(const owl_elastic2d* b, const ShotHostE2& H, int i, int j) {
    const int ic = std::min(std::max(i, 1), H.nx), jc = std::min(std::max(j, 1), H.nz);
    return size_t(H.nzbeg - 1 + jc - 1) + size_t(H.nxbeg - 1 + ic - 1) * size_t(b->cfg.nz);
}

// damp_coef (module_cfspml.f90:285-330) along one axis with ratio 0 and
// kmax = 1, for the distances |i-1|/pml, |i-1-0.5|/pml (i <= 1) and
// |i-n|/pml, |i-n-0.5|/pml (i >= n) formed in double (:466-575); zero
// elsewhere.  Appends ai, bi, ah, bh of n + 2 pml entries.
void damp_ab(float pmlvp, float nd, float d, int pml, float adt, float alphamax, float& a, float& b) {
    const float R0 = 1.0e-5f, kx = 1.0f;
    const float damp = pmlvp * (2 + 1.0f) * std::log(1.0f / R0) / (2.0f * float(pml) * d) * (nd * nd);
    const float alpha = std::max(alphamax * (1.0f - nd), 0.25f * alphamax);
    a = -2.0f * adt * damp / kx / (2.0f + adt * (alpha + damp / kx));
    b = (2.0f - adt * (alpha + damp / kx)) / (2.0f + adt * (alpha + damp / kx));
}

void damp_axis(int n, int pml, float d, float pmlvp, float adt, float alphamax, std::vector<float>& out) {
    const int len = n + 2 * pml;
    std::vector<float> ai(len, 0.f), bi(len, 0.f), ah(len, 0.f), bh(len, 0.f);
    auto ab = [&](float nd, float& a, float& b) { damp_ab(pmlvp, nd, d, pml, adt, alphamax, a, b); };
    for (int i = -pml + 1; i <= n + pml; ++i) {
        float ndi, ndh;
        if (i <= 1) {
            ndi = float(std::fabs((i - 1 + 0.0) / pml));
            ndh = float(std::fabs((i - 1 - 0.5) / pml));
        } else if (i >= n) {
            ndi = float(std::fabs((i - n + 0.0) / pml));
            ndh = float(std::fabs((i - n - 0.5) / pml));
        } else {
            continue;
        }
        const int q = i + pml - 1;
        ab(ndi, ai[q], bi[q]);
        ab(ndh, ah[q], bh[q]);
    }
    for (auto* v : {&ai, &bi, &ah, &bh}) out.insert(out.end(), v->begin(), v->end());
}

// The z profiles on the refined mesh: row i takes d = dz_i(i) for every
// grid (module_cfspml.f90:446-473); dzi at [i + pml].
void damp_axis_rows(int n, int pml, const std::vector<float>& dzi, float pmlvp, float adt, float alphamax,
                    std::vector<float>& out) {
    const int len = n + 2 * pml;
    std::vector<float> ai(len, 0.f), bi(len, 0.f), ah(len, 0.f), bh(len, 0.f);
    for (int i = -pml + 1; i <= n + pml; ++i) {
        float ndi, ndh;
        if (i <= 1) {
            ndi = float(std::fabs((i - 1 + 0.0) / pml));
            ndh = float(std::fabs((i - 1 - 0.5) / pml));
        } else if (i >= n) {
            ndi = float(std::fabs((i - n + 0.0) / pml));
            ndh = float(std::fabs((i - n - 0.5) / pml));
        } else {
            continue;
        }
        const float d = dzi[size_t(i + pml)];
        const int q = i + pml - 1;
        damp_ab(pmlvp, ndi, d, pml, adt, alphamax, ai[q], bi[q]);
        damp_ab(pmlvp, ndh, d, pml, adt, alphamax, ah[q], bh[q]);
    }
    for (auto* v : {&ai, &bi, &ah, &bh}) out.insert(out.end(), v->begin(), v->end());
}

void retile2(owl_elastic2d* b) {
    b->tile_shot.clear();
    for (size_t s = 0; s < b->sd.size(); ++s) {
        ShotE2& S = b->sd[s];
        S.ntx = (S.nx + 2 * b->pml + (64 - 2 * b->R) - 1) / (64 - 2 * b->R);  // sweeps update 64 - 2R columns
        S.nzc = (S.nz + (b->fs ? 1 : 2) * b->pml + b->kc - 1) / b->kc;  // free surface: rows 1 .. nz+pml
        S.tile0 = int(b->tile_shot.size());
        b->tile_shot.insert(b->tile_shot.end(), size_t(S.ntx) * S.nzc, int(s));
    }
    // whole blocks of cy tiles: pad with the last shot's tiles beyond its
    // range (their lanes find no work)
    while (b->tile_shot.size() % size_t(b->cy)) b->tile_shot.push_back(int(b->sd.size()) - 1);
}

void plan2(owl_elastic2d* b, const owl_acoustic_config* cfg, const float* vp, const float* vs, const float* rho,
           int nshot, const owl_shot* shots, const owl_elastic_surface* surf) {
    if (!cfg || !vp || !vs || !rho || nshot < 1 || !shots) fail(OWL_ERR_ARG, "null argument or no shots");
    if (!fd_coefficients(cfg->fdhalf, b->coef.c)) fail(OWL_ERR_ARG, "fdhalf must be 2..8");
    if (cfg->nx < 1 || cfg->nz < 1 || !(cfg->dx > 0) || !(cfg->dz > 0) || !(cfg->dt > 0))
        fail(OWL_ERR_ARG, "invalid grid or dt");
    if (cfg->pml < cfg->fdhalf || cfg->pml + cfg->fdhalf < kNkw) fail(OWL_ERR_ARG, "pml must be >= fdhalf");
    b->fs = cfg->free_surface != 0 || surf != nullptr;
    b->fs_refine = surf && surf->dz_refine > 0 ? surf->dz_refine : 4.0f;
    b->fs_dzmax = surf && surf->dz_max > 0 ? surf->dz_max : 1.5f * cfg->dz;
    if (cfg->data_dt != cfg->dt || cfg->data_tmax != cfg->tmax)
        fail(OWL_ERR_ARG, "data_dt/data_tmax must equal dt/tmax (resample the traces with owl_resample_traces)");
    b->cfg = *cfg;
    b->R = cfg->fdhalf;
    b->pml = cfg->pml;
    b->dt = cfg->dt;
    b->nt = time_steps(cfg->tmax, cfg->dt);
    if (b->nt < 2) fail(OWL_ERR_ARG, "tmax/dt gives fewer than 2 steps");
    b->idx = float(1.0 / double(cfg->dx));
    b->idz = float(1.0 / double(cfg->dz));
    const size_t nfull = size_t(cfg->nx) * cfg->nz;
    b->vp_full.assign(vp, vp + nfull);
    b->vs_full.assign(vs, vs + nfull);
    b->rho_full.assign(rho, rho + nfull);
    const float f0factor = cfg->f0factor > 0 ? cfg->f0factor : 1.0f;
    const int R = b->R, pml = b->pml, nt = b->nt, lo = 1 - pml - R;
    b->hs.assign(nshot, ShotHostE2{});
    b->sd.assign(nshot, ShotE2{});
    b->nrec_total = b->max_nr = b->max_nadj = 0;
    b->pts_full = 0;
    b->mesh.assign(b->fs ? size_t(nshot) : 0, fsm::SsgMesh{});
    b->rho_map.assign(b->fs ? size_t(nshot) : 0, std::vector<float>{});
    // free surface: per shot the window's c11 (= c33), c13, c55 and rho
    // mapped onto the refined rows, column i at [(i + pml) * nrj + j + pml]
    // for i, j = -pml .. n+pml+1
    std::vector<std::vector<float>> mapped(b->fs ? size_t(nshot) * 4 : 0);
    long long off = 0, trace = 0, strip = 0, grad = 0;
    int prof = 0;
    std::vector<float> profs, amp, rec_w, scal;
    std::vector<int> inj_beg{0}, adj_beg{0};
    std::vector<Tap> inj, adj_taps;
    std::vector<std::vector<Tap>> rec_taps;  // per (shot, component, receiver)
    for (int s = 0; s < nshot; ++s) {
        const owl_shot& in = shots[s];
        ShotHostE2& H = b->hs[s];
        ShotE2& S = b->sd[s];
        H.id = in.id;
        H.nxbeg = in.nxbeg > 0 ? in.nxbeg : 1;
        H.nzbeg = in.nzbeg > 0 ? in.nzbeg : 1;
        const int nxend = in.nxend > 0 ? in.nxend : cfg->nx, nzend = in.nzend > 0 ? in.nzend : cfg->nz;
        if (H.nxbeg > nxend || nxend > cfg->nx || H.nzbeg > nzend || nzend > cfg->nz)
            fail(OWL_ERR_ARG, "shot window outside the model");
        H.nx = nxend - H.nxbeg + 1;
        H.nz = nzend - H.nzbeg + 1;
        H.ox = float(H.nxbeg - 1) * cfg->dx + cfg->ox;
        H.oz = float(H.nzbeg - 1) * cfg->dz + cfg->oz;
        if (in.ns < 1 || !in.sources || in.nr < 1 || !in.receivers)
            fail(OWL_ERR_ARG, "shot without sources or receivers");
        H.src.assign(in.sources, in.sources + in.ns);
        H.rec.assign(in.receivers, in.receivers + in.nr);
        // pmlvp = sqrt(max c11/rho, c33/rho over the window's four edge
        // lines) (module_vars.f90:298-302); vmax for the CFL
        float pm = 0.f, vmax = 0.f, f0max = 0.f;
        for (int i = 1; i <= H.nx; ++i)
            for (int j = 1; j <= H.nz; ++j) {
                const size_t m = model_index2(b, H, i, j);
                const float r = b->rho_full[m], p = b->vp_full[m], v = b->vs_full[m];
                if (!(r > 0.f) || !(p > 0.f) || !(v > 0.f)) fail(OWL_ERR_ARG, "vp, vs and rho must be positive");
                const float c33 = r * (p * p);
                vmax = std::max(vmax, p);
                if (i == 1 || i == H.nx || j == 1 || j == H.nz) pm = std::max(pm, c33 / r);
            }
        const float pmlvp = std::sqrt(pm);
        float f0min = INFINITY;
        for (auto& sc : H.src) {
            f0max = std::max(f0max, sc.f0);
            f0min = std::min(f0min, sc.f0);
        }
        int nzr = H.nz;  // rows the shot steps (refined with the free surface)
        if (b->fs) {
            std::vector<float> wc(size_t(H.nx) * H.nz), wr(size_t(H.nx) * H.nz), wc13(wc.size()), wc55(wc.size());
            for (int i = 1; i <= H.nx; ++i)
                for (int j = 1; j <= H.nz; ++j) {
                    const size_t m = model_index2(b, H, i, j), q = size_t(j - 1) + size_t(i - 1) * H.nz;
                    const float r = b->rho_full[m], p = b->vp_full[m], v = b->vs_full[m];
                    const float c33 = r * (p * p), c55 = r * (v * v);
                    wc[q] = c33;
                    wc13[q] = c33 - 2 * c55;
                    wc55[q] = c55;
                    wr[q] = r;
                }
            fsm::SsgMesh& Mh = b->mesh[size_t(s)];
            Mh = fsm::build_ssg(H.nx, H.nz, cfg->dz, pml, b->fs_refine, b->fs_dzmax, wc55.data(), wr.data(), f0min);
            nzr = Mh.nz;
            const int nrj = nzr + 2 * pml + 2, nci = H.nx + 2 * pml + 2;
            const std::vector<float>* src4[4] = {&wc, &wc13, &wc55, &wr};
            for (int k = 0; k < 4; ++k) {
                auto& out = mapped[size_t(s) * 4 + k];
                out.resize(size_t(nci) * nrj);
                for (int i = -pml; i <= H.nx + pml + 1; ++i)
                    fsm::map_column_ssg(Mh, H.nx, H.nz, src4[k]->data(), i, out.data() + size_t(i + pml) * nrj);
            }
            // stability on the refined mesh (module_vars.f90:383-408)
            float csum = 0.0f;
            for (int m = 0; m < R; ++m) csum = csum + std::fabs(b->coef.c[m]);
            float dts = INFINITY;
            const auto& mc = mapped[size_t(s) * 4], &mr = mapped[size_t(s) * 4 + 3];
            for (int j = 1; j <= nzr; ++j) {
                const float dmin = std::min(Mh.dzi[size_t(j + pml)], Mh.dzh[size_t(j + pml)]);
                for (int i = 1; i <= H.nx; ++i) {
                    const size_t q = size_t(i + pml) * nrj + size_t(j + pml);
                    const float maxv = std::sqrt(mc[q] / mr[q]);  // c33 = c11 here
                    dts = std::min(dts, 1.0f / (csum * maxv *
                                               std::sqrt(1.0f / (cfg->dx * cfg->dx) + 1.0f / (dmin * dmin))));
                }
            }
            if (cfg->dt > dts)
                fail(OWL_ERR_CFL, "dt = " + std::to_string(cfg->dt) + " exceeds the stable dt = " +
                                      std::to_string(dts) + " of the free-surface mesh");
            auto& rm = b->rho_map[size_t(s)];
            rm.resize(size_t(H.nx) * nrj);
            for (int i = 1; i <= H.nx; ++i)
                std::copy(mr.begin() + long(size_t(i + pml) * nrj), mr.begin() + long(size_t(i + pml + 1) * nrj),
                          rm.begin() + long(size_t(i - 1) * nrj));
            S.sc0 = int(scal.size());
            scal.insert(scal.end(), Mh.sci.begin(), Mh.sci.end());
            scal.insert(scal.end(), Mh.sch.begin(), Mh.sch.end());
        } else {
            const float dts = dt_stable(R, 2, vmax, cfg->dx, 1.0f, cfg->dz);
            if (cfg->dt > dts)
                fail(OWL_ERR_CFL, "dt = " + std::to_string(cfg->dt) + " exceeds the stable dt = " + std::to_string(dts));
            S.sc0 = 0;
        }
        S.nx = H.nx;
        S.nz = nzr;
        S.px = (H.nx + 2 * (pml + R) + kPitchAlign - 1) / kPitchAlign * kPitchAlign;
        S.nzp = nzr + 2 * (pml + R);
        S.off = off;
        off += (long long)S.px * S.nzp;
        off = (off + 63) / 64 * 64;
        b->pts_full += double(H.nx + 2 * pml) * double(nzr + (b->fs ? 1 : 2) * pml);
        S.prof = prof;
        const float amax = float(double(1.0f * f0max) * M_PI), adt = std::fabs(cfg->dt);
        damp_axis(H.nx, pml, cfg->dx, pmlvp, adt, amax, profs);
        if (b->fs)
            damp_axis_rows(nzr, pml, b->mesh[size_t(s)].dzi, pmlvp, adt, amax, profs);
        else
            damp_axis(H.nz, pml, cfg->dz, pmlvp, adt, amax, profs);
        prof += 4 * (H.nx + 2 * pml) + 4 * (nzr + 2 * pml);
        S.nr = int(H.rec.size());
        S.trace0 = trace;
        trace += 2LL * nt * S.nr;
        S.strip0 = strip;
        strip += strip2_len(H.nx, nzr, R) * nt;
        S.g0 = grad;
        grad += (long long)H.nx * nzr;
        const fsm::SsgMesh* Ms = b->fs ? &b->mesh[size_t(s)] : nullptr;
        // a point's location: x on the regular mesh; z on the regular or the
        // refined rows
        auto place = [&](float x, float z) {
            Located L = locate(2, x, 0.0f, z, H.ox, 0.0f, H.oz, cfg->dx, 1.0f, cfg->dz);
            if (Ms) fsm::locate_ssg(*Ms, z, H.oz, L.gz, L.hz, L.iz, L.hzw);
            return L;
        };
        // the mapped density at refined node (i, j) (free surface)
        auto rho_at = [&](int i, int j) {
            const int nrj = nzr + 2 * pml + 2;
            const int ic = std::min(std::max(i, -pml), H.nx + pml + 1), jc = std::min(std::max(j, -pml), nzr + pml + 1);
            return mapped[size_t(s) * 4 + 3][size_t(ic + pml) * nrj + size_t(jc + pml)];
        };
        // sources (add_source, module_wavefield.f90:294-403): a force adds
        // sin(polar) amp / rho(hx, gz) w to vx at (hx, gz) and cos(polar)
        // amp / rho(gx, hz) w to vz at (gx, hz); an explosion subtracts
        // amp / (dx dz) w from sxx and szz at (gx, gz), its wavelet the time
        // derivative (time_integration = -1); amp = stf amp dt
        const float xmax = float(H.nx - 1) * cfg->dx + H.ox, zmax = float(H.nz - 1) * cfg->dz + H.oz;
        std::vector<Tap> st;
        for (const owl_source& sc0 : H.src) {
            owl_source sc = sc0;
            if (sc.mechanism != OWL_FORCE && sc.mechanism != OWL_EXPLOSION)
                fail(OWL_ERR_ARG, "the elastic solver takes force and explosion sources");
            // with the free surface a source above dz_i(1) moves to it with
            // its amplitude x 1.2 (module_vars.f90:459-466)
            if (Ms && sc.z < Ms->dzi[size_t(1 + pml)]) {
                sc.z = Ms->dzi[size_t(1 + pml)];
                sc.amp = sc.amp * 1.2f;
            }
            if (!(sc.x >= H.ox && sc.x <= xmax && sc.z >= H.oz && sc.z <= zmax))
                fail(OWL_ERR_ARG, "source outside its shot window");
            const Located L = place(sc.x, sc.z);
            int hnt = 0;
            std::vector<float> stf = source_time_function(sc.wavelet, sc.f0, cfg->dt, f0factor, &hnt);
            if (sc.mechanism == OWL_EXPLOSION) stf = stf_time_derivative(stf, cfg->dt, nt);
            if (&sc0 == &H.src[0]) {
                H.hnt = hnt;
                S.hnt = std::max(hnt, 1);
            }
            const int nbeg = fortran_nint(sc.t0 / cfg->dt) + 1, nend = nbeg + int(stf.size()) - 1;
            if (sc.mechanism == OWL_EXPLOSION) {
                // the stress drop divides by the local spacing dz_i(gz) on the refined mesh
                const float dzs = Ms ? Ms->dzi[size_t(L.gz + pml)] : cfg->dz;
                const int row = int(amp.size() / nt);
                for (int t = 1; t <= nt; ++t) {
                    const float a = (t >= nbeg && t <= nend) ? stf[t - nbeg] * sc.amp * cfg->dt : 0.0f;
                    amp.push_back(-(a / (cfg->dx * dzs)));
                }
                for (int c = 2; c < 4; ++c)
                    for (int irz = -kNkw; irz <= kNkw; ++irz)
                        for (int irx = -kNkw; irx <= kNkw; ++irx) {
                            const float w1 = L.ix[irx + kNkw], w2 = L.iz[irz + kNkw];
                            if (w1 == 0.f || w2 == 0.f) continue;
                            if (Ms && L.gz + irz < 2) continue;  // rows >= 2 take a source
                            st.push_back(Tap{(long long)cell2(b, S, L.gx + irx, L.gz + irz), w1, w2, row, nbeg, nend, c});
                        }
                continue;
            }
            const float rhx = Ms ? rho_at(L.hx, L.gz) : b->rho_full[model_index2(b, H, L.hx, L.gz)];
            const float rhz = Ms ? rho_at(L.gx, L.hz) : b->rho_full[model_index2(b, H, L.gx, L.hz)];
            const float sp = std::sin(sc.polar), cp = std::cos(sc.polar);
            const int gx[2] = {L.hx, L.gx}, gz[2] = {L.gz, L.hz};
            const float* wx[2] = {L.hxw, L.ix};
            const float* wz[2] = {L.iz, L.hzw};
            const float fac[2] = {sp, cp}, rr[2] = {rhx, rhz};
            for (int c = 0; c < 2; ++c) {
                const int row = int(amp.size() / nt);
                for (int t = 1; t <= nt; ++t) {
                    const float a = (t >= nbeg && t <= nend) ? stf[t - nbeg] * sc.amp * cfg->dt : 0.0f;
                    amp.push_back(fac[c] * a / rr[c]);
                }
                for (int irz = -kNkw; irz <= kNkw; ++irz)
                    for (int irx = -kNkw; irx <= kNkw; ++irx) {
                        const float w1 = wx[c][irx + kNkw], w2 = wz[c][irz + kNkw];
                        if (w1 == 0.f || w2 == 0.f) continue;
                        if (Ms && gz[c] + irz < 2) continue;
                        st.push_back(Tap{(long long)cell2(b, S, gx[c] + irx, gz[c] + irz), w1, w2, row, nbeg, nend, c});
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
        // receivers: vx at (hx, gz), vz at (gx, hz) (outside the window: weight 0)
        // With the free surface a recorded tap above it reads its even image
        // (vx(1-j) = vx(1+j), vz(2-j) = vz(1+j), submodule_forward_modeling
        // .f90:67-75), and the adjoint sources skip the rows above 2
        // (submodule_fwi_gradient.f90:223, 244).
        S.rec0 = b->nrec_total;
        std::vector<std::vector<Tap>> comp(2 * S.nr), acomp(2 * S.nr);
        for (int r = 0; r < S.nr; ++r) {
            const owl_receiver& rc = H.rec[r];
            const bool valid = rc.x >= H.ox && rc.x <= xmax && rc.z >= H.oz && rc.z <= zmax;
            const float w = (valid && rc.weight != 0.f) ? rc.weight : 0.f;
            rec_w.push_back(w);
            if (w == 0.f) continue;
            const Located L = place(rc.x, rc.z);
            const int gx[2] = {L.hx, L.gx}, gz[2] = {L.gz, L.hz};
            const float* wx[2] = {L.hxw, L.ix};
            const float* wz[2] = {L.iz, L.hzw};
            for (int c = 0; c < 2; ++c)
                for (int irz = -kNkw; irz <= kNkw; ++irz)
                    for (int irx = -kNkw; irx <= kNkw; ++irx) {
                        const float w1 = wx[c][irx + kNkw], w2 = wz[c][irz + kNkw];
                        if (w1 == 0.f || w2 == 0.f) continue;
                        int row = gz[c] + irz;
                        const Tap ta{(long long)cell2(b, S, gx[c] + irx, row), w1, w2, r, 0, 0, c};
                        if (Ms) {
                            if (row >= 2) acomp[size_t(c) * S.nr + r].push_back(ta);
                            if (c == 0 && row < 1 && row >= 1 - R) row = 2 - row;  // rows above 1-R stay zero
                            if (c == 1 && row < 2 && row >= 2 - R) row = 3 - row;
                        } else {
                            acomp[size_t(c) * S.nr + r].push_back(ta);
                        }
                        comp[size_t(c) * S.nr + r].push_back(
                            Tap{(long long)cell2(b, S, gx[c] + irx, row), w1, w2, r, 0, 0, c});
                    }
        }
        {
            std::vector<Tap> at;
            for (int c = 0; c < 2; ++c)
                for (int r = 0; r < S.nr; ++r)
                    for (const Tap& tp : acomp[size_t(c) * S.nr + r]) at.push_back(tp);
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
        for (auto& v : comp) rec_taps.push_back(std::move(v));
        b->nrec_total += S.nr;
        b->max_nr = std::max(b->max_nr, S.nr);
    }
    b->field_elems = size_t(off);
    if (b->field_elems >= (size_t(1) << 29))
        fail(OWL_ERR_ARG, "the shots' padded grids exceed 2 GiB per field; split the batch");
    b->trace_elems = size_t(trace);
    b->strip_elems = size_t(strip);
    b->grad_elems = size_t(grad);
    retile2(b);
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
    b->d_adj_beg.upload(adj_beg, stm);
    if (adj_taps.empty()) adj_taps.push_back(Tap{});
    b->d_adj_taps.upload(adj_taps, stm);
    {
        std::vector<int> win;
        for (auto& H : b->hs) win.insert(win.end(), {H.nxbeg, H.nzbeg});
        b->d_win.upload(win, stm);
    }
    // padded media (edge replication), then the effective values at the
    // staggered positions in OWL's operation order
    const size_t n = b->field_elems;
    std::vector<float> c11(n, 0.f), c13(n, 0.f), c55(n, 1.f), rh(n, 1.f);
    for (size_t s = 0; s < b->hs.size(); ++s) {
        const ShotHostE2& H = b->hs[s];
        const ShotE2& S = b->sd[s];
        const int nrj = S.nz + 2 * pml + 2;
        for (int j = lo; j < lo + S.nzp; ++j)
            for (int i = lo; i < lo + S.px; ++i) {
                const size_t q = size_t(S.off) + size_t(i - lo) + size_t(j - lo) * S.px;
                if (b->fs) {
                    // the mapped columns, edge values beyond the padded range
                    const int ic = std::min(std::max(i, -pml), H.nx + pml + 1);
                    const int jc = std::min(std::max(j, -pml), S.nz + pml + 1);
                    const size_t k = size_t(ic + pml) * nrj + size_t(jc + pml);
                    c11[q] = mapped[s * 4][k];
                    c13[q] = mapped[s * 4 + 1][k];
                    c55[q] = mapped[s * 4 + 2][k];
                    rh[q] = mapped[s * 4 + 3][k];
                    continue;
                }
                const size_t m = model_index2(b, H, i, j);
                const float r = b->rho_full[m], p = b->vp_full[m], v = b->vs_full[m];
                const float a = r * (p * p), mu = r * (v * v);
                c11[q] = a;
                c13[q] = a - 2 * mu;
                c55[q] = mu;
                rh[q] = r;
            }
    }
    if (scal.empty()) scal.push_back(1.0f);
    b->d_sc.upload(scal, stm);
    std::vector<float> mxz(n, 0.f), rx(n, 1.f), rz(n, 1.f);
    for (size_t s = 0; s < b->sd.size(); ++s) {
        const ShotE2& S = b->sd[s];
        const size_t sx = 1, sz = S.px;
        for (int j = lo + 1; j < lo + S.nzp; ++j)
            for (int i = lo + 1; i < lo + S.nx + 2 * (pml + R); ++i) {
                const size_t q = size_t(S.off) + size_t(i - lo) + size_t(j - lo) * S.px;
                mxz[q] = 4.0f / (1.0f / c55[q - sx - sz] + 1.0f / c55[q - sz] + 1.0f / c55[q - sx] + 1.0f / c55[q]);
                rx[q] = 0.5f * (rh[q - sx] + rh[q]);
                rz[q] = 0.5f * (rh[q - sz] + rh[q]);
            }
    }
    const std::vector<float>* media[6] = {&c11, &c13, &mxz, &rx, &rz, &c55};
    for (int k = 0; k < 6; ++k) b->med[k].upload(*media[k], stm);
    for (auto& f : b->fld) f.alloc(b->field_elems);
    b->d_traces.alloc(std::max<size_t>(b->trace_elems, 1));
    OWL_HIP(hipStreamSynchronize(stm));
}

EF2 fields2(DevBuf<float>* f) {
    EF2 F;
    float** p = &F.sxx;
    for (int k = 0; k < 13; ++k) p[k] = f[k].get();
    return F;
}
EV2 out_of(const EF2& F) { return EV2{F.sxx, F.szz, F.sxz, F.vx, F.vz}; }
EV2 ev_of(DevBuf<float>* f) { return EV2{f[0].get(), f[1].get(), f[2].get(), f[3].get(), f[4].get()}; }
EF2 with_fields(EF2 F, const EV2& V) {
    F.sxx = V.sxx; F.szz = V.szz; F.sxz = V.sxz; F.vx = V.vx; F.vz = V.vz;
    return F;
}

EM2 media2(owl_elastic2d* b) {
    return EM2{b->med[0].get(), b->med[1].get(), b->med[2].get(), b->med[3].get(), b->med[4].get()};
}

// Algorithmic bytes per sweep: stress sweep reads vx, vz, sxx, szz, sxz,
// c11, c13, mu and writes three stresses (44 B per point); velocity sweep
// reads three stresses, vx, vz, rho_x, rho_z and writes two velocities
// (36 B); 8 B per slab cell and damped direction (memory variable r+w).
void sweep_bytes2(const owl_elastic2d* b, double& bs, double& bv) {
    bs = bv = 0;
    for (auto& S : b->sd) {
        const double lx = S.nx + 2.0 * b->pml, lz = S.nz + (b->fs ? 1.0 : 2.0) * b->pml;
        const double slab = 2.0 * b->pml * (lx + lz);
        bs += 44.0 * lx * lz + 8.0 * 2.0 * slab;
        bv += 36.0 * lx * lz + 8.0 * 2.0 * slab;
    }
}

void maybe_collect(owl_elastic2d* b) {
    if (b->prof.pending() > 20000) {
        OWL_HIP(hipStreamSynchronize(b->stream));
        b->prof.collect();
    }
}

template <class Fn>
void dispatch_R2e(int R, Fn&& f) {
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

// One time step of F (memory variables in F) into O (O may be F's own
// fields): stresses, then velocities (module_wavefield.f90:42-134).
template <int R>
void step2(owl_elastic2d* b, const EF2& F, const EV2& O, float dt, double bs, double bv) {
    if (64 * b->cy > 256) fail(OWL_ERR_ARG, "launch shape exceeds the sweeps' 256-thread blocks");
    hipStream_t st = b->stream;
    const dim3 blk(64, unsigned(b->cy)), grd(unsigned(b->tile_shot.size() / size_t(b->cy)));
    const EM2 M = media2(b);
    int ev = b->prof.begin(st, 0, bs);
    if (b->fs)
        hipLaunchKernelGGL((k2e_stress<R, true>), grd, blk, 0, st, b->d_shots.get(), b->d_tile_shot.get(), F, O, M,
                           b->d_prof.get(), b->d_sc.get(), b->coef, dt, b->idx, b->idz, b->pml, b->kc);
    else
        hipLaunchKernelGGL((k2e_stress<R, false>), grd, blk, 0, st, b->d_shots.get(), b->d_tile_shot.get(), F, O, M,
                           b->d_prof.get(), b->d_sc.get(), b->coef, dt, b->idx, b->idz, b->pml, b->kc);
    b->prof.end(st, ev);
    if (b->fs) {
        ev = b->prof.begin(st, 5, 0);
        int w = 1;
        for (auto& S : b->sd) w = std::max(w, (S.nx + 2 * (b->pml + b->R)) * b->R);
        hipLaunchKernelGGL(k2e_fs_mirror, dim3(unsigned((w + 255) / 256), unsigned(b->sd.size())), dim3(256), 0, st,
                           b->d_shots.get(), O.sxx, O.szz, O.sxz, b->pml, b->R);
        b->prof.end(st, ev);
    }
    ev = b->prof.begin(st, 1, bv);
    if (b->fs)
        hipLaunchKernelGGL((k2e_velocity<R, true>), grd, blk, 0, st, b->d_shots.get(), b->d_tile_shot.get(), F, O,
                           M, b->d_prof.get(), b->d_sc.get(), b->coef, dt, b->idx, b->idz, b->pml, b->kc);
    else
        hipLaunchKernelGGL((k2e_velocity<R, false>), grd, blk, 0, st, b->d_shots.get(), b->d_tile_shot.get(), F, O,
                           M, b->d_prof.get(), b->d_sc.get(), b->coef, dt, b->idx, b->idz, b->pml, b->kc);
    b->prof.end(st, ev);
}

void strips2(owl_elastic2d* b, float* vx, float* vz, int t, int save) {
    long long mx = 1;
    for (auto& S : b->sd) mx = std::max(mx, strip2_len(S.nx, S.nz, b->R));
    const dim3 grd(unsigned(std::min<long long>(1024, (mx + 255) / 256)), unsigned(b->sd.size()));
    const RecStore st{b->d_strips.get(), b->d_rscale.get(), b->rec_store, b->nt};
    const int ev = b->prof.begin(b->stream, 4, 0);
    dispatch_R2e(b->R, [&](auto r) {
        if (save && b->rec_store != OWL_STORE_F32)
            hipLaunchKernelGGL(k2e_strips<decltype(r)::value>, grd, dim3(256), 0, b->stream, b->d_shots.get(), vx,
                               vz, st, t, b->pml, int(kRecAmax));
        hipLaunchKernelGGL(k2e_strips<decltype(r)::value>, grd, dim3(256), 0, b->stream, b->d_shots.get(), vx, vz,
                           st, t, b->pml, save ? int(kRecSave) : int(kRecRestore));
    });
    b->prof.end(b->stream, ev);
}

// Every device address and launch parameter the forward loop passes to its
// kernels: a captured loop is replayed only while these are unchanged.
std::vector<uintptr_t> forward_signature2(const owl_elastic2d* b) {
    std::vector<uintptr_t> v{uintptr_t(b->R), uintptr_t(b->nt), uintptr_t(b->kc), uintptr_t(b->cy),
                             uintptr_t(b->tile_shot.size()), uintptr_t(b->max_nr), uintptr_t(b->fs),
                             uintptr_t(float_bits(b->dt)), uintptr_t(float_bits(b->idx)),
                             uintptr_t(float_bits(b->idz)),
                             uintptr_t(b->d_shots.get()), uintptr_t(b->d_tile_shot.get()),
                             uintptr_t(b->d_rec_beg.get()), uintptr_t(b->d_rec.get()), uintptr_t(b->d_rec_w.get()),
                             uintptr_t(b->d_traces.get()), uintptr_t(b->d_inj_beg.get()), uintptr_t(b->d_inj.get()),
                             uintptr_t(b->d_amp.get()), uintptr_t(b->d_prof.get()), uintptr_t(b->d_sc.get())};
    for (auto& f : b->fld) v.push_back(uintptr_t(f.get()));
    for (auto& f : b->med) v.push_back(uintptr_t(f.get()));
    return v;
}

template <int R>
void run_forward2(owl_elastic2d* b, bool save_boundary) {
    hipStream_t st = b->stream;
    for (auto& f : b->fld) f.zero(st);
    b->d_traces.zero(st);
    if (save_boundary) alloc_records(b->d_strips, b->d_rscale, b->strip_elems, b->rec_store, b->sd.size(), b->nt, st);
    const EF2 F = fields2(b->fld);
    const unsigned nshot = unsigned(b->sd.size());
    double bs, bv;
    sweep_bytes2(b, bs, bv);
    const unsigned rg = unsigned(std::max(1, (2 * b->max_nr + 255) / 256));
    // step t: [boundary record of the state at the start of the step] ->
    // record(t) -> update -> source (submodule_forward_modeling.f90:31-180)
    auto loop = [&] {
        for (int t = 1; t <= b->nt; ++t) {
            if (save_boundary) strips2(b, F.vx, F.vz, t, 1);
            int ev = b->prof.begin(st, 2, 0);
            hipLaunchKernelGGL(k2e_record, dim3(rg, nshot), dim3(256), 0, st, b->d_shots.get(), F.vx, F.vz,
                               b->d_rec_beg.get(), b->d_rec.get(), b->d_rec_w.get(), b->d_traces.get(), b->nt, t);
            b->prof.end(st, ev);
            if (t == b->nt) break;  // the last update changes nothing that is recorded
            step2<R>(b, F, out_of(F), b->dt, bs, bv);
            ev = b->prof.begin(st, 2, 0);
            hipLaunchKernelGGL(k2e_inject, dim3(nshot), dim3(1024), 0, st, b->d_shots.get(), F.vx, F.vz, F.sxx,
                               F.szz, b->d_inj_beg.get(), b->d_inj.get(), b->d_amp.get(), b->nt, t);
            b->prof.end(st, ev);
            if (!b->graph.capturing) maybe_collect(b);
        }
    };
    const bool graph_ok = b->graph.enabled && !b->prof.on() && !save_boundary;
    OWL_HIP(hipEventRecord(b->ev0, st));
    if (graph_ok)
        b->graph.run(st, forward_signature2(b), loop);
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

// OWL's reverse pass (submodule_fwi_gradient.f90:138-270): t = nt .. hnt:
// prev = state; state = update(state, -dt); at t = nt the final state;
// boundary record t; adjoint = update(adjoint, -dt); adjoint sources;
// imaging.  The forward field steps out of place so the previous state
// stays in the other buffer; at t = nt the update of the zero field is the
// zero field, so the final state is used as is with prev = 0.  OWL's loop
// continues the adjoint below hnt without imaging; those steps do not touch
// the medium gradients and are left out.
template <int R>
void run_reverse2(owl_elastic2d* b) {
    hipStream_t st = b->stream;
    if (!b->have_boundary || !b->fwd_state_valid)
        fail(OWL_ERR_STATE, "the adjoint needs a preceding forward run with boundary saving");
    if (!b->d_adj.get() || b->d_adj.size() < b->trace_elems) fail(OWL_ERR_STATE, "no adjoint source");
    for (int k = 5; k < 13; ++k) b->fld[k].zero(st);  // the reverse run's memory variables start at zero
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
    sweep_bytes2(b, bs, bv);
    EF2 X = fields2(b->fld);
    EV2 Y = ev_of(b->fB);
    const EF2 A = fields2(b->aF);
    const EV2 adj = out_of(A);
    // imaging: 15 wavefields, c11, c13, c55, three FP64 accumulators r+w
    const double image_bytes = (15.0 * 4 + 12.0 + 48.0) * double(b->grad_elems);
    const unsigned ig = unsigned(std::max<size_t>(1, std::min<size_t>(2048, (b->grad_elems / nshot + 255) / 256)));
    OWL_HIP(hipEventRecord(b->ev0, st));
    for (int t = b->nt; t >= hnt; --t) {
        EV2 cur = out_of(X);
        const EV2 prev = out_of(X);
        if (t != b->nt) {
            step2<R>(b, X, Y, -b->dt, bs, bv);
            cur = Y;
        }
        strips2(b, cur.vx, cur.vz, t, 0);
        step2<R>(b, A, adj, -b->dt, bs, bv);
        int ev = b->prof.begin(st, 2, 0);
        if (b->max_nadj > 0)
            hipLaunchKernelGGL(k2e_inject_adjoint, dim3(unsigned((b->max_nadj + 255) / 256), nshot), dim3(256), 0, st,
                               b->d_shots.get(), adj.vx, adj.vz, b->d_adj_beg.get(), b->d_adj_taps.get(),
                               b->d_rec_w.get(), b->d_adj.get(), b->nt, t);
        b->prof.end(st, ev);
        if (t % cc == 0) {
            ev = b->prof.begin(st, 3, image_bytes);
            hipLaunchKernelGGL(k2e_image, dim3(ig, nshot), dim3(256), 0, st, b->d_shots.get(), cur, prev, adj,
                               b->med[0].get(), b->med[1].get(), b->med[5].get(), b->d_gA.get(), b->d_gB.get(),
                               b->d_gC.get(), t == b->nt ? 1 : 0, b->pml, R, t);
            b->prof.end(st, ev);
        }
        if (t != b->nt) {
            const EV2 old = out_of(X);
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

void traces_to_host2(const owl_elastic2d* b, const float* dev, float* out) {
    std::vector<float> h(b->trace_elems);
    OWL_HIP(hipMemcpy(h.data(), dev, h.size() * sizeof(float), hipMemcpyDeviceToHost));
    // device [2][nt][nr] per shot -> host [2][nr][nt] per shot
    for (auto& S : b->sd)
        for (int c = 0; c < 2; ++c)
            for (int r = 0; r < S.nr; ++r)
                for (int t = 0; t < b->nt; ++t)
                    out[size_t(S.trace0) + (size_t(c) * S.nr + r) * b->nt + t] =
                        h[size_t(S.trace0) + size_t(c) * b->nt * S.nr + size_t(t) * S.nr + r];
}

void traces_to_device2(owl_elastic2d* b, const float* in, DevBuf<float>& dev) {
    std::vector<float> h(b->trace_elems);
    for (auto& S : b->sd)
        for (int c = 0; c < 2; ++c)
            for (int r = 0; r < S.nr; ++r)
                for (int t = 0; t < b->nt; ++t)
                    h[size_t(S.trace0) + size_t(c) * b->nt * S.nr + size_t(t) * S.nr + r] =
                        in[size_t(S.trace0) + (size_t(c) * S.nr + r) * b->nt + t];
    dev.alloc(std::max<size_t>(h.size(), 1));
    dev.upload(h.data(), h.size(), b->stream);
    OWL_HIP(hipStreamSynchronize(b->stream));
}

// Chain rule (submodule_fwi_gradient.f90:330-360) with vp, vs in km/s and
// rho in g/cm^3: grad_vp = 2 rho vp A, grad_vs = 2 rho vs B, grad_rho = C.
inline void chain_rule2r(float vp, float vs, double r, double A, double B, double C, double& gv, double& gs,
                         double& gr) {
    const double p = double(vp * 1.0e-3f), v = double(vs * 1.0e-3f);
    gv = 2.0 * r * p * A;
    gs = 2.0 * r * v * B;
    gr = C;
}
inline void chain_rule2(float vp, float vs, float rho, double A, double B, double C, double& gv, double& gs,
                        double& gr) {
    chain_rule2r(vp, vs, double(rho * 1.0e-3f), A, B, C, gv, gs, gr);
}

// Free surface: shot s's accumulators on its refined rows to the window's
// rows (map_irregular_to_regular, submodule_fwi_gradient.f90:324-338): A,
// B, C column by column through z_i(1 .. nz), and rho x 1e-3 through all
// refined rows -pml .. nz+pml+1; then the chain rule with the window's vp,
// vs.  Outputs (nz_w, nx_w) z fastest.
void window_grads_fs(const owl_elastic2d* b, size_t s, const double* gA, const double* gB, const double* gC,
                     std::vector<double>& gv, std::vector<double>& gs, std::vector<double>& gr) {
    const ShotE2& S = b->sd[s];
    const ShotHostE2& H = b->hs[s];
    const fsm::SsgMesh& M = b->mesh[s];
    const int nx = H.nx, nz0 = H.nz, nzr = S.nz, pml = b->pml, nrj = nzr + 2 * pml + 2;
    gv.assign(size_t(nx) * nz0, 0.0);
    gs.assign(gv.size(), 0.0);
    gr.assign(gv.size(), 0.0);
    const size_t nr_ = static_cast<size_t>(nzr), n0 = static_cast<size_t>(nz0), nj = static_cast<size_t>(nrj);
    std::vector<double> ca(nr_), cb(nr_), cc(nr_), ma(n0), mb(n0), mc(n0), rr(nj), mr(n0);
    for (int i = 1; i <= nx; ++i) {
        for (int j = 1; j <= nzr; ++j) {
            const size_t o = size_t(S.g0) + size_t(i - 1) + size_t(j - 1) * nx;
            ca[size_t(j - 1)] = gA[o];
            cb[size_t(j - 1)] = gB[o];
            cc[size_t(j - 1)] = gC[o];
        }
        fsm::column_to_regular_ssg(M, nz0, ca.data(), ma.data());
        fsm::column_to_regular_ssg(M, nz0, cb.data(), mb.data());
        fsm::column_to_regular_ssg(M, nz0, cc.data(), mc.data());
        const std::vector<float>& rm = b->rho_map[s];
        for (int k = 0; k < nrj; ++k) rr[size_t(k)] = double(rm[size_t(i - 1) * nrj + k] * 1.0e-3f);
        fsm::spline_to_regular(M.zi.data(), rr.data(), nrj, M.dz, nz0, mr.data());
        for (int k = 1; k <= nz0; ++k) {
            const size_t m = model_index2(b, H, i, k), o = size_t(k - 1) + size_t(i - 1) * nz0;
            chain_rule2r(b->vp_full[m], b->vs_full[m], double(float(mr[size_t(k - 1)])), ma[size_t(k - 1)],
                         mb[size_t(k - 1)], mc[size_t(k - 1)], gv[o], gs[o], gr[o]);
        }
    }
}

}  // namespace

extern "C" {

owl_status owl_elastic2d_create(const owl_acoustic_config* cfg, const float* vp, const float* vs, const float* rho,
                                int nshot, const owl_shot* shots, owl_elastic2d** out) {
    return owl_elastic2d_create_surface(cfg, vp, vs, rho, nshot, shots, nullptr, out);
}

owl_status owl_elastic2d_create_surface(const owl_acoustic_config* cfg, const float* vp, const float* vs,
                                        const float* rho, int nshot, const owl_shot* shots,
                                        const owl_elastic_surface* surface, owl_elastic2d** out) {
    return guarded([&] {
        if (!out) fail(OWL_ERR_ARG, "out is null");
        *out = nullptr;
        auto* b = new owl_elastic2d();
        try {
            plan2(b, cfg, vp, vs, rho, nshot, shots, surface);
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

int owl_elastic2d_rows(const owl_elastic2d* b, int shot) {
    if (!b || shot < 0 || size_t(shot) >= b->sd.size()) return -1;
    return b->sd[size_t(shot)].nz;
}

void owl_elastic2d_destroy(owl_elastic2d* b) { delete b; }

int owl_elastic2d_nt(const owl_elastic2d* b) { return b ? b->nt : 0; }

owl_status owl_elastic2d_forward(owl_elastic2d* b, float* traces, int save_boundary) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        dispatch_R2e(b->R, [&](auto r) { run_forward2<decltype(r)::value>(b, save_boundary != 0); });
        if (traces) traces_to_host2(b, b->d_traces.get(), traces);
    });
}

owl_status owl_elastic2d_adjoint(owl_elastic2d* b, const float* adj, float* grad_vp, float* grad_vs,
                                 float* grad_rho) {
    return guarded([&] {
        if (!b || !adj) fail(OWL_ERR_ARG, "null argument");
        traces_to_device2(b, adj, b->d_adj);
        dispatch_R2e(b->R, [&](auto r) { run_reverse2<decltype(r)::value>(b); });
        std::vector<double> gA(b->grad_elems), gB(b->grad_elems), gC(b->grad_elems);
        OWL_HIP(hipMemcpy(gA.data(), b->d_gA.get(), gA.size() * sizeof(double), hipMemcpyDeviceToHost));
        OWL_HIP(hipMemcpy(gB.data(), b->d_gB.get(), gB.size() * sizeof(double), hipMemcpyDeviceToHost));
        OWL_HIP(hipMemcpy(gC.data(), b->d_gC.get(), gC.size() * sizeof(double), hipMemcpyDeviceToHost));
        // per-shot window gradients (nz_s, nx_s), z fastest, NaN/Inf -> 0
        size_t wo = 0;
        for (size_t s = 0; s < b->sd.size(); ++s) {
            const ShotE2& S = b->sd[s];
            const ShotHostE2& H = b->hs[s];
            if (b->fs) {
                std::vector<double> gv, gs, gr;
                window_grads_fs(b, s, gA.data(), gB.data(), gC.data(), gv, gs, gr);
                for (size_t e = 0; e < gv.size(); ++e) {
                    const float a = float(gv[e]), c = float(gs[e]), f = float(gr[e]);
                    if (grad_vp) grad_vp[wo + e] = std::isfinite(a) ? a : 0.f;
                    if (grad_vs) grad_vs[wo + e] = std::isfinite(c) ? c : 0.f;
                    if (grad_rho) grad_rho[wo + e] = std::isfinite(f) ? f : 0.f;
                }
                wo += gv.size();
                continue;
            }
            wo += size_t(S.nx) * S.nz;
            for (int k = 0; k < S.nz; ++k)
                for (int i = 0; i < S.nx; ++i) {
                    const size_t src = size_t(S.g0) + i + size_t(k) * S.nx;
                    const size_t dst = size_t(S.g0) + k + size_t(i) * S.nz;
                    const size_t m = model_index2(b, H, i + 1, k + 1);
                    double gv, gs, gr;
                    chain_rule2(b->vp_full[m], b->vs_full[m], b->rho_full[m], gA[src], gB[src], gC[src], gv, gs, gr);
                    const float a = float(gv), c = float(gs), e = float(gr);
                    if (grad_vp) grad_vp[dst] = std::isfinite(a) ? a : 0.f;
                    if (grad_vs) grad_vs[dst] = std::isfinite(c) ? c : 0.f;
                    if (grad_rho) grad_rho[dst] = std::isfinite(e) ? e : 0.f;
                }
        }
    });
}

owl_status owl_elastic2d_gradient(owl_elastic2d* b, const float* obs, double* misfit, double* grad_vp,
                                  double* grad_vs, double* grad_rho, float* traces_out) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        if (obs) {
            traces_to_device2(b, obs, b->d_obs);
            b->have_obs = true;
        }
        if (!b->have_obs) fail(OWL_ERR_STATE, "no observed data");
        dispatch_R2e(b->R, [&](auto r) { run_forward2<decltype(r)::value>(b, true); });
        const double fwd_ms = b->last_ms, fwd_up = b->last_updates;
        hipStream_t st = b->stream;
        const unsigned nshot = unsigned(b->sd.size());
        b->d_adj.alloc(std::max<size_t>(b->trace_elems, 1));
        b->d_misfit.alloc(2 * size_t(nshot));
        b->d_scratch.alloc(std::max<size_t>(4 * size_t(b->nrec_total), 1));
        hipLaunchKernelGGL(k_adjoint_source_2c<ShotE2>, dim3(nshot, 2), dim3(256), 0, st, b->d_shots.get(), b->d_traces.get(),
                           b->d_obs.get(), b->d_rec_w.get(), b->d_adj.get(), b->d_misfit.get(), b->nt, b->cfg.trace_discard_threshold,
                           b->d_scratch.get());
        dispatch_R2e(b->R, [&](auto r) { run_reverse2<decltype(r)::value>(b); });
        const size_t nfull = size_t(b->cfg.nx) * b->cfg.nz;
        if (!b->fs) {
            for (auto* g : {&b->d_fA, &b->d_fB, &b->d_fC}) g->alloc(nfull);
            const unsigned mg = unsigned(std::min<size_t>(8192, (nfull + 255) / 256));
            hipLaunchKernelGGL(k2e_merge, dim3(mg), dim3(256), 0, st, b->d_shots.get(), b->d_win.get(), int(nshot),
                               b->d_gA.get(), b->d_gB.get(), b->d_gC.get(), b->d_fA.get(), b->d_fB.get(),
                               b->d_fC.get(), b->cfg.nx, b->cfg.nz);
        }
        OWL_HIP(hipGetLastError());
        OWL_HIP(hipStreamSynchronize(st));
        if (misfit) {
            std::vector<double> m2(2 * size_t(nshot));
            OWL_HIP(hipMemcpy(m2.data(), b->d_misfit.get(), m2.size() * sizeof(double), hipMemcpyDeviceToHost));
            for (unsigned s = 0; s < nshot; ++s) misfit[s] = m2[2 * s] + m2[2 * s + 1];
        }
        if (b->fs && (grad_vp || grad_vs || grad_rho)) {
            // each shot mapped back to its window and through the chain rule
            // with its own mapped density, then summed on the model
            std::vector<double> gA(b->grad_elems), gB(b->grad_elems), gC(b->grad_elems);
            OWL_HIP(hipMemcpy(gA.data(), b->d_gA.get(), gA.size() * sizeof(double), hipMemcpyDeviceToHost));
            OWL_HIP(hipMemcpy(gB.data(), b->d_gB.get(), gB.size() * sizeof(double), hipMemcpyDeviceToHost));
            OWL_HIP(hipMemcpy(gC.data(), b->d_gC.get(), gC.size() * sizeof(double), hipMemcpyDeviceToHost));
            std::vector<double> sv(nfull, 0.0), ss(nfull, 0.0), sr(nfull, 0.0);
            for (size_t s = 0; s < b->sd.size(); ++s) {
                const ShotHostE2& H = b->hs[s];
                std::vector<double> gv, gs, gr;
                window_grads_fs(b, s, gA.data(), gB.data(), gC.data(), gv, gs, gr);
                for (int i = 1; i <= H.nx; ++i)
                    for (int k = 1; k <= H.nz; ++k) {
                        const size_t m = model_index2(b, H, i, k), o = size_t(k - 1) + size_t(i - 1) * H.nz;
                        sv[m] += gv[o];
                        ss[m] += gs[o];
                        sr[m] += gr[o];
                    }
            }
            for (size_t e = 0; e < nfull; ++e) {
                if (grad_vp) grad_vp[e] = std::isfinite(sv[e]) ? sv[e] : 0.0;
                if (grad_vs) grad_vs[e] = std::isfinite(ss[e]) ? ss[e] : 0.0;
                if (grad_rho) grad_rho[e] = std::isfinite(sr[e]) ? sr[e] : 0.0;
            }
        } else if (grad_vp || grad_vs || grad_rho) {
            std::vector<double> fA(nfull), fB(nfull), fC(nfull);
            OWL_HIP(hipMemcpy(fA.data(), b->d_fA.get(), nfull * sizeof(double), hipMemcpyDeviceToHost));
            OWL_HIP(hipMemcpy(fB.data(), b->d_fB.get(), nfull * sizeof(double), hipMemcpyDeviceToHost));
            OWL_HIP(hipMemcpy(fC.data(), b->d_fC.get(), nfull * sizeof(double), hipMemcpyDeviceToHost));
            for (size_t e = 0; e < nfull; ++e) {
                double gv, gs, gr;
                chain_rule2(b->vp_full[e], b->vs_full[e], b->rho_full[e], fA[e], fB[e], fC[e], gv, gs, gr);
                if (grad_vp) grad_vp[e] = std::isfinite(gv) ? gv : 0.0;
                if (grad_vs) grad_vs[e] = std::isfinite(gs) ? gs : 0.0;
                if (grad_rho) grad_rho[e] = std::isfinite(gr) ? gr : 0.0;
            }
        }
        if (traces_out) traces_to_host2(b, b->d_traces.get(), traces_out);
        b->last_ms += fwd_ms;
        b->last_updates += fwd_up;
    });
}

owl_status owl_elastic2d_set_graphs(owl_elastic2d* b, int on) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        b->graph.enabled = on != 0;
        b->graph.reset();
    });
}

// Launch-shape autotuner: times `steps` steps of the sweeps on the batch
// under each candidate (rows per tile column, tiles per block) and keeps the
// fastest; the fields are zeroed afterwards (every forward starts from
// zero).  Stencil timing does not depend on the field values.
owl_status owl_elastic2d_autotune(owl_elastic2d* b, int steps, int* planes, int* rows) {
    return guarded([&] {
        if (!b || steps < 1) fail(OWL_ERR_ARG, "null batch or no steps");
        static const int kPlanes[] = {4, 8, 12, 16, 32, 64, 128};
        static const int kRows[] = {1, 2, 4};
        hipStream_t st = b->stream;
        const bool prof_on = b->prof.on();
        b->prof.enable(false);
        for (auto& f : b->fld) f.zero(st);
        const EF2 F = fields2(b->fld);
        double bs, bv;
        sweep_bytes2(b, bs, bv);
        float best = -1.f;
        int bk = b->kc, by = b->cy;
        auto shape = [&](int kc, int cy) {
            b->kc = kc;
            b->cy = cy;
            retile2(b);
            b->d_shots.upload(b->sd, st);
            b->d_tile_shot.upload(b->tile_shot, st);
        };
        for (int kc : kPlanes)
            for (int cy : kRows) {
                shape(kc, cy);
                dispatch_R2e(b->R, [&](auto r) {
                    step2<decltype(r)::value>(b, F, out_of(F), b->dt, bs, bv);  // warm-up
                    OWL_HIP(hipEventRecord(b->ev0, st));
                    for (int t = 0; t < steps; ++t) step2<decltype(r)::value>(b, F, out_of(F), b->dt, bs, bv);
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
        shape(bk, by);
        for (auto& f : b->fld) f.zero(st);
        OWL_HIP(hipStreamSynchronize(st));
        b->prof.enable(prof_on);
        b->fwd_state_valid = false;
        b->have_boundary = false;
        if (planes) *planes = bk;
        if (rows) *rows = by;
    });
}

owl_status owl_elastic2d_set_launch(owl_elastic2d* b, int planes, int rows) {
    return guarded([&] {
        // blocks of 64 x rows threads, one tile per 64-thread row
        if (!b || planes < 1 || (rows != 1 && rows != 2 && rows != 4)) fail(OWL_ERR_ARG, "bad launch shape");
        b->kc = planes;
        b->cy = rows;
        retile2(b);
        b->d_shots.upload(b->sd, b->stream);
        b->d_tile_shot.upload(b->tile_shot, b->stream);
        OWL_HIP(hipStreamSynchronize(b->stream));
    });
}

owl_status owl_elastic2d_set_record_storage(owl_elastic2d* b, int kind) {
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

owl_status owl_elastic2d_set_profiling(owl_elastic2d* b, int on) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        b->prof.enable(on != 0);
    });
}

owl_status owl_elastic2d_kernel_stats(const owl_elastic2d* b, int kind, double* total_ms, long long* launches,
                                      double* bytes_total) {
    return guarded([&] {
        if (!b || kind < 0 || kind >= Profiler::kKinds) fail(OWL_ERR_ARG, "bad kernel kind");
        if (total_ms) *total_ms = b->prof.ms(kind);
        if (launches) *launches = b->prof.launches(kind);
        if (bytes_total) *bytes_total = b->prof.bytes(kind);
    });
}

owl_status owl_elastic2d_last_stats(const owl_elastic2d* b, double* ms, double* updates) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        if (ms) *ms = b->last_ms;
        if (updates) *updates = b->last_updates;
    });
}

}  // extern "C"
