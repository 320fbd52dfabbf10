// 2D Lebedev fully-staggered-grid (FSG) elastic time stepping for general
// (TTI) anisotropy and OWL's gradient on it, for MI355X (gfx950): OWL's
// elastic_tti_2d (module_wavefield.f90:35-252 update_wavefield, :593-736
// add_source, module_cfspml.f90:80-576 M-PML, submodule_forward_modeling.f90
// receivers, module_boundary_saving.f90, submodule_fwi_gradient.f90
// compute_adjoint and compute_gradient) for a batch of shots, each on its
// own window, in every launch.  Moduli given as c11, c13, c15, c33, c35,
// c55 (OWL's 'cij'), force and explosion sources, two-component velocity
// receivers (half of each velocity grid), CFS-MPML with the damping ratios
// of the Christoffel scan, and OWL's topography-following free surface on
// its depth-refined mesh (update_wavefield_free_surface,
// module_wavefield.f90:257-587; the mesh in fs_mesh.cpp).
//
// Grids, in storage index (i, j) of one padded array per field:
//   ii  stresses at (i, j)                       sxx_ii szz_ii sxz_ii
//   hh  stresses at (i - 1/2, j - 1/2)           sxx_hh szz_hh sxz_hh
//   hi  velocities at (i - 1/2, j)               vx_hi  vz_hi
//   ih  velocities at (i, j - 1/2)               vx_ih  vz_ih
// x fastest, rows padded to a multiple of 64 floats.  Each sweep tiles x
// with 64-wide thread rows and marches along z with register queues for the
// z stencils.  FP32 in the reference's operation order (derivatives divide
// by dx and dz, no FMA contraction); gradients in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "common.hpp"
#include "fs_mesh.hpp"
#include "host_prep.hpp"
#include "owl_amd.h"
#include "profiler.hpp"
#include "record_store.hpp"
#include "trace_misfit.hpp"

namespace owl {
namespace fsg {

constexpr int kPitchAlign = 64;

struct Coef {
    float c[8];
};

struct ShotF {
    int nx, nz;
    int px, nzp;           // row pitch (floats), padded rows
    long long off;         // element offset of node (lo, lo)
    int tile0, ntx, nzc;
    int nr, rec0;
    int ninj, inj0;
    int nadj, adj0;
    long long trace0;      // two [nt][nr] blocks (vx, vz)
    long long strip0;      // boundary records: nt blocks
    long long g0;          // window gradient accumulators (nx*nz, x fastest)
    int hnt;               // last reverse step OWL images for this shot (its first source's hnt)
    // free surface: the z spacing the z derivatives divide by, eta_max,
    // depth_max, and the offset of the shot's metric tables in the table
    // buffer: per storage row (eta_i, eta_h, sc_i, sc_h), then per storage
    // column (topo_i, slope_i, topo_h, slope_h)
    float fdz, eta_max, depth_max;
    long long met0;
};

// Wavefields (0..9) and CFS-PML memory variables (10..27), named as in OWL.
constexpr int kNW = 10, kNM = 18, kNF = kNW + kNM;
struct FF {
    float *sxx_ii, *szz_ii, *sxz_ii, *sxx_hh, *szz_hh, *sxz_hh, *vx_hi, *vz_hi, *vx_ih, *vz_ih;
    float *m_dxvx_ii, *m_dxvz_ii, *m_dzvx_ii, *m_dzvz_ii;  // stress sweep, integer grid
    float *m_dxvx_hh, *m_dxvz_hh, *m_dzvx_hh, *m_dzvz_hh;  // stress sweep, corner grid
    float *m_dxxx_hi, *m_dxxz_hi, *m_dzxz_hi, *m_dzzz_hi;  // velocity sweep, (i-1/2, j)
    float *m_dxxx_ih, *m_dxxz_ih, *m_dzxz_ih, *m_dzzz_ih;  // velocity sweep, (i, j-1/2)
    float *m_dzxx_hi, *m_dzxx_ih;                          // free surface only (memory_pdzxx_*)
};

// The ten wavefields a step writes.
struct FV {
    float *sxx_ii, *szz_ii, *sxz_ii, *sxx_hh, *szz_hh, *sxz_hh, *vx_hi, *vz_hi, *vx_ih, *vz_ih;
};

// Moduli on the integer grid and their 4-point means on the corner grid
// (module_wavefield.f90:166-171), |dt| over the 2-point densities at the
// velocity positions (:205, :238).
struct FM {
    const float* ci[6];
    const float* ch[6];
    const float *qhx, *qhz;
};

// Per-cell CFS-MPML coefficients (a, b) in x and z on the four grids
// (module_cfspml.f90:520-567), zero outside the damping frame.
struct FP {
    const float *axii, *bxii, *azii, *bzii;
    const float *axhh, *bxhh, *azhh, *bzhh;
    const float *axhi, *bxhi, *azhi, *bzhi;
    const float *axih, *bxih, *azih, *bzih;
};

struct Tap {
    long long cell;
    float w;   // w1 w2 for sources, w1 w2 weight 0.5 for receivers
    float w2;  // second weight of a source tap
    int row;
    int tbeg, tend;
    int field;  // 0 vx_hi, 1 vz_hi, 2 vx_ih, 3 vz_ih, 4 sxx_ii, 5 szz_ii, 6 sxx_hh, 7 szz_hh
};

struct Targets {
    float* f[8];
};

__device__ __forceinline__ size_t atf(const ShotF& S, int lo, int i, int j) {
    return size_t(S.off) + size_t(i - lo) + size_t(j - lo) * size_t(S.px);
}

// Load and store at a 32-bit byte offset from a kernel-argument base: a
// batch's buffers stay below 2 GiB (checked on the host), so every access of
// a sweep shares one offset register against scalar bases.
__device__ __forceinline__ float ldo(const float* base, unsigned off) {
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + off);
}
__device__ __forceinline__ void sto(float* base, unsigned off, float v) {
    *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + off) = v;
}

// m = a d + b m; d = (d + m) / kappa with kappa = 1
__device__ __forceinline__ float mpml(float* m, unsigned qb, float d, float a, float b) {
    const float v = a * d + b * ldo(m, qb);
    sto(m, qb, v);
    return d + v;
}

// the z stencil from a queue: slot k holds row j + k - R
template <int R>
__device__ __forceinline__ float sten_q(const float* qv, int off, const Coef& c) {
    float s = c.c[0] * (qv[R + off + 1] - qv[R + off]);
#pragma unroll
    for (int m = 2; m <= R; ++m) s = s + c.c[m - 1] * (qv[R + off + m] - qv[R + off + 1 - m]);
    return s;
}

// pdxw_stencil(w, i + off, j) (macro_fd_stencil.f90) from the lanes' own
// values: xv[k] is the value at column
// i - R + off + 1 + k, taken from lane l - R + off + 1 + k; returns
// sum_m c_m (w(i+off+m) - w(i+off-m+1)) in the macro's order.
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

// ---------------------------------------------------------------------------
// Stress sweep (module_wavefield.f90:87-186): the integer-grid stresses at
// (i, j) from dx of vx_hi, vz_hi and dz of vx_ih, vz_ih; the corner stresses
// at storage (i, j) from dx of vx_ih, vz_ih and dz of vx_hi, vz_hi, each
// one node back.  Reads velocities and old stresses from F, writes O.
// One wavefront per tile of 64 - 2R columns: its 64 lanes hold columns
// i0 - R .. i0 + 63 - R, each with register queues of the four velocities
// over rows j - R .. j + R; the x stencils take the neighbouring lanes'
// current rows by shuffles, so lanes R .. 63 - R update and the others only
// feed them.
// ---------------------------------------------------------------------------
// Free-surface metric of one cell (update_wavefield_free_surface,
// module_wavefield.f90:330-333): tpaz = max(0, eta_max - eta)/(topo +
// depth_max) * slope.  A zero slope gives the slope itself, as the product
// does, without the division.
__device__ __forceinline__ float fs_tpaz(float eta_max, float eta, float den, float slope) {
    if (slope == 0.0f) return slope;
    const float a = eta_max - eta > 0.0f ? eta_max - eta : 0.0f;
    return a / den * slope;
}

template <int R, bool FS>
__global__ __launch_bounds__(256) void k_fsg_stress(const ShotF* __restrict__ shots, const int* __restrict__ tile_shot,
                                                    FF F, FV O, FM M, FP P, Coef c, float dt, float dx, float dz,
                                                    int pml, int kc, const float* __restrict__ met) {
    constexpr int W = 64 - 2 * R;
    // one wavefront per tile (blockDim.x = 64): tile, shot and row indices
    // are wavefront-uniform and live in scalar registers
    const int tile = __builtin_amdgcn_readfirstlane(int(blockIdx.x) * int(blockDim.y) + int(threadIdx.y));
    const ShotF S = shots[tile_shot[tile]];
    const int lt = tile - S.tile0;
    const int tx = lt % S.ntx, tz = lt / S.ntx;
    const int nx = S.nx, nz = S.nz, lo = 1 - pml - R;
    const int l = int(threadIdx.x);
    const int i = -pml + 1 + tx * W - R + l;
    // with the free surface the rows start at the surface (j = 1)
    const int j0 = (FS ? 1 : -pml + 1) + tz * kc, j1 = min(j0 + kc - 1, nz + pml);
    if (j0 > j1) return;  // padding tile beyond the shot's rows (whole wavefront)
    const bool mcol = i <= nx + pml + R;                 // inside the padded grid
    const bool own = l >= R && l <= 63 - R && i <= nx + pml;
    const unsigned szb = unsigned(S.px) * 4u;
    unsigned q = unsigned(atf(S, lo, mcol ? i : lo, j0)) * 4u;  // byte offset
    // queues over rows j-R .. j+R
    float qxh[2 * R + 1], qzh[2 * R + 1], qxi[2 * R + 1], qzi[2 * R + 1];  // vx_hi vz_hi vx_ih vz_ih
#pragma unroll
    for (int m = 0; m < 2 * R; ++m) {
        const unsigned r = q + unsigned(m - R) * szb;
        qxh[m] = mcol ? ldo(F.vx_hi, r) : 0.0f;
        qzh[m] = mcol ? ldo(F.vz_hi, r) : 0.0f;
        qxi[m] = mcol ? ldo(F.vx_ih, r) : 0.0f;
        qzi[m] = mcol ? ldo(F.vz_ih, r) : 0.0f;
    }
    const bool xi = i <= 0 || i >= nx + 1, xh = i <= 1 || i >= nx + 1;
    // free surface: the column's topography terms; integer columns take
    // topo_i (ii grid), half columns topo_h (hh grid)
    float den_i = 1.0f, den_h = 1.0f, sl_i = 0.0f, sl_h = 0.0f, tpc_i = 1.0f, tpc_h = 1.0f;
    const float dzk = FS ? S.fdz : dz;
    if constexpr (FS) {
        if (own) {
            const float* col = met + S.met0 + 4LL * S.nzp + 4LL * (i - lo);
            den_i = col[0] + S.depth_max;
            sl_i = col[1];
            den_h = col[2] + S.depth_max;
            sl_h = col[3];
            tpc_i = S.eta_max / den_i;
            tpc_h = S.eta_max / den_h;
        }
    }
    // The inputs of one row: the queues' fronts (row j + R) and, on the
    // owned lanes, both grids' moduli and old stresses.  Row j + 1's are
    // loaded while row j is computed.
    struct Row {
        float fxh, fzh, fxi, fzi;
        float ci[6], ch[6], si[3], sh[3];
    };
    auto load = [&](unsigned qr, Row& w) {
        const unsigned r = qr + unsigned(R) * szb;
        w.fxh = mcol ? ldo(F.vx_hi, r) : 0.0f;
        w.fzh = mcol ? ldo(F.vz_hi, r) : 0.0f;
        w.fxi = mcol ? ldo(F.vx_ih, r) : 0.0f;
        w.fzi = mcol ? ldo(F.vz_ih, r) : 0.0f;
        if (own) {
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                w.ci[k] = ldo(M.ci[k], qr);
                w.ch[k] = ldo(M.ch[k], qr);
            }
            w.si[0] = ldo(F.sxx_ii, qr);
            w.si[1] = ldo(F.szz_ii, qr);
            w.si[2] = ldo(F.sxz_ii, qr);
            w.sh[0] = ldo(F.sxx_hh, qr);
            w.sh[1] = ldo(F.szz_hh, qr);
            w.sh[2] = ldo(F.sxz_hh, qr);
        }
    };
    Row cur{}, nxt{};
    load(q, cur);
    for (int j = j0; j <= j1; ++j, q += szb) {
        if (j < j1) load(q + szb, nxt);
        qxh[2 * R] = cur.fxh;
        qzh[2 * R] = cur.fzh;
        qxi[2 * R] = cur.fxi;
        qzi[2 * R] = cur.fzi;
        const float sxh = sten_lanes<R>(qxh[R], l, 0, c), szh = sten_lanes<R>(qzh[R], l, 0, c);
        const float sxi = sten_lanes<R>(qxi[R], l, -1, c), szi = sten_lanes<R>(qzi[R], l, -1, c);
        if (own) {
            {   // integer grid
                const float c11 = cur.ci[0], c13 = cur.ci[1], c15 = cur.ci[2];
                const float c33 = cur.ci[3], c35 = cur.ci[4], c55 = cur.ci[5];
                float pdxvx = sxh / dx;
                float pdxvz = szh / dx;
                if constexpr (FS) {
                    const float* rowt = met + S.met0 + 4LL * (j - lo);
                    const float tpaz = fs_tpaz(S.eta_max, rowt[0], den_i, sl_i), tpcz = tpc_i;
                    float pdzvx, pdzvz;
                    if (j == 1) {
                        if (xi) {
                            const float ax = ldo(P.axii, q), bx = ldo(P.bxii, q);
                            pdxvx = mpml(F.m_dxvx_ii, q, pdxvx, ax, bx);
                            pdxvz = mpml(F.m_dxvz_ii, q, pdxvz, ax, bx);
                        }
                        // z derivatives from the traction-free condition
                        const float hx = sl_i;
                        const float d11 = tpcz * (hx * hx * c11 + hx * 2.0f * c15 + c55);
                        const float d12 = tpcz * (hx * hx * c15 + hx * (c13 + c55) + c35);
                        const float d22 = tpcz * (hx * hx * c55 + hx * 2.0f * c35 + c33);
                        const float dd = d12 * d12 - d11 * d22;
                        const float f1 = (-(hx * c11) - c15) * pdxvx + (-(hx * c15) - c55) * pdxvz;
                        const float f2 = (-(hx * c15) - c13) * pdxvx + (-(hx * c55) - c35) * pdxvz;
                        pdzvx = (-(d22 * f1) + d12 * f2) / dd;
                        pdzvz = (-(d11 * f2) + d12 * f1) / dd;
                    } else {
                        const float sc = rowt[2];
                        pdzvx = sten_q<R>(qxi, 0, c) / dzk * sc;
                        pdzvz = sten_q<R>(qzi, 0, c) / dzk * sc;
                        if (xi || j >= nz + 1) {
                            const float ax = ldo(P.axii, q), bx = ldo(P.bxii, q), az = ldo(P.azii, q),
                                        bz = ldo(P.bzii, q);
                            pdxvx = mpml(F.m_dxvx_ii, q, pdxvx, ax, bx);
                            pdxvz = mpml(F.m_dxvz_ii, q, pdxvz, ax, bx);
                            pdzvx = mpml(F.m_dzvx_ii, q, pdzvx, az, bz);
                            pdzvz = mpml(F.m_dzvz_ii, q, pdzvz, az, bz);
                        }
                    }
                    const float pd1 = pdxvx + pdzvx * tpaz, pd2 = pdzvz * tpcz;
                    const float pd3 = pdzvx * tpcz + pdxvz + pdzvz * tpaz;
                    const float vxx = cur.si[0] + dt * (c11 * pd1 + c13 * pd2 + c15 * pd3);
                    const float vzz = cur.si[1] + dt * (c13 * pd1 + c33 * pd2 + c35 * pd3);
                    const float vxz = cur.si[2] + dt * (c15 * pd1 + c35 * pd2 + c55 * pd3);
                    sto(O.sxx_ii, q, vxx);
                    sto(O.szz_ii, q, vzz);
                    sto(O.sxz_ii, q, vxz);
                    if (j >= 2 && j <= R + 1) {  // odd mirror: row 2 - j
                        const unsigned qm = q - unsigned(2 * j - 2) * szb;
                        sto(O.sxx_ii, qm, -vxx);
                        sto(O.szz_ii, qm, -vzz);
                        sto(O.sxz_ii, qm, -vxz);
                    }
                } else {
                    float pdzvx = sten_q<R>(qxi, 0, c) / dz;
                    float pdzvz = sten_q<R>(qzi, 0, c) / dz;
                    if (xi || j <= 0 || j >= nz + 1) {
                        const float ax = ldo(P.axii, q), bx = ldo(P.bxii, q), az = ldo(P.azii, q), bz = ldo(P.bzii, q);
                        pdxvx = mpml(F.m_dxvx_ii, q, pdxvx, ax, bx);
                        pdxvz = mpml(F.m_dxvz_ii, q, pdxvz, ax, bx);
                        pdzvx = mpml(F.m_dzvx_ii, q, pdzvx, az, bz);
                        pdzvz = mpml(F.m_dzvz_ii, q, pdzvz, az, bz);
                    }
                    const float sh = pdxvz + pdzvx;
                    sto(O.sxx_ii, q, cur.si[0] + dt * (c11 * pdxvx + c13 * pdzvz + c15 * sh));
                    sto(O.szz_ii, q, cur.si[1] + dt * (c13 * pdxvx + c33 * pdzvz + c35 * sh));
                    sto(O.sxz_ii, q, cur.si[2] + dt * (c15 * pdxvx + c35 * pdzvz + c55 * sh));
                }
            }
            if (!FS || j >= 2) {   // corner grid (with the free surface from row 2; row 1 is the mirror of row 2)
                float pdxvx = sxi / dx;
                float pdxvz = szi / dx;
                const float c11 = cur.ch[0], c13 = cur.ch[1], c15 = cur.ch[2];
                const float c33 = cur.ch[3], c35 = cur.ch[4], c55 = cur.ch[5];
                if constexpr (FS) {
                    const float* rowt = met + S.met0 + 4LL * (j - lo);
                    const float sc = rowt[3];
                    float pdzvx = sten_q<R>(qxh, -1, c) / dzk * sc;
                    float pdzvz = sten_q<R>(qzh, -1, c) / dzk * sc;
                    if (xh || j >= nz + 1) {
                        const float ax = ldo(P.axhh, q), bx = ldo(P.bxhh, q), az = ldo(P.azhh, q), bz = ldo(P.bzhh, q);
                        pdxvx = mpml(F.m_dxvx_hh, q, pdxvx, ax, bx);
                        pdxvz = mpml(F.m_dxvz_hh, q, pdxvz, ax, bx);
                        pdzvx = mpml(F.m_dzvx_hh, q, pdzvx, az, bz);
                        pdzvz = mpml(F.m_dzvz_hh, q, pdzvz, az, bz);
                    }
                    const float tpaz = fs_tpaz(S.eta_max, rowt[1], den_h, sl_h), tpcz = tpc_h;
                    const float pd1 = pdxvx + pdzvx * tpaz, pd2 = pdzvz * tpcz;
                    const float pd3 = pdzvx * tpcz + pdxvz + pdzvz * tpaz;
                    const float vxx = cur.sh[0] + dt * (c11 * pd1 + c13 * pd2 + c15 * pd3);
                    const float vzz = cur.sh[1] + dt * (c13 * pd1 + c33 * pd2 + c35 * pd3);
                    const float vxz = cur.sh[2] + dt * (c15 * pd1 + c35 * pd2 + c55 * pd3);
                    sto(O.sxx_hh, q, vxx);
                    sto(O.szz_hh, q, vzz);
                    sto(O.sxz_hh, q, vxz);
                    if (j <= R + 1) {  // odd mirror: row 3 - j
                        const unsigned qm = q - unsigned(2 * j - 3) * szb;
                        sto(O.sxx_hh, qm, -vxx);
                        sto(O.szz_hh, qm, -vzz);
                        sto(O.sxz_hh, qm, -vxz);
                    }
                } else {
                    float pdzvx = sten_q<R>(qxh, -1, c) / dz;
                    float pdzvz = sten_q<R>(qzh, -1, c) / dz;
                    if (xh || j <= 1 || j >= nz + 1) {
                        const float ax = ldo(P.axhh, q), bx = ldo(P.bxhh, q), az = ldo(P.azhh, q), bz = ldo(P.bzhh, q);
                        pdxvx = mpml(F.m_dxvx_hh, q, pdxvx, ax, bx);
                        pdxvz = mpml(F.m_dxvz_hh, q, pdxvz, ax, bx);
                        pdzvx = mpml(F.m_dzvx_hh, q, pdzvx, az, bz);
                        pdzvz = mpml(F.m_dzvz_hh, q, pdzvz, az, bz);
                    }
                    const float sh = pdxvz + pdzvx;
                    sto(O.sxx_hh, q, cur.sh[0] + dt * (c11 * pdxvx + c13 * pdzvz + c15 * sh));
                    sto(O.szz_hh, q, cur.sh[1] + dt * (c13 * pdxvx + c33 * pdzvz + c35 * sh));
                    sto(O.sxz_hh, q, cur.sh[2] + dt * (c15 * pdxvx + c35 * pdzvz + c55 * sh));
                }
            }
        }
#pragma unroll
        for (int m = 0; m < 2 * R; ++m) {
            qxh[m] = qxh[m + 1];
            qzh[m] = qzh[m + 1];
            qxi[m] = qxi[m + 1];
            qzi[m] = qzi[m + 1];
        }
        cur = nxt;
    }
}

// ---------------------------------------------------------------------------
// Velocity sweep (module_wavefield.f90:188-250): (i - 1/2, j) from dx of the
// integer-grid stresses one node back and dz of the corner stresses;
// (i, j - 1/2) from dx of the corner stresses and dz of the integer-grid
// stresses one node back.  New stresses from O, old velocities from F.
// dt / rho_h is sgn * |dt| / rho_h, exactly.  Tiles as the stress sweep:
// the x stencils of sxx and sxz come from the lanes by shuffles.
// ---------------------------------------------------------------------------
template <int R, bool FS>
__global__ __launch_bounds__(256) void k_fsg_velocity(const ShotF* __restrict__ shots,
                                                      const int* __restrict__ tile_shot, FF F, FV O, FM M, FP P,
                                                      Coef c, float sgn, float dx, float dz, int pml, int kc,
                                                      const float* __restrict__ met) {
    constexpr int W = 64 - 2 * R;
    // one wavefront per tile (blockDim.x = 64): tile, shot and row indices
    // are wavefront-uniform and live in scalar registers
    const int tile = __builtin_amdgcn_readfirstlane(int(blockIdx.x) * int(blockDim.y) + int(threadIdx.y));
    const ShotF S = shots[tile_shot[tile]];
    const int lt = tile - S.tile0;
    const int tx = lt % S.ntx, tz = lt / S.ntx;
    const int nx = S.nx, nz = S.nz, lo = 1 - pml - R;
    const int l = int(threadIdx.x);
    const int i = -pml + 1 + tx * W - R + l;
    const int j0 = (FS ? 1 : -pml + 1) + tz * kc, j1 = min(j0 + kc - 1, nz + pml);
    if (j0 > j1) return;
    const bool mcol = i <= nx + pml + R;
    const bool own = l >= R && l <= 63 - R && i <= nx + pml;
    const unsigned szb = unsigned(S.px) * 4u;
    unsigned q = unsigned(atf(S, lo, mcol ? i : lo, j0)) * 4u;  // byte offset
    // sxz_hh szz_hh sxz_ii szz_ii; with the free surface also sxx_hh, sxx_ii
    constexpr int NB = FS ? 2 * R + 1 : 1;
    float qah[2 * R + 1], qch[2 * R + 1], qai[2 * R + 1], qci[2 * R + 1], qbh[NB], qbi[NB];
#pragma unroll
    for (int m = 0; m < 2 * R; ++m) {
        const unsigned r = q + unsigned(m - R) * szb;
        qah[m] = mcol ? ldo(O.sxz_hh, r) : 0.0f;
        qch[m] = mcol ? ldo(O.szz_hh, r) : 0.0f;
        qai[m] = mcol ? ldo(O.sxz_ii, r) : 0.0f;
        qci[m] = mcol ? ldo(O.szz_ii, r) : 0.0f;
        if constexpr (FS) {
            qbh[m] = mcol ? ldo(O.sxx_hh, r) : 0.0f;
            qbi[m] = mcol ? ldo(O.sxx_ii, r) : 0.0f;
        }
    }
    const bool xi = i <= 0 || i >= nx + 1, xh = i <= 1 || i >= nx + 1;
    float den_i = 1.0f, den_h = 1.0f, sl_i = 0.0f, sl_h = 0.0f, tpc_i = 1.0f, tpc_h = 1.0f;
    const float dzk = FS ? S.fdz : dz;
    if constexpr (FS) {
        if (own) {
            const float* col = met + S.met0 + 4LL * S.nzp + 4LL * (i - lo);
            den_i = col[0] + S.depth_max;
            sl_i = col[1];
            den_h = col[2] + S.depth_max;
            sl_h = col[3];
            tpc_i = S.eta_max / den_i;
            tpc_h = S.eta_max / den_h;
        }
    }
    // The inputs of one row: the queues' fronts (row j + R), both sxx, and on
    // the owned lanes the buoyancies and old velocities.  Row j + 1's are
    // loaded while row j is computed.
    struct Row {
        float fah, fch, fai, fci, xxi, xxh;
        float qhx, qhz, vxhi, vzhi, vxih, vzih;
    };
    auto load = [&](unsigned qr, Row& w) {
        const unsigned r = qr + unsigned(R) * szb;
        w.fah = mcol ? ldo(O.sxz_hh, r) : 0.0f;
        w.fch = mcol ? ldo(O.szz_hh, r) : 0.0f;
        w.fai = mcol ? ldo(O.sxz_ii, r) : 0.0f;
        w.fci = mcol ? ldo(O.szz_ii, r) : 0.0f;
        // sxx: the row itself, or with the free surface the queue fronts
        const unsigned rx = FS ? r : qr;
        w.xxi = mcol ? ldo(O.sxx_ii, rx) : 0.0f;
        w.xxh = mcol ? ldo(O.sxx_hh, rx) : 0.0f;
        if (own) {
            w.qhx = ldo(M.qhx, qr);
            w.qhz = ldo(M.qhz, qr);
            w.vxhi = ldo(F.vx_hi, qr);
            w.vzhi = ldo(F.vz_hi, qr);
            w.vxih = ldo(F.vx_ih, qr);
            w.vzih = ldo(F.vz_ih, qr);
        }
    };
    Row cur{}, nxt{};
    load(q, cur);
    for (int j = j0; j <= j1; ++j, q += szb) {
        if (j < j1) load(q + szb, nxt);
        qah[2 * R] = cur.fah;
        qch[2 * R] = cur.fch;
        qai[2 * R] = cur.fai;
        qci[2 * R] = cur.fci;
        float xxi = cur.xxi, xxh = cur.xxh;
        if constexpr (FS) {
            qbi[2 * R] = cur.xxi;
            qbh[2 * R] = cur.xxh;
            xxi = qbi[R];
            xxh = qbh[R];
        }
        const float d_xxi = sten_lanes<R>(xxi, l, -1, c), d_xzi = sten_lanes<R>(qai[R], l, -1, c);
        const float d_xxh = sten_lanes<R>(xxh, l, 0, c), d_xzh = sten_lanes<R>(qah[R], l, 0, c);
        if (own) {
            if constexpr (FS) {
                const float* rowt = met + S.met0 + 4LL * (j - lo);
                {   // (i - 1/2, j), rows 1 .. nz+pml
                    const float sc = rowt[2];
                    float pdxxx = d_xxi / dx;
                    float pdxxz = d_xzi / dx;
                    float pdzxx = sten_q<R>(qbh, 0, c) / dzk * sc;
                    float pdzxz = sten_q<R>(qah, 0, c) / dzk * sc;
                    float pdzzz = sten_q<R>(qch, 0, c) / dzk * sc;
                    if (xh || j >= nz + 1) {
                        const float ax = ldo(P.axhi, q), bx = ldo(P.bxhi, q), az = ldo(P.azhi, q), bz = ldo(P.bzhi, q);
                        pdxxx = mpml(F.m_dxxx_hi, q, pdxxx, ax, bx);
                        pdxxz = mpml(F.m_dxxz_hi, q, pdxxz, ax, bx);
                        pdzxx = mpml(F.m_dzxx_hi, q, pdzxx, az, bz);
                        pdzxz = mpml(F.m_dzxz_hi, q, pdzxz, az, bz);
                        pdzzz = mpml(F.m_dzzz_hi, q, pdzzz, az, bz);
                    }
                    const float tpaz = fs_tpaz(S.eta_max, rowt[0], den_h, sl_h), tpcz = tpc_h;
                    const float pd1 = pdxxx + tpaz * pdzxx + tpcz * pdzxz;
                    const float pd2 = pdxxz + tpaz * pdzxz + tpcz * pdzzz;
                    const float w = sgn * cur.qhx;
                    sto(O.vx_hi, q, cur.vxhi + w * pd1);
                    sto(O.vz_hi, q, cur.vzhi + w * pd2);
                }
                if (j >= 2) {   // (i, j - 1/2), rows 2 .. nz+pml
                    const float sc = rowt[3];
                    float pdxxx = d_xxh / dx;
                    float pdxxz = d_xzh / dx;
                    float pdzxx = sten_q<R>(qbi, -1, c) / dzk * sc;
                    float pdzxz = sten_q<R>(qai, -1, c) / dzk * sc;
                    float pdzzz = sten_q<R>(qci, -1, c) / dzk * sc;
                    if (xi || j >= nz + 1) {
                        const float ax = ldo(P.axih, q), bx = ldo(P.bxih, q), az = ldo(P.azih, q), bz = ldo(P.bzih, q);
                        pdxxx = mpml(F.m_dxxx_ih, q, pdxxx, ax, bx);
                        pdxxz = mpml(F.m_dxxz_ih, q, pdxxz, ax, bx);
                        pdzxx = mpml(F.m_dzxx_ih, q, pdzxx, az, bz);
                        pdzxz = mpml(F.m_dzxz_ih, q, pdzxz, az, bz);
                        pdzzz = mpml(F.m_dzzz_ih, q, pdzzz, az, bz);
                    }
                    const float tpaz = fs_tpaz(S.eta_max, rowt[1], den_i, sl_i), tpcz = tpc_i;
                    const float pd1 = pdxxx + tpaz * pdzxx + tpcz * pdzxz;
                    const float pd2 = pdxxz + tpaz * pdzxz + tpcz * pdzzz;
                    const float w = sgn * cur.qhz;
                    sto(O.vx_ih, q, cur.vxih + w * pd1);
                    sto(O.vz_ih, q, cur.vzih + w * pd2);
                }
            } else {
                {   // (i - 1/2, j)
                    float pdxxx = d_xxi / dx;
                    float pdxxz = d_xzi / dx;
                    float pdzxz = sten_q<R>(qah, 0, c) / dz;
                    float pdzzz = sten_q<R>(qch, 0, c) / dz;
                    if (xh || j <= 0 || j >= nz + 1) {
                        const float ax = ldo(P.axhi, q), bx = ldo(P.bxhi, q), az = ldo(P.azhi, q), bz = ldo(P.bzhi, q);
                        pdxxx = mpml(F.m_dxxx_hi, q, pdxxx, ax, bx);
                        pdxxz = mpml(F.m_dxxz_hi, q, pdxxz, ax, bx);
                        pdzxz = mpml(F.m_dzxz_hi, q, pdzxz, az, bz);
                        pdzzz = mpml(F.m_dzzz_hi, q, pdzzz, az, bz);
                    }
                    const float w = sgn * cur.qhx;
                    sto(O.vx_hi, q, cur.vxhi + w * (pdxxx + pdzxz));
                    sto(O.vz_hi, q, cur.vzhi + w * (pdxxz + pdzzz));
                }
                {   // (i, j - 1/2)
                    float pdxxx = d_xxh / dx;
                    float pdxxz = d_xzh / dx;
                    float pdzxz = sten_q<R>(qai, -1, c) / dz;
                    float pdzzz = sten_q<R>(qci, -1, c) / dz;
                    if (xi || j <= 1 || j >= nz + 1) {
                        const float ax = ldo(P.axih, q), bx = ldo(P.bxih, q), az = ldo(P.azih, q), bz = ldo(P.bzih, q);
                        pdxxx = mpml(F.m_dxxx_ih, q, pdxxx, ax, bx);
                        pdxxz = mpml(F.m_dxxz_ih, q, pdxxz, ax, bx);
                        pdzxz = mpml(F.m_dzxz_ih, q, pdzxz, az, bz);
                        pdzzz = mpml(F.m_dzzz_ih, q, pdzzz, az, bz);
                    }
                    const float w = sgn * cur.qhz;
                    sto(O.vx_ih, q, cur.vxih + w * (pdxxx + pdzxz));
                    sto(O.vz_ih, q, cur.vzih + w * (pdxxz + pdzzz));
                }
            }
        }
#pragma unroll
        for (int m = 0; m < 2 * R; ++m) {
            qah[m] = qah[m + 1];
            qch[m] = qch[m + 1];
            qai[m] = qai[m + 1];
            qci[m] = qci[m + 1];
            if constexpr (FS) {
                qbh[m] = qbh[m + 1];
                qbi[m] = qbi[m + 1];
            }
        }
        cur = nxt;
    }
}

// Sources of step t (add_source): one block per shot, taps grouped by
// (field, cell) in source order; amp_tab holds sin/cos(polar) amp / rho for
// forces and -amp / (dx dz) for explosions, so every tap adds amp w1 w2.
__global__ __launch_bounds__(1024) void k_fsg_inject(const ShotF* __restrict__ shots, Targets T,
                                                     const int* __restrict__ inj_beg, const Tap* __restrict__ inj,
                                                     const float* __restrict__ amp_tab, int nt, int t) {
    const ShotF S = shots[blockIdx.x];
    for (int c = threadIdx.x; c < S.ninj; c += blockDim.x) {
        const int b = inj_beg[S.inj0 + c], e = inj_beg[S.inj0 + c + 1];
        const long long cell = inj[b].cell;
        float* f = T.f[inj[b].field];
        float acc = f[cell];
        for (int k = b; k < e; ++k) {
            const Tap tp = inj[k];
            if (t < tp.tbeg || t > tp.tend) continue;
            acc = acc + amp_tab[size_t(tp.row) * nt + (t - 1)] * tp.w * tp.w2;
        }
        f[cell] = acc;
    }
}

// Receivers of step t: for each stencil tap (irz, irx), vx_hi at (hx, gz)
// times amp1 then vx_ih at (gx, hz) times amp2, amp = w1 w2 weight 0.5
// (submodule_forward_modeling.f90:78-118); likewise vz.
__global__ __launch_bounds__(256) void k_fsg_record(const ShotF* __restrict__ shots, Targets T,
                                                    const int* __restrict__ rec_beg, const Tap* __restrict__ rec,
                                                    float* traces, int nt, int t) {
    const ShotF S = shots[blockIdx.y];
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < 2 * S.nr; r += gridDim.x * blockDim.x) {
        const int comp = r / S.nr, ir = r % S.nr;
        const int g = S.rec0 + ir;
        float acc = 0.0f;
        for (int k = rec_beg[g]; k < rec_beg[g + 1]; ++k) {
            const Tap tp = rec[k];  // field 0 (hi grid) or 2 (ih grid)
            acc = acc + T.f[tp.field + comp][tp.cell] * tp.w;
        }
        traces[S.trace0 + size_t(comp) * nt * S.nr + size_t(t - 1) * S.nr + ir] = acc;
    }
}

// Boundary records (module_boundary_saving.f90:100-200): vx_hi, vz_hi,
// vx_ih, vz_ih, each left [2-R, 1] x [1-R, nz+R], right [nx+1, nx+R] x
// [1-R, nz+R], top [1-R, nx+R] x [2-R, 1], bottom [1-R, nx+R] x [nz+1, nz+R],
// x fastest.  save != 0: record <- fields; else fields <- record.
__host__ __device__ inline long long stripf_len(int nx, int nz, int R) {
    return 4LL * (2LL * R * (nz + 2 * R) + 2LL * R * (nx + 2 * R));
}

template <int R>
__global__ void k_fsg_strips(const ShotF* __restrict__ shots, Targets T, RecStore st, int t, int pml, int mode) {
    const ShotF S = shots[blockIdx.y];
    const int nx = S.nx, nz = S.nz, lo = 1 - pml - R;
    const long long ex = nx + 2 * R, ez = nz + 2 * R;
    const long long a = R * ez, b = ex * R, one = 2 * a + 2 * b, len = 4 * one;
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
        float* f = T.f[comp];
        const size_t q = atf(S, lo, i, j);
        if (mode == kRecAmax)
            m = fmaxf(m, fabsf(f[q]));
        else if (mode == kRecSave)
            rec_put(st, rec0 + e, f[q], inv);
        else
            f[q] = rec_get(st, rec0 + e, sc);
    }
    if (mode == kRecAmax) rec_amax(st, rid, m);
}

// Adjoint sources into both adjoint velocity grids (submodule_fwi_gradient.
// f90:348-390): taps grouped by (field, cell) in receiver order; adj holds
// two [nt][nr] blocks per shot.
__global__ void k_fsg_inject_adjoint(const ShotF* __restrict__ shots, Targets T, const int* __restrict__ cell_beg,
                                     const Tap* __restrict__ taps, const float* __restrict__ adj, int nt, int t) {
    const ShotF S = shots[blockIdx.y];
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < S.nadj; c += gridDim.x * blockDim.x) {
        const int b = cell_beg[S.adj0 + c], e = cell_beg[S.adj0 + c + 1];
        const long long cell = taps[b].cell;
        const int fld = taps[b].field;  // 0..3
        float* f = T.f[fld];
        const float* a_t = adj + S.trace0 + size_t(fld & 1) * nt * S.nr + size_t(t - 1) * S.nr;
        float acc = f[cell];
        for (int k = b; k < e; ++k) acc = acc + a_t[taps[k].row] * taps[k].w;
        f[cell] = acc;
    }
}

// Imaging (compute_gradient, submodule_fwi_gradient.f90:628-727) over the
// window interior: sub-grid sums of the stress increment and of the adjoint
// stresses, strains by the 2D compliance (cij_to_sij_2d), g_cij -= products
// of adjoint and forward strains, g_rho += velocity increment . adjoint
// velocity, into seven FP64 accumulators.
struct Comp {
    const float* s[6];  // s11 s13 s15 s33 s35 s55
};
struct Acc7 {
    double* g[7];
};

__device__ __forceinline__ float hsum(const float* a, size_t q, size_t sz) {
    return a[q] + a[q + 1] + a[q + sz] + a[q + 1 + sz];
}

__global__ __launch_bounds__(256) void k_fsg_image(const ShotF* __restrict__ shots, FV cur, FV prev, FV adj, Comp C,
                                                   Acc7 G, int prev_zero, int pml, int R, int t) {
    const ShotF S = shots[blockIdx.y];
    if (t < S.hnt) return;  // this shot's reverse loop has ended (whole block)
    const int nx = S.nx, nz = S.nz, lo = 1 - pml - R;
    const long long n = (long long)nx * nz;
    const size_t sz = S.px;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        const int i = 1 + int(e % nx), j = 1 + int(e / nx);
        const size_t q = atf(S, lo, i, j);
        float txx = cur.sxx_ii[q] + 0.25f * hsum(cur.sxx_hh, q, sz);
        float tzz = cur.szz_ii[q] + 0.25f * hsum(cur.szz_hh, q, sz);
        float txz = cur.sxz_ii[q] + 0.25f * hsum(cur.sxz_hh, q, sz);
        float svx = 0.5f * (cur.vx_hi[q] + cur.vx_hi[q + 1]) + 0.5f * (cur.vx_ih[q] + cur.vx_ih[q + sz]);
        float svz = 0.5f * (cur.vz_hi[q] + cur.vz_hi[q + 1]) + 0.5f * (cur.vz_ih[q] + cur.vz_ih[q + sz]);
        if (!prev_zero) {
            txx = txx - (prev.sxx_ii[q] + 0.25f * hsum(prev.sxx_hh, q, sz));
            tzz = tzz - (prev.szz_ii[q] + 0.25f * hsum(prev.szz_hh, q, sz));
            txz = txz - (prev.sxz_ii[q] + 0.25f * hsum(prev.sxz_hh, q, sz));
            svx = svx - (0.5f * (prev.vx_hi[q] + prev.vx_hi[q + 1]) + 0.5f * (prev.vx_ih[q] + prev.vx_ih[q + sz]));
            svz = svz - (0.5f * (prev.vz_hi[q] + prev.vz_hi[q + 1]) + 0.5f * (prev.vz_ih[q] + prev.vz_ih[q + sz]));
        }
        const float rxx = adj.sxx_ii[q] + 0.25f * hsum(adj.sxx_hh, q, sz);
        const float rzz = adj.szz_ii[q] + 0.25f * hsum(adj.szz_hh, q, sz);
        const float rxz = adj.sxz_ii[q] + 0.25f * hsum(adj.sxz_hh, q, sz);
        const float s11 = C.s[0][q], s13 = C.s[1][q], s15 = C.s[2][q];
        const float s33 = C.s[3][q], s35 = C.s[4][q], s55 = C.s[5][q];
        const float exx = s11 * txx + s13 * tzz + s15 * txz, ezz = s13 * txx + s33 * tzz + s35 * txz;
        const float exz = s15 * txx + s35 * tzz + s55 * txz;
        const float fxx = s11 * rxx + s13 * rzz + s15 * rxz, fzz = s13 * rxx + s33 * rzz + s35 * rxz;
        const float fxz = s15 * rxx + s35 * rzz + s55 * rxz;
        const float rvx = 0.5f * (adj.vx_hi[q] + adj.vx_hi[q + 1]) + 0.5f * (adj.vx_ih[q] + adj.vx_ih[q + sz]);
        const float rvz = 0.5f * (adj.vz_hi[q] + adj.vz_hi[q + 1]) + 0.5f * (adj.vz_ih[q] + adj.vz_ih[q + sz]);
        const size_t o = size_t(S.g0 + e);
        G.g[0][o] -= double(fxx * exx);
        G.g[1][o] -= double(fxx * ezz + fzz * exx);
        G.g[2][o] -= double(fxx * exz + exx * fxz);
        G.g[3][o] -= double(fzz * ezz);
        G.g[4][o] -= double(fzz * exz + ezz * fxz);
        G.g[5][o] -= double(fxz * exz);
        G.g[6][o] += double(svx * rvx) + double(svz * rvz);
    }
}

// Sum of the per-shot window accumulators on the full model (nz, nx), z
// fastest (module_gradient.f90:366-388), one parameter per launch.
__global__ void k_fsg_merge(const ShotF* __restrict__ shots, const int* __restrict__ win, int nshot,
                            const double* __restrict__ g, double* full, int nx, int nz) {
    const long long n = (long long)nx * nz;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        const int kz = int(e % nz), ix = int(e / nz);
        double a = 0.0;
        for (int s = 0; s < nshot; ++s) {
            const ShotF& S = shots[s];
            const int i = ix - (win[2 * s] - 1), k = kz - (win[2 * s + 1] - 1);
            if (i < 0 || i >= S.nx || k < 0 || k >= S.nz) continue;
            a += g[size_t(S.g0) + i + size_t(k) * S.nx];
        }
        full[e] = a;
    }
}

struct ShotHostF {
    int id = 0, nx = 0, nz = 0, nxbeg = 1, nzbeg = 1, hnt = 1;
    float ox = 0, oz = 0;
    long long out0 = 0;  // offset of the shot's window gradients in owl_fsg2d_adjoint's output
    std::vector<owl_source> src;
    std::vector<owl_receiver> rec;
    std::vector<float> ratio[4];  // left, right (j = -pml .. nz+pml+1), top, bottom (i = -pml .. nx+pml+1)
    // free surface: the shot's mesh and, while planning, its medium on it
    // (c11 .. c55, rho over i = -pml .. nx+pml+1, j = -pml .. nz+pml+1)
    fsm::Mesh mesh;
    std::vector<float> fmed[7];
};

}  // namespace fsg
}  // namespace owl

using namespace owl;
using namespace owl::fsg;

struct owl_fsg2d {
    owl_acoustic_config cfg{};
    int R = 4, pml = 15, nt = 0, kc = 32, cy = 4, aniso = 1, restore_all = 0;
    float dt = 0;
    bool fs = false;  // OWL's free surface on the refined mesh
    fsm::Surface surf;
    fsg::Coef coef{};
    std::vector<ShotHostF> hs;
    std::vector<ShotF> sd;
    std::vector<int> tile_shot;
    size_t field_elems = 0, trace_elems = 0, strip_elems = 0, grad_elems = 0;
    int nrec_total = 0, max_nr = 0, max_nadj = 0;
    double pts_full = 0;
    bool have_boundary = false, fwd_state_valid = false, have_obs = false;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0, last_updates = 0;
    Profiler prof;
    std::vector<float> cij_full[6], rho_full;
    DevBuf<ShotF> d_shots;
    DevBuf<int> d_tile_shot, d_inj_beg, d_rec_beg, d_adj_beg, d_win;
    DevBuf<float> d_amp, d_rec_w, d_traces, d_strips, d_adj, d_obs;
    DevBuf<unsigned> d_rscale;  // per (shot, step) boundary record: largest magnitude (16-bit storage)
    int rec_store = OWL_STORE_F32;
    DevBuf<float> d_met;  // free-surface metric tables
    DevBuf<float> med[14];  // ci[6], ch[6], qhx, qhz
    DevBuf<float> pmlc[16];
    DevBuf<float> comp[6];  // compliances at the integer nodes
    DevBuf<float> fld[kNF];
    DevBuf<float> fB[kNW];  // the other half of the reverse ping-pong
    DevBuf<float> aF[kNF];  // adjoint field
    DevBuf<Tap> d_inj, d_rec, d_adj_taps;
    DevBuf<double> d_g[7], d_full, d_misfit, d_scratch;
    // the forward time loop as one HIP graph (modelling without profiling
    // or boundary records), re-captured when a launch argument changes
    LoopGraph graph;
    ~owl_fsg2d() {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

inline size_t cellf(const owl_fsg2d* b, const ShotF& S, int i, int j) {
    const int lo = 1 - b->pml - b->R, h = b->pml + b->R;
    if (i < lo || i > S.nx + h || j < lo || j > S.nz + h)
        fail(OWL_ERR_ARG, "a source/receiver stencil reaches outside the padded grid; increase pml");
    return size_t(S.off) + size_t(i - lo) + size_t(j - lo) * size_t(S.px);
}

inline size_t model_indexf(const owl_fsg2d* b, const ShotHostF& H, int i, int j) {
    const int ic = std::min(std::max(i, 1), H.nx), jc = std::min(std::max(j, 1), H.nz);
    return size_t(H.nzbeg - 1 + jc - 1) + size_t(H.nxbeg - 1 + ic - 1) * size_t(b->cfg.nz);
}

// damp_coef (module_cfspml.f90:282-324) with the M-PML mixing wx = dampx +
// ratioz dampz, wz = dampz + ratiox dampx, kmax = 1.
void damp_coef_mpml(float pmlvp, float xd, float zd, float dx, float dz, int pml, float adt, float am, float rx,
                    float rz, float* ax, float* bx, float* az, float* bz) {
    const float R0 = 1.0e-5f, kx = 1.0f, kz = 1.0f;
    float dampx = pmlvp * (2 + 1.0f) * std::log(1.0f / R0) / (2.0f * float(pml) * dx) * (xd * xd);
    float dampz = pmlvp * (2 + 1.0f) * std::log(1.0f / R0) / (2.0f * float(pml) * dz) * (zd * zd);
    const float wx = dampx + rz * dampz, wz = dampz + rx * dampx;
    dampx = wx;
    dampz = wz;
    const float alx = std::max(am * (1.0f - xd), 0.25f * am);
    *ax = -2.0f * adt * dampx / kx / (2.0f + adt * (alx + dampx / kx));
    *bx = (2.0f - adt * (alx + dampx / kx)) / (2.0f + adt * (alx + dampx / kx));
    const float alz = std::max(am * (1.0f - zd), 0.25f * am);
    *az = -2.0f * adt * dampz / kz / (2.0f + adt * (alz + dampz / kz));
    *bz = (2.0f - adt * (alz + dampz / kz)) / (2.0f + adt * (alz + dampz / kz));
}

// Damping ratios of every shot's four padded edge lines
// (compute_damping_profile_ratio, module_cfspml.f90:222-276), memoised on
// the medium values and spread over host threads.
void damping_ratios(owl_fsg2d* b) {
    struct Job {
        float c[6], rho;
        bool dirx;
        float* out;
    };
    std::vector<Job> jobs;
    const int pml = b->pml;
    for (auto& H : b->hs) {
        const int nzr = b->fs ? H.mesh.nz : H.nz;
        H.ratio[0].assign(nzr + 2 * pml + 2, 0.f);
        H.ratio[1].assign(nzr + 2 * pml + 2, 0.f);
        H.ratio[2].assign(H.nx + 2 * pml + 2, 0.f);
        H.ratio[3].assign(H.nx + 2 * pml + 2, 0.f);
        if (!b->aniso) continue;
        auto add = [&](int i, int j, bool dirx, float* out) {
            Job J;
            if (b->fs) {
                // the refined mesh's medium (the ratios follow prepare_modeling)
                const size_t m = size_t(i + pml) * size_t(H.mesh.nrj()) + size_t(j + pml);
                for (int k = 0; k < 6; ++k) J.c[k] = H.fmed[k][m];
                J.rho = H.fmed[6][m];
            } else {
                const size_t m = model_indexf(b, H, i, j);
                for (int k = 0; k < 6; ++k) J.c[k] = b->cij_full[k][m];
                J.rho = b->rho_full[m];
            }
            J.dirx = dirx;
            J.out = out;
            jobs.push_back(J);
        };
        const int nzs = b->fs ? H.mesh.nz : H.nz;
        for (int j = -pml; j <= nzs + pml + 1; ++j) {
            add(1, j, true, &H.ratio[0][j + pml]);
            add(H.nx, j, true, &H.ratio[1][j + pml]);
        }
        for (int i = -pml; i <= H.nx + pml + 1; ++i) {
            add(i, 1, false, &H.ratio[2][i + pml]);
            add(i, nzs, false, &H.ratio[3][i + pml]);
        }
    }
    if (jobs.empty()) return;
    // distinct (medium, direction) keys
    using Key = std::tuple<float, float, float, float, float, float, float, bool>;
    std::map<Key, size_t> uniq;
    std::vector<size_t> slot(jobs.size());
    std::vector<const Job*> todo;
    for (size_t k = 0; k < jobs.size(); ++k) {
        const Job& J = jobs[k];
        const Key key{J.c[0], J.c[1], J.c[2], J.c[3], J.c[4], J.c[5], J.rho, J.dirx};
        auto it = uniq.find(key);
        if (it == uniq.end()) {
            it = uniq.emplace(key, todo.size()).first;
            todo.push_back(&J);
        }
        slot[k] = it->second;
    }
    std::vector<float> val(todo.size());
    const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < nth; ++w)
        pool.emplace_back([&, w] {
            for (size_t k = w; k < todo.size(); k += nth) val[k] = mpml_damp_ratio(todo[k]->c, todo[k]->rho, todo[k]->dirx);
        });
    for (auto& th : pool) th.join();
    for (size_t k = 0; k < jobs.size(); ++k) *jobs[k].out = val[slot[k]];
}

void retilef(owl_fsg2d* b) {
    b->tile_shot.clear();
    for (size_t s = 0; s < b->sd.size(); ++s) {
        ShotF& S = b->sd[s];
        const int W = 64 - 2 * b->R;  // columns a tile updates (the sweeps' halo lanes feed shuffles)
        S.ntx = (S.nx + 2 * b->pml + W - 1) / W;
        // rows -pml+1 .. nz+pml, or from the free surface 1 .. nz+pml
        const int rows = b->fs ? S.nz + b->pml : S.nz + 2 * b->pml;
        S.nzc = (rows + b->kc - 1) / b->kc;
        S.tile0 = int(b->tile_shot.size());
        b->tile_shot.insert(b->tile_shot.end(), size_t(S.ntx) * S.nzc, int(s));
    }
    while (b->tile_shot.size() % size_t(b->cy)) b->tile_shot.push_back(int(b->sd.size()) - 1);
}

// Largest phase velocity of the medium over 36 directions (the 2D
// Christoffel matrix's larger eigenvalue), for the stability check.
float max_phase_velocity(const float* c, float rho) {
    double vmax = 0.0;
    for (int k = 0; k < 36; ++k) {
        const double th = M_PI * k / 36.0, k1 = std::sin(th), k3 = std::cos(th);
        const double g11 = c[0] * k1 * k1 + 2.0 * c[2] * k1 * k3 + c[5] * k3 * k3;
        const double g12 = c[2] * k1 * k1 + (double(c[1]) + c[5]) * k1 * k3 + c[4] * k3 * k3;
        const double g22 = c[5] * k1 * k1 + 2.0 * c[4] * k1 * k3 + c[3] * k3 * k3;
        const double tr = g11 + g22, det = g11 * g22 - g12 * g12;
        const double lam = 0.5 * (tr + std::sqrt(std::max(0.0, tr * tr - 4.0 * det)));
        vmax = std::max(vmax, std::sqrt(std::max(0.0, lam / rho)));
    }
    return float(vmax);
}

// Stable dt of one free-surface shot (module_vars.f90:532-567): per cell of
// the refined mesh the largest phase velocity of its medium over 36
// directions, with OWL's metric-scaled spacing.  Rows are spread over host
// threads; a cell with the previous cell's medium reuses its velocity.
float surface_dt_stable(const owl_fsg2d* b, const ShotHostF& H) {
    const fsm::Mesh& m = H.mesh;
    const int pml = m.pml, nrj = m.nrj();
    float csum = 0.0f;
    for (int k = 0; k < b->R; ++k) csum = csum + std::fabs(b->coef.c[k]);
    const int j0 = 1, j1 = m.nz + pml + 1, nrow = j1 - j0 + 1;
    std::vector<float> vrow(size_t(nrow) * size_t(m.nx + 2 * pml), 0.f);
    const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < nth; ++w)
        pool.emplace_back([&, w] {
            float last[7] = {NAN, NAN, NAN, NAN, NAN, NAN, NAN}, lv = 0.f;
            for (int j = j0 + int(w); j <= j1; j += int(nth))
                for (int i = -pml + 1; i <= m.nx + pml; ++i) {
                    const size_t q = size_t(i + pml) * size_t(nrj) + size_t(j + pml);
                    float c[7];
                    for (int k = 0; k < 7; ++k) c[k] = H.fmed[k][q];
                    if (std::memcmp(c, last, sizeof(c)) != 0) {
                        lv = max_phase_velocity(c, c[6]);
                        std::memcpy(last, c, sizeof(c));
                    }
                    vrow[size_t(j - j0) * size_t(m.nx + 2 * pml) + size_t(i + pml - 1)] = lv;
                }
        });
    for (auto& t : pool) t.join();
    return fsm::dt_stable(m, csum, [&](int i, int j) {
        return vrow[size_t(j - j0) * size_t(m.nx + 2 * pml) + size_t(i + pml - 1)];
    });
}

void planf(owl_fsg2d* b, const owl_acoustic_config* cfg, const float* const* cij, const float* rho, int aniso,
           int nshot, const owl_shot* shots) {
    if (!cfg || !cij || !rho || nshot < 1 || !shots) fail(OWL_ERR_ARG, "null argument or no shots");
    for (int k = 0; k < 6; ++k)
        if (!cij[k]) fail(OWL_ERR_ARG, "null modulus array");
    if (!fd_coefficients(cfg->fdhalf, b->coef.c)) fail(OWL_ERR_ARG, "fdhalf must be 2..8");
    if (cfg->nx < 1 || cfg->nz < 1 || !(cfg->dx > 0) || !(cfg->dz > 0) || !(cfg->dt > 0))
        fail(OWL_ERR_ARG, "invalid grid or dt");
    if (cfg->pml < cfg->fdhalf || cfg->pml + cfg->fdhalf < kNkw) fail(OWL_ERR_ARG, "pml must be >= fdhalf");
    b->fs = cfg->free_surface != 0 || b->fs;
    if (b->fs && !(b->surf.dz_max > 0)) b->surf.dz_max = 1.5f * cfg->dz;  // module_model.f90:178
    if (cfg->data_dt != cfg->dt || cfg->data_tmax != cfg->tmax)
        fail(OWL_ERR_ARG, "data_dt/data_tmax must equal dt/tmax (resample the traces with owl_resample_traces)");
    b->cfg = *cfg;
    b->R = cfg->fdhalf;
    b->pml = cfg->pml;
    b->dt = cfg->dt;
    b->aniso = aniso ? 1 : 0;
    b->nt = time_steps(cfg->tmax, cfg->dt);
    if (b->nt < 2) fail(OWL_ERR_ARG, "tmax/dt gives fewer than 2 steps");
    const size_t nfull = size_t(cfg->nx) * cfg->nz;
    for (int k = 0; k < 6; ++k) b->cij_full[k].assign(cij[k], cij[k] + nfull);
    b->rho_full.assign(rho, rho + nfull);
    for (size_t e = 0; e < nfull; ++e)
        if (!(b->rho_full[e] > 0.f)) fail(OWL_ERR_ARG, "rho must be positive");
    const float f0factor = cfg->f0factor > 0 ? cfg->f0factor : 1.0f;
    const int R = b->R, pml = b->pml, nt = b->nt, lo = 1 - pml - R;
    b->hs.assign(nshot, ShotHostF{});
    b->sd.assign(nshot, ShotF{});
    b->nrec_total = b->max_nr = b->max_nadj = 0;
    b->pts_full = 0;
    long long off = 0, trace = 0, strip = 0, grad = 0, out = 0;
    std::vector<float> amp, rec_w;
    std::vector<int> inj_beg{0}, adj_beg{0};
    std::vector<Tap> inj, adj_taps;
    std::vector<std::vector<Tap>> rec_taps;  // per (shot, receiver)
    std::vector<float> vcell(nfull, -1.f);   // per-cell largest phase velocity, on demand
    for (int s = 0; s < nshot; ++s) {
        const owl_shot& in = shots[s];
        ShotHostF& H = b->hs[s];
        ShotF& S = b->sd[s];
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
        float dts;
        if (b->fs) {
            // the window's medium, its refined mesh and the medium on it
            std::vector<float> win[7];
            for (int k = 0; k < 7; ++k) {
                win[k].resize(size_t(H.nx) * H.nz);
                for (int i = 1; i <= H.nx; ++i)
                    for (int j = 1; j <= H.nz; ++j) {
                        const size_t m = model_indexf(b, H, i, j);
                        win[k][size_t(j - 1) + size_t(i - 1) * H.nz] = k < 6 ? b->cij_full[k][m] : b->rho_full[m];
                    }
            }
            float f0min = H.src[0].f0;
            for (auto& sc : H.src) f0min = std::min(f0min, sc.f0);
            H.mesh = fsm::build(H.nx, H.nz, cfg->dx, cfg->dz, H.ox, H.oz, pml, b->surf, win[5].data(), win[6].data(),
                                f0min);
            for (int k = 0; k < 7; ++k) H.fmed[k] = fsm::medium(H.mesh, win[k].data());
            dts = surface_dt_stable(b, H);
        } else {
            float vmax = 0.f;
            for (int i = 1; i <= H.nx; ++i)
                for (int j = 1; j <= H.nz; ++j) {
                    const size_t m = model_indexf(b, H, i, j);
                    if (vcell[m] < 0.f) {
                        float c[6];
                        for (int k = 0; k < 6; ++k) c[k] = b->cij_full[k][m];
                        vcell[m] = max_phase_velocity(c, b->rho_full[m]);
                    }
                    vmax = std::max(vmax, vcell[m]);
                }
            dts = dt_stable(R, 2, vmax, cfg->dx, 1.0f, cfg->dz);
        }
        if (cfg->dt > dts)
            fail(OWL_ERR_CFL, "dt = " + std::to_string(cfg->dt) + " exceeds the stable dt = " + std::to_string(dts));
        S.nx = H.nx;
        S.nz = b->fs ? H.mesh.nz : H.nz;
        S.fdz = b->fs ? H.mesh.dz : cfg->dz;
        S.eta_max = H.mesh.eta_max;
        S.depth_max = H.mesh.depth_max;
        S.px = (H.nx + 2 * (pml + R) + kPitchAlign - 1) / kPitchAlign * kPitchAlign;
        S.nzp = S.nz + 2 * (pml + R);
        S.off = off;
        off += (long long)S.px * S.nzp;
        off = (off + 63) / 64 * 64;
        b->pts_full += double(H.nx + 2 * pml) * double(b->fs ? S.nz + pml : S.nz + 2 * pml);
        S.nr = int(H.rec.size());
        S.trace0 = trace;
        trace += 2LL * nt * S.nr;
        S.strip0 = strip;
        strip += stripf_len(S.nx, S.nz, R) * nt;
        S.g0 = grad;
        grad += (long long)S.nx * S.nz;
        H.out0 = out;
        out += (long long)H.nx * H.nz;
        // sources (add_source): a force adds sin/cos(polar) amp / rho(gx, gz)
        // to vx/vz on (hx, gz) with hx iz and on (gx, hz) with ix hz; an
        // explosion subtracts amp / (dx dz) from sxx, szz on (gx, gz) with
        // ix iz and on (hx, hz) with hx hz; amp = stf amp dt
        const float xmax = float(H.nx - 1) * cfg->dx + H.ox, zmax = float(H.nz - 1) * cfg->dz + H.oz;
        std::vector<Tap> st;
        for (const owl_source& sc : H.src) {
            if (sc.mechanism != OWL_FORCE && sc.mechanism != OWL_EXPLOSION)
                fail(OWL_ERR_ARG, "the FSG solver takes force and explosion sources");
            Located L = locate(2, sc.x, 0.0f, sc.z, H.ox, 0.0f, H.oz, cfg->dx, 1.0f, cfg->dz);
            float zchk = sc.z, ascale = 1.0f, dz_s = cfg->dz;
            if (b->fs) {
                // depth on the refined mesh (module_vars.f90:639-680); a
                // source above the first row moves to it with amp * 1.2
                const fsm::ZLoc Z = fsm::locate_z(H.mesh, b->surf, sc.x, sc.z, true);
                L.gz = Z.gz;
                L.hz = Z.hz;
                std::copy(Z.iz, Z.iz + 9, L.iz);
                std::copy(Z.hzw, Z.hzw + 9, L.hzw);
                zchk = Z.zv;
                ascale = Z.amp_scale;
                dz_s = H.mesh.edz_i[size_t(Z.gz + pml)] * (H.mesh.topo_max + H.mesh.depth_max) / H.mesh.eta_max;
            }
            if (!(sc.x >= H.ox && sc.x <= xmax && zchk >= H.oz && zchk <= zmax))
                fail(OWL_ERR_ARG, "source outside its shot window");
            const float samp = sc.amp * ascale;
            int hnt = 0;
            std::vector<float> stf = source_time_function(sc.wavelet, sc.f0, cfg->dt, f0factor, &hnt);
            if (sc.mechanism == OWL_EXPLOSION) stf = stf_time_derivative(stf, cfg->dt, nt);
            if (&sc == &H.src[0]) {
                H.hnt = hnt;
                S.hnt = std::max(hnt, 1);
            }
            const int nbeg = fortran_nint(sc.t0 / cfg->dt) + 1, nend = nbeg + int(stf.size()) - 1;
            // with the free surface only rows >= 2 take the source (add_source)
            auto taps = [&](int gx, int gz, const float* wx, const float* wz, int field, int row) {
                for (int irz = -kNkw; irz <= kNkw; ++irz)
                    for (int irx = -kNkw; irx <= kNkw; ++irx) {
                        const float w1 = wx[irx + kNkw], w2 = wz[irz + kNkw];
                        if (w1 == 0.f || w2 == 0.f) continue;
                        if (b->fs && gz + irz < 2) continue;
                        st.push_back(Tap{(long long)cellf(b, S, gx + irx, gz + irz), w1, w2, row, nbeg, nend, field});
                    }
            };
            auto table = [&](float fac, bool divide_rho, float r) {
                const int row = int(amp.size() / nt);
                for (int t = 1; t <= nt; ++t) {
                    float a = (t >= nbeg && t <= nend) ? stf[t - nbeg] * samp * cfg->dt : 0.0f;
                    if (divide_rho) {
                        a = a / r;
                        amp.push_back(fac * a);
                    } else {
                        amp.push_back(-(a / (cfg->dx * dz_s)));
                    }
                }
                return row;
            };
            if (sc.mechanism == OWL_FORCE) {
                const float r = b->fs ? H.fmed[6][size_t(L.gx + pml) * H.mesh.nrj() + size_t(L.gz + pml)]
                                      : b->rho_full[model_indexf(b, H, L.gx, L.gz)];
                const int rs = table(std::sin(sc.polar), true, r), rc = table(std::cos(sc.polar), true, r);
                taps(L.hx, L.gz, L.hxw, L.iz, 0, rs);
                taps(L.hx, L.gz, L.hxw, L.iz, 1, rc);
                taps(L.gx, L.hz, L.ix, L.hzw, 2, rs);
                taps(L.gx, L.hz, L.ix, L.hzw, 3, rc);
            } else {
                const int re = table(0.f, false, 1.f);
                taps(L.gx, L.gz, L.ix, L.iz, 4, re);
                taps(L.gx, L.gz, L.ix, L.iz, 5, re);
                taps(L.hx, L.hz, L.hxw, L.hzw, 6, re);
                taps(L.hx, L.hz, L.hxw, L.hzw, 7, re);
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
        // receivers: per stencil tap the (hx, gz) cell with amp1, then the
        // (gx, hz) cell with amp2 (outside the window: weight 0, no taps)
        S.rec0 = b->nrec_total;
        std::vector<Tap> at;
        for (int r = 0; r < S.nr; ++r) {
            const owl_receiver& rc = H.rec[r];
            std::vector<Tap> taps, ataps;
            Located L = locate(2, rc.x, 0.0f, rc.z, H.ox, 0.0f, H.oz, cfg->dx, 1.0f, cfg->dz);
            float zchk = rc.z;
            if (b->fs) {
                const fsm::ZLoc Z = fsm::locate_z(H.mesh, b->surf, rc.x, rc.z, false);
                L.gz = Z.gz;
                L.hz = Z.hz;
                std::copy(Z.iz, Z.iz + 9, L.iz);
                std::copy(Z.hzw, Z.hzw + 9, L.hzw);
                zchk = Z.zv;
            }
            const bool valid = rc.x >= H.ox && rc.x <= xmax && zchk >= H.oz && zchk <= zmax;
            rec_w.push_back((valid && rc.weight != 0.f) ? rc.weight : 0.f);
            if (valid && rc.weight != 0.f) {
                // With the free surface the velocities above it are read as
                // their even mirror (submodule_forward_modeling.f90:66-75:
                // vx_hi row 1-j = row 1+j, vx_ih row 2-j = row 1+j, j = 1..R),
                // and the adjoint sources go to rows >= 2 only
                // (submodule_fwi_gradient.f90:350-380).
                auto rrow = [&](int row, bool half) {
                    if (!b->fs) return row;
                    if (!half && row >= 1 - R && row <= 0) return 2 - row;
                    if (half && row >= 2 - R && row <= 1) return 3 - row;
                    return row;
                };
                for (int irz = -kNkw; irz <= kNkw; ++irz)
                    for (int irx = -kNkw; irx <= kNkw; ++irx) {
                        const float a1 = L.hxw[irx + kNkw] * L.iz[irz + kNkw] * rc.weight * 0.5f;
                        const float a2 = L.ix[irx + kNkw] * L.hzw[irz + kNkw] * rc.weight * 0.5f;
                        const int r1 = L.gz + irz, r2 = L.hz + irz;
                        if (a1 != 0.f) {
                            taps.push_back(Tap{(long long)cellf(b, S, L.hx + irx, rrow(r1, false)), a1, 0.f, r, 0, 0, 0});
                            if (!b->fs || r1 >= 2)
                                ataps.push_back(Tap{(long long)cellf(b, S, L.hx + irx, r1), a1, 0.f, r, 0, 0, 0});
                        }
                        if (a2 != 0.f) {
                            taps.push_back(Tap{(long long)cellf(b, S, L.gx + irx, rrow(r2, true)), a2, 0.f, r, 0, 0, 2});
                            if (!b->fs || r2 >= 2)
                                ataps.push_back(Tap{(long long)cellf(b, S, L.gx + irx, r2), a2, 0.f, r, 0, 0, 2});
                        }
                    }
            }
            for (const Tap& tp : ataps)
                for (int c = 0; c < 2; ++c) {
                    Tap u = tp;
                    u.field = tp.field + c;
                    at.push_back(u);
                }
            rec_taps.push_back(std::move(taps));
        }
        // adjoint sources: grouped by (field, cell), receiver order within a cell
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
        b->nrec_total += S.nr;
        b->max_nr = std::max(b->max_nr, S.nr);
    }
    b->field_elems = size_t(off);
    if (b->field_elems >= (size_t(1) << 29))
        fail(OWL_ERR_ARG, "the shots' padded grids exceed 2 GiB per field; split the batch");
    b->trace_elems = size_t(trace);
    b->strip_elems = size_t(strip);
    b->grad_elems = size_t(grad);
    retilef(b);
    damping_ratios(b);
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
    // padded media (edge replication), corner means, |dt| / rho_h,
    // compliances, per-cell M-PML coefficients
    const size_t n = b->field_elems;
    std::vector<float> ci[6], ch[6], rh(n, 1.f), qhx(n, 0.f), qhz(n, 0.f), cs[6], pc[16];
    for (int k = 0; k < 6; ++k) {
        ci[k].assign(n, 0.f);
        ch[k].assign(n, 0.f);
        cs[k].assign(n, 0.f);
    }
    for (auto& v : pc) v.assign(n, 0.f);
    const float adt = std::fabs(cfg->dt);
    for (size_t s = 0; s < b->hs.size(); ++s) {
        const ShotHostF& H = b->hs[s];
        const ShotF& S = b->sd[s];
        auto at = [&](int i, int j) { return size_t(S.off) + size_t(i - lo) + size_t(j - lo) * S.px; };
        if (b->fs) {
            // the medium on the refined mesh, edge-replicated beyond it
            const int nrj = H.mesh.nrj();
            for (int j = lo; j < lo + S.nzp; ++j)
                for (int i = lo; i < lo + S.px; ++i) {
                    const int ic = std::min(std::max(i, -pml), H.nx + pml + 1);
                    const int jc = std::min(std::max(j, -pml), S.nz + pml + 1);
                    const size_t m = size_t(ic + pml) * size_t(nrj) + size_t(jc + pml), q = at(i, j);
                    for (int k = 0; k < 6; ++k) ci[k][q] = H.fmed[k][m];
                    rh[q] = H.fmed[6][m];
                }
        } else {
            for (int j = lo; j < lo + S.nzp; ++j)
                for (int i = lo; i < lo + S.px; ++i) {
                    const size_t m = model_indexf(b, H, i, j), q = at(i, j);
                    for (int k = 0; k < 6; ++k) ci[k][q] = b->cij_full[k][m];
                    rh[q] = b->rho_full[m];
                }
        }
        for (int j = lo + 1; j < lo + S.nzp; ++j)
            for (int i = lo + 1; i < lo + S.nx + 2 * (pml + R); ++i) {
                const size_t q = at(i, j), sz = S.px;
                for (int k = 0; k < 6; ++k)
                    ch[k][q] = 0.25f * (ci[k][q - 1 - sz] + ci[k][q - sz] + ci[k][q - 1] + ci[k][q]);
                qhx[q] = adt / (0.5f * (rh[q - 1] + rh[q]));
                qhz[q] = adt / (0.5f * (rh[q - sz] + rh[q]));
            }
        for (int j = 1; j <= S.nz; ++j)
            for (int i = 1; i <= S.nx; ++i) {
                const size_t q = at(i, j);
                const float c11 = ci[0][q], c13 = ci[1][q], c15 = ci[2][q], c33 = ci[3][q], c35 = ci[4][q],
                            c55 = ci[5][q];
                const float t1 = c15 * c15 * c33 - 2 * c13 * c15 * c35 + c11 * (c35 * c35) + c13 * c13 * c55 -
                                 c11 * c33 * c55;
                const float t2 = c15 * c15 * c33 - 2 * c13 * c15 * c35 + c13 * c13 * c55 + c11 * (c35 * c35 - c33 * c55);
                cs[0][q] = (c35 * c35 - c33 * c55) / t1;
                cs[1][q] = (-(c15 * c35) + c13 * c55) / t2;
                cs[2][q] = (c15 * c33 - c13 * c35) / t1;
                cs[3][q] = (c15 * c15 - c11 * c55) / t1;
                cs[4][q] = (-(c13 * c15) + c11 * c35) / t2;
                cs[5][q] = (c13 * c13 - c11 * c33) / t1;
            }
        // pmlvp = sqrt(max c11/rho, c33/rho over the window's edge lines)
        // (module_vars.f90:345-349); alphamax = max f0 pi
        float pm = 0.f, f0max = 0.f;
        for (int i = 1; i <= H.nx; ++i)
            for (int j = 1; j <= H.nz; ++j) {
                if (!(i == 1 || i == H.nx || j == 1 || j == H.nz)) continue;
                const size_t m = model_indexf(b, H, i, j);
                pm = std::max(pm, b->cij_full[0][m] / b->rho_full[m]);
                pm = std::max(pm, b->cij_full[3][m] / b->rho_full[m]);
            }
        for (auto& sc : H.src) f0max = std::max(f0max, sc.f0);
        const float pmlvp = std::sqrt(pm), am = float(double(1.0f * f0max) * M_PI);
        const int nx = H.nx, nz = S.nz;
        for (int j = -pml + 1; j <= nz + pml; ++j)
            for (int i = -pml + 1; i <= nx + pml; ++i) {
                if (i > 1 && i < nx && j > 1 && j < nz) continue;
                float xi = 0, xh = 0, zi = 0, zh = 0;
                if (i <= 1) {
                    xi = float(std::fabs((i - 1 + 0.0) / pml));
                    xh = float(std::fabs((i - 1 - 0.5) / pml));
                } else if (i >= nx) {
                    xi = float(std::fabs((i - nx + 0.0) / pml));
                    xh = float(std::fabs((i - nx - 0.5) / pml));
                }
                if (j <= 1) {
                    zi = float(std::fabs((j - 1 + 0.0) / pml));
                    zh = float(std::fabs((j - 1 - 0.5) / pml));
                } else if (j >= nz) {
                    zi = float(std::fabs((j - nz + 0.0) / pml));
                    zh = float(std::fabs((j - nz - 0.5) / pml));
                }
                float rx = 0.f, rz = 0.f;
                if (i <= 1)
                    rx = H.ratio[0][j + pml];
                else if (i >= nx)
                    rx = H.ratio[1][j + pml];
                if (j <= 1)
                    rz = H.ratio[2][i + pml];
                else if (j >= nz)
                    rz = H.ratio[3][i + pml];
                const size_t q = at(i, j);
                // with the free surface the z damping uses the local row
                // spacing eta_dz_i(j) on all four grids (module_cfspml.f90:395-480)
                const float dxf = cfg->dx, dzf = b->fs ? H.mesh.edz_i[size_t(j + pml)] : cfg->dz;
                // grids: ii (xi, zi), hh (xh, zh), hi (xh, zi), ih (xi, zh)
                damp_coef_mpml(pmlvp, xi, zi, dxf, dzf, pml, adt, am, rx, rz, &pc[0][q], &pc[1][q], &pc[2][q], &pc[3][q]);
                damp_coef_mpml(pmlvp, xh, zh, dxf, dzf, pml, adt, am, rx, rz, &pc[4][q], &pc[5][q], &pc[6][q], &pc[7][q]);
                damp_coef_mpml(pmlvp, xh, zi, dxf, dzf, pml, adt, am, rx, rz, &pc[8][q], &pc[9][q], &pc[10][q], &pc[11][q]);
                damp_coef_mpml(pmlvp, xi, zh, dxf, dzf, pml, adt, am, rx, rz, &pc[12][q], &pc[13][q], &pc[14][q],
                               &pc[15][q]);
            }
    }
    if (b->fs) {
        // per shot: (eta_i, eta_h, sc_i, sc_h) per storage row, then
        // (topo_i, slope_i, topo_h, slope_h) per storage column
        std::vector<float> met;
        for (size_t s = 0; s < b->hs.size(); ++s) {
            ShotF& S = b->sd[s];
            const fsm::Mesh& m = b->hs[s].mesh;
            S.met0 = (long long)met.size();
            for (int j = lo; j < lo + S.nzp; ++j) {
                const size_t jc = size_t(std::min(std::max(j, -pml), m.nz + pml + 1) + pml);
                met.insert(met.end(), {m.eta_i[jc], m.eta_h[jc], m.sc_i[jc], m.sc_h[jc]});
            }
            for (int i = lo; i < lo + S.px; ++i) {
                const size_t ic = size_t(std::min(std::max(i, -pml), m.nx + pml + 1) + pml);
                met.insert(met.end(), {m.topo_i[ic], m.slope_i[ic], m.topo_h[ic], m.slope_h[ic]});
            }
        }
        b->d_met.upload(met, stm);
        b->d_shots.upload(b->sd, stm);
        for (auto& H : b->hs)
            for (auto& v : H.fmed) std::vector<float>().swap(v);
    } else {
        b->d_met.upload(std::vector<float>(4, 0.f), stm);
    }
    for (int k = 0; k < 6; ++k) {
        b->med[k].upload(ci[k], stm);
        b->med[6 + k].upload(ch[k], stm);
        b->comp[k].upload(cs[k], stm);
    }
    b->med[12].upload(qhx, stm);
    b->med[13].upload(qhz, stm);
    for (int k = 0; k < 16; ++k) b->pmlc[k].upload(pc[k], stm);
    for (auto& f : b->fld) f.alloc(b->field_elems);
    b->d_traces.alloc(std::max<size_t>(b->trace_elems, 1));
    OWL_HIP(hipStreamSynchronize(stm));
}

FF fieldsf(DevBuf<float>* f) {
    FF F;
    float** p = &F.sxx_ii;
    for (int k = 0; k < kNF; ++k) p[k] = f[k].get();
    return F;
}
FV out_of(const FF& F) {
    FV V;
    float* const* s = &F.sxx_ii;
    float** d = &V.sxx_ii;
    for (int k = 0; k < kNW; ++k) d[k] = s[k];
    return V;
}
FV fv_of(DevBuf<float>* f) {
    FV V;
    float** d = &V.sxx_ii;
    for (int k = 0; k < kNW; ++k) d[k] = f[k].get();
    return V;
}
FF with_fields(FF F, const FV& V) {
    float** d = &F.sxx_ii;
    float* const* s = &V.sxx_ii;
    for (int k = 0; k < kNW; ++k) d[k] = s[k];
    return F;
}
Targets vel_targets(const FV& V) { return Targets{{V.vx_hi, V.vz_hi, V.vx_ih, V.vz_ih, V.sxx_ii, V.szz_ii, V.sxx_hh, V.szz_hh}}; }

FM mediaf(owl_fsg2d* b) {
    FM M;
    for (int k = 0; k < 6; ++k) {
        M.ci[k] = b->med[k].get();
        M.ch[k] = b->med[6 + k].get();
    }
    M.qhx = b->med[12].get();
    M.qhz = b->med[13].get();
    return M;
}
FP pmlf(owl_fsg2d* b) {
    FP P;
    const float** d = &P.axii;
    for (int k = 0; k < 16; ++k) d[k] = b->pmlc[k].get();
    return P;
}

// Algorithmic bytes per sweep: the stress sweep reads four velocities, six
// old stresses and twelve moduli and writes six stresses (112 B per point);
// the velocity sweep reads six stresses, four old velocities and two
// |dt| / rho and writes four velocities (64 B); in the damping frame each
// grid adds four memory variables (r+w) and four coefficients.
void sweep_bytesf(const owl_fsg2d* b, double& bs, double& bv) {
    bs = bv = 0;
    for (auto& S : b->sd) {
        // with the free surface the sweeps start at the surface row
        const double lx = S.nx + 2.0 * b->pml, lz = S.nz + (b->fs ? 1.0 : 2.0) * b->pml;
        const double frame = lx * lz - double(std::max(S.nx - 2, 0)) * double(std::max(S.nz - (b->fs ? 1 : 2), 0));
        bs += 112.0 * lx * lz + 2.0 * 48.0 * frame;
        bv += 64.0 * lx * lz + 2.0 * 48.0 * frame;
    }
}

void maybe_collectf(owl_fsg2d* b) {
    if (b->prof.pending() > 20000) {
        OWL_HIP(hipStreamSynchronize(b->stream));
        b->prof.collect();
    }
}

template <class Fn>
void dispatch_Rf(int R, Fn&& f) {
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
// fields): stresses, then velocities (update_wavefield).
template <int R>
void stepf(owl_fsg2d* b, const FF& F, const FV& O, float dt, double bs, double bv) {
    if (64 * b->cy > 256) fail(OWL_ERR_ARG, "launch shape exceeds the sweeps' 256-thread blocks");
    hipStream_t st = b->stream;
    const dim3 blk(64, unsigned(b->cy)), grd(unsigned(b->tile_shot.size() / size_t(b->cy)));
    const FM M = mediaf(b);
    const FP P = pmlf(b);
    const float* met = b->d_met.get();
    int ev = b->prof.begin(st, 0, bs);
    if (b->fs)
        hipLaunchKernelGGL((k_fsg_stress<R, true>), grd, blk, 0, st, b->d_shots.get(), b->d_tile_shot.get(), F, O, M,
                           P, b->coef, dt, b->cfg.dx, b->cfg.dz, b->pml, b->kc, met);
    else
        hipLaunchKernelGGL((k_fsg_stress<R, false>), grd, blk, 0, st, b->d_shots.get(), b->d_tile_shot.get(), F, O,
                           M, P, b->coef, dt, b->cfg.dx, b->cfg.dz, b->pml, b->kc, met);
    b->prof.end(st, ev);
    ev = b->prof.begin(st, 1, bv);
    if (b->fs)
        hipLaunchKernelGGL((k_fsg_velocity<R, true>), grd, blk, 0, st, b->d_shots.get(), b->d_tile_shot.get(), F, O,
                           M, P, b->coef, dt < 0 ? -1.0f : 1.0f, b->cfg.dx, b->cfg.dz, b->pml, b->kc, met);
    else
        hipLaunchKernelGGL((k_fsg_velocity<R, false>), grd, blk, 0, st, b->d_shots.get(), b->d_tile_shot.get(), F, O,
                           M, P, b->coef, dt < 0 ? -1.0f : 1.0f, b->cfg.dx, b->cfg.dz, b->pml, b->kc, met);
    b->prof.end(st, ev);
}

void stripsf(owl_fsg2d* b, const FV& V, int t, int save) {
    long long mx = 1;
    for (auto& S : b->sd) mx = std::max(mx, stripf_len(S.nx, S.nz, b->R));
    const dim3 grd(unsigned(std::min<long long>(1024, (mx + 255) / 256)), unsigned(b->sd.size()));
    const RecStore st{b->d_strips.get(), b->d_rscale.get(), b->rec_store, b->nt};
    const int ev = b->prof.begin(b->stream, 4, 0);
    dispatch_Rf(b->R, [&](auto r) {
        if (save && b->rec_store != OWL_STORE_F32)
            hipLaunchKernelGGL(k_fsg_strips<decltype(r)::value>, grd, dim3(256), 0, b->stream, b->d_shots.get(),
                               vel_targets(V), st, t, b->pml, int(kRecAmax));
        hipLaunchKernelGGL(k_fsg_strips<decltype(r)::value>, grd, dim3(256), 0, b->stream, b->d_shots.get(),
                           vel_targets(V), st, t, b->pml, save ? int(kRecSave) : int(kRecRestore));
    });
    b->prof.end(b->stream, ev);
}

// Every device address and launch parameter the forward loop passes to its
// kernels: a captured loop is replayed only while these are unchanged.
std::vector<uintptr_t> forward_signaturef(const owl_fsg2d* b) {
    std::vector<uintptr_t> v{uintptr_t(b->R), uintptr_t(b->nt), uintptr_t(b->kc), uintptr_t(b->cy),
                             uintptr_t(b->tile_shot.size()), uintptr_t(b->max_nr), uintptr_t(b->fs),
                             uintptr_t(float_bits(b->dt)),
                             uintptr_t(b->d_shots.get()), uintptr_t(b->d_tile_shot.get()),
                             uintptr_t(b->d_rec_beg.get()), uintptr_t(b->d_rec.get()), uintptr_t(b->d_traces.get()),
                             uintptr_t(b->d_inj_beg.get()), uintptr_t(b->d_inj.get()), uintptr_t(b->d_amp.get()),
                             uintptr_t(b->d_met.get())};
    for (auto& f : b->fld) v.push_back(uintptr_t(f.get()));
    for (auto& f : b->med) v.push_back(uintptr_t(f.get()));
    for (auto& f : b->pmlc) v.push_back(uintptr_t(f.get()));
    return v;
}

template <int R>
void run_forwardf(owl_fsg2d* b, bool save_boundary) {
    hipStream_t st = b->stream;
    for (auto& f : b->fld) f.zero(st);
    b->d_traces.zero(st);
    if (save_boundary) alloc_records(b->d_strips, b->d_rscale, b->strip_elems, b->rec_store, b->sd.size(), b->nt, st);
    const FF F = fieldsf(b->fld);
    const FV V = out_of(F);
    const unsigned nshot = unsigned(b->sd.size());
    double bs, bv;
    sweep_bytesf(b, bs, bv);
    const unsigned rg = unsigned(std::max(1, (2 * b->max_nr + 255) / 256));
    // step t: [boundary record] -> record(t) -> update -> source(t)
    // (submodule_forward_modeling.f90:55-240)
    auto loop = [&] {
        for (int t = 1; t <= b->nt; ++t) {
            if (save_boundary) stripsf(b, V, t, 1);
            int ev = b->prof.begin(st, 2, 0);
            hipLaunchKernelGGL(k_fsg_record, dim3(rg, nshot), dim3(256), 0, st, b->d_shots.get(), vel_targets(V),
                               b->d_rec_beg.get(), b->d_rec.get(), b->d_traces.get(), b->nt, t);
            b->prof.end(st, ev);
            if (t == b->nt) break;  // the last update changes nothing that is recorded
            stepf<R>(b, F, V, b->dt, bs, bv);
            ev = b->prof.begin(st, 2, 0);
            hipLaunchKernelGGL(k_fsg_inject, dim3(nshot), dim3(1024), 0, st, b->d_shots.get(), vel_targets(V),
                               b->d_inj_beg.get(), b->d_inj.get(), b->d_amp.get(), b->nt, t);
            b->prof.end(st, ev);
            if (!b->graph.capturing) maybe_collectf(b);
        }
    };
    const bool graph_ok = b->graph.enabled && !b->prof.on() && !save_boundary;
    OWL_HIP(hipEventRecord(b->ev0, st));
    if (graph_ok)
        b->graph.run(st, forward_signaturef(b), loop);
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

// OWL's reverse pass (submodule_fwi_gradient.f90:150-405): t = nt .. hnt:
// prev = state; state = update(state, -dt); at t = nt the final state;
// boundary record t; adjoint = update(adjoint, -dt); adjoint sources;
// imaging.  The forward field steps out of place so the previous state
// stays in the other buffer; at t = nt the update of the zero field is the
// zero field, so the final state is used as is with prev = 0.  OWL reads
// the final state back with vz_hxiz named twice
// (module_boundary_saving.f90:85-90): the saved vz_ih lands in vz_hi and
// vz_ih keeps the zero of the reconstruction; restore_all = 0 keeps that
// order, 1 restores every field.  OWL's loop continues the adjoint below
// hnt without imaging; those steps do not touch the gradients.
template <int R>
void run_reversef(owl_fsg2d* b) {
    hipStream_t st = b->stream;
    if (!b->have_boundary || !b->fwd_state_valid)
        fail(OWL_ERR_STATE, "the adjoint needs a preceding forward run with boundary saving");
    if (!b->d_adj.get() || b->d_adj.size() < b->trace_elems) fail(OWL_ERR_STATE, "no adjoint source");
    for (int k = kNW; k < kNF; ++k) b->fld[k].zero(st);  // the reverse run's memory variables start at zero
    for (auto& f : b->fB) {
        f.alloc(b->field_elems);
        f.zero(st);
    }
    for (auto& f : b->aF) {
        f.alloc(b->field_elems);
        f.zero(st);
    }
    const size_t ng = std::max<size_t>(b->grad_elems, 1);
    Acc7 G;
    for (int k = 0; k < 7; ++k) {
        b->d_g[k].alloc(ng);
        b->d_g[k].zero(st);
        G.g[k] = b->d_g[k].get();
    }
    Comp C;
    for (int k = 0; k < 6; ++k) C.s[k] = b->comp[k].get();
    const unsigned nshot = unsigned(b->sd.size());
    // each shot is imaged only while t >= its own hnt; the batch runs down to
    // the smallest one
    int hnt = b->nt;
    for (auto& H : b->hs) hnt = std::min(hnt, std::max(H.hnt, 1));
    const int cc = b->cfg.cc_step_interval > 0 ? b->cfg.cc_step_interval : 1;
    double bs, bv;
    sweep_bytesf(b, bs, bv);
    FF X = fieldsf(b->fld);
    FV Y = fv_of(b->fB);
    const FF A = fieldsf(b->aF);
    const FV adj = out_of(A);
    if (!b->restore_all) {
        // the final state as OWL reads it back
        OWL_HIP(hipMemcpyAsync(X.vz_hi, X.vz_ih, b->field_elems * sizeof(float), hipMemcpyDeviceToDevice, st));
        b->fld[9].zero(st);
    }
    // imaging: 30 wavefields, 6 compliances, seven FP64 accumulators r+w
    const double image_bytes = (30.0 * 4 + 24.0 + 112.0) * double(b->grad_elems);
    const unsigned ig = unsigned(std::max<size_t>(1, std::min<size_t>(2048, (b->grad_elems / nshot + 255) / 256)));
    OWL_HIP(hipEventRecord(b->ev0, st));
    for (int t = b->nt; t >= hnt; --t) {
        FV cur = out_of(X);
        const FV prev = out_of(X);
        if (t != b->nt) {
            stepf<R>(b, X, Y, -b->dt, bs, bv);
            cur = Y;
        }
        stripsf(b, cur, t, 0);
        stepf<R>(b, A, adj, -b->dt, bs, bv);
        int ev = b->prof.begin(st, 2, 0);
        if (b->max_nadj > 0)
            hipLaunchKernelGGL(k_fsg_inject_adjoint, dim3(unsigned((b->max_nadj + 255) / 256), nshot), dim3(256), 0,
                               st, b->d_shots.get(), vel_targets(adj), b->d_adj_beg.get(), b->d_adj_taps.get(),
                               b->d_adj.get(), b->nt, t);
        b->prof.end(st, ev);
        if (t % cc == 0) {
            ev = b->prof.begin(st, 3, image_bytes);
            hipLaunchKernelGGL(k_fsg_image, dim3(ig, nshot), dim3(256), 0, st, b->d_shots.get(), cur, prev, adj, C, G,
                               t == b->nt ? 1 : 0, b->pml, R, t);
            b->prof.end(st, ev);
        }
        if (t != b->nt) {
            const FV old = out_of(X);
            X = with_fields(X, Y);
            Y = old;
        }
        maybe_collectf(b);
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

void traces_to_hostf(const owl_fsg2d* b, const float* dev, float* out) {
    std::vector<float> h(b->trace_elems);
    OWL_HIP(hipMemcpy(h.data(), dev, h.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (auto& S : b->sd)
        for (int c = 0; c < 2; ++c)
            for (int r = 0; r < S.nr; ++r)
                for (int t = 0; t < b->nt; ++t)
                    out[size_t(S.trace0) + (size_t(c) * S.nr + r) * b->nt + t] =
                        h[size_t(S.trace0) + size_t(c) * b->nt * S.nr + size_t(t) * S.nr + r];
}

void traces_to_devicef(owl_fsg2d* b, const float* in, DevBuf<float>& dev) {
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

// The caller's surface description; zero dz_refine and dz_max take OWL's
// defaults (4 and 1.5 dz).
fsm::Surface surface_of(const owl_fsg_surface* sf, float dz) {
    fsm::Surface S;
    S.dz_max = 1.5f * dz;
    if (!sf) return S;
    if (sf->ntopo < 0 || (sf->ntopo > 0 && (!sf->topo_x || !sf->topo_z))) fail(OWL_ERR_ARG, "bad topography");
    S.tx.assign(sf->topo_x, sf->topo_x + sf->ntopo);
    S.tz.assign(sf->topo_z, sf->topo_z + sf->ntopo);
    if (sf->topo_interp != 0 && sf->topo_interp != 1) fail(OWL_ERR_ARG, "topo_interp must be 0 (linear) or 1 (cubic)");
    S.interp = sf->topo_interp;
    if (sf->dz_refine < 0 || sf->dz_max < 0) fail(OWL_ERR_ARG, "dz_refine and dz_max must not be negative");
    if (sf->dz_refine > 0) S.refine = sf->dz_refine;
    if (sf->dz_max > 0) S.dz_max = sf->dz_max;
    S.src_from_surface = sf->src_depth_from_surface != 0;
    S.rec_from_surface = sf->rec_depth_from_surface != 0;
    return S;
}

// One shot's refined-mesh accumulators (nx, nz refined, x fastest) on its
// regular window (nz0, nx) z fastest, NaN/Inf -> 0 first.
void shot_to_regular(const owl_fsg2d* b, size_t s, const double* g, float* out) {
    const ShotF& S = b->sd[s];
    std::vector<float> gf(size_t(S.nx) * S.nz);
    for (size_t e = 0; e < gf.size(); ++e) {
        const float v = float(g[size_t(S.g0) + e]);
        gf[e] = std::isfinite(v) ? v : 0.f;
    }
    fsm::to_regular(b->hs[s].mesh, gf.data(), out);
}

}  // namespace

extern "C" {

owl_status owl_fsg2d_create_surface(const owl_acoustic_config* cfg, const float* const* cij, const float* rho,
                                    int aniso, int nshot, const owl_shot* shots, const owl_fsg_surface* surface,
                                    owl_fsg2d** out) {
    return guarded([&] {
        if (!out || !cfg) fail(OWL_ERR_ARG, "null argument");
        *out = nullptr;
        auto* b = new owl_fsg2d();
        try {
            b->fs = surface != nullptr || cfg->free_surface != 0;
            b->surf = surface_of(surface, cfg->dz);
            planf(b, cfg, cij, rho, aniso, nshot, shots);
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

owl_status owl_fsg2d_surface_mesh(const owl_acoustic_config* cfg, const float* c55, const float* rho, float f0min,
                                  const owl_fsg_surface* surface, int cap, int* nz, float* scal, float* eta_i,
                                  float* eta_h, float* sc_i, float* sc_h, float* edz_i, float* topo_i,
                                  float* slope_i, float* topo_h, float* slope_h) {
    return guarded([&] {
        if (!cfg || !c55 || !rho || !nz || !scal) fail(OWL_ERR_ARG, "null argument");
        const fsm::Surface S = surface_of(surface, cfg->dz);
        const fsm::Mesh m =
            fsm::build(cfg->nx, cfg->nz, cfg->dx, cfg->dz, cfg->ox, cfg->oz, cfg->pml, S, c55, rho, f0min);
        *nz = m.nz;
        if (cap < m.nrj()) fail(OWL_ERR_ARG, "row tables need nz + 2 pml + 2 entries");
        scal[0] = m.dz;
        scal[1] = m.eta_max;
        scal[2] = m.depth_max;
        scal[3] = m.topo_max;
        const std::vector<float>* rows[5] = {&m.eta_i, &m.eta_h, &m.sc_i, &m.sc_h, &m.edz_i};
        float* rdst[5] = {eta_i, eta_h, sc_i, sc_h, edz_i};
        for (int k = 0; k < 5; ++k)
            if (rdst[k]) std::copy(rows[k]->begin(), rows[k]->end(), rdst[k]);
        const std::vector<float>* cols[4] = {&m.topo_i, &m.slope_i, &m.topo_h, &m.slope_h};
        float* cdst[4] = {topo_i, slope_i, topo_h, slope_h};
        for (int k = 0; k < 4; ++k)
            if (cdst[k]) std::copy(cols[k]->begin(), cols[k]->end(), cdst[k]);
    });
}

owl_status owl_fsg2d_create(const owl_acoustic_config* cfg, const float* const* cij, const float* rho, int aniso,
                            int nshot, const owl_shot* shots, owl_fsg2d** out) {
    return guarded([&] {
        if (!out) fail(OWL_ERR_ARG, "out is null");
        *out = nullptr;
        auto* b = new owl_fsg2d();
        try {
            planf(b, cfg, cij, rho, aniso, nshot, shots);
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

owl_status owl_fsg2d_set_model(owl_fsg2d* b, const float* const* cij, const float* rho) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        // the shots as the batch was created with (planning replaces hs)
        const std::vector<ShotHostF> old = b->hs;
        std::vector<owl_shot> shots;
        for (const ShotHostF& H : old)
            shots.push_back(owl_shot{H.id, int(H.src.size()), H.src.data(), int(H.rec.size()), H.rec.data(),
                                     H.nxbeg, H.nxbeg + H.nx - 1, 0, 0, H.nzbeg, H.nzbeg + H.nz - 1});
        const owl_acoustic_config cfg = b->cfg;
        b->have_boundary = false;
        b->fwd_state_valid = false;
        planf(b, &cfg, cij, rho, b->aniso, int(shots.size()), shots.data());
        OWL_HIP(hipStreamSynchronize(b->stream));
    });
}

void owl_fsg2d_destroy(owl_fsg2d* b) { delete b; }

int owl_fsg2d_nt(const owl_fsg2d* b) { return b ? b->nt : 0; }

int owl_fsg2d_rows(const owl_fsg2d* b, int shot) {
    if (!b || shot < 0 || shot >= int(b->sd.size())) return 0;
    return b->sd[size_t(shot)].nz;
}

owl_status owl_fsg2d_damping_ratios(const owl_fsg2d* b, int shot, float* left, float* right, float* top,
                                    float* bottom) {
    return guarded([&] {
        if (!b || shot < 0 || shot >= int(b->hs.size())) fail(OWL_ERR_ARG, "bad shot index");
        const ShotHostF& H = b->hs[size_t(shot)];
        float* dst[4] = {left, right, top, bottom};
        for (int k = 0; k < 4; ++k)
            if (dst[k]) std::copy(H.ratio[k].begin(), H.ratio[k].end(), dst[k]);
    });
}

owl_status owl_fsg2d_damp_ratio(const float* c, float rho, int dirx, float* ratio) {
    return guarded([&] {
        if (!c || !ratio) fail(OWL_ERR_ARG, "null argument");
        *ratio = mpml_damp_ratio(c, rho, dirx != 0);
    });
}

owl_status owl_fsg2d_set_restore_all(owl_fsg2d* b, int on) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        b->restore_all = on ? 1 : 0;
    });
}

owl_status owl_fsg2d_forward(owl_fsg2d* b, float* traces, int save_boundary) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        dispatch_Rf(b->R, [&](auto r) { run_forwardf<decltype(r)::value>(b, save_boundary != 0); });
        if (traces) traces_to_hostf(b, b->d_traces.get(), traces);
    });
}

owl_status owl_fsg2d_adjoint(owl_fsg2d* b, const float* adj, float* const* grads) {
    return guarded([&] {
        if (!b || !adj) fail(OWL_ERR_ARG, "null argument");
        traces_to_devicef(b, adj, b->d_adj);
        dispatch_Rf(b->R, [&](auto r) { run_reversef<decltype(r)::value>(b); });
        if (!grads) return;
        std::vector<double> g(b->grad_elems);
        for (int k = 0; k < 7; ++k) {
            if (!grads[k]) continue;
            OWL_HIP(hipMemcpy(g.data(), b->d_g[k].get(), g.size() * sizeof(double), hipMemcpyDeviceToHost));
            // per-shot window gradients (nz_s, nx_s), z fastest, NaN/Inf -> 0;
            // with the free surface mapped from the refined mesh
            for (size_t s = 0; s < b->sd.size(); ++s) {
                const ShotF& S = b->sd[s];
                float* dst = grads[k] + b->hs[s].out0;
                if (b->fs) {
                    shot_to_regular(b, s, g.data(), dst);
                    continue;
                }
                for (int kz = 0; kz < S.nz; ++kz)
                    for (int i = 0; i < S.nx; ++i) {
                        const float v = float(g[size_t(S.g0) + i + size_t(kz) * S.nx]);
                        dst[kz + size_t(i) * S.nz] = std::isfinite(v) ? v : 0.f;
                    }
            }
        }
    });
}

owl_status owl_fsg2d_gradient(owl_fsg2d* b, const float* obs, double* misfit, double* const* grads,
                              float* traces_out) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        if (obs) {
            traces_to_devicef(b, obs, b->d_obs);
            b->have_obs = true;
        }
        if (!b->have_obs) fail(OWL_ERR_STATE, "no observed data");
        dispatch_Rf(b->R, [&](auto r) { run_forwardf<decltype(r)::value>(b, true); });
        const double fwd_ms = b->last_ms, fwd_up = b->last_updates;
        hipStream_t st = b->stream;
        const unsigned nshot = unsigned(b->sd.size());
        b->d_adj.alloc(std::max<size_t>(b->trace_elems, 1));
        b->d_misfit.alloc(2 * size_t(nshot));
        b->d_scratch.alloc(std::max<size_t>(4 * size_t(b->nrec_total), 1));
        hipLaunchKernelGGL(k_adjoint_source_2c<ShotF>, dim3(nshot, 2), dim3(256), 0, st, b->d_shots.get(),
                           b->d_traces.get(), b->d_obs.get(), b->d_rec_w.get(), b->d_adj.get(),
                           b->d_misfit.get(), b->nt, b->cfg.trace_discard_threshold, b->d_scratch.get());
        dispatch_Rf(b->R, [&](auto r) { run_reversef<decltype(r)::value>(b); });
        const size_t nfull = size_t(b->cfg.nx) * b->cfg.nz;
        b->d_full.alloc(nfull);
        const unsigned mg = unsigned(std::min<size_t>(8192, (nfull + 255) / 256));
        std::vector<double> h(nfull);
        if (b->fs) {
            // map each shot's refined-mesh gradient to its window on the host,
            // then sum the windows on the model (module_gradient.f90:366-388)
            std::vector<double> g(b->grad_elems);
            for (int k = 0; k < 7; ++k) {
                if (!grads || !grads[k]) continue;
                OWL_HIP(hipMemcpy(g.data(), b->d_g[k].get(), g.size() * sizeof(double), hipMemcpyDeviceToHost));
                std::fill(grads[k], grads[k] + nfull, 0.0);
                for (size_t s = 0; s < b->sd.size(); ++s) {
                    const ShotHostF& H = b->hs[s];
                    std::vector<float> w(size_t(H.nx) * H.nz);
                    shot_to_regular(b, s, g.data(), w.data());
                    for (int i = 0; i < H.nx; ++i)
                        for (int kz = 0; kz < H.nz; ++kz)
                            grads[k][size_t(H.nzbeg - 1 + kz) + size_t(H.nxbeg - 1 + i) * b->cfg.nz] +=
                                double(w[size_t(kz) + size_t(i) * H.nz]);
                }
            }
        }
        for (int k = 0; k < 7 && !b->fs; ++k) {
            if (!grads || !grads[k]) continue;
            hipLaunchKernelGGL(k_fsg_merge, dim3(mg), dim3(256), 0, st, b->d_shots.get(), b->d_win.get(), int(nshot),
                               b->d_g[k].get(), b->d_full.get(), b->cfg.nx, b->cfg.nz);
            OWL_HIP(hipGetLastError());
            OWL_HIP(hipMemcpyAsync(h.data(), b->d_full.get(), nfull * sizeof(double), hipMemcpyDeviceToHost, st));
            OWL_HIP(hipStreamSynchronize(st));
            for (size_t e = 0; e < nfull; ++e) grads[k][e] = std::isfinite(h[e]) ? h[e] : 0.0;
        }
        OWL_HIP(hipStreamSynchronize(st));
        if (misfit) {
            std::vector<double> m2(2 * size_t(nshot));
            OWL_HIP(hipMemcpy(m2.data(), b->d_misfit.get(), m2.size() * sizeof(double), hipMemcpyDeviceToHost));
            for (unsigned s = 0; s < nshot; ++s) misfit[s] = m2[2 * s] + m2[2 * s + 1];
        }
        if (traces_out) traces_to_hostf(b, b->d_traces.get(), traces_out);
        b->last_ms += fwd_ms;
        b->last_updates += fwd_up;
    });
}

owl_status owl_fsg2d_set_graphs(owl_fsg2d* b, int on) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        b->graph.enabled = on != 0;
        b->graph.reset();
    });
}

// Launch-shape autotuner: times `steps` steps of the two sweeps on the
// batch under each candidate (rows per tile column, tiles per block) and
// keeps the fastest; the fields are zeroed afterwards (every forward starts
// from zero).  Stencil timing does not depend on the field values.
owl_status owl_fsg2d_autotune(owl_fsg2d* b, int steps, int* planes, int* rows) {
    return guarded([&] {
        if (!b || steps < 1) fail(OWL_ERR_ARG, "null batch or no steps");
        static const int kPlanes[] = {4, 8, 12, 16, 32, 64, 128};
        static const int kRows[] = {1, 2, 4};
        hipStream_t st = b->stream;
        const bool prof_on = b->prof.on();
        b->prof.enable(false);
        for (auto& f : b->fld) f.zero(st);
        const FF F = fieldsf(b->fld);
        const FV V = out_of(F);
        double bs, bv;
        sweep_bytesf(b, bs, bv);
        float best = -1.f;
        int bk = b->kc, by = b->cy;
        auto shape = [&](int kc, int cy) {
            b->kc = kc;
            b->cy = cy;
            retilef(b);
            b->d_shots.upload(b->sd, st);
            b->d_tile_shot.upload(b->tile_shot, st);
        };
        for (int kc : kPlanes)
            for (int cy : kRows) {
                shape(kc, cy);
                dispatch_Rf(b->R, [&](auto r) {
                    stepf<decltype(r)::value>(b, F, V, b->dt, bs, bv);  // warm-up
                    OWL_HIP(hipEventRecord(b->ev0, st));
                    for (int t = 0; t < steps; ++t) stepf<decltype(r)::value>(b, F, V, b->dt, bs, bv);
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

owl_status owl_fsg2d_set_launch(owl_fsg2d* b, int planes, int rows) {
    return guarded([&] {
        if (!b || planes < 1 || (rows != 1 && rows != 2 && rows != 4)) fail(OWL_ERR_ARG, "bad launch shape");
        b->kc = planes;
        b->cy = rows;
        retilef(b);
        b->d_shots.upload(b->sd, b->stream);
        b->d_tile_shot.upload(b->tile_shot, b->stream);
        OWL_HIP(hipStreamSynchronize(b->stream));
    });
}

owl_status owl_fsg2d_set_record_storage(owl_fsg2d* b, int kind) {
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

owl_status owl_fsg2d_set_profiling(owl_fsg2d* b, int on) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        b->prof.enable(on != 0);
    });
}

owl_status owl_fsg2d_kernel_stats(const owl_fsg2d* b, int kind, double* total_ms, long long* launches,
                                  double* bytes_total) {
    return guarded([&] {
        if (!b || kind < 0 || kind >= Profiler::kKinds) fail(OWL_ERR_ARG, "bad kernel kind");
        if (total_ms) *total_ms = b->prof.ms(kind);
        if (launches) *launches = b->prof.launches(kind);
        if (bytes_total) *bytes_total = b->prof.bytes(kind);
    });
}

owl_status owl_fsg2d_last_stats(const owl_fsg2d* b, double* ms, double* updates) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        if (ms) *ms = b->last_ms;
        if (updates) *updates = b->last_updates;
    });
}

}  // extern "C"
