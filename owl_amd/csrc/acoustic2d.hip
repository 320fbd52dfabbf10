// 2D acoustic SSG-FD time stepping and FWI gradient for MI355X (gfx950).
//
// The scheme is OWL's acoustic_iso_2d solver: pressure p on integer nodes,
// vx / vz on half nodes, optimised staggered coefficients of half-length R,
// CFS-PML memory variables in boundary slabs, 9x9 Kaiser-windowed sinc
// source injection and receiver sampling, and OWL's boundary-saving
// gradient (acoustic_iso_2d/module_wavefield.f90:40-263,
// submodule_forward_modeling.f90:31-132, submodule_fwi_gradient.f90:31-269).
//
// All shots of a batch live in one set of device buffers; every kernel
// launch covers every shot, each shot on its own window.  Arithmetic is
// FP32 in the reference's operation order (built with -ffp-contract=off),
// gradients and misfits accumulate in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "host_prep.hpp"
#include "owl_amd.h"

namespace owl {
namespace a2 {

constexpr int kBX = 64;     // threads along x (one wavefront per row)
constexpr int kBY = 4;      // rows of threads per block
constexpr int kRows = 16;   // z rows per block (kRows / kBY per thread)
constexpr int kPitchAlign = 64;  // row pitch multiple, floats (256 B)

struct Coef {
    float c[8];
};

// Device-side description of one shot.
struct Shot {
    int nx, nz;           // interior window
    int pitch;            // floats per padded row
    int nrow;             // padded rows
    long long off;        // element offset of node (lo, lo) in every field buffer
    long long mdel;       // model element = field element + mdel (shots with one window share the model)
    int tile0, ntx;       // first update tile, tiles per row
    int prof;             // offset of the shot's PML profiles
    int nr;               // receivers
    int rec0;             // first receiver of the shot in the receiver tables
    int inj0, ninj;       // source injection cells
    int adj0, nadj;       // adjoint-source injection cells
    long long strip0;     // offset of the shot's boundary records
    long long trace0;     // offset of the shot's [nt][nr] trace block
    long long g0;         // offset of the shot's window gradient (nx*nz)
    int edge0;            // offset of the shot's saved edge values (2*nz + 2*nx)
    int nxbeg, nzbeg;     // window origin in the full model (1-based)
    int ftile0, fntx;     // fused step: first column-strip tile, strips per row
    int atile0;           // fused adjoint: first column-strip tile (its own z-chunk rows)
    int hnt;              // last reverse step OWL images for this shot (its first source's hnt)
};

// Element of the model arrays (bk, rho) at field element q of shot S.
__device__ __forceinline__ size_t mq(const Shot& S, size_t q) { return size_t((long long)q + S.mdel); }

// One set of wavefields for a sweep.  Inputs and outputs may alias (in
// place) or not (ping-pong); the CFS-PML memory variables are m[0] dx vx,
// m[1] dz vz (integer nodes, p-sweep) and m[2] dx p, m[3] dz p (half nodes,
// v-sweep).
struct FieldSet {
    const float* p_in;
    float* p_out;
    const float* vx_in;
    float* vx_out;
    const float* vz_in;
    float* vz_out;
    const float* m_in[4];
    float* m_out[4];
};
struct Sets {
    FieldSet s[2];
};

__device__ __forceinline__ int shot_of_tile(const int* tile_shot, int tile) { return tile_shot[tile]; }

// ---------------------------------------------------------------------------
// p-sweep: module_wavefield.f90:51-71 (and :123-143 with a free surface).
// p = p - dt*bk*(dx+ vx + dz+ vz), PML memory update where i <= 0 or
// i >= nx+1 (resp. j).  Out of place when p_out != p_in.
// ---------------------------------------------------------------------------
template <int R>
__global__ __launch_bounds__(kBX * kBY) void k_update_p(const Shot* __restrict__ shots,
                                                        const int* __restrict__ tile_shot, Sets sets,
                                                        const float* __restrict__ bk,
                                                        const float* __restrict__ prof, Coef c,
                                                        float dt, float idx, float idz, int pml,
                                                        int free_surface) {
    const int tile = blockIdx.x;
    const Shot S = shots[shot_of_tile(tile_shot, tile)];
    const FieldSet F = sets.s[blockIdx.y];
    const int lt = tile - S.tile0;
    const int tx = lt % S.ntx, tz = lt / S.ntx;
    const int i = -pml + 1 + tx * kBX + int(threadIdx.x);
    const int nx = S.nx, nz = S.nz;
    if (i > nx + pml) return;
    const int lo = 1 - pml - R;
    const int lx = nx + 2 * pml, lz = nz + 2 * pml;
    const float* axi = prof + S.prof;
    const float* bxi = axi + lx;
    const float* azi = prof + S.prof + 4 * lx;
    const float* bzi = azi + lz;
    const bool xslab = (i <= 0 || i >= nx + 1);
    float ax = 0.f, bx = 0.f;
    if (xslab) {
        ax = axi[i + pml - 1];
        bx = bxi[i + pml - 1];
    }
    const size_t pitch = size_t(S.pitch);
#pragma unroll
    for (int k = 0; k < kRows / kBY; ++k) {
        // kBX = 64: the row is wavefront-uniform (scalar registers)
        const int j = -pml + 1 + tz * kRows + __builtin_amdgcn_readfirstlane(int(threadIdx.y)) + kBY * k;
        if (j > nz + pml) break;
        if (free_surface && j < 2) continue;
        const size_t q = size_t(S.off) + size_t(i - lo) + size_t(j - lo) * pitch;
        float sx = c.c[0] * (F.vx_in[q + 1] - F.vx_in[q]);
        float sz = c.c[0] * (F.vz_in[q + pitch] - F.vz_in[q]);
#pragma unroll
        for (int m = 2; m <= R; ++m) {
            sx = sx + c.c[m - 1] * (F.vx_in[q + m] - F.vx_in[q - (m - 1)]);
            sz = sz + c.c[m - 1] * (F.vz_in[q + m * pitch] - F.vz_in[q - (m - 1) * pitch]);
        }
        float pdx = idx * sx;
        float pdz = idz * sz;
        if (xslab) {
            const float mem = ax * pdx + bx * F.m_in[0][q];
            F.m_out[0][q] = mem;
            pdx = pdx + mem;  // division by kappa = 1 is exact and omitted
        }
        const bool zslab = free_surface ? (j >= nz + 1) : (j <= 0 || j >= nz + 1);
        if (zslab) {
            const float mem = azi[j + pml - 1] * pdz + bzi[j + pml - 1] * F.m_in[1][q];
            F.m_out[1][q] = mem;
            pdz = pdz + mem;
        }
        F.p_out[q] = F.p_in[q] - dt * bk[mq(S, q)] * (pdx + pdz);
    }
}

// Free-surface mirror p(i, j) = -p(i, 2 - j), j <= 0 (module_wavefield.f90:145-151).
template <int R>
__global__ void k_mirror_p(const Shot* __restrict__ shots, Sets sets, int pml) {
    const Shot S = shots[blockIdx.y];
    float* p = sets.s[blockIdx.z].p_out;
    const int lo = 1 - pml - R;
    const int lx = S.nx + 2 * pml;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < lx * pml; e += gridDim.x * blockDim.x) {
        const int i = -pml + 1 + e % lx;
        const int j = -pml + 1 + e / lx;  // j in [-pml+1, 0]
        const size_t pitch = size_t(S.pitch);
        p[size_t(S.off) + size_t(i - lo) + size_t(j - lo) * pitch] =
            -p[size_t(S.off) + size_t(i - lo) + size_t(2 - j - lo) * pitch];
    }
}

// ---------------------------------------------------------------------------
// v-sweeps: module_wavefield.f90:73-105 (and :153-185 with a free surface).
// vx(m) = vx(m) - dt/(0.5*(rho(m)+rho(m-1))) * dx- p, PML where m <= 1 or
// m >= nx+1; vz likewise along z.  One thread updates vx and vz at (i, j).
// ---------------------------------------------------------------------------
template <int R>
__global__ __launch_bounds__(kBX * kBY) void k_update_v(const Shot* __restrict__ shots,
                                                        const int* __restrict__ tile_shot, Sets sets,
                                                        const float* __restrict__ rho,
                                                        const float* __restrict__ prof, Coef c,
                                                        float dt, float idx, float idz, int pml,
                                                        int free_surface) {
    const int tile = blockIdx.x;
    const Shot S = shots[shot_of_tile(tile_shot, tile)];
    const FieldSet F = sets.s[blockIdx.y];
    const int lt = tile - S.tile0;
    const int tx = lt % S.ntx, tz = lt / S.ntx;
    const int i = -pml + 1 + tx * kBX + int(threadIdx.x);
    const int nx = S.nx, nz = S.nz;
    if (i > nx + pml) return;
    const int lo = 1 - pml - R;
    const int lx = nx + 2 * pml, lz = nz + 2 * pml;
    const float* axh = prof + S.prof + 2 * lx;
    const float* bxh = axh + lx;
    const float* azh = prof + S.prof + 4 * lx + 2 * lz;
    const float* bzh = azh + lz;
    const float* p = F.p_out;  // the p just updated
    const bool xslab = (i <= 1 || i >= nx + 1);
    float ax = 0.f, bx = 0.f;
    if (xslab) {
        ax = axh[i + pml - 1];
        bx = bxh[i + pml - 1];
    }
    const size_t pitch = size_t(S.pitch);
#pragma unroll
    for (int k = 0; k < kRows / kBY; ++k) {
        // kBX = 64: the row is wavefront-uniform (scalar registers)
        const int j = -pml + 1 + tz * kRows + __builtin_amdgcn_readfirstlane(int(threadIdx.y)) + kBY * k;
        if (j > nz + pml) break;
        const size_t q = size_t(S.off) + size_t(i - lo) + size_t(j - lo) * pitch;
        if (!free_surface || j >= 1) {
            // vx at half node m = i: difference p(m-1+k) - p(m-k)
            float s = c.c[0] * (p[q] - p[q - 1]);
#pragma unroll
            for (int m = 2; m <= R; ++m) s = s + c.c[m - 1] * (p[q + (m - 1)] - p[q - m]);
            float d = idx * s;
            if (xslab) {
                const float mem = ax * d + bx * F.m_in[2][q];
                F.m_out[2][q] = mem;
                d = d + mem;
            }
            const float reff = 0.5f * (rho[mq(S, q)] + rho[mq(S, q) - 1]);
            F.vx_out[q] = F.vx_in[q] - dt / reff * d;
        }
        if (!free_surface || j >= 2) {
            float s = c.c[0] * (p[q] - p[q - pitch]);
#pragma unroll
            for (int m = 2; m <= R; ++m) s = s + c.c[m - 1] * (p[q + (m - 1) * pitch] - p[q - m * pitch]);
            float d = idz * s;
            const bool zslab = free_surface ? (j >= nz + 1) : (j <= 1 || j >= nz + 1);
            if (zslab) {
                const float mem = azh[j + pml - 1] * d + bzh[j + pml - 1] * F.m_in[3][q];
                F.m_out[3][q] = mem;
                d = d + mem;
            }
            const float reff = 0.5f * (rho[mq(S, q)] + rho[mq(S, q) - pitch]);
            F.vz_out[q] = F.vz_in[q] - dt / reff * d;
        }
    }
}

// ---------------------------------------------------------------------------
// Fused forward step (p-sweep, then v-sweep with the new p) for the
// exact-adjoint path, out of place, no free surface.  One wavefront per
// tile: 64 lanes own 64 consecutive columns c = x0 - R + lane and march down
// z over the tile's rows [z0, z1], the front row r running R - 1 rows ahead
// of the v rows.  The new p of the last 2R rows stays in a register queue
// (the z-stencil of vz), its x-neighbours come from the other lanes
// (shuffles), and vz_in's z-stencil is a register queue too.  Lanes
// 0..R-1 and 65-R..63 compute p redundantly for their neighbours' vx; only
// lanes R..64-R write, and only the tile's own rows.  Every input must be
// distinct from every output (the redundant lanes read what a neighbouring
// tile may already have written); arithmetic as k_update_p / k_update_v,
// so the results are bit-identical to the two-sweep form.
// ---------------------------------------------------------------------------
// Load and store at a 32-bit byte offset from a kernel-argument base: the
// batch's buffers stay below 2 GiB (checked on the host), so the fused
// kernels address every array with one offset register against scalar
// bases instead of a 64-bit address pair per array.
__device__ __forceinline__ float ldo(const float* base, unsigned off) {
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + off);
}
__device__ __forceinline__ void sto(float* base, unsigned off, float v) {
    *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + off) = v;
}
__device__ __forceinline__ void addo(double* base, unsigned off8, double v) {
    double* p = reinterpret_cast<double*>(reinterpret_cast<char*>(base) + off8);
    *p += v;
}

template <int R>
__global__ __launch_bounds__(256) void k_step_fused(const Shot* __restrict__ shots,
                                                    const int* __restrict__ ftile_shot, int nftile,
                                                    FieldSet F, const float* __restrict__ bk,
                                                    const float* __restrict__ rho,
                                                    const float* __restrict__ prof, Coef c, float dt,
                                                    float idx, float idz, int pml, int zc) {
    constexpr int W = 64 - (2 * R - 1);
    // one wavefront per tile (blockDim.x = 64): the tile, its shot and every
    // row index are wavefront-uniform, so they live in scalar registers and
    // the per-row PML coefficients come through scalar loads
    const int tile = __builtin_amdgcn_readfirstlane(int(blockIdx.x) * int(blockDim.y) + int(threadIdx.y));
    if (tile >= nftile) return;
    const Shot S = shots[ftile_shot[tile]];
    const int lt = tile - S.ftile0;
    const int tx = lt % S.fntx, tz = lt / S.fntx;
    const int l = int(threadIdx.x);
    const int nx = S.nx, nz = S.nz;
    const int lo = 1 - pml - R;
    const int lx = nx + 2 * pml, lz = nz + 2 * pml;
    const int col = -pml + 1 + tx * W - R + l;
    const int z0 = -pml + 1 + tz * zc, z1 = min(z0 + zc - 1, nz + pml);
    const bool pcol = col >= -pml + 1 && col <= nx + pml;     // p is updated in this column
    const bool mcol = col <= nx + pml + R;                     // column inside the padded grid
    const bool own = l >= R && l <= 64 - R && col <= nx + pml;
    const float* axi = prof + S.prof;
    const float* bxi = axi + lx;
    const float* axh = prof + S.prof + 2 * lx;
    const float* bxh = axh + lx;
    const float* azi = prof + S.prof + 4 * lx;
    const float* bzi = azi + lz;
    const float* azh = azi + 2 * lz;
    const float* bzh = azh + lz;
    const bool xs_i = col <= 0 || col >= nx + 1, xs_h = col <= 1 || col >= nx + 1;
    float ax = 0.f, bx = 0.f, axh_ = 0.f, bxh_ = 0.f;
    if (pcol && xs_i) {
        ax = axi[col + pml - 1];
        bx = bxi[col + pml - 1];
    }
    if (own && xs_h) {
        axh_ = axh[col + pml - 1];
        bxh_ = bxh[col + pml - 1];
    }
    const unsigned pitchb = unsigned(S.pitch) * 4u;
    const unsigned base = unsigned(S.off + (col - lo)) * 4u;  // byte offset of (col, lo)
    const unsigned mdb = unsigned(S.mdel) * 4u;               // model minus field, bytes
    // Inputs of one loop iteration: the p row r and the v row j = r - R + 1.
    struct Row {
        float vx[2 * R];  // vx_in at row r, columns col - R + 1 .. col + R
        float vzf;        // vz_in at row r + R (the front of the vz queue)
        float p, bk, m0, m1;
        float rq, rl, ru, vxj, m2, m3;  // rho (j), rho (j, col - 1), rho (j - 1), vx_in (j)
    };
    const int rs = z0 - R, re = z1 + R - 1;
    auto load = [&](int r, Row& w) {
        const int vr = r + R;
        w.vzf = (mcol && vr >= lo && vr <= nz + pml + R) ? ldo(F.vz_in, base + unsigned(vr - lo) * pitchb) : 0.0f;
        if (pcol && r >= -pml + 1 && r <= nz + pml) {
            const unsigned q = base + unsigned(r - lo) * pitchb;
#pragma unroll
            for (int k = 0; k < 2 * R; ++k) w.vx[k] = ldo(F.vx_in, q - unsigned(R - 1) * 4u + unsigned(k) * 4u);
            w.p = ldo(F.p_in, q);
            w.bk = ldo(bk, q + mdb);
            w.m0 = xs_i ? ldo(F.m_in[0], q) : 0.0f;
            w.m1 = (r <= 0 || r >= nz + 1) ? ldo(F.m_in[1], q) : 0.0f;
        }
        const int j = r - R + 1;
        if (own && j >= z0) {
            const unsigned q = base + unsigned(j - lo) * pitchb;
            w.rq = ldo(rho, q + mdb);
            w.rl = ldo(rho, q + mdb - 4u);
            w.ru = ldo(rho, q + mdb - pitchb);
            w.vxj = ldo(F.vx_in, q);
            w.m2 = xs_h ? ldo(F.m_in[2], q) : 0.0f;
            w.m3 = (j <= 1 || j >= nz + 1) ? ldo(F.m_in[3], q) : 0.0f;
        }
    };
    auto vz_row = [&](int row) -> float {
        return (mcol && row >= lo && row <= nz + pml + R) ? ldo(F.vz_in, base + unsigned(row - lo) * pitchb) : 0.0f;
    };
    // V[k] = vz_in at row r - R + 1 + k;  P[k] = new p at row r - 2R + 1 + k
    float V[2 * R], P[2 * R];
#pragma unroll
    for (int k = 0; k < 2 * R - 1; ++k) V[k] = vz_row(rs - R + 1 + k);
#pragma unroll
    for (int k = 0; k < 2 * R; ++k) P[k] = 0.0f;
    Row cur, nxt;
    load(rs, cur);
    for (int r = rs; r <= re; ++r) {
        if (r < re) load(r + 1, nxt);  // in flight while row r is computed
        V[2 * R - 1] = cur.vzf;
        // ---- p at (col, r)
        float pn = 0.0f;
        if (pcol && r >= -pml + 1 && r <= nz + pml) {
            const unsigned q = base + unsigned(r - lo) * pitchb;
            float sx = c.c[0] * (cur.vx[R] - cur.vx[R - 1]);
            float sz = c.c[0] * (V[R] - V[R - 1]);
#pragma unroll
            for (int m = 2; m <= R; ++m) {
                sx = sx + c.c[m - 1] * (cur.vx[R - 1 + m] - cur.vx[R - m]);
                sz = sz + c.c[m - 1] * (V[R - 1 + m] - V[R - m]);
            }
            float pdx = idx * sx;
            float pdz = idz * sz;
            const bool wr = own && r >= z0 && r <= z1;
            if (xs_i) {
                const float mem = ax * pdx + bx * cur.m0;
                if (wr) sto(F.m_out[0], q, mem);
                pdx = pdx + mem;
            }
            if (r <= 0 || r >= nz + 1) {
                const float mem = azi[r + pml - 1] * pdz + bzi[r + pml - 1] * cur.m1;
                if (wr) sto(F.m_out[1], q, mem);
                pdz = pdz + mem;
            }
            pn = cur.p - dt * cur.bk * (pdx + pdz);
            if (wr) sto(F.p_out, q, pn);
        }
#pragma unroll
        for (int k = 0; k < 2 * R - 1; ++k) P[k] = P[k + 1];
        P[2 * R - 1] = pn;
        // ---- v at row j = r - R + 1
        const int j = r - R + 1;
        if (j >= z0) {
            float px[2 * R];  // new p of row j at columns col - R .. col + R - 1
#pragma unroll
            for (int k = 0; k < 2 * R; ++k) px[k] = __shfl(P[R], l - R + k, 64);
            if (own) {
                const unsigned q = base + unsigned(j - lo) * pitchb;
                float s = c.c[0] * (px[R] - px[R - 1]);
#pragma unroll
                for (int m = 2; m <= R; ++m) s = s + c.c[m - 1] * (px[R + m - 1] - px[R - m]);
                float d = idx * s;
                if (xs_h) {
                    const float mem = axh_ * d + bxh_ * cur.m2;
                    sto(F.m_out[2], q, mem);
                    d = d + mem;
                }
                float reff = 0.5f * (cur.rq + cur.rl);
                sto(F.vx_out, q, cur.vxj - dt / reff * d);
                s = c.c[0] * (P[R] - P[R - 1]);
#pragma unroll
                for (int m = 2; m <= R; ++m) s = s + c.c[m - 1] * (P[R + m - 1] - P[R - m]);
                d = idz * s;
                if (j <= 1 || j >= nz + 1) {
                    const float mem = azh[j + pml - 1] * d + bzh[j + pml - 1] * cur.m3;
                    sto(F.m_out[3], q, mem);
                    d = d + mem;
                }
                reff = 0.5f * (cur.rq + cur.ru);
                sto(F.vz_out, q, V[0] - dt / reff * d);
            }
        }
#pragma unroll
        for (int k = 0; k < 2 * R - 1; ++k) V[k] = V[k + 1];
        cur = nxt;
    }
}

// One tap of a point source or receiver: field cell and two weights.
struct Tap {
    long long cell;  // absolute element offset
    float w1, w2;
    int row;         // amplitude-table row (sources) or receiver index (adjoint)
    int field;       // 0 p, 1 vx, 2 vz
    int tbeg, tend;  // steps during which a source tap is active (nbeg..nend)
};

// ---------------------------------------------------------------------------
// After the update of step t: add_source(t) (module_wavefield.f90:192-263),
// then sample the receivers for step t+1
// (submodule_forward_modeling.f90:58-79) and save the boundary record of
// step t+1 (module_boundary_saving.f90:119-133).  One block per shot, so
// the samples see the injected source.
// ---------------------------------------------------------------------------
template <int R>
__global__ __launch_bounds__(1024) void k_post_step(
    const Shot* __restrict__ shots, float* p, float* vx, float* vz,
    const int* __restrict__ inj_cell_beg, const Tap* __restrict__ inj_taps,
    const float* __restrict__ amp_tab, int nt, int t, const int* __restrict__ rec_beg,
    const Tap* __restrict__ rec_taps, const float* __restrict__ rec_w, float* traces, float* strips,
    int pml) {
    const Shot S = shots[blockIdx.x];
    // source injection at step t (t in 1..nt); cells in CSR order
    for (int c = threadIdx.x; c < S.ninj; c += blockDim.x) {
        const int b = inj_cell_beg[S.inj0 + c], e = inj_cell_beg[S.inj0 + c + 1];
        const Tap first = inj_taps[b];
        float* f = first.field == 0 ? p : (first.field == 1 ? vx : vz);
        float acc = f[first.cell];
        for (int k = b; k < e; ++k) {
            const Tap tp = inj_taps[k];
            if (t < tp.tbeg || t > tp.tend) continue;
            acc = acc + amp_tab[size_t(tp.row) * nt + (t - 1)] * tp.w1 * tp.w2;
        }
        f[first.cell] = acc;
    }
    __syncthreads();
    if (t + 1 > nt) return;
    for (int r = threadIdx.x; traces && r < S.nr; r += blockDim.x) {
        const int g = S.rec0 + r;
        const float w = rec_w[g];
        float acc = 0.0f;
        for (int k = rec_beg[g]; k < rec_beg[g + 1]; ++k) {
            const Tap tp = rec_taps[k];
            acc = acc + p[tp.cell] * tp.w1 * tp.w2 * w;
        }
        traces[S.trace0 + size_t(t) * S.nr + r] = acc;
    }
    if (strips) {
        const int nx = S.nx, nz = S.nz, lo = 1 - pml - R;
        const int na = R * (nz + 2 * R), nc = (nx + 2 * R) * R;
        const int len = 2 * na + 2 * nc;
        float* rec = strips + S.strip0 + size_t(t) * len;
        for (int k = threadIdx.x; k < len; k += blockDim.x) {
            int i, j;
            const float* src;
            if (k < na) { i = 2 - R + k % R; j = 1 - R + k / R; src = vx; }
            else if (k < 2 * na) { const int kk = k - na; i = nx + 1 + kk % R; j = 1 - R + kk / R; src = vx; }
            else if (k < 2 * na + nc) { const int kk = k - 2 * na; i = 1 - R + kk % (nx + 2 * R); j = 2 - R + kk / (nx + 2 * R); src = vz; }
            else { const int kk = k - 2 * na - nc; i = 1 - R + kk % (nx + 2 * R); j = nz + 1 + kk / (nx + 2 * R); src = vz; }
            rec[k] = src[size_t(S.off) + size_t(i - lo) + size_t(j - lo) * size_t(S.pitch)];
        }
    }
}

// ---------------------------------------------------------------------------
// Reverse pass, OWL's algorithm.  Inject the boundary record of step t into
// the forward field (module_boundary_saving.f90:139-153) after keeping the
// pre-injection values of the edge columns/rows the imaging condition reads.
// ---------------------------------------------------------------------------
template <int R>
__global__ void k_inject_strips(const Shot* __restrict__ shots, float* vx, float* vz,
                                const float* __restrict__ strips, float* edge, int t, int pml) {
    const Shot S = shots[blockIdx.y];
    const int nx = S.nx, nz = S.nz, lo = 1 - pml - R;
    const int na = R * (nz + 2 * R), nc = (nx + 2 * R) * R;
    const int len = 2 * na + 2 * nc;
    const float* rec = strips + S.strip0 + size_t(t - 1) * len;
    const size_t pitch = size_t(S.pitch);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < len; k += gridDim.x * blockDim.x) {
        int i, j, e = -1;
        float* dst;
        if (k < na) {
            i = 2 - R + k % R; j = 1 - R + k / R; dst = vx;
            if (i == 1 && j >= 1 && j <= nz) e = j - 1;
        } else if (k < 2 * na) {
            const int kk = k - na; i = nx + 1 + kk % R; j = 1 - R + kk / R; dst = vx;
            if (i == nx + 1 && j >= 1 && j <= nz) e = nz + j - 1;
        } else if (k < 2 * na + nc) {
            const int kk = k - 2 * na; i = 1 - R + kk % (nx + 2 * R); j = 2 - R + kk / (nx + 2 * R); dst = vz;
            if (j == 1 && i >= 1 && i <= nx) e = 2 * nz + i - 1;
        } else {
            const int kk = k - 2 * na - nc; i = 1 - R + kk % (nx + 2 * R); j = nz + 1 + kk / (nx + 2 * R); dst = vz;
            if (j == nz + 1 && i >= 1 && i <= nx) e = 2 * nz + nx + i - 1;
        }
        const size_t q = size_t(S.off) + size_t(i - lo) + size_t(j - lo) * pitch;
        if (e >= 0) edge[S.edge0 + e] = dst[q];
        dst[q] = rec[k];
    }
}

// Adjoint source injection (submodule_fwi_gradient.f90:126-147): cells in CSR
// order, contributions of each cell in receiver order.
__device__ __forceinline__ void inject_adjoint_cells(const Shot& S, int c0, int dc, float* pr,
                                                     const int* __restrict__ cell_beg, const Tap* __restrict__ taps,
                                                     const float* __restrict__ rec_w, const float* __restrict__ adj,
                                                     int t) {
    for (int c = c0; c < S.nadj; c += dc) {
        const int b = cell_beg[S.adj0 + c], e = cell_beg[S.adj0 + c + 1];
        const long long cell = taps[b].cell;
        float acc = pr[cell];
        for (int k = b; k < e; ++k) {
            const Tap tp = taps[k];
            const float a = adj[S.trace0 + size_t(t - 1) * S.nr + tp.row];
            acc = acc + a * tp.w1 * tp.w2 * rec_w[S.rec0 + tp.row];
        }
        pr[cell] = acc;
    }
}

__global__ void k_inject_adjoint(const Shot* __restrict__ shots, float* pr,
                                 const int* __restrict__ cell_beg, const Tap* __restrict__ taps,
                                 const float* __restrict__ rec_w, const float* __restrict__ adj,
                                 int t) {
    inject_adjoint_cells(shots[blockIdx.y], int(blockIdx.x * blockDim.x + threadIdx.x), int(gridDim.x * blockDim.x),
                         pr, cell_beg, taps, rec_w, adj, t);
}

// Imaging condition of OWL's gradient (submodule_fwi_gradient.f90:201-254,
// energy preconditioning off): over the window interior
//   g_vp  -= (p - prev_p) * pr / (rho vp^3)
//   g_rho += (sum2 vx - sum2 prev_vx)(sum2 vxr) + (z analogue)
// with prev = 0 on the first reverse step.  FP64 accumulators.
template <int R>
__global__ __launch_bounds__(256) void k_image_owl(const Shot* __restrict__ shots,
                                                   const float* __restrict__ p,
                                                   const float* __restrict__ vx,
                                                   const float* __restrict__ vz,
                                                   const float* __restrict__ pp,
                                                   const float* __restrict__ pvx,
                                                   const float* __restrict__ pvz,
                                                   const float* __restrict__ edge,
                                                   const float* __restrict__ pr,
                                                   const float* __restrict__ vxr,
                                                   const float* __restrict__ vzr,
                                                   const float* __restrict__ den, double* gv,
                                                   double* gr, int prev_zero, int pml, double* energy,
                                                   size_t ne, int t) {
    const Shot S = shots[blockIdx.y];
    if (t < S.hnt) return;  // this shot's reverse loop has ended (whole block)
    const int nx = S.nx, nz = S.nz, lo = 1 - pml - R;
    const size_t pitch = size_t(S.pitch);
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nx * nz; e += gridDim.x * blockDim.x) {
        const int i = 1 + e % nx, j = 1 + e / nx;
        const size_t q = size_t(S.off) + size_t(i - lo) + size_t(j - lo) * pitch;
        const float pc = p[q];
        const float pv = prev_zero ? 0.0f : pp[q];
        const float sp = pc - pv;
        const float inc_v = sp * pr[q] / den[S.g0 + e];
        float px0 = 0.f, px1 = 0.f, pz0 = 0.f, pz1 = 0.f;
        if (!prev_zero) {
            px0 = (i == 1) ? edge[S.edge0 + j - 1] : pvx[q];
            px1 = (i == nx) ? edge[S.edge0 + nz + j - 1] : pvx[q + 1];
            pz0 = (j == 1) ? edge[S.edge0 + 2 * nz + i - 1] : pvz[q];
            pz1 = (j == nz) ? edge[S.edge0 + 2 * nz + nx + i - 1] : pvz[q + pitch];
        }
        const float sx = (vx[q] + vx[q + 1]) - (px0 + px1);
        const float rx = vxr[q] + vxr[q + 1];
        const float sz = (vz[q] + vz[q + pitch]) - (pz0 + pz1);
        const float rz = vzr[q] + vzr[q + pitch];
        gv[S.g0 + e] = gv[S.g0 + e] - double(inc_v);
        gr[S.g0 + e] = gr[S.g0 + e] + double(sx * rx) + double(sz * rz);
        if (energy) {
            // energy preconditioning (compute_gradient, :223-265): source and
            // receiver energies of the two imaging terms
            const size_t g = size_t(S.g0 + e);
            energy[g] += double(sp * sp / den[S.g0 + e]);
            energy[ne + g] += double(pr[q] * pr[q] / den[S.g0 + e]);
            energy[2 * ne + g] += double(sx * sx + sz * sz);
            energy[3 * ne + g] += double(rx * rx + rz * rz);
        }
    }
}

// Energy preconditioning of one shot window per block (compute_adjoint,
// submodule_fwi_gradient.f90:173-184): each energy plus 1e-3 of its largest
// value, the gradients divided by sqrt(source energy x receiver energy).
__global__ __launch_bounds__(256) void k_energy_precond(const Shot* __restrict__ shots, double* gv, double* gr,
                                                        const double* __restrict__ energy, size_t ne) {
    const Shot S = shots[blockIdx.x];
    const int n = S.nx * S.nz;
    __shared__ double red[4][256];
    double m[4] = {0, 0, 0, 0};
    for (int e = threadIdx.x; e < n; e += blockDim.x)
        for (int k = 0; k < 4; ++k) m[k] = fmax(m[k], energy[k * ne + S.g0 + e]);
    for (int k = 0; k < 4; ++k) red[k][threadIdx.x] = m[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w)
            for (int k = 0; k < 4; ++k) red[k][threadIdx.x] = fmax(red[k][threadIdx.x], red[k][threadIdx.x + w]);
        __syncthreads();
    }
    const double a0 = 1.0e-3 * red[0][0], a1 = 1.0e-3 * red[1][0], a2 = 1.0e-3 * red[2][0], a3 = 1.0e-3 * red[3][0];
    for (int e = threadIdx.x; e < n; e += blockDim.x) {
        const size_t g = size_t(S.g0 + e);
        gv[g] = gv[g] / sqrt((energy[g] + a0) * (energy[ne + g] + a1));
        gr[g] = gr[g] / sqrt((energy[2 * ne + g] + a2) * (energy[3 * ne + g] + a3));
    }
}

// L2 waveform misfit and adjoint source for one shot per block
// (compute_adjsrc_waveform, module_inversion_adjoint_source.f90:371-383;
// trace discard rule :1042-1054).  Traces are [nt][nr] per shot.
__global__ __launch_bounds__(512) void k_adjoint_source(const Shot* __restrict__ shots,
                                                        const float* __restrict__ syn,
                                                        const float* __restrict__ obs,
                                                        const float* __restrict__ rec_w,
                                                        float* adj, double* misfit, int nt,
                                                        float threshold, double* scratch) {
    const Shot S = shots[blockIdx.x];
    __shared__ double red[512];
    __shared__ double mo_s, ms_s;
    double* nobs = scratch + 2 * size_t(S.rec0);
    double* nsyn = nobs + S.nr;
    double lmo = 0, lms = 0;
    for (int r = threadIdx.x; r < S.nr; r += blockDim.x) {
        double a = 0, b = 0;
        for (int t = 0; t < nt; ++t) {
            const double o = obs[S.trace0 + size_t(t) * S.nr + r];
            const double s = syn[S.trace0 + size_t(t) * S.nr + r];
            a += o * o;
            b += s * s;
        }
        nobs[r] = sqrt(a);
        nsyn[r] = sqrt(b);
        lmo = fmax(lmo, nobs[r]);
        lms = fmax(lms, nsyn[r]);
    }
    red[threadIdx.x] = lmo;
    __syncthreads();
    for (int w = 256; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) mo_s = red[0];
    __syncthreads();
    red[threadIdx.x] = lms;
    __syncthreads();
    for (int w = 256; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) ms_s = red[0];
    __syncthreads();
    const double mo = mo_s, ms = ms_s;
    // Neumaier-compensated FP64 sum of the squared residuals
    double sum = 0.0, comp = 0.0;
    for (int r = threadIdx.x; r < S.nr; r += blockDim.x) {
        const bool keep = !(rec_w[S.rec0 + r] == 0.0f || nobs[r] < threshold * mo ||
                            nsyn[r] < threshold * ms || nobs[r] * nsyn[r] == 0.0);
        for (int t = 0; t < nt; ++t) {
            const size_t q = S.trace0 + size_t(t) * S.nr + r;
            if (keep) {
                const float d = obs[q] - syn[q];
                const double v = double(d) * double(d);
                const double tt = sum + v;
                comp += (fabs(sum) >= fabs(v)) ? (sum - tt) + v : (v - tt) + sum;
                sum = tt;
                adj[q] = syn[q] - obs[q];
            } else {
                adj[q] = 0.0f;
            }
        }
    }
    red[threadIdx.x] = sum + comp;
    __syncthreads();
    for (int w = 256; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) misfit[blockIdx.x] = red[0];
}

// Sum of the shots' window gradients into the full model, shot order fixed
// (module_gradient.f90:366-388 followed by the all-reduce at :411-413).
__global__ void k_merge(const Shot* __restrict__ shots, int nshot, const double* __restrict__ gw,
                        double* gfull, int nx_full, int nz_full) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nx_full * nz_full;
         e += gridDim.x * blockDim.x) {
        const int iz = e % nz_full, ix = e / nz_full;  // full model, z fastest
        double acc = 0.0;
        for (int s = 0; s < nshot; ++s) {
            const Shot S = shots[s];
            const int i = ix - (S.nxbeg - 1), j = iz - (S.nzbeg - 1);
            if (i >= 0 && i < S.nx && j >= 0 && j < S.nz) acc += gw[S.g0 + size_t(i) + size_t(j) * S.nx];
        }
        gfull[e] = acc;
    }
}


// ===========================================================================
// Exact discrete adjoint.  Transpose of one forward step u_t = V(P(u_{t-1}))
// (+ source), including the CFS-PML memory variables; no OWL counterpart.
// Adjoint variables: ph, vxh, vzh (of p, vx, vz), psx, psz (of the p-sweep
// memory), phx, phz (of the v-sweep memory); ex, ez = the transposed
// v-sweep derivative input, dx, dz = the transposed p-sweep input.
// Reverse step t:  p-pass  ph -= D+ e ; gK += (p_t - p_{t-1})/K ph ;
//                          d = -dt K ph (+ PML, psi update)
//                  v-pass  gX, gZ += (v_{t-1} - v_t)/rho_eff vh (old vh);
//                          vh -= D- d ; e = -(dt/rho_eff) vh (+ PML, phi update)
//                  receivers R^T r_t into ph.
// ===========================================================================
struct AdjSet {
    float* ph;
    float* vxh;
    float* vzh;
    float* psx;
    float* psz;
    float* phx;
    float* phz;
    float* ex;
    float* ez;
    float* dx;
    float* dz;
    double* gK;
    double* gX;
    double* gZ;
    double* gR;  // fused adjoint: neighbour-summed density term
};

template <int R>
__global__ __launch_bounds__(kBX * kBY) void k_adj_p(const Shot* __restrict__ shots,
                                                     const int* __restrict__ tile_shot, AdjSet A,
                                                     const float* __restrict__ bk,
                                                     const float* __restrict__ prof, Coef c, float dt,
                                                     float idx, float idz, int pml,
                                                     const float* __restrict__ p_cur,
                                                     const float* __restrict__ p_prev) {
    const int tile = blockIdx.x;
    const Shot S = shots[shot_of_tile(tile_shot, tile)];
    const int lt = tile - S.tile0;
    const int tx = lt % S.ntx, tz = lt / S.ntx;
    const int i = -pml + 1 + tx * kBX + int(threadIdx.x);
    const int nx = S.nx, nz = S.nz;
    if (i > nx + pml) return;
    const int lo = 1 - pml - R;
    const int lx = nx + 2 * pml, lz = nz + 2 * pml;
    const float* axi = prof + S.prof;
    const float* bxi = axi + lx;
    const float* azi = prof + S.prof + 4 * lx;
    const float* bzi = azi + lz;
    const bool xslab = (i <= 0 || i >= nx + 1);
    const size_t pitch = size_t(S.pitch);
#pragma unroll
    for (int k = 0; k < kRows / kBY; ++k) {
        // kBX = 64: the row is wavefront-uniform (scalar registers)
        const int j = -pml + 1 + tz * kRows + __builtin_amdgcn_readfirstlane(int(threadIdx.y)) + kBY * k;
        if (j > nz + pml) break;
        const size_t q = size_t(S.off) + size_t(i - lo) + size_t(j - lo) * pitch;
        float sx = 0.0f, sz = 0.0f;
#pragma unroll
        for (int m = 1; m <= R; ++m) {
            sx = sx + c.c[m - 1] * (A.ex[q + m] - A.ex[q + 1 - m]);
            sz = sz + c.c[m - 1] * (A.ez[q + m * pitch] - A.ez[q - (m - 1) * pitch]);
        }
        const float ph = A.ph[q] - (idx * sx + idz * sz);
        A.ph[q] = ph;
        const float K = bk[mq(S, q)];
        A.gK[q] += double((p_cur[q] - p_prev[q]) / K * ph);
        const float qv = -dt * K * ph;
        float dxv = qv, dzv = qv;
        if (xslab) {
            const float s = A.psx[q] + qv;
            dxv = qv + axi[i + pml - 1] * s;
            A.psx[q] = bxi[i + pml - 1] * s;
        }
        if (j <= 0 || j >= nz + 1) {
            const float s = A.psz[q] + qv;
            dzv = qv + azi[j + pml - 1] * s;
            A.psz[q] = bzi[j + pml - 1] * s;
        }
        A.dx[q] = dxv;
        A.dz[q] = dzv;
    }
}

template <int R>
__global__ __launch_bounds__(kBX * kBY) void k_adj_v(const Shot* __restrict__ shots,
                                                     const int* __restrict__ tile_shot, AdjSet A,
                                                     const float* __restrict__ rho,
                                                     const float* __restrict__ prof, Coef c, float dt,
                                                     float idx, float idz, int pml,
                                                     const float* __restrict__ vx_cur,
                                                     const float* __restrict__ vx_prev,
                                                     const float* __restrict__ vz_cur,
                                                     const float* __restrict__ vz_prev) {
    const int tile = blockIdx.x;
    const Shot S = shots[shot_of_tile(tile_shot, tile)];
    const int lt = tile - S.tile0;
    const int tx = lt % S.ntx, tz = lt / S.ntx;
    const int i = -pml + 1 + tx * kBX + int(threadIdx.x);
    const int nx = S.nx, nz = S.nz;
    if (i > nx + pml) return;
    const int lo = 1 - pml - R;
    const int lx = nx + 2 * pml, lz = nz + 2 * pml;
    const float* axh = prof + S.prof + 2 * lx;
    const float* bxh = axh + lx;
    const float* azh = prof + S.prof + 4 * lx + 2 * lz;
    const float* bzh = azh + lz;
    const bool xslab = (i <= 1 || i >= nx + 1);
    const size_t pitch = size_t(S.pitch);
#pragma unroll
    for (int k = 0; k < kRows / kBY; ++k) {
        // kBX = 64: the row is wavefront-uniform (scalar registers)
        const int j = -pml + 1 + tz * kRows + __builtin_amdgcn_readfirstlane(int(threadIdx.y)) + kBY * k;
        if (j > nz + pml) break;
        const size_t q = size_t(S.off) + size_t(i - lo) + size_t(j - lo) * pitch;
        const float rx = 0.5f * (rho[mq(S, q)] + rho[mq(S, q) - 1]);
        const float rz = 0.5f * (rho[mq(S, q)] + rho[mq(S, q) - pitch]);
        float vxh = A.vxh[q], vzh = A.vzh[q];
        A.gX[q] += double((vx_prev[q] - vx_cur[q]) / rx * vxh);
        A.gZ[q] += double((vz_prev[q] - vz_cur[q]) / rz * vzh);
        float sx = 0.0f, sz = 0.0f;
#pragma unroll
        for (int m = 1; m <= R; ++m) {
            sx = sx + c.c[m - 1] * (A.dx[q + (m - 1)] - A.dx[q - m]);
            sz = sz + c.c[m - 1] * (A.dz[q + (m - 1) * pitch] - A.dz[q - m * pitch]);
        }
        vxh = vxh - idx * sx;
        vzh = vzh - idz * sz;
        A.vxh[q] = vxh;
        A.vzh[q] = vzh;
        float r = -(dt / rx) * vxh;
        float e = r;
        if (xslab) {
            const float s = A.phx[q] + r;
            e = r + axh[i + pml - 1] * s;
            A.phx[q] = bxh[i + pml - 1] * s;
        }
        A.ex[q] = e;
        r = -(dt / rz) * vzh;
        e = r;
        if (j <= 1 || j >= nz + 1) {
            const float s = A.phz[q] + r;
            e = r + azh[j + pml - 1] * s;
            A.phz[q] = bzh[j + pml - 1] * s;
        }
        A.ez[q] = e;
    }
}

// Fused exact-adjoint step (the p-pass, then the v-pass with this step's
// dx, dz), organised like k_step_fused: one wavefront per tile of
// 64 - (2R - 1) columns, marching down z with the p-pass R - 1 rows ahead.
// dx of a row reaches the neighbouring lanes by shuffles, dz stays in a
// register queue, so neither is stored.  The p-pass is recomputed by the
// halo lanes and rows, so its inputs ph, psx, psz, ex, ez come from one
// buffer set and its outputs go to the other; vxh, vzh, phx, phz and the
// FP64 accumulators are touched by the owning lanes only and stay in place.
// Arithmetic as k_adj_p / k_adj_v, so the results are bit-identical.
struct AdjIO {
    const float *ph, *psx, *psz, *ex, *ez;   // p-pass inputs
    float *pho, *psxo, *pszo, *exo, *ezo;    // p-pass outputs, v-pass ex/ez outputs
};

template <int R>
__global__ __launch_bounds__(256) void k_adj_fused(const Shot* __restrict__ shots, const int* __restrict__ ftile_shot,
                                                   int nftile, AdjIO io, AdjSet A, const float* __restrict__ bk,
                                                   const float* __restrict__ rho, const float* __restrict__ prof,
                                                   Coef c, float dt, float idx, float idz, int pml, int zc,
                                                   const float* __restrict__ p_cur, const float* __restrict__ p_prev,
                                                   const float* __restrict__ vx_cur, const float* __restrict__ vx_prev,
                                                   const float* __restrict__ vz_cur, const float* __restrict__ vz_prev) {
    constexpr int W = 64 - (2 * R - 1);
    // one wavefront per tile (blockDim.x = 64): the tile, its shot and every
    // row index are wavefront-uniform, so they live in scalar registers and
    // the per-row PML coefficients come through scalar loads
    const int tile = __builtin_amdgcn_readfirstlane(int(blockIdx.x) * int(blockDim.y) + int(threadIdx.y));
    if (tile >= nftile) return;
    const Shot S = shots[ftile_shot[tile]];
    const int lt = tile - S.atile0;
    const int tx = lt % S.fntx, tz = lt / S.fntx;
    const int l = int(threadIdx.x);
    const int nx = S.nx, nz = S.nz;
    const int lo = 1 - pml - R;
    const int lx = nx + 2 * pml, lz = nz + 2 * pml;
    const int col = -pml + 1 + tx * W - R + l;
    const int z0 = -pml + 1 + tz * zc, z1 = min(z0 + zc - 1, nz + pml);
    const bool pcol = col >= -pml + 1 && col <= nx + pml;
    const bool mcol = col <= nx + pml + R;
    const bool own = l >= R && l <= 64 - R && col <= nx + pml;
    const float* axi = prof + S.prof;
    const float* bxi = axi + lx;
    const float* axh = prof + S.prof + 2 * lx;
    const float* bxh = axh + lx;
    const float* azi = prof + S.prof + 4 * lx;
    const float* bzi = azi + lz;
    const float* azh = azi + 2 * lz;
    const float* bzh = azh + lz;
    const bool xs_i = col <= 0 || col >= nx + 1, xs_h = col <= 1 || col >= nx + 1;
    float ax = 0.f, bx = 0.f, axh_ = 0.f, bxh_ = 0.f;
    if (pcol && xs_i) {
        ax = axi[col + pml - 1];
        bx = bxi[col + pml - 1];
    }
    if (own && xs_h) {
        axh_ = axh[col + pml - 1];
        bxh_ = bxh[col + pml - 1];
    }
    const unsigned pitchb = unsigned(S.pitch) * 4u;
    const unsigned base = unsigned(S.off + (col - lo)) * 4u;  // byte offset of (col, lo)
    const unsigned mdb = unsigned(S.mdel) * 4u;
    struct Row {
        float ex[2 * R];  // ex at row r, columns col - R + 1 .. col + R
        float ezf;        // ez at row r + R
        float ph, psx, psz, k, pc, pp;
        float rq, rl, ru, vxh, vzh, phx, phz, xc, xp, zc_, zp;
    };
    const int rs = z0 - R, re = z1 + R - 1;
    auto load = [&](int r, Row& w) {
        const int er = r + R;
        w.ezf = (mcol && er >= lo && er <= nz + pml + R) ? ldo(io.ez, base + unsigned(er - lo) * pitchb) : 0.0f;
        if (pcol && r >= -pml + 1 && r <= nz + pml) {
            const unsigned q = base + unsigned(r - lo) * pitchb;
#pragma unroll
            for (int k = 0; k < 2 * R; ++k) w.ex[k] = ldo(io.ex, q - unsigned(R - 1) * 4u + unsigned(k) * 4u);
            w.ph = ldo(io.ph, q);
            w.psx = xs_i ? ldo(io.psx, q) : 0.0f;
            w.psz = (r <= 0 || r >= nz + 1) ? ldo(io.psz, q) : 0.0f;
            w.k = ldo(bk, q + mdb);
            if (own && r >= z0 && r <= z1) {  // the K term is imaged on the tile's own cells only
                w.pc = ldo(p_cur, q);
                w.pp = ldo(p_prev, q);
            }
        }
        const int j = r - R + 1;
        if (own && j >= z0) {
            const unsigned q = base + unsigned(j - lo) * pitchb;
            w.rq = ldo(rho, q + mdb);
            w.rl = ldo(rho, q + mdb - 4u);
            w.ru = ldo(rho, q + mdb - pitchb);
            w.vxh = ldo(A.vxh, q);
            w.vzh = ldo(A.vzh, q);
            w.phx = xs_h ? ldo(A.phx, q) : 0.0f;
            w.phz = (j <= 1 || j >= nz + 1) ? ldo(A.phz, q) : 0.0f;
            w.xc = ldo(vx_cur, q);
            w.xp = ldo(vx_prev, q);
            w.zc_ = ldo(vz_cur, q);
            w.zp = ldo(vz_prev, q);
        }
    };
    auto ez_row = [&](int row) -> float {
        return (mcol && row >= lo && row <= nz + pml + R) ? ldo(io.ez, base + unsigned(row - lo) * pitchb) : 0.0f;
    };
    // E[k] = ez at row r - R + 1 + k;  D[k] = this step's dz at row r - 2R + 1 + k
    float E[2 * R], D[2 * R];
#pragma unroll
    for (int k = 0; k < 2 * R - 1; ++k) E[k] = ez_row(rs - R + 1 + k);
#pragma unroll
    for (int k = 0; k < 2 * R; ++k) D[k] = 0.0f;
    float dxq[R];  // this step's dx of rows r - R + 1 .. r (own column)
#pragma unroll
    for (int k = 0; k < R; ++k) dxq[k] = 0.0f;
    double gpend = 0.0;  // gR of row j - 1 without the cz of row j
    Row cur{}, nxt{};
    load(rs, cur);
    for (int r = rs; r <= re; ++r) {
        if (r < re) load(r + 1, nxt);
        E[2 * R - 1] = cur.ezf;
        // ---- p-pass at (col, r)
        float dxv = 0.0f, dzv = 0.0f;
        if (pcol && r >= -pml + 1 && r <= nz + pml) {
            const unsigned q = base + unsigned(r - lo) * pitchb;
            float sx = 0.0f, sz = 0.0f;
#pragma unroll
            for (int m = 1; m <= R; ++m) {
                sx = sx + c.c[m - 1] * (cur.ex[R - 1 + m] - cur.ex[R - m]);
                sz = sz + c.c[m - 1] * (E[R - 1 + m] - E[R - m]);
            }
            const float ph = cur.ph - (idx * sx + idz * sz);
            const bool wr = own && r >= z0 && r <= z1;
            if (wr) {
                sto(io.pho, q, ph);
                addo(A.gK, 2u * q, double((cur.pc - cur.pp) / cur.k * ph));
            }
            const float qv = -dt * cur.k * ph;
            dxv = qv;
            dzv = qv;
            if (xs_i) {
                const float sv = cur.psx + qv;
                dxv = qv + ax * sv;
                if (wr) sto(io.psxo, q, bx * sv);
            }
            if (r <= 0 || r >= nz + 1) {
                const float sv = cur.psz + qv;
                dzv = qv + azi[r + pml - 1] * sv;
                if (wr) sto(io.pszo, q, bzi[r + pml - 1] * sv);
            }
        }
#pragma unroll
        for (int k = 0; k < 2 * R - 1; ++k) D[k] = D[k + 1];
        D[2 * R - 1] = dzv;
#pragma unroll
        for (int k = 0; k < R - 1; ++k) dxq[k] = dxq[k + 1];
        dxq[R - 1] = dxv;
        // ---- v-pass at row j = r - R + 1
        const int j = r - R + 1;
        if (j >= z0) {
            float xs[2 * R];  // dx of row j at columns col - R .. col + R - 1
#pragma unroll
            for (int k = 0; k < 2 * R; ++k) xs[k] = __shfl(dxq[0], l - R + k, 64);
            // density terms of row j: cx at (col - 1/2, j), cz at (col, j - 1/2);
            // cxr = the next column's cx (zero unless its lane owns it)
            const float rx = 0.5f * (cur.rq + cur.rl);
            const float rz = 0.5f * (cur.rq + cur.ru);
            const float cx = own ? (cur.xp - cur.xc) / rx * cur.vxh : 0.0f;
            const float cxr = __shfl(cx, l + 1, 64);
            if (own) {
                const unsigned q = base + unsigned(j - lo) * pitchb;
                float vxh = cur.vxh, vzh = cur.vzh;
                const float cz = (cur.zp - cur.zc_) / rz * vzh;
                // gR(c) collects (cx(c) + cx(c+1) + cz(c) + cz(c+row)) / 2 where
                // this lane owns both cells; a tile's first column and a chunk's
                // first row also keep their own cx / cz in gX / gZ for the
                // neighbour outside (see k_fold_exact)
                const double s = 0.5 * double(cx) + 0.5 * double(cxr) + 0.5 * double(cz);
                if (j > z0) addo(A.gR, 2u * (q - pitchb), gpend + 0.5 * double(cz));
                else addo(A.gZ, 2u * q, double(cz));
                gpend = s;
                if (j == z1) addo(A.gR, 2u * q, gpend);
                if (l == R) addo(A.gX, 2u * q, double(cx));
                float sx = 0.0f, sz = 0.0f;
#pragma unroll
                for (int m = 1; m <= R; ++m) {
                    sx = sx + c.c[m - 1] * (xs[R + m - 1] - xs[R - m]);
                    sz = sz + c.c[m - 1] * (D[R + m - 1] - D[R - m]);
                }
                vxh = vxh - idx * sx;
                vzh = vzh - idz * sz;
                sto(A.vxh, q, vxh);
                sto(A.vzh, q, vzh);
                float rr = -(dt / rx) * vxh;
                float e = rr;
                if (xs_h) {
                    const float sv = cur.phx + rr;
                    e = rr + axh_ * sv;
                    sto(A.phx, q, bxh_ * sv);
                }
                sto(io.exo, q, e);
                rr = -(dt / rz) * vzh;
                e = rr;
                if (j <= 1 || j >= nz + 1) {
                    const float sv = cur.phz + rr;
                    e = rr + azh[j + pml - 1] * sv;
                    sto(A.phz, q, bzh[j + pml - 1] * sv);
                }
                sto(io.ezo, q, e);
            }
        }
#pragma unroll
        for (int k = 0; k < 2 * R - 1; ++k) E[k] = E[k + 1];
        cur = nxt;
    }
}

// The stored p_t includes the source injected at step t; the K gradient
// needs the increment of the update alone: remove source/K * ph at the
// source taps active at step t (one block per shot).
__device__ __forceinline__ void src_correction_cells(const Shot& S, const int* __restrict__ inj_cell_beg,
                                                     const Tap* __restrict__ inj_taps,
                                                     const float* __restrict__ amp_tab, const float* __restrict__ bk,
                                                     const float* ph, double* gK, int nt, int t) {
    for (int c = threadIdx.x; c < S.ninj; c += blockDim.x) {
        const int b = inj_cell_beg[S.inj0 + c], e = inj_cell_beg[S.inj0 + c + 1];
        const long long cell = inj_taps[b].cell;
        if (inj_taps[b].field != 0) continue;
        float src = 0.0f;
        for (int k = b; k < e; ++k) {
            const Tap tp = inj_taps[k];
            if (t < tp.tbeg || t > tp.tend) continue;
            src = src + amp_tab[size_t(tp.row) * nt + (t - 1)] * tp.w1 * tp.w2;
        }
        gK[cell] -= double(src / bk[mq(S, size_t(cell))] * ph[cell]);
    }
}

__global__ void k_src_correction(const Shot* __restrict__ shots, const int* __restrict__ inj_cell_beg,
                                 const Tap* __restrict__ inj_taps, const float* __restrict__ amp_tab,
                                 const float* __restrict__ bk, const float* __restrict__ ph, double* gK,
                                 int nt, int t) {
    src_correction_cells(shots[blockIdx.x], inj_cell_beg, inj_taps, amp_tab, bk, ph, gK, nt, t);
}

// The fused adjoint's tap work of reverse step t in one launch (one block
// per shot): the source correction (reads ph at the source taps), then the
// adjoint source injection into ph, in the order of the two kernels above.
__global__ __launch_bounds__(512) void k_adjoint_taps(const Shot* __restrict__ shots,
                                                      const int* __restrict__ inj_cell_beg,
                                                      const Tap* __restrict__ inj_taps,
                                                      const float* __restrict__ amp_tab,
                                                      const float* __restrict__ bk, float* ph, double* gK, int nt,
                                                      const int* __restrict__ adj_beg,
                                                      const Tap* __restrict__ adj_taps,
                                                      const float* __restrict__ rec_w, const float* __restrict__ adj,
                                                      int t) {
    const Shot S = shots[blockIdx.x];
    src_correction_cells(S, inj_cell_beg, inj_taps, amp_tab, bk, ph, gK, nt, t);
    __syncthreads();
    inject_adjoint_cells(S, int(threadIdx.x), int(blockDim.x), ph, adj_beg, adj_taps, rec_w, adj, t);
}

// Fold the padded-grid gradients onto the window (edge replication
// transposed) and change variables K = rho vp^2:
//   d/dvp = 2 rho vp dK,  d/drho = drho + vp^2 dK,
// with drho(c) = (gX(c) + gX(c+1) + gZ(c) + gZ(c+row)) / 2 after the
// two-pass adjoint.  After the fused adjoint (merged != 0) gR(c) already
// holds those sums where one lane owned both cells, and gX / gZ hold only a
// tile's first column and a chunk's first row: drho(c) = gR(c) +
// (gX(c+1) + gZ(c+row)) / 2.
template <int R>
__global__ void k_fold_exact(const Shot* __restrict__ shots, const double* __restrict__ gK,
                             const double* __restrict__ gX, const double* __restrict__ gZ,
                             const double* __restrict__ gR, int merged,
                             const float* __restrict__ vpw, const float* __restrict__ rhow, double* gv,
                             double* gr, int pml) {
    const Shot S = shots[blockIdx.y];
    const int nx = S.nx, nz = S.nz, lo = 1 - pml - R;
    const int hix = nx + pml + R, hiz = nz + pml + R;
    const size_t pitch = size_t(S.pitch);
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nx * nz; e += gridDim.x * blockDim.x) {
        const int i = 1 + e % nx, j = 1 + e / nx;
        const int i0 = (i == 1) ? lo : i, i1 = (i == nx) ? hix : i;
        const int j0 = (j == 1) ? lo : j, j1 = (j == nz) ? hiz : j;
        double k_acc = 0.0, r_acc = 0.0;
        for (int jj = j0; jj <= j1; ++jj)
            for (int ii = i0; ii <= i1; ++ii) {
                const size_t q = size_t(S.off) + size_t(ii - lo) + size_t(jj - lo) * pitch;
                k_acc += gK[q];
                const double gx1 = (ii + 1 <= nx + pml) ? gX[q + 1] : 0.0;
                const double gz1 = (jj + 1 <= nz + pml) ? gZ[q + pitch] : 0.0;
                if (merged) r_acc += gR[q] + (0.5 * gx1 + 0.5 * gz1);
                else r_acc += 0.5 * (gX[q] + gx1) + 0.5 * (gZ[q] + gz1);
            }
        const double v = vpw[S.g0 + e], r = rhow[S.g0 + e];
        gv[S.g0 + e] = 2.0 * r * v * k_acc;
        gr[S.g0 + e] = r_acc + v * v * k_acc;
    }
}

// Per-shot maximum |x| of one field buffer (bit pattern of a non-negative
// float, so unsigned max is float max).
// Zero n 32-bit words (stands in for a memset inside captured graphs).
__global__ void k_zero_words(unsigned* p, size_t n) {
    for (size_t e = blockIdx.x * size_t(blockDim.x) + threadIdx.x; e < n; e += size_t(gridDim.x) * blockDim.x) p[e] = 0u;
}

__global__ void k_fill_u32(unsigned* p, unsigned v, int n) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) p[e] = v;
}

// Largest |x| of each shot's field (as float bits, for atomicMax), and a
// flag raised when any value is NaN or infinite: a diverged run, whose
// snapshots cannot be scaled.
__global__ __launch_bounds__(256) void k_absmax(const Shot* __restrict__ shots, const float* __restrict__ f,
                                                unsigned* amax, unsigned* nonfinite) {
    const Shot S = shots[blockIdx.y];
    const size_t n = size_t(S.pitch) * S.nrow;
    float m = 0.0f;
    bool bad = false;
    for (size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n; e += size_t(gridDim.x) * blockDim.x) {
        const float a = fabsf(f[S.off + e]);
        bad |= !(a <= 3.4028235e38f);
        m = fmaxf(m, a);
    }
    if (__syncthreads_or(bad) && threadIdx.x == 0) *nonfinite = 1u;
    __shared__ float red[256];
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(amax + blockIdx.y, __float_as_uint(red[0]));
}

// 16-bit snapshot encodings: IEEE half and bfloat16 (round to nearest even),
// values scaled per shot so that max |x| maps to 2^14.
__device__ __forceinline__ unsigned short enc16(float x, int fmt) {
    if (fmt == 1) {
        _Float16 h = (_Float16)x;
        return __builtin_bit_cast(unsigned short, h);
    }
    unsigned u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float dec16(unsigned short h, int fmt) {
    if (fmt == 1) return (float)__builtin_bit_cast(_Float16, h);
    return __uint_as_float(unsigned(h) << 16);
}
// Scale so that max |x| maps to 2^14; capped at 2^100 so that a field whose
// maximum is subnormal (PML memory early in time) cannot overflow the scale.
__device__ __forceinline__ float snap_scale(unsigned amax_bits) {
    const float m = __uint_as_float(amax_bits);
    if (!(m > 0.0f) || !(m < 3.0e38f)) return 1.0f;
    const float s = 16384.0f / m;
    return s < 1.2676506e30f ? s : 1.2676506e30f;
}

__global__ void k_pack16(const Shot* __restrict__ shots, const float* __restrict__ f,
                         unsigned short* dst, const unsigned* __restrict__ amax, int fmt) {
    const Shot S = shots[blockIdx.y];
    const float sc = snap_scale(amax[blockIdx.y]);
    const size_t n = size_t(S.pitch) * S.nrow;
    for (size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n; e += size_t(gridDim.x) * blockDim.x)
        dst[S.off + e] = enc16(f[S.off + e] * sc, fmt);
}

__global__ void k_unpack16(const Shot* __restrict__ shots, const unsigned short* __restrict__ src, float* f,
                           const unsigned* __restrict__ amax, int fmt) {
    const Shot S = shots[blockIdx.y];
    const float inv = 1.0f / snap_scale(amax[blockIdx.y]);
    const size_t n = size_t(S.pitch) * S.nrow;
    for (size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n; e += size_t(gridDim.x) * blockDim.x)
        f[S.off + e] = dec16(src[S.off + e], fmt) * inv;
}

// ===========================================================================
// Host side
// ===========================================================================
struct ShotHost {
    int id = 0;
    int nx = 0, nz = 0;
    int nxbeg = 1, nzbeg = 1;
    float ox = 0, oz = 0;
    std::vector<owl_source> src;
    std::vector<owl_receiver> rec;
    int hnt = 1;
    float pmlvp = 0;
    float f0max = 0;
};

// ---------------------------------------------------------------------------
// Differentiable path with device-resident tensors: the model from device
// memory and the trace/gradient layout conversions, so nothing crosses PCIe.

// Padded bk = rho vp^2 and rho of each shot's model copy, and rho vp^3, vp,
// rho on the window (imaging arrays), from the full model in device memory
// (element (ix, kz) at kz + ix nz, or at ix + kz nx when xf); per shot the
// largest vp, the largest vp on the window's edge lines and the count of
// non-positive vp or rho (stats[3 s .. 3 s + 2], positive floats compared
// as their bit patterns).  One block row per shot; shots that share a
// window write the same values into the shared copy.
__global__ void k_model_from_device(const Shot* __restrict__ shots, const float* __restrict__ vp,
                                    const float* __restrict__ rho, int nxf, int nzf, int xf, int pml, int R,
                                    float* bk, float* rh, float* den, float* vpw, float* rhow, unsigned* stats) {
    const Shot S = shots[blockIdx.y];
    const int lo = 1 - pml - R, nxp = S.nx + 2 * (pml + R), nzp = S.nz + 2 * (pml + R);
    const long long n = (long long)nxp * nzp;
    unsigned vmax = 0, vedge = 0, bad = 0;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        const int i = lo + int(e % nxp), j = lo + int(e / nxp);
        const int ic = min(max(i, 1), S.nx), jc = min(max(j, 1), S.nz);
        const int ix = S.nxbeg - 1 + ic - 1, kz = S.nzbeg - 1 + jc - 1;
        const size_t m = xf ? size_t(ix) + size_t(kz) * size_t(nxf) : size_t(kz) + size_t(ix) * size_t(nzf);
        const float v = vp[m], r = rho[m];
        const size_t q = size_t((long long)S.off + S.mdel + (i - lo) + (long long)(j - lo) * S.pitch);
        rh[q] = r;
        bk[q] = r * (v * v);
        if (i >= 1 && i <= S.nx && j >= 1 && j <= S.nz) {
            const size_t g = size_t(S.g0) + size_t(i - 1) + size_t(j - 1) * S.nx;
            den[g] = r * (v * v * v);
            vpw[g] = v;
            rhow[g] = r;
            if (!(v > 0.0f) || !(r > 0.0f)) {
                ++bad;
            } else {
                vmax = max(vmax, __float_as_uint(v));
                if (i == 1 || i == S.nx || j == 1 || j == S.nz) vedge = max(vedge, __float_as_uint(v));
            }
        }
    }
    atomicMax(&stats[3 * blockIdx.y], vmax);
    atomicMax(&stats[3 * blockIdx.y + 1], vedge);
    if (bad) atomicAdd(&stats[3 * blockIdx.y + 2], bad);
}

// Traces between the internal [t][r] blocks and the C ABI's [r][t] blocks.
__global__ void k_traces_layout(const Shot* __restrict__ shots, const float* __restrict__ src, float* dst, int nt,
                                int to_abi) {
    const Shot S = shots[blockIdx.y];
    const long long n = (long long)S.nr * nt;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        const int r = int(e / nt), t = int(e % nt);
        const size_t a = size_t(S.trace0) + size_t(r) * nt + t, in = size_t(S.trace0) + size_t(t) * S.nr + r;
        if (to_abi)
            dst[a] = src[in];
        else
            dst[in] = src[a];
    }
}

// The FP64 model gradient (z fastest) as float, z fastest or x fastest.
__global__ void k_grad_to_float(const double* __restrict__ g, float* out, int nxf, int nzf, int xf) {
    const long long n = (long long)nxf * nzf;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        const int kz = int(e % nzf), ix = int(e / nzf);
        out[xf ? size_t(ix) + size_t(kz) * nxf : size_t(e)] = float(g[e]);
    }
}

}  // namespace a2
}  // namespace owl

using namespace owl;
using namespace owl::a2;

struct owl_acoustic2d {
    owl_acoustic_config cfg{};
    int R = 4, pml = 15, nt = 0;
    float dt = 0, idx = 0, idz = 0;
    Coef coef{};
    std::vector<ShotHost> hs;
    std::vector<Shot> sd;          // device descriptors (host copy)
    std::vector<int> tile_shot;
    size_t field_elems = 0;        // elements of one field buffer (all shots)
    size_t model_elems = 0;        // elements of bk / rho (one padded grid per distinct window)
    size_t grad_elems = 0;         // sum of nx*nz
    size_t trace_elems = 0;        // sum of nt*nr
    size_t strip_elems = 0;        // sum of nt*strip_len
    size_t edge_elems = 0;
    int nrec_total = 0;
    int max_ninj = 0, max_nadj = 0;
    int src_t0 = 1, src_t1 = 0;    // steps during which any source tap is active
    double pts_full = 0;           // sum over shots of (nx+2pml)(nz+2pml)
    double pts_model = 0;          // the same over distinct windows (model reads)
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0, last_updates = 0;
    bool have_boundary = false;
    bool fwd_state_valid = false;  // fA holds the final forward state
    bool have_obs = false;         // d_obs holds observed traces
    // per-kernel timing with HIP events on the batch stream
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    struct EvUse { int kind, first; double bytes; };
    std::vector<EvUse> ev_used;    // launches awaiting collection
    size_t ev_next = 0;
    double kern_ms[16] = {0};
    long long kern_launches[16] = {0};
    double kern_bytes[16] = {0};   // algorithmic bytes of all recorded launches, by kind
    std::vector<float> vp_full, rho_full;  // host copy of the full model

    DevBuf<Shot> d_shots;
    DevBuf<int> d_tile_shot;
    DevBuf<float> d_bk, d_rho, d_den, d_prof;
    std::vector<float> h_prof;  // host copy of the PML profiles (rebuilt when a model update moves pmlvp)
    DevBuf<unsigned> d_mstat;
    DevBuf<float> fA[3], fB[3], fM[4];  // forward field (ping-pong) and PML memory
    DevBuf<float> aF[3], aM[4];         // adjoint field and PML memory
    DevBuf<float> d_amp;                // source amplitude rows [row][nt]
    DevBuf<int> d_inj_beg, d_adj_beg, d_rec_beg;
    DevBuf<Tap> d_inj_taps, d_adj_taps, d_rec_taps;
    DevBuf<float> d_rec_w;
    DevBuf<float> d_traces, d_obs, d_adj, d_strips, d_edge;
    DevBuf<double> d_gv, d_gr, d_gfull_v, d_gfull_r, d_misfit, d_scratch;
    DevBuf<float> d_vpw, d_rhow;      // window models (x fastest, per shot at g0)
    // exact adjoint: second working set, adjoint state, FP64 accumulators
    DevBuf<float> fM2[4];
    DevBuf<float> fC[3], fM3[4];      // third working set (fused steps ping-pong)
    // fused forward step: 0 two sweeps, 1 fused, 2 fused for batches of
    // 4 shots or more; z rows per tile
    int step_kernel = 2, fzc = 16, nftile = 0;
    std::vector<int> ftile_shot;
    DevBuf<int> d_ftile_shot;
    // fused adjoint: z rows per tile and its tile list; the autotuner picks
    // fzc and azc on the first exact gradient unless the caller set them
    int azc = 8, natile = 0;
    std::vector<int> atile_shot;
    DevBuf<int> d_atile_shot;
    bool autotune = true, tuned = false;
    DevBuf<float> adj[11];            // ph vxh vzh psx psz phx phz ex ez dx dz
    DevBuf<float> adj2[5];            // fused adjoint: second ph psx psz ex ez
    int adj_cur = 0;                  // which of the two holds the current ph psx psz ex ez
    DevBuf<double> gK, gX, gZ, gR;
    struct Slots {                    // stored forward states
        int fmt = 0, nfield = 3, nslot = 0;
        DevBuf<float> f32;
        DevBuf<unsigned short> h16;
        DevBuf<unsigned> amax;
    } full, snaps, seg;
    int store_policy = -1, store_param = 0, store_fmt = 0;  // what the last forward stored
    bool have_store = false;
    long long store_id = 0, store_next = 0;  // identifies the stored states (0: none)
    DevBuf<unsigned> d_nonfinite;    // raised by k_absmax when a 16-bit snapshot saw NaN/Inf
    std::vector<std::vector<int>> binom_split;  // cached optimal splits [s][l]
    int binom_n = 0, binom_s = -1;
    bool binom_carry = false;
    long long binom_steps = 0;       // forward steps of the cached schedule
    DevBuf<float> fK[6];             // carried states of the binomial schedule (2 x p, vx, vz)
    // HIP graph of the exact-adjoint gradient: captured on the second call
    // with the same schedule and buffers, replayed afterwards
    int use_graphs = 1;
    bool capturing = false;
    // the reverse call of a split binomial gradient replays the schedule's
    // first descent on the host only: its results are still in the buffers
    bool skip_gpu = false;
    hipGraphExec_t gexec = nullptr;
    std::vector<uintptr_t> gsig, gready;

    bool borrowed_stream = false;    // stream and ev0/ev1 belong to the static-scratch context
    bool energy_precond = false;     // OWL's energy preconditioning of its gradient
    DevBuf<double> d_energy;         // source/receiver energies of the vp and rho terms

    ~owl_acoustic2d() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        for (auto e : ev_pool) (void)hipEventDestroy(e);
        if (borrowed_stream) return;
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

inline size_t cell_of(const owl_acoustic2d* b, const Shot& S, int i, int j) {
    const int lo = 1 - b->pml - b->R;
    // every host-built tap must address the shot's padded array
    if (i < lo || i > S.nx + b->pml + b->R || j < lo || j > S.nz + b->pml + b->R)
        fail(OWL_ERR_ARG, "a source/receiver stencil reaches outside the padded grid; increase pml");
    return size_t(S.off) + size_t(i - lo) + size_t(j - lo) * size_t(S.pitch);
}

inline float full_at(const owl_acoustic2d* b, const std::vector<float>& m, const ShotHost& H,
                     int i, int j) {
    // window node (i, j), clamped to the window (edge replication), in the
    // full (nz, nx) z-fast model
    const int ic = std::min(std::max(i, 1), H.nx), jc = std::min(std::max(j, 1), H.nz);
    return m[size_t(H.nzbeg - 1 + jc - 1) + size_t(H.nxbeg - 1 + ic - 1) * size_t(b->cfg.nz)];
}

// bk = rho*vp**2 and rho, padded by edge replication (prepare_modeling,
// acoustic_iso_2d/module_vars.f90:108-115), and rho*vp**3 for the imaging
// condition (submodule_fwi_gradient.f90:218).
void upload_model(owl_acoustic2d* b) {
    std::vector<float> bk(b->model_elems, 0.f), rh(b->model_elems, 0.f), den(b->grad_elems, 0.f);
    std::vector<float> vpw(b->grad_elems, 0.f), rhow(b->grad_elems, 0.f);
    const int lo = 1 - b->pml - b->R;
    for (size_t s = 0; s < b->hs.size(); ++s) {
        const ShotHost& H = b->hs[s];
        const Shot& S = b->sd[s];
        const int NXp = H.nx + 2 * (b->pml + b->R), NZp = H.nz + 2 * (b->pml + b->R);
        for (int jj = 0; jj < NZp; ++jj)
            for (int ii = 0; ii < NXp; ++ii) {
                const int i = lo + ii, j = lo + jj;
                const float v = full_at(b, b->vp_full, H, i, j), r = full_at(b, b->rho_full, H, i, j);
                const size_t q = size_t((long long)cell_of(b, S, i, j) + S.mdel);
                rh[q] = r;
                bk[q] = r * (v * v);
            }
        for (int j = 1; j <= H.nz; ++j)
            for (int i = 1; i <= H.nx; ++i) {
                const float v = full_at(b, b->vp_full, H, i, j), r = full_at(b, b->rho_full, H, i, j);
                den[size_t(S.g0) + size_t(i - 1) + size_t(j - 1) * H.nx] = r * (v * v * v);
                vpw[size_t(S.g0) + size_t(i - 1) + size_t(j - 1) * H.nx] = v;
                rhow[size_t(S.g0) + size_t(i - 1) + size_t(j - 1) * H.nx] = r;
            }
    }
    b->d_bk.upload(bk, b->stream);
    b->d_rho.upload(rh, b->stream);
    b->d_den.upload(den, b->stream);
    b->d_vpw.upload(vpw, b->stream);
    b->d_rhow.upload(rhow, b->stream);
    OWL_HIP(hipStreamSynchronize(b->stream));
}

// Column-strip tiles of the fused step: 64 - (2R - 1) output columns and
// fzc rows each, one wavefront per tile.
void fused_tiling(owl_acoustic2d* b) {
    const int W = 64 - (2 * b->R - 1);
    b->ftile_shot.clear();
    for (size_t s = 0; s < b->sd.size(); ++s) {
        Shot& S = b->sd[s];
        S.ftile0 = int(b->ftile_shot.size());
        S.fntx = (S.nx + 2 * b->pml + W - 1) / W;
        const int nzc = (S.nz + 2 * b->pml + b->fzc - 1) / b->fzc;
        b->ftile_shot.insert(b->ftile_shot.end(), size_t(S.fntx) * nzc, int(s));
    }
    b->nftile = int(b->ftile_shot.size());
    if (b->ftile_shot.empty()) b->ftile_shot.push_back(0);
    b->atile_shot.clear();
    for (size_t s = 0; s < b->sd.size(); ++s) {
        Shot& S = b->sd[s];
        S.atile0 = int(b->atile_shot.size());
        const int nzc = (S.nz + 2 * b->pml + b->azc - 1) / b->azc;
        b->atile_shot.insert(b->atile_shot.end(), size_t(S.fntx) * nzc, int(s));
    }
    b->natile = int(b->atile_shot.size());
    if (b->atile_shot.empty()) b->atile_shot.push_back(0);
}

// Re-tile the fused kernels and upload the descriptors (synchronous).
void upload_fused_tiling(owl_acoustic2d* b) {
    fused_tiling(b);
    b->d_shots.upload(b->sd, b->stream);
    b->d_ftile_shot.upload(b->ftile_shot, b->stream);
    b->d_atile_shot.upload(b->atile_shot, b->stream);
    OWL_HIP(hipStreamSynchronize(b->stream));
}

void plan(owl_acoustic2d* b, const owl_acoustic_config* cfg, const float* vp, const float* rho,
          int nshot, const owl_shot* shots) {
    if (!cfg || !vp || !rho || nshot < 1 || !shots) fail(OWL_ERR_ARG, "null argument or no shots");
    if (cfg->ny > 1) fail(OWL_ERR_ARG, "the 2D solver requires ny = 1");
    if (!fd_coefficients(cfg->fdhalf, b->coef.c)) fail(OWL_ERR_ARG, "fdhalf must be 2..8");
    if (cfg->nx < 1 || cfg->nz < 1 || !(cfg->dx > 0) || !(cfg->dz > 0) || !(cfg->dt > 0))
        fail(OWL_ERR_ARG, "invalid grid or dt");
    if (cfg->pml < cfg->fdhalf) fail(OWL_ERR_ARG, "pml must be >= fdhalf");
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
    const float f0factor = cfg->f0factor > 0 ? cfg->f0factor : 1.0f;
    b->vp_full.assign(vp, vp + size_t(cfg->nx) * cfg->nz);
    b->rho_full.assign(rho, rho + size_t(cfg->nx) * cfg->nz);
    const int R = b->R, pml = b->pml, nt = b->nt;
    const int fs = cfg->free_surface ? 1 : 0;

    b->hs.assign(nshot, ShotHost{});
    b->sd.assign(nshot, Shot{});
    b->tile_shot.clear();
    long long off = 0, moff = 0, strip = 0, trace = 0, g0 = 0;
    int edge = 0, prof = 0;
    std::vector<float> profs, amp;
    std::vector<int> inj_beg{0}, adj_beg{0}, rec_beg{0};
    std::vector<Tap> inj_taps, adj_taps, rec_taps;
    std::vector<float> rec_w;
    b->pts_full = 0;
    b->pts_model = 0;
    b->nrec_total = 0;   // every per-plan table is rebuilt from zero
    b->max_ninj = 0;
    b->max_nadj = 0;
    b->src_t0 = 1 << 30;
    b->src_t1 = -1;
    for (int s = 0; s < nshot; ++s) {
        const owl_shot& in = shots[s];
        ShotHost& H = b->hs[s];
        Shot& S = b->sd[s];
        H.id = in.id;
        H.nxbeg = in.nxbeg > 0 ? in.nxbeg : 1;
        H.nzbeg = in.nzbeg > 0 ? in.nzbeg : 1;
        const int nxend = in.nxend > 0 ? in.nxend : cfg->nx;
        const int nzend = in.nzend > 0 ? in.nzend : cfg->nz;
        if (H.nxbeg > nxend || nxend > cfg->nx || H.nzbeg > nzend || nzend > cfg->nz)
            fail(OWL_ERR_ARG, "shot window outside the model");
        H.nx = nxend - H.nxbeg + 1;
        H.nz = nzend - H.nzbeg + 1;
        // window origin shot_xbeg (module_utility.f90:430): (nxbeg-1)*dx + ox
        H.ox = float(H.nxbeg - 1) * cfg->dx + cfg->ox;
        H.oz = float(H.nzbeg - 1) * cfg->dz + cfg->oz;
        if (in.ns < 1 || !in.sources) fail(OWL_ERR_ARG, "shot without sources");
        if (in.nr < 1 || !in.receivers) fail(OWL_ERR_ARG, "shot without receivers");
        H.src.assign(in.sources, in.sources + in.ns);
        H.rec.assign(in.receivers, in.receivers + in.nr);

        float vmax = 0.f, pmlvp = 0.f;
        for (int i = 1; i <= H.nx; ++i)
            for (int j = 1; j <= H.nz; ++j) {
                const float v = full_at(b, b->vp_full, H, i, j);
                if (!(v > 0.f) || !(full_at(b, b->rho_full, H, i, j) > 0.f))
                    fail(OWL_ERR_ARG, "vp and rho must be positive");
                vmax = std::max(vmax, v);
                if (i == 1 || i == H.nx || j == 1 || j == H.nz) pmlvp = std::max(pmlvp, v);
            }
        // check_dt_f0 (module_utility.f90:363-404): OWL silently resets dt;
        // here the call is refused so that dt stays what the caller chose.
        const float dts = dt_stable(R, 2, vmax, cfg->dx, 1.0f, cfg->dz);
        if (cfg->dt > dts)
            fail(OWL_ERR_CFL, "dt = " + std::to_string(cfg->dt) + " exceeds the stable dt = " +
                                  std::to_string(dts));
        H.pmlvp = pmlvp;
        for (auto& sc : H.src) H.f0max = std::max(H.f0max, sc.f0);

        // ---- layout
        const int NXp = H.nx + 2 * (pml + R), NZp = H.nz + 2 * (pml + R);
        S.nx = H.nx;
        S.nz = H.nz;
        S.pitch = (NXp + kPitchAlign - 1) / kPitchAlign * kPitchAlign;
        S.nrow = NZp;
        S.off = off;
        off += (long long)S.pitch * NZp;
        off = (off + 63) / 64 * 64;
        // shots on the same window read one copy of the padded model
        {
            long long mo = -1;
            for (int u = 0; u < s; ++u)
                if (b->hs[u].nxbeg == H.nxbeg && b->hs[u].nx == H.nx && b->hs[u].nzbeg == H.nzbeg &&
                    b->hs[u].nz == H.nz) {
                    mo = b->sd[u].off + b->sd[u].mdel;
                    break;
                }
            if (mo < 0) {
                b->pts_model += double(H.nx + 2 * pml) * double(H.nz + 2 * pml);
                mo = moff;
                moff += (long long)S.pitch * NZp;
                moff = (moff + 63) / 64 * 64;
            }
            S.mdel = mo - S.off;
        }
        S.tile0 = int(b->tile_shot.size());
        S.ntx = (H.nx + 2 * pml + kBX - 1) / kBX;
        const int ntz = (H.nz + 2 * pml + kRows - 1) / kRows;
        b->tile_shot.insert(b->tile_shot.end(), size_t(S.ntx) * ntz, s);
        b->pts_full += double(H.nx + 2 * pml) * double(H.nz + 2 * pml);
        S.prof = prof;
        PmlAxis px = cfspml_axis(H.nx, pml, cfg->dx, cfg->dt, pmlvp, H.f0max);
        PmlAxis pz = cfspml_axis(H.nz, pml, cfg->dz, cfg->dt, pmlvp, H.f0max);
        for (auto* v : {&px.ai, &px.bi, &px.ah, &px.bh, &pz.ai, &pz.bi, &pz.ah, &pz.bh})
            profs.insert(profs.end(), v->begin(), v->end());
        prof += 4 * (H.nx + 2 * pml) + 4 * (H.nz + 2 * pml);
        S.nxbeg = H.nxbeg;
        S.nzbeg = H.nzbeg;
        S.g0 = g0;
        g0 += (long long)H.nx * H.nz;
        S.edge0 = edge;
        edge += 2 * H.nz + 2 * H.nx;
        const int strip_len = 2 * R * (H.nz + 2 * R) + 2 * (H.nx + 2 * R) * R;
        S.strip0 = strip;
        strip += (long long)strip_len * nt;
        S.trace0 = trace;
        S.nr = int(H.rec.size());
        trace += (long long)nt * S.nr;

        // ---- sources (prepare_geometry + prepare_stf + add_source)
        const float xmax = float(H.nx - 1) * cfg->dx + H.ox, zmax = float(H.nz - 1) * cfg->dz + H.oz;
        std::vector<Tap> st;  // taps of this shot, ordered by source then tap
        for (size_t k = 0; k < H.src.size(); ++k) {
            const owl_source& sc = H.src[k];
            if (!(sc.x >= H.ox && sc.x <= xmax && sc.z >= H.oz && sc.z <= zmax))
                fail(OWL_ERR_ARG, "source outside its shot window");
            if (sc.mechanism != OWL_EXPLOSION && sc.mechanism != OWL_FORCE)
                fail(OWL_ERR_ARG, "the acoustic solver takes explosion or force sources");
            const Located L = locate(2, sc.x, 0.f, sc.z, H.ox, 0.f, H.oz, cfg->dx, 1.f, cfg->dz);
            int hnt = 0;
            const std::vector<float> stf = source_time_function(sc.wavelet, sc.f0, cfg->dt, f0factor, &hnt);
            if (k == 0) {
                H.hnt = hnt;
                S.hnt = std::max(hnt, 1);
            }
            const int nbeg = fortran_nint(sc.t0 / cfg->dt) + 1;
            const int nend = nbeg + int(stf.size()) - 1;
            b->src_t0 = std::min(b->src_t0, nbeg);
            b->src_t1 = std::max(b->src_t1, nend);
            // amplitude rows: amp = stf(t-nbeg+1)*amp*dt
            std::vector<float> a(nt, 0.f);
            for (int t = std::max(nbeg, 1); t <= std::min(nend, nt); ++t)
                a[t - 1] = stf[t - nbeg] * sc.amp * cfg->dt;
            if (sc.mechanism == OWL_EXPLOSION) {
                const int row = int(amp.size() / nt);
                amp.insert(amp.end(), a.begin(), a.end());
                for (int irz = -kNkw; irz <= kNkw; ++irz)
                    for (int irx = -kNkw; irx <= kNkw; ++irx) {
                        if (fs && !(L.gz + irz >= 2)) continue;
                        const float w1 = L.ix[irx + kNkw], w2 = L.iz[irz + kNkw];
                        if (w1 == 0.f || w2 == 0.f) continue;
                        st.push_back(Tap{(long long)cell_of(b, S, L.gx + irx, L.gz + irz), w1, w2, row, 0, nbeg, nend});
                    }
            } else {
                // force: vx at (hx, gz) and vz at (gx, hz), scaled by sin/cos(polar)/rho
                const float rx = full_at(b, b->rho_full, H, L.hx, L.gz);
                const float rz = full_at(b, b->rho_full, H, L.gx, L.hz);
                const float sp = std::sin(sc.polar), cp = std::cos(sc.polar);
                const int rowx = int(amp.size() / nt);
                for (int t = 0; t < nt; ++t) amp.push_back(sp * a[t] / rx);
                const int rowz = int(amp.size() / nt);
                for (int t = 0; t < nt; ++t) amp.push_back(cp * a[t] / rz);
                for (int irz = -kNkw; irz <= kNkw; ++irz)
                    for (int irx = -kNkw; irx <= kNkw; ++irx) {
                        if (!fs || L.gz + irz >= 2) {
                            const float w1 = L.hxw[irx + kNkw], w2 = L.iz[irz + kNkw];
                            if (w1 != 0.f && w2 != 0.f)
                                st.push_back(Tap{(long long)cell_of(b, S, L.hx + irx, L.gz + irz), w1, w2, rowx, 1, nbeg, nend});
                        }
                        if (!fs || L.hz + irz >= 2) {
                            const float w1 = L.ix[irx + kNkw], w2 = L.hzw[irz + kNkw];
                            if (w1 != 0.f && w2 != 0.f)
                                st.push_back(Tap{(long long)cell_of(b, S, L.gx + irx, L.hz + irz), w1, w2, rowz, 2, nbeg, nend});
                        }
                    }
            }
        }
        // group taps by (field, cell), keeping source order within a cell
        std::stable_sort(st.begin(), st.end(), [](const Tap& a, const Tap& c) {
            return a.field != c.field ? a.field < c.field : a.cell < c.cell;
        });
        S.inj0 = int(inj_beg.size()) - 1;
        S.ninj = 0;
        for (size_t k = 0; k < st.size(); ++k) {
            if (k == 0 || st[k].cell != st[k - 1].cell || st[k].field != st[k - 1].field) {
                if (k) inj_beg.push_back(int(inj_taps.size()));
                ++S.ninj;
            }
            inj_taps.push_back(st[k]);
        }
        if (!st.empty()) inj_beg.push_back(int(inj_taps.size()));
        b->max_ninj = std::max(b->max_ninj, S.ninj);

        // ---- receivers (prepare_geometry; sampling and adjoint injection)
        S.rec0 = b->nrec_total;
        std::vector<Tap> at;
        for (int r = 0; r < S.nr; ++r) {
            const owl_receiver& rc = H.rec[r];
            const bool valid = rc.x >= H.ox && rc.x <= xmax && rc.z >= H.oz && rc.z <= zmax;
            const float w = (valid && rc.weight != 0.f) ? rc.weight : 0.f;
            rec_w.push_back(w);
            if (w != 0.f) {
                const Located L = locate(2, rc.x, 0.f, rc.z, H.ox, 0.f, H.oz, cfg->dx, 1.f, cfg->dz);
                for (int irz = -kNkw; irz <= kNkw; ++irz)
                    for (int irx = -kNkw; irx <= kNkw; ++irx) {
                        const float w1 = L.ix[irx + kNkw], w2 = L.iz[irz + kNkw];
                        if (w1 == 0.f || w2 == 0.f) continue;
                        const Tap tp{(long long)cell_of(b, S, L.gx + irx, L.gz + irz), w1, w2, r, 0, 0, 0};
                        rec_taps.push_back(tp);
                        at.push_back(tp);
                    }
            }
            rec_beg.push_back(int(rec_taps.size()));
        }
        b->nrec_total += S.nr;
        std::stable_sort(at.begin(), at.end(), [](const Tap& a, const Tap& c) { return a.cell < c.cell; });
        S.adj0 = int(adj_beg.size()) - 1;
        S.nadj = 0;
        for (size_t k = 0; k < at.size(); ++k) {
            if (k == 0 || at[k].cell != at[k - 1].cell) {
                if (k) adj_beg.push_back(int(adj_taps.size()));
                ++S.nadj;
            }
            adj_taps.push_back(at[k]);
        }
        if (!at.empty()) adj_beg.push_back(int(adj_taps.size()));
        b->max_nadj = std::max(b->max_nadj, S.nadj);
    }
    b->field_elems = size_t(off);
    b->model_elems = size_t(moff);
    b->grad_elems = size_t(g0);
    b->trace_elems = size_t(trace);
    b->strip_elems = size_t(strip);
    b->edge_elems = size_t(edge);

    // Every table a kernel indexes must cover the ranges the descriptors
    // name; checked here, on the host, before anything is launched.
    for (size_t s = 0; s < b->sd.size(); ++s) {
        const Shot& S = b->sd[s];
        const bool ok =
            S.rec0 >= 0 && size_t(S.rec0) + size_t(S.nr) <= rec_w.size() &&
            size_t(S.rec0) + size_t(S.nr) + 1 <= rec_beg.size() &&
            S.inj0 >= 0 && size_t(S.inj0) + size_t(S.ninj) + (S.ninj ? 1 : 0) <= inj_beg.size() &&
            S.adj0 >= 0 && size_t(S.adj0) + size_t(S.nadj) + (S.nadj ? 1 : 0) <= adj_beg.size() &&
            size_t(S.trace0) + size_t(nt) * size_t(S.nr) <= b->trace_elems &&
            size_t(S.strip0) + size_t(nt) * size_t(2 * R * (S.nz + 2 * R) + 2 * (S.nx + 2 * R) * R) <= b->strip_elems &&
            size_t(S.g0) + size_t(S.nx) * size_t(S.nz) <= b->grad_elems &&
            size_t(S.edge0) + size_t(2 * S.nx + 2 * S.nz) <= b->edge_elems &&
            size_t(S.off) + size_t(S.pitch) * size_t(S.nrow) <= b->field_elems &&
            S.off + S.mdel >= 0 && size_t(S.off + S.mdel) + size_t(S.pitch) * size_t(S.nrow) <= b->model_elems &&
            size_t(S.prof) + size_t(4 * (S.nx + 2 * pml) + 4 * (S.nz + 2 * pml)) <= profs.size();
        if (!ok) fail(OWL_ERR_ARG, "internal layout check failed for shot " + std::to_string(s));
    }
    for (const Tap& tp : inj_taps)
        if (tp.cell < 0 || size_t(tp.cell) >= b->field_elems || size_t(tp.row + 1) * nt > amp.size())
            fail(OWL_ERR_ARG, "internal layout check failed for a source tap");
    for (const Tap& tp : rec_taps)
        if (tp.cell < 0 || size_t(tp.cell) >= b->field_elems) fail(OWL_ERR_ARG, "internal layout check failed for a receiver tap");

    if (!b->stream) {
        OWL_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        OWL_HIP(hipEventCreate(&b->ev0));
        OWL_HIP(hipEventCreate(&b->ev1));
    }
    hipStream_t st = b->stream;
    fused_tiling(b);
    b->d_shots.upload(b->sd, st);
    b->d_tile_shot.upload(b->tile_shot, st);
    b->d_ftile_shot.upload(b->ftile_shot, st);
    b->d_atile_shot.upload(b->atile_shot, st);
    b->d_prof.upload(profs, st);
    b->h_prof = profs;
    if (amp.empty()) amp.push_back(0.f);
    b->d_amp.upload(amp, st);
    b->d_inj_beg.upload(inj_beg, st);
    if (inj_taps.empty()) inj_taps.push_back(Tap{});
    b->d_inj_taps.upload(inj_taps, st);
    b->d_rec_beg.upload(rec_beg, st);
    if (rec_taps.empty()) rec_taps.push_back(Tap{});
    b->d_rec_taps.upload(rec_taps, st);
    b->d_adj_beg.upload(adj_beg, st);
    if (adj_taps.empty()) adj_taps.push_back(Tap{});
    b->d_adj_taps.upload(adj_taps, st);
    b->d_rec_w.upload(rec_w, st);
    for (int k = 0; k < 3; ++k) b->fA[k].alloc(b->field_elems);
    for (int k = 0; k < 4; ++k) b->fM[k].alloc(b->field_elems);
    b->d_traces.alloc(b->trace_elems);
    upload_model(b);
}

FieldSet in_place(float* p, float* vx, float* vz, float* const* m) {
    FieldSet F{};
    F.p_in = F.p_out = p;
    F.vx_in = F.vx_out = vx;
    F.vz_in = F.vz_out = vz;
    for (int k = 0; k < 4; ++k) {
        F.m_in[k] = m[k];
        F.m_out[k] = m[k];
    }
    return F;
}

Sets one_set(float* p, float* vx, float* vz, float* const* m) {
    Sets S{};
    S.s[0] = in_place(p, vx, vz, m);
    S.s[1] = S.s[0];
    return S;
}

enum KernelKind { kKP = 0, kKV = 1, kKPost = 2, kKImage = 3, kKStrips = 4, kKAdj = 5, kKMisfit = 6, kKMerge = 7,
                  kKP2 = 8, kKV2 = 9, kKFused = 13, kKinds = 16 };

void prof_collect(owl_acoustic2d* b);
constexpr size_t kMaxPendingPairs = 4096;

// Begin/end a timed region around one launch when profiling is on.
int prof_begin(owl_acoustic2d* b, int kind, double bytes = 0.0) {
    if (!b->profiling) return -1;
    // keep at most kMaxPendingPairs event pairs in flight: fold the recorded
    // ones into the totals first (profiling runs eagerly, never under capture)
    if (b->ev_used.size() >= kMaxPendingPairs && !b->capturing) {
        OWL_HIP(hipStreamSynchronize(b->stream));
        prof_collect(b);
    }
    if (b->ev_next + 2 > b->ev_pool.size()) {
        for (int k = 0; k < 256; ++k) {
            hipEvent_t e;
            OWL_HIP(hipEventCreate(&e));
            b->ev_pool.push_back(e);
        }
    }
    const int first = int(b->ev_next);
    b->ev_next += 2;
    OWL_HIP(hipEventRecord(b->ev_pool[first], b->stream));
    b->ev_used.push_back({kind, first, bytes});
    return first;
}
void prof_end(owl_acoustic2d* b, int first) {
    if (first < 0) return;
    OWL_HIP(hipEventRecord(b->ev_pool[first + 1], b->stream));
}
// After a stream synchronisation: fold the recorded pairs into the totals.
void prof_collect(owl_acoustic2d* b) {
    for (auto& u : b->ev_used) {
        float ms = 0.f;
        OWL_HIP(hipEventElapsedTime(&ms, b->ev_pool[u.first], b->ev_pool[u.first + 1]));
        b->kern_ms[u.kind] += ms;
        b->kern_launches[u.kind] += 1;
        b->kern_bytes[u.kind] += u.bytes;
    }
    b->ev_used.clear();
    b->ev_next = 0;
}

// Algorithmic (compulsory) bytes of one launch of the update kernels over
// all shots for nsets field sets: each array element read or written once.
//   p-sweep: vx, vz, p read, p written (16 B/point) + bk (4 B per point of
//            each distinct window) + 8 B per slab cell and damped direction
//            (memory variable read + write);
//   v-sweep: p, vx, vz read, vx, vz written (20 B/point) + rho + slabs.
struct StepBytes { double p, v; };
StepBytes step_bytes(const owl_acoustic2d* b, int nsets) {
    double bp = 0, bv = 0;
    const int pml = b->pml;
    for (auto& S : b->sd) {
        const double lx = S.nx + 2.0 * pml, lz = S.nz + 2.0 * pml;
        const double pts = lx * lz;
        const double slab_p = 2.0 * pml * lz + 2.0 * pml * lx;              // i<=0|i>=nx+1 ; j likewise
        const double slab_v = (2.0 * pml + 1) * lz + (2.0 * pml + 1) * lx;  // m<=1|m>=nx+1 ; likewise
        bp += 16.0 * pts + 8.0 * slab_p;
        bv += 20.0 * pts + 8.0 * slab_v;
    }
    return StepBytes{(bp + 4.0 * b->pts_model) * nsets, (bv + 4.0 * b->pts_model) * nsets};
}

template <int R>
void launch_step(owl_acoustic2d* b, const Sets& sets, int nsets, float dt) {
    const int fs = b->cfg.free_surface ? 1 : 0;
    const dim3 blk(kBX, kBY), grd(unsigned(b->tile_shot.size()), unsigned(nsets));
    const StepBytes sb = b->profiling ? step_bytes(b, nsets) : StepBytes{0, 0};
    int ev = prof_begin(b, nsets == 2 ? kKP2 : kKP, sb.p);
    hipLaunchKernelGGL(k_update_p<R>, grd, blk, 0, b->stream, b->d_shots.get(), b->d_tile_shot.get(), sets,
                       b->d_bk.get(), b->d_prof.get(), b->coef, dt, b->idx, b->idz, b->pml, fs);
    prof_end(b, ev);
    if (fs) {
        const dim3 g2(64, unsigned(b->sd.size()), unsigned(nsets));
        hipLaunchKernelGGL(k_mirror_p<R>, g2, dim3(256), 0, b->stream, b->d_shots.get(), sets, b->pml);
    }
    ev = prof_begin(b, nsets == 2 ? kKV2 : kKV, sb.v);
    hipLaunchKernelGGL(k_update_v<R>, grd, blk, 0, b->stream, b->d_shots.get(), b->d_tile_shot.get(), sets,
                       b->d_rho.get(), b->d_prof.get(), b->coef, dt, b->idx, b->idz, b->pml, fs);
    prof_end(b, ev);
}

template <int R>
void run_forward(owl_acoustic2d* b, bool save) {
    hipStream_t st = b->stream;
    b->have_store = false;  // the working sets below hold a stored exact-adjoint state
    b->store_id = 0;
    for (int k = 0; k < 3; ++k) b->fA[k].zero(st);
    for (int k = 0; k < 4; ++k) b->fM[k].zero(st);
    b->d_traces.zero(st);
    if (save) {
        b->d_strips.alloc(b->strip_elems);
        b->d_strips.zero(st);
    }
    float* m[4] = {b->fM[0].get(), b->fM[1].get(), b->fM[2].get(), b->fM[3].get()};
    const Sets sets = one_set(b->fA[0].get(), b->fA[1].get(), b->fA[2].get(), m);
    const unsigned nshot = unsigned(b->sd.size());
    OWL_HIP(hipEventRecord(b->ev0, st));
    // steps 1..nt-1; the update and source of step nt change nothing that
    // OWL records, and skipping them leaves the state of the start of step nt
    // (OWL's output_final_step_wavefield) in place.
    for (int t = 1; t <= b->nt - 1; ++t) {
        launch_step<R>(b, sets, 1, b->dt);
        const int ev = prof_begin(b, kKPost);
        hipLaunchKernelGGL(k_post_step<R>, dim3(nshot), dim3(512), 0, st, b->d_shots.get(),
                           b->fA[0].get(), b->fA[1].get(), b->fA[2].get(), b->d_inj_beg.get(),
                           b->d_inj_taps.get(), b->d_amp.get(), b->nt, t, b->d_rec_beg.get(),
                           b->d_rec_taps.get(), b->d_rec_w.get(), b->d_traces.get(),
                           save ? b->d_strips.get() : nullptr, b->pml);
        prof_end(b, ev);
    }
    OWL_HIP(hipGetLastError());
    OWL_HIP(hipEventRecord(b->ev1, st));
    OWL_HIP(hipStreamSynchronize(st));
    prof_collect(b);
    float ms = 0.f;
    OWL_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_ms = ms;
    b->last_updates = double(b->nt - 1) * b->pts_full;
    b->have_boundary = save;
    b->fwd_state_valid = true;
}

// OWL's reverse loop t = nt .. hnt (submodule_fwi_gradient.f90:75-167).
template <int R>
void run_reverse_owl(owl_acoustic2d* b) {
    hipStream_t st = b->stream;
    b->have_store = false;
    b->store_id = 0;
    if (!b->have_boundary || !b->fwd_state_valid)
        fail(OWL_ERR_STATE, "the adjoint needs a preceding forward run with boundary saving");
    for (int k = 0; k < 3; ++k) {
        b->fB[k].alloc(b->field_elems);
        b->fB[k].zero(st);
        b->aF[k].alloc(b->field_elems);
        b->aF[k].zero(st);
    }
    for (int k = 0; k < 4; ++k) {
        b->fM[k].zero(st);
        b->aM[k].alloc(b->field_elems);
        b->aM[k].zero(st);
    }
    b->d_edge.alloc(std::max<size_t>(b->edge_elems, 1));
    b->d_gv.alloc(b->grad_elems);
    b->d_gr.alloc(b->grad_elems);
    b->d_gv.zero(st);
    b->d_gr.zero(st);
    if (b->energy_precond) {
        b->d_energy.alloc(4 * b->grad_elems);
        b->d_energy.zero(st);
    }
    const unsigned nshot = unsigned(b->sd.size());
    // OWL loops down to the first source's hnt of each shot; a batch whose
    // shots disagree runs down to the smallest one and images each shot only
    // while t >= its own hnt (the extra reverse steps of the other shots
    // touch no gradient).
    int hnt = b->nt;
    for (auto& H : b->hs) hnt = std::min(hnt, std::max(H.hnt, 1));
    DevBuf<float>* A = b->fA;
    DevBuf<float>* B = b->fB;
    const int cc = b->cfg.cc_step_interval > 0 ? b->cfg.cc_step_interval : 1;
    const unsigned ig = unsigned(std::max(1, std::min(1024, (int)((b->grad_elems / nshot + 255) / 256))));
    // imaging: p, prev p, pr, rho*vp^3 and six velocity fields read once
    // (40 B/point), two FP64 accumulators read and written (32 B/point)
    const double image_bytes = 72.0 * double(b->grad_elems);
    OWL_HIP(hipEventRecord(b->ev0, st));
    for (int t = b->nt; t >= hnt; --t) {
        int ev = prof_begin(b, kKStrips);
        hipLaunchKernelGGL(k_inject_strips<R>, dim3(16, nshot), dim3(256), 0, st, b->d_shots.get(),
                           A[1].get(), A[2].get(), b->d_strips.get(), b->d_edge.get(), t, b->pml);
        prof_end(b, ev);
        Sets sets{};
        float* fm[4] = {b->fM[0].get(), b->fM[1].get(), b->fM[2].get(), b->fM[3].get()};
        float* am[4] = {b->aM[0].get(), b->aM[1].get(), b->aM[2].get(), b->aM[3].get()};
        sets.s[0] = in_place(A[0].get(), A[1].get(), A[2].get(), fm);
        sets.s[0].p_out = B[0].get();
        sets.s[0].vx_out = B[1].get();
        sets.s[0].vz_out = B[2].get();
        sets.s[1] = in_place(b->aF[0].get(), b->aF[1].get(), b->aF[2].get(), am);
        launch_step<R>(b, sets, 2, -b->dt);
        ev = prof_begin(b, kKAdj);
        if (b->max_nadj > 0)
            hipLaunchKernelGGL(k_inject_adjoint, dim3(unsigned((b->max_nadj + 255) / 256), nshot), dim3(256), 0,
                               st, b->d_shots.get(), b->aF[0].get(), b->d_adj_beg.get(),
                               b->d_adj_taps.get(), b->d_rec_w.get(), b->d_adj.get(), t);
        prof_end(b, ev);
        ev = prof_begin(b, kKImage, (t % cc == 0) ? image_bytes : 0.0);
        if (t % cc == 0)
            hipLaunchKernelGGL(k_image_owl<R>, dim3(ig, nshot), dim3(256), 0, st, b->d_shots.get(),
                               B[0].get(), B[1].get(), B[2].get(), A[0].get(), A[1].get(), A[2].get(),
                               b->d_edge.get(), b->aF[0].get(), b->aF[1].get(), b->aF[2].get(),
                               b->d_den.get(), b->d_gv.get(), b->d_gr.get(), t == b->nt ? 1 : 0, b->pml,
                               b->energy_precond ? b->d_energy.get() : nullptr, b->grad_elems, t);
        prof_end(b, ev);
        std::swap(A, B);
    }
    if (b->energy_precond)
        hipLaunchKernelGGL(k_energy_precond, dim3(nshot), dim3(256), 0, st, b->d_shots.get(), b->d_gv.get(),
                           b->d_gr.get(), b->d_energy.get(), b->grad_elems);
    OWL_HIP(hipGetLastError());
    OWL_HIP(hipEventRecord(b->ev1, st));
    OWL_HIP(hipStreamSynchronize(st));
    prof_collect(b);
    // the forward state now lives in whichever buffer A points to; the
    // boundary records have been consumed
    if (A != b->fA)
        for (int k = 0; k < 3; ++k) std::swap(b->fA[k], b->fB[k]);
    float ms = 0.f;
    OWL_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_ms = ms;
    b->last_updates = 2.0 * double(b->nt - hnt + 1) * b->pts_full;
    b->fwd_state_valid = false;
}


// ---------------------------------------------------------------------------
// Exact adjoint: host side.
// ---------------------------------------------------------------------------
// A working set: p, vx, vz and the four PML memory arrays.
struct W7 {
    float* f[7];
};

W7 work1(owl_acoustic2d* b) {
    return W7{{b->fA[0].get(), b->fA[1].get(), b->fA[2].get(), b->fM[0].get(), b->fM[1].get(),
               b->fM[2].get(), b->fM[3].get()}};
}
W7 work2(owl_acoustic2d* b) {
    return W7{{b->fB[0].get(), b->fB[1].get(), b->fB[2].get(), b->fM2[0].get(), b->fM2[1].get(),
               b->fM2[2].get(), b->fM2[3].get()}};
}
W7 work3(owl_acoustic2d* b) {
    return W7{{b->fC[0].get(), b->fC[1].get(), b->fC[2].get(), b->fM3[0].get(), b->fM3[1].get(),
               b->fM3[2].get(), b->fM3[3].get()}};
}
// The other half of the w1/w3 ping-pong.
W7 other_of(owl_acoustic2d* b, const W7& w) { return w.f[0] == b->fA[0].get() ? work3(b) : work1(b); }

// Algorithmic bytes of one fused step over all shots: p, vx, vz read and
// written (24 B/point), bk and rho read once per distinct window (shots on
// one window share them), plus 8 B per slab cell and damped direction
// (memory variable read + write) for the p and the v parts.
double fused_bytes(const owl_acoustic2d* b) {
    double by = 0;
    for (auto& S : b->sd) {
        const double lx = S.nx + 2.0 * b->pml, lz = S.nz + 2.0 * b->pml;
        const double slab_p = 2.0 * b->pml * lz + 2.0 * b->pml * lx;
        const double slab_v = (2.0 * b->pml + 1) * lz + (2.0 * b->pml + 1) * lx;
        by += 24.0 * lx * lz + 8.0 * (slab_p + slab_v);
    }
    return by + 8.0 * b->pts_model;  // bk and rho once per distinct window
}

// The fused kernels address fields and model with 32-bit byte offsets
// (FP64 accumulators with twice the offset).
bool offsets32(const owl_acoustic2d* b) {
    return b->field_elems < (size_t(1) << 28) && b->model_elems < (size_t(1) << 28);
}

// The fused adjoint step (k_adj_fused) runs wherever the fused forward step
// may: batches of 4 shots or more by default, no free surface.
bool adj_fused(const owl_acoustic2d* b) {
    if (b->step_kernel == 0 || b->cfg.free_surface || !offsets32(b)) return false;
    return !(b->step_kernel == 2 && b->sd.size() < 4);
}

bool use_fused(const owl_acoustic2d* b, const W7& in, const W7& out) {
    if (b->step_kernel == 0 || b->cfg.free_surface || !offsets32(b)) return false;
    if (b->step_kernel == 2 && b->sd.size() < 4) return false;
    for (int k = 0; k < 7; ++k)
        for (int m = 0; m < 7; ++m)
            if (in.f[k] == out.f[m]) return false;
    return true;
}

// One forward step t from `in` into `out` (in place when they alias):
// sweeps, then source injection of step t (and receivers of step t+1 when
// recording).
template <int R>
void fwd_step(owl_acoustic2d* b, const W7& in, const W7& out, int t, bool record) {
    if (b->skip_gpu) return;
    Sets sets{};
    FieldSet& F = sets.s[0];
    F.p_in = in.f[0];
    F.p_out = out.f[0];
    F.vx_in = in.f[1];
    F.vx_out = out.f[1];
    F.vz_in = in.f[2];
    F.vz_out = out.f[2];
    for (int k = 0; k < 4; ++k) {
        F.m_in[k] = in.f[3 + k];
        F.m_out[k] = out.f[3 + k];
    }
    sets.s[1] = F;
    if (use_fused(b, in, out)) {
        const int ev = prof_begin(b, kKFused, b->profiling ? fused_bytes(b) : 0.0);
        hipLaunchKernelGGL(k_step_fused<R>, dim3(unsigned((b->nftile + 3) / 4)), dim3(64, 4), 0, b->stream,
                           b->d_shots.get(), b->d_ftile_shot.get(), b->nftile, F, b->d_bk.get(), b->d_rho.get(),
                           b->d_prof.get(), b->coef, b->dt, b->idx, b->idz, b->pml, b->fzc);
        prof_end(b, ev);
    } else {
        launch_step<R>(b, sets, 1, b->dt);
    }
    // without receivers to sample and outside every source's active steps the
    // post-step kernel would change nothing
    if (!record && (t < b->src_t0 || t > b->src_t1)) return;
    const int ev = prof_begin(b, kKPost);
    hipLaunchKernelGGL(k_post_step<R>, dim3(unsigned(b->sd.size())), dim3(512), 0, b->stream, b->d_shots.get(),
                       out.f[0], out.f[1], out.f[2], b->d_inj_beg.get(), b->d_inj_taps.get(), b->d_amp.get(),
                       b->nt, t, b->d_rec_beg.get(), b->d_rec_taps.get(), b->d_rec_w.get(),
                       record ? b->d_traces.get() : nullptr, nullptr, b->pml);
    prof_end(b, ev);
}

// Clear device memory on the batch stream: a memset, or a fill kernel while
// a graph is being captured (memset nodes are kept out of graphs).
void dzero(owl_acoustic2d* b, void* p, size_t bytes) {
    if (!p || !bytes || b->skip_gpu) return;
    if (b->capturing) {
        const size_t n = bytes / 4;
        const unsigned g = unsigned(std::min<size_t>(4096, (n + 255) / 256));
        hipLaunchKernelGGL(k_zero_words, dim3(g), dim3(256), 0, b->stream, static_cast<unsigned*>(p), n);
    } else {
        OWL_HIP(hipMemsetAsync(p, 0, bytes, b->stream));
    }
}

void zero_set(owl_acoustic2d* b, const W7& w) {
    for (int k = 0; k < 7; ++k) dzero(b, w.f[k], b->field_elems * sizeof(float));
}

void slots_alloc(owl_acoustic2d* b, owl_acoustic2d::Slots& S, int nslot, int nfield, int fmt) {
    const size_t fe = b->field_elems;
    const size_t n = size_t(nslot) * nfield * fe;
    const size_t bytes = n * (fmt == OWL_STORE_F32 ? 4 : 2);
    const size_t have = (fmt == OWL_STORE_F32 ? S.f32.size() * 4 : S.h16.size() * 2);
    if (bytes > have) {
        size_t freeb = 0, total = 0;
        OWL_HIP(hipMemGetInfo(&freeb, &total));
        if (bytes - have + (size_t(1) << 30) > freeb)
            fail(OWL_ERR_NOMEM, "state storage needs " + std::to_string(bytes >> 20) + " MiB; " +
                                    std::to_string(freeb >> 20) + " MiB free");
    }
    S.fmt = fmt;
    S.nfield = nfield;
    S.nslot = nslot;
    if (fmt == OWL_STORE_F32) {
        S.h16.release();
        if (S.f32.size() < n) S.f32.alloc(n);
    } else {
        S.f32.release();
        if (S.h16.size() < n) S.h16.alloc(n);
        const size_t na = size_t(nslot) * nfield * b->sd.size();
        if (S.amax.size() < na) S.amax.alloc(na);
    }
}

void slot_save(owl_acoustic2d* b, owl_acoustic2d::Slots& S, int slot, float* const* fields) {
    if (b->skip_gpu) return;
    const size_t fe = b->field_elems;
    const unsigned nshot = unsigned(b->sd.size());
    const int ev = prof_begin(b, 12);
    for (int k = 0; k < S.nfield; ++k) {
        const size_t o = (size_t(slot) * S.nfield + k) * fe;
        if (S.fmt == OWL_STORE_F32) {
            OWL_HIP(hipMemcpyAsync(S.f32.get() + o, fields[k], fe * sizeof(float), hipMemcpyDeviceToDevice, b->stream));
        } else {
            unsigned* am = S.amax.get() + (size_t(slot) * S.nfield + k) * nshot;
            hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, b->stream, am, 0u, int(nshot));
            hipLaunchKernelGGL(k_absmax, dim3(64, nshot), dim3(256), 0, b->stream, b->d_shots.get(), fields[k], am,
                               b->d_nonfinite.get());
            hipLaunchKernelGGL(k_pack16, dim3(256, nshot), dim3(256), 0, b->stream, b->d_shots.get(), fields[k],
                               S.h16.get() + o, am, S.fmt);
        }
    }
    prof_end(b, ev);
}

void slot_load(owl_acoustic2d* b, owl_acoustic2d::Slots& S, int slot, float* const* fields) {
    if (b->skip_gpu) return;
    const size_t fe = b->field_elems;
    const unsigned nshot = unsigned(b->sd.size());
    const int ev = prof_begin(b, 12);
    for (int k = 0; k < S.nfield; ++k) {
        const size_t o = (size_t(slot) * S.nfield + k) * fe;
        if (S.fmt == OWL_STORE_F32) {
            OWL_HIP(hipMemcpyAsync(fields[k], S.f32.get() + o, fe * sizeof(float), hipMemcpyDeviceToDevice, b->stream));
        } else {
            const unsigned* am = S.amax.get() + (size_t(slot) * S.nfield + k) * nshot;
            hipLaunchKernelGGL(k_unpack16, dim3(256, nshot), dim3(256), 0, b->stream, b->d_shots.get(),
                               S.h16.get() + o, fields[k], am, S.fmt);
        }
    }
    prof_end(b, ev);
}

// FP32 view of slot `slot` (fields p, vx, vz) when stored in FP32.
const float* slot_ptr(owl_acoustic2d* b, owl_acoustic2d::Slots& S, int slot, int k) {
    return S.f32.get() + (size_t(slot) * S.nfield + k) * b->field_elems;
}

AdjSet adj_set(owl_acoustic2d* b) {
    AdjSet A{};
    float** f = &A.ph;
    for (int k = 0; k < 11; ++k) f[k] = b->adj[k].get();
    A.gK = b->gK.get();
    A.gX = b->gX.get();
    A.gZ = b->gZ.get();
    A.gR = b->gR.get();
    return A;
}

void adjoint_alloc_zero(owl_acoustic2d* b) {
    for (int k = 0; k < 11; ++k) {
        b->adj[k].alloc(b->field_elems);
        dzero(b, b->adj[k].get(), b->field_elems * sizeof(float));
    }
    for (DevBuf<double>* g : {&b->gK, &b->gX, &b->gZ, &b->gR}) {
        g->alloc(b->field_elems);
        dzero(b, g->get(), b->field_elems * sizeof(double));
    }
    if (adj_fused(b))
        for (auto& a : b->adj2) {
            a.alloc(b->field_elems);
            dzero(b, a.get(), b->field_elems * sizeof(float));
        }
    b->adj_cur = 0;
}

// The adjoint pressure currently in use (the fused adjoint alternates it
// between adj[0] and adj2[0]).
float* cur_ph(owl_acoustic2d* b) { return b->adj_cur ? b->adj2[0].get() : b->adj[0].get(); }

void inject_residual(owl_acoustic2d* b, int t) {
    const unsigned nshot = unsigned(b->sd.size());
    if (b->max_nadj <= 0) return;
    const int ev = prof_begin(b, kKAdj);
    hipLaunchKernelGGL(k_inject_adjoint, dim3(unsigned((b->max_nadj + 255) / 256), nshot), dim3(256), 0, b->stream,
                       b->d_shots.get(), cur_ph(b), b->d_adj_beg.get(), b->d_adj_taps.get(),
                       b->d_rec_w.get(), b->d_adj.get(), t);
    prof_end(b, ev);
}

// Reverse step t of the exact adjoint, given u_t (cur) and u_{t-1} (prev).
template <int R>
void adjoint_step(owl_acoustic2d* b, int t, const float* const* cur, const float* const* prev) {
    const dim3 blk(kBX, kBY), grd(unsigned(b->tile_shot.size()));
    const AdjSet A = adj_set(b);
    double pts = 0;
    for (auto& S : b->sd) pts += double(S.nx + 2 * b->pml) * double(S.nz + 2 * b->pml);
    if (adj_fused(b)) {
        DevBuf<float>* a = b->adj;
        DevBuf<float>* a2 = b->adj2;
        float* s0[5] = {a[0].get(), a[3].get(), a[4].get(), a[7].get(), a[8].get()};
        float* s1[5] = {a2[0].get(), a2[1].get(), a2[2].get(), a2[3].get(), a2[4].get()};
        float** in = b->adj_cur ? s1 : s0;
        float** out = b->adj_cur ? s0 : s1;
        const AdjIO io{in[0], in[1], in[2], in[3], in[4], out[0], out[1], out[2], out[3], out[4]};
        // ex, ez, ph, p_t, p_{t-1}, vxh, vzh, 4 states read; ph, vxh, vzh, ex, ez
        // written; gK, gR r+w (96 B per point, PML memories aside); gX on a
        // tile's first column and gZ on a chunk's first row r+w; bk and rho
        // once per distinct window
        const double wcol = 64 - (2 * R - 1);
        const int ev = prof_begin(b, 14, pts * (96.0 + 16.0 / wcol + 16.0 / b->azc) + 8.0 * b->pts_model);
        hipLaunchKernelGGL(k_adj_fused<R>, dim3(unsigned((b->natile + 3) / 4)), dim3(64, 4), 0, b->stream,
                           b->d_shots.get(), b->d_atile_shot.get(), b->natile, io, A, b->d_bk.get(), b->d_rho.get(),
                           b->d_prof.get(), b->coef, b->dt, b->idx, b->idz, b->pml, b->azc, cur[0], prev[0], cur[1],
                           prev[1], cur[2], prev[2]);
        prof_end(b, ev);
        b->adj_cur = 1 - b->adj_cur;
        if (b->max_ninj > 0 || b->max_nadj > 0) {
            const int ev2 = prof_begin(b, kKAdj);
            hipLaunchKernelGGL(k_adjoint_taps, dim3(unsigned(b->sd.size())), dim3(512), 0, b->stream,
                               b->d_shots.get(), b->d_inj_beg.get(), b->d_inj_taps.get(), b->d_amp.get(),
                               b->d_bk.get(), cur_ph(b), b->gK.get(), b->nt, b->d_adj_beg.get(),
                               b->d_adj_taps.get(), b->d_rec_w.get(), b->d_adj.get(), t);
            prof_end(b, ev2);
        }
        return;
    }
    // p-pass: ex, ez, ph, bk, p_t, p_{t-1}, psx/psz read; ph, dx, dz written; gK r+w
    int ev = prof_begin(b, 10, pts * (5 * 4 + 3 * 4 + 16) + 4.0 * b->pts_model);
    hipLaunchKernelGGL(k_adj_p<R>, grd, blk, 0, b->stream, b->d_shots.get(), b->d_tile_shot.get(), A, b->d_bk.get(),
                       b->d_prof.get(), b->coef, b->dt, b->idx, b->idz, b->pml, cur[0], prev[0]);
    prof_end(b, ev);
    if (b->max_ninj > 0)
        hipLaunchKernelGGL(k_src_correction, dim3(unsigned(b->sd.size())), dim3(256), 0, b->stream, b->d_shots.get(),
                           b->d_inj_beg.get(), b->d_inj_taps.get(), b->d_amp.get(), b->d_bk.get(), b->adj[0].get(),
                           b->gK.get(), b->nt, t);
    // v-pass: dx, dz, vxh, vzh, rho, 4 states read; vxh, vzh, ex, ez written; gX, gZ r+w
    ev = prof_begin(b, 11, pts * (8 * 4 + 4 * 4 + 32) + 4.0 * b->pts_model);
    hipLaunchKernelGGL(k_adj_v<R>, grd, blk, 0, b->stream, b->d_shots.get(), b->d_tile_shot.get(), A, b->d_rho.get(),
                       b->d_prof.get(), b->coef, b->dt, b->idx, b->idz, b->pml, cur[1], prev[1], cur[2], prev[2]);
    prof_end(b, ev);
    inject_residual(b, t);
}

// Optimal binomial split tables.  cost[s][l] = fewest forward steps to run
// the adjoint over l steps from a stored u_a with s free snapshots;
// split[s][l] = steps to advance before the next snapshot.  The adjoint
// step t reads u_t and u_{t-1}.
//  * carry = false: each adjoint step computes its own u_t from u_{t-1}:
//    cost[s][1] = 1, cost[0][l] = l(l+1)/2.
//  * carry = true: u_t is kept from the adjoint step before, where it was
//    the older state, so only u_{t-1} is recomputed: cost[s][1] = 0,
//    cost[0][l] = l(l-1)/2, plus one step for u_N.  This needs the two
//    states to come out of one exact trajectory, so it is used with FP32
//    snapshots only: with 16-bit ones u_t and u_{t-1} could be recomputed
//    from different rounded snapshots, and the imaging terms use their
//    difference.
// Both: cost[s][l] = min_j j + cost[s-1][l-j] + cost[s][j].
void binomial_table(owl_acoustic2d* b, int N, int S, bool carry) {
    if (b->binom_n == N && b->binom_s == S && b->binom_carry == carry) return;
    std::vector<std::vector<long long>> cost(S + 1, std::vector<long long>(N + 1, 0));
    b->binom_split.assign(S + 1, std::vector<int>(N + 1, 1));
    for (int l = 1; l <= N; ++l) cost[0][l] = carry ? (long long)l * (l - 1) / 2 : (long long)l * (l + 1) / 2;
    for (int s = 1; s <= S; ++s) {
        cost[s][1] = carry ? 0 : 1;
        int jlo = 1;
        for (int l = 2; l <= N; ++l) {
            long long best = -1;
            int bj = 1;
            // the optimal split is non-decreasing in l; search from the last one
            for (int j = std::max(1, jlo - 2); j < l; ++j) {
                const long long c = j + cost[s - 1][l - j] + cost[s][j];
                if (best < 0 || c < best) { best = c; bj = j; }
                else if (c > best + (long long)l) break;
            }
            cost[s][l] = best;
            b->binom_split[s][l] = bj;
            jlo = bj;
        }
    }
    b->binom_n = N;
    b->binom_s = S;
    b->binom_carry = carry;
    b->binom_steps = cost[S][N] + (carry ? 1 : 0);
}

struct RevCtx {
    bool first = true;           // still on the first forward descent (recording)
    bool direct = false;         // FP32 snapshots: steps read and write the slots directly
    bool carry = false;          // u_t of each adjoint step is kept from the step before
    bool carried = false;        // `cur` holds u_t of the next adjoint step
    int nlog = 1;                // snapshots of the schedule
    std::vector<int> phys;       // snapshot -> physical slot
    int spare = -1;              // the physical slot outside `phys` (carry mode)
    W7 cur{};                    // u_t of the next adjoint step (p, vx, vz)
    int cur_phys = -1;           // its physical slot, or -1
    std::function<void()> on_first_adjoint;
};

// Where a forward state is: kW1 / kW3 = the working sets w1 / w3 (the
// steps ping-pong between them), kZero = the zero initial state
// (materialised by clearing w1), >= 0 = snapshot.
constexpr int kW1 = -2, kZero = -1, kW3 = -3;

W7 slot_set(owl_acoustic2d* b, int phys) {
    W7 w;
    for (int k = 0; k < 7; ++k) w.f[k] = const_cast<float*>(slot_ptr(b, b->snaps, phys, k));
    return w;
}

// The two carried-state sets (p, vx, vz; memory variables not kept).
W7 carry_set(owl_acoustic2d* b, int k) {
    W7 w{};
    for (int i = 0; i < 3; ++i) w.f[i] = b->fK[3 * k + i].get();
    return w;
}

// The physical slot a store into snapshot `s` writes.  When that slot still
// holds the carried state, the spare slot takes its place in the map.
int store_slot(RevCtx& ctx, int s) {
    if (ctx.phys[s] == ctx.cur_phys) std::swap(ctx.phys[s], ctx.spare);
    return ctx.phys[s];
}

// The state at `loc` as a readable set (clearing w1 for the zero state).
W7 state_at(owl_acoustic2d* b, int loc, const RevCtx& ctx) {
    if (loc >= 0) return slot_set(b, ctx.phys[loc]);
    if (loc == kW3) return work3(b);
    const W7 w1 = work1(b);
    if (loc == kZero) zero_set(b, w1);
    return w1;
}

int loc_of(owl_acoustic2d* b, const W7& w) { return w.f[0] == b->fA[0].get() ? kW1 : kW3; }

// Advance the state u_from (at `loc`) to u_to, every step out of place
// (w1 <-> w3).  In direct mode the first step reads the snapshot in place
// and, when `into` >= 0, the last step writes snapshot `into` directly;
// otherwise the state is restored into w1 and saved by copy.  Returns
// where u_to is.
template <int R>
int advance(owl_acoustic2d* b, int from, int to, int loc, int into, bool record, RevCtx& ctx) {
    if (!ctx.direct) {
        W7 cur = work1(b);
        if (loc >= 0) slot_load(b, b->snaps, ctx.phys[loc], cur.f);
        else if (loc == kZero) zero_set(b, cur);
        else if (loc == kW3) cur = work3(b);
        for (int t = from + 1; t <= to; ++t) {
            const W7 out = other_of(b, cur);
            fwd_step<R>(b, cur, out, t, record);
            cur = out;
        }
        if (into >= 0) slot_save(b, b->snaps, store_slot(ctx, into), cur.f);
        return loc_of(b, cur);
    }
    if (to == from) {
        if (into >= 0 && into != loc) {
            const W7 src = state_at(b, loc, ctx);
            slot_save(b, b->snaps, store_slot(ctx, into), src.f);
            return into;
        }
        return loc;
    }
    W7 cur = state_at(b, loc, ctx);
    bool in_slot = loc >= 0;
    for (int t = from + 1; t <= to; ++t) {
        W7 out;
        if (t == to && into >= 0) out = slot_set(b, store_slot(ctx, into));
        else out = in_slot ? work1(b) : other_of(b, cur);
        fwd_step<R>(b, cur, out, t, record);
        cur = out;
        in_slot = false;
    }
    return into >= 0 ? into : loc_of(b, cur);
}

// One forward step from the full state `in` to u_t, writing p, vx, vz into
// `dst` and the memory variables into the working set `in` does not use.
template <int R>
W7 step_into(owl_acoustic2d* b, const W7& in, const W7& dst, int t, bool record) {
    W7 out = in.f[3] == b->fM[0].get() ? work3(b) : work1(b);
    for (int i = 0; i < 3; ++i) out.f[i] = dst.f[i];
    fwd_step<R>(b, in, out, t, record);
    return out;
}

// Reverse sweep over steps b..a+1: u_a is stored at `slot` (snapshot, or
// kZero) and currently at `loc`; sfree free snapshots.
template <int R>
void rev(owl_acoustic2d* bt, int a, int b, int sfree, int slot, int loc, RevCtx& ctx) {
    // carry = false: u_{t-1} recomputed from u_a, then u_t by one step into w2
    auto last_step = [&](int t, int at) {
        const W7 w2 = work2(bt);
        const W7 prev = state_at(bt, at, ctx);
        fwd_step<R>(bt, prev, w2, t, ctx.first);
        if (ctx.first) {
            ctx.first = false;
            ctx.on_first_adjoint();
        }
        adjoint_step<R>(bt, t, w2.f, prev.f);
    };
    // carry = true: u_{t-1} recomputed from u_a (at `from`) into the carried
    // set t-1 (or read in place when it is u_a itself); u_t is the state the
    // adjoint step t+1 read as its older one
    auto reverse_step = [&](int t, int from) {
        int prev_phys = -1;
        W7 prev;
        if (t - 1 > a) {
            const int at = advance<R>(bt, a, t - 2, from, -1, ctx.first, ctx);
            prev = step_into<R>(bt, state_at(bt, at, ctx), carry_set(bt, (t - 1) & 1), t - 1, ctx.first);
        } else if (from >= 0) {
            prev_phys = ctx.phys[from];
            prev = slot_set(bt, prev_phys);
        } else if (from == kZero) {
            prev = state_at(bt, kZero, ctx);  // u_0; t = 1 is the last adjoint step
        } else {
            // a working set: copy the fields out of the next recomputation's way
            const W7 w = state_at(bt, from, ctx);
            prev = w;
            const W7 k = carry_set(bt, (t - 1) & 1);
            for (int i = 0; i < 3; ++i) {
                if (!bt->skip_gpu)
                    OWL_HIP(hipMemcpyAsync(k.f[i], w.f[i], bt->field_elems * sizeof(float),
                                           hipMemcpyDeviceToDevice, bt->stream));
                prev.f[i] = k.f[i];
            }
        }
        if (!ctx.carried) {  // the first adjoint step: u_t from u_{t-1}
            ctx.cur = step_into<R>(bt, prev, carry_set(bt, t & 1), t, ctx.first);
            ctx.cur_phys = -1;
            ctx.carried = true;
            if (ctx.first) {
                ctx.first = false;
                ctx.on_first_adjoint();
            }
        }
        adjoint_step<R>(bt, t, ctx.cur.f, prev.f);
        ctx.cur = prev;
        ctx.cur_phys = prev_phys;
    };
    if (b == a + 1) {
        if (ctx.carry) reverse_step(b, loc);
        else last_step(b, ctx.direct ? loc : advance<R>(bt, a, a, loc, -1, ctx.first, ctx));
        return;
    }
    if (sfree == 0) {
        for (int t = b; t >= a + 1; --t) {
            if (ctx.carry) {
                reverse_step(t, t == b ? loc : slot);
            } else {
                const int at = advance<R>(bt, a, t - 1, t == b ? loc : slot, -1, ctx.first, ctx);
                last_step(t, at);
            }
        }
        return;
    }
    const int j = bt->binom_split[sfree][b - a];
    const int mine = ctx.nlog - sfree;
    const int at = advance<R>(bt, a, a + j, loc, mine, ctx.first, ctx);
    rev<R>(bt, a + j, b, sfree - 1, mine, at, ctx);
    rev<R>(bt, a, a + j, sfree, slot, slot, ctx);
}

// Phases of an exact-adjoint call: forward and reverse, or one of them.
enum ExactPhase { kBoth = 0, kForwardOnly = 1, kReverseOnly = 2 };

// The stream work of one exact-adjoint call (everything between the timing
// events): the storage schedule, forward and adjoint steps, the residual,
// the fold onto the windows.  Issues no synchronisation, so it can be
// captured into a graph.
template <int R>
void exact_body(owl_acoustic2d* b, int ckpt, int param, int fmt, const std::function<void()>& residual, int phase) {
    hipStream_t st = b->stream;
    const int N = b->nt - 1;  // forward steps
    const W7 w1 = work1(b), w2 = work2(b);
    const bool do_fwd = phase != kReverseOnly, do_rev = phase != kForwardOnly;
    if (fmt != OWL_STORE_F32)
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, b->d_nonfinite.get(), 0u, 1);
    if (do_fwd) {
        zero_set(b, w1);
        zero_set(b, w2);
        dzero(b, b->d_traces.get(), b->d_traces.size() * sizeof(float));
    }
    // w3 is a step target: its halo must read as zero, its memory variables
    // start from zero wherever a schedule first writes them (a resumed
    // binomial schedule finds it as its forward call left it)
    const bool resume = phase == kReverseOnly && ckpt == OWL_CKPT_BINOMIAL;
    if (!resume) zero_set(b, work3(b));
    double fwd_steps = 0, adj_steps = 0;
    if (ckpt == OWL_CKPT_FULL) {
        if (do_fwd) {
            slots_alloc(b, b->full, N + 1, 3, fmt);
            if (fmt == OWL_STORE_F32) {
                // each step reads slot t-1 and writes slot t directly (memory
                // variables alternate between w1 and w3); the slots' outer
                // halo stays zero
                dzero(b, b->full.f32.get(), size_t(N + 1) * 3 * b->field_elems * sizeof(float));
                W7 mA = w1, mB = work3(b);
                for (int t = 1; t <= N; ++t) {
                    W7 in = mA, out = mB;
                    for (int k = 0; k < 3; ++k) {
                        in.f[k] = const_cast<float*>(slot_ptr(b, b->full, t - 1, k));
                        out.f[k] = const_cast<float*>(slot_ptr(b, b->full, t, k));
                    }
                    fwd_step<R>(b, in, out, t, true);
                    std::swap(mA, mB);
                }
            } else {
                slot_save(b, b->full, 0, w1.f);
                W7 cur = w1;
                for (int t = 1; t <= N; ++t) {
                    const W7 out = other_of(b, cur);
                    fwd_step<R>(b, cur, out, t, true);
                    slot_save(b, b->full, t, out.f);
                    cur = out;
                }
            }
            fwd_steps += N;
        }
        if (!do_rev) return;
        residual();
        adjoint_alloc_zero(b);
        inject_residual(b, b->nt);
        // FP32 slots are used in place; 16-bit slots are decoded into the two
        // working sets, alternating roles
        float* A3[3] = {w1.f[0], w1.f[1], w1.f[2]};
        float* B3[3] = {w2.f[0], w2.f[1], w2.f[2]};
        if (fmt != OWL_STORE_F32) slot_load(b, b->full, N, A3);
        for (int t = N; t >= 1; --t) {
            if (fmt == OWL_STORE_F32) {
                const float* cur[3] = {slot_ptr(b, b->full, t, 0), slot_ptr(b, b->full, t, 1), slot_ptr(b, b->full, t, 2)};
                const float* prv[3] = {slot_ptr(b, b->full, t - 1, 0), slot_ptr(b, b->full, t - 1, 1), slot_ptr(b, b->full, t - 1, 2)};
                adjoint_step<R>(b, t, cur, prv);
            } else {
                slot_load(b, b->full, t - 1, B3);
                adjoint_step<R>(b, t, A3, B3);
                for (int k = 0; k < 3; ++k) std::swap(A3[k], B3[k]);
            }
        }
        adj_steps += N;
    } else if (ckpt == OWL_CKPT_UNIFORM) {
        const int k = std::max(1, std::min(param > 0 ? param : 1, N));
        const int nseg = (N + k - 1) / k;
        if (do_fwd) {
            slots_alloc(b, b->snaps, nseg, 7, fmt);
            W7 cur = w1;
            for (int t = 0; t < N; ++t) {
                if (t % k == 0) slot_save(b, b->snaps, t / k, cur.f);
                const W7 out = other_of(b, cur);
                fwd_step<R>(b, cur, out, t + 1, true);
                cur = out;
            }
            fwd_steps += N;
        }
        if (!do_rev) return;
        slots_alloc(b, b->seg, k + 1, 3, OWL_STORE_F32);
        // the recompute writes each slot out of place over the update domain
        // only; the R-wide outer halo must read as zero, as in the forward
        dzero(b, b->seg.f32.get(), size_t(k + 1) * 3 * b->field_elems * sizeof(float));
        residual();
        adjoint_alloc_zero(b);
        inject_residual(b, b->nt);
        for (int sgi = nseg - 1; sgi >= 0; --sgi) {
            const int a = sgi * k, e = std::min(a + k, N);
            // restore u_a: p, vx, vz into segment slot 0, memory into w1
            float* into[7] = {const_cast<float*>(slot_ptr(b, b->seg, 0, 0)), const_cast<float*>(slot_ptr(b, b->seg, 0, 1)),
                              const_cast<float*>(slot_ptr(b, b->seg, 0, 2)), w1.f[3], w1.f[4], w1.f[5], w1.f[6]};
            slot_load(b, b->snaps, sgi, into);
            // memory variables alternate between w1 and w3
            W7 mA = w1, mB = work3(b);
            for (int t = a + 1; t <= e; ++t) {
                W7 in{{const_cast<float*>(slot_ptr(b, b->seg, t - 1 - a, 0)), const_cast<float*>(slot_ptr(b, b->seg, t - 1 - a, 1)),
                       const_cast<float*>(slot_ptr(b, b->seg, t - 1 - a, 2)), mA.f[3], mA.f[4], mA.f[5], mA.f[6]}};
                W7 out{{const_cast<float*>(slot_ptr(b, b->seg, t - a, 0)), const_cast<float*>(slot_ptr(b, b->seg, t - a, 1)),
                        const_cast<float*>(slot_ptr(b, b->seg, t - a, 2)), mB.f[3], mB.f[4], mB.f[5], mB.f[6]}};
                fwd_step<R>(b, in, out, t, false);
                std::swap(mA, mB);
            }
            fwd_steps += e - a;
            for (int t = e; t >= a + 1; --t) {
                const float* cur[3] = {slot_ptr(b, b->seg, t - a, 0), slot_ptr(b, b->seg, t - a, 1), slot_ptr(b, b->seg, t - a, 2)};
                const float* prv[3] = {slot_ptr(b, b->seg, t - 1 - a, 0), slot_ptr(b, b->seg, t - 1 - a, 1), slot_ptr(b, b->seg, t - 1 - a, 2)};
                adjoint_step<R>(b, t, cur, prv);
            }
        }
        adj_steps += N;
    } else if (ckpt == OWL_CKPT_BINOMIAL) {
        const int S = std::max(0, std::min(param, N));
        RevCtx ctx;
        ctx.first = true;
        ctx.direct = fmt == OWL_STORE_F32;
        ctx.carry = ctx.direct;
        binomial_table(b, N, S, ctx.carry);
        // carry mode: one physical slot more, for a carried state whose
        // snapshot is stored over before the adjoint step reads it
        ctx.nlog = std::max(S, 1);
        const int nphys = ctx.nlog + (ctx.carry ? 1 : 0);
        ctx.phys.resize(size_t(ctx.nlog));
        for (int k = 0; k < ctx.nlog; ++k) ctx.phys[size_t(k)] = k;
        ctx.spare = ctx.carry ? ctx.nlog : -1;
        slots_alloc(b, b->snaps, nphys, 7, fmt);
        // steps write FP32 snapshots and the carried sets over the update
        // domain only: their outer halo must read as zero
        if (do_fwd) {
            if (ctx.direct) dzero(b, b->snaps.f32.get(), size_t(nphys) * 7 * b->field_elems * sizeof(float));
            if (ctx.carry)
                for (auto& k : b->fK) dzero(b, k.get(), b->field_elems * sizeof(float));
        }
        // The schedule's first descent models the traces and records the
        // checkpoints.  A forward-only call stops where the first adjoint
        // step would start; the reverse-only call replays the same descent
        // on the host without GPU work (its checkpoints, carried states and
        // working sets are still in the buffers) and goes on from there, so
        // the split call does the work of the joint one.
        struct StopAtFirstAdjoint {};
        if (phase == kForwardOnly) {
            ctx.on_first_adjoint = [] { throw StopAtFirstAdjoint{}; };
            try {
                rev<R>(b, 0, N, S, kZero, kW1, ctx);
            } catch (const StopAtFirstAdjoint&) {
            }
            return;
        }
        ctx.on_first_adjoint = [&] {
            b->skip_gpu = false;
            residual();
            adjoint_alloc_zero(b);
            inject_residual(b, b->nt);
        };
        b->skip_gpu = !do_fwd;
        try {
            rev<R>(b, 0, N, S, kZero, kW1, ctx);
        } catch (...) {
            b->skip_gpu = false;
            throw;
        }
        b->skip_gpu = false;
        adj_steps += N;
        fwd_steps += 0;  // counted by the schedule; reported through the updates figure below
    } else {
        fail(OWL_ERR_ARG, "unknown checkpoint policy");
    }
    {
    // fold the padded-grid gradients onto the windows and merge
    const unsigned nshot = unsigned(b->sd.size());
    b->d_gv.alloc(b->grad_elems);
    b->d_gr.alloc(b->grad_elems);
    const unsigned fg = unsigned(std::max<size_t>(1, std::min<size_t>(1024, (b->grad_elems / nshot + 255) / 256)));
    hipLaunchKernelGGL(k_fold_exact<R>, dim3(fg, nshot), dim3(256), 0, st, b->d_shots.get(), b->gK.get(), b->gX.get(),
                       b->gZ.get(), b->gR.get(), int(adj_fused(b)), b->d_vpw.get(), b->d_rhow.get(), b->d_gv.get(),
                       b->d_gr.get(), b->pml);
    }
    (void)fwd_steps;
    (void)adj_steps;
}

// Autotuner of the fused kernels' z-chunk rows: times each candidate over
// three launches (after one warm-up) on the batch's own buffers and keeps
// the fastest for the forward step and for the adjoint step.  Results do
// not depend on the choice.  The buffers it writes (work3, the adjoint
// set) are cleared again before a gradient uses them.
template <int R>
void autotune_rows(owl_acoustic2d* b) {
    hipStream_t st = b->stream;
    adjoint_alloc_zero(b);
    const W7 in = work1(b), out = work3(b), w2 = work2(b);
    FieldSet F{};
    F.p_in = in.f[0];
    F.p_out = out.f[0];
    F.vx_in = in.f[1];
    F.vx_out = out.f[1];
    F.vz_in = in.f[2];
    F.vz_out = out.f[2];
    for (int k = 0; k < 4; ++k) {
        F.m_in[k] = in.f[3 + k];
        F.m_out[k] = out.f[3 + k];
    }
    DevBuf<float>* a = b->adj;
    DevBuf<float>* a2 = b->adj2;
    const AdjIO io{a[0].get(), a[3].get(), a[4].get(), a[7].get(), a[8].get(),
                   a2[0].get(), a2[1].get(), a2[2].get(), a2[3].get(), a2[4].get()};
    const AdjSet A = adj_set(b);
    auto timed = [&](auto&& launch) {
        launch();
        OWL_HIP(hipEventRecord(b->ev0, st));
        for (int k = 0; k < 3; ++k) launch();
        OWL_HIP(hipEventRecord(b->ev1, st));
        OWL_HIP(hipEventSynchronize(b->ev1));
        float ms = 0.f;
        OWL_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        return ms;
    };
    const int fwd_rows[] = {8, 12, 16, 24, 32}, adj_rows[] = {4, 6, 8, 12, 16};
    int best_f = b->fzc, best_a = b->azc;
    float tf = 1e30f, ta = 1e30f;
    for (int z : fwd_rows) {
        b->fzc = z;
        upload_fused_tiling(b);
        const float ms = timed([&] {
            hipLaunchKernelGGL(k_step_fused<R>, dim3(unsigned((b->nftile + 3) / 4)), dim3(64, 4), 0, st,
                               b->d_shots.get(), b->d_ftile_shot.get(), b->nftile, F, b->d_bk.get(), b->d_rho.get(),
                               b->d_prof.get(), b->coef, b->dt, b->idx, b->idz, b->pml, b->fzc);
        });
        if (ms < tf) tf = ms, best_f = z;
    }
    b->fzc = best_f;
    for (int z : adj_rows) {
        b->azc = z;
        upload_fused_tiling(b);
        const float ms = timed([&] {
            hipLaunchKernelGGL(k_adj_fused<R>, dim3(unsigned((b->natile + 3) / 4)), dim3(64, 4), 0, st,
                               b->d_shots.get(), b->d_atile_shot.get(), b->natile, io, A, b->d_bk.get(),
                               b->d_rho.get(), b->d_prof.get(), b->coef, b->dt, b->idx, b->idz, b->pml, b->azc,
                               in.f[0], w2.f[0], in.f[1], w2.f[1], in.f[2], w2.f[2]);
        });
        if (ms < ta) ta = ms, best_a = z;
    }
    b->azc = best_a;
    upload_fused_tiling(b);
    OWL_HIP(hipGetLastError());
    b->tuned = true;
}

// Every device pointer and launch parameter an exact-adjoint call bakes
// into its launches: a graph captured under one signature is replayed only
// under the same one.
std::vector<uintptr_t> exact_signature(owl_acoustic2d* b, int ckpt, int param, int fmt) {
    std::vector<uintptr_t> v{uintptr_t(ckpt), uintptr_t(param), uintptr_t(fmt), uintptr_t(b->R), uintptr_t(b->nt),
                             uintptr_t(b->step_kernel), uintptr_t(b->fzc), uintptr_t(b->nftile),
                             uintptr_t(b->azc), uintptr_t(b->natile),
                             uintptr_t(b->tile_shot.size()), uintptr_t(b->field_elems)};
    auto add = [&](const void* p) { v.push_back(reinterpret_cast<uintptr_t>(p)); };
    for (int k = 0; k < 3; ++k) { add(b->fA[k].get()); add(b->fB[k].get()); add(b->fC[k].get()); }
    for (int k = 0; k < 4; ++k) { add(b->fM[k].get()); add(b->fM2[k].get()); add(b->fM3[k].get()); }
    for (auto& a : b->adj) add(a.get());
    for (auto& a : b->adj2) add(a.get());
    for (auto& k : b->fK) add(k.get());
    for (auto* g : {&b->gK, &b->gX, &b->gZ, &b->gR, &b->d_gv, &b->d_gr, &b->d_misfit, &b->d_scratch}) add(g->get());
    for (auto* S : {&b->full, &b->snaps, &b->seg}) { add(S->f32.get()); add(S->h16.get()); add(S->amax.get()); }
    for (auto* f : {&b->d_traces, &b->d_obs, &b->d_adj, &b->d_bk, &b->d_rho, &b->d_prof, &b->d_amp, &b->d_rec_w,
                    &b->d_vpw, &b->d_rhow})
        add(f->get());
    add(b->d_shots.get()); add(b->d_tile_shot.get()); add(b->d_ftile_shot.get()); add(b->d_atile_shot.get());
    add(b->d_inj_beg.get()); add(b->d_adj_beg.get()); add(b->d_rec_beg.get());
    add(b->d_inj_taps.get()); add(b->d_adj_taps.get()); add(b->d_rec_taps.get());
    return v;
}

// Exact-adjoint gradient over all shots.  residual(): fills d_adj once all
// traces are known (misfit adjoint source, or the caller's trace gradient).
// A full gradient call (both phases, profiling off) runs eagerly the first
// time; the next call with the same schedule and buffers is captured into a
// HIP graph, which later calls replay.
template <int R>
void run_exact(owl_acoustic2d* b, int ckpt, int param, int fmt, const std::function<void()>& residual,
               int phase = kBoth) {
    hipStream_t st = b->stream;
    const int N = b->nt - 1;  // forward steps
    for (int k = 0; k < 3; ++k) b->fB[k].alloc(b->field_elems);
    for (int k = 0; k < 4; ++k) b->fM2[k].alloc(b->field_elems);
    for (int k = 0; k < 3; ++k) b->fC[k].alloc(b->field_elems);
    for (int k = 0; k < 4; ++k) b->fM3[k].alloc(b->field_elems);
    if (ckpt == OWL_CKPT_BINOMIAL && fmt == OWL_STORE_F32)
        for (auto& k : b->fK) k.alloc(b->field_elems);
    b->d_nonfinite.alloc(1);
    const bool do_fwd = phase != kReverseOnly, do_rev = phase != kForwardOnly;
    if (do_rev && !do_fwd &&
        (!b->have_store || b->store_policy != ckpt || b->store_param != param || b->store_fmt != fmt))
        fail(OWL_ERR_STATE, "backward needs a preceding forward_store with the same storage policy");
    if (b->autotune && !b->tuned && adj_fused(b)) autotune_rows<R>(b);
    const bool graph_ok = phase == kBoth && b->use_graphs && !b->profiling;
    std::vector<uintptr_t> sig;
    if (graph_ok) sig = exact_signature(b, ckpt, param, fmt);
    OWL_HIP(hipEventRecord(b->ev0, st));
    if (graph_ok && b->gexec && sig == b->gsig) {
        OWL_HIP(hipGraphLaunch(b->gexec, st));
    } else if (graph_ok && sig == b->gready) {
        if (b->gexec) {
            (void)hipGraphExecDestroy(b->gexec);
            b->gexec = nullptr;
            b->gsig.clear();
        }
        hipGraph_t g = nullptr;
        b->capturing = true;
        OWL_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            exact_body<R>(b, ckpt, param, fmt, residual, phase);
        } catch (...) {
            (void)hipStreamEndCapture(st, &g);
            if (g) (void)hipGraphDestroy(g);
            b->capturing = false;
            throw;
        }
        const hipError_t e = hipStreamEndCapture(st, &g);
        b->capturing = false;
        if (e != hipSuccess || !g) fail(OWL_ERR_DEVICE, "graph capture of the exact gradient failed");
        const hipError_t ei = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            b->gexec = nullptr;
            fail(OWL_ERR_DEVICE, "graph instantiation of the exact gradient failed");
        }
        b->gsig = sig;
        OWL_HIP(hipGraphLaunch(b->gexec, st));
    } else {
        exact_body<R>(b, ckpt, param, fmt, residual, phase);
        if (graph_ok) b->gready = exact_signature(b, ckpt, param, fmt);
    }
    OWL_HIP(hipGetLastError());
    OWL_HIP(hipEventRecord(b->ev1, st));
    OWL_HIP(hipStreamSynchronize(st));
    prof_collect(b);
    float ms = 0.f;
    OWL_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_ms = ms;
    // grid-point updates of the useful work: N forward steps + N adjoint steps
    b->last_updates = double(N) * (double(do_fwd) + double(do_rev)) * b->pts_full;
    b->fwd_state_valid = false;
    b->have_boundary = false;
    b->have_store = false;
    b->store_id = 0;
    if (fmt != OWL_STORE_F32) {
        unsigned nonfinite = 0;
        OWL_HIP(hipMemcpy(&nonfinite, b->d_nonfinite.get(), sizeof(unsigned), hipMemcpyDeviceToHost));
        if (nonfinite)
            fail(OWL_ERR_NONFINITE, "the wavefield became NaN or infinite (the run diverged): 16-bit snapshots "
                                    "cannot be scaled");
    }
    if (phase == kForwardOnly) {
        b->have_store = true;
        b->store_policy = ckpt;
        b->store_param = param;
        b->store_fmt = fmt;
        b->store_id = ++b->store_next;
    }
}

void merge_full(owl_acoustic2d* b) {
    hipStream_t st = b->stream;
    b->d_gfull_v.alloc(size_t(b->cfg.nx) * b->cfg.nz);
    b->d_gfull_r.alloc(size_t(b->cfg.nx) * b->cfg.nz);
    const unsigned mg = unsigned(std::min<size_t>(4096, (size_t(b->cfg.nx) * b->cfg.nz + 255) / 256));
    const int ev = prof_begin(b, kKMerge);
    hipLaunchKernelGGL(k_merge, dim3(mg), dim3(256), 0, st, b->d_shots.get(), int(b->sd.size()), b->d_gv.get(),
                       b->d_gfull_v.get(), b->cfg.nx, b->cfg.nz);
    hipLaunchKernelGGL(k_merge, dim3(mg), dim3(256), 0, st, b->d_shots.get(), int(b->sd.size()), b->d_gr.get(),
                       b->d_gfull_r.get(), b->cfg.nx, b->cfg.nz);
    prof_end(b, ev);
}

template <class F>
void dispatch_R(int R, F&& f) {
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

// Device [shot][t][nr] <-> host [shot][nr][nt]
void traces_to_host(const owl_acoustic2d* b, const std::vector<float>& dev, float* out) {
    for (size_t s = 0; s < b->sd.size(); ++s) {
        const Shot& S = b->sd[s];
        for (int r = 0; r < S.nr; ++r)
            for (int t = 0; t < b->nt; ++t)
                out[size_t(S.trace0) + size_t(r) * b->nt + t] = dev[size_t(S.trace0) + size_t(t) * S.nr + r];
    }
}
void traces_to_dev(const owl_acoustic2d* b, const float* in, std::vector<float>& dev) {
    dev.assign(b->trace_elems, 0.f);
    for (size_t s = 0; s < b->sd.size(); ++s) {
        const Shot& S = b->sd[s];
        for (int r = 0; r < S.nr; ++r)
            for (int t = 0; t < b->nt; ++t)
                dev[size_t(S.trace0) + size_t(t) * S.nr + r] = in[size_t(S.trace0) + size_t(r) * b->nt + t];
    }
}

// The model from device memory (owl_acoustic2d_set_model_device): refill
// the padded moduli and the imaging arrays on the device, check positivity
// and the CFL bound per shot as plan() does, rebuild the PML profiles of the
// shots whose edge velocity moved, keep the host copy of the model current,
// and drop every state that belonged to the old model.
void set_model_device(owl_acoustic2d* b, const float* vp, const float* rho, int xf) {
    hipStream_t st = b->stream;
    const unsigned nshot = unsigned(b->sd.size());
    b->d_mstat.alloc(3 * size_t(nshot));
    b->d_mstat.zero(st);
    long long mx = 1;
    for (const Shot& S : b->sd)
        mx = std::max(mx, (long long)(S.nx + 2 * (b->pml + b->R)) * (S.nz + 2 * (b->pml + b->R)));
    const dim3 grd(unsigned(std::min<long long>(1024, (mx + 255) / 256)), nshot);
    hipLaunchKernelGGL(k_model_from_device, grd, dim3(256), 0, st, b->d_shots.get(), vp, rho, b->cfg.nx, b->cfg.nz,
                       xf, b->pml, b->R, b->d_bk.get(), b->d_rho.get(), b->d_den.get(), b->d_vpw.get(),
                       b->d_rhow.get(), b->d_mstat.get());
    OWL_HIP(hipGetLastError());
    std::vector<unsigned> stat(3 * size_t(nshot));
    OWL_HIP(hipMemcpyAsync(stat.data(), b->d_mstat.get(), stat.size() * sizeof(unsigned), hipMemcpyDeviceToHost, st));
    // the host copy of the model follows the device one
    const size_t nfull = size_t(b->cfg.nx) * b->cfg.nz;
    std::vector<float> hv(nfull), hr(nfull);
    OWL_HIP(hipMemcpyAsync(hv.data(), vp, nfull * sizeof(float), hipMemcpyDeviceToHost, st));
    OWL_HIP(hipMemcpyAsync(hr.data(), rho, nfull * sizeof(float), hipMemcpyDeviceToHost, st));
    OWL_HIP(hipStreamSynchronize(st));
    b->have_boundary = false;
    b->fwd_state_valid = false;
    b->have_store = false;
    b->store_id = 0;
    for (size_t e = 0; e < nfull; ++e) {
        const size_t ix = e / size_t(b->cfg.nz), kz = e % size_t(b->cfg.nz);
        const size_t m = xf ? ix + kz * size_t(b->cfg.nx) : e;
        b->vp_full[e] = hv[m];
        b->rho_full[e] = hr[m];
    }
    bool moved = false;
    for (size_t s = 0; s < nshot; ++s) {
        if (stat[3 * s + 2]) fail(OWL_ERR_ARG, "vp and rho must be positive");
        float vmax, vedge;
        std::memcpy(&vmax, &stat[3 * s], sizeof(float));
        std::memcpy(&vedge, &stat[3 * s + 1], sizeof(float));
        const float dts = dt_stable(b->R, 2, vmax, b->cfg.dx, 1.0f, b->cfg.dz);
        if (b->cfg.dt > dts)
            fail(OWL_ERR_CFL, "dt = " + std::to_string(b->cfg.dt) + " exceeds the stable dt = " + std::to_string(dts));
        ShotHost& H = b->hs[s];
        if (vedge != H.pmlvp) {
            H.pmlvp = vedge;
            const Shot& S = b->sd[s];
            const PmlAxis px = cfspml_axis(H.nx, b->pml, b->cfg.dx, b->cfg.dt, vedge, H.f0max);
            const PmlAxis pz = cfspml_axis(H.nz, b->pml, b->cfg.dz, b->cfg.dt, vedge, H.f0max);
            size_t o = size_t(S.prof);
            for (const auto* v : {&px.ai, &px.bi, &px.ah, &px.bh, &pz.ai, &pz.bi, &pz.ah, &pz.bh}) {
                std::copy(v->begin(), v->end(), b->h_prof.begin() + o);
                o += v->size();
            }
            moved = true;
        }
    }
    if (moved) {
        b->d_prof.upload(b->h_prof, st);
        OWL_HIP(hipStreamSynchronize(st));
    }
}

void download_traces(owl_acoustic2d* b, float* out) {
    std::vector<float> h(b->trace_elems);
    OWL_HIP(hipMemcpyAsync(h.data(), b->d_traces.get(), h.size() * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    OWL_HIP(hipStreamSynchronize(b->stream));
    traces_to_host(b, h, out);
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

owl_status owl_acoustic2d_create(const owl_acoustic_config* cfg, const float* vp, const float* rho,
                                 int nshot, const owl_shot* shots, owl_acoustic2d** out) {
    return guarded([&] {
        if (!out) fail(OWL_ERR_ARG, "out is null");
        *out = nullptr;
        auto* b = new owl_acoustic2d();
        try {
            plan(b, cfg, vp, rho, nshot, shots);
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

void owl_acoustic2d_destroy(owl_acoustic2d* b) { delete b; }

int owl_acoustic2d_nt(const owl_acoustic2d* b) { return b ? b->nt : 0; }

owl_status owl_acoustic2d_set_model(owl_acoustic2d* b, const float* vp, const float* rho) {
    return guarded([&] {
        if (!b || !vp || !rho) fail(OWL_ERR_ARG, "null argument");
        std::vector<owl_shot> shots(b->hs.size());
        for (size_t s = 0; s < b->hs.size(); ++s) {
            const ShotHost& H = b->hs[s];
            shots[s] = owl_shot{H.id, int(H.src.size()), H.src.data(), int(H.rec.size()), H.rec.data(),
                                H.nxbeg, H.nxbeg + H.nx - 1, 0, 0, H.nzbeg, H.nzbeg + H.nz - 1};
        }
        std::vector<ShotHost> keep = b->hs;  // shots reference these vectors
        for (size_t s = 0; s < keep.size(); ++s) {
            shots[s].sources = keep[s].src.data();
            shots[s].receivers = keep[s].rec.data();
        }
        const owl_acoustic_config cfg = b->cfg;
        plan(b, &cfg, vp, rho, int(shots.size()), shots.data());
        b->have_boundary = false;
        b->fwd_state_valid = false;
        // stored forward states belong to the old model
        b->have_store = false;
        b->store_id = 0;
    });
}

owl_status owl_acoustic2d_store_id(const owl_acoustic2d* b, long long* id) {
    return guarded([&] {
        if (!b || !id) fail(OWL_ERR_ARG, "null argument");
        *id = b->have_store ? b->store_id : 0;
    });
}

owl_status owl_acoustic2d_forward(owl_acoustic2d* b, float* traces, int save_boundary) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        dispatch_R(b->R, [&](auto r) { run_forward<decltype(r)::value>(b, save_boundary != 0); });
        if (traces) download_traces(b, traces);
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Forward modelling that allocates no device memory.  Every device buffer of
// the call (fields, PML memory, model, tables, traces) is carved from the
// file-scope array below, so a call fits if its plan needs at most
// kStaticScratchBytes.  Each GPU holds its own copy of the array until the
// process ends; its device address is looked up once per device.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kStaticScratchBytes = size_t(32) << 20;
constexpr int kMaxDevices = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {
struct StaticContext {
    std::once_flag once;
    hipError_t init_rc = hipSuccess;
    void* ptr = nullptr;             // device address of this GPU's scratch
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::mutex mu;                   // one call at a time per GPU in this process
};
StaticContext g_static[kMaxDevices];

StaticContext& static_context() {
    int dev = 0;
    OWL_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) fail(OWL_ERR_DEVICE, "device index out of range");
    StaticContext& C = g_static[dev];
    std::call_once(C.once, [&] {
        // with this device current: its copy of the array, and a stream and
        // two events the calls share
        C.init_rc = hipGetSymbolAddress(&C.ptr, HIP_SYMBOL(scratch));
        if (C.init_rc == hipSuccess) C.init_rc = hipStreamCreateWithFlags(&C.stream, hipStreamNonBlocking);
        if (C.init_rc == hipSuccess) C.init_rc = hipEventCreate(&C.ev0);
        if (C.init_rc == hipSuccess) C.init_rc = hipEventCreate(&C.ev1);
    });
    OWL_HIP(C.init_rc);
    if (!C.ptr) fail(OWL_ERR_DEVICE, "no device address for the static scratch");
    return C;
}
}  // namespace

extern "C" {

owl_status owl_acoustic2d_forward_static(const owl_acoustic_config* cfg, const float* vp, const float* rho,
                                         int nshot, const owl_shot* shots, float* traces) {
    return guarded([&] {
        if (!traces) fail(OWL_ERR_ARG, "traces is null");
        StaticContext& C = static_context();
        std::lock_guard<std::mutex> lock(C.mu);
        DeviceArena arena;
        arena.base = static_cast<char*>(C.ptr);
        arena.size = kStaticScratchBytes;
        DeviceArenaScope scope(&arena);
        owl_acoustic2d b;
        b.stream = C.stream;
        b.ev0 = C.ev0;
        b.ev1 = C.ev1;
        b.borrowed_stream = true;
        b.use_graphs = 0;
        plan(&b, cfg, vp, rho, nshot, shots);
        dispatch_R(b.R, [&](auto r) { run_forward<decltype(r)::value>(&b, false); });
        download_traces(&b, traces);
    });
}

owl_status owl_acoustic2d_adjoint(owl_acoustic2d* b, const float* adj, float* grad_vp, float* grad_rho) {
    return guarded([&] {
        if (!b || !adj) fail(OWL_ERR_ARG, "null argument");
        std::vector<float> h;
        traces_to_dev(b, adj, h);
        b->d_adj.upload(h, b->stream);
        dispatch_R(b->R, [&](auto r) { run_reverse_owl<decltype(r)::value>(b); });
        std::vector<double> gv(b->grad_elems), gr(b->grad_elems);
        OWL_HIP(hipMemcpy(gv.data(), b->d_gv.get(), gv.size() * sizeof(double), hipMemcpyDeviceToHost));
        OWL_HIP(hipMemcpy(gr.data(), b->d_gr.get(), gr.size() * sizeof(double), hipMemcpyDeviceToHost));
        // .grd arrays: transpose to (nz, nx) z-fast; NaN/Inf -> 0 (return_normal)
        for (size_t s = 0; s < b->sd.size(); ++s) {
            const Shot& S = b->sd[s];
            for (int i = 0; i < S.nx; ++i)
                for (int j = 0; j < S.nz; ++j) {
                    const size_t src = size_t(S.g0) + size_t(i) + size_t(j) * S.nx;
                    const size_t dst = size_t(S.g0) + size_t(j) + size_t(i) * S.nz;
                    const float a = float(gv[src]), c = float(gr[src]);
                    if (grad_vp) grad_vp[dst] = std::isfinite(a) ? a : 0.f;
                    if (grad_rho) grad_rho[dst] = std::isfinite(c) ? c : 0.f;
                }
        }
    });
}

owl_status owl_acoustic2d_gradient(owl_acoustic2d* b, const float* obs, int mode, int ckpt, int ckpt_param,
                                   int store, double* misfit, double* grad_vp, double* grad_rho,
                                   float* traces_out) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        if (mode != OWL_GRAD_OWL && mode != OWL_GRAD_EXACT) fail(OWL_ERR_ARG, "unknown gradient mode");
        if (mode == OWL_GRAD_EXACT) {
            if (b->cfg.free_surface) fail(OWL_ERR_ARG, "the exact adjoint does not cover the free surface");
            for (auto& H : b->hs)
                for (auto& sc : H.src)
                    if (sc.mechanism != OWL_EXPLOSION)
                        fail(OWL_ERR_ARG, "the exact adjoint covers explosion sources only");
            if (store != OWL_STORE_F32 && store != OWL_STORE_F16 && store != OWL_STORE_BF16)
                fail(OWL_ERR_ARG, "unknown snapshot storage type");
        }
        hipStream_t st = b->stream;
        if (obs) {
            std::vector<float> h;
            traces_to_dev(b, obs, h);
            b->d_obs.upload(h, st);
            b->have_obs = true;
        }
        if (!b->have_obs) fail(OWL_ERR_STATE, "no observed data: pass obs or call owl_acoustic2d_set_observed");
        b->d_adj.alloc(b->trace_elems);
        b->d_misfit.alloc(b->sd.size());
        b->d_scratch.alloc(2 * size_t(std::max(1, b->nrec_total)));
        auto adjoint_source = [&] {
            const int ev = prof_begin(b, kKMisfit);
            hipLaunchKernelGGL(k_adjoint_source, dim3(unsigned(b->sd.size())), dim3(512), 0, st, b->d_shots.get(),
                               b->d_traces.get(), b->d_obs.get(), b->d_rec_w.get(), b->d_adj.get(),
                               b->d_misfit.get(), b->nt, b->cfg.trace_discard_threshold, b->d_scratch.get());
            prof_end(b, ev);
        };
        double total_ms = 0, total_up = 0;
        if (mode == OWL_GRAD_EXACT) {
            dispatch_R(b->R, [&](auto r) {
                run_exact<decltype(r)::value>(b, ckpt, ckpt_param, store, adjoint_source, kBoth);
            });
            total_ms += b->last_ms;
            total_up += b->last_updates;
        } else {
            dispatch_R(b->R, [&](auto r) { run_forward<decltype(r)::value>(b, true); });
            total_ms += b->last_ms;
            total_up += b->last_updates;
            OWL_HIP(hipEventRecord(b->ev0, st));
            adjoint_source();
            OWL_HIP(hipEventRecord(b->ev1, st));
            OWL_HIP(hipStreamSynchronize(st));
            prof_collect(b);
            float ms = 0;
            OWL_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
            total_ms += ms;
            dispatch_R(b->R, [&](auto r) { run_reverse_owl<decltype(r)::value>(b); });
            total_ms += b->last_ms;
            total_up += b->last_updates;
        }
        OWL_HIP(hipEventRecord(b->ev0, st));
        merge_full(b);
        OWL_HIP(hipEventRecord(b->ev1, st));
        OWL_HIP(hipStreamSynchronize(st));
        prof_collect(b);
        float ms = 0;
        OWL_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        total_ms += ms;
        b->last_ms = total_ms;
        b->last_updates = total_up;
        const size_t nfull = size_t(b->cfg.nx) * b->cfg.nz;
        if (grad_vp) OWL_HIP(hipMemcpy(grad_vp, b->d_gfull_v.get(), nfull * sizeof(double), hipMemcpyDeviceToHost));
        if (grad_rho) OWL_HIP(hipMemcpy(grad_rho, b->d_gfull_r.get(), nfull * sizeof(double), hipMemcpyDeviceToHost));
        if (misfit) OWL_HIP(hipMemcpy(misfit, b->d_misfit.get(), b->sd.size() * sizeof(double), hipMemcpyDeviceToHost));
        if (traces_out) download_traces(b, traces_out);
    });
}

owl_status owl_acoustic2d_set_energy_precond(owl_acoustic2d* b, int on) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        b->energy_precond = on != 0;
    });
}

owl_status owl_acoustic2d_misfit(owl_acoustic2d* b, const float* obs, double* misfit, float* traces_out) {
    return guarded([&] {
        if (!b || !misfit) fail(OWL_ERR_ARG, "null argument");
        hipStream_t st = b->stream;
        if (obs) {
            std::vector<float> h;
            traces_to_dev(b, obs, h);
            b->d_obs.upload(h, st);
            b->have_obs = true;
        }
        if (!b->have_obs) fail(OWL_ERR_STATE, "no observed data: pass obs or call owl_acoustic2d_set_observed");
        dispatch_R(b->R, [&](auto r) { run_forward<decltype(r)::value>(b, false); });
        b->d_adj.alloc(b->trace_elems);
        b->d_misfit.alloc(b->sd.size());
        b->d_scratch.alloc(2 * size_t(std::max(1, b->nrec_total)));
        hipLaunchKernelGGL(k_adjoint_source, dim3(unsigned(b->sd.size())), dim3(512), 0, st, b->d_shots.get(),
                           b->d_traces.get(), b->d_obs.get(), b->d_rec_w.get(), b->d_adj.get(), b->d_misfit.get(),
                           b->nt, b->cfg.trace_discard_threshold, b->d_scratch.get());
        OWL_HIP(hipGetLastError());
        OWL_HIP(hipMemcpyAsync(misfit, b->d_misfit.get(), b->sd.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        OWL_HIP(hipStreamSynchronize(st));
        if (traces_out) download_traces(b, traces_out);
    });
}

owl_status owl_acoustic2d_forward_store(owl_acoustic2d* b, float* traces, int ckpt, int ckpt_param, int store) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        if (b->cfg.free_surface) fail(OWL_ERR_ARG, "the exact adjoint does not cover the free surface");
        for (auto& H : b->hs)
            for (auto& sc : H.src)
                if (sc.mechanism != OWL_EXPLOSION) fail(OWL_ERR_ARG, "the exact adjoint covers explosion sources only");
        dispatch_R(b->R, [&](auto r) {
            run_exact<decltype(r)::value>(b, ckpt, ckpt_param, store, [] {}, kForwardOnly);
        });
        if (traces) download_traces(b, traces);
    });
}

owl_status owl_acoustic2d_backward(owl_acoustic2d* b, const float* dtraces, double* grad_vp, double* grad_rho) {
    return guarded([&] {
        if (!b || !dtraces) fail(OWL_ERR_ARG, "null argument");
        if (!b->have_store) fail(OWL_ERR_STATE, "backward needs a preceding owl_acoustic2d_forward_store");
        std::vector<float> h;
        traces_to_dev(b, dtraces, h);
        b->d_adj.upload(h, b->stream);
        dispatch_R(b->R, [&](auto r) {
            run_exact<decltype(r)::value>(b, b->store_policy, b->store_param, b->store_fmt, [] {}, kReverseOnly);
        });
        merge_full(b);
        OWL_HIP(hipStreamSynchronize(b->stream));
        const size_t nfull = size_t(b->cfg.nx) * b->cfg.nz;
        if (grad_vp) OWL_HIP(hipMemcpy(grad_vp, b->d_gfull_v.get(), nfull * sizeof(double), hipMemcpyDeviceToHost));
        if (grad_rho) OWL_HIP(hipMemcpy(grad_rho, b->d_gfull_r.get(), nfull * sizeof(double), hipMemcpyDeviceToHost));
    });
}

owl_status owl_acoustic2d_set_model_device(owl_acoustic2d* b, const float* vp, const float* rho, int x_fastest) {
    return guarded([&] {
        if (!b || !vp || !rho) fail(OWL_ERR_ARG, "null argument");
        for (auto& H : b->hs)
            for (auto& sc : H.src)
                if (sc.mechanism != OWL_EXPLOSION)
                    fail(OWL_ERR_ARG, "device model updates cover explosion sources only (a force's amplitude "
                                      "depends on rho)");
        set_model_device(b, vp, rho, x_fastest != 0);
    });
}

owl_status owl_acoustic2d_forward_store_device(owl_acoustic2d* b, float* traces, int ckpt, int ckpt_param,
                                               int store) {
    return guarded([&] {
        if (!b || !traces) fail(OWL_ERR_ARG, "null argument");
        if (b->cfg.free_surface) fail(OWL_ERR_ARG, "the exact adjoint does not cover the free surface");
        for (auto& H : b->hs)
            for (auto& sc : H.src)
                if (sc.mechanism != OWL_EXPLOSION) fail(OWL_ERR_ARG, "the exact adjoint covers explosion sources only");
        dispatch_R(b->R, [&](auto r) {
            run_exact<decltype(r)::value>(b, ckpt, ckpt_param, store, [] {}, kForwardOnly);
        });
        int mnr = 1;
        for (const Shot& S : b->sd) mnr = std::max(mnr, S.nr);
        const unsigned g = unsigned(std::min<long long>(256, (mnr * (long long)b->nt + 255) / 256));
        hipLaunchKernelGGL(k_traces_layout, dim3(std::max(g, 1u), unsigned(b->sd.size())), dim3(256), 0, b->stream,
                           b->d_shots.get(), b->d_traces.get(), traces, b->nt, 1);
        OWL_HIP(hipGetLastError());
        OWL_HIP(hipStreamSynchronize(b->stream));
    });
}

owl_status owl_acoustic2d_backward_device(owl_acoustic2d* b, const float* dtraces, float* grad_vp, float* grad_rho,
                                          int x_fastest) {
    return guarded([&] {
        if (!b || !dtraces) fail(OWL_ERR_ARG, "null argument");
        if (!b->have_store) fail(OWL_ERR_STATE, "backward needs a preceding owl_acoustic2d_forward_store");
        hipStream_t st = b->stream;
        b->d_adj.alloc(std::max<size_t>(b->trace_elems, 1));
        int mnr = 1;
        for (const Shot& S : b->sd) mnr = std::max(mnr, S.nr);
        const unsigned g = unsigned(std::min<long long>(256, (mnr * (long long)b->nt + 255) / 256));
        hipLaunchKernelGGL(k_traces_layout, dim3(std::max(g, 1u), unsigned(b->sd.size())), dim3(256), 0, st,
                           b->d_shots.get(), dtraces, b->d_adj.get(), b->nt, 0);
        dispatch_R(b->R, [&](auto r) {
            run_exact<decltype(r)::value>(b, b->store_policy, b->store_param, b->store_fmt, [] {}, kReverseOnly);
        });
        merge_full(b);
        const size_t nfull = size_t(b->cfg.nx) * b->cfg.nz;
        const unsigned mg = unsigned(std::min<size_t>(4096, (nfull + 255) / 256));
        if (grad_vp)
            hipLaunchKernelGGL(k_grad_to_float, dim3(mg), dim3(256), 0, st, b->d_gfull_v.get(), grad_vp, b->cfg.nx,
                               b->cfg.nz, x_fastest != 0 ? 1 : 0);
        if (grad_rho)
            hipLaunchKernelGGL(k_grad_to_float, dim3(mg), dim3(256), 0, st, b->d_gfull_r.get(), grad_rho, b->cfg.nx,
                               b->cfg.nz, x_fastest != 0 ? 1 : 0);
        OWL_HIP(hipGetLastError());
        OWL_HIP(hipStreamSynchronize(st));
    });
}

owl_status owl_acoustic2d_shot_dims(const owl_acoustic2d* b, int s, int* nx, int* nz, int* nr) {
    return guarded([&] {
        if (!b || s < 0 || s >= int(b->sd.size())) fail(OWL_ERR_ARG, "bad shot index");
        if (nx) *nx = b->sd[s].nx;
        if (nz) *nz = b->sd[s].nz;
        if (nr) *nr = b->sd[s].nr;
    });
}

owl_status owl_acoustic2d_set_observed(owl_acoustic2d* b, const float* obs) {
    return guarded([&] {
        if (!b || !obs) fail(OWL_ERR_ARG, "null argument");
        std::vector<float> h;
        traces_to_dev(b, obs, h);
        b->d_obs.upload(h, b->stream);
        OWL_HIP(hipStreamSynchronize(b->stream));
        b->have_obs = true;
    });
}

owl_status owl_acoustic2d_observe(owl_acoustic2d* b) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        dispatch_R(b->R, [&](auto r) { run_forward<decltype(r)::value>(b, false); });
        b->d_obs.alloc(b->trace_elems);
        OWL_HIP(hipMemcpyAsync(b->d_obs.get(), b->d_traces.get(), b->trace_elems * sizeof(float),
                               hipMemcpyDeviceToDevice, b->stream));
        OWL_HIP(hipStreamSynchronize(b->stream));
        b->have_obs = true;
        b->fwd_state_valid = false;
    });
}

owl_status owl_acoustic2d_set_graphs(owl_acoustic2d* b, int on) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        b->use_graphs = on != 0;
        if (b->gexec) {
            OWL_HIP(hipGraphExecDestroy(b->gexec));
            b->gexec = nullptr;
        }
        b->gsig.clear();
        b->gready.clear();
    });
}

owl_status owl_acoustic2d_set_step_kernel(owl_acoustic2d* b, int mode, int zrows) {
    return guarded([&] {
        if (!b || mode < 0 || mode > 2 || zrows < 1) fail(OWL_ERR_ARG, "bad step kernel choice");
        b->step_kernel = mode;
        b->autotune = false;
        if (zrows != b->fzc) {
            b->fzc = zrows;
            upload_fused_tiling(b);
        }
    });
}

owl_status owl_acoustic2d_set_adjoint_rows(owl_acoustic2d* b, int zrows) {
    return guarded([&] {
        if (!b || zrows < 1) fail(OWL_ERR_ARG, "bad adjoint row count");
        b->autotune = false;
        if (zrows != b->azc) {
            b->azc = zrows;
            upload_fused_tiling(b);
        }
    });
}

owl_status owl_acoustic2d_set_autotune(owl_acoustic2d* b, int on) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        b->autotune = on != 0;
        b->tuned = false;
    });
}

owl_status owl_acoustic2d_fused_rows(const owl_acoustic2d* b, int* zrows, int* adjoint_rows) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        if (zrows) *zrows = b->fzc;
        if (adjoint_rows) *adjoint_rows = b->azc;
    });
}

owl_status owl_acoustic2d_set_profiling(owl_acoustic2d* b, int on) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        b->profiling = on != 0;
        for (int k = 0; k < 16; ++k) {
            b->kern_ms[k] = 0;
            b->kern_launches[k] = 0;
            b->kern_bytes[k] = 0;
        }
    });
}

owl_status owl_acoustic2d_kernel_stats(const owl_acoustic2d* b, int kind, double* total_ms,
                                       long long* launches, double* bytes_total) {
    return guarded([&] {
        if (!b || kind < 0 || kind > 15) fail(OWL_ERR_ARG, "bad kernel kind");
        if (total_ms) *total_ms = b->kern_ms[kind];
        if (launches) *launches = b->kern_launches[kind];
        if (bytes_total) *bytes_total = b->kern_bytes[kind];
    });
}

owl_status owl_acoustic2d_last_stats(const owl_acoustic2d* b, double* ms, double* updates) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        if (ms) *ms = b->last_ms;
        if (updates) *updates = b->last_updates;
    });
}

}  // extern "C"
