// 3D acoustic SSG-FD time stepping for MI355X (gfx950): OWL's
// acoustic_iso_3d solver (acoustic_iso_3d/module_wavefield.f90:40-444,
// submodule_forward_modeling.f90:31-183) and OWL's gradient
// (submodule_fwi_gradient.f90:33-361) for a batch of shots, each on its own
// window, in every launch.
//
// Layout: per shot a padded grid x fastest, then y, then z; rows padded to a
// multiple of 64 floats and started on a float4 boundary.  The sweeps tile
// x-y and march along z, keeping the z neighbours in a per-thread register
// queue of 2R values so that each plane is loaded once per column.  Variants
// differ in how x and y neighbours are shared (LDS plane, L1/L2, float4
// columns).  FP32 in the reference's operation order (no FMA contraction).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "host_prep.hpp"
#include "owl_amd.h"
#include "profiler.hpp"

namespace owl {
namespace a3 {

constexpr int kBX = 64;          // threads along x
constexpr int kBY = 4;           // thread rows
constexpr int kRY = 4;           // y rows per thread
constexpr int kTY = kBY * kRY;   // y extent of a tile
constexpr int kPitchAlign = 64;

constexpr int kXcd = 8;          // XCDs of an MI355X (L2 domains)

struct Coef {
    float c[8];
};

struct Shot3 {
    int nx, ny, nz;        // window interior
    int px;                // row pitch (floats)
    int nyp, nzp;          // padded rows per plane, padded planes
    long long pxy;         // plane pitch (floats)
    long long off;         // element offset of node (lo, lo, lo)
    int tile0, ntx, nty, nzc;
    int prof;              // PML profiles: x (ai bi ah bh), y, z
    int nr, rec0;          // receivers
    int ninj, inj0;        // source cells
    long long trace0;      // [nt][nr] block
    int nxbeg, nybeg, nzbeg;
    int xb;                // x index stored at row position 0 (<= lo; 16-byte aligned interior)
    int nadj, adj0;        // adjoint-source cells (CSR over receiver taps)
    long long strip0;      // boundary records of the shot ([nt][strip length])
    long long edge0;       // saved interior-face velocities
    long long g0;          // window gradient (nx*ny*nz, x fastest)
    int ftile0, fntx, fnty, fnzc;  // tiles of the fused step
    int hnt;               // last reverse step OWL images for this shot (its first source's hnt)
};

// Wavefields of one sweep.  The p-sweep reads vx, vy, vz (stencils) and p
// (own node) and writes po; the v-sweep reads p (the new pressure, stencils)
// and vx, vy, vz (own node) and writes vxo, vyo, vzo.  In place when the
// pointers coincide; out of place (ping-pong) otherwise.  No sweep reads a
// node that another thread of the same sweep writes.  Memory variables m[0..2]
// for the p-sweep (dx vx, dy vy, dz vz on integer nodes), m[3..5] for the
// v-sweep (dx p, dy p, dz p on half nodes): full-size, updated in place and
// touched only in the slabs.
struct F3 {
    const float* p;
    float* po;
    const float* vx;
    float* vxo;
    const float* vy;
    float* vyo;
    const float* vz;
    float* vzo;
    float* m[6];
};

__host__ __device__ inline F3 f3_in_place(float* p, float* vx, float* vy, float* vz, float* const* m) {
    F3 F{p, p, vx, vx, vy, vy, vz, vz, {m[0], m[1], m[2], m[3], m[4], m[5]}};
    return F;
}

// First plane a sweep updates: with a free surface at z index 1 the p-sweep
// starts at k = 2 and the v-sweep at k = 1 (vz at k = 2)
// (update_wavefield_free_surface, module_wavefield.f90:205-317).
__device__ __forceinline__ bool zslab_p(int k, int nz, int fs) { return fs ? (k >= nz + 1) : (k <= 0 || k >= nz + 1); }
__device__ __forceinline__ bool zslab_v(int k, int nz, int fs) { return fs ? (k >= nz + 1) : (k <= 1 || k >= nz + 1); }

struct Tap3 {
    long long cell;
    float w1, w2, w3;
    int row;        // amplitude row (source) or receiver index
    int tbeg, tend; // active steps of a source tap
    int field;      // 0 p, 1 vx, 2 vy, 3 vz (sources)
};

// Hardware hands consecutive workgroups to the XCDs round-robin; remap so
// that each XCD walks a contiguous range of tiles (neighbouring tiles share
// halos through that XCD's L2).  Returns -1 for the padding workgroups.
__device__ __forceinline__ int xcd_tile(int ntiles, int remap) {
    if (!remap) return int(blockIdx.x) < ntiles ? int(blockIdx.x) : -1;
    const int per = (ntiles + kXcd - 1) / kXcd;
    const int t = int(blockIdx.x % kXcd) * per + int(blockIdx.x / kXcd);
    return t < ntiles ? t : -1;
}

__device__ __forceinline__ size_t at3(const Shot3& S, int lo, int i, int j, int k) {
    return size_t(S.off) + size_t(i - S.xb) + size_t(j - lo) * size_t(S.px) + size_t(k - lo) * size_t(S.pxy);
}

// ---------------------------------------------------------------------------
// p-sweep (module_wavefield.f90:66-102): p -= dt*bk*(dx+ vx + dy+ vy + dz+ vz)
// with CFS-PML memory updates in the slabs.
// ---------------------------------------------------------------------------
template <int R>
__global__ __launch_bounds__(kBX * kBY) void k3_update_p(const Shot3* __restrict__ shots,
                                                         const int* __restrict__ tile_shot, F3 F,
                                                         const float* __restrict__ bk,
                                                         const float* __restrict__ prof, Coef c,
                                                         float dt, float idx, float idy, float idz,
                                                         int pml, int kc, int fs) {
    const int tile = blockIdx.x;
    const Shot3 S = shots[tile_shot[tile]];
    int lt = tile - S.tile0;
    const int tx = lt % S.ntx;
    lt /= S.ntx;
    const int ty = lt % S.nty, tzc = lt / S.nty;
    const int nx = S.nx, ny = S.ny, nz = S.nz;
    const int lo = 1 - pml - R, hx = nx + pml + R, hy = ny + pml + R;
    const int i0 = -pml + 1 + tx * kBX, j0 = -pml + 1 + ty * kTY;
    const int kt = -pml + 1 + tzc * kc, k1 = min(kt + kc - 1, nz + pml);
    const int k0 = fs ? max(kt, 2) : kt;
    const int lane = threadIdx.x, tid = threadIdx.y * kBX + threadIdx.x;
    const int i = i0 + lane;
    __shared__ float svx[kTY][kBX + 2 * R];
    __shared__ float svy[kTY + 2 * R][kBX];
    const int lx = nx + 2 * pml, ly = ny + 2 * pml, lz = nz + 2 * pml;
    const float* ax_i = prof + S.prof;
    const float* bx_i = ax_i + lx;
    const float* ay_i = prof + S.prof + 4 * lx;
    const float* by_i = ay_i + ly;
    const float* az_i = prof + S.prof + 4 * lx + 4 * ly;
    const float* bz_i = az_i + lz;
    bool ok[kRY];
    float q[kRY][2 * R];
#pragma unroll
    for (int r = 0; r < kRY; ++r) {
        const int j = j0 + threadIdx.y + kBY * r;
        ok[r] = (i <= nx + pml) && (j <= ny + pml);
#pragma unroll
        for (int m = 0; m < 2 * R - 1; ++m) q[r][m] = ok[r] ? F.vz[at3(S, lo, i, j, k0 - R + 1 + m)] : 0.0f;
    }
    const bool xslab = (i <= 0 || i >= nx + 1);
    const float axv = xslab && i <= nx + pml ? ax_i[i + pml - 1] : 0.0f;
    const float bxv = xslab && i <= nx + pml ? bx_i[i + pml - 1] : 0.0f;
    for (int k = k0; k <= k1; ++k) {
#pragma unroll
        for (int r = 0; r < kRY; ++r) {
            const int j = j0 + threadIdx.y + kBY * r;
            q[r][2 * R - 1] = ok[r] ? F.vz[at3(S, lo, i, j, k + R)] : 0.0f;
        }
        __syncthreads();
        // vx with its x halo: columns x = i0-R+1 .. i0+63+R at col = x - (i0-R+1)
#pragma unroll
        for (int r = 0; r < kRY; ++r) {
            const int row = threadIdx.y + kBY * r, j = j0 + row;
            svx[row][lane + R - 1] = (i <= hx && j <= hy) ? F.vx[at3(S, lo, i, j, k)] : 0.0f;
        }
        if (tid < (2 * R - 1) * kTY) {
            const int hc = tid % (2 * R - 1), row = tid / (2 * R - 1);
            const int col = hc < R - 1 ? hc : kBX + hc;
            const int xg = i0 - R + 1 + col, j = j0 + row;
            svx[row][col] = (xg >= lo && xg <= hx && j <= hy) ? F.vx[at3(S, lo, xg, j, k)] : 0.0f;
        }
        // vy with its y halo: rows y = j0-R+1 .. j0+15+R at row = y - (j0-R+1)
        for (int e = tid; e < (kTY + 2 * R - 1) * kBX; e += kBX * kBY) {
            const int row = e / kBX, col = e % kBX;
            const int yg = j0 - R + 1 + row, xg = i0 + col;
            svy[row][col] = (yg >= lo && yg <= hy && xg <= hx) ? F.vy[at3(S, lo, xg, yg, k)] : 0.0f;
        }
        __syncthreads();
        const bool zslab = zslab_p(k, nz, fs);
        const float azv = zslab ? az_i[k + pml - 1] : 0.0f, bzv = zslab ? bz_i[k + pml - 1] : 0.0f;
#pragma unroll
        for (int r = 0; r < kRY; ++r) {
            if (ok[r]) {
                const int row = threadIdx.y + kBY * r, j = j0 + row;
                const int cx = lane + R - 1, cy = row + R - 1;
                float sx = c.c[0] * (svx[row][cx + 1] - svx[row][cx]);
                float sy = c.c[0] * (svy[cy + 1][lane] - svy[cy][lane]);
                float sz = c.c[0] * (q[r][R] - q[r][R - 1]);
#pragma unroll
                for (int m = 2; m <= R; ++m) {
                    sx = sx + c.c[m - 1] * (svx[row][cx + m] - svx[row][cx - m + 1]);
                    sy = sy + c.c[m - 1] * (svy[cy + m][lane] - svy[cy - m + 1][lane]);
                    sz = sz + c.c[m - 1] * (q[r][R - 1 + m] - q[r][R - m]);
                }
                float dxv = idx * sx, dyv = idy * sy, dzv = idz * sz;
                const size_t qq = at3(S, lo, i, j, k);
                if (xslab) {
                    const float mem = axv * dxv + bxv * F.m[0][qq];
                    F.m[0][qq] = mem;
                    dxv = dxv + mem;
                }
                if (j <= 0 || j >= ny + 1) {
                    const float mem = ay_i[j + pml - 1] * dyv + by_i[j + pml - 1] * F.m[1][qq];
                    F.m[1][qq] = mem;
                    dyv = dyv + mem;
                }
                if (zslab) {
                    const float mem = azv * dzv + bzv * F.m[2][qq];
                    F.m[2][qq] = mem;
                    dzv = dzv + mem;
                }
                F.po[qq] = F.p[qq] - dt * bk[qq] * (dxv + dyv + dzv);
            }
#pragma unroll
            for (int m = 0; m < 2 * R - 1; ++m) q[r][m] = q[r][m + 1];
        }
    }
}

// ---------------------------------------------------------------------------
// v-sweeps (module_wavefield.f90:106-170): v -= dt/rho_eff * d- p with
// rho_eff the mean of the two nodes, CFS-PML in the half-node slabs.
// ---------------------------------------------------------------------------
template <int R>
__global__ __launch_bounds__(kBX * kBY) void k3_update_v(const Shot3* __restrict__ shots,
                                                         const int* __restrict__ tile_shot, F3 F,
                                                         const float* __restrict__ rho,
                                                         const float* __restrict__ prof, Coef c,
                                                         float dt, float idx, float idy, float idz,
                                                         int pml, int kc, int fs) {
    const int tile = blockIdx.x;
    const Shot3 S = shots[tile_shot[tile]];
    int lt = tile - S.tile0;
    const int tx = lt % S.ntx;
    lt /= S.ntx;
    const int ty = lt % S.nty, tzc = lt / S.nty;
    const int nx = S.nx, ny = S.ny, nz = S.nz;
    const int lo = 1 - pml - R, hx = nx + pml + R, hy = ny + pml + R;
    const int i0 = -pml + 1 + tx * kBX, j0 = -pml + 1 + ty * kTY;
    const int kt = -pml + 1 + tzc * kc, k1 = min(kt + kc - 1, nz + pml);
    const int k0 = fs ? max(kt, 1) : kt;
    const int lane = threadIdx.x, tid = threadIdx.y * kBX + threadIdx.x;
    const int i = i0 + lane;
    // p plane: x = i0-R .. i0+63+R-1 at col = x - (i0-R); y = j0-R .. j0+15+R-1 at row = y - (j0-R)
    __shared__ float sp[kTY + 2 * R][kBX + 2 * R];
    const int lx = nx + 2 * pml, ly = ny + 2 * pml, lz = nz + 2 * pml;
    const float* ax_h = prof + S.prof + 2 * lx;
    const float* bx_h = ax_h + lx;
    const float* ay_h = prof + S.prof + 4 * lx + 2 * ly;
    const float* by_h = ay_h + ly;
    const float* az_h = prof + S.prof + 4 * lx + 4 * ly + 2 * lz;
    const float* bz_h = az_h + lz;
    bool ok[kRY];
    float q[kRY][2 * R];
    float rprev[kRY];
#pragma unroll
    for (int r = 0; r < kRY; ++r) {
        const int j = j0 + threadIdx.y + kBY * r;
        ok[r] = (i <= nx + pml) && (j <= ny + pml);
#pragma unroll
        for (int m = 0; m < 2 * R - 1; ++m) q[r][m] = ok[r] ? F.p[at3(S, lo, i, j, k0 - R + m)] : 0.0f;
        rprev[r] = ok[r] ? rho[at3(S, lo, i, j, k0 - 1)] : 0.0f;
    }
    const bool xslab = (i <= 1 || i >= nx + 1);
    const float axv = xslab && i <= nx + pml ? ax_h[i + pml - 1] : 0.0f;
    const float bxv = xslab && i <= nx + pml ? bx_h[i + pml - 1] : 0.0f;
    for (int k = k0; k <= k1; ++k) {
#pragma unroll
        for (int r = 0; r < kRY; ++r) {
            const int j = j0 + threadIdx.y + kBY * r;
            q[r][2 * R - 1] = ok[r] ? F.p[at3(S, lo, i, j, k + R - 1)] : 0.0f;
        }
        __syncthreads();
        // rows of the tile with the x halo (R left, R-1 right)
#pragma unroll
        for (int r = 0; r < kRY; ++r) {
            const int row = threadIdx.y + kBY * r, j = j0 + row;
            sp[row + R][lane + R] = (i <= hx && j <= hy) ? F.p[at3(S, lo, i, j, k)] : 0.0f;
        }
        if (tid < (2 * R - 1) * kTY) {
            const int hc = tid % (2 * R - 1), row = tid / (2 * R - 1);
            const int col = hc < R ? hc : kBX + hc;  // 0..R-1 and 64+R..64+2R-2
            const int xg = i0 - R + col, j = j0 + row;
            sp[row + R][col] = (xg >= lo && xg <= hx && j <= hy) ? F.p[at3(S, lo, xg, j, k)] : 0.0f;
        }
        // y halo rows (R above, R-1 below) over the tile's columns
        for (int e = tid; e < (2 * R - 1) * kBX; e += kBX * kBY) {
            const int hr = e / kBX, col = e % kBX;
            const int row = hr < R ? hr : kTY + hr;  // 0..R-1 and 16+R..16+2R-2
            const int yg = j0 - R + row, xg = i0 + col;
            sp[row][col + R] = (yg >= lo && yg <= hy && xg <= hx) ? F.p[at3(S, lo, xg, yg, k)] : 0.0f;
        }
        __syncthreads();
        const bool zslab = zslab_v(k, nz, fs);
        const bool vz_on = !fs || k >= 2;
        const float azv = zslab ? az_h[k + pml - 1] : 0.0f, bzv = zslab ? bz_h[k + pml - 1] : 0.0f;
#pragma unroll
        for (int r = 0; r < kRY; ++r) {
            if (ok[r]) {
                const int row = threadIdx.y + kBY * r, j = j0 + row;
                const int cx = lane + R, cy = row + R;
                float sx = c.c[0] * (sp[cy][cx] - sp[cy][cx - 1]);
                float sy = c.c[0] * (sp[cy][cx] - sp[cy - 1][cx]);
                float sz = c.c[0] * (q[r][R] - q[r][R - 1]);
#pragma unroll
                for (int m = 2; m <= R; ++m) {
                    sx = sx + c.c[m - 1] * (sp[cy][cx + m - 1] - sp[cy][cx - m]);
                    sy = sy + c.c[m - 1] * (sp[cy + m - 1][cx] - sp[cy - m][cx]);
                    sz = sz + c.c[m - 1] * (q[r][R - 1 + m] - q[r][R - m]);
                }
                float dxv = idx * sx, dyv = idy * sy, dzv = idz * sz;
                const size_t qq = at3(S, lo, i, j, k);
                const float r0 = rho[qq];
                const float rx = 0.5f * (r0 + rho[qq - 1]);
                const float ry = 0.5f * (r0 + rho[qq - size_t(S.px)]);
                const float rz = 0.5f * (r0 + rprev[r]);
                rprev[r] = r0;
                if (xslab) {
                    const float mem = axv * dxv + bxv * F.m[3][qq];
                    F.m[3][qq] = mem;
                    dxv = dxv + mem;
                }
                if (j <= 1 || j >= ny + 1) {
                    const float mem = ay_h[j + pml - 1] * dyv + by_h[j + pml - 1] * F.m[4][qq];
                    F.m[4][qq] = mem;
                    dyv = dyv + mem;
                }
                if (zslab) {
                    const float mem = azv * dzv + bzv * F.m[5][qq];
                    F.m[5][qq] = mem;
                    dzv = dzv + mem;
                }
                F.vxo[qq] = F.vx[qq] - dt / rx * dxv;
                F.vyo[qq] = F.vy[qq] - dt / ry * dyv;
                if (vz_on) F.vzo[qq] = F.vz[qq] - dt / rz * dzv;
            }
#pragma unroll
            for (int m = 0; m < 2 * R - 1; ++m) q[r][m] = q[r][m + 1];
        }
    }
}


// ---------------------------------------------------------------------------
// Barrier-free variant: one thread per (x, y) column segment of kc planes,
// z neighbours from the register queue, x and y neighbours read directly
// (L1/L2 hits: the wavefront covers 64 consecutive x nodes, the block 4
// rows).  No LDS, no barriers, so many waves keep loads in flight.
// ---------------------------------------------------------------------------
template <int R>
__global__ __launch_bounds__(1024) void k3c_update_p(const Shot3* __restrict__ shots,
                                                    const int* __restrict__ tile_shot, F3 F,
                                                    const float* __restrict__ bk,
                                                    const float* __restrict__ prof, Coef c, float dt,
                                                    float idx, float idy, float idz, int pml, int kc,
                                                    int ntiles, int remap, int fs) {
    const int tile = xcd_tile(ntiles, remap);
    if (tile < 0) return;
    const Shot3 S = shots[tile_shot[tile]];
    int lt = tile - S.tile0;
    const int tx = lt % S.ntx;
    lt /= S.ntx;
    const int ty = lt % S.nty, tzc = lt / S.nty;
    const int nx = S.nx, ny = S.ny, nz = S.nz;
    const int lo = 1 - pml - R;
    const int i = -pml + 1 + tx * 64 + int(threadIdx.x);
    // blockDim.x = 64: the row is wavefront-uniform (scalar registers)
    const int j = -pml + 1 + ty * int(blockDim.y) + __builtin_amdgcn_readfirstlane(int(threadIdx.y));
    if (i > nx + pml || j > ny + pml) return;
    const int kt = -pml + 1 + tzc * kc, k1 = min(kt + kc - 1, nz + pml);
    const int k0 = fs ? max(kt, 2) : kt;
    const float* __restrict__ vx = F.vx;
    const float* __restrict__ vy = F.vy;
    const float* __restrict__ vz = F.vz;
    const int lx = nx + 2 * pml, ly = ny + 2 * pml, lz = nz + 2 * pml;
    const float* az_i = prof + S.prof + 4 * lx + 4 * ly;
    const float* bz_i = az_i + lz;
    const bool xslab = (i <= 0 || i >= nx + 1), yslab = (j <= 0 || j >= ny + 1);
    const float axv = xslab ? prof[S.prof + i + pml - 1] : 0.0f;
    const float bxv = xslab ? prof[S.prof + lx + i + pml - 1] : 0.0f;
    const float ayv = yslab ? prof[S.prof + 4 * lx + j + pml - 1] : 0.0f;
    const float byv = yslab ? prof[S.prof + 4 * lx + ly + j + pml - 1] : 0.0f;
    const size_t px = S.px, pxy = S.pxy;
    size_t q = at3(S, lo, i, j, k0);
    float qz[2 * R];
#pragma unroll
    for (int m = 0; m < 2 * R - 1; ++m) qz[m] = vz[q + (m - R + 1) * pxy];
    for (int k = k0; k <= k1; ++k, q += pxy) {
        qz[2 * R - 1] = vz[q + R * pxy];
        float sx = c.c[0] * (vx[q + 1] - vx[q]);
        float sy = c.c[0] * (vy[q + px] - vy[q]);
        float sz = c.c[0] * (qz[R] - qz[R - 1]);
#pragma unroll
        for (int m = 2; m <= R; ++m) {
            sx = sx + c.c[m - 1] * (vx[q + m] - vx[q - (m - 1)]);
            sy = sy + c.c[m - 1] * (vy[q + m * px] - vy[q - (m - 1) * px]);
            sz = sz + c.c[m - 1] * (qz[R - 1 + m] - qz[R - m]);
        }
        float dxv = idx * sx, dyv = idy * sy, dzv = idz * sz;
        if (xslab) {
            const float mem = axv * dxv + bxv * F.m[0][q];
            F.m[0][q] = mem;
            dxv = dxv + mem;
        }
        if (yslab) {
            const float mem = ayv * dyv + byv * F.m[1][q];
            F.m[1][q] = mem;
            dyv = dyv + mem;
        }
        if (zslab_p(k, nz, fs)) {
            const float mem = az_i[k + pml - 1] * dzv + bz_i[k + pml - 1] * F.m[2][q];
            F.m[2][q] = mem;
            dzv = dzv + mem;
        }
        F.po[q] = F.p[q] - dt * bk[q] * (dxv + dyv + dzv);
#pragma unroll
        for (int m = 0; m < 2 * R - 1; ++m) qz[m] = qz[m + 1];
    }
}

template <int R>
__global__ __launch_bounds__(1024) void k3c_update_v(const Shot3* __restrict__ shots,
                                                    const int* __restrict__ tile_shot, F3 F,
                                                    const float* __restrict__ rho,
                                                    const float* __restrict__ prof, Coef c, float dt,
                                                    float idx, float idy, float idz, int pml, int kc,
                                                    int ntiles, int remap, int fs) {
    const int tile = xcd_tile(ntiles, remap);
    if (tile < 0) return;
    const Shot3 S = shots[tile_shot[tile]];
    int lt = tile - S.tile0;
    const int tx = lt % S.ntx;
    lt /= S.ntx;
    const int ty = lt % S.nty, tzc = lt / S.nty;
    const int nx = S.nx, ny = S.ny, nz = S.nz;
    const int lo = 1 - pml - R;
    const int i = -pml + 1 + tx * 64 + int(threadIdx.x);
    // blockDim.x = 64: the row is wavefront-uniform (scalar registers)
    const int j = -pml + 1 + ty * int(blockDim.y) + __builtin_amdgcn_readfirstlane(int(threadIdx.y));
    if (i > nx + pml || j > ny + pml) return;
    const int kt = -pml + 1 + tzc * kc, k1 = min(kt + kc - 1, nz + pml);
    const int k0 = fs ? max(kt, 1) : kt;
    const float* __restrict__ pn = F.p;
    const int lx = nx + 2 * pml, ly = ny + 2 * pml, lz = nz + 2 * pml;
    const float* az_h = prof + S.prof + 4 * lx + 4 * ly + 2 * lz;
    const float* bz_h = az_h + lz;
    const bool xslab = (i <= 1 || i >= nx + 1), yslab = (j <= 1 || j >= ny + 1);
    const float axv = xslab ? prof[S.prof + 2 * lx + i + pml - 1] : 0.0f;
    const float bxv = xslab ? prof[S.prof + 3 * lx + i + pml - 1] : 0.0f;
    const float ayv = yslab ? prof[S.prof + 4 * lx + 2 * ly + j + pml - 1] : 0.0f;
    const float byv = yslab ? prof[S.prof + 4 * lx + 3 * ly + j + pml - 1] : 0.0f;
    const size_t px = S.px, pxy = S.pxy;
    size_t q = at3(S, lo, i, j, k0);
    float qz[2 * R];
#pragma unroll
    for (int m = 0; m < 2 * R - 1; ++m) qz[m] = pn[q + (m - R) * pxy];
    float rprev = rho[q - pxy];
    for (int k = k0; k <= k1; ++k, q += pxy) {
        qz[2 * R - 1] = pn[q + (R - 1) * pxy];
        float sx = c.c[0] * (pn[q] - pn[q - 1]);
        float sy = c.c[0] * (pn[q] - pn[q - px]);
        float sz = c.c[0] * (qz[R] - qz[R - 1]);
#pragma unroll
        for (int m = 2; m <= R; ++m) {
            sx = sx + c.c[m - 1] * (pn[q + (m - 1)] - pn[q - m]);
            sy = sy + c.c[m - 1] * (pn[q + (m - 1) * px] - pn[q - m * px]);
            sz = sz + c.c[m - 1] * (qz[R - 1 + m] - qz[R - m]);
        }
        float dxv = idx * sx, dyv = idy * sy, dzv = idz * sz;
        const float r0 = rho[q];
        const float rx = 0.5f * (r0 + rho[q - 1]);
        const float ry = 0.5f * (r0 + rho[q - px]);
        const float rz = 0.5f * (r0 + rprev);
        rprev = r0;
        if (xslab) {
            const float mem = axv * dxv + bxv * F.m[3][q];
            F.m[3][q] = mem;
            dxv = dxv + mem;
        }
        if (yslab) {
            const float mem = ayv * dyv + byv * F.m[4][q];
            F.m[4][q] = mem;
            dyv = dyv + mem;
        }
        if (zslab_v(k, nz, fs)) {
            const float mem = az_h[k + pml - 1] * dzv + bz_h[k + pml - 1] * F.m[5][q];
            F.m[5][q] = mem;
            dzv = dzv + mem;
        }
        F.vxo[q] = F.vx[q] - dt / rx * dxv;
        F.vyo[q] = F.vy[q] - dt / ry * dyv;
        if (!fs || k >= 2) F.vzo[q] = F.vz[q] - dt / rz * dzv;
#pragma unroll
        for (int m = 0; m < 2 * R - 1; ++m) qz[m] = qz[m + 1];
    }
}


// ---------------------------------------------------------------------------
// Software-pipelined columns: as the barrier-free columns, but every operand
// of plane k+1 (the x and y stencil rows, the z queue's new plane, the node's
// own values) is loaded before plane k is computed and stored, so each wave
// keeps two planes of loads in flight.
// ---------------------------------------------------------------------------
template <int R>
__global__ __launch_bounds__(1024) void k3q_update_p(const Shot3* __restrict__ shots,
                                                    const int* __restrict__ tile_shot, F3 F,
                                                    const float* __restrict__ bk,
                                                    const float* __restrict__ prof, Coef c, float dt,
                                                    float idx, float idy, float idz, int pml, int kc,
                                                    int ntiles, int remap, int fs) {
    const int tile = xcd_tile(ntiles, remap);
    if (tile < 0) return;
    const Shot3 S = shots[tile_shot[tile]];
    int lt = tile - S.tile0;
    const int tx = lt % S.ntx;
    lt /= S.ntx;
    const int ty = lt % S.nty, tzc = lt / S.nty;
    const int nx = S.nx, ny = S.ny, nz = S.nz;
    const int lo = 1 - pml - R;
    const int i = -pml + 1 + tx * 64 + int(threadIdx.x);
    // blockDim.x = 64: the row is wavefront-uniform (scalar registers)
    const int j = -pml + 1 + ty * int(blockDim.y) + __builtin_amdgcn_readfirstlane(int(threadIdx.y));
    if (i > nx + pml || j > ny + pml) return;
    const int kt = -pml + 1 + tzc * kc, k1 = min(kt + kc - 1, nz + pml);
    const int k0 = fs ? max(kt, 2) : kt;
    if (k0 > k1) return;
    const float* __restrict__ vx = F.vx;
    const float* __restrict__ vy = F.vy;
    const float* __restrict__ vz = F.vz;
    const float* __restrict__ pin = F.p;
    const int lx = nx + 2 * pml, ly = ny + 2 * pml, lz = nz + 2 * pml;
    const float* az_i = prof + S.prof + 4 * lx + 4 * ly;
    const float* bz_i = az_i + lz;
    const bool xslab = (i <= 0 || i >= nx + 1), yslab = (j <= 0 || j >= ny + 1);
    const float axv = xslab ? prof[S.prof + i + pml - 1] : 0.0f;
    const float bxv = xslab ? prof[S.prof + lx + i + pml - 1] : 0.0f;
    const float ayv = yslab ? prof[S.prof + 4 * lx + j + pml - 1] : 0.0f;
    const float byv = yslab ? prof[S.prof + 4 * lx + ly + j + pml - 1] : 0.0f;
    const size_t px = S.px, pxy = S.pxy;
    size_t q = at3(S, lo, i, j, k0);
    float qz[2 * R];
#pragma unroll
    for (int m = 0; m < 2 * R - 1; ++m) qz[m] = vz[q + (m - R + 1) * pxy];
    // operands of the current plane: X[m] = vx(i - R + 1 + m), Y[m] = vy(j - R + 1 + m)
    float X[2 * R], Y[2 * R], Zn, P, K;
#pragma unroll
    for (int m = 0; m < 2 * R; ++m) {
        X[m] = vx[q + m - (R - 1)];
        Y[m] = vy[q + (m - (R - 1)) * px];
    }
    Zn = vz[q + R * pxy];
    P = pin[q];
    K = bk[q];
    for (int k = k0; k <= k1; ++k, q += pxy) {
        float Xn[2 * R], Yn[2 * R], Zn2 = 0.f, Pn = 0.f, Kn = 0.f;
        const bool more = k < k1;
        if (more) {
            const size_t qn = q + pxy;
#pragma unroll
            for (int m = 0; m < 2 * R; ++m) {
                Xn[m] = vx[qn + m - (R - 1)];
                Yn[m] = vy[qn + (m - (R - 1)) * px];
            }
            Zn2 = vz[qn + R * pxy];
            Pn = pin[qn];
            Kn = bk[qn];
        }
        qz[2 * R - 1] = Zn;
        float sx = c.c[0] * (X[R] - X[R - 1]);
        float sy = c.c[0] * (Y[R] - Y[R - 1]);
        float sz = c.c[0] * (qz[R] - qz[R - 1]);
#pragma unroll
        for (int m = 2; m <= R; ++m) {
            sx = sx + c.c[m - 1] * (X[R - 1 + m] - X[R - m]);
            sy = sy + c.c[m - 1] * (Y[R - 1 + m] - Y[R - m]);
            sz = sz + c.c[m - 1] * (qz[R - 1 + m] - qz[R - m]);
        }
        float dxv = idx * sx, dyv = idy * sy, dzv = idz * sz;
        if (xslab) {
            const float mem = axv * dxv + bxv * F.m[0][q];
            F.m[0][q] = mem;
            dxv = dxv + mem;
        }
        if (yslab) {
            const float mem = ayv * dyv + byv * F.m[1][q];
            F.m[1][q] = mem;
            dyv = dyv + mem;
        }
        if (zslab_p(k, nz, fs)) {
            const float mem = az_i[k + pml - 1] * dzv + bz_i[k + pml - 1] * F.m[2][q];
            F.m[2][q] = mem;
            dzv = dzv + mem;
        }
        F.po[q] = P - dt * K * (dxv + dyv + dzv);
#pragma unroll
        for (int m = 0; m < 2 * R - 1; ++m) qz[m] = qz[m + 1];
        if (more) {
#pragma unroll
            for (int m = 0; m < 2 * R; ++m) {
                X[m] = Xn[m];
                Y[m] = Yn[m];
            }
            Zn = Zn2;
            P = Pn;
            K = Kn;
        }
    }
}

template <int R>
__global__ __launch_bounds__(1024) void k3q_update_v(const Shot3* __restrict__ shots,
                                                    const int* __restrict__ tile_shot, F3 F,
                                                    const float* __restrict__ rho,
                                                    const float* __restrict__ prof, Coef c, float dt,
                                                    float idx, float idy, float idz, int pml, int kc,
                                                    int ntiles, int remap, int fs) {
    const int tile = xcd_tile(ntiles, remap);
    if (tile < 0) return;
    const Shot3 S = shots[tile_shot[tile]];
    int lt = tile - S.tile0;
    const int tx = lt % S.ntx;
    lt /= S.ntx;
    const int ty = lt % S.nty, tzc = lt / S.nty;
    const int nx = S.nx, ny = S.ny, nz = S.nz;
    const int lo = 1 - pml - R;
    const int i = -pml + 1 + tx * 64 + int(threadIdx.x);
    // blockDim.x = 64: the row is wavefront-uniform (scalar registers)
    const int j = -pml + 1 + ty * int(blockDim.y) + __builtin_amdgcn_readfirstlane(int(threadIdx.y));
    if (i > nx + pml || j > ny + pml) return;
    const int kt = -pml + 1 + tzc * kc, k1 = min(kt + kc - 1, nz + pml);
    const int k0 = fs ? max(kt, 1) : kt;
    if (k0 > k1) return;
    const float* __restrict__ pn = F.p;
    const float* __restrict__ vx = F.vx;
    const float* __restrict__ vy = F.vy;
    const float* __restrict__ vz = F.vz;
    const int lx = nx + 2 * pml, ly = ny + 2 * pml, lz = nz + 2 * pml;
    const float* az_h = prof + S.prof + 4 * lx + 4 * ly + 2 * lz;
    const float* bz_h = az_h + lz;
    const bool xslab = (i <= 1 || i >= nx + 1), yslab = (j <= 1 || j >= ny + 1);
    const float axv = xslab ? prof[S.prof + 2 * lx + i + pml - 1] : 0.0f;
    const float bxv = xslab ? prof[S.prof + 3 * lx + i + pml - 1] : 0.0f;
    const float ayv = yslab ? prof[S.prof + 4 * lx + 2 * ly + j + pml - 1] : 0.0f;
    const float byv = yslab ? prof[S.prof + 4 * lx + 3 * ly + j + pml - 1] : 0.0f;
    const size_t px = S.px, pxy = S.pxy;
    size_t q = at3(S, lo, i, j, k0);
    float qz[2 * R];
#pragma unroll
    for (int m = 0; m < 2 * R - 1; ++m) qz[m] = pn[q + (m - R) * pxy];
    float rprev = rho[q - pxy];
    // X[m] = p(i - R + m), Y[m] = p(j - R + m); m = 0 .. 2R-1
    float X[2 * R], Y[2 * R], Zn, r0, rl, rym, ux, uy, uz;
#pragma unroll
    for (int m = 0; m < 2 * R; ++m) {
        X[m] = pn[q + m - R];
        Y[m] = pn[q + (m - R) * px];
    }
    Zn = pn[q + (R - 1) * pxy];
    r0 = rho[q];
    rl = rho[q - 1];
    rym = rho[q - px];
    ux = vx[q];
    uy = vy[q];
    uz = vz[q];
    for (int k = k0; k <= k1; ++k, q += pxy) {
        float Xn[2 * R], Yn[2 * R], Zn2 = 0.f, r0n = 0.f, rln = 0.f, rymn = 0.f, uxn = 0.f, uyn = 0.f, uzn = 0.f;
        const bool more = k < k1;
        if (more) {
            const size_t qn = q + pxy;
#pragma unroll
            for (int m = 0; m < 2 * R; ++m) {
                Xn[m] = pn[qn + m - R];
                Yn[m] = pn[qn + (m - R) * px];
            }
            Zn2 = pn[qn + (R - 1) * pxy];
            r0n = rho[qn];
            rln = rho[qn - 1];
            rymn = rho[qn - px];
            uxn = vx[qn];
            uyn = vy[qn];
            uzn = vz[qn];
        }
        qz[2 * R - 1] = Zn;
        float sx = c.c[0] * (X[R] - X[R - 1]);
        float sy = c.c[0] * (Y[R] - Y[R - 1]);
        float sz = c.c[0] * (qz[R] - qz[R - 1]);
#pragma unroll
        for (int m = 2; m <= R; ++m) {
            sx = sx + c.c[m - 1] * (X[R + m - 1] - X[R - m]);
            sy = sy + c.c[m - 1] * (Y[R + m - 1] - Y[R - m]);
            sz = sz + c.c[m - 1] * (qz[R - 1 + m] - qz[R - m]);
        }
        float dxv = idx * sx, dyv = idy * sy, dzv = idz * sz;
        const float rx = 0.5f * (r0 + rl);
        const float ry = 0.5f * (r0 + rym);
        const float rz = 0.5f * (r0 + rprev);
        rprev = r0;
        if (xslab) {
            const float mem = axv * dxv + bxv * F.m[3][q];
            F.m[3][q] = mem;
            dxv = dxv + mem;
        }
        if (yslab) {
            const float mem = ayv * dyv + byv * F.m[4][q];
            F.m[4][q] = mem;
            dyv = dyv + mem;
        }
        if (zslab_v(k, nz, fs)) {
            const float mem = az_h[k + pml - 1] * dzv + bz_h[k + pml - 1] * F.m[5][q];
            F.m[5][q] = mem;
            dzv = dzv + mem;
        }
        F.vxo[q] = ux - dt / rx * dxv;
        F.vyo[q] = uy - dt / ry * dyv;
        if (!fs || k >= 2) F.vzo[q] = uz - dt / rz * dzv;
#pragma unroll
        for (int m = 0; m < 2 * R - 1; ++m) qz[m] = qz[m + 1];
        if (more) {
#pragma unroll
            for (int m = 0; m < 2 * R; ++m) {
                X[m] = Xn[m];
                Y[m] = Yn[m];
            }
            Zn = Zn2;
            r0 = r0n;
            rl = rln;
            rym = rymn;
            ux = uxn;
            uy = uyn;
            uz = uzn;
        }
    }
}

// ---------------------------------------------------------------------------
// Four-wide column variant: each thread updates 4 consecutive x nodes with
// 16-byte loads and stores (x neighbours from A/4*2+1 aligned float4s, y
// neighbours one float4 per row, z neighbours from a float4 register
// queue).  Block = 16 x 16 threads = a 64 x 16 x-y tile.
// ---------------------------------------------------------------------------
constexpr int kWX = 16, kWY = 16;

__device__ __forceinline__ float4 ld4(const float* p, size_t q) { return *reinterpret_cast<const float4*>(p + q); }
__device__ __forceinline__ void st4(float* p, size_t q, float4 v) { *reinterpret_cast<float4*>(p + q) = v; }
__device__ __forceinline__ float c4(const float4& v, int e) { return e == 0 ? v.x : (e == 1 ? v.y : (e == 2 ? v.z : v.w)); }
__device__ __forceinline__ void s4(float4& v, int e, float x) {
    if (e == 0) v.x = x; else if (e == 1) v.y = x; else if (e == 2) v.z = x; else v.w = x;
}

template <int R>
__global__ __launch_bounds__(kWX * kWY) void k3w_update_p(const Shot3* __restrict__ shots,
                                                          const int* __restrict__ tile_shot, F3 F,
                                                          const float* __restrict__ bk,
                                                          const float* __restrict__ prof, Coef c, float dt,
                                                          float idx, float idy, float idz, int pml, int kc,
                                                          int ntiles, int remap, int fs) {
    constexpr int A = (R + 3) / 4 * 4, NW = 2 * A / 4 + 1;
    const int tile = xcd_tile(ntiles, remap);
    if (tile < 0) return;
    const Shot3 S = shots[tile_shot[tile]];
    int lt = tile - S.tile0;
    const int tx = lt % S.ntx;
    lt /= S.ntx;
    const int ty = lt % S.nty, tzc = lt / S.nty;
    const int nx = S.nx, ny = S.ny, nz = S.nz;
    const int lo = 1 - pml - R;
    const int x0 = -pml + 1 + tx * (4 * kWX) + 4 * int(threadIdx.x);
    const int j = -pml + 1 + ty * kWY + int(threadIdx.y);
    if (x0 > nx + pml || j > ny + pml) return;
    const int kt = -pml + 1 + tzc * kc, k1 = min(kt + kc - 1, nz + pml);
    const int k0 = fs ? max(kt, 2) : kt;
    const int lx = nx + 2 * pml, ly = ny + 2 * pml, lz = nz + 2 * pml;
    const float* ax_i = prof + S.prof;
    const float* bx_i = ax_i + lx;
    const float* az_i = prof + S.prof + 4 * lx + 4 * ly;
    const float* bz_i = az_i + lz;
    bool xs[4];
    float axv[4], bxv[4];
    bool anyx = false;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int x = x0 + e;
        xs[e] = (x <= 0 || x >= nx + 1) && x <= nx + pml;
        axv[e] = xs[e] ? ax_i[x + pml - 1] : 0.0f;
        bxv[e] = xs[e] ? bx_i[x + pml - 1] : 0.0f;
        anyx = anyx || xs[e];
    }
    const bool yslab = (j <= 0 || j >= ny + 1);
    const float ayv = yslab ? prof[S.prof + 4 * lx + j + pml - 1] : 0.0f;
    const float byv = yslab ? prof[S.prof + 4 * lx + ly + j + pml - 1] : 0.0f;
    const size_t px = S.px, pxy = S.pxy;
    size_t q = at3(S, lo, x0, j, k0);
    float4 qz[2 * R];
#pragma unroll
    for (int m = 0; m < 2 * R - 1; ++m) qz[m] = ld4(F.vz, q + (m - R + 1) * pxy);
    for (int k = k0; k <= k1; ++k, q += pxy) {
        qz[2 * R - 1] = ld4(F.vz, q + R * pxy);
        float w[4 * NW];
#pragma unroll
        for (int h = 0; h < NW; ++h) {
            const float4 v = ld4(F.vx, q - A + 4 * h);
            w[4 * h] = v.x; w[4 * h + 1] = v.y; w[4 * h + 2] = v.z; w[4 * h + 3] = v.w;
        }
        float4 yp[R], ym[R];  // vy(j+m) and vy(j-m+1), m = 1..R
#pragma unroll
        for (int m = 1; m <= R; ++m) {
            yp[m - 1] = ld4(F.vy, q + m * px);
            ym[m - 1] = ld4(F.vy, q - (m - 1) * px);
        }
        float4 pv = ld4(F.p, q);
        const float4 kv = ld4(bk, q);
        const bool zslab = zslab_p(k, nz, fs);
        float4 mx = make_float4(0, 0, 0, 0), my = mx, mz = mx;
        if (anyx) mx = ld4(F.m[0], q);
        if (yslab) my = ld4(F.m[1], q);
        if (zslab) mz = ld4(F.m[2], q);
        const float azv = zslab ? az_i[k + pml - 1] : 0.0f, bzv = zslab ? bz_i[k + pml - 1] : 0.0f;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            float sx = c.c[0] * (w[A + e + 1] - w[A + e]);
            float sy = c.c[0] * (c4(yp[0], e) - c4(ym[0], e));
            float sz = c.c[0] * (c4(qz[R], e) - c4(qz[R - 1], e));
#pragma unroll
            for (int m = 2; m <= R; ++m) {
                sx = sx + c.c[m - 1] * (w[A + e + m] - w[A + e - m + 1]);
                sy = sy + c.c[m - 1] * (c4(yp[m - 1], e) - c4(ym[m - 1], e));
                sz = sz + c.c[m - 1] * (c4(qz[R - 1 + m], e) - c4(qz[R - m], e));
            }
            float dxv = idx * sx, dyv = idy * sy, dzv = idz * sz;
            if (xs[e]) {
                const float mem = axv[e] * dxv + bxv[e] * c4(mx, e);
                s4(mx, e, mem);
                dxv = dxv + mem;
            }
            if (yslab) {
                const float mem = ayv * dyv + byv * c4(my, e);
                s4(my, e, mem);
                dyv = dyv + mem;
            }
            if (zslab) {
                const float mem = azv * dzv + bzv * c4(mz, e);
                s4(mz, e, mem);
                dzv = dzv + mem;
            }
            s4(pv, e, c4(pv, e) - dt * c4(kv, e) * (dxv + dyv + dzv));
        }
        // nodes past nx+pml in the last float4 are halo: they hold zeros and
        // the update of a zero stencil input leaves them zero only if the
        // stencil is zero there, so mask the write
        if (x0 + 3 > nx + pml) {
            const float4 old = ld4(F.p, q);
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (x0 + e > nx + pml) s4(pv, e, c4(old, e));
        }
        st4(F.po, q, pv);
        if (anyx) st4(F.m[0], q, mx);
        if (yslab) st4(F.m[1], q, my);
        if (zslab) st4(F.m[2], q, mz);
#pragma unroll
        for (int m = 0; m < 2 * R - 1; ++m) qz[m] = qz[m + 1];
    }
}

template <int R>
__global__ __launch_bounds__(kWX * kWY) void k3w_update_v(const Shot3* __restrict__ shots,
                                                          const int* __restrict__ tile_shot, F3 F,
                                                          const float* __restrict__ rho,
                                                          const float* __restrict__ prof, Coef c, float dt,
                                                          float idx, float idy, float idz, int pml, int kc,
                                                          int ntiles, int remap, int fs) {
    constexpr int A = (R + 3) / 4 * 4, NW = 2 * A / 4 + 1;
    const int tile = xcd_tile(ntiles, remap);
    if (tile < 0) return;
    const Shot3 S = shots[tile_shot[tile]];
    int lt = tile - S.tile0;
    const int tx = lt % S.ntx;
    lt /= S.ntx;
    const int ty = lt % S.nty, tzc = lt / S.nty;
    const int nx = S.nx, ny = S.ny, nz = S.nz;
    const int lo = 1 - pml - R;
    const int x0 = -pml + 1 + tx * (4 * kWX) + 4 * int(threadIdx.x);
    const int j = -pml + 1 + ty * kWY + int(threadIdx.y);
    if (x0 > nx + pml || j > ny + pml) return;
    const int kt = -pml + 1 + tzc * kc, k1 = min(kt + kc - 1, nz + pml);
    const int k0 = fs ? max(kt, 1) : kt;
    const int lx = nx + 2 * pml, ly = ny + 2 * pml, lz = nz + 2 * pml;
    const float* ax_h = prof + S.prof + 2 * lx;
    const float* bx_h = ax_h + lx;
    const float* az_h = prof + S.prof + 4 * lx + 4 * ly + 2 * lz;
    const float* bz_h = az_h + lz;
    bool xs[4];
    float axv[4], bxv[4];
    bool anyx = false;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int x = x0 + e;
        xs[e] = (x <= 1 || x >= nx + 1) && x <= nx + pml;
        axv[e] = xs[e] ? ax_h[x + pml - 1] : 0.0f;
        bxv[e] = xs[e] ? bx_h[x + pml - 1] : 0.0f;
        anyx = anyx || xs[e];
    }
    const bool yslab = (j <= 1 || j >= ny + 1);
    const float ayv = yslab ? prof[S.prof + 4 * lx + 2 * ly + j + pml - 1] : 0.0f;
    const float byv = yslab ? prof[S.prof + 4 * lx + 3 * ly + j + pml - 1] : 0.0f;
    const size_t px = S.px, pxy = S.pxy;
    size_t q = at3(S, lo, x0, j, k0);
    float4 qz[2 * R];
#pragma unroll
    for (int m = 0; m < 2 * R - 1; ++m) qz[m] = ld4(F.p, q + (m - R) * pxy);
    float4 rprev = ld4(rho, q - pxy);
    for (int k = k0; k <= k1; ++k, q += pxy) {
        qz[2 * R - 1] = ld4(F.p, q + (R - 1) * pxy);
        float w[4 * NW];
#pragma unroll
        for (int h = 0; h < NW; ++h) {
            const float4 v = ld4(F.p, q - A + 4 * h);
            w[4 * h] = v.x; w[4 * h + 1] = v.y; w[4 * h + 2] = v.z; w[4 * h + 3] = v.w;
        }
        float4 yp[R], ym[R];  // p(j-1+m) and p(j-m), m = 1..R
#pragma unroll
        for (int m = 1; m <= R; ++m) {
            yp[m - 1] = ld4(F.p, q + (m - 1) * px);
            ym[m - 1] = ld4(F.p, q - m * px);
        }
        const float4 r0 = ld4(rho, q), rl = ld4(rho, q - 4), rym = ld4(rho, q - px);
        float4 vx = ld4(F.vx, q), vy = ld4(F.vy, q), vz = ld4(F.vz, q);
        const bool zslab = zslab_v(k, nz, fs);
        const bool vz_on = !fs || k >= 2;
        float4 mx = make_float4(0, 0, 0, 0), my = mx, mz = mx;
        if (anyx) mx = ld4(F.m[3], q);
        if (yslab) my = ld4(F.m[4], q);
        if (zslab) mz = ld4(F.m[5], q);
        const float azv = zslab ? az_h[k + pml - 1] : 0.0f, bzv = zslab ? bz_h[k + pml - 1] : 0.0f;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            float sx = c.c[0] * (w[A + e] - w[A + e - 1]);
            float sy = c.c[0] * (c4(yp[0], e) - c4(ym[0], e));
            float sz = c.c[0] * (c4(qz[R], e) - c4(qz[R - 1], e));
#pragma unroll
            for (int m = 2; m <= R; ++m) {
                sx = sx + c.c[m - 1] * (w[A + e + m - 1] - w[A + e - m]);
                sy = sy + c.c[m - 1] * (c4(yp[m - 1], e) - c4(ym[m - 1], e));
                sz = sz + c.c[m - 1] * (c4(qz[R - 1 + m], e) - c4(qz[R - m], e));
            }
            float dxv = idx * sx, dyv = idy * sy, dzv = idz * sz;
            const float rr = c4(r0, e);
            const float rx = 0.5f * (rr + (e == 0 ? rl.w : c4(r0, e - 1)));
            const float ry = 0.5f * (rr + c4(rym, e));
            const float rz = 0.5f * (rr + c4(rprev, e));
            if (xs[e]) {
                const float mem = axv[e] * dxv + bxv[e] * c4(mx, e);
                s4(mx, e, mem);
                dxv = dxv + mem;
            }
            if (yslab) {
                const float mem = ayv * dyv + byv * c4(my, e);
                s4(my, e, mem);
                dyv = dyv + mem;
            }
            if (zslab) {
                const float mem = azv * dzv + bzv * c4(mz, e);
                s4(mz, e, mem);
                dzv = dzv + mem;
            }
            if (x0 + e <= nx + pml) {
                s4(vx, e, c4(vx, e) - dt / rx * dxv);
                s4(vy, e, c4(vy, e) - dt / ry * dyv);
                if (vz_on) s4(vz, e, c4(vz, e) - dt / rz * dzv);
            }
        }
        rprev = r0;
        st4(F.vxo, q, vx);
        st4(F.vyo, q, vy);
        st4(F.vzo, q, vz);
        if (anyx) st4(F.m[3], q, mx);
        if (yslab) st4(F.m[4], q, my);
        if (zslab) st4(F.m[5], q, mz);
#pragma unroll
        for (int m = 0; m < 2 * R - 1; ++m) qz[m] = qz[m + 1];
    }
}


// ---------------------------------------------------------------------------
// Fused step (no free surface): the p-sweep and the v-sweep of one time step
// in one z-marching kernel, out of place (reads set A, writes set B), so a
// step reads p, vx, vy, vz, bk, rho once and writes p, vx, vy, vz once: 40 B
// per point instead of the two sweeps' 24 + 32.
//
// A block of kFW wavefronts owns a tile of W = 65 - 2R columns by TY rows
// and marches down a chunk of kc planes.  Its lanes compute the new pressure
// on 64 columns (the tile and R - 1 + R halo columns) and TY + 2R - 1 rows,
// R - 1 planes ahead of the velocities, which need the new pressure R - 1
// planes below.  Each new-pressure plane goes to an LDS ring of R + 1
// planes, where the velocity update reads its x and y stencils (one barrier
// per plane); the z stencils stay in per-thread register queues.  Row slots
// s < NO of wavefront w hold the owned rows Y0 + w + kFW*s, the other slots
// the halo rows, so every wavefront owns the same number of rows.  The halo
// pressure (and its PML memory, read only) is recomputed by both neighbouring
// tiles; only the owner writes it.  FP32 in the sweeps' operation order:
// results are bit-identical to k3c_update_p + k3c_update_v.
// ---------------------------------------------------------------------------
constexpr int kFW = 8, kFS = 4;

template <int R>
struct FusedShape {
    static constexpr int NH = (2 * R - 1 + kFW - 1) / kFW;  // halo slots per wavefront
    static constexpr int NO = kFS - NH;                      // owned slots per wavefront
    static constexpr int TY = kFW * NO;                      // owned rows per tile
    static constexpr int W = 65 - 2 * R;                     // owned columns per tile
    static constexpr int NROW = TY + 2 * R;                  // ring rows (p rows Y0-R .. Y0+TY+R-1)
};

// Field sets of the fused step: inputs (A) and outputs (B) of p, vx, vy, vz
// and the p-sweep's memory variables; the v-sweep's memory variables are
// touched by their owners only and stay in place.
struct F3F {
    const float* p;
    const float* vx;
    const float* vy;
    const float* vz;
    const float* pm[3];
    float* po;
    float* vxo;
    float* vyo;
    float* vzo;
    float* pmo[3];
    float* vm[3];
};

template <int R>
__global__ __launch_bounds__(64 * kFW) void k3f_step(const Shot3* __restrict__ shots,
                                                     const int* __restrict__ ftile_shot, F3F F,
                                                     const float* __restrict__ bk,
                                                     const float* __restrict__ rho,
                                                     const float* __restrict__ prof, Coef c, float dt, float idx,
                                                     float idy, float idz, int pml, int kc, int ntiles, int remap) {
    using FS = FusedShape<R>;
    constexpr int NO = FS::NO, TY = FS::TY, W = FS::W, NROW = FS::NROW, NSL = R + 1;
    __shared__ float ring[NSL][NROW][64];
    const int tile = xcd_tile(ntiles, remap);
    if (tile < 0) return;  // whole block
    const Shot3 S = shots[ftile_shot[tile]];
    int lt = tile - S.ftile0;
    const int tx = lt % S.fntx;
    lt /= S.fntx;
    const int ty = lt % S.fnty, tzc = lt / S.fnty;
    const int nx = S.nx, ny = S.ny, nz = S.nz;
    const int lo = 1 - pml - R;
    const int lane = int(threadIdx.x);
    const int w = __builtin_amdgcn_readfirstlane(int(threadIdx.y));
    const int X0 = -pml + 1 + tx * W, Y0 = -pml + 1 + ty * TY;
    const int x = X0 - R + lane;
    const bool xin = x >= -pml + 1 && x <= nx + pml;
    const bool xown = lane >= R && lane < R + W && xin;
    const int K0 = -pml + 1 + tzc * kc, K1 = min(K0 + kc - 1, nz + pml);
    const int lx = nx + 2 * pml, ly = ny + 2 * pml, lz = nz + 2 * pml;
    const float* az_i = prof + S.prof + 4 * lx + 4 * ly;
    const float* bz_i = az_i + lz;
    const float* az_h = az_i + 2 * lz;
    const float* bz_h = az_i + 3 * lz;
    const bool xsp = xin && (x <= 0 || x >= nx + 1), xsv = xown && (x <= 1 || x >= nx + 1);
    const float axi = xsp ? prof[S.prof + x + pml - 1] : 0.0f;
    const float bxi = xsp ? prof[S.prof + lx + x + pml - 1] : 0.0f;
    const float axh = xsv ? prof[S.prof + 2 * lx + x + pml - 1] : 0.0f;
    const float bxh = xsv ? prof[S.prof + 3 * lx + x + pml - 1] : 0.0f;
    const unsigned px = unsigned(S.px), pxy = unsigned(S.pxy);
    // row of each slot and its position u in the ring (u = y - (Y0 - R))
    int ys[kFS], us[kFS];
    bool yin[kFS];
    float ayi[kFS], byi[kFS];
    bool ysp[kFS];
#pragma unroll
    for (int s2 = 0; s2 < kFS; ++s2) {
        int u;
        if (s2 < NO) {
            u = R + w + kFW * s2;
        } else {
            const int h = w + kFW * (s2 - NO);
            u = h < R ? h : (h < 2 * R - 1 ? TY + h : -1);  // -1: a spare slot, no row
        }
        us[s2] = u;
        ys[s2] = Y0 - R + u;
        yin[s2] = u >= 0 && ys[s2] >= -pml + 1 && ys[s2] <= ny + pml;
        ysp[s2] = yin[s2] && (ys[s2] <= 0 || ys[s2] >= ny + 1);
        ayi[s2] = ysp[s2] ? prof[S.prof + 4 * lx + ys[s2] + pml - 1] : 0.0f;
        byi[s2] = ysp[s2] ? prof[S.prof + 4 * lx + ly + ys[s2] + pml - 1] : 0.0f;
    }
    // element offsets of (x, y, lo) per slot; meaningful only where the
    // slot's row and the lane's column are inside the update region
    // (batches are limited to 2^31 elements per field, so 32-bit offsets do)
    unsigned base[kFS];
#pragma unroll
    for (int s2 = 0; s2 < kFS; ++s2)
        base[s2] = unsigned(S.off) + unsigned(x - S.xb) + unsigned(max(ys[s2], lo) - lo) * unsigned(px);
    const int kfirst = max(K0 - R, -pml + 1);  // first plane with a pressure update
    float qz[kFS][2 * R];                      // vz(k-R+1 .. k+R) of each slot
#pragma unroll
    for (int s2 = 0; s2 < kFS; ++s2)
#pragma unroll
        for (int m = 0; m < 2 * R; ++m) qz[s2][m] = 0.0f;
#pragma unroll
    for (int s2 = 0; s2 < kFS; ++s2)
        if (yin[s2] && xin) {
#pragma unroll
            for (int m = 0; m < 2 * R - 1; ++m) qz[s2][m] = F.vz[base[s2] + unsigned(kfirst - R + 1 + m - lo) * pxy];
        }
    float pq[NO][2 * R];  // new p(m-R .. m+R-1) of the owned rows, m = k-R+1
    float rprev[NO];
#pragma unroll
    for (int s2 = 0; s2 < NO; ++s2) {
#pragma unroll
        for (int m = 0; m < 2 * R; ++m) pq[s2][m] = 0.0f;
        rprev[s2] = (yin[s2] && xown) ? rho[base[s2] + unsigned(K0 - 1 - lo) * pxy] : 0.0f;
    }
    int slot = 0;  // ring slot of plane k
    for (int k = K0 - R; k <= K1 + R - 1; ++k) {
        // ---- new pressure at plane k on every slot
        const bool kin = k >= kfirst && k <= nz + pml;
        const bool kown = k >= K0 && k <= K1;
        const bool zsp = kin && (k <= 0 || k >= nz + 1);
        const float azi = zsp ? az_i[k + pml - 1] : 0.0f, bzi = zsp ? bz_i[k + pml - 1] : 0.0f;
#pragma unroll
        for (int s2 = 0; s2 < kFS; ++s2) {
            float pn = 0.0f;
            if (kin && yin[s2] && xin) {
                const unsigned q = base[s2] + unsigned(k - lo) * pxy;
                qz[s2][2 * R - 1] = F.vz[q + R * pxy];
                const float* __restrict__ vx = F.vx;
                const float* __restrict__ vy = F.vy;
                float sx = c.c[0] * (vx[q + 1] - vx[q]);
                float sy = c.c[0] * (vy[q + px] - vy[q]);
                float sz = c.c[0] * (qz[s2][R] - qz[s2][R - 1]);
#pragma unroll
                for (int m = 2; m <= R; ++m) {
                    sx = sx + c.c[m - 1] * (vx[q + m] - vx[q - (m - 1)]);
                    sy = sy + c.c[m - 1] * (vy[q + m * px] - vy[q - (m - 1) * px]);
                    sz = sz + c.c[m - 1] * (qz[s2][R - 1 + m] - qz[s2][R - m]);
                }
                float dxv = idx * sx, dyv = idy * sy, dzv = idz * sz;
                const bool own = kown && xown && s2 < NO;
                if (xsp) {
                    const float mem = axi * dxv + bxi * F.pm[0][q];
                    if (own) F.pmo[0][q] = mem;
                    dxv = dxv + mem;
                }
                if (ysp[s2]) {
                    const float mem = ayi[s2] * dyv + byi[s2] * F.pm[1][q];
                    if (own) F.pmo[1][q] = mem;
                    dyv = dyv + mem;
                }
                if (zsp) {
                    const float mem = azi * dzv + bzi * F.pm[2][q];
                    if (own) F.pmo[2][q] = mem;
                    dzv = dzv + mem;
                }
                pn = F.p[q] - dt * bk[q] * (dxv + dyv + dzv);
                if (own) F.po[q] = pn;
#pragma unroll
                for (int m = 0; m < 2 * R - 1; ++m) qz[s2][m] = qz[s2][m + 1];
            }
            if (us[s2] >= 0) ring[slot][us[s2]][lane] = pn;
            if (s2 < NO) pq[s2][2 * R - 1] = pn;
            __builtin_amdgcn_sched_barrier(0);  // one slot's loads in flight at a time
        }
        __syncthreads();
        // ---- velocities at plane m = k - R + 1 on the owned rows
        const int m = k - R + 1;
        if (m >= K0) {
            const int sm = slot >= R - 1 ? slot - (R - 1) : slot + 2;  // ring slot of plane m
            const bool zsv = m <= 1 || m >= nz + 1;
            const float azv = zsv ? az_h[m + pml - 1] : 0.0f, bzv = zsv ? bz_h[m + pml - 1] : 0.0f;
#pragma unroll
            for (int s2 = 0; s2 < NO; ++s2) {
                if (!(yin[s2] && xown)) continue;
                const int u = us[s2];
                const float(*rg)[64] = ring[sm];
                float sx = c.c[0] * (rg[u][lane] - rg[u][lane - 1]);
                float sy = c.c[0] * (rg[u][lane] - rg[u - 1][lane]);
                float sz = c.c[0] * (pq[s2][R] - pq[s2][R - 1]);
#pragma unroll
                for (int mm = 2; mm <= R; ++mm) {
                    sx = sx + c.c[mm - 1] * (rg[u][lane + mm - 1] - rg[u][lane - mm]);
                    sy = sy + c.c[mm - 1] * (rg[u + mm - 1][lane] - rg[u - mm][lane]);
                    sz = sz + c.c[mm - 1] * (pq[s2][R - 1 + mm] - pq[s2][R - mm]);
                }
                float dxv = idx * sx, dyv = idy * sy, dzv = idz * sz;
                const unsigned q = base[s2] + unsigned(m - lo) * pxy;
                const float r0 = rho[q];
                const float rx = 0.5f * (r0 + rho[q - 1]);
                const float ry = 0.5f * (r0 + rho[q - px]);
                const float rz = 0.5f * (r0 + rprev[s2]);
                rprev[s2] = r0;
                if (xsv) {
                    const float mem = axh * dxv + bxh * F.vm[0][q];
                    F.vm[0][q] = mem;
                    dxv = dxv + mem;
                }
                const int y = ys[s2];
                if (y <= 1 || y >= ny + 1) {
                    const float mem = prof[S.prof + 4 * lx + 2 * ly + y + pml - 1] * dyv +
                                      prof[S.prof + 4 * lx + 3 * ly + y + pml - 1] * F.vm[1][q];
                    F.vm[1][q] = mem;
                    dyv = dyv + mem;
                }
                if (zsv) {
                    const float mem = azv * dzv + bzv * F.vm[2][q];
                    F.vm[2][q] = mem;
                    dzv = dzv + mem;
                }
                F.vxo[q] = F.vx[q] - dt / rx * dxv;
                F.vyo[q] = F.vy[q] - dt / ry * dyv;
                F.vzo[q] = F.vz[q] - dt / rz * dzv;
                __builtin_amdgcn_sched_barrier(0);
            }
        }
#pragma unroll
        for (int s2 = 0; s2 < NO; ++s2)
#pragma unroll
            for (int mm = 0; mm < 2 * R - 1; ++mm) pq[s2][mm] = pq[s2][mm + 1];
        slot = slot == R ? 0 : slot + 1;
    }
}

// Source injection of step t (add_source, module_wavefield.f90:322-444:
// explosions into p, forces into vx, vy, vz) then receivers of step t+1;
// one block per shot so that the samples see the injected source.  Taps are
// grouped by (field, cell); the contributions to one cell are added in
// source order.
__global__ __launch_bounds__(1024) void k3_post_step(const Shot3* __restrict__ shots, float* p, float* vx,
                                                     float* vy, float* vz, const int* __restrict__ inj_beg,
                                                     const Tap3* __restrict__ inj,
                                                     const float* __restrict__ amp_tab, int nt, int t,
                                                     const int* __restrict__ rec_beg,
                                                     const Tap3* __restrict__ rec,
                                                     const float* __restrict__ rec_w, float* traces) {
    const Shot3 S = shots[blockIdx.x];
    for (int c = threadIdx.x; c < S.ninj; c += blockDim.x) {
        const int b = inj_beg[S.inj0 + c], e = inj_beg[S.inj0 + c + 1];
        const long long cell = inj[b].cell;
        const int fld = inj[b].field;
        float* f = fld == 0 ? p : (fld == 1 ? vx : (fld == 2 ? vy : vz));
        float acc = f[cell];
        for (int k = b; k < e; ++k) {
            const Tap3 tp = inj[k];
            if (t < tp.tbeg || t > tp.tend) continue;
            acc = acc + amp_tab[size_t(tp.row) * nt + (t - 1)] * tp.w1 * tp.w2 * tp.w3;
        }
        f[cell] = acc;
    }
    __syncthreads();
    if (t + 1 > nt || !traces) return;
    for (int r = threadIdx.x; r < S.nr; r += blockDim.x) {
        const int g = S.rec0 + r;
        const float w = rec_w[g];
        float acc = 0.0f;
        for (int k = rec_beg[g]; k < rec_beg[g + 1]; ++k) {
            const Tap3 tp = rec[k];
            acc = acc + p[tp.cell] * tp.w1 * tp.w2 * tp.w3 * w;
        }
        traces[S.trace0 + size_t(t) * S.nr + r] = acc;
    }
}

// Free surface: p(:,:,k) = -p(:,:,2-k) for k <= 0 (module_wavefield.f90:240-250).
template <int R>
__global__ void k3_mirror_p(const Shot3* __restrict__ shots, float* p, int pml) {
    const Shot3 S = shots[blockIdx.y];
    const int lo = 1 - pml - R, lx = S.nx + 2 * pml, ly = S.ny + 2 * pml;
    const long long n = (long long)lx * ly * pml;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const int i = -pml + 1 + int(e % lx), j = -pml + 1 + int((e / lx) % ly), k = -pml + 1 + int(e / (lx * ly));
        p[at3(S, lo, i, j, k)] = -p[at3(S, lo, i, j, 2 - k)];
    }
}

// Boundary strips (acoustic_iso_3d/module_boundary_saving.f90:176-207): vx
// on x in [2-R, 1] and [nx+1, nx+R], vy on the y faces, vz on the z faces,
// each over [1-R, n+R] on the other two axes.  Record of step t: the six
// sections in order, first index fastest.
struct StripGeo {
    long long a, b, c;  // section sizes: x faces R*ey*ez, y faces ex*R*ez, z faces ex*ey*R
    int ex, ey;
};

__device__ __forceinline__ StripGeo strip_geo(const Shot3& S, int R) {
    StripGeo g;
    g.ex = S.nx + 2 * R;
    g.ey = S.ny + 2 * R;
    const int ez = S.nz + 2 * R;
    g.a = (long long)R * g.ey * ez;
    g.b = (long long)g.ex * R * ez;
    g.c = (long long)g.ex * g.ey * R;
    return g;
}

// Node of strip element e; returns the field (1 vx, 2 vy, 3 vz) and, in
// *edge, the slot of the node in the saved interior faces (or -1).
__device__ __forceinline__ int strip_node(const Shot3& S, const StripGeo& g, int R, long long e, int& i, int& j,
                                          int& k, long long& edge) {
    const int nx = S.nx, ny = S.ny, nz = S.nz;
    edge = -1;
    if (e < 2 * g.a) {
        const bool hi = e >= g.a;
        const long long r = hi ? e - g.a : e;
        i = (hi ? nx + 1 : 2 - R) + int(r % R);
        j = 1 - R + int((r / R) % g.ey);
        k = 1 - R + int(r / (R * (long long)g.ey));
        if (i == (hi ? nx + 1 : 1) && j >= 1 && j <= ny && k >= 1 && k <= nz)
            edge = (hi ? (long long)ny * nz : 0) + (j - 1) + (long long)(k - 1) * ny;
        return 1;
    }
    e -= 2 * g.a;
    if (e < 2 * g.b) {
        const bool hi = e >= g.b;
        const long long r = hi ? e - g.b : e;
        i = 1 - R + int(r % g.ex);
        j = (hi ? ny + 1 : 2 - R) + int((r / g.ex) % R);
        k = 1 - R + int(r / (g.ex * (long long)R));
        if (j == (hi ? ny + 1 : 1) && i >= 1 && i <= nx && k >= 1 && k <= nz)
            edge = 2LL * ny * nz + (hi ? (long long)nx * nz : 0) + (i - 1) + (long long)(k - 1) * nx;
        return 2;
    }
    e -= 2 * g.b;
    const bool hi = e >= g.c;
    const long long r = hi ? e - g.c : e;
    i = 1 - R + int(r % g.ex);
    j = 1 - R + int((r / g.ex) % g.ey);
    k = (hi ? nz + 1 : 2 - R) + int(r / (g.ex * (long long)g.ey));
    if (k == (hi ? nz + 1 : 1) && i >= 1 && i <= nx && j >= 1 && j <= ny)
        edge = 2LL * ny * nz + 2LL * nx * nz + (hi ? (long long)nx * ny : 0) + (i - 1) + (long long)(j - 1) * nx;
    return 3;
}

// save != 0: strips of step t <- fields; else fields <- strips of step t,
// keeping the overwritten interior-face values in edge (the imaging needs
// the state before injection there).
template <int R>
__global__ void k3_strips(const Shot3* __restrict__ shots, float* vx, float* vy, float* vz, float* strips,
                          float* edge, int t, int pml, int save) {
    const Shot3 S = shots[blockIdx.y];
    const StripGeo g = strip_geo(S, R);
    const long long len = 2 * (g.a + g.b + g.c);
    const int lo = 1 - pml - R;
    float* rec = strips + S.strip0 + (long long)(t - 1) * len;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < len; e += (long long)gridDim.x * blockDim.x) {
        int i, j, k;
        long long ed;
        const int f = strip_node(S, g, R, e, i, j, k, ed);
        float* dst = f == 1 ? vx : (f == 2 ? vy : vz);
        const size_t q = at3(S, lo, i, j, k);
        if (save) {
            rec[e] = dst[q];
        } else {
            if (ed >= 0) edge[S.edge0 + ed] = dst[q];
            dst[q] = rec[e];
        }
    }
}

// Adjoint source injection (submodule_fwi_gradient.f90:170-197): cells in
// CSR order, contributions of one cell in receiver order.
__global__ void k3_inject_adjoint(const Shot3* __restrict__ shots, float* pr, const int* __restrict__ cell_beg,
                                  const Tap3* __restrict__ taps, const float* __restrict__ rec_w,
                                  const float* __restrict__ adj, int t) {
    const Shot3 S = shots[blockIdx.y];
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < S.nadj; c += gridDim.x * blockDim.x) {
        const int b = cell_beg[S.adj0 + c], e = cell_beg[S.adj0 + c + 1];
        const long long cell = taps[b].cell;
        float acc = pr[cell];
        for (int k = b; k < e; ++k) {
            const Tap3 tp = taps[k];
            const float a = adj[S.trace0 + size_t(t - 1) * S.nr + tp.row];
            acc = acc + a * tp.w1 * tp.w2 * tp.w3 * rec_w[S.rec0 + tp.row];
        }
        pr[cell] = acc;
    }
}

// Imaging condition (compute_gradient, submodule_fwi_gradient.f90:276-361,
// energy terms off) over the window interior:
//   g_vp  -= (p - prev_p) pr / (rho vp^3)
//   g_rho += sum over x, y, z of (v(i)+v(i+1) - prev)(vr(i)+vr(i+1))
// prev = 0 on the first reverse step; interior-face prev velocities from
// edge.  FP64 accumulators; window gradient x fastest.
template <int R>
__global__ __launch_bounds__(256) void k3_image_owl(const Shot3* __restrict__ shots, F3 cur, F3 prev,
                                                    const float* __restrict__ edge, F3 adj,
                                                    const float* __restrict__ den, double* gv, double* gr,
                                                    int prev_zero, int pml, int t) {
    const Shot3 S = shots[blockIdx.y];
    if (t < S.hnt) return;  // this shot's reverse loop has ended (whole block)
    const int nx = S.nx, ny = S.ny, nz = S.nz, lo = 1 - pml - R;
    const long long n = (long long)nx * ny * nz;
    const size_t px = S.px, pxy = S.pxy;
    const float* eg = edge + S.edge0;
    const long long oy = 2LL * ny * nz, oz = oy + 2LL * nx * nz;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const int i = 1 + int(e % nx), j = 1 + int((e / nx) % ny), k = 1 + int(e / ((long long)nx * ny));
        const size_t q = at3(S, lo, i, j, k);
        const float pv = prev_zero ? 0.0f : prev.p[q];
        const float sp = cur.p[q] - pv;
        const float inc_v = sp * adj.p[q] / den[S.g0 + e];
        float x0 = 0.f, x1 = 0.f, y0 = 0.f, y1 = 0.f, z0 = 0.f, z1 = 0.f;
        if (!prev_zero) {
            x0 = (i == 1) ? eg[(j - 1) + (long long)(k - 1) * ny] : prev.vx[q];
            x1 = (i == nx) ? eg[(long long)ny * nz + (j - 1) + (long long)(k - 1) * ny] : prev.vx[q + 1];
            y0 = (j == 1) ? eg[oy + (i - 1) + (long long)(k - 1) * nx] : prev.vy[q];
            y1 = (j == ny) ? eg[oy + (long long)nx * nz + (i - 1) + (long long)(k - 1) * nx] : prev.vy[q + px];
            z0 = (k == 1) ? eg[oz + (i - 1) + (long long)(j - 1) * nx] : prev.vz[q];
            z1 = (k == nz) ? eg[oz + (long long)nx * ny + (i - 1) + (long long)(j - 1) * nx] : prev.vz[q + pxy];
        }
        const float sx = (cur.vx[q] + cur.vx[q + 1]) - (x0 + x1);
        const float rx = adj.vx[q] + adj.vx[q + 1];
        const float sy = (cur.vy[q] + cur.vy[q + px]) - (y0 + y1);
        const float ry = adj.vy[q] + adj.vy[q + px];
        const float sz = (cur.vz[q] + cur.vz[q + pxy]) - (z0 + z1);
        const float rz = adj.vz[q] + adj.vz[q + pxy];
        gv[S.g0 + e] = gv[S.g0 + e] - double(inc_v);
        gr[S.g0 + e] = gr[S.g0 + e] + double(sx * rx) + double(sy * ry) + double(sz * rz);
    }
}

// L2 misfit and adjoint source per shot (compute_adjsrc_waveform,
// module_inversion_adjoint_source.f90:371-383; trace discard rule
// :1042-1054): traces [nt][nr] per shot; FP64 Neumaier sums.
__global__ __launch_bounds__(256) void k3_adjoint_source(const Shot3* __restrict__ shots,
                                                         const float* __restrict__ syn,
                                                         const float* __restrict__ obs,
                                                         const float* __restrict__ rec_w, float* adj,
                                                         double* misfit, int nt, float threshold,
                                                         double* scratch) {
    const Shot3 S = shots[blockIdx.x];
    __shared__ double red[256];
    __shared__ double mo_s, ms_s;
    double* nobs = scratch + 2 * size_t(S.rec0);
    double* nsyn = nobs + S.nr;
    double lmo = 0, lms = 0;
    for (int r = threadIdx.x; r < S.nr; r += blockDim.x) {
        double a = 0, b = 0;
        for (int t = 0; t < nt; ++t) {
            const double o = obs[S.trace0 + size_t(t) * S.nr + r];
            const double sv = syn[S.trace0 + size_t(t) * S.nr + r];
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
    double sum = 0, comp = 0;
    for (int r = threadIdx.x; r < S.nr; r += blockDim.x) {
        const bool keep = !(rec_w[S.rec0 + r] == 0.0f || nobs[r] < threshold * mo || nsyn[r] < threshold * ms ||
                            nobs[r] * nsyn[r] == 0.0);
        for (int t = 0; t < nt; ++t) {
            const size_t q = S.trace0 + size_t(t) * S.nr + r;
            if (keep) {
                const float d = obs[q] - syn[q];
                const double d2 = double(d) * double(d);
                const double y = sum + d2;
                comp += (fabs(sum) >= fabs(d2)) ? (sum - y) + d2 : (d2 - y) + sum;
                sum = y;
                adj[q] = syn[q] - obs[q];
            } else {
                adj[q] = 0.0f;
            }
        }
    }
    red[threadIdx.x] = sum + comp;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) misfit[blockIdx.x] = red[0];
}

// Sum of the per-shot window gradients on the full model (nz, ny, nx), z
// fastest (module_gradient.f90:366-388 in 3D).
__global__ void k3_merge(const Shot3* __restrict__ shots, int nshot, const double* __restrict__ gw, double* full,
                         int nx, int ny, int nz) {
    const long long n = (long long)nx * ny * nz;
    for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n; e += (long long)gridDim.x * blockDim.x) {
        const int kz = int(e % nz), jy = int((e / nz) % ny), ix = int(e / ((long long)nz * ny));
        double acc = 0.0;
        for (int s = 0; s < nshot; ++s) {
            const Shot3& S = shots[s];
            const int i = ix - (S.nxbeg - 1), j = jy - (S.nybeg - 1), k = kz - (S.nzbeg - 1);
            if (i < 0 || i >= S.nx || j < 0 || j >= S.ny || k < 0 || k >= S.nz) continue;
            acc += gw[S.g0 + i + (long long)j * S.nx + (long long)k * S.nx * S.ny];
        }
        full[e] = acc;
    }
}

struct ShotHost3 {
    int id = 0;
    int nx = 0, ny = 0, nz = 0, nxbeg = 1, nybeg = 1, nzbeg = 1;
    float ox = 0, oy = 0, oz = 0;
    int hnt = 1;
    std::vector<owl_source> src;
    std::vector<owl_receiver> rec;
};

}  // namespace a3
}  // namespace owl

using namespace owl;
using namespace owl::a3;

struct owl_acoustic3d {
    owl_acoustic_config cfg{};
    int R = 4, pml = 15, nt = 0, kc = 64;
    int variant = 1;   // 0: LDS planes with barriers, 1: barrier-free columns, 2: four-wide, 3: pipelined columns
    int cy = 4;        // rows per block of the column variant
    int remap = 0;     // XCD-contiguous tile order
    int launch_code = 1;  // the public variant code of the current shape
    bool fused = false;   // forward steps by the fused kernel (k3f_step; no free surface)
    int fkc = 64;         // z planes per tile of the fused kernel
    std::vector<int> ftile_shot;
    DevBuf<int> d_ftile_shot;
    DevBuf<float> mB[3];  // second set of the p-sweep memory variables (fused ping-pong)
    int fs = 0;        // free surface at z index 1
    float dt = 0, idx = 0, idy = 0, idz = 0;
    a3::Coef coef{};
    std::vector<ShotHost3> hs;
    std::vector<Shot3> sd;
    std::vector<int> tile_shot;
    size_t field_elems = 0, trace_elems = 0, strip_elems = 0, edge_elems = 0, grad_elems = 0;
    int nrec_total = 0, max_ninj = 0, max_nadj = 0;
    double pts_full = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0, last_updates = 0;
    Profiler prof;
    std::vector<float> vp_full, rho_full;
    DevBuf<Shot3> d_shots;
    DevBuf<int> d_tile_shot, d_inj_beg, d_rec_beg, d_adj_beg;
    DevBuf<float> d_bk, d_rho, d_den, d_prof, d_amp, d_rec_w, d_traces, d_obs, d_adj;
    DevBuf<float> f[4], m[6];            // forward state (p, vx, vy, vz) and its memory variables
    DevBuf<float> fB[4];                 // second forward state of the reverse pass (ping-pong)
    DevBuf<float> aF[4], aM[6];          // adjoint state
    DevBuf<float> d_strips, d_edge;
    DevBuf<double> d_gv, d_gr, d_gfull_v, d_gfull_r, d_misfit, d_scratch;
    DevBuf<Tap3> d_inj, d_rec, d_adj_taps;
    bool have_boundary = false, fwd_state_valid = false, have_obs = false;
    // HIP graph of the forward time loop (modelling without boundary
    // records, profiling off): captured by the first such call, replayed by
    // the next ones while the launch shape and buffers stay the same
    LoopGraph graph;
    ~owl_acoustic3d() {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

inline size_t cell3(const owl_acoustic3d* b, const Shot3& S, int i, int j, int k) {
    const int lo = 1 - b->pml - b->R;
    const int h = b->pml + b->R;
    if (i < lo || i > S.nx + h || j < lo || j > S.ny + h || k < lo || k > S.nz + h)
        fail(OWL_ERR_ARG, "a source/receiver stencil reaches outside the padded grid; increase pml");
    return size_t(S.off) + size_t(i - S.xb) + size_t(j - lo) * size_t(S.px) + size_t(k - lo) * size_t(S.pxy);
}

inline float model_at(const owl_acoustic3d* b, const std::vector<float>& mdl, const ShotHost3& H, int i, int j,
                      int k) {
    const int ic = std::min(std::max(i, 1), H.nx), jc = std::min(std::max(j, 1), H.ny),
              kc = std::min(std::max(k, 1), H.nz);
    const size_t nz = size_t(b->cfg.nz), ny = size_t(b->cfg.ny);
    return mdl[size_t(H.nzbeg - 1 + kc - 1) + size_t(H.nybeg - 1 + jc - 1) * nz +
               size_t(H.nxbeg - 1 + ic - 1) * nz * ny];
}

// Group taps by (field, cell) in CSR form, keeping their order within a cell.
int csr_by_cell(std::vector<Tap3>& st, std::vector<int>& beg, std::vector<Tap3>& out) {
    std::stable_sort(st.begin(), st.end(), [](const Tap3& x, const Tap3& y) {
        return x.field != y.field ? x.field < y.field : x.cell < y.cell;
    });
    int ncell = 0;
    for (size_t k = 0; k < st.size(); ++k) {
        if (k == 0 || st[k].cell != st[k - 1].cell || st[k].field != st[k - 1].field) {
            if (k) beg.push_back(int(out.size()));
            ++ncell;
        }
        out.push_back(st[k]);
    }
    if (!st.empty()) beg.push_back(int(out.size()));
    return ncell;
}

// Owned columns and rows per tile of the fused step for radius R.
void fused_shape(int R, int& W, int& TY) {
    switch (R) {
#define OWL_FS_CASE(RR)                  \
    case RR:                             \
        W = FusedShape<RR>::W;           \
        TY = FusedShape<RR>::TY;         \
        break;
        OWL_FS_CASE(2)
        OWL_FS_CASE(3)
        OWL_FS_CASE(4)
        OWL_FS_CASE(5)
        OWL_FS_CASE(6)
        OWL_FS_CASE(7)
        OWL_FS_CASE(8)
#undef OWL_FS_CASE
        default: fail(OWL_ERR_ARG, "fdhalf must be 2..8");
    }
}

// Tile table of the current launch shape: per shot x-y tiles of the
// variant's block times z segments of kc planes, all shots in one grid.
void retile(owl_acoustic3d* b) {
    const int pml = b->pml;
    b->tile_shot.clear();
    const int ty_rows = b->variant == 0 ? kTY : (b->variant == 2 ? kWY : b->cy);
    const int tx_cols = b->variant == 2 ? 4 * kWX : kBX;
    for (size_t s = 0; s < b->sd.size(); ++s) {
        Shot3& S = b->sd[s];
        S.nty = (S.ny + 2 * pml + ty_rows - 1) / ty_rows;
        S.ntx = (S.nx + 2 * pml + tx_cols - 1) / tx_cols;
        S.nzc = (S.nz + 2 * pml + b->kc - 1) / b->kc;
        S.tile0 = int(b->tile_shot.size());
        b->tile_shot.insert(b->tile_shot.end(), size_t(S.ntx) * S.nty * S.nzc, int(s));
    }
    b->ftile_shot.clear();
    if (b->R <= 8) {
        int W = 0, TY = 0;
        fused_shape(b->R, W, TY);
        for (size_t s = 0; s < b->sd.size(); ++s) {
            Shot3& S = b->sd[s];
            S.fntx = (S.nx + 2 * pml + W - 1) / W;
            S.fnty = (S.ny + 2 * pml + TY - 1) / TY;
            S.fnzc = (S.nz + 2 * pml + b->fkc - 1) / b->fkc;
            S.ftile0 = int(b->ftile_shot.size());
            b->ftile_shot.insert(b->ftile_shot.end(), size_t(S.fntx) * S.fnty * S.fnzc, int(s));
        }
    }
    if (b->ftile_shot.empty()) b->ftile_shot.push_back(0);
    b->d_shots.upload(b->sd, b->stream);
    b->d_tile_shot.upload(b->tile_shot, b->stream);
    b->d_ftile_shot.upload(b->ftile_shot, b->stream);
    OWL_HIP(hipStreamSynchronize(b->stream));
}

void plan3(owl_acoustic3d* b, const owl_acoustic_config* cfg, const float* vp, const float* rho, int nshot,
           const owl_shot* shots) {
    if (!cfg || !vp || !rho || nshot < 1 || !shots) fail(OWL_ERR_ARG, "null argument or no shots");
    if (!fd_coefficients(cfg->fdhalf, b->coef.c)) fail(OWL_ERR_ARG, "fdhalf must be 2..8");
    if (cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 1 || !(cfg->dx > 0) || !(cfg->dy > 0) || !(cfg->dz > 0) ||
        !(cfg->dt > 0))
        fail(OWL_ERR_ARG, "invalid grid or dt");
    if (cfg->pml < cfg->fdhalf) fail(OWL_ERR_ARG, "pml must be >= fdhalf");
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
    b->rho_full.assign(rho, rho + nfull);
    const float f0factor = cfg->f0factor > 0 ? cfg->f0factor : 1.0f;
    const int R = b->R, pml = b->pml, nt = b->nt;
    b->hs.assign(nshot, ShotHost3{});
    b->sd.assign(nshot, Shot3{});
    b->tile_shot.clear();
    b->nrec_total = 0;
    b->max_ninj = b->max_nadj = 0;
    b->pts_full = 0;
    b->have_boundary = b->fwd_state_valid = b->have_obs = false;
    long long off = 0, trace = 0, strip = 0, edge = 0, grad = 0;
    int prof = 0;
    std::vector<float> profs, amp, rec_w;
    std::vector<int> inj_beg{0}, rec_beg{0}, adj_beg{0};
    std::vector<Tap3> inj, rec, adjt;
    for (int s = 0; s < nshot; ++s) {
        const owl_shot& in = shots[s];
        ShotHost3& H = b->hs[s];
        Shot3& S = b->sd[s];
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
        float vmax = 0.f, pmlvp = 0.f, f0max = 0.f;
        for (int i = 1; i <= H.nx; ++i)
            for (int j = 1; j <= H.ny; ++j)
                for (int k = 1; k <= H.nz; ++k) {
                    const float v = model_at(b, b->vp_full, H, i, j, k);
                    if (!(v > 0.f) || !(model_at(b, b->rho_full, H, i, j, k) > 0.f))
                        fail(OWL_ERR_ARG, "vp and rho must be positive");
                    vmax = std::max(vmax, v);
                    if (i == 1 || i == H.nx || j == 1 || j == H.ny || k == 1 || k == H.nz) pmlvp = std::max(pmlvp, v);
                }
        const float dts = dt_stable(R, 3, vmax, cfg->dx, cfg->dy, cfg->dz);
        if (cfg->dt > dts)
            fail(OWL_ERR_CFL, "dt = " + std::to_string(cfg->dt) + " exceeds the stable dt = " + std::to_string(dts));
        for (auto& sc : H.src) f0max = std::max(f0max, sc.f0);
        const int NYp = H.ny + 2 * (pml + R), NZp = H.nz + 2 * (pml + R);
        S.nx = H.nx;
        S.ny = H.ny;
        S.nz = H.nz;
        // x positions start at xb = -pml+1-A (A = R rounded up to 4): the first
        // updated node sits 16-byte aligned and float4 reads of the halo stay
        // in the row; 4 spare floats on the right
        const int A = (R + 3) / 4 * 4;
        S.xb = -pml + 1 - A;
        S.px = (H.nx + 2 * pml + 2 * A + 4 + kPitchAlign - 1) / kPitchAlign * kPitchAlign;
        S.nyp = NYp;
        S.nzp = NZp;
        S.pxy = (long long)S.px * NYp;
        S.off = off;
        off += S.pxy * NZp;
        off = (off + 63) / 64 * 64;
        b->pts_full += double(H.nx + 2 * pml) * double(H.ny + 2 * pml) * double(H.nz + 2 * pml);
        S.prof = prof;
        const PmlAxis PX = cfspml_axis(H.nx, pml, cfg->dx, cfg->dt, pmlvp, f0max);
        const PmlAxis PY = cfspml_axis(H.ny, pml, cfg->dy, cfg->dt, pmlvp, f0max);
        const PmlAxis PZ = cfspml_axis(H.nz, pml, cfg->dz, cfg->dt, pmlvp, f0max);
        for (const PmlAxis* P : {&PX, &PY, &PZ})
            for (auto* v : {&P->ai, &P->bi, &P->ah, &P->bh}) profs.insert(profs.end(), v->begin(), v->end());
        prof += 4 * (H.nx + 2 * pml) + 4 * (H.ny + 2 * pml) + 4 * (H.nz + 2 * pml);
        S.nxbeg = H.nxbeg;
        S.nybeg = H.nybeg;
        S.nzbeg = H.nzbeg;
        S.nr = int(H.rec.size());
        S.trace0 = trace;
        trace += (long long)nt * S.nr;
        // boundary records, saved interior faces, window gradient
        const long long ex = H.nx + 2 * R, ey = H.ny + 2 * R, ez = H.nz + 2 * R;
        S.strip0 = strip;
        strip += (long long)nt * 2 * R * (ey * ez + ex * ez + ex * ey);
        S.edge0 = edge;
        edge += 2LL * H.ny * H.nz + 2LL * H.nx * H.nz + 2LL * H.nx * H.ny;
        S.g0 = grad;
        grad += (long long)H.nx * H.ny * H.nz;
        // sources (add_source, acoustic_iso_3d/module_wavefield.f90:322-444)
        const float xmax = float(H.nx - 1) * cfg->dx + H.ox, ymax = float(H.ny - 1) * cfg->dy + H.oy,
                    zmax = float(H.nz - 1) * cfg->dz + H.oz;
        std::vector<Tap3> st;
        H.hnt = 0;
        for (size_t si = 0; si < H.src.size(); ++si) {
            const owl_source& sc = H.src[si];
            if (sc.mechanism != OWL_EXPLOSION && sc.mechanism != OWL_FORCE)
                fail(OWL_ERR_ARG, "the acoustic solver takes explosion and force sources");
            if (!(sc.x >= H.ox && sc.x <= xmax && sc.y >= H.oy && sc.y <= ymax && sc.z >= H.oz && sc.z <= zmax))
                fail(OWL_ERR_ARG, "source outside its shot window");
            const Located L = locate(3, sc.x, sc.y, sc.z, H.ox, H.oy, H.oz, cfg->dx, cfg->dy, cfg->dz);
            int hnt = 0;
            const std::vector<float> stf = source_time_function(sc.wavelet, sc.f0, cfg->dt, f0factor, &hnt);
            if (si == 0) {
                H.hnt = hnt;
                S.hnt = std::max(hnt, 1);
            }
            const int nbeg = fortran_nint(sc.t0 / cfg->dt) + 1, nend = nbeg + int(stf.size()) - 1;
            std::vector<float> a(nt, 0.f);
            for (int t = std::max(nbeg, 1); t <= std::min(nend, nt); ++t) a[t - 1] = stf[t - nbeg] * sc.amp * cfg->dt;
            auto add_taps = [&](int field, int gx, int gy, int gz, const float* wx, const float* wy, const float* wz,
                                int row, int zcell_min) {
                for (int irz = -kNkw; irz <= kNkw; ++irz)
                    for (int iry = -kNkw; iry <= kNkw; ++iry)
                        for (int irx = -kNkw; irx <= kNkw; ++irx) {
                            const float w1 = wx[irx + kNkw], w2 = wy[iry + kNkw], w3 = wz[irz + kNkw];
                            if (w1 == 0.f || w2 == 0.f || w3 == 0.f) continue;
                            if (b->fs && gz + irz < zcell_min) continue;
                            st.push_back(Tap3{(long long)cell3(b, S, gx + irx, gy + iry, gz + irz), w1, w2, w3, row,
                                              nbeg, nend, field});
                        }
            };
            if (sc.mechanism == OWL_EXPLOSION) {
                const int row = int(amp.size() / nt);
                amp.insert(amp.end(), a.begin(), a.end());
                add_taps(0, L.gx, L.gy, L.gz, L.ix, L.iy, L.iz, row, 2);
            } else {
                // amp / rho(gx, gy, gz), then sin(polar) cos(azimuth), sin(polar)
                // sin(azimuth), cos(polar) for vx, vy, vz (left to right)
                const float rs = model_at(b, b->rho_full, H, L.gx, L.gy, L.gz);
                const float sp = std::sin(sc.polar), cp = std::cos(sc.polar);
                const float sa = std::sin(sc.azimuth), ca = std::cos(sc.azimuth);
                const float fac[3] = {sp * ca, sp * sa, cp};
                int rows[3];
                for (int c = 0; c < 3; ++c) {
                    rows[c] = int(amp.size() / nt);
                    for (int t = 0; t < nt; ++t) amp.push_back(fac[c] * (a[t] / rs));
                }
                add_taps(1, L.hx, L.gy, L.gz, L.hxw, L.iy, L.iz, rows[0], 2);
                add_taps(2, L.gx, L.hy, L.gz, L.ix, L.hyw, L.iz, rows[1], 2);
                add_taps(3, L.gx, L.gy, L.hz, L.ix, L.iy, L.hzw, rows[2], 2);
            }
        }
        S.inj0 = int(inj_beg.size()) - 1;
        S.ninj = csr_by_cell(st, inj_beg, inj);
        b->max_ninj = std::max(b->max_ninj, S.ninj);
        // receivers (outside the window: weight 0)
        S.rec0 = b->nrec_total;
        std::vector<Tap3> at;
        for (int r = 0; r < S.nr; ++r) {
            const owl_receiver& rc = H.rec[r];
            const bool valid = rc.x >= H.ox && rc.x <= xmax && rc.y >= H.oy && rc.y <= ymax && rc.z >= H.oz &&
                               rc.z <= zmax;
            const float w = (valid && rc.weight != 0.f) ? rc.weight : 0.f;
            rec_w.push_back(w);
            if (w != 0.f) {
                const Located L = locate(3, rc.x, rc.y, rc.z, H.ox, H.oy, H.oz, cfg->dx, cfg->dy, cfg->dz);
                for (int irz = -kNkw; irz <= kNkw; ++irz)
                    for (int iry = -kNkw; iry <= kNkw; ++iry)
                        for (int irx = -kNkw; irx <= kNkw; ++irx) {
                            const float w1 = L.ix[irx + kNkw], w2 = L.iy[iry + kNkw], w3 = L.iz[irz + kNkw];
                            if (w1 == 0.f || w2 == 0.f || w3 == 0.f) continue;
                            const Tap3 tp{(long long)cell3(b, S, L.gx + irx, L.gy + iry, L.gz + irz), w1, w2, w3, r,
                                          0, 0, 0};
                            rec.push_back(tp);
                            at.push_back(tp);
                        }
            }
            rec_beg.push_back(int(rec.size()));
        }
        // adjoint-source taps: the receivers' cells, contributions in receiver order
        S.adj0 = int(adj_beg.size()) - 1;
        S.nadj = csr_by_cell(at, adj_beg, adjt);
        b->max_nadj = std::max(b->max_nadj, S.nadj);
        b->nrec_total += S.nr;
    }
    b->field_elems = size_t(off);
    b->trace_elems = size_t(trace);
    b->strip_elems = size_t(strip);
    b->edge_elems = size_t(edge);
    b->grad_elems = size_t(grad);
    for (auto& S : b->sd) {
        const bool ok = size_t(S.rec0) + size_t(S.nr) <= rec_w.size() && size_t(S.rec0) + S.nr + 1 <= rec_beg.size() &&
                        size_t(S.inj0) + size_t(S.ninj) + (S.ninj ? 1 : 0) <= inj_beg.size() &&
                        size_t(S.adj0) + size_t(S.nadj) + (S.nadj ? 1 : 0) <= adj_beg.size() &&
                        size_t(S.trace0) + size_t(nt) * S.nr <= b->trace_elems &&
                        size_t(S.off) + size_t(S.pxy) * S.nzp <= b->field_elems;
        if (!ok) fail(OWL_ERR_ARG, "internal layout check failed");
    }
    if (!b->stream) {
        OWL_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        OWL_HIP(hipEventCreate(&b->ev0));
        OWL_HIP(hipEventCreate(&b->ev1));
    }
    hipStream_t stm = b->stream;
    retile(b);
    b->d_prof.upload(profs, stm);
    if (amp.empty()) amp.push_back(0.f);
    b->d_amp.upload(amp, stm);
    b->d_inj_beg.upload(inj_beg, stm);
    if (inj.empty()) inj.push_back(Tap3{});
    b->d_inj.upload(inj, stm);
    b->d_rec_beg.upload(rec_beg, stm);
    if (rec.empty()) rec.push_back(Tap3{});
    b->d_rec.upload(rec, stm);
    b->d_rec_w.upload(rec_w, stm);
    b->d_adj_beg.upload(adj_beg, stm);
    if (adjt.empty()) adjt.push_back(Tap3{});
    b->d_adj_taps.upload(adjt, stm);
    // padded models (edge replication) -- permute(...,321): x fastest; and
    // rho vp^3 on each window's interior for the imaging condition
    std::vector<float> bk(b->field_elems, 0.f), rh(b->field_elems, 0.f), den(std::max<size_t>(b->grad_elems, 1), 0.f);
    const int lo = 1 - pml - R;
    for (size_t s = 0; s < b->hs.size(); ++s) {
        const ShotHost3& H = b->hs[s];
        const Shot3& S = b->sd[s];
        for (int k = lo; k < lo + S.nzp; ++k)
            for (int j = lo; j < lo + S.nyp; ++j)
                for (int i = S.xb; i < S.xb + S.px; ++i) {
                    const float v = model_at(b, b->vp_full, H, i, j, k), r = model_at(b, b->rho_full, H, i, j, k);
                    const size_t q = size_t(S.off) + size_t(i - S.xb) + size_t(j - lo) * size_t(S.px) +
                                     size_t(k - lo) * size_t(S.pxy);
                    rh[q] = r;
                    bk[q] = r * (v * v);
                }
        for (int k = 1; k <= H.nz; ++k)
            for (int j = 1; j <= H.ny; ++j)
                for (int i = 1; i <= H.nx; ++i) {
                    const float v = model_at(b, b->vp_full, H, i, j, k), r = model_at(b, b->rho_full, H, i, j, k);
                    den[size_t(S.g0) + size_t(i - 1) + size_t(j - 1) * H.nx + size_t(k - 1) * H.nx * H.ny] =
                        r * (v * v * v);
                }
    }
    b->d_bk.upload(bk, stm);
    b->d_rho.upload(rh, stm);
    b->d_den.upload(den, stm);
    for (auto& x : b->f) x.alloc(b->field_elems);
    for (auto& x : b->m) x.alloc(b->field_elems);
    b->d_traces.alloc(b->trace_elems);
    OWL_HIP(hipStreamSynchronize(stm));
}

// Algorithmic bytes of one sweep over every shot: the p-sweep reads vx, vy,
// vz, p, bk and writes p (24 B per point); the v-sweep reads p, vx, vy, vz,
// rho and writes vx, vy, vz (32 B); each damped slab direction adds a memory
// variable read + write (8 B).
void sweep_bytes(const owl_acoustic3d* b, double& bp, double& bv) {
    bp = bv = 0;
    for (auto& S : b->sd) {
        const double lx = S.nx + 2.0 * b->pml, ly = S.ny + 2.0 * b->pml, lz = S.nz + 2.0 * b->pml;
        const double pts = lx * ly * lz;
        const double sp = 2.0 * b->pml * (ly * lz + lx * lz + lx * ly);
        const double sv = (2.0 * b->pml + 1) * (ly * lz + lx * lz + lx * ly);
        bp += 24.0 * pts + 8.0 * sp;
        bv += 32.0 * pts + 8.0 * sv;
    }
}

// One time step of OWL's update_wavefield on field set (in -> out) with
// step dt: p-sweep, free-surface mirror, v-sweep.
template <int R>
void sweeps3(owl_acoustic3d* b, const F3& in_place_or_pair, float dt) {
    hipStream_t st = b->stream;
    const F3& F = in_place_or_pair;
    const dim3 blk(kBX, kBY), grd(unsigned(b->tile_shot.size()));
    const int ntiles = int(b->tile_shot.size());
    const dim3 cblk(64, unsigned(b->cy)), cgrd(unsigned(b->remap ? (ntiles + kXcd - 1) / kXcd * kXcd : ntiles));
    double bp, bv;
    sweep_bytes(b, bp, bv);
    int ev = b->prof.begin(st, 0, bp);
    if (b->variant == 2)
        hipLaunchKernelGGL(k3w_update_p<R>, cgrd, dim3(kWX, kWY), 0, st, b->d_shots.get(), b->d_tile_shot.get(), F,
                           b->d_bk.get(), b->d_prof.get(), b->coef, dt, b->idx, b->idy, b->idz, b->pml, b->kc, ntiles,
                           b->remap, b->fs);
    else if (b->variant == 3)
        hipLaunchKernelGGL(k3q_update_p<R>, cgrd, cblk, 0, st, b->d_shots.get(), b->d_tile_shot.get(), F,
                           b->d_bk.get(), b->d_prof.get(), b->coef, dt, b->idx, b->idy, b->idz, b->pml, b->kc, ntiles,
                           b->remap, b->fs);
    else if (b->variant == 0)
        hipLaunchKernelGGL(k3_update_p<R>, grd, blk, 0, st, b->d_shots.get(), b->d_tile_shot.get(), F, b->d_bk.get(),
                           b->d_prof.get(), b->coef, dt, b->idx, b->idy, b->idz, b->pml, b->kc, b->fs);
    else
        hipLaunchKernelGGL(k3c_update_p<R>, cgrd, cblk, 0, st, b->d_shots.get(), b->d_tile_shot.get(), F,
                           b->d_bk.get(), b->d_prof.get(), b->coef, dt, b->idx, b->idy, b->idz, b->pml, b->kc, ntiles,
                           b->remap, b->fs);
    b->prof.end(st, ev);
    if (b->fs) {
        const unsigned nshot = unsigned(b->sd.size());
        hipLaunchKernelGGL(k3_mirror_p<R>, dim3(256, nshot), dim3(256), 0, st, b->d_shots.get(), F.po, b->pml);
    }
    // the v-sweep reads the new pressure
    F3 V = F;
    V.p = F.po;
    ev = b->prof.begin(st, 1, bv);
    if (b->variant == 2)
        hipLaunchKernelGGL(k3w_update_v<R>, cgrd, dim3(kWX, kWY), 0, st, b->d_shots.get(), b->d_tile_shot.get(), V,
                           b->d_rho.get(), b->d_prof.get(), b->coef, dt, b->idx, b->idy, b->idz, b->pml, b->kc, ntiles,
                           b->remap, b->fs);
    else if (b->variant == 3)
        hipLaunchKernelGGL(k3q_update_v<R>, cgrd, cblk, 0, st, b->d_shots.get(), b->d_tile_shot.get(), V,
                           b->d_rho.get(), b->d_prof.get(), b->coef, dt, b->idx, b->idy, b->idz, b->pml, b->kc, ntiles,
                           b->remap, b->fs);
    else if (b->variant == 0)
        hipLaunchKernelGGL(k3_update_v<R>, grd, blk, 0, st, b->d_shots.get(), b->d_tile_shot.get(), V,
                           b->d_rho.get(), b->d_prof.get(), b->coef, dt, b->idx, b->idy, b->idz, b->pml, b->kc, b->fs);
    else
        hipLaunchKernelGGL(k3c_update_v<R>, cgrd, cblk, 0, st, b->d_shots.get(), b->d_tile_shot.get(), V,
                           b->d_rho.get(), b->d_prof.get(), b->coef, dt, b->idx, b->idy, b->idz, b->pml, b->kc, ntiles,
                           b->remap, b->fs);
    b->prof.end(st, ev);
}

// Algorithmic bytes of one fused step: p, vx, vy, vz read and written and
// bk, rho read (40 B per point), 8 B per slab cell and damped direction for
// each of the two sets of memory variables.
double fused_bytes(const owl_acoustic3d* b) {
    double bp, bv;
    sweep_bytes(b, bp, bv);
    double pts = 0;
    for (auto& S : b->sd)
        pts += (S.nx + 2.0 * b->pml) * (S.ny + 2.0 * b->pml) * (S.nz + 2.0 * b->pml);
    return bp + bv - 16.0 * pts;
}

struct Set3 {
    float* f[4];    // p, vx, vy, vz
    float* pm[3];   // p-sweep memory variables
};

// One time step by the fused kernel from set `in` to set `out`.
template <int R>
void fused_step3(owl_acoustic3d* b, const Set3& in, const Set3& out, float dt) {
    F3F F{};
    F.p = in.f[0];
    F.vx = in.f[1];
    F.vy = in.f[2];
    F.vz = in.f[3];
    F.po = out.f[0];
    F.vxo = out.f[1];
    F.vyo = out.f[2];
    F.vzo = out.f[3];
    for (int k = 0; k < 3; ++k) {
        F.pm[k] = in.pm[k];
        F.pmo[k] = out.pm[k];
        F.vm[k] = b->m[3 + k].get();
    }
    const int ntiles = int(b->ftile_shot.size());
    const unsigned grid = unsigned((ntiles + kXcd - 1) / kXcd * kXcd);
    const int ev = b->prof.begin(b->stream, 6, fused_bytes(b));
    hipLaunchKernelGGL(k3f_step<R>, dim3(grid), dim3(64, kFW), 0, b->stream, b->d_shots.get(), b->d_ftile_shot.get(),
                       F, b->d_bk.get(), b->d_rho.get(), b->d_prof.get(), b->coef, dt, b->idx, b->idy, b->idz,
                       b->pml, b->fkc, ntiles, 1);
    b->prof.end(b->stream, ev);
}

Set3 set_a(owl_acoustic3d* b) {
    return Set3{{b->f[0].get(), b->f[1].get(), b->f[2].get(), b->f[3].get()},
                {b->m[0].get(), b->m[1].get(), b->m[2].get()}};
}
Set3 set_b(owl_acoustic3d* b) {
    for (auto& x : b->fB) x.alloc(b->field_elems);
    for (auto& x : b->mB) x.alloc(b->field_elems);
    return Set3{{b->fB[0].get(), b->fB[1].get(), b->fB[2].get(), b->fB[3].get()},
                {b->mB[0].get(), b->mB[1].get(), b->mB[2].get()}};
}

void maybe_collect(owl_acoustic3d* b) {
    if (b->prof.pending() > 20000) {
        OWL_HIP(hipStreamSynchronize(b->stream));
        b->prof.collect();
    }
}

void strips_launch(owl_acoustic3d* b, int R, float* vx, float* vy, float* vz, int t, int save);

// Every launch parameter and device pointer the forward loop bakes into its
// launches: a graph captured under one signature is replayed only under the
// same one.
std::vector<uintptr_t> forward_signature(owl_acoustic3d* b) {
    std::vector<uintptr_t> v{uintptr_t(b->nt), uintptr_t(b->R), uintptr_t(b->fused), uintptr_t(b->fkc),
                             uintptr_t(b->kc), uintptr_t(b->variant), uintptr_t(b->cy), uintptr_t(b->remap),
                             uintptr_t(b->tile_shot.size()), uintptr_t(b->ftile_shot.size()),
                             uintptr_t(b->sd.size())};
    auto add = [&](const void* p) { v.push_back(reinterpret_cast<uintptr_t>(p)); };
    for (auto& x : b->f) add(x.get());
    for (auto& x : b->m) add(x.get());
    for (auto& x : b->fB) add(x.get());
    for (auto& x : b->mB) add(x.get());
    for (auto* x : {&b->d_traces, &b->d_bk, &b->d_rho, &b->d_prof, &b->d_amp, &b->d_rec_w}) add(x->get());
    add(b->d_shots.get()); add(b->d_tile_shot.get()); add(b->d_ftile_shot.get());
    add(b->d_inj_beg.get()); add(b->d_inj.get()); add(b->d_rec_beg.get()); add(b->d_rec.get());
    return v;
}

template <int R>
void run_forward3(owl_acoustic3d* b, bool save_boundary) {
    hipStream_t st = b->stream;
    for (auto& x : b->f) x.zero(st);
    for (auto& x : b->m) x.zero(st);
    b->d_traces.zero(st);
    if (save_boundary) {
        size_t freeb = 0, total = 0;
        OWL_HIP(hipMemGetInfo(&freeb, &total));
        const size_t need = b->strip_elems * sizeof(float);
        if (b->d_strips.size() < b->strip_elems && need + (size_t(1) << 30) > freeb)
            fail(OWL_ERR_NOMEM, "boundary records need " + std::to_string(need >> 20) + " MiB; " +
                                    std::to_string(freeb >> 20) + " MiB free");
        b->d_strips.alloc(std::max<size_t>(b->strip_elems, 1));
        b->d_edge.alloc(std::max<size_t>(b->edge_elems, 1));
    }
    float* mm[6];
    for (int k = 0; k < 6; ++k) mm[k] = b->m[k].get();
    const F3 F = f3_in_place(b->f[0].get(), b->f[1].get(), b->f[2].get(), b->f[3].get(), mm);
    const bool fused = b->fused && !b->fs;
    Set3 cur = set_a(b), nxt{};
    if (fused) {
        nxt = set_b(b);
        for (auto& x : b->fB) x.zero(st);
        for (auto& x : b->mB) x.zero(st);
    }
    // step order (compute_forward): save strips(t) -> record -> update -> source;
    // the record of step 1 is the zero initial state.  The state at the start
    // of step nt is the final state the reverse pass starts from.
    auto loop = [&] {
        for (int t = 1; t <= b->nt - 1; ++t) {
            if (save_boundary) strips_launch(b, R, cur.f[1], cur.f[2], cur.f[3], t, 1);
            if (fused) {
                fused_step3<R>(b, cur, nxt, b->dt);
                std::swap(cur, nxt);
            } else {
                sweeps3<R>(b, F, b->dt);
            }
            const int ev = b->prof.begin(st, 2, 0);
            hipLaunchKernelGGL(k3_post_step, dim3(unsigned(b->sd.size())), dim3(1024), 0, st, b->d_shots.get(),
                               cur.f[0], cur.f[1], cur.f[2], cur.f[3], b->d_inj_beg.get(),
                               b->d_inj.get(), b->d_amp.get(), b->nt, t, b->d_rec_beg.get(), b->d_rec.get(),
                               b->d_rec_w.get(), b->d_traces.get());
            b->prof.end(st, ev);
            if (!b->graph.capturing) maybe_collect(b);
        }
    };
    const bool graph_ok = b->graph.enabled && !b->prof.on() && !save_boundary;
    OWL_HIP(hipEventRecord(b->ev0, st));
    if (graph_ok)
        b->graph.run(st, forward_signature(b), loop);
    else
        loop();
    if (save_boundary) strips_launch(b, R, cur.f[1], cur.f[2], cur.f[3], b->nt, 1);
    // the final state lives in b->f for the reverse pass (modelling alone
    // keeps the buffers where they are, so a captured loop stays valid)
    if (save_boundary && fused && cur.f[0] != b->f[0].get()) {
        for (int k = 0; k < 4; ++k) std::swap(b->f[k], b->fB[k]);
        for (int k = 0; k < 3; ++k) std::swap(b->m[k], b->mB[k]);
    }
    OWL_HIP(hipGetLastError());
    OWL_HIP(hipEventRecord(b->ev1, st));
    OWL_HIP(hipStreamSynchronize(st));
    b->prof.collect();
    float ms = 0.f;
    OWL_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_ms = ms;
    b->last_updates = double(b->nt - 1) * b->pts_full;
    b->have_boundary = save_boundary;
    b->fwd_state_valid = save_boundary;
}

void strips_launch(owl_acoustic3d* b, int R, float* vx, float* vy, float* vz, int t, int save) {
    const unsigned nshot = unsigned(b->sd.size());
    const int ev = b->prof.begin(b->stream, 4, 0);
    switch (R) {
#define OWL_STRIPS_CASE(RR)                                                                                         \
    case RR:                                                                                                        \
        hipLaunchKernelGGL(k3_strips<RR>, dim3(128, nshot), dim3(256), 0, b->stream, b->d_shots.get(), vx, vy, vz,   \
                           b->d_strips.get(), b->d_edge.get(), t, b->pml, save);                                    \
        break;
        OWL_STRIPS_CASE(2)
        OWL_STRIPS_CASE(3)
        OWL_STRIPS_CASE(4)
        OWL_STRIPS_CASE(5)
        OWL_STRIPS_CASE(6)
        OWL_STRIPS_CASE(7)
        OWL_STRIPS_CASE(8)
#undef OWL_STRIPS_CASE
        default: fail(OWL_ERR_ARG, "fdhalf must be 2..8");
    }
    b->prof.end(b->stream, ev);
}

// OWL's reverse pass (compute_adjoint, acoustic_iso_3d/submodule_fwi_gradient.f90:
// 33-274): t = nt .. hnt; the forward field runs backwards out of place
// (the previous state stays in the other buffer for the imaging), the
// adjoint field in place.
template <int R>
void run_reverse3(owl_acoustic3d* b) {
    hipStream_t st = b->stream;
    if (!b->have_boundary || !b->fwd_state_valid)
        fail(OWL_ERR_STATE, "the adjoint needs a preceding forward run with boundary saving");
    if (!b->d_adj.get() || b->d_adj.size() < b->trace_elems) fail(OWL_ERR_STATE, "no adjoint source");
    // OWL's loop ends at the first source's hnt of the shot
    // (submodule_fwi_gradient.f90:81); the batch runs down to the smallest one
    // and images each shot only while t >= its own
    for (int k = 0; k < 4; ++k) {
        b->fB[k].alloc(b->field_elems);
        b->fB[k].zero(st);
        b->aF[k].alloc(b->field_elems);
        b->aF[k].zero(st);
    }
    for (int k = 0; k < 6; ++k) {
        b->m[k].zero(st);   // the reverse run starts its memory variables from zero
        b->aM[k].alloc(b->field_elems);
        b->aM[k].zero(st);
    }
    b->d_gv.alloc(std::max<size_t>(b->grad_elems, 1));
    b->d_gr.alloc(std::max<size_t>(b->grad_elems, 1));
    b->d_gv.zero(st);
    b->d_gr.zero(st);
    const unsigned nshot = unsigned(b->sd.size());
    int hnt = b->nt;
    for (auto& H : b->hs) hnt = std::min(hnt, std::max(H.hnt, 1));
    const int cc = b->cfg.cc_step_interval > 0 ? b->cfg.cc_step_interval : 1;
    DevBuf<float>* A = b->f;
    DevBuf<float>* Bf = b->fB;
    float* fm[6];
    float* am[6];
    for (int k = 0; k < 6; ++k) {
        fm[k] = b->m[k].get();
        am[k] = b->aM[k].get();
    }
    const F3 adj = f3_in_place(b->aF[0].get(), b->aF[1].get(), b->aF[2].get(), b->aF[3].get(), am);
    // imaging: p, prev p, pr, rho vp^3 and nine velocity fields (52 B per
    // point) and two FP64 accumulators read and written (32 B)
    const double image_bytes = 84.0 * double(b->grad_elems);
    const unsigned ig = unsigned(std::max<size_t>(1, std::min<size_t>(2048, (b->grad_elems / nshot + 255) / 256)));
    OWL_HIP(hipEventRecord(b->ev0, st));
    for (int t = b->nt; t >= hnt; --t) {
        strips_launch(b, R, A[1].get(), A[2].get(), A[3].get(), t, 0);
        F3 F = f3_in_place(A[0].get(), A[1].get(), A[2].get(), A[3].get(), fm);
        F.po = Bf[0].get();
        F.vxo = Bf[1].get();
        F.vyo = Bf[2].get();
        F.vzo = Bf[3].get();
        sweeps3<R>(b, F, -b->dt);
        sweeps3<R>(b, adj, -b->dt);
        int ev = b->prof.begin(st, 5, 0);
        if (b->max_nadj > 0)
            hipLaunchKernelGGL(k3_inject_adjoint, dim3(unsigned((b->max_nadj + 255) / 256), nshot), dim3(256), 0, st,
                               b->d_shots.get(), b->aF[0].get(), b->d_adj_beg.get(), b->d_adj_taps.get(),
                               b->d_rec_w.get(), b->d_adj.get(), t);
        b->prof.end(st, ev);
        if (t % cc == 0) {
            const F3 cur = f3_in_place(Bf[0].get(), Bf[1].get(), Bf[2].get(), Bf[3].get(), fm);
            const F3 prv = f3_in_place(A[0].get(), A[1].get(), A[2].get(), A[3].get(), fm);
            ev = b->prof.begin(st, 3, image_bytes);
            hipLaunchKernelGGL(k3_image_owl<R>, dim3(ig, nshot), dim3(256), 0, st, b->d_shots.get(), cur, prv,
                               b->d_edge.get(), adj, b->d_den.get(), b->d_gv.get(), b->d_gr.get(),
                               t == b->nt ? 1 : 0, b->pml, t);
            b->prof.end(st, ev);
        }
        std::swap(A, Bf);
        maybe_collect(b);
    }
    OWL_HIP(hipGetLastError());
    OWL_HIP(hipEventRecord(b->ev1, st));
    OWL_HIP(hipStreamSynchronize(st));
    b->prof.collect();
    if (A != b->f)
        for (int k = 0; k < 4; ++k) std::swap(b->f[k], b->fB[k]);
    float ms = 0.f;
    OWL_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_ms = ms;
    b->last_updates = 2.0 * double(b->nt - hnt + 1) * b->pts_full;
    b->fwd_state_valid = false;
}

template <class F>
void dispatch_R3(int R, F&& f) {
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

// Device traces [nt][nr] per shot -> host [nr][nt] per shot.
void traces_to_host(owl_acoustic3d* b, const float* dev, float* out) {
    std::vector<float> h(b->trace_elems);
    OWL_HIP(hipMemcpy(h.data(), dev, h.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (auto& S : b->sd)
        for (int r = 0; r < S.nr; ++r)
            for (int t = 0; t < b->nt; ++t)
                out[size_t(S.trace0) + size_t(r) * b->nt + t] = h[size_t(S.trace0) + size_t(t) * S.nr + r];
}

void set_shape(owl_acoustic3d* b, int planes, int variant);

// Launch-shape autotuner: times `steps` forward steps (p-sweep, v-sweep) of
// the batch under each candidate shape and keeps the fastest.  The state is
// zeroed afterwards (every forward starts from zero anyway).  Stencil timing
// does not depend on the field values.
template <int R>
void autotune3(owl_acoustic3d* b, int steps) {
    static const int kVariants[] = {1, 2, 3, 5, 6, 8, 24, 9, 10, 11, 13, 0};
    static const int kPlanes[] = {32, 64, 128};
    static const int kFusedPlanes[] = {32, 64, 96, 128, 192};
    hipStream_t st = b->stream;
    for (auto& x : b->f) x.zero(st);
    for (auto& x : b->m) x.zero(st);
    float* mm[6];
    for (int k = 0; k < 6; ++k) mm[k] = b->m[k].get();
    const F3 F = f3_in_place(b->f[0].get(), b->f[1].get(), b->f[2].get(), b->f[3].get(), mm);
    const bool prof_on = b->prof.on();
    b->prof.enable(false);
    b->fused = false;
    auto timed = [&](auto&& step) {
        step();  // warm-up
        OWL_HIP(hipEventRecord(b->ev0, st));
        for (int t = 0; t < steps; ++t) step();
        OWL_HIP(hipEventRecord(b->ev1, st));
        OWL_HIP(hipStreamSynchronize(st));
        float ms = 0.f;
        OWL_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
        return ms;
    };
    float best = -1.f;
    int best_v = b->launch_code, best_k = b->kc;
    for (int v : kVariants)
        for (int kc : kPlanes) {
            set_shape(b, kc, v);
            const float ms = timed([&] { sweeps3<R>(b, F, b->dt); });
            if (best < 0 || ms < best) {
                best = ms;
                best_v = v;
                best_k = kc;
            }
        }
    set_shape(b, best_k, best_v);
    b->launch_code = best_v;
    if (!b->fs && b->field_elems < (size_t(1) << 32)) {
        // the fused step against the best pair of sweeps
        const Set3 A = set_a(b), B = set_b(b);
        for (auto& x : b->fB) x.zero(st);
        for (auto& x : b->mB) x.zero(st);
        float fbest = -1.f;
        int fk = b->fkc;
        for (int kc : kFusedPlanes) {
            b->fkc = kc;
            retile(b);
            const float ms = timed([&] { fused_step3<R>(b, A, B, b->dt); });
            if (fbest < 0 || ms < fbest) {
                fbest = ms;
                fk = kc;
            }
        }
        b->fkc = fk;
        retile(b);
        if (fbest < best) {
            b->fused = true;
            b->launch_code = 32;
        }
        for (auto& x : b->fB) x.zero(st);
        for (auto& x : b->mB) x.zero(st);
    }
    for (auto& x : b->f) x.zero(st);
    for (auto& x : b->m) x.zero(st);
    OWL_HIP(hipStreamSynchronize(st));
    b->prof.enable(prof_on);
}

void traces_to_device(owl_acoustic3d* b, const float* in, DevBuf<float>& dev) {
    std::vector<float> h(b->trace_elems);
    for (auto& S : b->sd)
        for (int r = 0; r < S.nr; ++r)
            for (int t = 0; t < b->nt; ++t)
                h[size_t(S.trace0) + size_t(t) * S.nr + r] = in[size_t(S.trace0) + size_t(r) * b->nt + t];
    dev.alloc(std::max<size_t>(h.size(), 1));
    dev.upload(h.data(), h.size(), b->stream);
    OWL_HIP(hipStreamSynchronize(b->stream));
}

}  // namespace

extern "C" {

owl_status owl_acoustic3d_create(const owl_acoustic_config* cfg, const float* vp, const float* rho, int nshot,
                                 const owl_shot* shots, owl_acoustic3d** out) {
    return guarded([&] {
        if (!out) fail(OWL_ERR_ARG, "out is null");
        *out = nullptr;
        auto* b = new owl_acoustic3d();
        try {
            plan3(b, cfg, vp, rho, nshot, shots);
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

void owl_acoustic3d_destroy(owl_acoustic3d* b) { delete b; }

int owl_acoustic3d_nt(const owl_acoustic3d* b) { return b ? b->nt : 0; }

}  // extern "C"

namespace {
// Launch shape from the public variant code: 0 LDS planes; 1, 2, 3 columns
// with 4, 8, 16 rows per block; 8 four-wide columns; 9, 10, 11 pipelined
// columns with 4, 8, 16 rows; +4 (or +16 for 8): XCD-contiguous tile order.
void set_shape(owl_acoustic3d* b, int planes, int variant) {
    if (planes < 1 || variant < 0) fail(OWL_ERR_ARG, "bad launch shape");
    if (variant == 32) {
        // fused forward step; the reverse pass keeps the sweeps' shape
        if (b->fs) fail(OWL_ERR_ARG, "the fused step does not cover the free surface");
        if (b->field_elems >= (size_t(1) << 32))
            fail(OWL_ERR_ARG, "the fused step addresses fields of fewer than 2^32 elements");
        b->fused = true;
        if (planes != b->fkc) {
            b->fkc = planes;
            retile(b);
        }
        return;
    }
    b->fused = false;
    int kind, cy = b->cy, rm;
    if (variant == 8 || variant == 24) {
        kind = 2;
        rm = variant == 24;
    } else if (variant >= 9 && variant <= 15 && variant != 12) {
        const int v = variant & 3;
        kind = 3;
        cy = v == 1 ? 4 : (v == 2 ? 8 : 16);
        rm = (variant >> 2) & 1;
    } else if (variant <= 7) {
        const int v = variant & 3;
        kind = v == 0 ? 0 : 1;
        cy = v == 1 ? 4 : (v == 2 ? 8 : 16);
        rm = variant >> 2;
    } else {
        fail(OWL_ERR_ARG, "bad launch variant");
    }
    if (planes == b->kc && kind == b->variant && cy == b->cy && rm == b->remap) return;
    b->variant = kind;
    b->cy = cy;
    b->remap = rm;
    b->kc = planes;
    retile(b);
}
}  // namespace

extern "C" {

owl_status owl_acoustic3d_set_graphs(owl_acoustic3d* b, int on) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        b->graph.enabled = on != 0;
        b->graph.reset();
    });
}

owl_status owl_acoustic3d_set_launch(owl_acoustic3d* b, int planes, int variant) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        set_shape(b, planes, variant);
        b->launch_code = variant;
    });
}

owl_status owl_acoustic3d_autotune(owl_acoustic3d* b, int steps, int* planes, int* variant) {
    return guarded([&] {
        if (!b || steps < 1) fail(OWL_ERR_ARG, "bad argument");
        dispatch_R3(b->R, [&](auto r) { autotune3<decltype(r)::value>(b, steps); });
        if (planes) *planes = b->fused ? b->fkc : b->kc;
        if (variant) *variant = b->launch_code;
    });
}

owl_status owl_acoustic3d_forward(owl_acoustic3d* b, float* traces, int save_boundary) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        dispatch_R3(b->R, [&](auto r) { run_forward3<decltype(r)::value>(b, save_boundary != 0); });
        if (traces) traces_to_host(b, b->d_traces.get(), traces);
    });
}

owl_status owl_acoustic3d_adjoint(owl_acoustic3d* b, const float* adj, float* grad_vp, float* grad_rho) {
    return guarded([&] {
        if (!b || !adj) fail(OWL_ERR_ARG, "null argument");
        traces_to_device(b, adj, b->d_adj);
        dispatch_R3(b->R, [&](auto r) { run_reverse3<decltype(r)::value>(b); });
        // per-shot window gradients (nz_s, ny_s, nx_s), z fastest: the .grd
        // arrays (permute 321 of the x-fast accumulators), NaN/Inf -> 0
        std::vector<double> gv(b->grad_elems), gr(b->grad_elems);
        OWL_HIP(hipMemcpy(gv.data(), b->d_gv.get(), gv.size() * sizeof(double), hipMemcpyDeviceToHost));
        OWL_HIP(hipMemcpy(gr.data(), b->d_gr.get(), gr.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (auto& S : b->sd)
            for (int k = 0; k < S.nz; ++k)
                for (int j = 0; j < S.ny; ++j)
                    for (int i = 0; i < S.nx; ++i) {
                        const size_t src = size_t(S.g0) + i + size_t(j) * S.nx + size_t(k) * S.nx * S.ny;
                        const size_t dst = size_t(S.g0) + k + size_t(j) * S.nz + size_t(i) * S.nz * S.ny;
                        const float a = float(gv[src]), c = float(gr[src]);
                        if (grad_vp) grad_vp[dst] = std::isfinite(a) ? a : 0.f;
                        if (grad_rho) grad_rho[dst] = std::isfinite(c) ? c : 0.f;
                    }
    });
}

owl_status owl_acoustic3d_gradient(owl_acoustic3d* b, const float* obs, double* misfit, double* grad_vp,
                                   double* grad_rho, float* traces_out) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        if (obs) {
            traces_to_device(b, obs, b->d_obs);
            b->have_obs = true;
        }
        if (!b->have_obs) fail(OWL_ERR_STATE, "no observed data");
        dispatch_R3(b->R, [&](auto r) { run_forward3<decltype(r)::value>(b, true); });
        const double fwd_ms = b->last_ms, fwd_up = b->last_updates;
        hipStream_t st = b->stream;
        const unsigned nshot = unsigned(b->sd.size());
        b->d_adj.alloc(std::max<size_t>(b->trace_elems, 1));
        b->d_misfit.alloc(nshot);
        b->d_scratch.alloc(std::max<size_t>(2 * size_t(b->nrec_total), 1));
        hipLaunchKernelGGL(k3_adjoint_source, dim3(nshot), dim3(256), 0, st, b->d_shots.get(), b->d_traces.get(),
                           b->d_obs.get(), b->d_rec_w.get(), b->d_adj.get(), b->d_misfit.get(), b->nt, b->cfg.trace_discard_threshold,
                           b->d_scratch.get());
        dispatch_R3(b->R, [&](auto r) { run_reverse3<decltype(r)::value>(b); });
        const size_t nfull = size_t(b->cfg.nx) * b->cfg.ny * b->cfg.nz;
        b->d_gfull_v.alloc(nfull);
        b->d_gfull_r.alloc(nfull);
        const unsigned mg = unsigned(std::min<size_t>(8192, (nfull + 255) / 256));
        hipLaunchKernelGGL(k3_merge, dim3(mg), dim3(256), 0, st, b->d_shots.get(), int(nshot), b->d_gv.get(),
                           b->d_gfull_v.get(), b->cfg.nx, b->cfg.ny, b->cfg.nz);
        hipLaunchKernelGGL(k3_merge, dim3(mg), dim3(256), 0, st, b->d_shots.get(), int(nshot), b->d_gr.get(),
                           b->d_gfull_r.get(), b->cfg.nx, b->cfg.ny, b->cfg.nz);
        OWL_HIP(hipGetLastError());
        OWL_HIP(hipStreamSynchronize(st));
        if (misfit) OWL_HIP(hipMemcpy(misfit, b->d_misfit.get(), nshot * sizeof(double), hipMemcpyDeviceToHost));
        if (grad_vp) OWL_HIP(hipMemcpy(grad_vp, b->d_gfull_v.get(), nfull * sizeof(double), hipMemcpyDeviceToHost));
        if (grad_rho) OWL_HIP(hipMemcpy(grad_rho, b->d_gfull_r.get(), nfull * sizeof(double), hipMemcpyDeviceToHost));
        if (traces_out) traces_to_host(b, b->d_traces.get(), traces_out);
        b->last_ms += fwd_ms;
        b->last_updates += fwd_up;
    });
}

owl_status owl_acoustic3d_set_profiling(owl_acoustic3d* b, int on) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        b->prof.enable(on != 0);
    });
}

owl_status owl_acoustic3d_kernel_stats(const owl_acoustic3d* b, int kind, double* total_ms, long long* launches,
                                       double* bytes_total) {
    return guarded([&] {
        if (!b || kind < 0 || kind >= Profiler::kKinds) fail(OWL_ERR_ARG, "bad kernel kind");
        if (total_ms) *total_ms = b->prof.ms(kind);
        if (launches) *launches = b->prof.launches(kind);
        if (bytes_total) *bytes_total = b->prof.bytes(kind);
    });
}

owl_status owl_acoustic3d_last_stats(const owl_acoustic3d* b, double* ms, double* updates) {
    return guarded([&] {
        if (!b) fail(OWL_ERR_ARG, "null batch");
        if (ms) *ms = b->last_ms;
        if (updates) *updates = b->last_updates;
    });
}

}  // extern "C"
