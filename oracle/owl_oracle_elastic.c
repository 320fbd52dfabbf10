/*
 * owl_oracle_elastic.c -- CPU restatement of OWL's isotropic elastic
 * staggered-grid velocity-stress scheme in 3D (elastic_vhtiort_3d with
 * anisotropy_type = 'iso').
 *
 * TEST INFRASTRUCTURE ONLY (see owl_oracle.c): only tests/, smoke() and the
 * CPU baseline may load it, as the checker.
 *
 * Restated from (paths relative to the OWL source tree):
 *   src/solver/elastic_vhtiort_3d/module_vars.f90:225-300  iso c_ij, pmlvp, padding
 *   src/solver/elastic_vhtiort_3d/module_cfspml.f90:428-625 CFS-MPML profiles
 *   src/solver/elastic_vhtiort_3d/module_wavefield.f90:52-545 update_wavefield
 *   src/solver/elastic_vhtiort_3d/module_wavefield.f90:549-1090 free surface
 *   src/solver/elastic_vhtiort_3d/module_wavefield.f90:1095-1200 add_source (force)
 *   src/solver/elastic_vhtiort_3d/submodule_forward_modeling.f90:31-140
 *   src/solver/elastic_vhtiort_3d/macro_fd_stencil.f90 (stencil offsets)
 *
 * Scope and stated assumptions:
 *   - isotropic media: the M-PML damping ratios are zero, kappa = 1;
 *   - uniform mesh (free_surface_dz_refine = 1, dz_max = dz): the refined
 *     mesh's dz_scaling factors are 1 and the model is used as given
 *     (FLIT map_regular_to_irregular taken as the identity there);
 *   - force sources (explosion/moment-tensor sources need FLIT's deriv of
 *     the STF, whose end treatment is unpinned);
 *   - the 4-term sums of the harmonic shear moduli follow Fortran array
 *     element order, left to right;
 *   - the free-surface mirrors (v above the surface zeroed before the stress
 *     update, stresses mirrored after it, v mirrored before recording) are
 *     applied.  In the 3D Fortran they sit behind tests against the block's
 *     z range from FLIT's domain decomposition (module_wavefield.f90:580-590,
 *     823-845; submodule_forward_modeling.f90:64-73), which is absent; the
 *     2D solver applies the same mirrors unconditionally
 *     (elastic_vhtiort_2d/module_wavefield.f90:159-163, 225-244).
 * Compiled in float and double with owl_oracle.c's conventions.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef ORA_REAL
#define ORA_REAL float
#endif
typedef ORA_REAL real;
#if defined(ORA_F64)
#define ORA_SYM(name) ora64_##name
#else
#define ORA_SYM(name) ora32_##name
#endif
#define NKW 4

/* shared with owl_oracle.c */
typedef struct {
    float x, y, z;
    int gx, gy, gz, hx, hy, hz;
    float ix[9], iy[9], iz[9], hxw[9], hyw[9], hzw[9];
} ora_point;
int ORA_SYM(fdcoefs)(int R, float *out);
float ORA_SYM(alphamax)(float f0max);
void ORA_SYM(locate_point)(int dims, float x, float y, float z, float ox, float oy, float oz, float dx, float dy,
                           float dz, ora_point *pt);

typedef struct {
    int nx, ny, nz;
    float dx, dy, dz, ox, oy, oz;
    float dt;
    int nt, pml, R;
    int free_surface;
    int cc_step_interval;
} ora_egrid3d;

typedef struct {
    float x, y, z, amp, t0, polar, azimuth; /* radians */
    int mechanism;                          /* 0 explosion, 1 force */
    int nt_stf;
    const real *stf;
} ora_esource3d;

typedef struct {
    float x, y, z, weight;
} ora_ereceiver3d;

typedef struct {
    int lo, NX, NY, NZ, nx, ny, nz, pml, R, fs;
    size_t sy, sz;
} elay;

#define IE(L, i, j, k) \
    ((size_t)((i) - (L).lo) + (size_t)((j) - (L).lo) * (L).sy + (size_t)((k) - (L).lo) * (L).sz)

/* 1D M-PML profile of one axis (module_cfspml.f90:498-620): distance to
 * the first/last interior node in units of pml cells, d0 = ln(1/R0) (n+1)
 * pmlvp / (2 pml d) nd^2, alpha = max(alphamax (1 - nd), alphamax/4); set
 * for i <= 1 and i >= n (only i >= n when `top_off`).  Index i + pml - 1. */
typedef struct {
    float *di, *dh, *ai, *ah;
    int n, pml;
} eprof;

static void eprof_init(eprof *P, int n, int pml, float d, float pmlvp, float alphamax, int top_off) {
    const int len = n + 2 * pml;
    P->n = n;
    P->pml = pml;
    P->di = (float *)calloc(4 * (size_t)len, sizeof(float));
    P->dh = P->di + len;
    P->ai = P->dh + len;
    P->ah = P->ai + len;
    const float R0 = 1.0e-5f, extra = 0.25f * alphamax;
    const float lg = logf(1.0f / R0);
    for (int i = -pml + 1; i <= n + pml; ++i) {
        const int q = i + pml - 1;
        int side = 0;
        if (!top_off && i <= 1) side = -1;
        else if (i >= n) side = 1;
        if (!side) continue;
        const int ref = side < 0 ? 1 : n;
        const float nd = (float)abs(i - ref) * 1.0f / (float)pml;
        const float ndh = fabsf((float)(i - ref) - 0.5f) * 1.0f / (float)pml;
        P->di[q] = lg * (2 + 1.0f) * pmlvp / (2.0f * (float)pml * d) * (nd * nd);
        P->dh[q] = lg * (2 + 1.0f) * pmlvp / (2.0f * (float)pml * d) * (ndh * ndh);
        P->ai[q] = fmaxf(alphamax * (1.0f - nd), extra);
        P->ah[q] = fmaxf(alphamax * (1.0f - ndh), extra);
    }
}

/* alpha at a node whose own axis index is interior (1 < i < n): alphamax
 * outside the interior box, 0 inside (module_cfspml.f90:488-512).  Its
 * damping profile there is 0, which makes the memory update a no-op, but
 * it is restated as written. */
static inline float ealpha(const eprof *P, int i, int outside_box, float alphamax, int half) {
    const int q = i + P->pml - 1;
    if (i <= 1 || i >= P->n) {
        if (half ? (P->dh[q] != 0.0f || P->ah[q] != 0.0f) : (P->di[q] != 0.0f || P->ai[q] != 0.0f))
            return half ? P->ah[q] : P->ai[q];
    }
    return outside_box ? alphamax : 0.0f;
}

/* damp_coef + damp_wavefield (module_wavefield.f90:40-47), kappa = 1 */
static inline real damp(real *w, real v, float d, float alpha, real dt) {
    const real adt = (real)fabs((double)dt);
    const real a_ = (real)d / (real)1.0f;
    const real b_ = (real)d / (real)1.0f + (real)alpha;
    const real ca = -((real)2.0f * adt * a_ / ((real)2.0f + adt * b_));
    const real cb = ((real)2.0f - adt * b_) / ((real)2.0f + adt * b_);
    *w = ca * v + cb * (*w);
    return (v + *w) / (real)1.0f;
}

typedef struct {
    elay L;
    real *c11, *c12, *c44, *rho;   /* padded iso moduli (c22 = c33 = c11, c13 = c23 = c12, c55 = c66 = c44) */
    eprof px, py, pz;
    float alphamax;
    real idx, idy, idz;
    float dx, dy, dz;
    float cf[8];
} emed;

typedef struct {
    real *sxx, *syy, *szz, *syz, *sxz, *sxy, *vx, *vy, *vz;
    /* memory variables, named by derivative and field as in OWL */
    real *m_dxvx, *m_dyvy, *m_dzvz, *m_dyvz, *m_dzvy, *m_dxvz, *m_dzvx, *m_dxvy, *m_dyvx;
    real *m_dxxx, *m_dyxy, *m_dzxz, *m_dxxy, *m_dyyy, *m_dzyz, *m_dxxz, *m_dyyz, *m_dzzz;
} efld;

static real *ezalloc(size_t n) { return (real *)calloc(n, sizeof(real)); }

static void efld_alloc(efld *F, const elay *L) {
    const size_t n = (size_t)L->NX * L->NY * L->NZ;
    real **p = &F->sxx;
    for (int k = 0; k < 27; ++k) p[k] = ezalloc(n);
}
static void efld_free(efld *F) {
    real **p = &F->sxx;
    for (int k = 0; k < 27; ++k) free(p[k]);
}

static void emed_init(emed *M, const ora_egrid3d *g, const float *vp, const float *vs, const float *rho, float f0max) {
    elay L;
    L.pml = g->pml; L.R = g->R; L.nx = g->nx; L.ny = g->ny; L.nz = g->nz; L.fs = g->free_surface;
    L.lo = 1 - g->pml - g->R;
    L.NX = g->nx + 2 * (g->pml + g->R);
    L.NY = g->ny + 2 * (g->pml + g->R);
    L.NZ = g->nz + 2 * (g->pml + g->R);
    L.sy = (size_t)L.NX;
    L.sz = (size_t)L.NX * L.NY;
    M->L = L;
    M->dx = g->dx; M->dy = g->dy; M->dz = g->dz;
    const size_t n = (size_t)L.NX * L.NY * L.NZ;
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    M->c11 = ezalloc(n); M->c12 = ezalloc(n); M->c44 = ezalloc(n); M->rho = ezalloc(n);
    /* c11 = rho vp^2, c44 = rho vs^2, c12 = c11 - 2 c55 (module_vars.f90:284-293) */
    float pm = 0.0f;
    for (int k = L.lo; k < L.lo + L.NZ; ++k)
        for (int j = L.lo; j < L.lo + L.NY; ++j)
            for (int i = L.lo; i < L.lo + L.NX; ++i) {
                const int ic = i < 1 ? 1 : (i > nx ? nx : i), jc = j < 1 ? 1 : (j > ny ? ny : j),
                          kc = k < 1 ? 1 : (k > nz ? nz : k);
                const size_t s = (size_t)(kc - 1) + (size_t)(jc - 1) * nz + (size_t)(ic - 1) * nz * ny;
                const float r = rho[s], p = vp[s], sv = vs[s];
                const float c11 = r * (p * p), c44 = r * (sv * sv);
                const float c12 = c11 - 2.0f * c44;
                const size_t q = IE(L, i, j, k);
                M->c11[q] = c11; M->c12[q] = c12; M->c44[q] = c44; M->rho[q] = r;
                if ((i == 1 || i == nx || j == 1 || j == ny || k == 1 || k == nz) && i >= 1 && i <= nx && j >= 1 &&
                    j <= ny && k >= 1 && k <= nz) {
                    const float v = c11 / r;
                    if (v > pm) pm = v;
                }
            }
    const float pmlvp = sqrtf(pm);
    M->alphamax = ORA_SYM(alphamax)(f0max);
    eprof_init(&M->px, nx, g->pml, g->dx, pmlvp, M->alphamax, 0);
    eprof_init(&M->py, ny, g->pml, g->dy, pmlvp, M->alphamax, 0);
    eprof_init(&M->pz, nz, g->pml, g->dz, pmlvp, M->alphamax, g->free_surface);
#if defined(ORA_F64)
    M->idx = 1.0 / (double)g->dx; M->idy = 1.0 / (double)g->dy; M->idz = 1.0 / (double)g->dz;
#else
    M->idx = (float)(1.0 / (double)g->dx); M->idy = (float)(1.0 / (double)g->dy);
    M->idz = (float)(1.0 / (double)g->dz);
#endif
    ORA_SYM(fdcoefs)(g->R, M->cf);
}

static void emed_free(emed *M) {
    free(M->c11); free(M->c12); free(M->c44); free(M->rho);
    free(M->px.di); free(M->py.di); free(M->pz.di);
}

/* forward difference at f(base): sum_m c_m (f(base + m s) - f(base - (m-1) s)) */
static inline real dplus(const real *f, size_t b, size_t s, const float *c, int R) {
    real a = (real)c[0] * (f[b + s] - f[b]);
    for (int m = 2; m <= R; ++m) a = a + (real)c[m - 1] * (f[b + m * s] - f[b - (m - 1) * s]);
    return a;
}
/* backward difference at node b: sum_m c_m (f(b + (m-1) s) - f(b - m s)) */
static inline real dminus(const real *f, size_t b, size_t s, const float *c, int R) {
    real a = (real)c[0] * (f[b] - f[b - s]);
    for (int m = 2; m <= R; ++m) a = a + (real)c[m - 1] * (f[b + (m - 1) * s] - f[b - m * s]);
    return a;
}

static inline int inbox(const elay *L, int i, int j, int k) {
    return i >= 1 && i <= L->nx && j >= 1 && j <= L->ny && k >= 1 && k <= L->nz;
}

/* x/y/z memory update with this axis's profile at its own index */
#define DAMP_X(w, v, q, ii, half, pt) \
    v = damp(&(w)[q], v, (half) ? M->px.dh[(ii) + pml - 1] : M->px.di[(ii) + pml - 1], ealpha(&M->px, ii, !(pt), M->alphamax, half), dt)
#define DAMP_Y(w, v, q, jj, half, pt) \
    v = damp(&(w)[q], v, (half) ? M->py.dh[(jj) + pml - 1] : M->py.di[(jj) + pml - 1], ealpha(&M->py, jj, !(pt), M->alphamax, half), dt)
#define DAMP_Z(w, v, q, kk, half, pt) \
    v = damp(&(w)[q], v, (half) ? M->pz.dh[(kk) + pml - 1] : M->pz.di[(kk) + pml - 1], ealpha(&M->pz, kk, !(pt), M->alphamax, half), dt)

/* One update_wavefield / update_wavefield_free_surface step. */
static void eupdate(const emed *M, efld *F, real dt) {
    const elay *Lp = &M->L;
    const elay L = *Lp;
    const int nx = L.nx, ny = L.ny, nz = L.nz, pml = L.pml, R = L.R, fs = L.fs;
    const float *c = M->cf;
    const size_t sx = 1, sy = L.sy, sz = L.sz;
    if (fs) {
        /* velocities above the free surface are zero (:580-590) */
        for (int kk = 1; kk <= R; ++kk)
            for (int j = L.lo; j < L.lo + L.NY; ++j)
                for (int i = L.lo; i < L.lo + L.NX; ++i) {
                    F->vx[IE(L, i, j, 1 - kk)] = 0;
                    F->vy[IE(L, i, j, 1 - kk)] = 0;
                    F->vz[IE(L, i, j, 2 - kk)] = 0;
                }
    }
    const int k0 = fs ? 1 : -pml + 1;
    /* normal stresses at integer nodes (:58-158, :592-705) */
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = k0; k <= nz + pml; ++k)
        for (int j = -pml + 1; j <= ny + pml; ++j)
            for (int i = -pml + 1; i <= nx + pml; ++i) {
                const size_t q = IE(L, i, j, k);
                real pdxvx = M->idx * dplus(F->vx, q, sx, c, R);
                real pdyvy = M->idy * dplus(F->vy, q, sy, c, R);
                real pdzvz = 0;
                const int pt = inbox(Lp, i, j, k);
                if (fs && k == 1) {
                    if (i <= 0 || i >= nx + 1 || j <= 0 || j >= ny + 1) {
                        DAMP_X(F->m_dxvx, pdxvx, q, i, 0, pt);
                        DAMP_Y(F->m_dyvy, pdyvy, q, j, 0, pt);
                    }
                    pdzvz = -(M->c12[q] / M->c11[q] * pdxvx) - M->c12[q] / M->c11[q] * pdyvy;
                } else {
                    pdzvz = M->idz * dplus(F->vz, q, sz, c, R);
                    if (i <= 0 || i >= nx + 1 || j <= 0 || j >= ny + 1 || (!fs && k <= 0) || k >= nz + 1) {
                        DAMP_X(F->m_dxvx, pdxvx, q, i, 0, pt);
                        DAMP_Y(F->m_dyvy, pdyvy, q, j, 0, pt);
                        DAMP_Z(F->m_dzvz, pdzvz, q, k, 0, pt);
                    }
                }
                const real c11 = M->c11[q], c12 = M->c12[q];
                F->sxx[q] = F->sxx[q] + dt * (c11 * pdxvx + c12 * pdyvy + c12 * pdzvz);
                F->syy[q] = F->syy[q] + dt * (c12 * pdxvx + c11 * pdyvy + c12 * pdzvz);
                F->szz[q] = F->szz[q] + dt * (c12 * pdxvx + c12 * pdyvy + c11 * pdzvz);
            }
    /* syz at (i, Y, Z) (:160-210, :708-745): mu = 4 / sum 1/c44(i, Y-1:Y, Z-1:Z) */
    const int z0s = fs ? 2 : -pml + 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int Z = z0s; Z <= nz + pml; ++Z)
        for (int Y = -pml + 1; Y <= ny + pml; ++Y)
            for (int i = -pml + 1; i <= nx + pml; ++i) {
                const size_t q = IE(L, i, Y, Z);
                real pdzvy = M->idz * dminus(F->vy, q, sz, c, R);
                real pdyvz = M->idy * dminus(F->vz, q, sy, c, R);
                const int pt = inbox(Lp, i, Y, Z);
                if (i <= 0 || i >= nx + 1 || Y <= 1 || Y >= ny + 1 || (!fs && Z <= 1) || Z >= nz + 1) {
                    DAMP_Y(F->m_dyvz, pdyvz, q, Y, 1, pt);
                    DAMP_Z(F->m_dzvy, pdzvy, q, Z, 1, pt);
                }
                const real mu = (real)4.0f / ((real)1.0f / M->c44[q - sy - sz] + (real)1.0f / M->c44[q - sz] +
                                              (real)1.0f / M->c44[q - sy] + (real)1.0f / M->c44[q]);
                F->syz[q] = F->syz[q] + dt * mu * (pdzvy + pdyvz);
            }
    /* sxz at (X, j, Z) (:212-258, :747-783) */
#pragma omp parallel for collapse(2) schedule(static)
    for (int Z = z0s; Z <= nz + pml; ++Z)
        for (int j = -pml + 1; j <= ny + pml; ++j)
            for (int X = -pml + 1; X <= nx + pml; ++X) {
                const size_t q = IE(L, X, j, Z);
                real pdzvx = M->idz * dminus(F->vx, q, sz, c, R);
                real pdxvz = M->idx * dminus(F->vz, q, sx, c, R);
                const int pt = inbox(Lp, X, j, Z);
                if (X <= 1 || X >= nx + 1 || j <= 0 || j >= ny + 1 || (!fs && Z <= 1) || Z >= nz + 1) {
                    DAMP_X(F->m_dxvz, pdxvz, q, X, 1, pt);
                    DAMP_Z(F->m_dzvx, pdzvx, q, Z, 1, pt);
                }
                const real mu = (real)4.0f / ((real)1.0f / M->c44[q - sx - sz] + (real)1.0f / M->c44[q - sz] +
                                              (real)1.0f / M->c44[q - sx] + (real)1.0f / M->c44[q]);
                F->sxz[q] = F->sxz[q] + dt * mu * (pdzvx + pdxvz);
            }
    /* sxy at (X, Y, k) (:260-306, :785-821) */
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = k0; k <= nz + pml; ++k)
        for (int Y = -pml + 1; Y <= ny + pml; ++Y)
            for (int X = -pml + 1; X <= nx + pml; ++X) {
                const size_t q = IE(L, X, Y, k);
                real pdyvx = M->idy * dminus(F->vx, q, sy, c, R);
                real pdxvy = M->idx * dminus(F->vy, q, sx, c, R);
                const int pt = inbox(Lp, X, Y, k);
                if (X <= 1 || X >= nx + 1 || Y <= 1 || Y >= ny + 1 || (!fs && (k <= 0 || k >= nz + 1))) {
                    DAMP_X(F->m_dxvy, pdxvy, q, X, 1, pt);
                    DAMP_Y(F->m_dyvx, pdyvx, q, Y, 1, pt);
                }
                const real mu = (real)4.0f / ((real)1.0f / M->c44[q - sx - sy] + (real)1.0f / M->c44[q - sy] +
                                              (real)1.0f / M->c44[q - sx] + (real)1.0f / M->c44[q]);
                F->sxy[q] = F->sxy[q] + dt * mu * (pdyvx + pdxvy);
            }
    if (fs) {
        /* stress mirror about the free surface (:823-845) */
        for (int kk = 1; kk <= R; ++kk)
            for (int j = L.lo; j < L.lo + L.NY; ++j)
                for (int i = L.lo; i < L.lo + L.NX; ++i) {
                    F->szz[IE(L, i, j, 1 - kk)] = -F->szz[IE(L, i, j, 1 + kk)];
                    F->szz[IE(L, i, j, 1)] = 0;
                    F->sxx[IE(L, i, j, 1 - kk)] = -F->sxx[IE(L, i, j, 1 + kk)];
                    F->syy[IE(L, i, j, 1 - kk)] = -F->syy[IE(L, i, j, 1 + kk)];
                    F->sxy[IE(L, i, j, 1 - kk)] = -F->sxy[IE(L, i, j, 1 + kk)];
                    F->syz[IE(L, i, j, 2 - kk)] = -F->syz[IE(L, i, j, 1 + kk)];
                    F->sxz[IE(L, i, j, 2 - kk)] = -F->sxz[IE(L, i, j, 1 + kk)];
                }
    }
    /* vx at (X, j, k) (:316-377, :852-900): rho_eff = 0.5 (rho(X-1) + rho(X)) */
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = k0; k <= nz + pml; ++k)
        for (int j = -pml + 1; j <= ny + pml; ++j)
            for (int X = -pml + 1; X <= nx + pml; ++X) {
                const size_t q = IE(L, X, j, k);
                real pdxxx = M->idx * dminus(F->sxx, q, sx, c, R);
                real pdyxy = M->idy * dplus(F->sxy, q, sy, c, R);
                real pdzxz = M->idz * dplus(F->sxz, q, sz, c, R);
                const int pt = inbox(Lp, X, j, k);
                if (X <= 1 || X >= nx + 1 || j <= 0 || j >= ny + 1 || (!fs && k <= 0) || k >= nz + 1) {
                    DAMP_X(F->m_dxxx, pdxxx, q, X, 1, pt);
                    DAMP_Y(F->m_dyxy, pdyxy, q, j, 0, pt);
                    DAMP_Z(F->m_dzxz, pdzxz, q, k, 0, pt);
                }
                const real re = (real)0.5f * (M->rho[q - sx] + M->rho[q]);
                F->vx[q] = F->vx[q] + dt / re * (pdxxx + pdyxy + pdzxz);
            }
    /* vy at (i, Y, k) (:379-440) */
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = k0; k <= nz + pml; ++k)
        for (int Y = -pml + 1; Y <= ny + pml; ++Y)
            for (int i = -pml + 1; i <= nx + pml; ++i) {
                const size_t q = IE(L, i, Y, k);
                real pdxxy = M->idx * dplus(F->sxy, q, sx, c, R);
                real pdyyy = M->idy * dminus(F->syy, q, sy, c, R);
                real pdzyz = M->idz * dplus(F->syz, q, sz, c, R);
                const int pt = inbox(Lp, i, Y, k);
                if (i <= 0 || i >= nx + 1 || Y <= 1 || Y >= ny + 1 || (!fs && k <= 0) || k >= nz + 1) {
                    DAMP_X(F->m_dxxy, pdxxy, q, i, 0, pt);
                    DAMP_Y(F->m_dyyy, pdyyy, q, Y, 1, pt);
                    DAMP_Z(F->m_dzyz, pdzyz, q, k, 0, pt);
                }
                const real re = (real)0.5f * (M->rho[q - sy] + M->rho[q]);
                F->vy[q] = F->vy[q] + dt / re * (pdxxy + pdyyy + pdzyz);
            }
    /* vz at (i, j, Z) (:442-503) */
    const int z0v = fs ? 2 : -pml + 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int Z = z0v; Z <= nz + pml; ++Z)
        for (int j = -pml + 1; j <= ny + pml; ++j)
            for (int i = -pml + 1; i <= nx + pml; ++i) {
                const size_t q = IE(L, i, j, Z);
                real pdxxz = M->idx * dplus(F->sxz, q, sx, c, R);
                real pdyyz = M->idy * dplus(F->syz, q, sy, c, R);
                real pdzzz = M->idz * dminus(F->szz, q, sz, c, R);
                const int pt = inbox(Lp, i, j, Z);
                if (i <= 0 || i >= nx + 1 || j <= 0 || j >= ny + 1 || (!fs && Z <= 1) || Z >= nz + 1) {
                    DAMP_X(F->m_dxxz, pdxxz, q, i, 0, pt);
                    DAMP_Y(F->m_dyyz, pdyyz, q, j, 0, pt);
                    DAMP_Z(F->m_dzzz, pdzzz, q, Z, 1, pt);
                }
                const real re = (real)0.5f * (M->rho[q - sz] + M->rho[q]);
                F->vz[q] = F->vz[q] + dt / re * (pdxxz + pdyyz + pdzzz);
            }
}

/* add_source (module_wavefield.f90:1115-1228): a force adds to vx, vy, vz
 * with amp / rho(gx, gy, gz); an explosion subtracts amp / (dx dy dz) from
 * sxx, syy, szz on (gx, gy, gz) with ix iy iz (its stf is the time
 * derivative, time_integration = -1); amp = stf amp dt. */
static void eadd_source(const emed *M, efld *F, const ora_esource3d *src, const ora_point *pts, int ns, int t,
                        real dt) {
    const elay L = M->L;
    const int fs = L.fs;
    for (int s = 0; s < ns; ++s) {
        const int nbeg = (int)roundf(src[s].t0 / (float)dt) + 1, nend = nbeg + src[s].nt_stf - 1;
        if (t < nbeg || t > nend) continue;
        real amp = src[s].stf[t - nbeg] * (real)src[s].amp * dt;
        const ora_point *P = &pts[s];
        if (src[s].mechanism == 0) {
            amp = amp / ((real)M->dx * (real)M->dy * (real)M->dz);
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int iry = -NKW; iry <= NKW; ++iry)
                    for (int irx = -NKW; irx <= NKW; ++irx)
                        if (!fs || P->gz + irz >= 2) {
                            const size_t q = IE(L, P->gx + irx, P->gy + iry, P->gz + irz);
                            const real w = amp * (real)P->ix[irx + NKW] * (real)P->iy[iry + NKW] *
                                           (real)P->iz[irz + NKW];
                            F->sxx[q] = F->sxx[q] - w;
                            F->syy[q] = F->syy[q] - w;
                            F->szz[q] = F->szz[q] - w;
                        }
            continue;
        }
        amp = amp / M->rho[IE(L, P->gx, P->gy, P->gz)];
#if defined(ORA_F64)
        const real sp = sin((double)src[s].polar), cp = cos((double)src[s].polar);
        const real sa = sin((double)src[s].azimuth), ca = cos((double)src[s].azimuth);
#else
        const real sp = sinf(src[s].polar), cp = cosf(src[s].polar);
        const real sa = sinf(src[s].azimuth), ca = cosf(src[s].azimuth);
#endif
        for (int irz = -NKW; irz <= NKW; ++irz)
            for (int iry = -NKW; iry <= NKW; ++iry)
                for (int irx = -NKW; irx <= NKW; ++irx)
                    if (!fs || P->gz + irz >= 2) {
                        const size_t q = IE(L, P->hx + irx, P->gy + iry, P->gz + irz);
                        F->vx[q] = F->vx[q] + sp * ca * amp * (real)P->hxw[irx + NKW] * (real)P->iy[iry + NKW] *
                                                  (real)P->iz[irz + NKW];
                    }
        for (int irz = -NKW; irz <= NKW; ++irz)
            for (int iry = -NKW; iry <= NKW; ++iry)
                for (int irx = -NKW; irx <= NKW; ++irx)
                    if (!fs || P->gz + irz >= 2) {
                        const size_t q = IE(L, P->gx + irx, P->hy + iry, P->gz + irz);
                        F->vy[q] = F->vy[q] + sp * sa * amp * (real)P->ix[irx + NKW] * (real)P->hyw[iry + NKW] *
                                                  (real)P->iz[irz + NKW];
                    }
        for (int irz = -NKW; irz <= NKW; ++irz)
            for (int iry = -NKW; iry <= NKW; ++iry)
                for (int irx = -NKW; irx <= NKW; ++irx)
                    if (!fs || P->hz + irz >= 2) {
                        const size_t q = IE(L, P->gx + irx, P->gy + iry, P->hz + irz);
                        F->vz[q] = F->vz[q] + cp * amp * (real)P->ix[irx + NKW] * (real)P->iy[iry + NKW] *
                                                  (real)P->hzw[irz + NKW];
                    }
    }
}

/* compute_forward (submodule_forward_modeling.f90:31-140): per step
 * [free surface: mirror the velocities above it] -> record vx, vy, vz ->
 * update -> source.  traces: three [nr][nt] blocks (x, y, z). */
int ORA_SYM(elastic3d_forward)(const ora_egrid3d *g, const float *vp, const float *vs, const float *rho, int ns,
                               const ora_esource3d *src, float f0max, int nr, const ora_ereceiver3d *rec,
                               real *traces) {
    if (g->R < 2 || g->R > 8 || g->pml < g->R) return -1;
    for (int s = 0; s < ns; ++s)
        if (src[s].mechanism != 0 && src[s].mechanism != 1) return -2;
    emed M;
    emed_init(&M, g, vp, vs, rho, f0max);
    const elay L = M.L;
    efld F;
    efld_alloc(&F, &L);
    ora_point *sp = (ora_point *)calloc(ns > 0 ? ns : 1, sizeof(ora_point));
    ora_point *rp = (ora_point *)calloc(nr > 0 ? nr : 1, sizeof(ora_point));
    /* with a free surface a source above the first node below it moves to
     * z = dz with amplitude x 1.2 (module_vars.f90:650-655) */
    ora_esource3d *s2 = (ora_esource3d *)calloc(ns > 0 ? ns : 1, sizeof(ora_esource3d));
    for (int s = 0; s < ns; ++s) {
        s2[s] = src[s];
        if (g->free_surface && s2[s].z < g->dz) {
            s2[s].z = g->dz;
            s2[s].amp = s2[s].amp * 1.2f;
        }
        ORA_SYM(locate_point)(3, s2[s].x, s2[s].y, s2[s].z, g->ox, g->oy, g->oz, g->dx, g->dy, g->dz, &sp[s]);
    }
    for (int r = 0; r < nr; ++r)
        ORA_SYM(locate_point)(3, rec[r].x, rec[r].y, rec[r].z, g->ox, g->oy, g->oz, g->dx, g->dy, g->dz, &rp[r]);
    const int nt = g->nt;
    memset(traces, 0, sizeof(real) * 3 * (size_t)nr * nt);
    real *tx = traces, *ty = traces + (size_t)nr * nt, *tz = traces + 2 * (size_t)nr * nt;
    const real dt = (real)g->dt;
    for (int t = 1; t <= nt; ++t) {
        if (g->free_surface)
            for (int kk = 1; kk <= NKW; ++kk)
                for (int j = L.lo; j < L.lo + L.NY; ++j)
                    for (int i = L.lo; i < L.lo + L.NX; ++i) {
                        F.vx[IE(L, i, j, 1 - kk)] = F.vx[IE(L, i, j, 1 + kk)];
                        F.vy[IE(L, i, j, 1 - kk)] = F.vy[IE(L, i, j, 1 + kk)];
                        F.vz[IE(L, i, j, 2 - kk)] = F.vz[IE(L, i, j, 1 + kk)];
                    }
#pragma omp parallel for schedule(static)
        for (int r = 0; r < nr; ++r) {
            if (rec[r].weight == 0.0f) continue;
            const ora_point *P = &rp[r];
            const real w = (real)rec[r].weight;
            real ax = 0, ay = 0, az = 0;
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int iry = -NKW; iry <= NKW; ++iry)
                    for (int irx = -NKW; irx <= NKW; ++irx)
                        ax = ax + F.vx[IE(L, P->hx + irx, P->gy + iry, P->gz + irz)] * (real)P->hxw[irx + NKW] *
                                      (real)P->iy[iry + NKW] * (real)P->iz[irz + NKW] * w;
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int iry = -NKW; iry <= NKW; ++iry)
                    for (int irx = -NKW; irx <= NKW; ++irx)
                        ay = ay + F.vy[IE(L, P->gx + irx, P->hy + iry, P->gz + irz)] * (real)P->ix[irx + NKW] *
                                      (real)P->hyw[iry + NKW] * (real)P->iz[irz + NKW] * w;
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int iry = -NKW; iry <= NKW; ++iry)
                    for (int irx = -NKW; irx <= NKW; ++irx)
                        az = az + F.vz[IE(L, P->gx + irx, P->gy + iry, P->hz + irz)] * (real)P->ix[irx + NKW] *
                                      (real)P->iy[iry + NKW] * (real)P->hzw[irz + NKW] * w;
            tx[(size_t)r * nt + (t - 1)] = ax;
            ty[(size_t)r * nt + (t - 1)] = ay;
            tz[(size_t)r * nt + (t - 1)] = az;
        }
        eupdate(&M, &F, dt);
        eadd_source(&M, &F, s2, sp, ns, t, dt);
    }
    free(sp); free(rp); free(s2);
    efld_free(&F);
    emed_free(&M);
    return 0;
}

/* ======================================================================== */
/* OWL's elastic gradient (isotropic): compute_adjoint + compute_gradient,   */
/* elastic_vhtiort_3d/submodule_fwi_gradient.f90:31-420, 702-827, energy   */
/* preconditioning off, medium parameters only.                             */
/* ======================================================================== */

/* Boundary strips of the three velocities over the six face regions
 * (elastic_vhtiort_3d/module_boundary_saving.f90:83-236): per velocity the
 * x faces [2-R, 1] and [nx+1, nx+R], y faces, z faces, each over
 * [1-R, n+R] on the other two axes; vx's six sections, then vy's, vz's. */
static size_t estrip_len(const elay *L) {
    const size_t R = (size_t)L->R, ex = L->nx + 2 * R, ey = L->ny + 2 * R, ez = L->nz + 2 * R;
    return 3 * 2 * R * (ey * ez + ex * ez + ex * ey);
}

static void estrips(const elay *L, efld *F, real *rec, int save) {
    const int R = L->R, nx = L->nx, ny = L->ny, nz = L->nz;
    size_t k = 0;
    real *v3[3] = {F->vx, F->vy, F->vz};
    for (int c = 0; c < 3; ++c) {
        real *f = v3[c];
#define XF(i, j, kk)                         \
    do {                                     \
        size_t q_ = IE(*L, i, j, kk);        \
        if (save) rec[k] = f[q_];            \
        else f[q_] = rec[k];                 \
        ++k;                                 \
    } while (0)
        for (int kk = 1 - R; kk <= nz + R; ++kk)
            for (int j = 1 - R; j <= ny + R; ++j)
                for (int i = 2 - R; i <= 1; ++i) XF(i, j, kk);
        for (int kk = 1 - R; kk <= nz + R; ++kk)
            for (int j = 1 - R; j <= ny + R; ++j)
                for (int i = nx + 1; i <= nx + R; ++i) XF(i, j, kk);
        for (int kk = 1 - R; kk <= nz + R; ++kk)
            for (int j = 2 - R; j <= 1; ++j)
                for (int i = 1 - R; i <= nx + R; ++i) XF(i, j, kk);
        for (int kk = 1 - R; kk <= nz + R; ++kk)
            for (int j = ny + 1; j <= ny + R; ++j)
                for (int i = 1 - R; i <= nx + R; ++i) XF(i, j, kk);
        for (int kk = 2 - R; kk <= 1; ++kk)
            for (int j = 1 - R; j <= ny + R; ++j)
                for (int i = 1 - R; i <= nx + R; ++i) XF(i, j, kk);
        for (int kk = nz + 1; kk <= nz + R; ++kk)
            for (int j = 1 - R; j <= ny + R; ++j)
                for (int i = 1 - R; i <= nx + R; ++i) XF(i, j, kk);
#undef XF
    }
}

/* Isotropic compliances at a node (submodule_fwi_gradient.f90:73-95), with
 * c22 = c33 = c11 and c13 = c23 = c12, in the Fortran's expression order. */
typedef struct {
    real s11, s12, s13, s22, s23, s33, s44;
} ecompl;

static ecompl compliances(real c11, real c12, real c44) {
    const real c13 = c12, c22 = c11, c23 = c12, c33 = c11;
    ecompl s;
    const real den1 = c13 * c13 * c22 - 2 * c12 * c13 * c23 + c11 * (c23 * c23) + c12 * c12 * c33 - c11 * c22 * c33;
    const real den2 = c13 * c13 * c22 - 2 * c12 * c13 * c23 + c12 * c12 * c33 + c11 * (c23 * c23 - c22 * c33);
    s.s11 = (c23 * c23 - c22 * c33) / den1;
    s.s12 = (-c13 * c23 + c12 * c33) / den2;
    s.s13 = (c13 * c22 - c12 * c23) / den1;
    s.s22 = (c13 * c13 - c11 * c33) / den1;
    s.s23 = (-c12 * c13 + c11 * c23) / den2;
    s.s33 = (c12 * c12 - c11 * c22) / den1;
    s.s44 = c44 != 0 ? (real)1.0f / c44 : (real)0;
    return s;
}

/* adj: three [nr][nt] blocks (x, y, z).  Outputs (nz, ny, nx) z-fast:
 * grad_vp, grad_vs (chain rule with vp, vs in km/s and rho in g/cm^3,
 * :486-495) and grad_rho, NaN/Inf mapped to zero. */
int ORA_SYM(elastic3d_gradient_owl)(const ora_egrid3d *g, const float *vp, const float *vs, const float *rho,
                                    int ns, const ora_esource3d *src, float f0max, int hnt, int nr,
                                    const ora_ereceiver3d *rec, const real *adj, real *grad_vp, real *grad_vs,
                                    real *grad_rho, real *traces_out) {
    if (g->R < 2 || g->R > 8 || g->pml < g->R || ns < 1) return -1;
    for (int s = 0; s < ns; ++s)
        if (src[s].mechanism != 0 && src[s].mechanism != 1) return -2;
    emed M;
    emed_init(&M, g, vp, vs, rho, f0max);
    const elay L = M.L;
    const size_t n = (size_t)L.NX * L.NY * L.NZ, sl = estrip_len(&L);
    const int nt = g->nt, fs = g->free_surface;
    real *strips = ezalloc(sl * (size_t)nt);
    efld F, A, P;
    efld_alloc(&F, &L);
    efld_alloc(&A, &L);
    efld_alloc(&P, &L);
    ora_point *sp = (ora_point *)calloc((size_t)ns, sizeof(ora_point));
    ora_point *rp = (ora_point *)calloc(nr > 0 ? nr : 1, sizeof(ora_point));
    ora_esource3d *s2 = (ora_esource3d *)calloc((size_t)ns, sizeof(ora_esource3d));
    for (int s = 0; s < ns; ++s) {
        s2[s] = src[s];
        if (fs && s2[s].z < g->dz) {
            s2[s].z = g->dz;
            s2[s].amp = s2[s].amp * 1.2f;
        }
        ORA_SYM(locate_point)(3, s2[s].x, s2[s].y, s2[s].z, g->ox, g->oy, g->oz, g->dx, g->dy, g->dz, &sp[s]);
    }
    for (int r = 0; r < nr; ++r)
        ORA_SYM(locate_point)(3, rec[r].x, rec[r].y, rec[r].z, g->ox, g->oy, g->oz, g->dx, g->dy, g->dz, &rp[r]);
    real *tr = traces_out ? traces_out : ezalloc(3 * (size_t)nr * nt);
    memset(tr, 0, sizeof(real) * 3 * (size_t)nr * nt);
    const real dt = (real)g->dt;
    /* forward with boundary saving; the final state stays in F */
    for (int t = 1; t <= nt; ++t) {
        estrips(&L, &F, strips + (size_t)(t - 1) * sl, 1);
        if (t == nt) break;   /* final state: the state at the start of step nt */
        if (fs)
            for (int kk = 1; kk <= NKW; ++kk)
                for (int j = L.lo; j < L.lo + L.NY; ++j)
                    for (int i = L.lo; i < L.lo + L.NX; ++i) {
                        F.vx[IE(L, i, j, 1 - kk)] = F.vx[IE(L, i, j, 1 + kk)];
                        F.vy[IE(L, i, j, 1 - kk)] = F.vy[IE(L, i, j, 1 + kk)];
                        F.vz[IE(L, i, j, 2 - kk)] = F.vz[IE(L, i, j, 1 + kk)];
                    }
        for (int r = 0; r < nr; ++r) {
            if (rec[r].weight == 0.0f) continue;
            const ora_point *Q = &rp[r];
            const real w = (real)rec[r].weight;
            real ax = 0, ay = 0, az = 0;
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int iry = -NKW; iry <= NKW; ++iry)
                    for (int irx = -NKW; irx <= NKW; ++irx) {
                        ax = ax + F.vx[IE(L, Q->hx + irx, Q->gy + iry, Q->gz + irz)] * (real)Q->hxw[irx + NKW] *
                                      (real)Q->iy[iry + NKW] * (real)Q->iz[irz + NKW] * w;
                    }
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int iry = -NKW; iry <= NKW; ++iry)
                    for (int irx = -NKW; irx <= NKW; ++irx)
                        ay = ay + F.vy[IE(L, Q->gx + irx, Q->hy + iry, Q->gz + irz)] * (real)Q->ix[irx + NKW] *
                                      (real)Q->hyw[iry + NKW] * (real)Q->iz[irz + NKW] * w;
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int iry = -NKW; iry <= NKW; ++iry)
                    for (int irx = -NKW; irx <= NKW; ++irx)
                        az = az + F.vz[IE(L, Q->gx + irx, Q->gy + iry, Q->hz + irz)] * (real)Q->ix[irx + NKW] *
                                      (real)Q->iy[iry + NKW] * (real)Q->hzw[irz + NKW] * w;
            tr[(size_t)r * nt + (t - 1)] = ax;
            tr[((size_t)nr + r) * nt + (t - 1)] = ay;
            tr[(2 * (size_t)nr + r) * nt + (t - 1)] = az;
        }
        eupdate(&M, &F, dt);
        eadd_source(&M, &F, s2, sp, ns, t, dt);
    }
    /* final state: the nine wavefields at the start of step nt, before the
     * receivers' mirror (module_boundary_saving.f90:62-70) */
    real *fin = ezalloc(9 * n);
    real **fw = &F.sxx;
    for (int c = 0; c < 9; ++c) memcpy(fin + (size_t)c * n, fw[c], n * sizeof(real));
    /* the record of step nt (state nt-1, FS-mirrored) */
    if (fs)
        for (int kk = 1; kk <= NKW; ++kk)
            for (int j = L.lo; j < L.lo + L.NY; ++j)
                for (int i = L.lo; i < L.lo + L.NX; ++i) {
                    F.vx[IE(L, i, j, 1 - kk)] = F.vx[IE(L, i, j, 1 + kk)];
                    F.vy[IE(L, i, j, 1 - kk)] = F.vy[IE(L, i, j, 1 + kk)];
                    F.vz[IE(L, i, j, 2 - kk)] = F.vz[IE(L, i, j, 1 + kk)];
                }
    for (int r = 0; r < nr; ++r) {
        if (rec[r].weight == 0.0f) continue;
        const ora_point *Q = &rp[r];
        const real w = (real)rec[r].weight;
        real ax = 0, ay = 0, az = 0;
        for (int irz = -NKW; irz <= NKW; ++irz)
            for (int iry = -NKW; iry <= NKW; ++iry)
                for (int irx = -NKW; irx <= NKW; ++irx)
                    ax = ax + F.vx[IE(L, Q->hx + irx, Q->gy + iry, Q->gz + irz)] * (real)Q->hxw[irx + NKW] *
                                  (real)Q->iy[iry + NKW] * (real)Q->iz[irz + NKW] * w;
        for (int irz = -NKW; irz <= NKW; ++irz)
            for (int iry = -NKW; iry <= NKW; ++iry)
                for (int irx = -NKW; irx <= NKW; ++irx)
                    ay = ay + F.vy[IE(L, Q->gx + irx, Q->hy + iry, Q->gz + irz)] * (real)Q->ix[irx + NKW] *
                                  (real)Q->hyw[iry + NKW] * (real)Q->iz[irz + NKW] * w;
        for (int irz = -NKW; irz <= NKW; ++irz)
            for (int iry = -NKW; iry <= NKW; ++iry)
                for (int irx = -NKW; irx <= NKW; ++irx)
                    az = az + F.vz[IE(L, Q->gx + irx, Q->gy + iry, Q->hz + irz)] * (real)Q->ix[irx + NKW] *
                                  (real)Q->iy[iry + NKW] * (real)Q->hzw[irz + NKW] * w;
        tr[(size_t)r * nt + (nt - 1)] = ax;
        tr[((size_t)nr + r) * nt + (nt - 1)] = ay;
        tr[(2 * (size_t)nr + r) * nt + (nt - 1)] = az;
    }
    efld_free(&F);
    efld_alloc(&F, &L);
    fw = &F.sxx;
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t ni = (size_t)nx * ny * nz;
    real *gc[10];
    for (int c = 0; c < 10; ++c) gc[c] = ezalloc(ni);  /* c11 c12 c13 c22 c23 c33 c44 c55 c66 rho */
    const real mdt = -dt;
    const int cc = g->cc_step_interval > 0 ? g->cc_step_interval : 1;
    const size_t sx = 1, sy = L.sy, sz = L.sz;
    for (int t = nt; t >= hnt; --t) {
        real **pw = &P.sxx;
        for (int c = 0; c < 9; ++c) memcpy(pw[c], fw[c], n * sizeof(real));
        eupdate(&M, &F, mdt);
        if (t == nt)
            for (int c = 0; c < 9; ++c) memcpy(fw[c], fin + (size_t)c * n, n * sizeof(real));
        estrips(&L, &F, strips + (size_t)(t - 1) * sl, 0);
        eupdate(&M, &A, mdt);
        /* adjoint sources into the adjoint velocities (:256-330) */
        for (int r = 0; r < nr; ++r) {
            if (rec[r].weight == 0.0f) continue;
            const ora_point *Q = &rp[r];
            const real w = (real)rec[r].weight;
            const real ax = adj[(size_t)r * nt + (t - 1)], ay = adj[((size_t)nr + r) * nt + (t - 1)],
                       az = adj[(2 * (size_t)nr + r) * nt + (t - 1)];
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int iry = -NKW; iry <= NKW; ++iry)
                    for (int irx = -NKW; irx <= NKW; ++irx)
                        if (!fs || Q->gz + irz >= 2) {
                            const size_t q = IE(L, Q->hx + irx, Q->gy + iry, Q->gz + irz);
                            A.vx[q] = A.vx[q] + ax * (real)Q->hxw[irx + NKW] * (real)Q->iy[iry + NKW] *
                                                    (real)Q->iz[irz + NKW] * w;
                        }
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int iry = -NKW; iry <= NKW; ++iry)
                    for (int irx = -NKW; irx <= NKW; ++irx)
                        if (!fs || Q->gz + irz >= 2) {
                            const size_t q = IE(L, Q->gx + irx, Q->hy + iry, Q->gz + irz);
                            A.vy[q] = A.vy[q] + ay * (real)Q->ix[irx + NKW] * (real)Q->hyw[iry + NKW] *
                                                    (real)Q->iz[irz + NKW] * w;
                        }
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int iry = -NKW; iry <= NKW; ++iry)
                    for (int irx = -NKW; irx <= NKW; ++irx)
                        if (!fs || Q->hz + irz >= 2) {
                            const size_t q = IE(L, Q->gx + irx, Q->gy + iry, Q->hz + irz);
                            A.vz[q] = A.vz[q] + az * (real)Q->ix[irx + NKW] * (real)Q->iy[iry + NKW] *
                                                    (real)Q->hzw[irz + NKW] * w;
                        }
        }
        if (t % cc != 0) continue;
        /* compute_gradient (:702-827) over the interior */
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = 1; k <= nz; ++k)
            for (int j = 1; j <= ny; ++j)
                for (int i = 1; i <= nx; ++i) {
                    const size_t q = IE(L, i, j, k);
                    const size_t o = (size_t)(i - 1) + (size_t)(j - 1) * nx + (size_t)(k - 1) * nx * ny;
                    const ecompl s = compliances(M.c11[q], M.c12[q], M.c44[q]);
                    const real txx = F.sxx[q] - P.sxx[q], tyy = F.syy[q] - P.syy[q], tzz = F.szz[q] - P.szz[q];
                    const real tyz = (real)0.25f * (F.syz[q] + F.syz[q + sy] + F.syz[q + sz] + F.syz[q + sy + sz]) -
                                     (real)0.25f * (P.syz[q] + P.syz[q + sy] + P.syz[q + sz] + P.syz[q + sy + sz]);
                    const real txz = (real)0.25f * (F.sxz[q] + F.sxz[q + sx] + F.sxz[q + sz] + F.sxz[q + sx + sz]) -
                                     (real)0.25f * (P.sxz[q] + P.sxz[q + sx] + P.sxz[q + sz] + P.sxz[q + sx + sz]);
                    const real txy = (real)0.25f * (F.sxy[q] + F.sxy[q + sx] + F.sxy[q + sy] + F.sxy[q + sx + sy]) -
                                     (real)0.25f * (P.sxy[q] + P.sxy[q + sx] + P.sxy[q + sy] + P.sxy[q + sx + sy]);
                    const real rxx = A.sxx[q], ryy = A.syy[q], rzz = A.szz[q];
                    const real ryz = (real)0.25f * (A.syz[q] + A.syz[q + sy] + A.syz[q + sz] + A.syz[q + sy + sz]);
                    const real rxz = (real)0.25f * (A.sxz[q] + A.sxz[q + sx] + A.sxz[q + sz] + A.sxz[q + sx + sz]);
                    const real rxy = (real)0.25f * (A.sxy[q] + A.sxy[q + sx] + A.sxy[q + sy] + A.sxy[q + sx + sy]);
                    const real exx = s.s11 * txx + s.s12 * tyy + s.s13 * tzz;
                    const real eyy = s.s12 * txx + s.s22 * tyy + s.s23 * tzz;
                    const real ezz = s.s13 * txx + s.s23 * tyy + s.s33 * tzz;
                    const real eyz = s.s44 * tyz, exz = s.s44 * txz, exy = s.s44 * txy;
                    const real fxx = s.s11 * rxx + s.s12 * ryy + s.s13 * rzz;
                    const real fyy = s.s12 * rxx + s.s22 * ryy + s.s23 * rzz;
                    const real fzz = s.s13 * rxx + s.s23 * ryy + s.s33 * rzz;
                    const real fyz = s.s44 * ryz, fxz = s.s44 * rxz, fxy = s.s44 * rxy;
                    gc[0][o] = gc[0][o] - exx * fxx;
                    gc[1][o] = gc[1][o] - exx * fyy - eyy * fxx;
                    gc[2][o] = gc[2][o] - exx * fzz - ezz * fxx;
                    gc[3][o] = gc[3][o] - eyy * fyy;
                    gc[4][o] = gc[4][o] - eyy * fzz - ezz * fyy;
                    gc[5][o] = gc[5][o] - ezz * fzz;
                    gc[6][o] = gc[6][o] - eyz * fyz;
                    gc[7][o] = gc[7][o] - exz * fxz;
                    gc[8][o] = gc[8][o] - exy * fxy;
                    const real svx = (F.vx[q] + F.vx[q + sx]) - (P.vx[q] + P.vx[q + sx]);
                    const real rvx = A.vx[q] + A.vx[q + sx];
                    const real svy = (F.vy[q] + F.vy[q + sy]) - (P.vy[q] + P.vy[q + sy]);
                    const real rvy = A.vy[q] + A.vy[q + sy];
                    const real svz = (F.vz[q] + F.vz[q + sz]) - (P.vz[q] + P.vz[q + sz]);
                    const real rvz = A.vz[q] + A.vz[q + sz];
                    gc[9][o] = gc[9][o] + svx * rvx + svy * rvy + svz * rvz;
                }
    }
    /* chain rule (:486-495) and output (nz, ny, nx) z-fast */
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                const size_t o = (size_t)(i - 1) + (size_t)(j - 1) * nx + (size_t)(k - 1) * nx * ny;
                const size_t m = (size_t)(k - 1) + (size_t)(j - 1) * nz + (size_t)(i - 1) * nz * ny;
                const real r = (real)rho[m] * (real)1.0e-3f, p = (real)vp[m] * (real)1.0e-3f,
                           v = (real)vs[m] * (real)1.0e-3f;
                real a = gc[0][o] + gc[1][o] + gc[2][o] + gc[3][o] + gc[4][o] + gc[5][o];
                a = 2 * r * p * a;
                real b = gc[6][o] + gc[7][o] + gc[8][o] - 2 * gc[1][o] - 2 * gc[2][o] - 2 * gc[4][o];
                b = 2 * r * v * b;
                const real c = gc[9][o];
                grad_vp[m] = isfinite(a) ? a : 0;
                grad_vs[m] = isfinite(b) ? b : 0;
                grad_rho[m] = isfinite(c) ? c : 0;
            }
    for (int c = 0; c < 10; ++c) free(gc[c]);
    if (!traces_out) free(tr);
    free(fin); free(strips); free(sp); free(rp); free(s2);
    efld_free(&F); efld_free(&A); efld_free(&P);
    emed_free(&M);
    return 0;
}
