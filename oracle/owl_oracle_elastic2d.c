/*
 * owl_oracle_elastic2d.c -- CPU restatement of OWL's isotropic elastic
 * staggered-grid velocity-stress scheme in 2D (elastic_vhtiort_2d with
 * anisotropy_type = 'iso') and of OWL's elastic gradient on it.
 *
 * TEST INFRASTRUCTURE ONLY (see owl_oracle.c): only tests/, smoke() and the
 * CPU baseline may load it, as the checker.
 *
 * Restated from (paths relative to the OWL source tree):
 *   src/solver/elastic_vhtiort_2d/module_vars.f90:208-302   iso c_ij, pmlvp, padding
 *   src/solver/elastic_vhtiort_2d/module_cfspml.f90:285-578 damp_coef, coefficient arrays
 *   src/solver/elastic_vhtiort_2d/module_wavefield.f90:42-134 update_wavefield
 *   src/solver/elastic_vhtiort_2d/module_wavefield.f90 add_source (force)
 *   src/solver/elastic_vhtiort_2d/macro_fd_stencil.f90 (stencil offsets)
 *   src/solver/elastic_vhtiort_2d/submodule_forward_modeling.f90:31-180
 *   src/solver/elastic_vhtiort_2d/module_boundary_saving.f90:53-160
 *   src/solver/elastic_vhtiort_2d/submodule_fwi_gradient.f90:60-100, 138-270, 330-360, 440-520
 *
 * Scope and stated assumptions:
 *   - isotropic media: the M-PML damping ratios are zero and kappa = 1, so
 *     the per-cell coefficient arrays reduce to per-axis profiles (a cell
 *     whose own-axis distance is zero has a = 0 and keeps a zero memory
 *     variable, exactly as in the per-cell form);
 *   - the free surface runs, as in OWL, on the depth-refined mesh
 *     (module_vars.f90:310-371: a layer of spacing dz/refine over half a
 *     Rayleigh wavelength, then 2.5 % growth per row up to dz_max), with the
 *     medium mapped linearly onto it (map_regular_to_irregular, :139-170),
 *     z damping on the local spacing (module_cfspml.f90:398-480), the
 *     traction-free update and mirrors (update_wavefield_free_surface,
 *     module_wavefield.f90:143-289), sources and receivers located on the
 *     refined rows (prepare_geometry's depth-varying branch,
 *     src/lib/module_source_receiver.f90:596-680) and the gradients mapped
 *     back with a cubic column spline (map_irregular_to_regular,
 *     :175-203).  FLIT's regspace, deriv and ginterp come from
 *     owl_oracle_fsg2d.c (same stated assumptions);
 *   - force and explosion sources (the explosion's wavelet is the time
 *     derivative, made by the caller: stf_time_derivative);
 *   - Fortran sum() over 2x2 sections in array element order.
 * Compiled in float and double with owl_oracle.c's conventions.
 */
#include <float.h>
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

typedef struct {
    float x, y, z;
    int gx, gy, gz, hx, hy, hz;
    float ix[9], iy[9], iz[9], hxw[9], hyw[9], hzw[9];
} ora_point;
int ORA_SYM(fdcoefs)(int R, float *out);
float ORA_SYM(alphamax)(float f0max);
void ORA_SYM(kaiser_sinc)(float x, float *h);
float *ORA_SYM(flit_regspace)(float a, float d, float b, int *n);
void ORA_SYM(flit_deriv)(const float *x, int n, float *d);
void ORA_SYM(flit_ginterp)(const float *x, const float *y, int n, const float *xi, int m, float *yi, int cubic);
void ORA_SYM(locate_point)(int dims, float x, float y, float z, float ox, float oy, float oz, float dx, float dy,
                           float dz, ora_point *pt);

typedef struct {
    int nx, nz;
    float dx, dz, ox, oz;
    float dt;
    int nt, pml, R;
    int cc_step_interval;
} ora_egrid2d;

typedef struct {
    float x, z, amp, t0, polar; /* polar in radians */
    int mechanism;              /* 0 explosion, 1 force */
    int nt_stf;
    const real *stf;
} ora_esource2d;

typedef struct {
    float x, z, weight;
} ora_ereceiver2d;

/* free surface: free_surface_dz_refine, the caller's dz_max (OWL's
 * default 1.5 dz, module_model.f90:176-178) and the smallest source f0 of
 * the shot (the Rayleigh wavelength of the refined layer) */
typedef struct {
    float refine, dz_max, f0min;
} ora_esurf2d;

/* The depth-refined mesh of one window: rows j = -pml .. nz + pml + 1 at
 * [j + pml] (module_vars.f90:310-360). */
typedef struct {
    int nz, pml;
    float dz, rr, dzm;
    float *zi, *zh, *dzi, *dzh, *sci, *sch;
} emesh2;

static void emesh2_free(emesh2 *S) {
    free(S->zi);
    S->zi = S->zh = S->dzi = S->dzh = S->sci = S->sch = NULL;
}

/* rl = sqrt(mean(c55/rho over the padded first row)) / min f0; refine_nz =
 * ceiling(rl/2/dz); z = regspace(0, dz/rr, (refine_nz-1) dz), then rows
 * spaced min(1.025 dz_prev, dz_max) until reaching (nz-1) dz; half rows at
 * the midpoints with zh(1) = -dz/(2 rr); dz_i = deriv(z_i), dz_h =
 * deriv(z_h), scalings dz/dz_i, dz/dz_h, all padded by edge replication;
 * z_i, z_h extended by dz/rr above and by the last spacing below. */
static int emesh2_build(emesh2 *S, const ora_egrid2d *g, const ora_esurf2d *fs, const float *vs, const float *rho,
                        float f0min) {
    const int nx = g->nx, nz0 = g->nz, pml = g->pml;
    const float dz = g->dz;
    if (!(fs->refine > 0) || !(fs->dz_max > 0) || !(f0min > 0) || nz0 < 3) return -3;
    float sm = 0.0f;
    for (int i = -pml; i <= nx + pml + 1; ++i) {
        const int ic = i < 1 ? 1 : (i > nx ? nx : i);
        const size_t m = (size_t)(ic - 1) * nz0;
        const float r = rho[m], c55 = r * (vs[m] * vs[m]);
        sm = sm + c55 / r;
    }
    const float rl = sqrtf(sm / (float)(nx + 2 * pml + 2)) / f0min;
    const int rnz = (int)ceilf(0.5f * rl / dz);
    const float rr = fs->refine;
    float dzm = dz / rr;
    int n = 0;
    float *z0 = ORA_SYM(flit_regspace)(0.0f, dzm, (float)(rnz - 1) * dz, &n);
    if (n < 1) {
        free(z0);
        return -3;
    }
    size_t cap = (size_t)n + 1024;
    float *z = (float *)malloc(cap * sizeof(float));
    memcpy(z, z0, (size_t)n * sizeof(float));
    free(z0);
    while (z[n - 1] < (float)(nz0 - 1) * dz) {
        dzm = fminf(dzm * 1.025f, fs->dz_max);
        if ((size_t)n == cap) {
            cap *= 2;
            z = (float *)realloc(z, cap * sizeof(float));
        }
        z[n] = z[n - 1] + dzm;
        ++n;
        if (n > 10000000) {
            free(z);
            return -3;
        }
    }
    const int nz = n, nrj = nz + 2 * pml + 2;
    float *zh = (float *)malloc((size_t)nz * sizeof(float)), *di = (float *)malloc((size_t)nz * sizeof(float)),
          *dh = (float *)malloc((size_t)nz * sizeof(float));
    zh[0] = -0.5f * dz / rr;
    for (int i = 0; i < nz - 1; ++i) zh[i + 1] = 0.5f * (z[i] + z[i + 1]);
    ORA_SYM(flit_deriv)(z, nz, di);
    ORA_SYM(flit_deriv)(zh, nz, dh);
    S->zi = (float *)calloc(6 * (size_t)nrj, sizeof(float));
    S->zh = S->zi + nrj;
    S->dzi = S->zh + nrj;
    S->dzh = S->dzi + nrj;
    S->sci = S->dzh + nrj;
    S->sch = S->sci + nrj;
    for (int j = -pml; j <= nz + pml + 1; ++j) {
        const int jc = j < 1 ? 1 : (j > nz ? nz : j), k = j + pml;
        S->dzi[k] = di[jc - 1];
        S->dzh[k] = dh[jc - 1];
        S->sci[k] = dz / di[jc - 1];
        S->sch[k] = dz / dh[jc - 1];
        if (j >= 1 && j <= nz) {
            S->zi[k] = z[j - 1];
            S->zh[k] = zh[j - 1];
        }
    }
    for (int i = 1; i <= pml + 1; ++i) {
        S->zi[1 - i + pml] = S->zi[1 + pml] - (float)i * dz / rr;
        S->zi[nz + i + pml] = S->zi[nz + pml] + (float)i * dzm;
        S->zh[1 - i + pml] = S->zh[1 + pml] - (float)i * dz / rr;
        S->zh[nz + i + pml] = S->zh[nz + pml] + (float)i * dzm;
    }
    S->nz = nz;
    S->pml = pml;
    S->dz = dz;
    S->rr = rr;
    S->dzm = dzm;
    free(z);
    free(zh);
    free(di);
    free(dh);
    return 0;
}

typedef struct {
    int lo, NX, NZ, nx, nz, pml, R;
} lay2;

#define I2(L, i, j) ((size_t)((i) - (L).lo) + (size_t)((j) - (L).lo) * (size_t)(L).NX)

/* damp_coef (module_cfspml.f90:285-330) with ratio 0 and kmax = 1 along one
 * axis: damp = vp (n+1) ln(1/R0) / (2 pml d) nd^2, alpha = max(alphamax (1 -
 * nd), alphamax / 4), a = -2|dt| damp / k / (2 + |dt| (alpha + damp / k)),
 * b = (2 - |dt| (alpha + damp / k)) / (2 + |dt| (alpha + damp / k)). */
static void damp_ab(float pmlvp, float nd, float d, int pml, float adt, float alphamax, float *a, float *b) {
    const float R0 = 1.0e-5f, kx = 1.0f;
    const float damp = pmlvp * (2 + 1.0f) * logf(1.0f / R0) / (2.0f * (float)pml * d) * (nd * nd);
    const float alpha = fmaxf(alphamax * (1.0f - nd), 0.25f * alphamax);
    *a = -2.0f * adt * damp / kx / (2.0f + adt * (alpha + damp / kx));
    *b = (2.0f - adt * (alpha + damp / kx)) / (2.0f + adt * (alpha + damp / kx));
}

/* Per-axis profiles a/b at integer and half positions, index i + pml - 1,
 * set for i <= 1 and i >= n (module_cfspml.f90:466-575: distances
 * |i-1|/pml, |i-1-0.5|/pml, |i-n|/pml, |i-n-0.5|/pml formed in double). */
typedef struct {
    float *ai, *bi, *ah, *bh;
} prof2;

static void prof2_init(prof2 *P, int n, int pml, float d, float pmlvp, float adt, float alphamax) {
    const int len = n + 2 * pml;
    P->ai = (float *)calloc(4 * (size_t)len, sizeof(float));
    P->bi = P->ai + len;
    P->ah = P->bi + len;
    P->bh = P->ah + len;
    for (int i = -pml + 1; i <= n + pml; ++i) {
        float ndi, ndh;
        if (i <= 1) {
            ndi = (float)fabs((i - 1 + 0.0) / pml);
            ndh = (float)fabs((i - 1 - 0.5) / pml);
        } else if (i >= n) {
            ndi = (float)fabs((i - n + 0.0) / pml);
            ndh = (float)fabs((i - n - 0.5) / pml);
        } else {
            continue;
        }
        const int q = i + pml - 1;
        damp_ab(pmlvp, ndi, d, pml, adt, alphamax, &P->ai[q], &P->bi[q]);
        damp_ab(pmlvp, ndh, d, pml, adt, alphamax, &P->ah[q], &P->bh[q]);
    }
}

/* z profiles on the refined mesh: row j takes d = dz_i(j) for every grid
 * (module_cfspml.f90:446-473 pass dz_i(j) to all four damp_coef calls). */
static void prof2_init_rows(prof2 *P, int n, int pml, const float *dzi, float pmlvp, float adt, float alphamax) {
    const int len = n + 2 * pml;
    P->ai = (float *)calloc(4 * (size_t)len, sizeof(float));
    P->bi = P->ai + len;
    P->ah = P->bi + len;
    P->bh = P->ah + len;
    for (int i = -pml + 1; i <= n + pml; ++i) {
        float ndi, ndh;
        if (i <= 1) {
            ndi = (float)fabs((i - 1 + 0.0) / pml);
            ndh = (float)fabs((i - 1 - 0.5) / pml);
        } else if (i >= n) {
            ndi = (float)fabs((i - n + 0.0) / pml);
            ndh = (float)fabs((i - n - 0.5) / pml);
        } else {
            continue;
        }
        const int q = i + pml - 1;
        damp_ab(pmlvp, ndi, dzi[i + pml], pml, adt, alphamax, &P->ai[q], &P->bi[q]);
        damp_ab(pmlvp, ndh, dzi[i + pml], pml, adt, alphamax, &P->ah[q], &P->bh[q]);
    }
}

typedef struct {
    lay2 L;
    float c[8];
    real idx, idz;
    float dx, dz;
    real *c11, *c13, *c33, *c55, *mu, *rx, *rz, *rho;
    prof2 px, pz;
    const emesh2 *S; /* free surface: the refined mesh, else NULL */
} emed2;

typedef struct {
    real *sxx, *szz, *sxz, *vx, *vz;
    real *m_dxvx, *m_dzvz, *m_dxvz, *m_dzvx, *m_dxxx, *m_dzxz, *m_dxxz, *m_dzzz;
} efld2;

static real *zal(size_t n) { return (real *)calloc(n, sizeof(real)); }

static void efld2_alloc(efld2 *F, const lay2 *L) {
    const size_t n = (size_t)L->NX * L->NZ;
    real **p = &F->sxx;
    for (int k = 0; k < 13; ++k) p[k] = zal(n);
}
static void efld2_free(efld2 *F) {
    real **p = &F->sxx;
    for (int k = 0; k < 13; ++k) free(p[k]);
}

/* vp, vs, rho: (nz, nx) z fastest.  c33 = rho vp^2, c55 = rho vs^2,
 * c11 = c33, c13 = c33 - 2 c55, padded by edge replication; pmlvp = sqrt of
 * the largest c11/rho, c33/rho on the four edge lines. */
static void emed2_init(emed2 *M, const ora_egrid2d *g, const float *vp, const float *vs, const float *rho,
                       float f0max) {
    const int R = g->R, pml = g->pml, nx = g->nx, nz = g->nz;
    lay2 L = {1 - pml - R, nx + 2 * (pml + R), nz + 2 * (pml + R), nx, nz, pml, R};
    M->L = L;
    M->dx = g->dx;
    M->dz = g->dz;
    M->S = NULL;
    ORA_SYM(fdcoefs)(R, M->c);
    M->idx = (real)(float)(1.0 / (double)g->dx);
    M->idz = (real)(float)(1.0 / (double)g->dz);
    const size_t n = (size_t)L.NX * L.NZ;
    M->c11 = zal(n); M->c13 = zal(n); M->c33 = zal(n); M->c55 = zal(n);
    M->mu = zal(n); M->rx = zal(n); M->rz = zal(n); M->rho = zal(n);
    float pm = 0.0f;
    for (int i = 1; i <= nx; ++i)
        for (int j = 1; j <= nz; ++j) {
            if (!(i == 1 || i == nx || j == 1 || j == nz)) continue;
            const size_t m = (size_t)(j - 1) + (size_t)(i - 1) * nz;
            const float r = rho[m], c33 = r * (vp[m] * vp[m]), c55 = r * (vs[m] * vs[m]), c11 = c33;
            if (c11 / r > pm) pm = c11 / r;
            if (c33 / r > pm) pm = c33 / r;
            (void)c55;
        }
    const float pmlvp = sqrtf(pm);
    for (int j = L.lo; j < L.lo + L.NZ; ++j)
        for (int i = L.lo; i < L.lo + L.NX; ++i) {
            const int ic = i < 1 ? 1 : (i > nx ? nx : i), jc = j < 1 ? 1 : (j > nz ? nz : j);
            const size_t m = (size_t)(jc - 1) + (size_t)(ic - 1) * nz, q = I2(L, i, j);
            const float r = rho[m];
            const float c33 = r * (vp[m] * vp[m]), c55 = r * (vs[m] * vs[m]);
            M->c33[q] = c33;
            M->c55[q] = c55;
            M->c11[q] = c33;
            M->c13[q] = c33 - 2 * c55;
            M->rho[q] = r;
        }
    for (int j = L.lo + 1; j < L.lo + L.NZ; ++j)
        for (int i = L.lo + 1; i < L.lo + L.NX; ++i) {
            const size_t q = I2(L, i, j), sx = 1, sz = (size_t)L.NX;
            /* sxz at storage (m, n) sits between nodes m-1..m, n-1..n */
            M->mu[q] = (real)4.0f / ((real)1.0f / M->c55[q - sx - sz] + (real)1.0f / M->c55[q - sz] +
                                     (real)1.0f / M->c55[q - sx] + (real)1.0f / M->c55[q]);
            M->rx[q] = (real)0.5f * (M->rho[q - sx] + M->rho[q]);
            M->rz[q] = (real)0.5f * (M->rho[q - sz] + M->rho[q]);
        }
    const float am = ORA_SYM(alphamax)(f0max), adt = fabsf(g->dt);
    prof2_init(&M->px, nx, pml, g->dx, pmlvp, adt, am);
    prof2_init(&M->pz, nz, pml, g->dz, pmlvp, adt, am);
}

/* The window's medium on the refined mesh: c11, c13, c33, c55 and rho of
 * the padded regular window (rows k = -pml .. nz+pml+1 at (k-1) dz) mapped
 * column by column onto z_i by FLIT's linear ginterp
 * (map_regular_to_irregular); the effective shear modulus and densities
 * from the mapped values as on the regular mesh; pmlvp from the regular
 * window's edges; z damping on dz_i(j). */
static void emed2_init_fs(emed2 *M, const ora_egrid2d *g, const emesh2 *S, const float *vp, const float *vs,
                          const float *rho, float f0max) {
    const int R = g->R, pml = g->pml, nx = g->nx, nz0 = g->nz, nz = S->nz;
    lay2 L = {1 - pml - R, nx + 2 * (pml + R), nz + 2 * (pml + R), nx, nz, pml, R};
    M->L = L;
    M->dx = g->dx;
    M->dz = g->dz;
    M->S = S;
    ORA_SYM(fdcoefs)(R, M->c);
    M->idx = (real)(float)(1.0 / (double)g->dx);
    M->idz = (real)(float)(1.0 / (double)g->dz);
    const size_t n = (size_t)L.NX * L.NZ;
    M->c11 = zal(n); M->c13 = zal(n); M->c33 = zal(n); M->c55 = zal(n);
    M->mu = zal(n); M->rx = zal(n); M->rz = zal(n); M->rho = zal(n);
    float pm = 0.0f;
    for (int i = 1; i <= nx; ++i)
        for (int j = 1; j <= nz0; ++j) {
            if (!(i == 1 || i == nx || j == 1 || j == nz0)) continue;
            const size_t m = (size_t)(j - 1) + (size_t)(i - 1) * nz0;
            const float r = rho[m], c33 = r * (vp[m] * vp[m]);
            if (c33 / r > pm) pm = c33 / r;
        }
    const float pmlvp = sqrtf(pm);
    const int nr = nz0 + 2 * pml + 2, nq = nz + 2 * pml + 2;
    float *zr = (float *)malloc(sizeof(float) * (size_t)nr), *col = (float *)malloc(sizeof(float) * 5 * (size_t)nr);
    float *out = (float *)malloc(sizeof(float) * 5 * (size_t)nq);
    for (int k = -pml; k <= nz0 + pml + 1; ++k) zr[k + pml] = (float)(k - 1) * g->dz;
    real *dst[5] = {M->c11, M->c13, M->c33, M->c55, M->rho};
    for (int i = L.lo; i < L.lo + L.NX; ++i) {
        const int ic = i < 1 ? 1 : (i > nx ? nx : i);
        for (int k = -pml; k <= nz0 + pml + 1; ++k) {
            const int kc = k < 1 ? 1 : (k > nz0 ? nz0 : k);
            const size_t m = (size_t)(kc - 1) + (size_t)(ic - 1) * nz0;
            const float r = rho[m], c33 = r * (vp[m] * vp[m]), c55 = r * (vs[m] * vs[m]);
            const int q = k + pml;
            col[q] = c33;
            col[nr + q] = c33 - 2 * c55;
            col[2 * nr + q] = c33;
            col[3 * nr + q] = c55;
            col[4 * nr + q] = r;
        }
        for (int c = 0; c < 5; ++c)
            ORA_SYM(flit_ginterp)(zr, col + (size_t)c * nr, nr, S->zi, nq, out + (size_t)c * nq, 0);
        for (int j = L.lo; j < L.lo + L.NZ; ++j) {
            const int jc = j < -pml ? -pml : (j > nz + pml + 1 ? nz + pml + 1 : j);
            for (int c = 0; c < 5; ++c) dst[c][I2(L, i, j)] = out[(size_t)c * nq + (size_t)(jc + pml)];
        }
    }
    free(zr);
    free(col);
    free(out);
    for (int j = L.lo + 1; j < L.lo + L.NZ; ++j)
        for (int i = L.lo + 1; i < L.lo + L.NX; ++i) {
            const size_t q = I2(L, i, j), sx = 1, sz = (size_t)L.NX;
            M->mu[q] = (real)4.0f / ((real)1.0f / M->c55[q - sx - sz] + (real)1.0f / M->c55[q - sz] +
                                     (real)1.0f / M->c55[q - sx] + (real)1.0f / M->c55[q]);
            M->rx[q] = (real)0.5f * (M->rho[q - sx] + M->rho[q]);
            M->rz[q] = (real)0.5f * (M->rho[q - sz] + M->rho[q]);
        }
    const float am = ORA_SYM(alphamax)(f0max), adt = fabsf(g->dt);
    prof2_init(&M->px, nx, pml, g->dx, pmlvp, adt, am);
    prof2_init_rows(&M->pz, nz, pml, S->dzi, pmlvp, adt, am);
}

static void emed2_free(emed2 *M) {
    free(M->c11); free(M->c13); free(M->c33); free(M->c55);
    free(M->mu); free(M->rx); free(M->rz); free(M->rho);
    free(M->px.ai); free(M->pz.ai);
}

/* update_wavefield (module_wavefield.f90:42-134): normal stresses, shear
 * stress, vx, vz, each over [-pml+1, n+pml] in storage index, memory
 * variables m = a d + b m, d = d + m inside the masks. */
static void eupdate2(const emed2 *M, efld2 *F, real dt) {
    const lay2 L = M->L;
    const int R = L.R, pml = L.pml, nx = L.nx, nz = L.nz;
    const size_t sx = 1, sz = (size_t)L.NX;
    const float *c = M->c;
#pragma omp parallel for schedule(static)
    for (int j = -pml + 1; j <= nz + pml; ++j)
        for (int i = -pml + 1; i <= nx + pml; ++i) {
            const size_t q = I2(L, i, j);
            real a = (real)c[0] * (F->vx[q + sx] - F->vx[q]);
            real b = (real)c[0] * (F->vz[q + sz] - F->vz[q]);
            for (int m = 2; m <= R; ++m) {
                a = a + (real)c[m - 1] * (F->vx[q + m * sx] - F->vx[q - (m - 1) * sx]);
                b = b + (real)c[m - 1] * (F->vz[q + m * sz] - F->vz[q - (m - 1) * sz]);
            }
            real pdxvx = M->idx * a, pdzvz = M->idz * b;
            if (i <= 0 || i >= nx + 1 || j <= 0 || j >= nz + 1) {
                const int ii = i + pml - 1, jj = j + pml - 1;
                F->m_dxvx[q] = (real)M->px.ai[ii] * pdxvx + (real)M->px.bi[ii] * F->m_dxvx[q];
                pdxvx = (pdxvx + F->m_dxvx[q]) / (real)1.0f;
                F->m_dzvz[q] = (real)M->pz.ai[jj] * pdzvz + (real)M->pz.bi[jj] * F->m_dzvz[q];
                pdzvz = (pdzvz + F->m_dzvz[q]) / (real)1.0f;
            }
            F->sxx[q] = F->sxx[q] + dt * (M->c11[q] * pdxvx + M->c13[q] * pdzvz);
            F->szz[q] = F->szz[q] + dt * (M->c13[q] * pdxvx + M->c33[q] * pdzvz);
        }
#pragma omp parallel for schedule(static)
    for (int n = -pml + 1; n <= nz + pml; ++n)
        for (int m_ = -pml + 1; m_ <= nx + pml; ++m_) {
            const size_t q = I2(L, m_, n);
            real a = (real)c[0] * (F->vx[q] - F->vx[q - sz]);
            real b = (real)c[0] * (F->vz[q] - F->vz[q - sx]);
            for (int k = 2; k <= R; ++k) {
                a = a + (real)c[k - 1] * (F->vx[q + (k - 1) * sz] - F->vx[q - k * sz]);
                b = b + (real)c[k - 1] * (F->vz[q + (k - 1) * sx] - F->vz[q - k * sx]);
            }
            real pdzvx = M->idz * a, pdxvz = M->idx * b;
            if (m_ <= 1 || m_ >= nx + 1 || n <= 1 || n >= nz + 1) {
                const int ii = m_ + pml - 1, jj = n + pml - 1;
                F->m_dxvz[q] = (real)M->px.ah[ii] * pdxvz + (real)M->px.bh[ii] * F->m_dxvz[q];
                pdxvz = (pdxvz + F->m_dxvz[q]) / (real)1.0f;
                F->m_dzvx[q] = (real)M->pz.ah[jj] * pdzvx + (real)M->pz.bh[jj] * F->m_dzvx[q];
                pdzvx = (pdzvx + F->m_dzvx[q]) / (real)1.0f;
            }
            F->sxz[q] = F->sxz[q] + dt * M->mu[q] * (pdzvx + pdxvz);
        }
#pragma omp parallel for schedule(static)
    for (int j = -pml + 1; j <= nz + pml; ++j)
        for (int m_ = -pml + 1; m_ <= nx + pml; ++m_) {
            const size_t q = I2(L, m_, j);
            real a = (real)c[0] * (F->sxx[q] - F->sxx[q - sx]);
            real b = (real)c[0] * (F->sxz[q + sz] - F->sxz[q]);
            for (int k = 2; k <= R; ++k) {
                a = a + (real)c[k - 1] * (F->sxx[q + (k - 1) * sx] - F->sxx[q - k * sx]);
                b = b + (real)c[k - 1] * (F->sxz[q + k * sz] - F->sxz[q - (k - 1) * sz]);
            }
            real pdxxx = M->idx * a, pdzxz = M->idz * b;
            if (m_ <= 1 || m_ >= nx + 1 || j <= 0 || j >= nz + 1) {
                const int ii = m_ + pml - 1, jj = j + pml - 1;
                F->m_dxxx[q] = (real)M->px.ah[ii] * pdxxx + (real)M->px.bh[ii] * F->m_dxxx[q];
                pdxxx = (pdxxx + F->m_dxxx[q]) / (real)1.0f;
                F->m_dzxz[q] = (real)M->pz.ai[jj] * pdzxz + (real)M->pz.bi[jj] * F->m_dzxz[q];
                pdzxz = (pdzxz + F->m_dzxz[q]) / (real)1.0f;
            }
            F->vx[q] = F->vx[q] + dt / M->rx[q] * (pdxxx + pdzxz);
        }
#pragma omp parallel for schedule(static)
    for (int n = -pml + 1; n <= nz + pml; ++n)
        for (int i = -pml + 1; i <= nx + pml; ++i) {
            const size_t q = I2(L, i, n);
            real a = (real)c[0] * (F->sxz[q + sx] - F->sxz[q]);
            real b = (real)c[0] * (F->szz[q] - F->szz[q - sz]);
            for (int k = 2; k <= R; ++k) {
                a = a + (real)c[k - 1] * (F->sxz[q + k * sx] - F->sxz[q - (k - 1) * sx]);
                b = b + (real)c[k - 1] * (F->szz[q + (k - 1) * sz] - F->szz[q - k * sz]);
            }
            real pdxxz = M->idx * a, pdzzz = M->idz * b;
            if (i <= 0 || i >= nx + 1 || n <= 1 || n >= nz + 1) {
                const int ii = i + pml - 1, jj = n + pml - 1;
                F->m_dxxz[q] = (real)M->px.ai[ii] * pdxxz + (real)M->px.bi[ii] * F->m_dxxz[q];
                pdxxz = (pdxxz + F->m_dxxz[q]) / (real)1.0f;
                F->m_dzzz[q] = (real)M->pz.ah[jj] * pdzzz + (real)M->pz.bh[jj] * F->m_dzzz[q];
                pdzzz = (pdzzz + F->m_dzzz[q]) / (real)1.0f;
            }
            F->vz[q] = F->vz[q] + dt / M->rz[q] * (pdxxz + pdzzz);
        }
}

/* update_wavefield_free_surface (module_wavefield.f90:143-289): the
 * velocities above the surface are zeroed; normal stresses on rows 1 ..
 * nz+pml, the surface row taking dz vz = -c13/c33 dx vx (x damping only);
 * z derivatives scaled by dz/dz_i (integer rows) or dz/dz_h (half rows);
 * no damping above the bottom slab in z; shear stress on half rows 2 ..
 * nz+pml; then the odd stress mirrors (szz(1) = 0, szz(1-j) = -szz(1+j),
 * sxx(1-j) = -sxx(1+j), sxz(2-j) = -sxz(1+j)); vx on rows 1 .. nz+pml, vz
 * on half rows 2 .. nz+pml. */
static void eupdate2_fs(const emed2 *M, efld2 *F, real dt) {
    const lay2 L = M->L;
    const emesh2 *S = M->S;
    const int R = L.R, pml = L.pml, nx = L.nx, nz = L.nz;
    const size_t sx = 1, sz = (size_t)L.NX;
    const float *c = M->c;
    for (int j = 1; j <= R; ++j)
        for (int i = L.lo; i < L.lo + L.NX; ++i) {
            F->vx[I2(L, i, 1 - j)] = 0;
            F->vz[I2(L, i, 2 - j)] = 0;
        }
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= nz + pml; ++j)
        for (int i = -pml + 1; i <= nx + pml; ++i) {
            const size_t q = I2(L, i, j);
            const int ii = i + pml - 1, jj = j + pml - 1;
            real a = (real)c[0] * (F->vx[q + sx] - F->vx[q]);
            for (int m = 2; m <= R; ++m) a = a + (real)c[m - 1] * (F->vx[q + m * sx] - F->vx[q - (m - 1) * sx]);
            real pdxvx = M->idx * a, pdzvz;
            if (j == 1) {
                if (i <= 0 || i >= nx + 1) {
                    F->m_dxvx[q] = (real)M->px.ai[ii] * pdxvx + (real)M->px.bi[ii] * F->m_dxvx[q];
                    pdxvx = (pdxvx + F->m_dxvx[q]) / (real)1.0f;
                }
                pdzvz = -(M->c13[q] / M->c33[q] * pdxvx);
            } else {
                real b = (real)c[0] * (F->vz[q + sz] - F->vz[q]);
                for (int m = 2; m <= R; ++m) b = b + (real)c[m - 1] * (F->vz[q + m * sz] - F->vz[q - (m - 1) * sz]);
                pdzvz = M->idz * b * (real)S->sci[j + pml];
                if (i <= 0 || i >= nx + 1 || j >= nz + 1) {
                    F->m_dxvx[q] = (real)M->px.ai[ii] * pdxvx + (real)M->px.bi[ii] * F->m_dxvx[q];
                    pdxvx = (pdxvx + F->m_dxvx[q]) / (real)1.0f;
                    F->m_dzvz[q] = (real)M->pz.ai[jj] * pdzvz + (real)M->pz.bi[jj] * F->m_dzvz[q];
                    pdzvz = (pdzvz + F->m_dzvz[q]) / (real)1.0f;
                }
            }
            F->sxx[q] = F->sxx[q] + dt * (M->c11[q] * pdxvx + M->c13[q] * pdzvz);
            F->szz[q] = F->szz[q] + dt * (M->c13[q] * pdxvx + M->c33[q] * pdzvz);
        }
#pragma omp parallel for schedule(static)
    for (int n = 2; n <= nz + pml; ++n)
        for (int m_ = -pml + 1; m_ <= nx + pml; ++m_) {
            const size_t q = I2(L, m_, n);
            real a = (real)c[0] * (F->vx[q] - F->vx[q - sz]);
            real b = (real)c[0] * (F->vz[q] - F->vz[q - sx]);
            for (int k = 2; k <= R; ++k) {
                a = a + (real)c[k - 1] * (F->vx[q + (k - 1) * sz] - F->vx[q - k * sz]);
                b = b + (real)c[k - 1] * (F->vz[q + (k - 1) * sx] - F->vz[q - k * sx]);
            }
            real pdzvx = M->idz * a * (real)S->sch[n + pml], pdxvz = M->idx * b;
            if (m_ <= 1 || m_ >= nx + 1 || n >= nz + 1) {
                const int ii = m_ + pml - 1, jj = n + pml - 1;
                F->m_dxvz[q] = (real)M->px.ah[ii] * pdxvz + (real)M->px.bh[ii] * F->m_dxvz[q];
                pdxvz = (pdxvz + F->m_dxvz[q]) / (real)1.0f;
                F->m_dzvx[q] = (real)M->pz.ah[jj] * pdzvx + (real)M->pz.bh[jj] * F->m_dzvx[q];
                pdzvx = (pdzvx + F->m_dzvx[q]) / (real)1.0f;
            }
            F->sxz[q] = F->sxz[q] + dt * M->mu[q] * (pdzvx + pdxvz);
        }
    for (int j = 1; j <= R; ++j)
        for (int i = L.lo; i < L.lo + L.NX; ++i) {
            F->szz[I2(L, i, 1 - j)] = -F->szz[I2(L, i, 1 + j)];
            F->szz[I2(L, i, 1)] = 0;
            F->sxx[I2(L, i, 1 - j)] = -F->sxx[I2(L, i, 1 + j)];
            F->sxz[I2(L, i, 2 - j)] = -F->sxz[I2(L, i, 1 + j)];
        }
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= nz + pml; ++j)
        for (int m_ = -pml + 1; m_ <= nx + pml; ++m_) {
            const size_t q = I2(L, m_, j);
            real a = (real)c[0] * (F->sxx[q] - F->sxx[q - sx]);
            real b = (real)c[0] * (F->sxz[q + sz] - F->sxz[q]);
            for (int k = 2; k <= R; ++k) {
                a = a + (real)c[k - 1] * (F->sxx[q + (k - 1) * sx] - F->sxx[q - k * sx]);
                b = b + (real)c[k - 1] * (F->sxz[q + k * sz] - F->sxz[q - (k - 1) * sz]);
            }
            real pdxxx = M->idx * a, pdzxz = M->idz * b * (real)S->sci[j + pml];
            if (m_ <= 1 || m_ >= nx + 1 || j >= nz + 1) {
                const int ii = m_ + pml - 1, jj = j + pml - 1;
                F->m_dxxx[q] = (real)M->px.ah[ii] * pdxxx + (real)M->px.bh[ii] * F->m_dxxx[q];
                pdxxx = (pdxxx + F->m_dxxx[q]) / (real)1.0f;
                F->m_dzxz[q] = (real)M->pz.ai[jj] * pdzxz + (real)M->pz.bi[jj] * F->m_dzxz[q];
                pdzxz = (pdzxz + F->m_dzxz[q]) / (real)1.0f;
            }
            F->vx[q] = F->vx[q] + dt / M->rx[q] * (pdxxx + pdzxz);
        }
#pragma omp parallel for schedule(static)
    for (int n = 2; n <= nz + pml; ++n)
        for (int i = -pml + 1; i <= nx + pml; ++i) {
            const size_t q = I2(L, i, n);
            real a = (real)c[0] * (F->sxz[q + sx] - F->sxz[q]);
            real b = (real)c[0] * (F->szz[q] - F->szz[q - sz]);
            for (int k = 2; k <= R; ++k) {
                a = a + (real)c[k - 1] * (F->sxz[q + k * sx] - F->sxz[q - (k - 1) * sx]);
                b = b + (real)c[k - 1] * (F->szz[q + (k - 1) * sz] - F->szz[q - k * sz]);
            }
            real pdxxz = M->idx * a, pdzzz = M->idz * b * (real)S->sch[n + pml];
            if (i <= 0 || i >= nx + 1 || n >= nz + 1) {
                const int ii = i + pml - 1, jj = n + pml - 1;
                F->m_dxxz[q] = (real)M->px.ai[ii] * pdxxz + (real)M->px.bi[ii] * F->m_dxxz[q];
                pdxxz = (pdxxz + F->m_dxxz[q]) / (real)1.0f;
                F->m_dzzz[q] = (real)M->pz.ah[jj] * pdzzz + (real)M->pz.bh[jj] * F->m_dzzz[q];
                pdzzz = (pdzzz + F->m_dzzz[q]) / (real)1.0f;
            }
            F->vz[q] = F->vz[q] + dt / M->rz[q] * (pdxxz + pdzzz);
        }
}

static void eupdate2_any(const emed2 *M, efld2 *F, real dt) {
    if (M->S) eupdate2_fs(M, F, dt);
    else eupdate2(M, F, dt);
}

/* before recording (submodule_forward_modeling.f90:67-75): even mirror of
 * the velocities above the surface, so taps above it read the image */
static void emirror_v(const lay2 *L, efld2 *F) {
    for (int j = 1; j <= L->R; ++j)
        for (int i = L->lo; i < L->lo + L->NX; ++i) {
            F->vx[I2(*L, i, 1 - j)] = F->vx[I2(*L, i, 1 + j)];
            F->vz[I2(*L, i, 2 - j)] = F->vz[I2(*L, i, 1 + j)];
        }
}

/* add_source (module_wavefield.f90:294-403): a force adds to vx at (hx, gz)
 * sin(polar) amp / rho(hx, gz) * hx_w * iz_w and to vz at (gx, hz)
 * cos(polar) amp / rho(gx, hz) * ix_w * hz_w; an explosion subtracts
 * amp / (dx dz) * ix_w * iz_w from sxx and szz at (gx, gz) (its stf is the
 * time derivative, time_integration = -1); amp = stf amp dt. */
static void eadd_source2(const emed2 *M, efld2 *F, const ora_esource2d *src, const ora_point *pts, int ns, int t,
                         real dt) {
    const lay2 L = M->L;
    for (int s = 0; s < ns; ++s) {
        const ora_point *P = &pts[s];
        const int nb = (int)roundf(src[s].t0 / (float)dt) + 1;  /* Fortran nint: half away from zero */
        const int ne = nb + src[s].nt_stf - 1;
        if (t < nb || t > ne) continue;
        const real amp = src[s].stf[t - nb] * (real)src[s].amp * dt;
        /* with a free surface only rows >= 2 take a source; an explosion
         * divides by the local spacing dz_i(gz) */
        const int fs = M->S != NULL;
        if (src[s].mechanism == 0) {
            const float dzs = fs ? M->S->dzi[P->gz + M->S->pml] : M->dz;
            const real ae = amp / ((real)M->dx * (real)dzs);
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int irx = -NKW; irx <= NKW; ++irx) {
                    if (fs && P->gz + irz < 2) continue;
                    const size_t q = I2(L, P->gx + irx, P->gz + irz);
                    F->sxx[q] = F->sxx[q] - ae * (real)P->ix[irx + NKW] * (real)P->iz[irz + NKW];
                    F->szz[q] = F->szz[q] - ae * (real)P->ix[irx + NKW] * (real)P->iz[irz + NKW];
                }
            continue;
        }
#if defined(ORA_F64)
        const real sp = sin((double)src[s].polar), cp = cos((double)src[s].polar);
#else
        const real sp = sinf(src[s].polar), cp = cosf(src[s].polar);
#endif
        const real rx = M->rho[I2(L, P->hx, P->gz)], rz = M->rho[I2(L, P->gx, P->hz)];
        for (int irz = -NKW; irz <= NKW; ++irz)
            for (int irx = -NKW; irx <= NKW; ++irx) {
                if (!fs || P->gz + irz >= 2) {
                    const size_t qx = I2(L, P->hx + irx, P->gz + irz);
                    F->vx[qx] = F->vx[qx] + sp * amp / rx * (real)P->hxw[irx + NKW] * (real)P->iz[irz + NKW];
                }
                if (!fs || P->hz + irz >= 2) {
                    const size_t qz = I2(L, P->gx + irx, P->hz + irz);
                    F->vz[qz] = F->vz[qz] + cp * amp / rz * (real)P->ix[irx + NKW] * (real)P->hzw[irz + NKW];
                }
            }
    }
}

static void erecord2(const lay2 *L, const efld2 *F, const ora_ereceiver2d *rec, const ora_point *rp, int nr,
                     int nt, int t, real *tr) {
    for (int r = 0; r < nr; ++r) {
        if (rec[r].weight == 0.0f) continue;
        const ora_point *Q = &rp[r];
        const real w = (real)rec[r].weight;
        real ax = 0, az = 0;
        for (int irz = -NKW; irz <= NKW; ++irz)
            for (int irx = -NKW; irx <= NKW; ++irx)
                ax = ax + F->vx[I2(*L, Q->hx + irx, Q->gz + irz)] * (real)Q->hxw[irx + NKW] * (real)Q->iz[irz + NKW] * w;
        for (int irz = -NKW; irz <= NKW; ++irz)
            for (int irx = -NKW; irx <= NKW; ++irx)
                az = az + F->vz[I2(*L, Q->gx + irx, Q->hz + irz)] * (real)Q->ix[irx + NKW] * (real)Q->hzw[irz + NKW] * w;
        tr[(size_t)r * nt + (t - 1)] = ax;
        tr[((size_t)nr + r) * nt + (t - 1)] = az;
    }
}

/* Boundary records (module_boundary_saving.f90:89-160): for vx, then vz:
 * left [2-R, 1] x [1-R, nz+R], right [nx+1, nx+R] x [1-R, nz+R], top
 * [1-R, nx+R] x [2-R, 1], bottom [1-R, nx+R] x [nz+1, nz+R], first index
 * fastest. */
static size_t estrip2_len(const lay2 *L) {
    const size_t R = (size_t)L->R, ex = L->nx + 2 * R, ez = L->nz + 2 * R;
    return 2 * (2 * R * ez + 2 * R * ex);
}

static void estrips2(const lay2 *L, efld2 *F, real *rec, int save) {
    const int R = L->R, nx = L->nx, nz = L->nz;
    size_t k = 0;
    real *v2[2] = {F->vx, F->vz};
    for (int c = 0; c < 2; ++c) {
        real *f = v2[c];
#define XF(i, j)                       \
    do {                               \
        const size_t q_ = I2(*L, i, j); \
        if (save) rec[k] = f[q_];      \
        else f[q_] = rec[k];           \
        ++k;                           \
    } while (0)
        for (int j = 1 - R; j <= nz + R; ++j)
            for (int i = 2 - R; i <= 1; ++i) XF(i, j);
        for (int j = 1 - R; j <= nz + R; ++j)
            for (int i = nx + 1; i <= nx + R; ++i) XF(i, j);
        for (int j = 2 - R; j <= 1; ++j)
            for (int i = 1 - R; i <= nx + R; ++i) XF(i, j);
        for (int j = nz + 1; j <= nz + R; ++j)
            for (int i = 1 - R; i <= nx + R; ++i) XF(i, j);
#undef XF
    }
}

/* prepare_geometry's depth-varying branch (module_source_receiver.f90:
 * 596-680): x as on the regular mesh; gz, hz the first rows nearest to
 * z - oz on z_i and (z + float_small at z = 0) - oz on z_h, over j = -pml ..
 * nz+pml+1; weights Kaiser-sinc of the offset over dz_i(gz), dz_h(hz). */
static void locate_fs(const emesh2 *S, const ora_egrid2d *g, float x, float z, ora_point *pt) {
    ORA_SYM(locate_point)(2, x, 0.0f, z, g->ox, 0.0f, g->oz, g->dx, 1.0f, g->dz, pt);
    const int pml = S->pml;
    const float zs = z + (z == 0.0f ? FLT_EPSILON : 0.0f);
    float bi = INFINITY, bh = INFINITY;
    for (int j = -pml; j <= S->nz + pml + 1; ++j) {
        const float di = fabsf(z - g->oz - S->zi[j + pml]);
        if (di < bi) {
            bi = di;
            pt->gz = j;
        }
        const float dh = fabsf(zs - g->oz - S->zh[j + pml]);
        if (dh < bh) {
            bh = dh;
            pt->hz = j;
        }
    }
    ORA_SYM(kaiser_sinc)((z - g->oz - S->zi[pt->gz + pml]) / S->dzi[pt->gz + pml], pt->iz);
    ORA_SYM(kaiser_sinc)((z - g->oz - S->zh[pt->hz + pml]) / S->dzh[pt->hz + pml], pt->hzw);
}

/* With a free surface a source above dz_i(1) moves to dz_i(1) with its
 * amplitude x 1.2 (module_vars.f90:459-466); src2 holds the moved copies. */
static int prepare(const ora_egrid2d *g, const emesh2 *S, const ora_esource2d *src, int ns,
                   const ora_ereceiver2d *rec, int nr, ora_point **sp, ora_point **rp, ora_esource2d **src2) {
    if (g->R < 2 || g->R > 8 || g->pml < g->R || ns < 1) return -1;
    *sp = (ora_point *)calloc((size_t)ns, sizeof(ora_point));
    *rp = (ora_point *)calloc(nr > 0 ? (size_t)nr : 1, sizeof(ora_point));
    *src2 = (ora_esource2d *)malloc(sizeof(ora_esource2d) * (size_t)ns);
    memcpy(*src2, src, sizeof(ora_esource2d) * (size_t)ns);
    for (int s = 0; s < ns; ++s) {
        ora_esource2d *c = &(*src2)[s];
        if (S) {
            const float d1 = S->dzi[1 + S->pml];
            if (c->z < d1) {
                c->z = d1;
                c->amp = c->amp * 1.2f;
            }
            locate_fs(S, g, c->x, c->z, &(*sp)[s]);
        } else {
            ORA_SYM(locate_point)(2, c->x, 0.0f, c->z, g->ox, 0.0f, g->oz, g->dx, 1.0f, g->dz, &(*sp)[s]);
        }
    }
    for (int r = 0; r < nr; ++r) {
        if (S) locate_fs(S, g, rec[r].x, rec[r].z, &(*rp)[r]);
        else ORA_SYM(locate_point)(2, rec[r].x, 0.0f, rec[r].z, g->ox, 0.0f, g->oz, g->dx, 1.0f, g->dz, &(*rp)[r]);
    }
    return 0;
}

/* compute_forward (submodule_forward_modeling.f90:31-180): step t =
 * [boundary record] -> record(t) -> update -> source(t); traces two
 * [nr][nt] blocks (vx, vz).  strips/final (nullable) receive the boundary
 * records of every step and the five wavefields at the start of step nt. */
static int forward2(const ora_egrid2d *g, const emed2 *M, const ora_esource2d *src0, int ns, int nr,
                    const ora_ereceiver2d *rec, real *tr, real *strips, real *fin) {
    ora_point *sp, *rp;
    ora_esource2d *src;
    if (prepare(g, M->S, src0, ns, rec, nr, &sp, &rp, &src)) return -1;
    const lay2 L = M->L;
    const size_t n = (size_t)L.NX * L.NZ, sl = estrip2_len(&L);
    efld2 F;
    efld2_alloc(&F, &L);
    memset(tr, 0, sizeof(real) * 2 * (size_t)nr * g->nt);
    for (int t = 1; t <= g->nt; ++t) {
        if (strips) estrips2(&L, &F, strips + (size_t)(t - 1) * sl, 1);
        if (t == g->nt && fin) {
            real **fw = &F.sxx;
            for (int c = 0; c < 5; ++c) memcpy(fin + (size_t)c * n, fw[c], n * sizeof(real));
        }
        if (M->S) emirror_v(&L, &F);
        erecord2(&L, &F, rec, rp, nr, g->nt, t, tr);
        if (t == g->nt) break;
        eupdate2_any(M, &F, (real)g->dt);
        eadd_source2(M, &F, src, sp, ns, t, (real)g->dt);
    }
    efld2_free(&F);
    free(sp);
    free(rp);
    free(src);
    return 0;
}

int ORA_SYM(elastic2d_forward)(const ora_egrid2d *g, const float *vp, const float *vs, const float *rho, int ns,
                               const ora_esource2d *src, float f0max, int nr, const ora_ereceiver2d *rec,
                               real *traces) {
    emed2 M;
    emed2_init(&M, g, vp, vs, rho, f0max);
    const int rc = forward2(g, &M, src, ns, nr, rec, traces, NULL, NULL);
    emed2_free(&M);
    return rc;
}

/* OWL's elastic gradient (submodule_fwi_gradient.f90:138-270, 440-520,
 * chain rule :330-360): reverse t = nt .. hnt: prev = state; state =
 * update(state, -dt); at t = nt the final state; boundary record t; adjoint
 * = update(adjoint, -dt); adjoint sources (vx at (hx, gz), vz at (gx, hz));
 * imaging over the interior when t % cc == 0.  The adjoint's own steps
 * below hnt do not touch the medium gradients and are left out.  adj: two
 * [nr][nt] blocks.  Outputs (nz, nx) z fastest: grad_vp, grad_vs (vp, vs
 * in km/s, rho in g/cm^3), grad_rho. */
/* one column's natural cubic spline (FLIT's ginterp 'cubic', evaluated in
 * double) from the n knots x, y to the m points xi */
static void cubic_column(const float *x, const real *y, int n, const float *xi, int m, real *out) {
    double *M2 = (double *)calloc((size_t)n, sizeof(double)), *cp = (double *)calloc((size_t)n, sizeof(double));
    double *dp = (double *)calloc((size_t)n, sizeof(double));
    for (int i = 1; i < n - 1; ++i) {
        const double h0 = (double)x[i] - x[i - 1], h1 = (double)x[i + 1] - x[i];
        const double r = 6.0 * (((double)y[i + 1] - y[i]) / h1 - ((double)y[i] - y[i - 1]) / h0);
        const double den = 2.0 * (h0 + h1) - h0 * cp[i - 1];
        cp[i] = h1 / den;
        dp[i] = (r - h0 * dp[i - 1]) / den;
    }
    for (int i = n - 2; i >= 1; --i) M2[i] = dp[i] - cp[i] * M2[i + 1];
    for (int k = 0; k < m; ++k) {
        int s0;
        if (xi[k] < x[1]) s0 = 0;
        else if (xi[k] >= x[n - 2]) s0 = n - 2;
        else {
            int lo = 1, hi = n - 2;
            while (hi - lo > 1) {
                const int mid = (lo + hi) / 2;
                if (x[mid] <= xi[k]) lo = mid;
                else hi = mid;
            }
            s0 = lo;
        }
        const double h = (double)x[s0 + 1] - x[s0];
        const double A = ((double)x[s0 + 1] - xi[k]) / h, B = ((double)xi[k] - x[s0]) / h;
        out[k] = (real)(A * y[s0] + B * y[s0 + 1] +
                        ((A * A * A - A) * M2[s0] + (B * B * B - B) * M2[s0 + 1]) * h * h / 6.0);
    }
    free(M2);
    free(cp);
    free(dp);
}

/* OWL's elastic gradient core on medium M (regular or refined mesh):
 * reverse t = nt .. hnt: prev = state; state = update(state, -dt); at t =
 * nt the final state; boundary record t; adjoint = update(adjoint, -dt);
 * adjoint sources (vx at (hx, gz), vz at (gx, hz); with a free surface rows
 * >= 2 only); imaging over the interior rows of M when t % cc == 0.  The
 * adjoint's own steps below hnt do not touch the medium gradients and are
 * left out.  gc[5]: c11, c13, c33, c55, rho gradients (nx, nz of M) x
 * fastest. */
static int gradient2(const ora_egrid2d *g, const emed2 *M, int ns, const ora_esource2d *src, int hnt, int nr,
                     const ora_ereceiver2d *rec, const real *adj, real **gc, real *tr) {
    const lay2 L = M->L;
    const size_t n = (size_t)L.NX * L.NZ, sl = estrip2_len(&L);
    const int nt = g->nt;
    real *strips = zal(sl * (size_t)nt), *fin = zal(5 * n);
    int rc = forward2(g, M, src, ns, nr, rec, tr, strips, fin);
    if (rc) {
        free(strips);
        free(fin);
        return rc;
    }
    ora_point *sp, *rp;
    ora_esource2d *src2;
    prepare(g, M->S, src, ns, rec, nr, &sp, &rp, &src2);
    efld2 F, A, P;
    efld2_alloc(&F, &L);
    efld2_alloc(&A, &L);
    efld2_alloc(&P, &L);
    const int nx = L.nx, nz = L.nz, fs = M->S != NULL;
    const real mdt = -(real)g->dt;
    const int cc = g->cc_step_interval > 0 ? g->cc_step_interval : 1;
    const size_t sx = 1, sz = (size_t)L.NX;
    for (int t = nt; t >= hnt; --t) {
        real **fw = &F.sxx, **pw = &P.sxx;
        for (int c = 0; c < 5; ++c) memcpy(pw[c], fw[c], n * sizeof(real));
        eupdate2_any(M, &F, mdt);
        if (t == nt)
            for (int c = 0; c < 5; ++c) memcpy(fw[c], fin + (size_t)c * n, n * sizeof(real));
        estrips2(&L, &F, strips + (size_t)(t - 1) * sl, 0);
        eupdate2_any(M, &A, mdt);
        for (int r = 0; r < nr; ++r) {
            if (rec[r].weight == 0.0f) continue;
            const ora_point *Q = &rp[r];
            const real w = (real)rec[r].weight;
            const real ax = adj[(size_t)r * nt + (t - 1)], az = adj[((size_t)nr + r) * nt + (t - 1)];
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int irx = -NKW; irx <= NKW; ++irx) {
                    if (fs && Q->gz + irz < 2) continue;
                    const size_t q = I2(L, Q->hx + irx, Q->gz + irz);
                    A.vx[q] = A.vx[q] + ax * (real)Q->hxw[irx + NKW] * (real)Q->iz[irz + NKW] * w;
                }
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int irx = -NKW; irx <= NKW; ++irx) {
                    if (fs && Q->hz + irz < 2) continue;
                    const size_t q = I2(L, Q->gx + irx, Q->hz + irz);
                    A.vz[q] = A.vz[q] + az * (real)Q->ix[irx + NKW] * (real)Q->hzw[irz + NKW] * w;
                }
        }
        if (t % cc != 0) continue;
#pragma omp parallel for schedule(static)
        for (int j = 1; j <= nz; ++j)
            for (int i = 1; i <= nx; ++i) {
                const size_t q = I2(L, i, j), o = (size_t)(i - 1) + (size_t)(j - 1) * nx;
                const real c11 = M->c11[q], c13 = M->c13[q], c33 = M->c33[q], c55 = M->c55[q];
                real s11 = 0, s13 = 0, s33 = 0, s55 = 0;
                if (c13 * c13 - c11 * c33 != 0) {
                    s11 = c33 / (-(c13 * c13) + c11 * c33);
                    s13 = c13 / (c13 * c13 - c11 * c33);
                    s33 = c11 / (-(c13 * c13) + c11 * c33);
                }
                if (c55 != 0) s55 = (real)1.0f / c55;
                else s11 = s13 = s33 = s55 = 0;
                const real txx = F.sxx[q] - P.sxx[q], tzz = F.szz[q] - P.szz[q];
                const real txz = (real)0.25f * (F.sxz[q] + F.sxz[q + sx] + F.sxz[q + sz] + F.sxz[q + sx + sz]) -
                                 (real)0.25f * (P.sxz[q] + P.sxz[q + sx] + P.sxz[q + sz] + P.sxz[q + sx + sz]);
                const real rxx = A.sxx[q], rzz = A.szz[q];
                const real rxz = (real)0.25f * (A.sxz[q] + A.sxz[q + sx] + A.sxz[q + sz] + A.sxz[q + sx + sz]);
                const real exx = s11 * txx + s13 * tzz, ezz = s13 * txx + s33 * tzz, exz = s55 * txz;
                const real fxx = s11 * rxx + s13 * rzz, fzz = s13 * rxx + s33 * rzz, fxz = s55 * rxz;
                gc[0][o] = gc[0][o] - (fxx * exx);
                gc[1][o] = gc[1][o] - (fxx * ezz + fzz * exx);
                gc[2][o] = gc[2][o] - (fzz * ezz);
                gc[3][o] = gc[3][o] - (fxz * exz);
                const real svx = (F.vx[q] + F.vx[q + sx]) - (P.vx[q] + P.vx[q + sx]);
                const real rvx = A.vx[q] + A.vx[q + sx];
                const real svz = (F.vz[q] + F.vz[q + sz]) - (P.vz[q] + P.vz[q + sz]);
                const real rvz = A.vz[q] + A.vz[q + sz];
                gc[4][o] = gc[4][o] + svx * rvx + svz * rvz;
            }
    }
    free(strips); free(fin); free(sp); free(rp); free(src2);
    efld2_free(&F); efld2_free(&A); efld2_free(&P);
    return 0;
}

/* chain rule (submodule_fwi_gradient.f90:330-360): vp, vs in km/s and rho
 * in g/cm^3; rho_k: the density per output point (regular or mapped back),
 * already x 1e-3; NaN/Inf to zero.  gc and outputs (nz, nx) z fastest. */
static void chain_rule2(int nx, int nz, const float *vp, const float *vs, const real *rho_k, real *const *gc,
                        real *grad_vp, real *grad_vs, real *grad_rho) {
    for (int j = 1; j <= nz; ++j)
        for (int i = 1; i <= nx; ++i) {
            const size_t m = (size_t)(j - 1) + (size_t)(i - 1) * nz;
            const real r = rho_k[m], p = (real)(vp[m] * 1.0e-3f), v = (real)(vs[m] * 1.0e-3f);
            real a = gc[0][m] + gc[1][m] + gc[2][m];
            a = 2 * r * p * a;
            real b = gc[3][m] - 2 * gc[1][m];
            b = 2 * r * v * b;
            const real c = gc[4][m];
            grad_vp[m] = isfinite(a) ? a : 0;
            grad_vs[m] = isfinite(b) ? b : 0;
            grad_rho[m] = isfinite(c) ? c : 0;
        }
}

/* OWL's elastic gradient (submodule_fwi_gradient.f90:138-270, 440-520,
 * chain rule :330-360).  adj: two [nr][nt] blocks.  Outputs (nz, nx) z
 * fastest: grad_vp, grad_vs (vp, vs in km/s, rho in g/cm^3), grad_rho. */
int ORA_SYM(elastic2d_gradient_owl)(const ora_egrid2d *g, const float *vp, const float *vs, const float *rho,
                                    int ns, const ora_esource2d *src, float f0max, int hnt, int nr,
                                    const ora_ereceiver2d *rec, const real *adj, real *grad_vp, real *grad_vs,
                                    real *grad_rho, real *traces_out) {
    emed2 M;
    emed2_init(&M, g, vp, vs, rho, f0max);
    const int nx = g->nx, nz = g->nz, nt = g->nt;
    const size_t ni = (size_t)nx * nz;
    real *gx[5], *gz[5], *rk = zal(ni);
    for (int c = 0; c < 5; ++c) {
        gx[c] = zal(ni);
        gz[c] = zal(ni);
    }
    real *tr = traces_out ? traces_out : zal(2 * (size_t)nr * nt);
    const int rc = gradient2(g, &M, ns, src, hnt, nr, rec, adj, gx, tr);
    if (!rc) {
        for (int c = 0; c < 5; ++c)
            for (int j = 1; j <= nz; ++j)
                for (int i = 1; i <= nx; ++i)
                    gz[c][(size_t)(j - 1) + (size_t)(i - 1) * nz] = gx[c][(size_t)(i - 1) + (size_t)(j - 1) * nx];
        for (size_t m = 0; m < ni; ++m) rk[m] = (real)(rho[m] * 1.0e-3f);
        chain_rule2(nx, nz, vp, vs, rk, gz, grad_vp, grad_vs, grad_rho);
    }
    for (int c = 0; c < 5; ++c) {
        free(gx[c]);
        free(gz[c]);
    }
    free(rk);
    if (!traces_out) free(tr);
    emed2_free(&M);
    return rc;
}

/* The free-surface forms: OWL's elastic_vhtiort_2d with free_surface on
 * its refined mesh (see the header).  The forward records as OWL does; the
 * gradients c11, c13, c33, c55, rho on the refined rows 1 .. nz are mapped
 * back to the window's rows by the column spline through z_i(1 .. nz), and
 * rho x 1e-3 on the refined mesh by the spline through all its rows -pml ..
 * nz+pml+1 (map_irregular_to_regular, submodule_fwi_gradient.f90:324-338),
 * before the chain rule. */
static int fs_setup(const ora_egrid2d *g, const ora_esurf2d *fs, const float *vp, const float *vs, const float *rho,
                    int ns, const ora_esource2d *src, float f0max, emesh2 *S, emed2 *M) {
    (void)ns;
    (void)src;
    const int rc = emesh2_build(S, g, fs, vs, rho, fs->f0min > 0 ? fs->f0min : f0max);
    if (rc) return rc;
    emed2_init_fs(M, g, S, vp, vs, rho, f0max);
    return 0;
}

int ORA_SYM(elastic2d_forward_fs)(const ora_egrid2d *g, const ora_esurf2d *fs, const float *vp, const float *vs,
                                  const float *rho, int ns, const ora_esource2d *src, float f0max, int nr,
                                  const ora_ereceiver2d *rec, real *traces) {
    emesh2 S;
    emed2 M;
    int rc = fs_setup(g, fs, vp, vs, rho, ns, src, f0max, &S, &M);
    if (rc) return rc;
    rc = forward2(g, &M, src, ns, nr, rec, traces, NULL, NULL);
    emed2_free(&M);
    emesh2_free(&S);
    return rc;
}

int ORA_SYM(elastic2d_gradient_owl_fs)(const ora_egrid2d *g, const ora_esurf2d *fs, const float *vp, const float *vs,
                                       const float *rho, int ns, const ora_esource2d *src, float f0max, int hnt,
                                       int nr, const ora_ereceiver2d *rec, const real *adj, real *grad_vp,
                                       real *grad_vs, real *grad_rho, real *traces_out) {
    emesh2 S;
    emed2 M;
    int rc = fs_setup(g, fs, vp, vs, rho, ns, src, f0max, &S, &M);
    if (rc) return rc;
    const int nx = g->nx, nz0 = g->nz, nzr = S.nz, nt = g->nt, pml = g->pml;
    const size_t nr_ = (size_t)nx * nzr, n0 = (size_t)nx * nz0;
    real *gx[5], *gz[5], *rk = zal(n0);
    for (int c = 0; c < 5; ++c) {
        gx[c] = zal(nr_);
        gz[c] = zal(n0);
    }
    real *tr = traces_out ? traces_out : zal(2 * (size_t)nr * nt);
    rc = gradient2(g, &M, ns, src, hnt, nr, rec, adj, gx, tr);
    if (!rc) {
        float *zq = (float *)malloc(sizeof(float) * (size_t)nz0);
        for (int j = 1; j <= nz0; ++j) zq[j - 1] = (float)(j - 1) * g->dz;
        real *col = zal((size_t)(nzr + 2 * pml + 2)), *o = zal((size_t)nz0);
        for (int i = 1; i <= nx; ++i) {
            for (int c = 0; c < 5; ++c) {
                for (int j = 1; j <= nzr; ++j) col[j - 1] = gx[c][(size_t)(i - 1) + (size_t)(j - 1) * nx];
                cubic_column(S.zi + pml + 1, col, nzr, zq, nz0, o);
                for (int j = 1; j <= nz0; ++j) gz[c][(size_t)(j - 1) + (size_t)(i - 1) * nz0] = o[j - 1];
            }
            for (int j = -pml; j <= nzr + pml + 1; ++j)
                col[j + pml] = M.rho[I2(M.L, i, j)] * (real)1.0e-3f;
            cubic_column(S.zi, col, nzr + 2 * pml + 2, zq, nz0, o);
            for (int j = 1; j <= nz0; ++j) rk[(size_t)(j - 1) + (size_t)(i - 1) * nz0] = o[j - 1];
        }
        free(zq);
        free(col);
        free(o);
        chain_rule2(nx, nz0, vp, vs, rk, gz, grad_vp, grad_vs, grad_rho);
    }
    for (int c = 0; c < 5; ++c) {
        free(gx[c]);
        free(gz[c]);
    }
    free(rk);
    if (!traces_out) free(tr);
    emed2_free(&M);
    emesh2_free(&S);
    return rc;
}

/* the refined mesh of a window, for the tests: nz and the rows z_i, z_h,
 * dz_i, dz_h over j = -pml .. nz+pml+1 (cap entries each) */
int ORA_SYM(elastic2d_mesh)(const ora_egrid2d *g, const ora_esurf2d *fs, const float *vs, const float *rho,
                            float f0min, int cap, float *zi, float *zh, float *dzi, float *dzh) {
    emesh2 S;
    const int rc = emesh2_build(&S, g, fs, vs, rho, f0min);
    if (rc) return rc;
    const int n = S.nz + 2 * g->pml + 2;
    if (n <= cap) {
        memcpy(zi, S.zi, sizeof(float) * (size_t)n);
        memcpy(zh, S.zh, sizeof(float) * (size_t)n);
        memcpy(dzi, S.dzi, sizeof(float) * (size_t)n);
        memcpy(dzh, S.dzh, sizeof(float) * (size_t)n);
    }
    emesh2_free(&S);
    return S.nz;
}
