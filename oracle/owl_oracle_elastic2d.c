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
 *   - no free surface (the 2D elastic free surface always runs on OWL's
 *     depth-refined mesh, built with FLIT's deriv, which is unpinned);
 *   - force sources (explosion/moment-tensor sources need FLIT's deriv of
 *     the STF);
 *   - Fortran sum() over 2x2 sections in array element order.
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

typedef struct {
    lay2 L;
    float c[8];
    real idx, idz;
    float dx, dz;
    real *c11, *c13, *c33, *c55, *mu, *rx, *rz, *rho;
    prof2 px, pz;
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
        if (src[s].mechanism == 0) {
            const real ae = amp / ((real)M->dx * (real)M->dz);
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int irx = -NKW; irx <= NKW; ++irx) {
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
                const size_t qx = I2(L, P->hx + irx, P->gz + irz);
                F->vx[qx] = F->vx[qx] + sp * amp / rx * (real)P->hxw[irx + NKW] * (real)P->iz[irz + NKW];
                const size_t qz = I2(L, P->gx + irx, P->hz + irz);
                F->vz[qz] = F->vz[qz] + cp * amp / rz * (real)P->ix[irx + NKW] * (real)P->hzw[irz + NKW];
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

static int prepare(const ora_egrid2d *g, const ora_esource2d *src, int ns, const ora_ereceiver2d *rec, int nr,
                   ora_point **sp, ora_point **rp) {
    if (g->R < 2 || g->R > 8 || g->pml < g->R || ns < 1) return -1;
    *sp = (ora_point *)calloc((size_t)ns, sizeof(ora_point));
    *rp = (ora_point *)calloc(nr > 0 ? (size_t)nr : 1, sizeof(ora_point));
    for (int s = 0; s < ns; ++s)
        ORA_SYM(locate_point)(2, src[s].x, 0.0f, src[s].z, g->ox, 0.0f, g->oz, g->dx, 1.0f, g->dz, &(*sp)[s]);
    for (int r = 0; r < nr; ++r)
        ORA_SYM(locate_point)(2, rec[r].x, 0.0f, rec[r].z, g->ox, 0.0f, g->oz, g->dx, 1.0f, g->dz, &(*rp)[r]);
    return 0;
}

/* compute_forward (submodule_forward_modeling.f90:31-180): step t =
 * [boundary record] -> record(t) -> update -> source(t); traces two
 * [nr][nt] blocks (vx, vz).  strips/final (nullable) receive the boundary
 * records of every step and the five wavefields at the start of step nt. */
static int forward2(const ora_egrid2d *g, const emed2 *M, const ora_esource2d *src, int ns, int nr,
                    const ora_ereceiver2d *rec, real *tr, real *strips, real *fin) {
    ora_point *sp, *rp;
    if (prepare(g, src, ns, rec, nr, &sp, &rp)) return -1;
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
        erecord2(&L, &F, rec, rp, nr, g->nt, t, tr);
        if (t == g->nt) break;
        eupdate2(M, &F, (real)g->dt);
        eadd_source2(M, &F, src, sp, ns, t, (real)g->dt);
    }
    efld2_free(&F);
    free(sp);
    free(rp);
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
int ORA_SYM(elastic2d_gradient_owl)(const ora_egrid2d *g, const float *vp, const float *vs, const float *rho,
                                    int ns, const ora_esource2d *src, float f0max, int hnt, int nr,
                                    const ora_ereceiver2d *rec, const real *adj, real *grad_vp, real *grad_vs,
                                    real *grad_rho, real *traces_out) {
    emed2 M;
    emed2_init(&M, g, vp, vs, rho, f0max);
    const lay2 L = M.L;
    const size_t n = (size_t)L.NX * L.NZ, sl = estrip2_len(&L);
    const int nt = g->nt;
    real *strips = zal(sl * (size_t)nt), *fin = zal(5 * n);
    real *tr = traces_out ? traces_out : zal(2 * (size_t)nr * nt);
    int rc = forward2(g, &M, src, ns, nr, rec, tr, strips, fin);
    if (rc) {
        free(strips); free(fin);
        if (!traces_out) free(tr);
        emed2_free(&M);
        return rc;
    }
    ora_point *sp, *rp;
    prepare(g, src, ns, rec, nr, &sp, &rp);
    efld2 F, A, P;
    efld2_alloc(&F, &L);
    efld2_alloc(&A, &L);
    efld2_alloc(&P, &L);
    const int nx = g->nx, nz = g->nz;
    const size_t ni = (size_t)nx * nz;
    real *gc[5];
    for (int c = 0; c < 5; ++c) gc[c] = zal(ni); /* c11 c13 c33 c55 rho */
    const real mdt = -(real)g->dt;
    const int cc = g->cc_step_interval > 0 ? g->cc_step_interval : 1;
    const size_t sx = 1, sz = (size_t)L.NX;
    for (int t = nt; t >= hnt; --t) {
        real **fw = &F.sxx, **pw = &P.sxx;
        for (int c = 0; c < 5; ++c) memcpy(pw[c], fw[c], n * sizeof(real));
        eupdate2(&M, &F, mdt);
        if (t == nt)
            for (int c = 0; c < 5; ++c) memcpy(fw[c], fin + (size_t)c * n, n * sizeof(real));
        estrips2(&L, &F, strips + (size_t)(t - 1) * sl, 0);
        eupdate2(&M, &A, mdt);
        for (int r = 0; r < nr; ++r) {
            if (rec[r].weight == 0.0f) continue;
            const ora_point *Q = &rp[r];
            const real w = (real)rec[r].weight;
            const real ax = adj[(size_t)r * nt + (t - 1)], az = adj[((size_t)nr + r) * nt + (t - 1)];
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int irx = -NKW; irx <= NKW; ++irx) {
                    const size_t q = I2(L, Q->hx + irx, Q->gz + irz);
                    A.vx[q] = A.vx[q] + ax * (real)Q->hxw[irx + NKW] * (real)Q->iz[irz + NKW] * w;
                }
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int irx = -NKW; irx <= NKW; ++irx) {
                    const size_t q = I2(L, Q->gx + irx, Q->hz + irz);
                    A.vz[q] = A.vz[q] + az * (real)Q->ix[irx + NKW] * (real)Q->hzw[irz + NKW] * w;
                }
        }
        if (t % cc != 0) continue;
#pragma omp parallel for schedule(static)
        for (int j = 1; j <= nz; ++j)
            for (int i = 1; i <= nx; ++i) {
                const size_t q = I2(L, i, j), o = (size_t)(i - 1) + (size_t)(j - 1) * nx;
                const real c11 = M.c11[q], c13 = M.c13[q], c33 = M.c33[q], c55 = M.c55[q];
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
    for (int j = 1; j <= nz; ++j)
        for (int i = 1; i <= nx; ++i) {
            const size_t o = (size_t)(i - 1) + (size_t)(j - 1) * nx, m = (size_t)(j - 1) + (size_t)(i - 1) * nz;
            const real r = (real)(rho[m] * 1.0e-3f), p = (real)(vp[m] * 1.0e-3f), v = (real)(vs[m] * 1.0e-3f);
            real a = gc[0][o] + gc[1][o] + gc[2][o];
            a = 2 * r * p * a;
            real b = gc[3][o] - 2 * gc[1][o];
            b = 2 * r * v * b;
            const real c = gc[4][o];
            grad_vp[m] = isfinite(a) ? a : 0;
            grad_vs[m] = isfinite(b) ? b : 0;
            grad_rho[m] = isfinite(c) ? c : 0;
        }
    for (int c = 0; c < 5; ++c) free(gc[c]);
    if (!traces_out) free(tr);
    free(strips); free(fin); free(sp); free(rp);
    efld2_free(&F); efld2_free(&A); efld2_free(&P);
    emed2_free(&M);
    return 0;
}
