/*
 * owl_oracle_fsg2d.c -- CPU restatement of OWL's 2D Lebedev fully staggered
 * grid (FSG) elastic scheme for general (TTI) anisotropy (elastic_tti_2d)
 * and of OWL's elastic gradient on it.
 *
 * TEST INFRASTRUCTURE ONLY (see owl_oracle.c): only tests/, smoke() and the
 * CPU baseline may load it, as the checker.
 *
 * Restated from (paths relative to the OWL source tree):
 *   src/solver/elastic_tti_2d/module_wavefield.f90:35-252     update_wavefield
 *   src/solver/elastic_tti_2d/module_wavefield.f90:593-736    add_source (force)
 *   src/solver/elastic_tti_2d/module_cfspml.f90:80-280        M-PML damping ratio (Christoffel scan)
 *   src/solver/elastic_tti_2d/module_cfspml.f90:282-576       damp_coef, per-cell coefficient arrays
 *   src/solver/elastic_tti_2d/module_vars.f90:345-349, 522-525 pmlvp, loop bounds
 *   src/solver/elastic_tti_2d/submodule_forward_modeling.f90:80-110 receivers (half of each sub-grid)
 *   src/solver/elastic_tti_2d/module_boundary_saving.f90:100-200
 *   src/solver/elastic_tti_2d/submodule_fwi_gradient.f90:60-80, 130-400, 628-727
 *   src/lib/module_anisotropy.f90:58-79                       cij_to_sij_2d
 *
 * Scope and stated assumptions:
 *   - the elastic moduli come as c11, c13, c15, c33, c35, c55 (OWL's 'cij'
 *     anisotropy type); the Thomsen conversion with FLIT's Bond-matrix
 *     rotation is not part of this restatement;
 *   - regular mesh without a free surface (OWL's FSG free surface runs on a
 *     depth-refined, topography-following mesh built with FLIT's deriv,
 *     ginterp and rescale, which are unpinned);
 *   - force sources (explosion and moment-tensor sources use FLIT's deriv
 *     of the STF);
 *   - the M-PML damping ratios are 0 when `aniso` is 0 (OWL's 'iso') and
 *     from the Christoffel scan otherwise.  The scan's eigenvalues are
 *     purely imaginary for physical media, so its complex arithmetic is
 *     evaluated as the equivalent real arithmetic in the same order;
 *   - Fortran sum() over 2x2 sections in array element order; const_pi as
 *     float pi.
 * Compiled in float and double with owl_oracle.c's conventions (the
 * coefficient and ratio set-up is float in both, as in OWL).
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
/* FLIT's const_pi taken as double precision pi, as in owl_oracle.c */
#define FSG_PI 3.14159265358979323846264338327950288

typedef struct {
    float x, y, z;
    int gx, gy, gz, hx, hy, hz;
    float ix[9], iy[9], iz[9], hxw[9], hyw[9], hzw[9];
} ora_point;
int ORA_SYM(fdcoefs)(int R, float *out);
float ORA_SYM(alphamax)(float f0max);
void ORA_SYM(locate_point)(int dims, float x, float y, float z, float ox, float oy, float oz, float dx, float dy,
                           float dz, ora_point *pt);
void ORA_SYM(kaiser_sinc)(float x, float *h);

typedef struct {
    int nx, nz;
    float dx, dz, ox, oz;
    float dt;
    int nt, pml, R;
    int cc_step_interval;
    int aniso; /* 0: damping ratios 0 (OWL 'iso'); 1: Christoffel scan */
} ora_fgrid2d;

typedef struct {
    float x, z, amp, t0, polar;
    int mechanism; /* 0 force (polar from the z axis), 1 explosion */
    int nt_stf;
    const real *stf;
} ora_fsource2d;

typedef struct {
    float x, z, weight;
} ora_freceiver2d;

typedef struct {
    int lo, NX, NZ, nx, nz, pml, R;
} flay;

#define IF(L, i, j) ((size_t)((i) - (L).lo) + (size_t)((j) - (L).lo) * (size_t)(L).NX)

/* ---- M-PML damping ratio (module_cfspml.f90:80-280) --------------------- */

/* Fortran integer powers as gfortran expands them: x**3 = (x*x)*x,
 * x**4 = (x*x)*(x*x). */
static inline float P2(float x) { return x * x; }
static inline float P3(float x) { return (x * x) * x; }
static inline float P4(float x) { const float y = x * x; return y * y; }

/* A = trace of the 2D Christoffel matrix, B = minus its determinant
 * (mpml_eigenvalue's two polynomials) */
static inline float chr_a(float c11, float c15, float c33, float c35, float c55, float k1, float k3) {
    return (c11 + c55) * P2(k1) + 2 * (c15 + c35) * k1 * k3 + (c33 + c55) * P2(k3);
}
static inline float chr_b(float c11, float c13, float c15, float c33, float c35, float c55, float k1, float k3) {
    return (P2(c15) - c11 * c55) * P4(k1) + 2 * (c13 * c15 - c11 * c35) * P3(k1) * k3 +
           (P2(c13) - c11 * c33 - 2 * c15 * c35 + 2 * c13 * c55) * P2(k1) * P2(k3) +
           2 * (-(c15 * c33) + c13 * c35) * k1 * P3(k3) + (P2(c35) - c33 * c55) * P4(k3);
}

/* mpml_eigenvalue: -sqrt(cmplx(x)) / sqrt(2) with x < 0 for a stable
 * medium (B < 0 and A^2 + 4B >= 0), so the eigenvalue is -i b with b =
 * sqrt(-x) / sqrt(2).  Returns its square, -b^2 (the derivative needs only
 * eigval**2 and eigval**4, whose real parts are -b*b and (b*b)*(b*b)). */
static float eig2_of(float c11, float c13, float c15, float c33, float c35, float c55, float rho, int mode, float k1,
                     float k3) {
    const float A = chr_a(c11, c15, c33, c35, c55, k1, k3);
    const float B = chr_b(c11, c13, c15, c33, c35, c55, k1, k3);
    const float s = sqrtf((P2(A) + 4 * B) * P2(rho));
    float negx;
    if (mode == 1)
        negx = (A * rho + s) / P2(rho);
    else
        negx = -((-(A * rho) + s) / P2(rho));
    const float b = sqrtf(negx) / sqrtf(2.0f);
    return -(b * b);
}

/* real part of mpml_eigenvalue_derivative; with purely imaginary
 * eigenvalues every imaginary part is zero and the complex arithmetic
 * reduces to this real arithmetic in the same order */
static float eig_deriv(float c11, float c13, float c15, float c33, float c35, float c55, float rho, int dirx, float e2,
                       float ratio, float k1, float k3) {
    const float e4 = e2 * e2;
    const float den = 2 * chr_b(c11, c13, c15, c33, c35, c55, k1, k3) -
                      3 * e2 * chr_a(c11, c15, c33, c35, c55, k1, k3) * rho - 4 * e4 * P2(rho);
    float num;
    if (dirx) {
        num = 2 * c15 * c35 * P2(k1) * P2(k3) + 3 * c15 * c33 * k1 * P3(k3) - 2 * P2(c35) * P4(k3) +
              2 * c33 * c55 * P4(k3) - 2 * P2(c15) * P4(k1) * ratio + 2 * c15 * c35 * P2(k1) * P2(k3) * ratio +
              c15 * c33 * k1 * P3(k3) * ratio - P2(c13) * P2(k1) * P2(k3) * (1 + ratio) -
              c13 * k1 * k3 * (c15 * P2(k1) * (1 + 3 * ratio) + k3 * (2 * c55 * k1 * (1 + ratio) + c35 * k3 * (3 + ratio))) +
              e2 * (k3 * (3 * c15 * k1 * (1 + ratio) + 3 * c35 * k1 * (1 + ratio) + c33 * k3 * (2 + ratio)) +
                    c55 * (P2(k3) * (2 + ratio) + P2(k1) * (1 + 2 * ratio))) * rho +
              2 * e4 * (1 + ratio) * P2(rho) +
              c11 * P2(k1) * (2 * c55 * P2(k1) * ratio + c33 * P2(k3) * (1 + ratio) + c35 * k1 * k3 * (1 + 3 * ratio) +
                              e2 * (1 + 2 * ratio) * rho);
    } else {
        num = -2 * P2(c15) * P4(k1) +
              P2(k3) * (-2 * (P2(c35) - c33 * c55) * P2(k3) * ratio - P2(c13) * P2(k1) * (1 + ratio) -
                        c13 * k1 * (2 * c55 * k1 * (1 + ratio) + c35 * k3 * (1 + 3 * ratio))) +
              e2 * (k3 * (3 * c35 * k1 * (1 + ratio) + c33 * k3 * (1 + 2 * ratio)) +
                    c55 * (P2(k1) * (2 + ratio) + P2(k3) * (1 + 2 * ratio))) * rho +
              2 * e4 * (1 + ratio) * P2(rho) +
              c15 * k1 * k3 * (-(c13 * P2(k1) * (3 + ratio)) + k3 * (2 * c35 * k1 * (1 + ratio) + c33 * k3 * (1 + 3 * ratio)) +
                               3 * e2 * (1 + ratio) * rho) +
              c11 * P2(k1) * (2 * c55 * P2(k1) + k3 * (c33 * k3 * (1 + ratio) + c35 * k1 * (3 + ratio)) +
                              e2 * (2 + ratio) * rho);
    }
    return num / den;
}

/* damp_ratio: at each of 180 angles, raise the ratio in steps of 0.002
 * (never lowering it between angles) until the eigenvalue derivative of
 * both modes is below -0.01, or the ratio reaches 1.  The derivative takes
 * the moduli and density scaled by 1e-9 and the unscaled eigenvalues. */
float ORA_SYM(fsg_damp_ratio)(float c11, float c13, float c15, float c33, float c35, float c55, float rho, int dirx) {
    const float dtheta = (float)(FSG_PI / 180.0), alpha = -0.01f, dratio = 0.002f;
    const float s = 1.0e-9f;
    const float d11 = c11 * s, d13 = c13 * s, d15 = c15 * s, d33 = c33 * s, d35 = c35 * s, d55 = c55 * s, dr = rho * s;
    float ratio = 0.0f;
    for (int i = 1; i <= 180; ++i) {
        const float theta = ((float)i + 0.5f) * dtheta;
        const float kx = sinf(theta), kz = cosf(theta);
        const float e1 = eig2_of(c11, c13, c15, c33, c35, c55, rho, 1, kx, kz);
        const float e2 = eig2_of(c11, c13, c15, c33, c35, c55, rho, 2, kx, kz);
        float dg = fmaxf(eig_deriv(d11, d13, d15, d33, d35, d55, dr, dirx, e1, ratio, kx, kz),
                         eig_deriv(d11, d13, d15, d33, d35, d55, dr, dirx, e2, ratio, kx, kz));
        while (dg >= alpha && ratio < 1.0f) {
            ratio = ratio + dratio;
            dg = fmaxf(eig_deriv(d11, d13, d15, d33, d35, d55, dr, dirx, e1, ratio, kx, kz),
                       eig_deriv(d11, d13, d15, d33, d35, d55, dr, dirx, e2, ratio, kx, kz));
        }
    }
    return ratio;
}

/* For checking: squared eigenvalues of both modes at direction (k1, k3)
 * and their M-PML derivatives at the given ratio, as damp_ratio forms them. */
void ORA_SYM(fsg_mpml_eig)(const float *c, float rho, int dirx, float ratio, float k1, float k3, float *out) {
    const float s = 1.0e-9f;
    out[0] = eig2_of(c[0], c[1], c[2], c[3], c[4], c[5], rho, 1, k1, k3);
    out[1] = eig2_of(c[0], c[1], c[2], c[3], c[4], c[5], rho, 2, k1, k3);
    out[2] = eig_deriv(c[0] * s, c[1] * s, c[2] * s, c[3] * s, c[4] * s, c[5] * s, rho * s, dirx, out[0], ratio, k1, k3);
    out[3] = eig_deriv(c[0] * s, c[1] * s, c[2] * s, c[3] * s, c[4] * s, c[5] * s, rho * s, dirx, out[1], ratio, k1, k3);
}

/* ---- medium, coefficients --------------------------------------------- */

typedef struct {
    float *a, *b; /* per cell, padded layout */
} abpair;

typedef struct {
    flay L;
    float c[8];
    real dx, dz;
    real *c11, *c13, *c15, *c33, *c35, *c55, *rho;
    abpair xii, zii, xih, zih, xhi, zhi, xhh, zhh;
} fmed;

typedef struct {
    real *sxx_ii, *szz_ii, *sxz_ii, *sxx_hh, *szz_hh, *sxz_hh, *vx_ih, *vz_ih, *vx_hi, *vz_hi;
    real *m_dxvx_hh, *m_dxvz_hh, *m_dxvx_ii, *m_dxvz_ii, *m_dxxx_hi, *m_dxxz_hi, *m_dxxx_ih, *m_dxxz_ih;
    real *m_dzvx_hh, *m_dzvz_hh, *m_dzvx_ii, *m_dzvz_ii, *m_dzxz_hi, *m_dzzz_hi, *m_dzxz_ih, *m_dzzz_ih;
    real *m_dzxx_hi, *m_dzxx_ih; /* free-surface update only (memory_pdzxx_hxiz, _ixhz) */
} ffld;

#define NFF 28
static real *fz(size_t n) { return (real *)calloc(n, sizeof(real)); }
static void ffld_alloc(ffld *F, const flay *L) {
    const size_t n = (size_t)L->NX * L->NZ;
    real **p = &F->sxx_ii;
    for (int k = 0; k < NFF; ++k) p[k] = fz(n);
}
static void ffld_free(ffld *F) {
    real **p = &F->sxx_ii;
    for (int k = 0; k < NFF; ++k) free(p[k]);
}

/* damp_coef (module_cfspml.f90:282-330) with the M-PML mixing wx = dampx +
 * ratioz dampz, wz = dampz + ratiox dampx and kmax = 1. */
static void damp_coef2(float pmlvp, float xd, float zd, float dx, float dz, int pml, float adt, float am, float rx,
                       float rz, float *ax, float *bx, float *az, float *bz) {
    const float R0 = 1.0e-5f, kx = 1.0f, kz = 1.0f;
    float dampx = pmlvp * (2 + 1.0f) * logf(1.0f / R0) / (2.0f * (float)pml * dx) * (xd * xd);
    float dampz = pmlvp * (2 + 1.0f) * logf(1.0f / R0) / (2.0f * (float)pml * dz) * (zd * zd);
    const float wx = dampx + rz * dampz, wz = dampz + rx * dampx;
    dampx = wx;
    dampz = wz;
    const float alx = fmaxf(am * (1.0f - xd), 0.25f * am);
    *ax = -2.0f * adt * dampx / kx / (2.0f + adt * (alx + dampx / kx));
    *bx = (2.0f - adt * (alx + dampx / kx)) / (2.0f + adt * (alx + dampx / kx));
    const float alz = fmaxf(am * (1.0f - zd), 0.25f * am);
    *az = -2.0f * adt * dampz / kz / (2.0f + adt * (alz + dampz / kz));
    *bz = (2.0f - adt * (alz + dampz / kz)) / (2.0f + adt * (alz + dampz / kz));
}

/* OWL's depth-refined free-surface mesh of one shot window
 * (elastic_tti_2d/module_vars.f90:360-519).  Row tables over j = -pml ..
 * nz+pml+1 (index j + pml), column tables over i = -pml .. nx+pml+1 (index
 * i + pml); eta_i and eta_h are clipped to [-pml dz, eta_max] as OWL clips
 * them before the time loop. */
typedef struct fsmesh {
    int nz;               /* refined rows */
    int nx, nz0, pml, nci, nrj;
    float dx, dz0, dz;    /* dz = eta_max / (nz0 - 1), the z spacing the update divides by */
    float eta_max, depth_max, topo_max;
    float *topo_i, *slope_i, *topo_h, *slope_h;
    float *eta_i, *eta_h, *edz_i, *edz_h, *sc_i, *sc_h;
    float *zz_i;          /* physical depth of the integer rows, [i + pml][j + pml], j fastest */
} fsmesh;
static void fsmesh_medium(const fsmesh *S, const ora_fgrid2d *g, const float *v, real *out, const flay *L);
static float fsmesh_edz(const fsmesh *S, int j);

/* S: the refined free-surface mesh (NULL for the regular mesh).  With S the
 * medium is OWL's map_regular_to_irregular of the window and the z damping
 * uses the local row spacing eta_dz_i(j) (module_cfspml.f90:395-480). */
static void fmed_init(fmed *M, const ora_fgrid2d *g, const float *const *cij, const float *rho, float f0max,
                      const fsmesh *S);
static void fmed_init(fmed *M, const ora_fgrid2d *g, const float *const *cij, const float *rho, float f0max,
                      const fsmesh *S) {
    const int R = g->R, pml = g->pml, nx = g->nx, nz0 = g->nz;
    const int nz = S ? S->nz : nz0;
    flay L = {1 - pml - R, nx + 2 * (pml + R), nz + 2 * (pml + R), nx, nz, pml, R};
    M->L = L;
    ORA_SYM(fdcoefs)(R, M->c);
    M->dx = (real)g->dx;
    M->dz = (real)g->dz;
    const size_t n = (size_t)L.NX * L.NZ;
    real **cs[7] = {&M->c11, &M->c13, &M->c15, &M->c33, &M->c35, &M->c55, &M->rho};
    for (int k = 0; k < 7; ++k) *cs[k] = fz(n);
    if (S) {
        for (int k = 0; k < 6; ++k) fsmesh_medium(S, g, cij[k], *cs[k], &L);
        fsmesh_medium(S, g, rho, M->rho, &L);
    } else {
        for (int j = L.lo; j < L.lo + L.NZ; ++j)
            for (int i = L.lo; i < L.lo + L.NX; ++i) {
                const int ic = i < 1 ? 1 : (i > nx ? nx : i), jc = j < 1 ? 1 : (j > nz ? nz : j);
                const size_t m = (size_t)(jc - 1) + (size_t)(ic - 1) * nz, q = IF(L, i, j);
                for (int k = 0; k < 6; ++k) (*cs[k])[q] = cij[k][m];
                M->rho[q] = rho[m];
            }
    }
    /* pmlvp (module_vars.f90:345-349): c11/rho and c33/rho on the four edge
     * lines of the regular window, before any refinement */
    float pm = 0.0f;
    for (int i = 1; i <= nx; ++i)
        for (int j = 1; j <= nz0; ++j) {
            if (!(i == 1 || i == nx || j == 1 || j == nz0)) continue;
            const size_t m = (size_t)(j - 1) + (size_t)(i - 1) * nz0;
            const float a = cij[0][m] / rho[m], c = cij[3][m] / rho[m];
            if (a > pm) pm = a;
            if (c > pm) pm = c;
        }
    const float pmlvp = sqrtf(pm), am = ORA_SYM(alphamax)(f0max), adt = fabsf(g->dt);
    /* damping ratios on the four sides (compute_damping_profile_ratio) */
    const int n1 = -pml, n2 = nz + pml + 1, m1 = -pml, m2 = nx + pml + 1;
    float *left = (float *)calloc((size_t)(n2 - n1 + 1), sizeof(float));
    float *right = (float *)calloc((size_t)(n2 - n1 + 1), sizeof(float));
    float *top = (float *)calloc((size_t)(m2 - m1 + 1), sizeof(float));
    float *bottom = (float *)calloc((size_t)(m2 - m1 + 1), sizeof(float));
    if (g->aniso) {
#pragma omp parallel for schedule(dynamic)
        for (int j = n1; j <= n2; ++j) {
            const size_t ql = IF(L, 1, j), qr = IF(L, nx, j);
            left[j - n1] = ORA_SYM(fsg_damp_ratio)((float)M->c11[ql], (float)M->c13[ql], (float)M->c15[ql], (float)M->c33[ql],
                                                   (float)M->c35[ql], (float)M->c55[ql], (float)M->rho[ql], 1);
            right[j - n1] = ORA_SYM(fsg_damp_ratio)((float)M->c11[qr], (float)M->c13[qr], (float)M->c15[qr], (float)M->c33[qr],
                                                    (float)M->c35[qr], (float)M->c55[qr], (float)M->rho[qr], 1);
        }
#pragma omp parallel for schedule(dynamic)
        for (int i = m1; i <= m2; ++i) {
            const size_t qt = IF(L, i, 1), qb = IF(L, i, nz);
            top[i - m1] = ORA_SYM(fsg_damp_ratio)((float)M->c11[qt], (float)M->c13[qt], (float)M->c15[qt], (float)M->c33[qt],
                                                  (float)M->c35[qt], (float)M->c55[qt], (float)M->rho[qt], 0);
            bottom[i - m1] = ORA_SYM(fsg_damp_ratio)((float)M->c11[qb], (float)M->c13[qb], (float)M->c15[qb], (float)M->c33[qb],
                                                     (float)M->c35[qb], (float)M->c55[qb], (float)M->rho[qb], 0);
        }
    }
    abpair *sets[8] = {&M->xii, &M->zii, &M->xih, &M->zih, &M->xhi, &M->zhi, &M->xhh, &M->zhh};
    for (int k = 0; k < 8; ++k) {
        sets[k]->a = (float *)calloc(n, sizeof(float));
        sets[k]->b = (float *)calloc(n, sizeof(float));
    }
    for (int j = -pml + 1; j <= nz + pml; ++j)
        for (int i = -pml + 1; i <= nx + pml; ++i) {
            float xi = 0, xh = 0, zi = 0, zh = 0;
            if (i <= 1) {
                xi = (float)fabs((i - 1 + 0.0) / pml);
                xh = (float)fabs((i - 1 - 0.5) / pml);
            } else if (i >= nx) {
                xi = (float)fabs((i - nx + 0.0) / pml);
                xh = (float)fabs((i - nx - 0.5) / pml);
            }
            if (j <= 1) {
                zi = (float)fabs((j - 1 + 0.0) / pml);
                zh = (float)fabs((j - 1 - 0.5) / pml);
            } else if (j >= nz) {
                zi = (float)fabs((j - nz + 0.0) / pml);
                zh = (float)fabs((j - nz - 0.5) / pml);
            }
            if (i > 1 && i < nx && j > 1 && j < nz) continue;
            float rx = 0.0f, rz = 0.0f;
            if (i <= 1) rx = left[j - n1];
            else if (i >= nx) rx = right[j - n1];
            if (j <= 1) rz = top[i - m1];
            else if (j >= nz) rz = bottom[i - m1];
            const size_t q = IF(L, i, j);
            const float dzj = S ? fsmesh_edz(S, j) : g->dz;
            damp_coef2(pmlvp, xi, zi, g->dx, dzj, pml, adt, am, rx, rz, &M->xii.a[q], &M->xii.b[q], &M->zii.a[q], &M->zii.b[q]);
            damp_coef2(pmlvp, xi, zh, g->dx, dzj, pml, adt, am, rx, rz, &M->xih.a[q], &M->xih.b[q], &M->zih.a[q], &M->zih.b[q]);
            damp_coef2(pmlvp, xh, zi, g->dx, dzj, pml, adt, am, rx, rz, &M->xhi.a[q], &M->xhi.b[q], &M->zhi.a[q], &M->zhi.b[q]);
            damp_coef2(pmlvp, xh, zh, g->dx, dzj, pml, adt, am, rx, rz, &M->xhh.a[q], &M->xhh.b[q], &M->zhh.a[q], &M->zhh.b[q]);
        }
    free(left); free(right); free(top); free(bottom);
}

static void fmed_free(fmed *M) {
    real **cs[7] = {&M->c11, &M->c13, &M->c15, &M->c33, &M->c35, &M->c55, &M->rho};
    for (int k = 0; k < 7; ++k) free(*cs[k]);
    abpair *sets[8] = {&M->xii, &M->zii, &M->xih, &M->zih, &M->xhi, &M->zhi, &M->xhh, &M->zhh};
    for (int k = 0; k < 8; ++k) { free(sets[k]->a); free(sets[k]->b); }
}

/* m = a d + b m; d = (d + m) / kappa with kappa = 1 */
#define MP(mem, q, d, A, B)                                             \
    do {                                                                \
        (mem)[q] = (real)(A)[q] * (d) + (real)(B)[q] * (mem)[q];        \
        d = ((d) + (mem)[q]) / (real)1.0f;                              \
    } while (0)

/* x / z derivative stencils: sum_m c_m (w(i+m) - w(i-m+1)) with the first
 * point of the difference at offset off (0 or -1 in storage index) */
static inline real dxs(const real *w, size_t q, int off, const float *c, int R) {
    real s = (real)c[0] * (w[q + off + 1] - w[q + off]);
    for (int m = 2; m <= R; ++m) s = s + (real)c[m - 1] * (w[q + off + m] - w[q + off - (m - 1)]);
    return s;
}
static inline real dzs(const real *w, size_t q, long off, size_t sz, const float *c, int R) {
    real s = (real)c[0] * (w[q + (off + 1) * (long)sz] - w[q + off * (long)sz]);
    for (int m = 2; m <= R; ++m) s = s + (real)c[m - 1] * (w[q + (off + m) * (long)sz] - w[q + (off - (m - 1)) * (long)sz]);
    return s;
}

/* update_wavefield (module_wavefield.f90:35-252) in storage index: stress
 * (i, j) on the integer grid and at (m, n) = (i+1/2, j+1/2); vx, vz at
 * (m, j) = (i+1/2, j) and (i, n) = (i, j+1/2); each over [-pml+1, n+pml]. */
static void fupdate(const fmed *M, ffld *F, real dt) {
    const flay L = M->L;
    const int R = L.R, pml = L.pml, nx = L.nx, nz = L.nz;
    const size_t sz = (size_t)L.NX;
    const float *c = M->c;
    const real dx = M->dx, dz = M->dz;
#pragma omp parallel for schedule(static)
    for (int j = -pml + 1; j <= nz + pml; ++j)
        for (int i = -pml + 1; i <= nx + pml; ++i) {
            const size_t q = IF(L, i, j);
            /* integer-grid stress: x derivatives on the (i+1/2, j) grid, z on (i, j+1/2) */
            real pdxvx = dxs(F->vx_hi, q, 0, c, R) / dx, pdxvz = dxs(F->vz_hi, q, 0, c, R) / dx;
            real pdzvx = dzs(F->vx_ih, q, 0, sz, c, R) / dz, pdzvz = dzs(F->vz_ih, q, 0, sz, c, R) / dz;
            if (i <= 0 || i >= nx + 1 || j <= 0 || j >= nz + 1) {
                MP(F->m_dxvx_ii, q, pdxvx, M->xii.a, M->xii.b);
                MP(F->m_dxvz_ii, q, pdxvz, M->xii.a, M->xii.b);
                MP(F->m_dzvx_ii, q, pdzvx, M->zii.a, M->zii.b);
                MP(F->m_dzvz_ii, q, pdzvz, M->zii.a, M->zii.b);
            }
            const real c11 = M->c11[q], c13 = M->c13[q], c15 = M->c15[q], c33 = M->c33[q], c35 = M->c35[q], c55 = M->c55[q];
            F->sxx_ii[q] = F->sxx_ii[q] + dt * (c11 * pdxvx + c13 * pdzvz + c15 * (pdxvz + pdzvx));
            F->szz_ii[q] = F->szz_ii[q] + dt * (c13 * pdxvx + c33 * pdzvz + c35 * (pdxvz + pdzvx));
            F->sxz_ii[q] = F->sxz_ii[q] + dt * (c15 * pdxvx + c35 * pdzvz + c55 * (pdxvz + pdzvx));
        }
#pragma omp parallel for schedule(static)
    for (int n = -pml + 1; n <= nz + pml; ++n)
        for (int m = -pml + 1; m <= nx + pml; ++m) {
            const size_t q = IF(L, m, n);
            /* corner stress: x derivatives on the (i, j+1/2) grid, z on (i+1/2, j) */
            real pdxvx = dxs(F->vx_ih, q, -1, c, R) / dx, pdxvz = dxs(F->vz_ih, q, -1, c, R) / dx;
            real pdzvx = dzs(F->vx_hi, q, -1, sz, c, R) / dz, pdzvz = dzs(F->vz_hi, q, -1, sz, c, R) / dz;
            if (m <= 1 || m >= nx + 1 || n <= 1 || n >= nz + 1) {
                MP(F->m_dxvx_hh, q, pdxvx, M->xhh.a, M->xhh.b);
                MP(F->m_dxvz_hh, q, pdxvz, M->xhh.a, M->xhh.b);
                MP(F->m_dzvx_hh, q, pdzvx, M->zhh.a, M->zhh.b);
                MP(F->m_dzvz_hh, q, pdzvz, M->zhh.a, M->zhh.b);
            }
            const size_t q00 = q - 1 - sz, q10 = q - sz, q01 = q - 1, q11 = q;
#define AV4(a) ((real)0.25f * ((a)[q00] + (a)[q10] + (a)[q01] + (a)[q11]))
            const real c11 = AV4(M->c11), c13 = AV4(M->c13), c15 = AV4(M->c15), c33 = AV4(M->c33), c35 = AV4(M->c35),
                       c55 = AV4(M->c55);
#undef AV4
            F->sxx_hh[q] = F->sxx_hh[q] + dt * (c11 * pdxvx + c13 * pdzvz + c15 * (pdxvz + pdzvx));
            F->szz_hh[q] = F->szz_hh[q] + dt * (c13 * pdxvx + c33 * pdzvz + c35 * (pdxvz + pdzvx));
            F->sxz_hh[q] = F->sxz_hh[q] + dt * (c15 * pdxvx + c35 * pdzvz + c55 * (pdxvz + pdzvx));
        }
#pragma omp parallel for schedule(static)
    for (int j = -pml + 1; j <= nz + pml; ++j)
        for (int m = -pml + 1; m <= nx + pml; ++m) {
            const size_t q = IF(L, m, j);
            /* velocity on (i+1/2, j): x from integer-grid stress, z from corner stress */
            real pdxxx = dxs(F->sxx_ii, q, -1, c, R) / dx, pdxxz = dxs(F->sxz_ii, q, -1, c, R) / dx;
            real pdzxz = dzs(F->sxz_hh, q, 0, sz, c, R) / dz, pdzzz = dzs(F->szz_hh, q, 0, sz, c, R) / dz;
            if (m <= 1 || m >= nx + 1 || j <= 0 || j >= nz + 1) {
                MP(F->m_dxxx_hi, q, pdxxx, M->xhi.a, M->xhi.b);
                MP(F->m_dxxz_hi, q, pdxxz, M->xhi.a, M->xhi.b);
                MP(F->m_dzxz_hi, q, pdzxz, M->zhi.a, M->zhi.b);
                MP(F->m_dzzz_hi, q, pdzzz, M->zhi.a, M->zhi.b);
            }
            const real rh = (real)0.5f * (M->rho[q - 1] + M->rho[q]);
            F->vx_hi[q] = F->vx_hi[q] + dt / rh * (pdxxx + pdzxz);
            F->vz_hi[q] = F->vz_hi[q] + dt / rh * (pdxxz + pdzzz);
        }
#pragma omp parallel for schedule(static)
    for (int n = -pml + 1; n <= nz + pml; ++n)
        for (int i = -pml + 1; i <= nx + pml; ++i) {
            const size_t q = IF(L, i, n);
            /* velocity on (i, j+1/2): x from corner stress, z from integer-grid stress */
            real pdxxx = dxs(F->sxx_hh, q, 0, c, R) / dx, pdxxz = dxs(F->sxz_hh, q, 0, c, R) / dx;
            real pdzxz = dzs(F->sxz_ii, q, -1, sz, c, R) / dz, pdzzz = dzs(F->szz_ii, q, -1, sz, c, R) / dz;
            if (i <= 0 || i >= nx + 1 || n <= 1 || n >= nz + 1) {
                MP(F->m_dxxx_ih, q, pdxxx, M->xih.a, M->xih.b);
                MP(F->m_dxxz_ih, q, pdxxz, M->xih.a, M->xih.b);
                MP(F->m_dzxz_ih, q, pdzxz, M->zih.a, M->zih.b);
                MP(F->m_dzzz_ih, q, pdzzz, M->zih.a, M->zih.b);
            }
            const real rh = (real)0.5f * (M->rho[q - sz] + M->rho[q]);
            F->vx_ih[q] = F->vx_ih[q] + dt / rh * (pdxxx + pdzxz);
            F->vz_ih[q] = F->vz_ih[q] + dt / rh * (pdxxz + pdzzz);
        }
}

/* add_source (module_wavefield.f90:593-736) without a free surface: amp =
 * stf amp dt; a force adds amp / rho(gx, gz) along (sin polar, cos polar)
 * to both velocity grids; an explosion subtracts amp / (dx dz) from the
 * normal stresses of both stress grids. */
static void fadd_source(const fmed *M, ffld *F, const ora_fsource2d *src, const ora_point *pts, int ns, int t,
                        real dt) {
    const flay L = M->L;
    for (int s = 0; s < ns; ++s) {
        const ora_point *P = &pts[s];
        const int nb = (int)roundf(src[s].t0 / (float)dt) + 1, ne = nb + src[s].nt_stf - 1;
        if (t < nb || t > ne) continue;
        real amp = src[s].stf[t - nb] * (real)src[s].amp * dt;
        if (src[s].mechanism == 1) {
            amp = amp / (M->dx * M->dz);
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int irx = -NKW; irx <= NKW; ++irx) {
                    const size_t q1 = IF(L, P->gx + irx, P->gz + irz);
                    F->sxx_ii[q1] = F->sxx_ii[q1] - amp * (real)P->ix[irx + NKW] * (real)P->iz[irz + NKW];
                    F->szz_ii[q1] = F->szz_ii[q1] - amp * (real)P->ix[irx + NKW] * (real)P->iz[irz + NKW];
                    const size_t q2 = IF(L, P->hx + irx, P->hz + irz);
                    F->sxx_hh[q2] = F->sxx_hh[q2] - amp * (real)P->hxw[irx + NKW] * (real)P->hzw[irz + NKW];
                    F->szz_hh[q2] = F->szz_hh[q2] - amp * (real)P->hxw[irx + NKW] * (real)P->hzw[irz + NKW];
                }
            continue;
        }
        amp = amp / M->rho[IF(L, P->gx, P->gz)];
#if defined(ORA_F64)
        const real sp = sin((double)src[s].polar), cp = cos((double)src[s].polar);
#else
        const real sp = sinf(src[s].polar), cp = cosf(src[s].polar);
#endif
        for (int irz = -NKW; irz <= NKW; ++irz)
            for (int irx = -NKW; irx <= NKW; ++irx) {
                const size_t q1 = IF(L, P->hx + irx, P->gz + irz);
                F->vx_hi[q1] = F->vx_hi[q1] + sp * amp * (real)P->hxw[irx + NKW] * (real)P->iz[irz + NKW];
                F->vz_hi[q1] = F->vz_hi[q1] + cp * amp * (real)P->hxw[irx + NKW] * (real)P->iz[irz + NKW];
                const size_t q2 = IF(L, P->gx + irx, P->hz + irz);
                F->vx_ih[q2] = F->vx_ih[q2] + sp * amp * (real)P->ix[irx + NKW] * (real)P->hzw[irz + NKW];
                F->vz_ih[q2] = F->vz_ih[q2] + cp * amp * (real)P->ix[irx + NKW] * (real)P->hzw[irz + NKW];
            }
    }
}

/* receivers: half of each velocity grid (submodule_forward_modeling.f90:80-110) */
static void frecord(const flay *L, const ffld *F, const ora_freceiver2d *rec, const ora_point *rp, int nr, int nt, int t,
                    real *tr) {
    for (int r = 0; r < nr; ++r) {
        if (rec[r].weight == 0.0f) continue;
        const ora_point *Q = &rp[r];
        const float w = rec[r].weight;
        real ax = 0, az = 0;
        for (int irz = -NKW; irz <= NKW; ++irz)
            for (int irx = -NKW; irx <= NKW; ++irx) {
                const real a1 = (real)(Q->hxw[irx + NKW] * Q->iz[irz + NKW] * w * 0.5f);
                const real a2 = (real)(Q->ix[irx + NKW] * Q->hzw[irz + NKW] * w * 0.5f);
                ax = ax + F->vx_hi[IF(*L, Q->hx + irx, Q->gz + irz)] * a1 + F->vx_ih[IF(*L, Q->gx + irx, Q->hz + irz)] * a2;
            }
        for (int irz = -NKW; irz <= NKW; ++irz)
            for (int irx = -NKW; irx <= NKW; ++irx) {
                const real a1 = (real)(Q->hxw[irx + NKW] * Q->iz[irz + NKW] * w * 0.5f);
                const real a2 = (real)(Q->ix[irx + NKW] * Q->hzw[irz + NKW] * w * 0.5f);
                az = az + F->vz_hi[IF(*L, Q->hx + irx, Q->gz + irz)] * a1 + F->vz_ih[IF(*L, Q->gx + irx, Q->hz + irz)] * a2;
            }
        tr[(size_t)r * nt + (t - 1)] = ax;
        tr[((size_t)nr + r) * nt + (t - 1)] = az;
    }
}

/* boundary records (module_boundary_saving.f90:100-200): vx_hi, vz_hi,
 * vx_ih, vz_ih; each left [2-R, 1] x [1-R, nz+R], right [nx+1, nx+R] x
 * [1-R, nz+R], top [1-R, nx+R] x [2-R, 1], bottom [1-R, nx+R] x [nz+1, nz+R]. */
static size_t fstrip_len(const flay *L) {
    const size_t R = (size_t)L->R, ex = L->nx + 2 * R, ez = L->nz + 2 * R;
    return 4 * (2 * R * ez + 2 * R * ex);
}

static void fstrips(const flay *L, ffld *F, real *rec, int save) {
    const int R = L->R, nx = L->nx, nz = L->nz;
    size_t k = 0;
    real *v4[4] = {F->vx_hi, F->vz_hi, F->vx_ih, F->vz_ih};
    for (int c = 0; c < 4; ++c) {
        real *f = v4[c];
#define XF(i, j)                         \
    do {                                 \
        const size_t q_ = IF(*L, i, j);  \
        if (save) rec[k] = f[q_];        \
        else f[q_] = rec[k];             \
        ++k;                             \
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

static int fprepare(const ora_fgrid2d *g, const ora_fsource2d *src, int ns, const ora_freceiver2d *rec, int nr,
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

#define NSTATE 10 /* the ten wavefields (memory variables are not part of a state) */

static int fforward(const ora_fgrid2d *g, const fmed *M, const ora_fsource2d *src, int ns, int nr,
                    const ora_freceiver2d *rec, real *tr, real *strips, real *fin) {
    ora_point *sp, *rp;
    if (fprepare(g, src, ns, rec, nr, &sp, &rp)) return -1;
    const flay L = M->L;
    const size_t n = (size_t)L.NX * L.NZ, sl = fstrip_len(&L);
    ffld F;
    ffld_alloc(&F, &L);
    memset(tr, 0, sizeof(real) * 2 * (size_t)nr * g->nt);
    for (int t = 1; t <= g->nt; ++t) {
        if (strips) fstrips(&L, &F, strips + (size_t)(t - 1) * sl, 1);
        if (t == g->nt && fin) {
            real **fw = &F.sxx_ii;
            for (int k = 0; k < NSTATE; ++k) memcpy(fin + (size_t)k * n, fw[k], n * sizeof(real));
        }
        frecord(&L, &F, rec, rp, nr, g->nt, t, tr);
        if (t == g->nt) break;
        fupdate(M, &F, (real)g->dt);
        fadd_source(M, &F, src, sp, ns, t, (real)g->dt);
    }
    ffld_free(&F);
    free(sp);
    free(rp);
    return 0;
}

/* cij: c11, c13, c15, c33, c35, c55; all (nz, nx) z fastest */
int ORA_SYM(fsg2d_forward)(const ora_fgrid2d *g, const float *const *cij, const float *rho, int ns,
                           const ora_fsource2d *src, float f0max, int nr, const ora_freceiver2d *rec, real *traces) {
    fmed M;
    fmed_init(&M, g, cij, rho, f0max, NULL);
    const int rc = fforward(g, &M, src, ns, nr, rec, traces, NULL, NULL);
    fmed_free(&M);
    return rc;
}

/* Damping-ratio profiles of the four sides (left, right over j = -pml ..
 * nz+pml+1; top, bottom over i = -pml .. nx+pml+1), for checking. */
int ORA_SYM(fsg2d_ratios)(const ora_fgrid2d *g, const float *const *cij, const float *rho, float *left, float *right,
                          float *top, float *bottom) {
    fmed M;
    fmed_init(&M, g, cij, rho, 10.0f, NULL);
    const flay L = M.L;
    for (int j = -g->pml; j <= g->nz + g->pml + 1; ++j) {
        const size_t ql = IF(L, 1, j), qr = IF(L, g->nx, j);
        left[j + g->pml] = ORA_SYM(fsg_damp_ratio)((float)M.c11[ql], (float)M.c13[ql], (float)M.c15[ql], (float)M.c33[ql],
                                                   (float)M.c35[ql], (float)M.c55[ql], (float)M.rho[ql], 1);
        right[j + g->pml] = ORA_SYM(fsg_damp_ratio)((float)M.c11[qr], (float)M.c13[qr], (float)M.c15[qr], (float)M.c33[qr],
                                                    (float)M.c35[qr], (float)M.c55[qr], (float)M.rho[qr], 1);
    }
    for (int i = -g->pml; i <= g->nx + g->pml + 1; ++i) {
        const size_t qt = IF(L, i, 1), qb = IF(L, i, g->nz);
        top[i + g->pml] = ORA_SYM(fsg_damp_ratio)((float)M.c11[qt], (float)M.c13[qt], (float)M.c15[qt], (float)M.c33[qt],
                                                  (float)M.c35[qt], (float)M.c55[qt], (float)M.rho[qt], 0);
        bottom[i + g->pml] = ORA_SYM(fsg_damp_ratio)((float)M.c11[qb], (float)M.c13[qb], (float)M.c15[qb], (float)M.c33[qb],
                                                     (float)M.c35[qb], (float)M.c55[qb], (float)M.rho[qb], 0);
    }
    fmed_free(&M);
    return 0;
}

/* OWL's FSG gradient (submodule_fwi_gradient.f90:130-400, 628-727):
 * reverse t = nt .. hnt: prev = state; state = update(state, -dt); at t = nt
 * the final state; boundary record t; adjoint = update(adjoint, -dt);
 * adjoint sources into both velocity grids (half each); imaging with the
 * sub-grid sums of the stress increments and velocities and the 2D
 * compliance (cij_to_sij_2d).  adj: two [nr][nt] blocks (x, z).
 * restore_all 0 restores the final state in OWL's read order, 1 restores
 * all ten fields.  Outputs
 * (nz, nx) z fastest: gradients of c11, c13, c15, c33, c35, c55, rho. */
int ORA_SYM(fsg2d_gradient_owl)(const ora_fgrid2d *g, const float *const *cij, const float *rho, int ns,
                                const ora_fsource2d *src, float f0max, int hnt, int nr, const ora_freceiver2d *rec,
                                const real *adj, int restore_all, real *const *grad, real *traces_out) {
    fmed M;
    fmed_init(&M, g, cij, rho, f0max, NULL);
    const flay L = M.L;
    const size_t n = (size_t)L.NX * L.NZ, sl = fstrip_len(&L);
    const int nt = g->nt;
    real *strips = fz(sl * (size_t)nt), *fin = fz((size_t)NSTATE * n);
    real *tr = traces_out ? traces_out : fz(2 * (size_t)nr * nt);
    int rc = fforward(g, &M, src, ns, nr, rec, tr, strips, fin);
    if (rc) {
        free(strips); free(fin);
        if (!traces_out) free(tr);
        fmed_free(&M);
        return rc;
    }
    ora_point *sp, *rp;
    fprepare(g, src, ns, rec, nr, &sp, &rp);
    ffld F, A, P;
    ffld_alloc(&F, &L);
    ffld_alloc(&A, &L);
    ffld_alloc(&P, &L);
    const int nx = g->nx, nz = g->nz;
    const size_t ni = (size_t)nx * nz;
    real *gc[7];
    for (int c = 0; c < 7; ++c) gc[c] = fz(ni);
    const real mdt = -(real)g->dt;
    const int cc = g->cc_step_interval > 0 ? g->cc_step_interval : 1;
    const size_t sx = 1, sz = (size_t)L.NX;
    for (int t = nt; t >= hnt; --t) {
        real **fw = &F.sxx_ii, **pw = &P.sxx_ii;
        for (int k = 0; k < NSTATE; ++k) memcpy(pw[k], fw[k], n * sizeof(real));
        fupdate(&M, &F, mdt);
        if (t == nt) {
            /* input_final_step_wavefield (module_boundary_saving.f90:77-95)
             * reads the ten saved fields back into vx_hi, vx_ih, vz_hi,
             * vz_hi, ...: the saved vz_ih lands in vz_hi and vz_ih keeps
             * its reconstructed value.  restore_all puts every field back. */
            for (int k = 0; k < NSTATE; ++k) {
                if (!restore_all && (fw[k] == F.vz_hi || fw[k] == F.vz_ih)) continue;
                memcpy(fw[k], fin + (size_t)k * n, n * sizeof(real));
            }
            if (!restore_all) memcpy(F.vz_hi, fin + (size_t)7 * n, n * sizeof(real)); /* saved vz_ih */
        }
        fstrips(&L, &F, strips + (size_t)(t - 1) * sl, 0);
        fupdate(&M, &A, mdt);
        for (int r = 0; r < nr; ++r) {
            if (rec[r].weight == 0.0f) continue;
            const ora_point *Q = &rp[r];
            const float w = rec[r].weight;
            const real ax = adj[(size_t)r * nt + (t - 1)], az = adj[((size_t)nr + r) * nt + (t - 1)];
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int irx = -NKW; irx <= NKW; ++irx) {
                    const real a1 = (real)(Q->hxw[irx + NKW] * Q->iz[irz + NKW] * w * 0.5f);
                    const size_t q = IF(L, Q->hx + irx, Q->gz + irz);
                    A.vx_hi[q] = A.vx_hi[q] + ax * a1;
                    A.vz_hi[q] = A.vz_hi[q] + az * a1;
                }
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int irx = -NKW; irx <= NKW; ++irx) {
                    const real a2 = (real)(Q->ix[irx + NKW] * Q->hzw[irz + NKW] * w * 0.5f);
                    const size_t q = IF(L, Q->gx + irx, Q->hz + irz);
                    A.vx_ih[q] = A.vx_ih[q] + ax * a2;
                    A.vz_ih[q] = A.vz_ih[q] + az * a2;
                }
        }
        if (t % cc != 0) continue;
#pragma omp parallel for schedule(static)
        for (int j = 1; j <= nz; ++j)
            for (int i = 1; i <= nx; ++i) {
                const size_t q = IF(L, i, j), o = (size_t)(i - 1) + (size_t)(j - 1) * nx;
                const real c11 = M.c11[q], c13 = M.c13[q], c15 = M.c15[q], c33 = M.c33[q], c35 = M.c35[q], c55 = M.c55[q];
                const real t1 = c15 * c15 * c33 - 2 * c13 * c15 * c35 + c11 * (c35 * c35) + c13 * c13 * c55 - c11 * c33 * c55;
                const real t2 = c15 * c15 * c33 - 2 * c13 * c15 * c35 + c13 * c13 * c55 + c11 * (c35 * c35 - c33 * c55);
                const real s11 = (c35 * c35 - c33 * c55) / t1, s13 = (-(c15 * c35) + c13 * c55) / t2;
                const real s15 = (c15 * c33 - c13 * c35) / t1, s33 = (c15 * c15 - c11 * c55) / t1;
                const real s35 = (-(c13 * c15) + c11 * c35) / t2, s55 = (c13 * c13 - c11 * c33) / t1;
                /* corner stresses around the node: storage (i..i+1, j..j+1) */
#define H4(a) ((a)[q] + (a)[q + sx] + (a)[q + sz] + (a)[q + sx + sz])
                const real txx = (F.sxx_ii[q] + (real)0.25f * H4(F.sxx_hh)) - (P.sxx_ii[q] + (real)0.25f * H4(P.sxx_hh));
                const real tzz = (F.szz_ii[q] + (real)0.25f * H4(F.szz_hh)) - (P.szz_ii[q] + (real)0.25f * H4(P.szz_hh));
                const real txz = (F.sxz_ii[q] + (real)0.25f * H4(F.sxz_hh)) - (P.sxz_ii[q] + (real)0.25f * H4(P.sxz_hh));
                const real rxx = A.sxx_ii[q] + (real)0.25f * H4(A.sxx_hh);
                const real rzz = A.szz_ii[q] + (real)0.25f * H4(A.szz_hh);
                const real rxz = A.sxz_ii[q] + (real)0.25f * H4(A.sxz_hh);
#undef H4
                const real exx = s11 * txx + s13 * tzz + s15 * txz, ezz = s13 * txx + s33 * tzz + s35 * txz,
                           exz = s15 * txx + s35 * tzz + s55 * txz;
                const real fxx = s11 * rxx + s13 * rzz + s15 * rxz, fzz = s13 * rxx + s33 * rzz + s35 * rxz,
                           fxz = s15 * rxx + s35 * rzz + s55 * rxz;
                gc[0][o] = gc[0][o] - (fxx * exx);
                gc[1][o] = gc[1][o] - (fxx * ezz + fzz * exx);
                gc[2][o] = gc[2][o] - (fxx * exz + exx * fxz);
                gc[3][o] = gc[3][o] - (fzz * ezz);
                gc[4][o] = gc[4][o] - (fzz * exz + ezz * fxz);
                gc[5][o] = gc[5][o] - (fxz * exz);
                const real svx = ((real)0.5f * (F.vx_hi[q] + F.vx_hi[q + sx]) + (real)0.5f * (F.vx_ih[q] + F.vx_ih[q + sz])) -
                                 ((real)0.5f * (P.vx_hi[q] + P.vx_hi[q + sx]) + (real)0.5f * (P.vx_ih[q] + P.vx_ih[q + sz]));
                const real rvx = (real)0.5f * (A.vx_hi[q] + A.vx_hi[q + sx]) + (real)0.5f * (A.vx_ih[q] + A.vx_ih[q + sz]);
                const real svz = ((real)0.5f * (F.vz_hi[q] + F.vz_hi[q + sx]) + (real)0.5f * (F.vz_ih[q] + F.vz_ih[q + sz])) -
                                 ((real)0.5f * (P.vz_hi[q] + P.vz_hi[q + sx]) + (real)0.5f * (P.vz_ih[q] + P.vz_ih[q + sz]));
                const real rvz = (real)0.5f * (A.vz_hi[q] + A.vz_hi[q + sx]) + (real)0.5f * (A.vz_ih[q] + A.vz_ih[q + sz]);
                gc[6][o] = gc[6][o] + svx * rvx + svz * rvz;
            }
    }
    for (int j = 1; j <= nz; ++j)
        for (int i = 1; i <= nx; ++i) {
            const size_t o = (size_t)(i - 1) + (size_t)(j - 1) * nx, m = (size_t)(j - 1) + (size_t)(i - 1) * nz;
            for (int c = 0; c < 7; ++c) grad[c][m] = isfinite(gc[c][o]) ? gc[c][o] : 0;
        }
    for (int c = 0; c < 7; ++c) free(gc[c]);
    if (!traces_out) free(tr);
    free(strips); free(fin); free(sp); free(rp);
    ffld_free(&F); ffld_free(&A); ffld_free(&P);
    fmed_free(&M);
    return 0;
}

/* ======================================================================
 * Free surface on OWL's depth-refined, topography-following mesh
 *   module_vars.f90:360-519 (mesh), 169-249 (map_regular_to_irregular,
 *   map_irregular_to_regular), 639-680 (source/receiver depths)
 *   module_wavefield.f90:257-587 update_wavefield_free_surface, :593-736
 *   add_source (rows >= 2)
 *   module_cfspml.f90:395-480 damping with the local row spacing
 *   submodule_forward_modeling.f90:66-75 velocity mirror before recording
 *   submodule_fwi_gradient.f90:163, 300, 350-380, 456-470
 *   src/lib/module_source_receiver.f90:556-760 nearest rows on the mesh
 *
 * FLIT helpers restated here (FLIT is not part of the OWL tree; these are
 * assumptions, parity unpinned beyond the physical checks of the tests):
 *   regspace(a, d, b): nint((b - a)/d) + 1 values a + k d;
 *   deriv(x, 'center'): (x(i+1) - x(i-1))/2 inside, one-sided differences
 *     at the two ends;
 *   rescale(x, [r1, r2]): (x - min)/(max - min)*(r2 - r1) + r1;
 *   ginterp 'linear': piecewise linear, y0 + (y1 - y0)*t on the segment
 *     x0 <= v < x1, end segments extended;
 *   ginterp 'cubic': natural cubic spline in double, end polynomials
 *     extended;
 *   pad_array: edge replication; float_small = FLT_EPSILON.
 * ====================================================================== */

typedef struct {
    int ntopo;            /* topography control points; 0: flat at elevation 0 */
    const float *tx, *tz; /* x and elevation (positive up) */
    int interp;           /* topo_interp: 0 'linear', 1 'cubic' */
    float refine;         /* free_surface_dz_refine */
    float dz_max;         /* the caller's dz_max (module_model.f90:176-178) */
    int src_from_surface, rec_from_surface;
} ora_fsurf2d;

static float *flit_regspace(float a, float d, float b, int *n) {
    const int m = (int)lroundf((b - a) / d) + 1;
    float *x = (float *)malloc(sizeof(float) * (size_t)(m > 0 ? m : 1));
    for (int k = 0; k < m; ++k) x[k] = a + (float)k * d;
    *n = m;
    return x;
}

static void flit_deriv(const float *x, int n, float *d) {
    if (n < 2) {
        if (n == 1) d[0] = 0.0f;
        return;
    }
    d[0] = x[1] - x[0];
    d[n - 1] = x[n - 1] - x[n - 2];
    for (int i = 1; i < n - 1; ++i) d[i] = 0.5f * (x[i + 1] - x[i - 1]);
}

static void flit_rescale(const float *x, int n, float r1, float r2, float *y) {
    float lo = x[0], hi = x[0];
    for (int i = 1; i < n; ++i) {
        if (x[i] < lo) lo = x[i];
        if (x[i] > hi) hi = x[i];
    }
    for (int i = 0; i < n; ++i) y[i] = (x[i] - lo) / (hi - lo) * (r2 - r1) + r1;
}

/* segment k with x[k] <= v < x[k+1] of an increasing x, clamped to the ends */
static int seg_of(const float *x, int n, float v) {
    if (v < x[1]) return 0;
    if (v >= x[n - 2]) return n - 2;
    int lo = 1, hi = n - 2; /* x[lo] <= v < x[hi] */
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (x[mid] <= v) lo = mid;
        else hi = mid;
    }
    return lo;
}

static void flit_ginterp(const float *x, const float *y, int n, const float *xi, int m, float *yi, int cubic) {
    if (n == 1) {
        for (int k = 0; k < m; ++k) yi[k] = y[0];
        return;
    }
    if (!cubic || n < 3) {
        for (int k = 0; k < m; ++k) {
            const int s = seg_of(x, n, xi[k]);
            const float t = (xi[k] - x[s]) / (x[s + 1] - x[s]);
            yi[k] = y[s] + (y[s + 1] - y[s]) * t;
        }
        return;
    }
    /* natural spline: second derivatives by the tridiagonal system */
    double *M = (double *)calloc((size_t)n, sizeof(double)), *cp = (double *)calloc((size_t)n, sizeof(double));
    double *dp = (double *)calloc((size_t)n, sizeof(double));
    for (int i = 1; i < n - 1; ++i) {
        const double h0 = (double)x[i] - x[i - 1], h1 = (double)x[i + 1] - x[i];
        const double a = h0, b = 2.0 * (h0 + h1), c = h1;
        const double r = 6.0 * (((double)y[i + 1] - y[i]) / h1 - ((double)y[i] - y[i - 1]) / h0);
        const double den = b - a * cp[i - 1];
        cp[i] = c / den;
        dp[i] = (r - a * dp[i - 1]) / den;
    }
    for (int i = n - 2; i >= 1; --i) M[i] = dp[i] - cp[i] * M[i + 1];
    for (int k = 0; k < m; ++k) {
        const int s = seg_of(x, n, xi[k]);
        const double h = (double)x[s + 1] - x[s];
        const double A = ((double)x[s + 1] - xi[k]) / h, B = ((double)xi[k] - x[s]) / h;
        const double v = A * y[s] + B * y[s + 1] + ((A * A * A - A) * M[s] + (B * B * B - B) * M[s + 1]) * h * h / 6.0;
        yi[k] = (float)v;
    }
    free(M);
    free(cp);
    free(dp);
}

/* the same FLIT restatements for the 2D standard-staggered free surface
 * (owl_oracle_elastic2d.c) */
float *ORA_SYM(flit_regspace)(float a, float d, float b, int *n) { return flit_regspace(a, d, b, n); }
void ORA_SYM(flit_deriv)(const float *x, int n, float *d) { flit_deriv(x, n, d); }
void ORA_SYM(flit_ginterp)(const float *x, const float *y, int n, const float *xi, int m, float *yi, int cubic) {
    flit_ginterp(x, y, n, xi, m, yi, cubic);
}

static void fsmesh_free(fsmesh *S) {
    float **p[11] = {&S->topo_i, &S->slope_i, &S->topo_h, &S->slope_h, &S->eta_i, &S->eta_h,
                     &S->edz_i,  &S->edz_h,   &S->sc_i,   &S->sc_h,    &S->zz_i};
    for (int k = 0; k < 11; ++k) {
        free(*p[k]);
        *p[k] = NULL;
    }
}

static float fsmesh_edz(const fsmesh *S, int j) {
    int jc = j < -S->pml ? -S->pml : j;
    if (jc > S->nz + S->pml + 1) jc = S->nz + S->pml + 1;
    return S->edz_i[jc + S->pml];
}

#define ZZ(S, i, j) ((S)->zz_i[(size_t)((i) + (S)->pml) * (size_t)(S)->nrj + (size_t)((j) + (S)->pml)])

/* prepare_modeling's free-surface block (module_vars.f90:360-519).  c55 and
 * rho: the window (nz0, nx) z fastest; f0min: the shot's smallest f0. */
static int fsmesh_build(fsmesh *S, const ora_fgrid2d *g, const ora_fsurf2d *fs, const float *c55, const float *rho,
                        float f0min) {
    memset(S, 0, sizeof(*S));
    const int nx = g->nx, nz0 = g->nz, pml = g->pml, nci = nx + 2 * pml + 2;
    const float dx = g->dx, dz0 = g->dz, ox = g->ox;
    if (nz0 < 3 || !(fs->refine > 0) || !(fs->dz_max > 0) || !(f0min > 0)) return -2;
    S->nx = nx;
    S->nz0 = nz0;
    S->pml = pml;
    S->nci = nci;
    S->dx = dx;
    S->dz0 = dz0;
    int n1;
    float *topox = flit_regspace(ox - (float)(pml + 1) * dx, dx, ox + (float)(nx + pml) * dx, &n1);
    if (n1 != nci) {
        free(topox);
        return -2;
    }
    S->topo_i = (float *)calloc((size_t)nci, sizeof(float));
    S->slope_i = (float *)calloc((size_t)nci, sizeof(float));
    S->topo_h = (float *)calloc((size_t)nci, sizeof(float));
    S->slope_h = (float *)calloc((size_t)nci, sizeof(float));
    if (fs->ntopo > 0) {
        flit_ginterp(fs->tx, fs->tz, fs->ntopo, topox, nci, S->topo_i, fs->interp);
        flit_deriv(S->topo_i, nci, S->slope_i);
        for (int k = 0; k < nci; ++k) S->slope_i[k] = S->slope_i[k] / dx;
        float *xh = (float *)malloc(sizeof(float) * (size_t)nci);
        for (int k = 0; k < nci; ++k) xh[k] = topox[k] - 0.5f * dx;
        flit_ginterp(topox, S->topo_i, nci, xh, nci, S->topo_h, fs->interp);
        flit_deriv(S->topo_h, nci, S->slope_h);
        for (int k = 0; k < nci; ++k) S->slope_h[k] = S->slope_h[k] / dx;
        free(xh);
    }
    free(topox);
    float tmax = S->topo_i[0];
    for (int k = 1; k < nci; ++k)
        if (S->topo_i[k] > tmax) tmax = S->topo_i[k];
    const float depth_max = (float)(nz0 - 1) * dz0 - tmax;
    float eta_max = depth_max;
    int bad = 0;
    for (int k = 0; k < nci; ++k)
        if ((S->topo_i[k] + depth_max) / eta_max < fmaxf(1.0f, fabsf(S->slope_i[k]))) bad = 1;
    if (bad) {
        float mn = INFINITY;
        for (int k = 0; k < nci; ++k) {
            const float v = (S->topo_i[k] + depth_max) / fmaxf(1.0f, fabsf(S->slope_i[k]));
            if (v < mn) mn = v;
        }
        eta_max = 0.975f * mn;
    }
    const float dz = eta_max / (float)(nz0 - 1);
    /* sqrt(mean(c55(:, 1)/rho(:, 1)))/minval(f0) over the padded first row */
    float sm = 0.0f;
    for (int i = -pml; i <= nx + pml + 1; ++i) {
        const int ic = i < 1 ? 1 : (i > nx ? nx : i);
        const size_t m = (size_t)(ic - 1) * nz0;
        sm = sm + c55[m] / rho[m];
    }
    const float rl = sqrtf(sm / (float)nci) / f0min;
    const int refine_nz = (int)ceilf(0.5f * rl / dz);
    const float rr = fs->refine, cap = fs->dz_max * eta_max / (depth_max + tmax);
    float dzm = fminf(cap, dz / rr);
    if (!(dzm > 0) || refine_nz < 1) return -2;
    int ne;
    float *e0 = flit_regspace(0.0f, dzm, (float)(refine_nz - 1) * dz, &ne);
    if (ne < 1) {
        free(e0);
        return -2;
    }
    int capn = ne + 64;
    float *eta = (float *)malloc(sizeof(float) * (size_t)capn);
    memcpy(eta, e0, sizeof(float) * (size_t)ne);
    free(e0);
    while (eta[ne - 1] < eta_max) {
        dzm = fminf(dzm * 1.025f, cap);
        if (ne == capn) {
            capn *= 2;
            eta = (float *)realloc(eta, sizeof(float) * (size_t)capn);
        }
        eta[ne] = eta[ne - 1] + dzm;
        ++ne;
        if (ne > 1000000) {
            free(eta);
            return -2;
        }
    }
    const int nz = ne, nrj = nz + 2 * pml + 2;
    S->nz = nz;
    S->nrj = nrj;
    S->dz = dz;
    S->eta_max = eta_max;
    S->depth_max = depth_max;
    S->topo_max = tmax;
    float *ei = (float *)calloc((size_t)nrj, sizeof(float)), *eh = (float *)calloc((size_t)nrj, sizeof(float));
    for (int j = 1; j <= nz; ++j) ei[j + pml] = eta[j - 1];
    eh[1 + pml] = -0.5f * dz / rr;
    for (int j = 1; j <= nz - 1; ++j) eh[j + 1 + pml] = 0.5f * (eta[j - 1] + eta[j]);
    for (int i = 1; i <= pml + 1; ++i) {
        ei[1 - i + pml] = ei[1 + pml] - (float)i * dz / rr;
        ei[nz + i + pml] = ei[nz + pml] + (float)i * dzm;
        eh[1 - i + pml] = eh[1 + pml] - (float)i * dz / rr;
        eh[nz + i + pml] = eh[nz + pml] + (float)i * dzm;
    }
    S->edz_i = (float *)calloc((size_t)nrj, sizeof(float));
    S->edz_h = (float *)calloc((size_t)nrj, sizeof(float));
    S->sc_i = (float *)calloc((size_t)nrj, sizeof(float));
    S->sc_h = (float *)calloc((size_t)nrj, sizeof(float));
    flit_deriv(ei, nrj, S->edz_i);
    flit_deriv(eh, nrj, S->edz_h);
    for (int k = 0; k < nrj; ++k) {
        S->sc_i[k] = dz / S->edz_i[k];
        S->sc_h[k] = dz / S->edz_h[k];
    }
    /* physical depths of the integer rows, per column */
    S->zz_i = (float *)calloc((size_t)nci * (size_t)nrj, sizeof(float));
    float *col = (float *)malloc(sizeof(float) * (size_t)nz);
    for (int i = -pml; i <= nx + pml + 1; ++i) {
        flit_rescale(eta, nz, -S->topo_i[i + pml], depth_max, col);
        for (int j = 1; j <= nz; ++j) ZZ(S, i, j) = col[j - 1];
    }
    free(col);
    float dsum = 0.0f;
    for (int i = -pml; i <= nx + pml + 1; ++i) dsum = dsum + (ZZ(S, i, nz) - ZZ(S, i, nz - 1));
    const float dzmean = dsum / (float)nci;
    for (int i = -pml; i <= nx + pml + 1; ++i) {
        for (int k = 1; k <= pml + 1; ++k) ZZ(S, i, nz + k) = ZZ(S, i, nz) + (float)k * dzmean;
        const float d1 = ZZ(S, i, 2) - ZZ(S, i, 1);
        for (int k = 1; k <= pml + 1; ++k) ZZ(S, i, 1 - k) = ZZ(S, i, 1) - (float)k * d1;
    }
    free(eta);
    /* clip(eta, -pml dz, eta_max) */
    const float clo = (float)(-pml) * dz;
    for (int k = 0; k < nrj; ++k) {
        ei[k] = fminf(fmaxf(ei[k], clo), eta_max);
        eh[k] = fminf(fmaxf(eh[k], clo), eta_max);
    }
    S->eta_i = ei;
    S->eta_h = eh;
    return 0;
}

/* map_regular_to_irregular (module_vars.f90:169-203): each padded column
 * of the window, at depths (k - 1) dz - max(topo) for k = -pml .. nz0+pml+1,
 * linearly interpolated at the column's refined depths; written to the
 * padded oracle layout with edge replication beyond the mesh. */
static void fsmesh_medium(const fsmesh *S, const ora_fgrid2d *g, const float *v, real *out, const flay *L) {
    const int nx = S->nx, nz0 = S->nz0, pml = S->pml, nz = S->nz, nr = nz0 + 2 * pml + 2;
    float *zr = (float *)malloc(sizeof(float) * (size_t)nr), *vr = (float *)malloc(sizeof(float) * (size_t)nr);
    float *zq = (float *)malloc(sizeof(float) * (size_t)S->nrj), *vq = (float *)malloc(sizeof(float) * (size_t)S->nrj);
    float *mapped = (float *)malloc(sizeof(float) * (size_t)S->nci * (size_t)S->nrj);
    for (int k = -pml; k <= nz0 + pml + 1; ++k) zr[k + pml] = (float)(k - 1) * g->dz - S->topo_max;
    for (int i = -pml; i <= nx + pml + 1; ++i) {
        const int ic = i < 1 ? 1 : (i > nx ? nx : i);
        for (int k = -pml; k <= nz0 + pml + 1; ++k) {
            const int kc = k < 1 ? 1 : (k > nz0 ? nz0 : k);
            vr[k + pml] = v[(size_t)(kc - 1) + (size_t)(ic - 1) * nz0];
        }
        for (int j = -pml; j <= nz + pml + 1; ++j) zq[j + pml] = ZZ(S, i, j);
        flit_ginterp(zr, vr, nr, zq, S->nrj, vq, 0);
        memcpy(mapped + (size_t)(i + pml) * S->nrj, vq, sizeof(float) * (size_t)S->nrj);
    }
    for (int j = L->lo; j < L->lo + L->NZ; ++j)
        for (int i = L->lo; i < L->lo + L->NX; ++i) {
            const int ic = i < -pml ? -pml : (i > nx + pml + 1 ? nx + pml + 1 : i);
            const int jc = j < -pml ? -pml : (j > nz + pml + 1 ? nz + pml + 1 : j);
            out[IF(*L, i, j)] = (real)mapped[(size_t)(ic + pml) * S->nrj + (size_t)(jc + pml)];
        }
    free(zr); free(vr); free(zq); free(vq); free(mapped);
}

/* update_wavefield_free_surface (module_wavefield.f90:257-587) */
static void fupdate_fs(const fmed *M, const fsmesh *S, ffld *F, real dt) {
    const flay L = M->L;
    const int R = L.R, pml = L.pml, nx = L.nx, nz = L.nz;
    const size_t sz = (size_t)L.NX;
    const float *c = M->c;
    const real dx = M->dx, dz = (real)S->dz;
    const real eta_max = (real)S->eta_max, depth_max = (real)S->depth_max;
    /* particle velocities above the free surface are zero */
    for (int j = 1; j <= R; ++j)
        for (int i = L.lo; i < L.lo + L.NX; ++i) {
            F->vx_hi[IF(L, i, 1 - j)] = 0;
            F->vz_hi[IF(L, i, 1 - j)] = 0;
            F->vx_ih[IF(L, i, 2 - j)] = 0;
            F->vz_ih[IF(L, i, 2 - j)] = 0;
        }
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= nz + pml; ++j)
        for (int i = -pml + 1; i <= nx + pml; ++i) {
            const size_t q = IF(L, i, j);
            const real eta = (real)S->eta_i[j + pml];
            const real ti = (real)S->topo_i[i + pml], si = (real)S->slope_i[i + pml];
            const real tpaz = (eta_max - eta > 0 ? eta_max - eta : (real)0) / (ti + depth_max) * si;
            const real tpcz = eta_max / (ti + depth_max);
            const real c11 = M->c11[q], c13 = M->c13[q], c15 = M->c15[q], c33 = M->c33[q], c35 = M->c35[q], c55 = M->c55[q];
            real pdxvx = dxs(F->vx_hi, q, 0, c, R) / dx, pdxvz = dxs(F->vz_hi, q, 0, c, R) / dx, pdzvx, pdzvz;
            if (j == 1) {
                if (i <= 0 || i >= nx + 1) {
                    MP(F->m_dxvx_ii, q, pdxvx, M->xii.a, M->xii.b);
                    MP(F->m_dxvz_ii, q, pdxvz, M->xii.a, M->xii.b);
                }
                /* z derivatives from sigma . n = 0 */
                const real hx = si;
                const real d11 = tpcz * (hx * hx * c11 + hx * (real)2.0f * c15 + c55);
                const real d12 = tpcz * (hx * hx * c15 + hx * (c13 + c55) + c35);
                const real d22 = tpcz * (hx * hx * c55 + hx * (real)2.0f * c35 + c33);
                const real dd = d12 * d12 - d11 * d22;
                const real f1 = (-(hx * c11) - c15) * pdxvx + (-(hx * c15) - c55) * pdxvz;
                const real f2 = (-(hx * c15) - c13) * pdxvx + (-(hx * c55) - c35) * pdxvz;
                pdzvx = (-(d22 * f1) + d12 * f2) / dd;
                pdzvz = (-(d11 * f2) + d12 * f1) / dd;
            } else {
                pdzvx = dzs(F->vx_ih, q, 0, sz, c, R) / dz * (real)S->sc_i[j + pml];
                pdzvz = dzs(F->vz_ih, q, 0, sz, c, R) / dz * (real)S->sc_i[j + pml];
                if (i <= 0 || i >= nx + 1 || j >= nz + 1) {
                    MP(F->m_dxvx_ii, q, pdxvx, M->xii.a, M->xii.b);
                    MP(F->m_dxvz_ii, q, pdxvz, M->xii.a, M->xii.b);
                    MP(F->m_dzvx_ii, q, pdzvx, M->zii.a, M->zii.b);
                    MP(F->m_dzvz_ii, q, pdzvz, M->zii.a, M->zii.b);
                }
            }
            const real pd1 = pdxvx + pdzvx * tpaz, pd2 = pdzvz * tpcz, pd3 = pdzvx * tpcz + pdxvz + pdzvz * tpaz;
            F->sxx_ii[q] = F->sxx_ii[q] + dt * (c11 * pd1 + c13 * pd2 + c15 * pd3);
            F->szz_ii[q] = F->szz_ii[q] + dt * (c13 * pd1 + c33 * pd2 + c35 * pd3);
            F->sxz_ii[q] = F->sxz_ii[q] + dt * (c15 * pd1 + c35 * pd2 + c55 * pd3);
        }
#pragma omp parallel for schedule(static)
    for (int n = 2; n <= nz + pml; ++n)
        for (int m = -pml + 1; m <= nx + pml; ++m) {
            const size_t q = IF(L, m, n);
            const real eta = (real)S->eta_h[n + pml];
            const real th = (real)S->topo_h[m + pml], sh = (real)S->slope_h[m + pml];
            const real tpaz = (eta_max - eta > 0 ? eta_max - eta : (real)0) / (th + depth_max) * sh;
            const real tpcz = eta_max / (th + depth_max);
            const size_t q00 = q - 1 - sz, q10 = q - sz, q01 = q - 1, q11 = q;
#define AV4(a) ((real)0.25f * ((a)[q00] + (a)[q10] + (a)[q01] + (a)[q11]))
            const real c11 = AV4(M->c11), c13 = AV4(M->c13), c15 = AV4(M->c15), c33 = AV4(M->c33), c35 = AV4(M->c35),
                       c55 = AV4(M->c55);
#undef AV4
            real pdxvx = dxs(F->vx_ih, q, -1, c, R) / dx, pdxvz = dxs(F->vz_ih, q, -1, c, R) / dx;
            real pdzvx = dzs(F->vx_hi, q, -1, sz, c, R) / dz * (real)S->sc_h[n + pml];
            real pdzvz = dzs(F->vz_hi, q, -1, sz, c, R) / dz * (real)S->sc_h[n + pml];
            if (m <= 1 || m >= nx + 1 || n >= nz + 1) {
                MP(F->m_dxvx_hh, q, pdxvx, M->xhh.a, M->xhh.b);
                MP(F->m_dxvz_hh, q, pdxvz, M->xhh.a, M->xhh.b);
                MP(F->m_dzvx_hh, q, pdzvx, M->zhh.a, M->zhh.b);
                MP(F->m_dzvz_hh, q, pdzvz, M->zhh.a, M->zhh.b);
            }
            const real pd1 = pdxvx + pdzvx * tpaz, pd2 = pdzvz * tpcz, pd3 = pdzvx * tpcz + pdxvz + pdzvz * tpaz;
            F->sxx_hh[q] = F->sxx_hh[q] + dt * (c11 * pd1 + c13 * pd2 + c15 * pd3);
            F->szz_hh[q] = F->szz_hh[q] + dt * (c13 * pd1 + c33 * pd2 + c35 * pd3);
            F->sxz_hh[q] = F->sxz_hh[q] + dt * (c15 * pd1 + c35 * pd2 + c55 * pd3);
        }
    /* odd mirror of the stresses about the free surface */
    for (int j = 1; j <= R; ++j)
        for (int i = L.lo; i < L.lo + L.NX; ++i) {
            F->sxx_ii[IF(L, i, 1 - j)] = -F->sxx_ii[IF(L, i, 1 + j)];
            F->szz_ii[IF(L, i, 1 - j)] = -F->szz_ii[IF(L, i, 1 + j)];
            F->sxz_ii[IF(L, i, 1 - j)] = -F->sxz_ii[IF(L, i, 1 + j)];
            F->sxx_hh[IF(L, i, 2 - j)] = -F->sxx_hh[IF(L, i, 1 + j)];
            F->szz_hh[IF(L, i, 2 - j)] = -F->szz_hh[IF(L, i, 1 + j)];
            F->sxz_hh[IF(L, i, 2 - j)] = -F->sxz_hh[IF(L, i, 1 + j)];
        }
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= nz + pml; ++j)
        for (int m = -pml + 1; m <= nx + pml; ++m) {
            const size_t q = IF(L, m, j);
            const real eta = (real)S->eta_i[j + pml];
            const real th = (real)S->topo_h[m + pml], sh = (real)S->slope_h[m + pml];
            const real tpaz = (eta_max - eta > 0 ? eta_max - eta : (real)0) / (th + depth_max) * sh;
            const real tpcz = eta_max / (th + depth_max);
            const real sc = (real)S->sc_i[j + pml];
            real pdxxx = dxs(F->sxx_ii, q, -1, c, R) / dx, pdxxz = dxs(F->sxz_ii, q, -1, c, R) / dx;
            real pdzxx = dzs(F->sxx_hh, q, 0, sz, c, R) / dz * sc;
            real pdzxz = dzs(F->sxz_hh, q, 0, sz, c, R) / dz * sc;
            real pdzzz = dzs(F->szz_hh, q, 0, sz, c, R) / dz * sc;
            if (m <= 1 || m >= nx + 1 || j >= nz + 1) {
                MP(F->m_dxxx_hi, q, pdxxx, M->xhi.a, M->xhi.b);
                MP(F->m_dxxz_hi, q, pdxxz, M->xhi.a, M->xhi.b);
                MP(F->m_dzxx_hi, q, pdzxx, M->zhi.a, M->zhi.b);
                MP(F->m_dzxz_hi, q, pdzxz, M->zhi.a, M->zhi.b);
                MP(F->m_dzzz_hi, q, pdzzz, M->zhi.a, M->zhi.b);
            }
            const real pd1 = pdxxx + tpaz * pdzxx + tpcz * pdzxz, pd2 = pdxxz + tpaz * pdzxz + tpcz * pdzzz;
            const real rh = (real)0.5f * (M->rho[q - 1] + M->rho[q]);
            F->vx_hi[q] = F->vx_hi[q] + dt / rh * pd1;
            F->vz_hi[q] = F->vz_hi[q] + dt / rh * pd2;
        }
#pragma omp parallel for schedule(static)
    for (int n = 2; n <= nz + pml; ++n)
        for (int i = -pml + 1; i <= nx + pml; ++i) {
            const size_t q = IF(L, i, n);
            const real eta = (real)S->eta_h[n + pml];
            const real ti = (real)S->topo_i[i + pml], si = (real)S->slope_i[i + pml];
            const real tpaz = (eta_max - eta > 0 ? eta_max - eta : (real)0) / (ti + depth_max) * si;
            const real tpcz = eta_max / (ti + depth_max);
            const real sc = (real)S->sc_h[n + pml];
            real pdxxx = dxs(F->sxx_hh, q, 0, c, R) / dx, pdxxz = dxs(F->sxz_hh, q, 0, c, R) / dx;
            real pdzxx = dzs(F->sxx_ii, q, -1, sz, c, R) / dz * sc;
            real pdzxz = dzs(F->sxz_ii, q, -1, sz, c, R) / dz * sc;
            real pdzzz = dzs(F->szz_ii, q, -1, sz, c, R) / dz * sc;
            if (i <= 0 || i >= nx + 1 || n >= nz + 1) {
                MP(F->m_dxxx_ih, q, pdxxx, M->xih.a, M->xih.b);
                MP(F->m_dxxz_ih, q, pdxxz, M->xih.a, M->xih.b);
                MP(F->m_dzxx_ih, q, pdzxx, M->zih.a, M->zih.b);
                MP(F->m_dzxz_ih, q, pdzxz, M->zih.a, M->zih.b);
                MP(F->m_dzzz_ih, q, pdzzz, M->zih.a, M->zih.b);
            }
            const real pd1 = pdxxx + tpaz * pdzxx + tpcz * pdzxz, pd2 = pdxxz + tpaz * pdzxz + tpcz * pdzzz;
            const real rh = (real)0.5f * (M->rho[q - sz] + M->rho[q]);
            F->vx_ih[q] = F->vx_ih[q] + dt / rh * pd1;
            F->vz_ih[q] = F->vz_ih[q] + dt / rh * pd2;
        }
}

/* even mirror of the velocities before recording
 * (submodule_forward_modeling.f90:66-75) */
static void fmirror_v(const flay *L, ffld *F) {
    for (int j = 1; j <= L->R; ++j)
        for (int i = L->lo; i < L->lo + L->NX; ++i) {
            F->vx_hi[IF(*L, i, 1 - j)] = F->vx_hi[IF(*L, i, 1 + j)];
            F->vz_hi[IF(*L, i, 1 - j)] = F->vz_hi[IF(*L, i, 1 + j)];
            F->vx_ih[IF(*L, i, 2 - j)] = F->vx_ih[IF(*L, i, 1 + j)];
            F->vz_ih[IF(*L, i, 2 - j)] = F->vz_ih[IF(*L, i, 1 + j)];
        }
}

/* Source and receiver depths on the mesh (module_vars.f90:639-680 and the
 * depth-varying branch of prepare_geometry): the surface elevation under x
 * by topo_interp, depth below the surface mapped to the virtual depth
 * (z - stp)/(depth_max - stp) eta_max, nearest integer and half rows by
 * the first minimum of |z - oz - eta|, Kaiser-sinc weights at the offset
 * over the local row spacing.  A source above the first row moves to
 * eta_dz_i(1) with its amplitude times 1.2. */
static void locate_fs(const fsmesh *S, const ora_fgrid2d *g, const ora_fsurf2d *fs, float x, float z, int is_src,
                      ora_point *P, float *amp_scale) {
    const int pml = S->pml, nci = S->nci;
    int n2;
    float *topox = flit_regspace(0.0f, S->dx, (float)(S->nx + 2 * pml + 1) * S->dx, &n2);
    float *surf = (float *)malloc(sizeof(float) * (size_t)nci);
    for (int k = 0; k < nci; ++k) {
        topox[k] = topox[k] - (float)(pml + 1) * S->dx;
        surf[k] = ZZ(S, k - pml, 1);
    }
    float stp;
    flit_ginterp(topox, surf, nci, &x, 1, &stp, fs->interp);
    free(topox);
    free(surf);
    if (is_src ? fs->src_from_surface : fs->rec_from_surface) z = z + stp;
    float zv = (z - stp) / (-stp + S->depth_max) * S->eta_max;
    *amp_scale = 1.0f;
    if (is_src && zv < S->edz_i[1 + pml]) {
        zv = S->edz_i[1 + pml];
        *amp_scale = 1.2f;
    }
    ORA_SYM(locate_point)(2, x, 0.0f, zv, g->ox, 0.0f, g->oz, g->dx, 1.0f, g->dz, P);
    int gz = -pml, hz = -pml;
    float bi = INFINITY, bh = INFINITY;
    const float zs = zv + (zv == 0.0f ? 1.1920928955078125e-7f : 0.0f);
    for (int j = -pml; j <= S->nz + pml + 1; ++j) {
        const float di = fabsf(zv - g->oz - S->eta_i[j + pml]);
        if (di < bi) {
            bi = di;
            gz = j;
        }
        const float dh = fabsf(zs - g->oz - S->eta_h[j + pml]);
        if (dh < bh) {
            bh = dh;
            hz = j;
        }
    }
    P->gz = gz;
    P->hz = hz;
    ORA_SYM(kaiser_sinc)((zv - g->oz - S->eta_i[gz + pml]) / S->edz_i[gz + pml], P->iz);
    ORA_SYM(kaiser_sinc)((zv - g->oz - S->eta_h[hz + pml]) / S->edz_h[hz + pml], P->hzw);
}

/* add_source with a free surface: taps on rows >= 2 only; an explosion
 * divides by dx dz_s with dz_s = eta_dz_i(gz)(topo_max + depth_max)/eta_max */
static void fadd_source_fs(const fmed *M, const fsmesh *S, ffld *F, const ora_fsource2d *src, const ora_point *pts,
                           const float *ascale, int ns, int t, real dt) {
    const flay L = M->L;
    for (int s = 0; s < ns; ++s) {
        const ora_point *P = &pts[s];
        const int nb = (int)roundf(src[s].t0 / (float)dt) + 1, ne = nb + src[s].nt_stf - 1;
        if (t < nb || t > ne) continue;
        real amp = src[s].stf[t - nb] * (real)(src[s].amp * ascale[s]) * dt;
        if (src[s].mechanism == 1) {
            const float dzs_ = S->edz_i[P->gz + S->pml] * (S->topo_max + S->depth_max) / S->eta_max;
            amp = amp / (M->dx * (real)dzs_);
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int irx = -NKW; irx <= NKW; ++irx) {
                    if (P->gz + irz >= 2) {
                        const size_t q1 = IF(L, P->gx + irx, P->gz + irz);
                        F->sxx_ii[q1] = F->sxx_ii[q1] - amp * (real)P->ix[irx + NKW] * (real)P->iz[irz + NKW];
                        F->szz_ii[q1] = F->szz_ii[q1] - amp * (real)P->ix[irx + NKW] * (real)P->iz[irz + NKW];
                    }
                    if (P->hz + irz >= 2) {
                        const size_t q2 = IF(L, P->hx + irx, P->hz + irz);
                        F->sxx_hh[q2] = F->sxx_hh[q2] - amp * (real)P->hxw[irx + NKW] * (real)P->hzw[irz + NKW];
                        F->szz_hh[q2] = F->szz_hh[q2] - amp * (real)P->hxw[irx + NKW] * (real)P->hzw[irz + NKW];
                    }
                }
            continue;
        }
        amp = amp / M->rho[IF(L, P->gx, P->gz)];
#if defined(ORA_F64)
        const real sp = sin((double)src[s].polar), cp = cos((double)src[s].polar);
#else
        const real sp = sinf(src[s].polar), cp = cosf(src[s].polar);
#endif
        for (int irz = -NKW; irz <= NKW; ++irz)
            for (int irx = -NKW; irx <= NKW; ++irx) {
                if (P->gz + irz >= 2) {
                    const size_t q1 = IF(L, P->hx + irx, P->gz + irz);
                    F->vx_hi[q1] = F->vx_hi[q1] + sp * amp * (real)P->hxw[irx + NKW] * (real)P->iz[irz + NKW];
                    F->vz_hi[q1] = F->vz_hi[q1] + cp * amp * (real)P->hxw[irx + NKW] * (real)P->iz[irz + NKW];
                }
                if (P->hz + irz >= 2) {
                    const size_t q2 = IF(L, P->gx + irx, P->hz + irz);
                    F->vx_ih[q2] = F->vx_ih[q2] + sp * amp * (real)P->ix[irx + NKW] * (real)P->hzw[irz + NKW];
                    F->vz_ih[q2] = F->vz_ih[q2] + cp * amp * (real)P->ix[irx + NKW] * (real)P->hzw[irz + NKW];
                }
            }
    }
}

typedef struct {
    fsmesh S;
    fmed M;
    ora_point *sp, *rp;
    float *ascale;
} fsctx;

static float min_f0(const float *f0, int ns) {
    float m = f0[0];
    for (int s = 1; s < ns; ++s)
        if (f0[s] < m) m = f0[s];
    return m;
}

static int fsctx_init(fsctx *X, const ora_fgrid2d *g, const ora_fsurf2d *fs, const float *const *cij, const float *rho,
                      int ns, const ora_fsource2d *src, const float *f0, float f0max, int nr, const ora_freceiver2d *rec) {
    memset(X, 0, sizeof(*X));
    if (g->R < 2 || g->R > 8 || g->pml < g->R || ns < 1) return -1;
    const int rc = fsmesh_build(&X->S, g, fs, cij[5], rho, min_f0(f0, ns));
    if (rc) {
        fsmesh_free(&X->S);
        return rc;
    }
    fmed_init(&X->M, g, cij, rho, f0max, &X->S);
    X->sp = (ora_point *)calloc((size_t)ns, sizeof(ora_point));
    X->rp = (ora_point *)calloc(nr > 0 ? (size_t)nr : 1, sizeof(ora_point));
    X->ascale = (float *)calloc((size_t)ns, sizeof(float));
    float unused;
    for (int s = 0; s < ns; ++s) locate_fs(&X->S, g, fs, src[s].x, src[s].z, 1, &X->sp[s], &X->ascale[s]);
    for (int r = 0; r < nr; ++r) locate_fs(&X->S, g, fs, rec[r].x, rec[r].z, 0, &X->rp[r], &unused);
    return 0;
}

static void fsctx_free(fsctx *X) {
    fmed_free(&X->M);
    fsmesh_free(&X->S);
    free(X->sp);
    free(X->rp);
    free(X->ascale);
}

static void fforward_fs(const ora_fgrid2d *g, fsctx *X, const ora_fsource2d *src, int ns, int nr,
                        const ora_freceiver2d *rec, real *tr, real *strips, real *fin) {
    const flay L = X->M.L;
    const size_t n = (size_t)L.NX * L.NZ, sl = fstrip_len(&L);
    ffld F;
    ffld_alloc(&F, &L);
    memset(tr, 0, sizeof(real) * 2 * (size_t)nr * g->nt);
    for (int t = 1; t <= g->nt; ++t) {
        if (strips) fstrips(&L, &F, strips + (size_t)(t - 1) * sl, 1);
        if (t == g->nt && fin) {
            real **fw = &F.sxx_ii;
            for (int k = 0; k < NSTATE; ++k) memcpy(fin + (size_t)k * n, fw[k], n * sizeof(real));
        }
        fmirror_v(&L, &F);
        frecord(&L, &F, rec, X->rp, nr, g->nt, t, tr);
        if (t == g->nt) break;
        fupdate_fs(&X->M, &X->S, &F, (real)g->dt);
        fadd_source_fs(&X->M, &X->S, &F, src, X->sp, X->ascale, ns, t, (real)g->dt);
    }
    ffld_free(&F);
}

/* The refined mesh of one window, for checking: row tables over j = -pml ..
 * nz+pml+1 (cap entries available), column tables over i = -pml ..
 * nx+pml+1; scal = dz, eta_max, depth_max, topo_max.  Returns -3 with *nz
 * set when cap is too small. */
int ORA_SYM(fsg2d_mesh)(const ora_fgrid2d *g, const ora_fsurf2d *fs, const float *c55, const float *rho, float f0min,
                        int cap, int *nz, float *scal, float *eta_i, float *eta_h, float *sc_i, float *sc_h,
                        float *edz_i, float *topo_i, float *slope_i, float *topo_h, float *slope_h) {
    fsmesh S;
    const int rc = fsmesh_build(&S, g, fs, c55, rho, f0min);
    if (rc) {
        fsmesh_free(&S);
        return rc;
    }
    *nz = S.nz;
    if (cap < S.nrj) {
        fsmesh_free(&S);
        return -3;
    }
    scal[0] = S.dz;
    scal[1] = S.eta_max;
    scal[2] = S.depth_max;
    scal[3] = S.topo_max;
    memcpy(eta_i, S.eta_i, sizeof(float) * (size_t)S.nrj);
    memcpy(eta_h, S.eta_h, sizeof(float) * (size_t)S.nrj);
    memcpy(sc_i, S.sc_i, sizeof(float) * (size_t)S.nrj);
    memcpy(sc_h, S.sc_h, sizeof(float) * (size_t)S.nrj);
    memcpy(edz_i, S.edz_i, sizeof(float) * (size_t)S.nrj);
    memcpy(topo_i, S.topo_i, sizeof(float) * (size_t)S.nci);
    memcpy(slope_i, S.slope_i, sizeof(float) * (size_t)S.nci);
    memcpy(topo_h, S.topo_h, sizeof(float) * (size_t)S.nci);
    memcpy(slope_h, S.slope_h, sizeof(float) * (size_t)S.nci);
    fsmesh_free(&S);
    return 0;
}

/* Forward traces with OWL's free surface; f0: the sources' f0 (the mesh
 * refinement takes the smallest). */
int ORA_SYM(fsg2d_forward_fs)(const ora_fgrid2d *g, const ora_fsurf2d *fs, const float *const *cij, const float *rho,
                              int ns, const ora_fsource2d *src, const float *f0, float f0max, int nr,
                              const ora_freceiver2d *rec, real *traces) {
    fsctx X;
    const int rc = fsctx_init(&X, g, fs, cij, rho, ns, src, f0, f0max, nr, rec);
    if (rc) return rc;
    fforward_fs(g, &X, src, ns, nr, rec, traces, NULL, NULL);
    fsctx_free(&X);
    return 0;
}

/* map_irregular_to_regular (module_vars.f90:208-249) of one refined-mesh
 * gradient (i = 1..nx, j = 1..nz refined, x fastest) to the window (nz0, nx)
 * z fastest: the column's cubic interpolant at (j - 1) dz0 - max(topo),
 * zero above the free surface. */
static void fs_to_regular(const fsmesh *S, const real *gin, float *out) {
    const int nx = S->nx, nz = S->nz, nz0 = S->nz0;
    float *x = (float *)malloc(sizeof(float) * (size_t)nz), *y = (float *)malloc(sizeof(float) * (size_t)nz);
    float *xq = (float *)malloc(sizeof(float) * (size_t)nz0), *yq = (float *)malloc(sizeof(float) * (size_t)nz0);
    for (int i = 1; i <= nx; ++i) {
        for (int j = 1; j <= nz; ++j) {
            x[j - 1] = ZZ(S, i, j);
            y[j - 1] = (float)gin[(size_t)(i - 1) + (size_t)(j - 1) * nx];
        }
        for (int j = 1; j <= nz0; ++j) xq[j - 1] = (float)(j - 1) * S->dz0 - S->topo_max;
        flit_ginterp(x, y, nz, xq, nz0, yq, 1);
        for (int j = 1; j <= nz0; ++j) {
            const float zz = (float)(j - 1) * S->dz0;
            out[(size_t)(j - 1) + (size_t)(i - 1) * nz0] = zz < S->topo_max - S->topo_i[i + S->pml] ? 0.0f : yq[j - 1];
        }
    }
    free(x); free(y); free(xq); free(yq);
}

/* OWL's FSG gradient with the free surface (as fsg2d_gradient_owl, on the
 * refined mesh; adjoint sources on rows >= 2), mapped to the regular window.
 * grad: 7 float arrays (nz0, nx) z fastest. */
int ORA_SYM(fsg2d_gradient_owl_fs)(const ora_fgrid2d *g, const ora_fsurf2d *fs, const float *const *cij,
                                   const float *rho, int ns, const ora_fsource2d *src, const float *f0, float f0max,
                                   int hnt, int nr, const ora_freceiver2d *rec, const real *adj, int restore_all,
                                   float *const *grad, real *traces_out) {
    fsctx X;
    int rc = fsctx_init(&X, g, fs, cij, rho, ns, src, f0, f0max, nr, rec);
    if (rc) return rc;
    const fmed *M = &X.M;
    const flay L = M->L;
    const size_t n = (size_t)L.NX * L.NZ, sl = fstrip_len(&L);
    const int nt = g->nt;
    real *strips = fz(sl * (size_t)nt), *fin = fz((size_t)NSTATE * n);
    real *tr = traces_out ? traces_out : fz(2 * (size_t)nr * nt);
    fforward_fs(g, &X, src, ns, nr, rec, tr, strips, fin);
    ffld F, A, P;
    ffld_alloc(&F, &L);
    ffld_alloc(&A, &L);
    ffld_alloc(&P, &L);
    const int nx = g->nx, nz = L.nz;
    const size_t ni = (size_t)nx * nz;
    real *gc[7];
    for (int c = 0; c < 7; ++c) gc[c] = fz(ni);
    const real mdt = -(real)g->dt;
    const int cc = g->cc_step_interval > 0 ? g->cc_step_interval : 1;
    const size_t sx = 1, sz = (size_t)L.NX;
    for (int t = nt; t >= hnt; --t) {
        real **fw = &F.sxx_ii, **pw = &P.sxx_ii;
        for (int k = 0; k < NSTATE; ++k) memcpy(pw[k], fw[k], n * sizeof(real));
        fupdate_fs(M, &X.S, &F, mdt);
        if (t == nt) {
            for (int k = 0; k < NSTATE; ++k) {
                if (!restore_all && (fw[k] == F.vz_hi || fw[k] == F.vz_ih)) continue;
                memcpy(fw[k], fin + (size_t)k * n, n * sizeof(real));
            }
            if (!restore_all) memcpy(F.vz_hi, fin + (size_t)7 * n, n * sizeof(real));
        }
        fstrips(&L, &F, strips + (size_t)(t - 1) * sl, 0);
        fupdate_fs(M, &X.S, &A, mdt);
        for (int r = 0; r < nr; ++r) {
            if (rec[r].weight == 0.0f) continue;
            const ora_point *Q = &X.rp[r];
            const float w = rec[r].weight;
            const real ax = adj[(size_t)r * nt + (t - 1)], az = adj[((size_t)nr + r) * nt + (t - 1)];
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int irx = -NKW; irx <= NKW; ++irx) {
                    if (Q->gz + irz < 2) continue;
                    const real a1 = (real)(Q->hxw[irx + NKW] * Q->iz[irz + NKW] * w * 0.5f);
                    const size_t q = IF(L, Q->hx + irx, Q->gz + irz);
                    A.vx_hi[q] = A.vx_hi[q] + ax * a1;
                    A.vz_hi[q] = A.vz_hi[q] + az * a1;
                }
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int irx = -NKW; irx <= NKW; ++irx) {
                    if (Q->hz + irz < 2) continue;
                    const real a2 = (real)(Q->ix[irx + NKW] * Q->hzw[irz + NKW] * w * 0.5f);
                    const size_t q = IF(L, Q->gx + irx, Q->hz + irz);
                    A.vx_ih[q] = A.vx_ih[q] + ax * a2;
                    A.vz_ih[q] = A.vz_ih[q] + az * a2;
                }
        }
        if (t % cc != 0) continue;
#pragma omp parallel for schedule(static)
        for (int j = 1; j <= nz; ++j)
            for (int i = 1; i <= nx; ++i) {
                const size_t q = IF(L, i, j), o = (size_t)(i - 1) + (size_t)(j - 1) * nx;
                const real c11 = M->c11[q], c13 = M->c13[q], c15 = M->c15[q], c33 = M->c33[q], c35 = M->c35[q],
                           c55 = M->c55[q];
                const real t1 = c15 * c15 * c33 - 2 * c13 * c15 * c35 + c11 * (c35 * c35) + c13 * c13 * c55 - c11 * c33 * c55;
                const real t2 = c15 * c15 * c33 - 2 * c13 * c15 * c35 + c13 * c13 * c55 + c11 * (c35 * c35 - c33 * c55);
                const real s11 = (c35 * c35 - c33 * c55) / t1, s13 = (-(c15 * c35) + c13 * c55) / t2;
                const real s15 = (c15 * c33 - c13 * c35) / t1, s33 = (c15 * c15 - c11 * c55) / t1;
                const real s35 = (-(c13 * c15) + c11 * c35) / t2, s55 = (c13 * c13 - c11 * c33) / t1;
#define H4(a) ((a)[q] + (a)[q + sx] + (a)[q + sz] + (a)[q + sx + sz])
                const real txx = (F.sxx_ii[q] + (real)0.25f * H4(F.sxx_hh)) - (P.sxx_ii[q] + (real)0.25f * H4(P.sxx_hh));
                const real tzz = (F.szz_ii[q] + (real)0.25f * H4(F.szz_hh)) - (P.szz_ii[q] + (real)0.25f * H4(P.szz_hh));
                const real txz = (F.sxz_ii[q] + (real)0.25f * H4(F.sxz_hh)) - (P.sxz_ii[q] + (real)0.25f * H4(P.sxz_hh));
                const real rxx = A.sxx_ii[q] + (real)0.25f * H4(A.sxx_hh);
                const real rzz = A.szz_ii[q] + (real)0.25f * H4(A.szz_hh);
                const real rxz = A.sxz_ii[q] + (real)0.25f * H4(A.sxz_hh);
#undef H4
                const real exx = s11 * txx + s13 * tzz + s15 * txz, ezz = s13 * txx + s33 * tzz + s35 * txz,
                           exz = s15 * txx + s35 * tzz + s55 * txz;
                const real fxx = s11 * rxx + s13 * rzz + s15 * rxz, fzz = s13 * rxx + s33 * rzz + s35 * rxz,
                           fxz = s15 * rxx + s35 * rzz + s55 * rxz;
                gc[0][o] = gc[0][o] - (fxx * exx);
                gc[1][o] = gc[1][o] - (fxx * ezz + fzz * exx);
                gc[2][o] = gc[2][o] - (fxx * exz + exx * fxz);
                gc[3][o] = gc[3][o] - (fzz * ezz);
                gc[4][o] = gc[4][o] - (fzz * exz + ezz * fxz);
                gc[5][o] = gc[5][o] - (fxz * exz);
                const real svx = ((real)0.5f * (F.vx_hi[q] + F.vx_hi[q + sx]) + (real)0.5f * (F.vx_ih[q] + F.vx_ih[q + sz])) -
                                 ((real)0.5f * (P.vx_hi[q] + P.vx_hi[q + sx]) + (real)0.5f * (P.vx_ih[q] + P.vx_ih[q + sz]));
                const real rvx = (real)0.5f * (A.vx_hi[q] + A.vx_hi[q + sx]) + (real)0.5f * (A.vx_ih[q] + A.vx_ih[q + sz]);
                const real svz = ((real)0.5f * (F.vz_hi[q] + F.vz_hi[q + sx]) + (real)0.5f * (F.vz_ih[q] + F.vz_ih[q + sz])) -
                                 ((real)0.5f * (P.vz_hi[q] + P.vz_hi[q + sx]) + (real)0.5f * (P.vz_ih[q] + P.vz_ih[q + sz]));
                const real rvz = (real)0.5f * (A.vz_hi[q] + A.vz_hi[q + sx]) + (real)0.5f * (A.vz_ih[q] + A.vz_ih[q + sz]);
                gc[6][o] = gc[6][o] + svx * rvx + svz * rvz;
            }
    }
    for (int c = 0; c < 7; ++c) {
        for (size_t e = 0; e < ni; ++e)
            if (!isfinite(gc[c][e])) gc[c][e] = 0;
        fs_to_regular(&X.S, gc[c], grad[c]);
        free(gc[c]);
    }
    if (!traces_out) free(tr);
    free(strips); free(fin);
    ffld_free(&F); ffld_free(&A); ffld_free(&P);
    fsctx_free(&X);
    return 0;
}
