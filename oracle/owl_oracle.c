/*
 * owl_oracle.c -- CPU restatement of OWL's acoustic SSG-FD time stepping and
 * its boundary-saving FWI gradient.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle (and the
 * timed CPU baseline).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker.  The
 * product library (owl_amd/) never links, loads or calls anything here.
 *
 * The file is compiled twice: with -DORA_REAL=float (the reference's own
 * precision) and with -DORA_REAL=double (an FP64 build of the same scheme,
 * used to state the FP32 error).  Integer-valued quantities (grid indices,
 * STF length, hnt, nt) are always derived in float32 arithmetic exactly as
 * the Fortran writes them, so both builds share them bit for bit.
 *
 * Each function cites the reference file:line it restates; paths are
 * relative to the OWL source tree.  Operation order follows the Fortran
 * left-to-right association; build with -ffp-contract=off.
 *
 * Parity pinning (what checks each part independently of this file):
 *   fd coefficients ........ literal constants, consistency sum c_m(2m-1)~1
 *   nint / grid indices .... half-way cases vs round-half-away-from-zero
 *   Kaiser-windowed sinc ... numpy sinc * numpy i0 closed form
 *   Ricker STF ............. numpy closed form
 *   CFS-PML coefficients ... closed-form numpy evaluation in float64
 *   forward propagation .... 2D Green's function (Hankel), linearity,
 *                            reciprocity, mirror symmetry
 *   OWL gradient ........... dot-product style C-consistency across probes,
 *                            mirror symmetry, finite-difference sign/scale
 *   exact adjoint .......... dot-product test <J dm, dd> = <dm, J^T dd>
 *                            and central finite differences
 * FLIT internals (const_pi precision, bessel_i0, sinc) are restated from
 * their mathematical definitions; OWL's FLIT version is unpinned.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#ifndef ORA_REAL
#define ORA_REAL float
#endif
typedef ORA_REAL real;

#if defined(ORA_F64)
#define ORA_SYM(name) ora64_##name
#else
#define ORA_SYM(name) ora32_##name
#endif

#define ORA_PI 3.14159265358979323846264338327950288
#define NKW 4

/* ------------------------------------------------------------------------ */
/* FD coefficients: src/lib/module_fdcoef.f90:24-71 (R = fdhalf = 2..8).     */
/* ------------------------------------------------------------------------ */
static const float FDC[9][8] = {
    {0}, {0},
    {0.1129042e+1f, -0.4301412e-1f},
    {0.1185991e+1f, -0.7249965e-1f, 0.6301572e-2f},
    {0.1217990e+1f, -0.9382142e-1f, 0.1507536e-1f, -0.1700324e-2f},
    {0.1236607e+1f, -0.1082265e+0f, 0.2343440e-1f, -0.5033546e-2f, 0.6817483e-3f},
    {0.1247662e+1f, -0.1175538e+0f, 0.2997970e-1f, -0.8719078e-2f, 0.2215897e-2f,
     -0.3462075e-3f},
    {0.1254799e+1f, -0.1238928e+0f, 0.3494371e-1f, -0.1208897e-1f, 0.4132531e-2f,
     -0.1197110e-2f, 0.2122227e-3f},
    {0.1259312e+1f, -0.1280347e+0f, 0.3841945e-1f, -0.1473229e-1f, 0.5924913e-2f,
     -0.2248618e-2f, 0.7179226e-3f, -0.1400855e-3f},
};

int ORA_SYM(fdcoefs)(int R, float *out) {
    if (R < 2 || R > 8) return -1;
    for (int m = 0; m < R; ++m) out[m] = FDC[R][m];
    return 0;
}

/* Fortran nint: round half away from zero (float argument). */
static int fnint(float x) { return (int)roundf(x); }

/* ------------------------------------------------------------------------ */
/* Kaiser-windowed sinc: module_source_receiver.f90:789-828.                  */
/* kaiser_b0 is a double parameter initialised from the default-real literal */
/* 6.31 (:147), i.e. the float 6.31 widened to double.                       */
/* bessel_i0: modified Bessel function I0 by its power series (FLIT).       */
/* ------------------------------------------------------------------------ */
static double bessel_i0(double x) {
    double sum = 1.0, term = 1.0, q = 0.25 * x * x;
    for (int k = 1; k < 500; ++k) {
        term *= q / ((double)k * (double)k);
        sum += term;
        if (term < 1e-18 * sum) break;
    }
    return sum;
}

static double sinc_d(double x) { return x == 0.0 ? 1.0 : sin(x) / x; }

void ORA_SYM(kaiser_sinc)(float x, float *h /* [9] for -4..4 */) {
    const double b0 = (double)6.31f;
    for (int i = 0; i < 2 * NKW + 1; ++i) h[i] = 0.0f;
    if (x == 0.0f) { h[NKW] = 1.0f; return; }
    for (int i = -NKW; i <= NKW; ++i) {
        float t = (float)i - x;
        float hi = (float)sinc_d(ORA_PI * (double)t);
        float r;
        if (t >= -(float)NKW && t <= (float)NKW) {
            float q = t / (float)NKW;
            r = sqrtf(1.0f - q * q);
            r = (float)(bessel_i0(b0 * (double)r) / bessel_i0(b0));
        } else {
            r = 0.0f;
        }
        h[i + NKW] = hi * r;
    }
}

/* ------------------------------------------------------------------------ */
/* Geometry.  prepare_geometry, module_source_receiver.f90:556-785 (uniform  */
/* mesh branch).  One record per point; indices are OWL's 1-based grid      */
/* indices within the shot window.                                          */
/* ------------------------------------------------------------------------ */
typedef struct {
    float x, y, z;
    int gx, gy, gz, hx, hy, hz;
    float ix[9], iy[9], iz[9], hxw[9], hyw[9], hzw[9];
} ora_point;

/* dims = 2 or 3; ox..dz describe the window's first grid node and spacing. */
void ORA_SYM(locate_point)(int dims, float x, float y, float z, float ox, float oy,
                           float oz, float dx, float dy, float dz, ora_point *pt) {
    memset(pt, 0, sizeof(*pt));
    pt->x = x; pt->y = y; pt->z = z;
    pt->gx = fnint((x - ox) / dx) + 1;
    pt->gz = fnint((z - oz) / dz) + 1;
    pt->hx = fnint((x - ox + 0.5f * dx) / dx) + 1;
    pt->hz = fnint((z - oz + 0.5f * dz) / dz) + 1;
    float dis;
    dis = (x - ox) / dx - (float)(pt->gx - 1);
    ORA_SYM(kaiser_sinc)(dis, pt->ix);
    dis = (z - oz) / dz - (float)(pt->gz - 1);
    ORA_SYM(kaiser_sinc)(dis, pt->iz);
    dis = (x - ox + 0.5f * dx) / dx - (float)(pt->hx - 1);
    ORA_SYM(kaiser_sinc)(dis, pt->hxw);
    dis = (z - oz + 0.5f * dz) / dz - (float)(pt->hz - 1);
    ORA_SYM(kaiser_sinc)(dis, pt->hzw);
    if (dims == 3) {
        pt->gy = fnint((y - oy) / dy) + 1;
        pt->hy = fnint((y - oy + 0.5f * dy) / dy) + 1;
        dis = (y - oy) / dy - (float)(pt->gy - 1);
        ORA_SYM(kaiser_sinc)(dis, pt->iy);
        dis = (y - oy + 0.5f * dy) / dy - (float)(pt->hy - 1);
        ORA_SYM(kaiser_sinc)(dis, pt->hyw);
    } else {
        pt->gy = pt->hy = 1;
        pt->iy[NKW] = pt->hyw[NKW] = 1.0f;
    }
}

/* ------------------------------------------------------------------------ */
/* Source time function.  prepare_stf, module_source_receiver.f90:1002-1097, */
/* ricker_wavelet :883-891, ricker_deriv :896-904, gaussian :847-854.       */
/* const_pi (FLIT) is taken as double precision pi.                         */
/* wavelet: 0 ricker, 1 gaussian, 2 gaussian_deriv, 3 ricker_deriv          */
/* Returns nt_stf; writes hnt.  stf must hold nt_stf values (query with     */
/* stf = NULL first).                                                       */
/* ------------------------------------------------------------------------ */
int ORA_SYM(stf_length)(float f0, float dt, float f0factor) {
    return (int)ceilf(2.0f * f0factor / f0 / dt);
}

int ORA_SYM(prepare_stf)(int wavelet, float f0, float dt, float f0factor, real *stf, int *hnt) {
    int n = ORA_SYM(stf_length)(f0, dt, f0factor);
    if (hnt) *hnt = fnint((float)n / 2.0f);
    if (!stf) return n;
    real mx = 0;
    for (int t = 1; t <= n; ++t) {
        real tt = (real)((float)(t - 1) * dt - f0factor / f0);
        real w, v;
        switch (wavelet) {
        case 1: /* gaussian */
            w = (real)((ORA_PI * (double)f0 * (double)tt) * (ORA_PI * (double)f0 * (double)tt));
            v = (real)exp(-(double)w);
#if !defined(ORA_F64)
            v = expf(-w);
#endif
            break;
        case 2: /* gaussian_deriv */
            w = (real)((ORA_PI * (double)f0 * (double)tt) * (ORA_PI * (double)f0 * (double)tt));
#if defined(ORA_F64)
            v = -tt * exp(-w);
#else
            v = -tt * expf(-w);
#endif
            break;
        case 3: /* ricker_deriv */
            w = (real)((ORA_PI * (double)f0 * (double)tt) * (ORA_PI * (double)f0 * (double)tt));
#if defined(ORA_F64)
            v = (((real)2.0 * w - (real)3.0) * exp(-w)) * tt;
#else
            v = ((2.0f * w - 3.0f) * expf(-w)) * tt;
#endif
            break;
        default: /* ricker */
            w = (real)((ORA_PI * (double)f0 * (double)tt) * (ORA_PI * (double)f0 * (double)tt));
#if defined(ORA_F64)
            v = ((real)1 - (real)2.0 * w) * exp(-w);
#else
            v = (1.0f - 2.0f * w) * expf(-w);
#endif
            break;
        }
        stf[t - 1] = v;
        real a = v < 0 ? -v : v;
        if (a > mx) mx = a;
    }
    for (int t = 0; t < n; ++t) stf[t] = stf[t] / mx;
    return n;
}

/* ------------------------------------------------------------------------ */
/* Stress-drop (explosion) wavelet of the elastic solvers: time_integration */
/* = -1 (elastic_vhtiort_2d/module_vars.f90:476, elastic_tti_2d/            */
/* module_vars.f90:689) makes prepare_stf take deriv(stf)/dt                */
/* (module_source_receiver.f90:1111-1124).  FLIT's deriv taken as centred   */
/* inside and one-sided at the ends; if the last sample keeps more than a   */
/* tenth of the peak, the wavelet runs to nt samples with its last value    */
/* (pad_array, edge replication).  out holds max(n, nt) values; returns the */
/* new length.                                                              */
/* ------------------------------------------------------------------------ */
int ORA_SYM(stf_time_derivative)(const real *stf, int n, float dt, int nt, real *out) {
    real peak = 0;
    for (int i = 0; i < n; ++i) {
        real d;
        if (n < 2) d = 0;
        else if (i == 0) d = stf[1] - stf[0];
        else if (i == n - 1) d = stf[n - 1] - stf[n - 2];
        else d = (real)0.5 * (stf[i + 1] - stf[i - 1]);
        out[i] = d / (real)dt;
        real a = out[i] < 0 ? -out[i] : out[i];
        if (a > peak) peak = a;
    }
    if (n > 0 && peak > 0) {
        real last = out[n - 1] < 0 ? -out[n - 1] : out[n - 1];
        if (last / peak > (real)0.1f && nt > n) {
            for (int i = n; i < nt; ++i) out[i] = out[n - 1];
            return nt;
        }
    }
    return n;
}

/* ------------------------------------------------------------------------ */
/* CFL: prepare_modeling, acoustic_iso_2d/module_vars.f90:121-126 and       */
/* acoustic_iso_3d/module_vars.f90:174-181.  Computed in float32.           */
/* ------------------------------------------------------------------------ */
float ORA_SYM(dt_stable)(int R, int dims, float vmax, float dx, float dy, float dz) {
    float temp = 0.0f;
    for (int m = 0; m < R; ++m) temp = temp + fabsf(FDC[R][m]);
    float s = 1.0f / (dx * dx) + ((dims == 3) ? 1.0f / (dy * dy) : 0.0f);
    if (dims == 3) s = 1.0f / (dx * dx) + 1.0f / (dy * dy) + 1.0f / (dz * dz);
    else s = 1.0f / (dx * dx) + 1.0f / (dz * dz);
    return 1.0f / (temp * vmax * sqrtf(s));
}

/* nt = nint(tmax/dt + 1): acoustic_iso_2d/module_vars.f90:134 */
int ORA_SYM(nsteps)(float tmax, float dt) { return fnint(tmax / dt + 1.0f); }

/* ------------------------------------------------------------------------ */
/* CFS-PML damping profile for one axis: damp_coef + compute_cfspml_damping  */
/* _coef, acoustic_iso_2d/module_cfspml.f90:37-133.  Arrays are indexed     */
/* i = -pml+1 .. n+pml (length n+2*pml); entries outside the slabs hold     */
/* a = b = 0, k = 1 (alloc_array zero-initialises; k set to 1 at :88-91).   */
/* ------------------------------------------------------------------------ */
static void damp_coef(float bdvp, float d, float dist, int pml, float dt, float alphamax,
                      float *a, float *b, float *k) {
    const float R0 = 1.0e-5f, kmax = 1.0f;
    float nd = dist / ((float)pml * d);
    float alpha = alphamax * (1.0f - nd);
    float alpha_min = 0.25f * alphamax;
    if (alpha_min > alpha) alpha = alpha_min;
    float damp = bdvp * (2.0f + 1.0f) * logf(1.0f / R0) / (2.0f * (float)pml * d) * (nd * nd);
    float kk = 1.0f + (kmax - 1.0f) * (nd * nd);
    float adt = fabsf(dt);
    *a = -(2.0f * adt * damp / kk / (2.0f + adt * (alpha + damp / kk)));
    *b = (2.0f - adt * (alpha + damp / kk)) / (2.0f + adt * (alpha + damp / kk));
    *k = kk;
}

float ORA_SYM(alphamax)(float f0max) { return (float)((double)(1.0f * f0max) * ORA_PI); }

void ORA_SYM(cfspml_axis)(int n, int pml, float d, float dt, float pmlvp, float f0max,
                          float *ai, float *bi, float *ki, float *ah, float *bh, float *kh) {
    float alphamax = ORA_SYM(alphamax)(f0max);
    int len = n + 2 * pml;
    for (int q = 0; q < len; ++q) {
        ai[q] = bi[q] = ah[q] = bh[q] = 0.0f;
        ki[q] = kh[q] = 1.0f;
    }
    for (int i = -pml + 1; i <= n + pml; ++i) {
        float di = 0.0f, dh = 0.0f;
        if (i <= 1) {
            di = fabsf((float)(i - 1) * d);
            dh = fabsf(((float)(i - 1) - 0.5f) * d);
        } else if (i >= n + 1) {
            di = fabsf((float)(i - n) * d);
            dh = fabsf(((float)(i - n) - 0.5f) * d);
        }
        if (i <= 1 || i >= n + 1) {
            int q = i + pml - 1;
            damp_coef(pmlvp, d, di, pml, dt, alphamax, &ai[q], &bi[q], &ki[q]);
            damp_coef(pmlvp, d, dh, pml, dt, alphamax, &ah[q], &bh[q], &kh[q]);
        }
    }
}

/* ======================================================================== */
/* 2D acoustic solver                                                        */
/* ======================================================================== */

typedef struct {
    int nx, nz;            /* shot window, interior points */
    float dx, dz, ox, oz;  /* window origin (shot_xbeg, shot_zbeg) */
    float dt;              /* time step (after the caller's CFL policy) */
    int nt;                /* number of steps */
    int pml;               /* npml */
    int R;                 /* fdhalf */
    int free_surface;
    int cc_step_interval;
} ora_grid2d;

typedef struct {
    float x, z, amp, t0, polar; /* polar in radians */
    int mechanism;              /* 0 explosion, 1 force */
    int nt_stf;
    const real *stf;            /* nt_stf samples */
} ora_source2d;

typedef struct {
    float x, z, weight;
} ora_receiver2d;

typedef struct {
    int lo, NX, NZ; /* index lo..lo+N-1 on both axes, x fastest */
    int nx, nz, pml, R;
} lay2;

#define I2(L, i, j) ((size_t)((i) - (L).lo) + (size_t)((j) - (L).lo) * (size_t)(L).NX)

static lay2 make_lay2(const ora_grid2d *g) {
    lay2 L;
    L.lo = 1 - g->pml - g->R;
    L.NX = g->nx + 2 * (g->pml + g->R);
    L.NZ = g->nz + 2 * (g->pml + g->R);
    L.nx = g->nx; L.nz = g->nz; L.pml = g->pml; L.R = g->R;
    return L;
}

typedef struct {
    lay2 L;
    real *bk, *rho, *vp;                       /* padded, edge-replicated */
    float *axi, *bxi, *kxi, *axh, *bxh, *kxh;  /* index i+pml-1 */
    float *azi, *bzi, *kzi, *azh, *bzh, *kzh;
    real idx, idz;
    float cf[8];
} med2;

typedef struct {
    real *p, *vx, *vz;
    real *m_dxvx, *m_dzvz, *m_dxp, *m_dzp;
} fld2;

static real *zalloc(size_t n) { return (real *)calloc(n, sizeof(real)); }

static void fld2_alloc(fld2 *f, const lay2 *L) {
    size_t n = (size_t)L->NX * L->NZ;
    f->p = zalloc(n); f->vx = zalloc(n); f->vz = zalloc(n);
    f->m_dxvx = zalloc(n); f->m_dzvz = zalloc(n); f->m_dxp = zalloc(n); f->m_dzp = zalloc(n);
}
static void fld2_free(fld2 *f) {
    free(f->p); free(f->vx); free(f->vz);
    free(f->m_dxvx); free(f->m_dzvz); free(f->m_dxp); free(f->m_dzp);
}

/* prepare_modeling (module_vars.f90:108-115): transpose the (nz,nx) z-fast
 * model to x-fast, bk = rho*vp**2, then pad by edge replication (FLIT
 * pad_array).  PML coefficients: compute_cfspml_damping_coef. */
static void med2_init(med2 *M, const ora_grid2d *g, const float *vp_zfast, const float *rho_zfast,
                      float f0max) {
    lay2 L = make_lay2(g);
    M->L = L;
    size_t n = (size_t)L.NX * L.NZ;
    M->bk = zalloc(n); M->rho = zalloc(n); M->vp = zalloc(n);
    for (int j = L.lo; j < L.lo + L.NZ; ++j)
        for (int i = L.lo; i < L.lo + L.NX; ++i) {
            int ic = i < 1 ? 1 : (i > g->nx ? g->nx : i);
            int jc = j < 1 ? 1 : (j > g->nz ? g->nz : j);
            size_t s = (size_t)(jc - 1) + (size_t)(ic - 1) * (size_t)g->nz;
            real v = (real)vp_zfast[s], r = (real)rho_zfast[s];
            M->vp[I2(L, i, j)] = v;
            M->rho[I2(L, i, j)] = r;
            M->bk[I2(L, i, j)] = r * (v * v);
        }
    M->idx = (real)(1.0 / (double)g->dx);
    M->idz = (real)(1.0 / (double)g->dz);
#if !defined(ORA_F64)
    M->idx = (float)(1.0 / (double)g->dx);
    M->idz = (float)(1.0 / (double)g->dz);
#endif
    /* pmlvp: max of vp on the four edge lines (module_cfspml.f90:93) */
    float pmlvp = 0.0f;
    for (int j = 1; j <= g->nz; ++j) {
        float a = vp_zfast[(size_t)(j - 1)];
        float b = vp_zfast[(size_t)(j - 1) + (size_t)(g->nx - 1) * g->nz];
        if (a > pmlvp) pmlvp = a;
        if (b > pmlvp) pmlvp = b;
    }
    for (int i = 1; i <= g->nx; ++i) {
        float a = vp_zfast[(size_t)(i - 1) * g->nz];
        float b = vp_zfast[(size_t)(g->nz - 1) + (size_t)(i - 1) * g->nz];
        if (a > pmlvp) pmlvp = a;
        if (b > pmlvp) pmlvp = b;
    }
    int lx = g->nx + 2 * g->pml, lz = g->nz + 2 * g->pml;
    float *buf = (float *)calloc((size_t)6 * (lx + lz), sizeof(float));
    M->axi = buf; M->bxi = buf + lx; M->kxi = buf + 2 * lx;
    M->axh = buf + 3 * lx; M->bxh = buf + 4 * lx; M->kxh = buf + 5 * lx;
    float *bz = buf + 6 * lx;
    M->azi = bz; M->bzi = bz + lz; M->kzi = bz + 2 * lz;
    M->azh = bz + 3 * lz; M->bzh = bz + 4 * lz; M->kzh = bz + 5 * lz;
    ORA_SYM(cfspml_axis)(g->nx, g->pml, g->dx, g->dt, pmlvp, f0max, M->axi, M->bxi, M->kxi,
                         M->axh, M->bxh, M->kxh);
    ORA_SYM(cfspml_axis)(g->nz, g->pml, g->dz, g->dt, pmlvp, f0max, M->azi, M->bzi, M->kzi,
                         M->azh, M->bzh, M->kzh);
    for (int m = 0; m < g->R; ++m) M->cf[m] = FDC[g->R][m];
}

static void med2_free(med2 *M) {
    free(M->bk); free(M->rho); free(M->vp); free(M->axi);
}

/* Staggered differences: macro_fd_stencil.f90:20-183.  Forward difference
 * at integer node i uses f(i+m) - f(i-m+1); accumulated left to right. */
static inline real dplus_x(const real *f, const lay2 *L, const float *c, int R, int i, int j) {
    size_t b = I2(*L, i, j);
    real s = (real)c[0] * (f[b + 1] - f[b]);
    for (int m = 2; m <= R; ++m) s = s + (real)c[m - 1] * (f[b + m] - f[b - (m - 1)]);
    return s;
}
static inline real dplus_z(const real *f, const lay2 *L, const float *c, int R, int i, int j) {
    size_t b = I2(*L, i, j), st = (size_t)L->NX;
    real s = (real)c[0] * (f[b + st] - f[b]);
    for (int m = 2; m <= R; ++m) s = s + (real)c[m - 1] * (f[b + m * st] - f[b - (m - 1) * st]);
    return s;
}

/* update_wavefield (module_wavefield.f90:40-107) and its free-surface form
 * update_wavefield_free_surface (:112-187).  Three separate sweeps, in
 * place, OpenMP collapse(2) like the reference. */
static void update2(const med2 *M, fld2 *F, real dt, int free_surface) {
    const lay2 *L = &M->L;
    const int nx = L->nx, nz = L->nz, pml = L->pml, R = L->R;
    const float *c = M->cf;
    const int j0p = free_surface ? 2 : -pml + 1;
    /* p-sweep */
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = j0p; j <= nz + pml; ++j)
        for (int i = -pml + 1; i <= nx + pml; ++i) {
            real pdxvx = M->idx * dplus_x(F->vx, L, c, R, i, j);
            real pdzvz = M->idz * dplus_z(F->vz, L, c, R, i, j);
            size_t q = I2(*L, i, j);
            if (i <= 0 || i >= nx + 1) {
                int a = i + pml - 1;
                F->m_dxvx[q] = (real)M->axi[a] * pdxvx + (real)M->bxi[a] * F->m_dxvx[q];
                pdxvx = (pdxvx + F->m_dxvx[q]) / (real)M->kxi[a];
            }
            int zp = free_surface ? (j >= nz + 1) : (j <= 0 || j >= nz + 1);
            if (zp) {
                int a = j + pml - 1;
                F->m_dzvz[q] = (real)M->azi[a] * pdzvz + (real)M->bzi[a] * F->m_dzvz[q];
                pdzvz = (pdzvz + F->m_dzvz[q]) / (real)M->kzi[a];
            }
            F->p[q] = F->p[q] - dt * M->bk[q] * (pdxvx + pdzvz);
        }
    if (free_surface) {
        /* p(i, j) = -p(i, 2 - j) for j <= 0 (:145-151) */
#pragma omp parallel for collapse(2) schedule(static)
        for (int j = -pml + 1; j <= 0; ++j)
            for (int i = -pml + 1; i <= nx + pml; ++i) F->p[I2(*L, i, j)] = -F->p[I2(*L, i, 2 - j)];
    }
    const int j0v = free_surface ? 1 : -pml + 1;
    /* vx-sweep */
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = j0v; j <= nz + pml; ++j)
        for (int i = -pml; i <= nx + pml - 1; ++i) {
            real pdxp = M->idx * dplus_x(F->p, L, c, R, i, j);
            size_t q = I2(*L, i + 1, j);
            int m = i + 1;
            if (m <= 1 || m >= nx + 1) {
                int a = m + pml - 1;
                F->m_dxp[q] = (real)M->axh[a] * pdxp + (real)M->bxh[a] * F->m_dxp[q];
                pdxp = (pdxp + F->m_dxp[q]) / (real)M->kxh[a];
            }
            real reff = (real)0.5 * (M->rho[I2(*L, i + 1, j)] + M->rho[I2(*L, i, j)]);
            F->vx[q] = F->vx[q] - dt / reff * pdxp;
        }
    /* vz-sweep */
    const int j0z = free_surface ? 1 : -pml;
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = j0z; j <= nz + pml - 1; ++j)
        for (int i = -pml + 1; i <= nx + pml; ++i) {
            real pdzp = M->idz * dplus_z(F->p, L, c, R, i, j);
            size_t q = I2(*L, i, j + 1);
            int m = j + 1;
            int zp = free_surface ? (m >= nz + 1) : (m <= 1 || m >= nz + 1);
            if (zp) {
                int a = m + pml - 1;
                F->m_dzp[q] = (real)M->azh[a] * pdzp + (real)M->bzh[a] * F->m_dzp[q];
                pdzp = (pdzp + F->m_dzp[q]) / (real)M->kzh[a];
            }
            real reff = (real)0.5 * (M->rho[I2(*L, i, j + 1)] + M->rho[I2(*L, i, j)]);
            F->vz[q] = F->vz[q] - dt / reff * pdzp;
        }
}

/* add_source: module_wavefield.f90:192-263 */
static void add_source2(const med2 *M, fld2 *F, const ora_source2d *src, const ora_point *pts,
                        int ns, int t, real dt, int free_surface) {
    const lay2 *L = &M->L;
    for (int k = 0; k < ns; ++k) {
        const ora_source2d *s = &src[k];
        const ora_point *P = &pts[k];
        int nbeg = fnint(s->t0 / (float)dt) + 1;
        int nend = nbeg + s->nt_stf - 1;
        if (t < nbeg || t > nend) continue;
        real amp = s->stf[t - nbeg] * (real)s->amp * dt;
        if (s->mechanism == 1) {
            real sp = (real)sin((double)s->polar), cp = (real)cos((double)s->polar);
#if !defined(ORA_F64)
            sp = sinf(s->polar); cp = cosf(s->polar);
#endif
            real rx = M->rho[I2(*L, P->hx, P->gz)];
            real rz = M->rho[I2(*L, P->gx, P->hz)];
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int irx = -NKW; irx <= NKW; ++irx) {
                    if (!free_surface || P->gz + irz >= 2) {
                        size_t q = I2(*L, P->hx + irx, P->gz + irz);
                        F->vx[q] = F->vx[q] + sp * amp / rx * (real)P->hxw[irx + NKW] * (real)P->iz[irz + NKW];
                    }
                    if (!free_surface || P->hz + irz >= 2) {
                        size_t q = I2(*L, P->gx + irx, P->hz + irz);
                        F->vz[q] = F->vz[q] + cp * amp / rz * (real)P->ix[irx + NKW] * (real)P->hzw[irz + NKW];
                    }
                }
        } else {
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int irx = -NKW; irx <= NKW; ++irx) {
                    if (!free_surface || P->gz + irz >= 2) {
                        size_t q = I2(*L, P->gx + irx, P->gz + irz);
                        F->p[q] = F->p[q] + amp * (real)P->ix[irx + NKW] * (real)P->iz[irz + NKW];
                    }
                }
        }
    }
}

/* Boundary strips: module_boundary_saving.f90:88-153.  Record layout is the
 * Fortran column-major order of the four array sections. */
static size_t strip_len2(const lay2 *L) {
    size_t R = (size_t)L->R, nx = (size_t)L->nx, nz = (size_t)L->nz;
    return 2 * R * (nz + 2 * R) + 2 * (nx + 2 * R) * R;
}

static void strips2(const lay2 *L, fld2 *F, real *rec, int save) {
    const int R = L->R, nx = L->nx, nz = L->nz;
    size_t k = 0;
#define XFER(arr, i, j)                                 \
    do {                                                \
        size_t q_ = I2(*L, i, j);                       \
        if (save) rec[k] = (arr)[q_];                   \
        else (arr)[q_] = rec[k];                        \
        ++k;                                            \
    } while (0)
    for (int j = 1 - R; j <= nz + R; ++j)
        for (int i = 2 - R; i <= 1; ++i) XFER(F->vx, i, j);
    for (int j = 1 - R; j <= nz + R; ++j)
        for (int i = nx + 1; i <= nx + R; ++i) XFER(F->vx, i, j);
    for (int j = 2 - R; j <= 1; ++j)
        for (int i = 1 - R; i <= nx + R; ++i) XFER(F->vz, i, j);
    for (int j = nz + 1; j <= nz + R; ++j)
        for (int i = 1 - R; i <= nx + R; ++i) XFER(F->vz, i, j);
#undef XFER
}

/* Receiver gather: submodule_forward_modeling.f90:58-79 */
static void record2(const lay2 *L, const real *p, const ora_receiver2d *rec, const ora_point *rp,
                    int nr, int t, int nt, real *traces) {
#pragma omp parallel for schedule(static)
    for (int ir = 0; ir < nr; ++ir) {
        if (rec[ir].weight == 0.0f) continue;
        const ora_point *P = &rp[ir];
        real acc = traces[(size_t)ir * nt + (t - 1)];
        for (int irz = -NKW; irz <= NKW; ++irz)
            for (int irx = -NKW; irx <= NKW; ++irx)
                acc = acc + p[I2(*L, P->gx + irx, P->gz + irz)] * (real)P->ix[irx + NKW] *
                                (real)P->iz[irz + NKW] * (real)rec[ir].weight;
        traces[(size_t)ir * nt + (t - 1)] = acc;
    }
}


static float f0max_of(const float *f0s, int ns) {
    float m = 0.0f;
    for (int k = 0; k < ns; ++k) if (f0s[k] > m) m = f0s[k];
    return m;
}

/* Forward modelling: compute_forward, submodule_forward_modeling.f90:31-132.
 * traces: [nr][nt] (receiver-major), zero-initialised by this call.
 * strips (optional): [nt][strip_len] boundary records, saved at the start
 * of each step t (reconstruct = .true.).  final_state (optional): vx, vz, p
 * full padded arrays at the start of step nt (output_final_step_wavefield).
 * snap_t/snap_p (optional): p interior (nz,nx) z-fast at the start of step
 * snap_t (before the update), like the snapshot output at :81-92. */
int ORA_SYM(acoustic2d_forward)(const ora_grid2d *g, const float *vp, const float *rho,
                                int ns, const ora_source2d *src, const float *f0s,
                                int nr, const ora_receiver2d *rec, real *traces,
                                real *strips, real *final_state) {
    if (g->R < 2 || g->R > 8) return -1;
    med2 M;
    med2_init(&M, g, vp, rho, f0max_of(f0s, ns));
    const lay2 *L = &M.L;
    fld2 F;
    fld2_alloc(&F, L);
    ora_point *sp = (ora_point *)calloc(ns > 0 ? ns : 1, sizeof(ora_point));
    ora_point *rp = (ora_point *)calloc(nr > 0 ? nr : 1, sizeof(ora_point));
    for (int k = 0; k < ns; ++k)
        ORA_SYM(locate_point)(2, src[k].x, 0.0f, src[k].z, g->ox, 0.0f, g->oz, g->dx, 1.0f, g->dz, &sp[k]);
    for (int k = 0; k < nr; ++k)
        ORA_SYM(locate_point)(2, rec[k].x, 0.0f, rec[k].z, g->ox, 0.0f, g->oz, g->dx, 1.0f, g->dz, &rp[k]);
    memset(traces, 0, sizeof(real) * (size_t)nr * g->nt);
    size_t sl = strip_len2(L), n = (size_t)L->NX * L->NZ;
    const real dt = (real)g->dt;
    for (int t = 1; t <= g->nt; ++t) {
        if (strips) {
            strips2(L, &F, strips + (size_t)(t - 1) * sl, 1);
            if (t == g->nt && final_state) {
                memcpy(final_state, F.vx, n * sizeof(real));
                memcpy(final_state + n, F.vz, n * sizeof(real));
                memcpy(final_state + 2 * n, F.p, n * sizeof(real));
            }
        }
        record2(L, F.p, rec, rp, nr, t, g->nt, traces);
        update2(&M, &F, dt, g->free_surface);
        add_source2(&M, &F, src, sp, ns, t, dt, g->free_surface);
    }
    free(sp); free(rp);
    fld2_free(&F);
    med2_free(&M);
    return 0;
}

/* energy_precond of the 2D gradient below (submodule_fwi_gradient.f90:
 * 173-184, 223-265); off by default. */
static int g_energy_precond2 = 0;
void ORA_SYM(set_energy_precond2)(int on) { g_energy_precond2 = on != 0; }

/* OWL's gradient: compute_adjoint + compute_gradient,
 * acoustic_iso_2d/submodule_fwi_gradient.f90:31-269 (energy
 * preconditioning as set by set_energy_precond2).  Re-runs the forward with boundary saving, then the reverse loop
 * t = nt .. hnt.  adj: [nr][nt] adjoint source traces (syn - obs).
 * grad_vp, grad_rho: (nz, nx) z-fast outputs = the transposed .grd arrays
 * (:187-195), NaN/Inf mapped to zero (FLIT return_normal).
 * traces_out (optional): the forward traces of the same run. */
int ORA_SYM(acoustic2d_gradient_owl)(const ora_grid2d *g, const float *vp, const float *rho,
                                     int ns, const ora_source2d *src, const float *f0s, int hnt,
                                     int nr, const ora_receiver2d *rec, const real *adj,
                                     real *grad_vp, real *grad_rho, real *traces_out) {
    if (g->R < 2 || g->R > 8 || ns < 1) return -1;
    lay2 L0;
    {
        ora_grid2d gg = *g;
        L0 = make_lay2(&gg);
    }
    size_t n = (size_t)L0.NX * L0.NZ, sl = strip_len2(&L0);
    real *strips = zalloc(sl * (size_t)g->nt);
    real *final_state = zalloc(3 * n);
    real *tr = traces_out ? traces_out : zalloc((size_t)nr * g->nt);
    int rc = ORA_SYM(acoustic2d_forward)(g, vp, rho, ns, src, f0s, nr, rec, tr, strips, final_state);
    if (!traces_out) free(tr);
    if (rc) { free(strips); free(final_state); return rc; }

    med2 M;
    med2_init(&M, g, vp, rho, f0max_of(f0s, ns));
    const lay2 *L = &M.L;
    fld2 F, A;
    fld2_alloc(&F, L);
    fld2_alloc(&A, L);
    real *prev_p = zalloc(n), *prev_vx = zalloc(n), *prev_vz = zalloc(n);
    ora_point *rp = (ora_point *)calloc(nr > 0 ? nr : 1, sizeof(ora_point));
    for (int k = 0; k < nr; ++k)
        ORA_SYM(locate_point)(2, rec[k].x, 0.0f, rec[k].z, g->ox, 0.0f, g->oz, g->dx, 1.0f, g->dz, &rp[k]);
    const int nx = g->nx, nz = g->nz;
    real *gv = zalloc((size_t)nx * nz), *gr = zalloc((size_t)nx * nz);
    const int ep = g_energy_precond2;
    real *esv = zalloc((size_t)nx * nz), *erv = zalloc((size_t)nx * nz);
    real *esa = zalloc((size_t)nx * nz), *era = zalloc((size_t)nx * nz);
    const real mdt = -(real)g->dt;
    const int cc = g->cc_step_interval > 0 ? g->cc_step_interval : 1;
    for (int t = g->nt; t >= hnt; --t) {
        memcpy(prev_p, F.p, n * sizeof(real));
        memcpy(prev_vx, F.vx, n * sizeof(real));
        memcpy(prev_vz, F.vz, n * sizeof(real));
        if (t == g->nt) {
            memcpy(F.vx, final_state, n * sizeof(real));
            memcpy(F.vz, final_state + n, n * sizeof(real));
            memcpy(F.p, final_state + 2 * n, n * sizeof(real));
        }
        strips2(L, &F, strips + (size_t)(t - 1) * sl, 0);
        update2(&M, &F, mdt, g->free_surface);
        update2(&M, &A, mdt, g->free_surface);
        /* adjoint source: :126-147, receivers serialised */
        for (int ir = 0; ir < nr; ++ir) {
            if (rec[ir].weight == 0.0f) continue;
            const ora_point *P = &rp[ir];
            real a = adj[(size_t)ir * g->nt + (t - 1)];
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int irx = -NKW; irx <= NKW; ++irx) {
                    size_t q = I2(*L, P->gx + irx, P->gz + irz);
                    A.p[q] = A.p[q] + a * (real)P->ix[irx + NKW] * (real)P->iz[irz + NKW] * (real)rec[ir].weight;
                }
        }
        if (t % cc == 0) {
#pragma omp parallel for collapse(2) schedule(static)
            for (int j = 1; j <= nz; ++j)
                for (int i = 1; i <= nx; ++i) {
                    size_t q = I2(*L, i, j);
                    size_t o = (size_t)(i - 1) + (size_t)(j - 1) * nx;
                    real v = M.vp[q];
                    real sp_ = F.p[q] - prev_p[q];
                    real rp_ = A.p[q];
                    gv[o] = gv[o] - sp_ * rp_ / (M.rho[q] * (v * v * v));
                    real sx = (F.vx[q] + F.vx[q + 1]) - (prev_vx[q] + prev_vx[q + 1]);
                    real rx = A.vx[q] + A.vx[q + 1];
                    real sz = (F.vz[q] + F.vz[q + L->NX]) - (prev_vz[q] + prev_vz[q + L->NX]);
                    real rz = A.vz[q] + A.vz[q + L->NX];
                    gr[o] = gr[o] + sx * rx + sz * rz;
                    if (ep) {
                        esv[o] = esv[o] + sp_ * sp_ / (M.rho[q] * (v * v * v));
                        erv[o] = erv[o] + rp_ * rp_ / (M.rho[q] * (v * v * v));
                        esa[o] = esa[o] + (sx * sx + sz * sz);
                        era[o] = era[o] + (rx * rx + rz * rz);
                    }
                }
        }
    }
    if (ep) {
        real m[4] = {0, 0, 0, 0};
        real *e[4] = {esv, erv, esa, era};
        for (int k = 0; k < 4; ++k)
            for (size_t o = 0; o < (size_t)nx * nz; ++o)
                if (e[k][o] > m[k]) m[k] = e[k][o];
        for (size_t o = 0; o < (size_t)nx * nz; ++o) {
            for (int k = 0; k < 4; ++k) e[k][o] = e[k][o] + (real)1.0e-3 * m[k];
            gv[o] = gv[o] / sqrt(esv[o] * erv[o]);
            gr[o] = gr[o] / sqrt(esa[o] * era[o]);
        }
    }
    free(esv); free(erv); free(esa); free(era);
    for (int j = 1; j <= nz; ++j)
        for (int i = 1; i <= nx; ++i) {
            real a = gv[(size_t)(i - 1) + (size_t)(j - 1) * nx];
            real b = gr[(size_t)(i - 1) + (size_t)(j - 1) * nx];
            if (!isfinite(a)) a = 0;
            if (!isfinite(b)) b = 0;
            grad_vp[(size_t)(j - 1) + (size_t)(i - 1) * nz] = a;
            grad_rho[(size_t)(j - 1) + (size_t)(i - 1) * nz] = b;
        }
    free(gv); free(gr); free(rp);
    free(prev_p); free(prev_vx); free(prev_vz);
    fld2_free(&F); fld2_free(&A);
    med2_free(&M);
    free(strips); free(final_state);
    return 0;
}

/* L2 waveform misfit and adjoint source: compute_adjsrc_waveform,
 * module_inversion_adjoint_source.f90:371-383, with the trace-discard rule
 * of compute_adjoint_source (:1042-1054, threshold 1e-6).  syn, obs, adj:
 * [nr][nt].  Returns the misfit sum (:1167), accumulated in double. */
double ORA_SYM(adjoint_source_l2)(int nr, int nt, const float *weight, const real *syn,
                                  const real *obs, real *adj, float threshold) {
    double *no = (double *)calloc(nr > 0 ? nr : 1, sizeof(double));
    double *nsy = (double *)calloc(nr > 0 ? nr : 1, sizeof(double));
    double mo = 0, ms = 0;
    for (int i = 0; i < nr; ++i) {
        double a = 0, b = 0;
        for (int t = 0; t < nt; ++t) {
            a += (double)obs[(size_t)i * nt + t] * obs[(size_t)i * nt + t];
            b += (double)syn[(size_t)i * nt + t] * syn[(size_t)i * nt + t];
        }
        no[i] = sqrt(a); nsy[i] = sqrt(b);
        if (no[i] > mo) mo = no[i];
        if (nsy[i] > ms) ms = nsy[i];
    }
    double mis = 0;
    for (int i = 0; i < nr; ++i) {
        int keep = !(weight[i] == 0.0f || no[i] < threshold * mo || nsy[i] < threshold * ms ||
                     no[i] * nsy[i] == 0.0);
        for (int t = 0; t < nt; ++t) {
            size_t q = (size_t)i * nt + t;
            if (keep) {
                real d = obs[q] - syn[q];
                mis += (double)d * (double)d;
                adj[q] = syn[q] - obs[q];
            } else {
                adj[q] = 0;
            }
        }
    }
    free(no); free(nsy);
    return mis;
}

/* ======================================================================== */
/* 3D acoustic solver: acoustic_iso_3d/module_wavefield.f90:40-444 with no  */
/* domain decomposition (nx1 = -pml+1 .. nx2 = nx+pml, etc.).  The separate */
/* min/max memory arrays are folded into one array per derivative (their   */
/* index ranges never overlap).                                             */
/* ======================================================================== */
typedef struct {
    int nx, ny, nz;
    float dx, dy, dz, ox, oy, oz;
    float dt;
    int nt, pml, R;
    int free_surface;
    int cc_step_interval;
} ora_grid3d;

typedef struct {
    float x, y, z, amp, t0, polar, azimuth; /* angles in radians */
    int mechanism;                          /* 0 explosion, 1 force */
    int nt_stf;
    const real *stf;
} ora_source3d;

typedef struct {
    float x, y, z, weight;
} ora_receiver3d;

typedef struct {
    int lo, NX, NY, NZ;
    int nx, ny, nz, pml, R;
} lay3;

#define I3L(L, i, j, k)                                                                \
    ((size_t)((i) - (L).lo) + (size_t)((j) - (L).lo) * (size_t)(L).NX +                \
     (size_t)((k) - (L).lo) * (size_t)(L).NX * (size_t)(L).NY)

typedef struct {
    lay3 L;
    real *bk, *rho, *vp;
    float *prof;                                   /* owns the profile arrays */
    float *axi, *bxi, *kxi, *axh, *bxh, *kxh;
    float *ayi, *byi, *kyi, *ayh, *byh, *kyh;
    float *azi, *bzi, *kzi, *azh, *bzh, *kzh;
    real idx, idy, idz;
    float cf[8];
} med3;

typedef struct {
    real *p, *vx, *vy, *vz;
    real *mvx, *mvy, *mvz, *mpx, *mpy, *mpz;
} fld3;

static size_t lay3_n(const lay3 *L) { return (size_t)L->NX * L->NY * L->NZ; }

static void fld3_alloc(fld3 *f, const lay3 *L) {
    size_t n = lay3_n(L);
    f->p = zalloc(n); f->vx = zalloc(n); f->vy = zalloc(n); f->vz = zalloc(n);
    f->mvx = zalloc(n); f->mvy = zalloc(n); f->mvz = zalloc(n);
    f->mpx = zalloc(n); f->mpy = zalloc(n); f->mpz = zalloc(n);
}
static void fld3_free(fld3 *f) {
    free(f->p); free(f->vx); free(f->vy); free(f->vz);
    free(f->mvx); free(f->mvy); free(f->mvz); free(f->mpx); free(f->mpy); free(f->mpz);
}

/* prepare_modeling, acoustic_iso_3d/module_vars.f90:108-168: the (nz,ny,nx)
 * z-fast model permuted to x-fast (permute(...,321)), bk = rho*vp**2, padded
 * by edge replication; PML profiles from the max vp on the six faces. */
static void med3_init(med3 *M, const ora_grid3d *g, const float *vp, const float *rho, float f0max) {
    lay3 L;
    L.lo = 1 - g->pml - g->R;
    L.NX = g->nx + 2 * (g->pml + g->R);
    L.NY = g->ny + 2 * (g->pml + g->R);
    L.NZ = g->nz + 2 * (g->pml + g->R);
    L.nx = g->nx; L.ny = g->ny; L.nz = g->nz; L.pml = g->pml; L.R = g->R;
    M->L = L;
    const int nx = g->nx, ny = g->ny, nz = g->nz, pml = g->pml;
    size_t n = lay3_n(&L);
    M->bk = zalloc(n); M->rho = zalloc(n); M->vp = zalloc(n);
    for (int k = L.lo; k < L.lo + L.NZ; ++k)
        for (int j = L.lo; j < L.lo + L.NY; ++j)
            for (int i = L.lo; i < L.lo + L.NX; ++i) {
                int ic = i < 1 ? 1 : (i > nx ? nx : i);
                int jc = j < 1 ? 1 : (j > ny ? ny : j);
                int kc = k < 1 ? 1 : (k > nz ? nz : k);
                size_t s = (size_t)(kc - 1) + (size_t)(jc - 1) * nz + (size_t)(ic - 1) * nz * ny;
                real v = (real)vp[s], r = (real)rho[s];
                size_t q = I3L(L, i, j, k);
                M->vp[q] = v;
                M->rho[q] = r;
                M->bk[q] = r * (v * v);
            }
    float pmlvp = 0.0f;
    for (int i = 1; i <= nx; ++i)
        for (int j = 1; j <= ny; ++j)
            for (int k = 1; k <= nz; ++k)
                if (i == 1 || i == nx || j == 1 || j == ny || k == 1 || k == nz) {
                    float v = vp[(size_t)(k - 1) + (size_t)(j - 1) * nz + (size_t)(i - 1) * nz * ny];
                    if (v > pmlvp) pmlvp = v;
                }
    int lx = nx + 2 * pml, ly = ny + 2 * pml, lz = nz + 2 * pml;
    float *ax = (float *)calloc(6 * (size_t)(lx + ly + lz), sizeof(float));
    M->prof = ax;
    M->axi = ax; M->bxi = ax + lx; M->kxi = ax + 2 * lx; M->axh = ax + 3 * lx; M->bxh = ax + 4 * lx; M->kxh = ax + 5 * lx;
    float *ay = ax + 6 * lx;
    M->ayi = ay; M->byi = ay + ly; M->kyi = ay + 2 * ly; M->ayh = ay + 3 * ly; M->byh = ay + 4 * ly; M->kyh = ay + 5 * ly;
    float *az = ay + 6 * ly;
    M->azi = az; M->bzi = az + lz; M->kzi = az + 2 * lz; M->azh = az + 3 * lz; M->bzh = az + 4 * lz; M->kzh = az + 5 * lz;
    ORA_SYM(cfspml_axis)(nx, pml, g->dx, g->dt, pmlvp, f0max, M->axi, M->bxi, M->kxi, M->axh, M->bxh, M->kxh);
    ORA_SYM(cfspml_axis)(ny, pml, g->dy, g->dt, pmlvp, f0max, M->ayi, M->byi, M->kyi, M->ayh, M->byh, M->kyh);
    ORA_SYM(cfspml_axis)(nz, pml, g->dz, g->dt, pmlvp, f0max, M->azi, M->bzi, M->kzi, M->azh, M->bzh, M->kzh);
#if defined(ORA_F64)
    M->idx = 1.0 / (double)g->dx; M->idy = 1.0 / (double)g->dy; M->idz = 1.0 / (double)g->dz;
#else
    M->idx = (float)(1.0 / (double)g->dx); M->idy = (float)(1.0 / (double)g->dy);
    M->idz = (float)(1.0 / (double)g->dz);
#endif
    for (int m = 0; m < g->R; ++m) M->cf[m] = FDC[g->R][m];
}

static void med3_free(med3 *M) {
    free(M->bk); free(M->rho); free(M->vp); free(M->prof);
}

/* update_wavefield (module_wavefield.f90:40-172) and
 * update_wavefield_free_surface (:177-317): p-sweep, then vx, vy, vz sweeps
 * with the new p.  With a free surface p starts at k = 2 and is mirrored,
 * p(:,:,k) = -p(:,:,2-k) for k <= 0; v starts at k = 1 (vz at k = 2) and
 * the top PML is off. */
static void update3(const med3 *M, fld3 *F, real dt, int fs) {
    const lay3 L = M->L;
    const int nx = L.nx, ny = L.ny, nz = L.nz, pml = L.pml, R = L.R;
    const float *c = M->cf;
    const size_t sx = 1, sy = (size_t)L.NX, sz = (size_t)L.NX * L.NY;
    const int k0p = fs ? 2 : -pml + 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = k0p; k <= nz + pml; ++k)
        for (int j = -pml + 1; j <= ny + pml; ++j)
            for (int i = -pml + 1; i <= nx + pml; ++i) {
                size_t b = I3L(L, i, j, k);
                real s1 = (real)c[0] * (F->vx[b + sx] - F->vx[b]);
                real s2 = (real)c[0] * (F->vy[b + sy] - F->vy[b]);
                real s3 = (real)c[0] * (F->vz[b + sz] - F->vz[b]);
                for (int m = 2; m <= R; ++m) {
                    s1 = s1 + (real)c[m - 1] * (F->vx[b + m * sx] - F->vx[b - (m - 1) * sx]);
                    s2 = s2 + (real)c[m - 1] * (F->vy[b + m * sy] - F->vy[b - (m - 1) * sy]);
                    s3 = s3 + (real)c[m - 1] * (F->vz[b + m * sz] - F->vz[b - (m - 1) * sz]);
                }
                real dxv = M->idx * s1, dyv = M->idy * s2, dzv = M->idz * s3;
                if (i <= 0 || i >= nx + 1) {
                    int a = i + pml - 1;
                    F->mvx[b] = (real)M->axi[a] * dxv + (real)M->bxi[a] * F->mvx[b];
                    dxv = (dxv + F->mvx[b]) / (real)M->kxi[a];
                }
                if (j <= 0 || j >= ny + 1) {
                    int a = j + pml - 1;
                    F->mvy[b] = (real)M->ayi[a] * dyv + (real)M->byi[a] * F->mvy[b];
                    dyv = (dyv + F->mvy[b]) / (real)M->kyi[a];
                }
                if (fs ? (k >= nz + 1) : (k <= 0 || k >= nz + 1)) {
                    int a = k + pml - 1;
                    F->mvz[b] = (real)M->azi[a] * dzv + (real)M->bzi[a] * F->mvz[b];
                    dzv = (dzv + F->mvz[b]) / (real)M->kzi[a];
                }
                F->p[b] = F->p[b] - dt * M->bk[b] * (dxv + dyv + dzv);
            }
    if (fs) {
#pragma omp parallel for collapse(2) schedule(static)
        for (int k = -pml + 1; k <= 0; ++k)
            for (int j = -pml + 1; j <= ny + pml; ++j)
                for (int i = -pml + 1; i <= nx + pml; ++i) F->p[I3L(L, i, j, k)] = -F->p[I3L(L, i, j, 2 - k)];
    }
    const int k0v = fs ? 1 : -pml + 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = k0v; k <= nz + pml; ++k)
        for (int j = -pml + 1; j <= ny + pml; ++j)
            for (int i = -pml; i <= nx + pml - 1; ++i) {
                size_t b = I3L(L, i, j, k);
                real s = (real)c[0] * (F->p[b + sx] - F->p[b]);
                for (int m = 2; m <= R; ++m) s = s + (real)c[m - 1] * (F->p[b + m * sx] - F->p[b - (m - 1) * sx]);
                real d = M->idx * s;
                int m1 = i + 1;
                if (m1 <= 1 || m1 >= nx + 1) {
                    int a = m1 + pml - 1;
                    F->mpx[b + sx] = (real)M->axh[a] * d + (real)M->bxh[a] * F->mpx[b + sx];
                    d = (d + F->mpx[b + sx]) / (real)M->kxh[a];
                }
                real reff = (real)0.5 * (M->rho[b + sx] + M->rho[b]);
                F->vx[b + sx] = F->vx[b + sx] - dt / reff * d;
            }
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = k0v; k <= nz + pml; ++k)
        for (int j = -pml; j <= ny + pml - 1; ++j)
            for (int i = -pml + 1; i <= nx + pml; ++i) {
                size_t b = I3L(L, i, j, k);
                real s = (real)c[0] * (F->p[b + sy] - F->p[b]);
                for (int m = 2; m <= R; ++m) s = s + (real)c[m - 1] * (F->p[b + m * sy] - F->p[b - (m - 1) * sy]);
                real d = M->idy * s;
                int m1 = j + 1;
                if (m1 <= 1 || m1 >= ny + 1) {
                    int a = m1 + pml - 1;
                    F->mpy[b + sy] = (real)M->ayh[a] * d + (real)M->byh[a] * F->mpy[b + sy];
                    d = (d + F->mpy[b + sy]) / (real)M->kyh[a];
                }
                real reff = (real)0.5 * (M->rho[b + sy] + M->rho[b]);
                F->vy[b + sy] = F->vy[b + sy] - dt / reff * d;
            }
    const int k0z = fs ? 1 : -pml;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = k0z; k <= nz + pml - 1; ++k)
        for (int j = -pml + 1; j <= ny + pml; ++j)
            for (int i = -pml + 1; i <= nx + pml; ++i) {
                size_t b = I3L(L, i, j, k);
                real s = (real)c[0] * (F->p[b + sz] - F->p[b]);
                for (int m = 2; m <= R; ++m) s = s + (real)c[m - 1] * (F->p[b + m * sz] - F->p[b - (m - 1) * sz]);
                real d = M->idz * s;
                int m1 = k + 1;
                if (fs ? (m1 >= nz + 1) : (m1 <= 1 || m1 >= nz + 1)) {
                    int a = m1 + pml - 1;
                    F->mpz[b + sz] = (real)M->azh[a] * d + (real)M->bzh[a] * F->mpz[b + sz];
                    d = (d + F->mpz[b + sz]) / (real)M->kzh[a];
                }
                real reff = (real)0.5 * (M->rho[b + sz] + M->rho[b]);
                F->vz[b + sz] = F->vz[b + sz] - dt / reff * d;
            }
}

/* add_source, module_wavefield.f90:322-444.  Force: amp / rho(gx,gy,gz),
 * then vx at (hx, gy, gz) scaled by sin(polar) cos(azimuth), vy at
 * (gx, hy, gz) by sin(polar) sin(azimuth), vz at (gx, gy, hz) by
 * cos(polar).  With a free surface, only cells with z index >= 2. */
static void add_source3(const med3 *M, fld3 *F, const ora_source3d *src, const ora_point *pts, int ns, int t,
                        real dt, int fs) {
    const lay3 L = M->L;
    for (int s = 0; s < ns; ++s) {
        int nbeg = fnint(src[s].t0 / (float)dt) + 1, nend = nbeg + src[s].nt_stf - 1;
        if (t < nbeg || t > nend) continue;
        real amp = src[s].stf[t - nbeg] * (real)src[s].amp * dt;
        const ora_point *P = &pts[s];
        if (src[s].mechanism == 1) {
            amp = amp / M->rho[I3L(L, P->gx, P->gy, P->gz)];
#if defined(ORA_F64)
            real sp = sin((double)src[s].polar), cp = cos((double)src[s].polar);
            real sa = sin((double)src[s].azimuth), ca = cos((double)src[s].azimuth);
#else
            real sp = sinf(src[s].polar), cp = cosf(src[s].polar);
            real sa = sinf(src[s].azimuth), ca = cosf(src[s].azimuth);
#endif
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int iry = -NKW; iry <= NKW; ++iry)
                    for (int irx = -NKW; irx <= NKW; ++irx)
                        if (!fs || P->gz + irz >= 2) {
                            size_t q = I3L(L, P->hx + irx, P->gy + iry, P->gz + irz);
                            F->vx[q] = F->vx[q] + sp * ca * amp * (real)P->hxw[irx + NKW] * (real)P->iy[iry + NKW] *
                                                      (real)P->iz[irz + NKW];
                        }
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int iry = -NKW; iry <= NKW; ++iry)
                    for (int irx = -NKW; irx <= NKW; ++irx)
                        if (!fs || P->gz + irz >= 2) {
                            size_t q = I3L(L, P->gx + irx, P->hy + iry, P->gz + irz);
                            F->vy[q] = F->vy[q] + sp * sa * amp * (real)P->ix[irx + NKW] * (real)P->hyw[iry + NKW] *
                                                      (real)P->iz[irz + NKW];
                        }
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int iry = -NKW; iry <= NKW; ++iry)
                    for (int irx = -NKW; irx <= NKW; ++irx)
                        if (!fs || P->hz + irz >= 2) {
                            size_t q = I3L(L, P->gx + irx, P->gy + iry, P->hz + irz);
                            F->vz[q] = F->vz[q] + cp * amp * (real)P->ix[irx + NKW] * (real)P->iy[iry + NKW] *
                                                      (real)P->hzw[irz + NKW];
                        }
        } else {
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int iry = -NKW; iry <= NKW; ++iry)
                    for (int irx = -NKW; irx <= NKW; ++irx)
                        if (!fs || P->gz + irz >= 2) {
                            size_t q = I3L(L, P->gx + irx, P->gy + iry, P->gz + irz);
                            F->p[q] = F->p[q] + amp * (real)P->ix[irx + NKW] * (real)P->iy[iry + NKW] *
                                                    (real)P->iz[irz + NKW];
                        }
        }
    }
}

/* Receiver gather of state t-1 (submodule_forward_modeling.f90:74-110 in 3D). */
static void record3(const lay3 *L, const real *p, const ora_receiver3d *rec, const ora_point *rp, int nr, int t,
                    int nt, real *traces) {
#pragma omp parallel for schedule(static)
    for (int ir = 0; ir < nr; ++ir) {
        if (rec[ir].weight == 0.0f) continue;
        const ora_point *P = &rp[ir];
        real acc = traces[(size_t)ir * nt + (t - 1)];
        for (int irz = -NKW; irz <= NKW; ++irz)
            for (int iry = -NKW; iry <= NKW; ++iry)
                for (int irx = -NKW; irx <= NKW; ++irx)
                    acc = acc + p[I3L(*L, P->gx + irx, P->gy + iry, P->gz + irz)] * (real)P->ix[irx + NKW] *
                                    (real)P->iy[iry + NKW] * (real)P->iz[irz + NKW] * (real)rec[ir].weight;
        traces[(size_t)ir * nt + (t - 1)] = acc;
    }
}

/* Boundary strips, acoustic_iso_3d/module_boundary_saving.f90:83-207: vx on
 * x in [2-R, 1] and [nx+1, nx+R], vy on the y faces, vz on the z faces, each
 * over [1-R, n+R] on the other two axes; record layout = the Fortran order
 * of the six array sections. */
static size_t strip_len3(const lay3 *L) {
    size_t R = (size_t)L->R, ex = L->nx + 2 * R, ey = L->ny + 2 * R, ez = L->nz + 2 * R;
    return 2 * R * (ey * ez + ex * ez + ex * ey);
}

static void strips3(const lay3 *L, fld3 *F, real *rec, int save) {
    const int R = L->R, nx = L->nx, ny = L->ny, nz = L->nz;
    size_t k = 0;
#define XFER3(arr, i, j, kk)                         \
    do {                                             \
        size_t q_ = I3L(*L, i, j, kk);               \
        if (save) rec[k] = (arr)[q_];                \
        else (arr)[q_] = rec[k];                     \
        ++k;                                         \
    } while (0)
    for (int kk = 1 - R; kk <= nz + R; ++kk)
        for (int j = 1 - R; j <= ny + R; ++j)
            for (int i = 2 - R; i <= 1; ++i) XFER3(F->vx, i, j, kk);
    for (int kk = 1 - R; kk <= nz + R; ++kk)
        for (int j = 1 - R; j <= ny + R; ++j)
            for (int i = nx + 1; i <= nx + R; ++i) XFER3(F->vx, i, j, kk);
    for (int kk = 1 - R; kk <= nz + R; ++kk)
        for (int j = 2 - R; j <= 1; ++j)
            for (int i = 1 - R; i <= nx + R; ++i) XFER3(F->vy, i, j, kk);
    for (int kk = 1 - R; kk <= nz + R; ++kk)
        for (int j = ny + 1; j <= ny + R; ++j)
            for (int i = 1 - R; i <= nx + R; ++i) XFER3(F->vy, i, j, kk);
    for (int kk = 2 - R; kk <= 1; ++kk)
        for (int j = 1 - R; j <= ny + R; ++j)
            for (int i = 1 - R; i <= nx + R; ++i) XFER3(F->vz, i, j, kk);
    for (int kk = nz + 1; kk <= nz + R; ++kk)
        for (int j = 1 - R; j <= ny + R; ++j)
            for (int i = 1 - R; i <= nx + R; ++i) XFER3(F->vz, i, j, kk);
#undef XFER3
}

static void locate3_all(const ora_grid3d *g, int ns, const ora_source3d *src, ora_point *sp, int nr,
                        const ora_receiver3d *rec, ora_point *rp) {
    for (int s = 0; s < ns; ++s)
        ORA_SYM(locate_point)(3, src[s].x, src[s].y, src[s].z, g->ox, g->oy, g->oz, g->dx, g->dy, g->dz, &sp[s]);
    for (int r = 0; r < nr; ++r)
        ORA_SYM(locate_point)(3, rec[r].x, rec[r].y, rec[r].z, g->ox, g->oy, g->oz, g->dx, g->dy, g->dz, &rp[r]);
}

/* compute_forward, acoustic_iso_3d/submodule_forward_modeling.f90:31-183.
 * Step order: [save strips(t); at t = nt the final vx, vy, vz, p] ->
 * record -> update -> source.  traces [nr][nt]; strips [nt][strip_len3] and
 * final_state (vx, vy, vz, p padded arrays) optional. */
static int forward3(const ora_grid3d *g, const float *vp, const float *rho, int ns, const ora_source3d *src,
                    float f0max, int nr, const ora_receiver3d *rec, real *traces, real *strips, real *final_state) {
    if (g->R < 2 || g->R > 8) return -1;
    med3 M;
    med3_init(&M, g, vp, rho, f0max);
    fld3 F;
    fld3_alloc(&F, &M.L);
    ora_point *sp = (ora_point *)calloc(ns > 0 ? ns : 1, sizeof(ora_point));
    ora_point *rp = (ora_point *)calloc(nr > 0 ? nr : 1, sizeof(ora_point));
    locate3_all(g, ns, src, sp, nr, rec, rp);
    memset(traces, 0, sizeof(real) * (size_t)nr * g->nt);
    const size_t n = lay3_n(&M.L), sl = strip_len3(&M.L);
    const real dt = (real)g->dt;
    for (int t = 1; t <= g->nt; ++t) {
        if (strips) {
            strips3(&M.L, &F, strips + (size_t)(t - 1) * sl, 1);
            if (t == g->nt && final_state) {
                memcpy(final_state, F.vx, n * sizeof(real));
                memcpy(final_state + n, F.vy, n * sizeof(real));
                memcpy(final_state + 2 * n, F.vz, n * sizeof(real));
                memcpy(final_state + 3 * n, F.p, n * sizeof(real));
            }
        }
        record3(&M.L, F.p, rec, rp, nr, t, g->nt, traces);
        update3(&M, &F, dt, g->free_surface);
        add_source3(&M, &F, src, sp, ns, t, dt, g->free_surface);
    }
    free(sp); free(rp);
    fld3_free(&F);
    med3_free(&M);
    return 0;
}

int ORA_SYM(acoustic3d_forward)(const ora_grid3d *g, const float *vp, const float *rho, int ns,
                                const ora_source3d *src, float f0max, int nr,
                                const ora_receiver3d *rec, real *traces) {
    return forward3(g, vp, rho, ns, src, f0max, nr, rec, traces, NULL, NULL);
}

/* OWL's 3D gradient: compute_adjoint + compute_gradient,
 * acoustic_iso_3d/submodule_fwi_gradient.f90:33-361 (energy preconditioning
 * off).  Forward with boundary saving, then t = nt .. hnt: prev <- state;
 * at t = nt load the final state; inject strips(t); update(-dt) on the
 * forward and the adjoint fields; add adj(t) to pr; image over the interior.
 * grad_vp, grad_rho: (nz, ny, nx) z-fast (the .grd arrays, permute 321),
 * NaN/Inf mapped to zero. */
int ORA_SYM(acoustic3d_gradient_owl)(const ora_grid3d *g, const float *vp, const float *rho, int ns,
                                     const ora_source3d *src, float f0max, int hnt, int nr,
                                     const ora_receiver3d *rec, const real *adj, real *grad_vp, real *grad_rho,
                                     real *traces_out) {
    if (g->R < 2 || g->R > 8 || ns < 1) return -1;
    med3 M;
    med3_init(&M, g, vp, rho, f0max);
    const lay3 L = M.L;
    const size_t n = lay3_n(&L), sl = strip_len3(&L);
    real *strips = zalloc(sl * (size_t)g->nt);
    real *final_state = zalloc(4 * n);
    real *tr = traces_out ? traces_out : zalloc((size_t)nr * g->nt);
    int rc = forward3(g, vp, rho, ns, src, f0max, nr, rec, tr, strips, final_state);
    if (!traces_out) free(tr);
    if (rc) { free(strips); free(final_state); med3_free(&M); return rc; }
    fld3 F, A;
    fld3_alloc(&F, &L);
    fld3_alloc(&A, &L);
    real *pp = zalloc(n), *pvx = zalloc(n), *pvy = zalloc(n), *pvz = zalloc(n);
    ora_point *sp = (ora_point *)calloc((size_t)ns, sizeof(ora_point));
    ora_point *rp = (ora_point *)calloc(nr > 0 ? nr : 1, sizeof(ora_point));
    locate3_all(g, ns, src, sp, nr, rec, rp);
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const size_t ni = (size_t)nx * ny * nz;
    real *gv = zalloc(ni), *gr = zalloc(ni);
    const real mdt = -(real)g->dt;
    const int cc = g->cc_step_interval > 0 ? g->cc_step_interval : 1;
    const size_t sy = (size_t)L.NX, sz = (size_t)L.NX * L.NY;
    for (int t = g->nt; t >= hnt; --t) {
        memcpy(pp, F.p, n * sizeof(real));
        memcpy(pvx, F.vx, n * sizeof(real));
        memcpy(pvy, F.vy, n * sizeof(real));
        memcpy(pvz, F.vz, n * sizeof(real));
        if (t == g->nt) {
            memcpy(F.vx, final_state, n * sizeof(real));
            memcpy(F.vy, final_state + n, n * sizeof(real));
            memcpy(F.vz, final_state + 2 * n, n * sizeof(real));
            memcpy(F.p, final_state + 3 * n, n * sizeof(real));
        }
        strips3(&L, &F, strips + (size_t)(t - 1) * sl, 0);
        update3(&M, &F, mdt, g->free_surface);
        update3(&M, &A, mdt, g->free_surface);
        for (int ir = 0; ir < nr; ++ir) {
            if (rec[ir].weight == 0.0f) continue;
            const ora_point *P = &rp[ir];
            real a = adj[(size_t)ir * g->nt + (t - 1)];
            for (int irz = -NKW; irz <= NKW; ++irz)
                for (int iry = -NKW; iry <= NKW; ++iry)
                    for (int irx = -NKW; irx <= NKW; ++irx) {
                        size_t q = I3L(L, P->gx + irx, P->gy + iry, P->gz + irz);
                        A.p[q] = A.p[q] + a * (real)P->ix[irx + NKW] * (real)P->iy[iry + NKW] *
                                              (real)P->iz[irz + NKW] * (real)rec[ir].weight;
                    }
        }
        if (t % cc == 0) {
#pragma omp parallel for collapse(2) schedule(static)
            for (int k = 1; k <= nz; ++k)
                for (int j = 1; j <= ny; ++j)
                    for (int i = 1; i <= nx; ++i) {
                        size_t q = I3L(L, i, j, k);
                        size_t o = (size_t)(i - 1) + (size_t)(j - 1) * nx + (size_t)(k - 1) * nx * ny;
                        real v = M.vp[q];
                        real spp = F.p[q] - pp[q];
                        gv[o] = gv[o] - spp * A.p[q] / (M.rho[q] * (v * v * v));
                        real sxv = (F.vx[q] + F.vx[q + 1]) - (pvx[q] + pvx[q + 1]);
                        real rxv = A.vx[q] + A.vx[q + 1];
                        real syv = (F.vy[q] + F.vy[q + sy]) - (pvy[q] + pvy[q + sy]);
                        real ryv = A.vy[q] + A.vy[q + sy];
                        real szv = (F.vz[q] + F.vz[q + sz]) - (pvz[q] + pvz[q + sz]);
                        real rzv = A.vz[q] + A.vz[q + sz];
                        gr[o] = gr[o] + sxv * rxv + syv * ryv + szv * rzv;
                    }
        }
    }
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                size_t o = (size_t)(i - 1) + (size_t)(j - 1) * nx + (size_t)(k - 1) * nx * ny;
                size_t d = (size_t)(k - 1) + (size_t)(j - 1) * nz + (size_t)(i - 1) * nz * ny;
                real a = gv[o], b = gr[o];
                grad_vp[d] = isfinite(a) ? a : 0;
                grad_rho[d] = isfinite(b) ? b : 0;
            }
    free(gv); free(gr); free(sp); free(rp);
    free(pp); free(pvx); free(pvy); free(pvz);
    fld3_free(&F); fld3_free(&A);
    med3_free(&M);
    free(strips); free(final_state);
    return 0;
}

/* ======================================================================== */
/* Exact discrete adjoint of the 2D forward scheme (no OWL counterpart: OWL */
/* reconstructs the forward field backwards; this is the transpose of the  */
/* forward step of acoustic_iso_2d/module_wavefield.f90:40-107 including   */
/* the CFS-PML memory variables).                                           */
/*                                                                          */
/* Forward: u_t = V(P(u_{t-1})) + s_t for t = 1..nt-1, d_t = R u_{t-1}.     */
/* Misfit phi = 1/2 sum_t |d_t - obs_t|^2; returns d phi / d(vp, rho) on   */
/* the (nz, nx) model, with the edge-replicated padding folded back onto   */
/* the edge nodes.  PML profiles are treated as constants (they depend on  */
/* the model only through a max over the edge lines).  Explosion sources.  */
/* misfit_out receives sum_t |d_t - obs_t|^2 (OWL's convention, no 1/2).    */
/* ======================================================================== */
int ORA_SYM(acoustic2d_gradient_exact)(const ora_grid2d *g, const float *vp, const float *rho,
                                       int ns, const ora_source2d *src, const float *f0s,
                                       int nr, const ora_receiver2d *rec, const real *obs,
                                       real *grad_vp, real *grad_rho, double *misfit_out,
                                       real *traces_out) {
    if (g->R < 2 || g->R > 8 || g->free_surface) return -1;
    for (int k = 0; k < ns; ++k) if (src[k].mechanism != 0) return -2;
    med2 M;
    med2_init(&M, g, vp, rho, f0max_of(f0s, ns));
    const lay2 *L = &M.L;
    const int nx = g->nx, nz = g->nz, pml = g->pml, R = g->R, nt = g->nt;
    const size_t n = (size_t)L->NX * L->NZ;
    const real dt = (real)g->dt;
    const float *c = M.cf;
    ora_point *sp = (ora_point *)calloc(ns > 0 ? ns : 1, sizeof(ora_point));
    ora_point *rp = (ora_point *)calloc(nr > 0 ? nr : 1, sizeof(ora_point));
    for (int k = 0; k < ns; ++k)
        ORA_SYM(locate_point)(2, src[k].x, 0.0f, src[k].z, g->ox, 0.0f, g->oz, g->dx, 1.0f, g->dz, &sp[k]);
    for (int k = 0; k < nr; ++k)
        ORA_SYM(locate_point)(2, rec[k].x, 0.0f, rec[k].z, g->ox, 0.0f, g->oz, g->dx, 1.0f, g->dz, &rp[k]);

    /* forward, keeping every state: P[t], PP[t] (p before the source), VX[t], VZ[t] */
    real *P = zalloc(n * nt), *PP = zalloc(n * nt), *VX = zalloc(n * nt), *VZ = zalloc(n * nt);
    real *tr = zalloc((size_t)nr * nt);
    fld2 F;
    fld2_alloc(&F, L);
    for (int t = 1; t <= nt; ++t) {
        record2(L, F.p, rec, rp, nr, t, nt, tr);
        if (t == nt) break;
        update2(&M, &F, dt, 0);
        memcpy(PP + n * t, F.p, n * sizeof(real));
        add_source2(&M, &F, src, sp, ns, t, dt, 0);
        memcpy(P + n * t, F.p, n * sizeof(real));
        memcpy(VX + n * t, F.vx, n * sizeof(real));
        memcpy(VZ + n * t, F.vz, n * sizeof(real));
    }
    if (traces_out) memcpy(traces_out, tr, sizeof(real) * (size_t)nr * nt);
    /* residual r_t = d_t - obs_t */
    double mis = 0.0;
    real *res = zalloc((size_t)nr * nt);
    for (int ir = 0; ir < nr; ++ir) {
        if (rec[ir].weight == 0.0f) continue;
        for (int t = 0; t < nt; ++t) {
            size_t q = (size_t)ir * nt + t;
            res[q] = tr[q] - obs[q];
            mis += (double)res[q] * (double)res[q];
        }
    }
    if (misfit_out) *misfit_out = mis;

    /* adjoint state */
    real *ph = zalloc(n), *vxh = zalloc(n), *vzh = zalloc(n);
    real *psx = zalloc(n), *psz = zalloc(n), *phx = zalloc(n), *phz = zalloc(n);
    real *ex = zalloc(n), *ez = zalloc(n), *dxa = zalloc(n), *dza = zalloc(n);
    double *gK = (double *)calloc(n, sizeof(double)), *gR = (double *)calloc(n, sizeof(double));
    real *gxa = zalloc(n), *gza = zalloc(n);
#define INJECT_RT(tt)                                                                         \
    for (int ir = 0; ir < nr; ++ir) {                                                         \
        if (rec[ir].weight == 0.0f) continue;                                                 \
        const ora_point *Pt = &rp[ir];                                                        \
        real a_ = res[(size_t)ir * nt + ((tt) - 1)];                                          \
        for (int irz = -NKW; irz <= NKW; ++irz)                                               \
            for (int irx = -NKW; irx <= NKW; ++irx) {                                         \
                size_t q_ = I2(*L, Pt->gx + irx, Pt->gz + irz);                               \
                ph[q_] += a_ * (real)Pt->ix[irx + NKW] * (real)Pt->iz[irz + NKW] * (real)rec[ir].weight; \
            }                                                                                 \
    }
    INJECT_RT(nt);
    for (int t = nt - 1; t >= 1; --t) {
        const real *pc = PP + n * t, *pv = P + n * (t - 1);
        const real *vxc = VX + n * t, *vxp = VX + n * (t - 1), *vzc = VZ + n * t, *vzp = VZ + n * (t - 1);
        /* (a) transpose of the v-update: e = -(dt/rho_eff) v^ (+ PML), phi^ update.
         * The density terms of a cell go to gx/gz first and are summed into
         * gR by the pass after, so that rows can run in parallel. */
#pragma omp parallel for schedule(static)
        for (int j = -pml + 1; j <= nz + pml; ++j)
            for (int i = -pml + 1; i <= nx + pml; ++i) {
                size_t q = I2(*L, i, j);
                real rx = (real)0.5 * (M.rho[q] + M.rho[q - 1]);
                real r = -(dt / rx) * vxh[q];
                if (i <= 1 || i >= nx + 1) {
                    int a = i + pml - 1;
                    real s = phx[q] + r;
                    ex[q] = r + (real)M.axh[a] * s;
                    phx[q] = (real)M.bxh[a] * s;
                } else ex[q] = r;
                /* rho gradient from vx: (vx_{t-1} - vx_t)/rho_eff * v^x */
                gxa[q] = (vxp[q] - vxc[q]) / rx * vxh[q];
                real rz = (real)0.5 * (M.rho[q] + M.rho[q - L->NX]);
                r = -(dt / rz) * vzh[q];
                if (j <= 1 || j >= nz + 1) {
                    int a = j + pml - 1;
                    real s = phz[q] + r;
                    ez[q] = r + (real)M.azh[a] * s;
                    phz[q] = (real)M.bzh[a] * s;
                } else ez[q] = r;
                gza[q] = (vzp[q] - vzc[q]) / rz * vzh[q];
            }
        /* rho gradient: (v_{t-1} - v_t)/rho_eff * v^ of the faces at c and
         * c+1 (x), c and c+row (z) of each cell; gx, gz are zero outside
         * the update domain */
#pragma omp parallel for schedule(static)
        for (size_t q = 0; q < n; ++q) {
            double a = 0.5 * (double)gxa[q] + 0.5 * (double)gza[q];
            if (q + 1 < n) a += 0.5 * (double)gxa[q + 1];
            if (q + (size_t)L->NX < n) a += 0.5 * (double)gza[q + L->NX];
            gR[q] += a;
        }
#pragma omp parallel for schedule(static)
        for (int j = -pml + 1; j <= nz + pml; ++j)
            for (int i = -pml + 1; i <= nx + pml; ++i) {
                /* p^ += -D+^T... = -(idx * sum c (e(i+m) - e(i+1-m)) + z) */
                size_t q = I2(*L, i, j);
                real sx = 0, sz = 0;
                for (int m = 1; m <= R; ++m) {
                    int ia = i + m, ib = i + 1 - m, ja = j + m, jb = j + 1 - m;
                    real xa = (ia <= nx + pml) ? ex[I2(*L, ia, j)] : 0, xb = (ib >= -pml + 1) ? ex[I2(*L, ib, j)] : 0;
                    real za = (ja <= nz + pml) ? ez[I2(*L, i, ja)] : 0, zb = (jb >= -pml + 1) ? ez[I2(*L, i, jb)] : 0;
                    sx += (real)c[m - 1] * (xa - xb);
                    sz += (real)c[m - 1] * (za - zb);
                }
                ph[q] = ph[q] - (M.idx * sx + M.idz * sz);
            }
        /* (b) transpose of the p-update: K gradient, d = -dt K p^ (+ PML), v^ update */
#pragma omp parallel for schedule(static)
        for (int j = -pml + 1; j <= nz + pml; ++j)
            for (int i = -pml + 1; i <= nx + pml; ++i) {
                size_t q = I2(*L, i, j);
                gK[q] += (double)((pc[q] - pv[q]) / M.bk[q] * ph[q]);
                real qv = -dt * M.bk[q] * ph[q];
                if (i <= 0 || i >= nx + 1) {
                    int a = i + pml - 1;
                    real s = psx[q] + qv;
                    dxa[q] = qv + (real)M.axi[a] * s;
                    psx[q] = (real)M.bxi[a] * s;
                } else dxa[q] = qv;
                if (j <= 0 || j >= nz + 1) {
                    int a = j + pml - 1;
                    real s = psz[q] + qv;
                    dza[q] = qv + (real)M.azi[a] * s;
                    psz[q] = (real)M.bzi[a] * s;
                } else dza[q] = qv;
            }
#pragma omp parallel for schedule(static)
        for (int j = -pml + 1; j <= nz + pml; ++j)
            for (int i = -pml + 1; i <= nx + pml; ++i) {
                size_t q = I2(*L, i, j);
                real sx = 0, sz = 0;
                for (int m = 1; m <= R; ++m) {
                    int ia = i - 1 + m, ib = i - m, ja = j - 1 + m, jb = j - m;
                    real xa = (ia <= nx + pml) ? dxa[I2(*L, ia, j)] : 0, xb = (ib >= -pml + 1) ? dxa[I2(*L, ib, j)] : 0;
                    real za = (ja <= nz + pml) ? dza[I2(*L, i, ja)] : 0, zb = (jb >= -pml + 1) ? dza[I2(*L, i, jb)] : 0;
                    sx += (real)c[m - 1] * (xa - xb);
                    sz += (real)c[m - 1] * (za - zb);
                }
                vxh[q] = vxh[q] - M.idx * sx;
                vzh[q] = vzh[q] - M.idz * sz;
            }
        /* (c) receivers of d_t = R u_{t-1} */
        INJECT_RT(t);
    }
#undef INJECT_RT
    /* fold the edge-replicated padding back onto the window and change
     * variables (K = rho vp^2): d/dvp = 2 rho vp dK, d/drho = drho + vp^2 dK */
    double *fK = (double *)calloc((size_t)nx * nz, sizeof(double));
    double *fR = (double *)calloc((size_t)nx * nz, sizeof(double));
    for (int j = L->lo; j < L->lo + L->NZ; ++j)
        for (int i = L->lo; i < L->lo + L->NX; ++i) {
            int ic = i < 1 ? 1 : (i > nx ? nx : i), jc = j < 1 ? 1 : (j > nz ? nz : j);
            size_t o = (size_t)(ic - 1) + (size_t)(jc - 1) * nx;
            fK[o] += gK[I2(*L, i, j)];
            fR[o] += gR[I2(*L, i, j)];
        }
    for (int j = 1; j <= nz; ++j)
        for (int i = 1; i <= nx; ++i) {
            size_t o = (size_t)(i - 1) + (size_t)(j - 1) * nx;
            double v = vp[(size_t)(j - 1) + (size_t)(i - 1) * nz], r = rho[(size_t)(j - 1) + (size_t)(i - 1) * nz];
            grad_vp[(size_t)(j - 1) + (size_t)(i - 1) * nz] = (real)(2.0 * r * v * fK[o]);
            grad_rho[(size_t)(j - 1) + (size_t)(i - 1) * nz] = (real)(fR[o] + v * v * fK[o]);
        }
    free(fK); free(fR); free(gK); free(gR); free(gxa); free(gza);
    free(ph); free(vxh); free(vzh); free(psx); free(psz); free(phx); free(phz);
    free(ex); free(ez); free(dxa); free(dza);
    free(P); free(PP); free(VX); free(VZ); free(tr); free(res);
    free(sp); free(rp);
    fld2_free(&F);
    med2_free(&M);
    return 0;
}
