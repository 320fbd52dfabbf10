/*
 * Oracle (test infrastructure only): OWL's per-shot aperture window.
 *
 * Restates, statement by statement in float32:
 *   - the absolute offset and its validity test of prepare_geometry
 *     (src/lib/module_source_receiver.f90:421-436; FLIT's mean taken as the
 *     float32 sum in index order over the array size);
 *   - set_adaptive_range (src/lib/module_utility.f90:410-481): the two
 *     offset `where`s, then for x, y, z in turn the masked minval/maxval,
 *     the clip to the model extent, floor/ceiling onto the grid and the
 *     `where`s that drop receivers outside the range, then the node ranges
 *     clip(int((beg-o)/d)+1, 1, n).
 * The pins are hand-derived windows (tests/test_window.py).  Nothing in
 * owl_amd/ links this file.
 */
#include <float.h>
#include <math.h>

typedef struct {
    int nx, ny, nz;
    float dx, dy, dz, ox, oy, oz;
    float xmin, xmax, ymin, ymax, zmin, zmax;
    int yn_adpx, yn_adpy, yn_adpz;
    float adp_extrax, adp_extray, adp_extraz;
    float offset_min, offset_max;
} ora_window_par;

static float ora_mean(const float *a, int n) {
    float s = 0.0f;
    for (int i = 0; i < n; ++i) s = s + a[i];
    return s / (float)n;
}

static int ora_clip(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* One `if (yn_adp?) then ... else ... end if` block of set_adaptive_range. */
static void ora_axis(int on, const float *rc, const float *sc, float *w, int nr, int ns, float extra,
                     float lo_lim, float hi_lim, float o, float d, int n, float *beg_out, float *end_out) {
    float beg, end;
    if (on) {
        float mn = FLT_MAX, mx = -FLT_MAX, smn = sc[0], smx = sc[0];
        for (int i = 0; i < nr; ++i)
            if (w[i] != 0.0f) {
                if (rc[i] < mn) mn = rc[i];
                if (rc[i] > mx) mx = rc[i];
            }
        for (int i = 1; i < ns; ++i) {
            if (sc[i] < smn) smn = sc[i];
            if (sc[i] > smx) smx = sc[i];
        }
        float t = (mn < smn ? mn : smn) - extra;
        beg = lo_lim > t ? lo_lim : t;
        t = (mx > smx ? mx : smx) + extra;
        end = hi_lim < t ? hi_lim : t;
        int kf = (int)floorf((beg - o) / d);
        int kc = (int)ceilf((end - o) / d);
        beg = (float)kf * d + o;
        end = (float)kc * d + o;
        for (int i = 0; i < nr; ++i)
            if (rc[i] < beg) w[i] = 0.0f;
        for (int i = 0; i < nr; ++i)
            if (rc[i] > end) w[i] = 0.0f;
    } else {
        beg = o;
        end = o + (float)(n - 1) * d;
    }
    *beg_out = beg;
    *end_out = end;
}

/* rx, ry, rz, w: nr receivers (w updated); sx, sy, sz: ns sources.
 * win[6] = nxbeg, nxend, nybeg, nyend, nzbeg, nzend; rng[6] the ranges. */
int ora_set_adaptive_range(const ora_window_par *p, int ns, const float *sx, const float *sy, const float *sz,
                           int nr, const float *rx, const float *ry, const float *rz, float *w, int *win,
                           float *rng) {
    if (ns < 1) return -1;
    const float mx = ora_mean(sx, ns), my = ora_mean(sy, ns), mz = ora_mean(sz, ns);
    for (int i = 0; i < nr; ++i) {
        const float aoff = sqrtf((rx[i] - mx) * (rx[i] - mx) + (ry[i] - my) * (ry[i] - my) +
                                 (rz[i] - mz) * (rz[i] - mz));
        if (!(aoff >= p->offset_min && aoff <= p->offset_max)) w[i] = 0.0f;
    }
    ora_axis(p->yn_adpx, rx, sx, w, nr, ns, p->adp_extrax, p->xmin, p->xmax, p->ox, p->dx, p->nx, &rng[0], &rng[1]);
    ora_axis(p->yn_adpy, ry, sy, w, nr, ns, p->adp_extray, p->ymin, p->ymax, p->oy, p->dy, p->ny, &rng[2], &rng[3]);
    ora_axis(p->yn_adpz, rz, sz, w, nr, ns, p->adp_extraz, p->zmin, p->zmax, p->oz, p->dz, p->nz, &rng[4], &rng[5]);
    win[0] = ora_clip((int)((rng[0] - p->ox) / p->dx) + 1, 1, p->nx);
    win[1] = ora_clip((int)((rng[1] - p->ox) / p->dx) + 1, 1, p->nx);
    win[2] = ora_clip((int)((rng[2] - p->oy) / p->dy) + 1, 1, p->ny);
    win[3] = ora_clip((int)((rng[3] - p->oy) / p->dy) + 1, 1, p->ny);
    win[4] = ora_clip((int)((rng[4] - p->oz) / p->dz) + 1, 1, p->nz);
    win[5] = ora_clip((int)((rng[5] - p->oz) / p->dz) + 1, 1, p->nz);
    return 0;
}
