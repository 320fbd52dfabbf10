// OWL's free-surface mesh for the 2D FSG solver (see fs_mesh.hpp).
#include "fs_mesh.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>

#include "common.hpp"
#include "host_prep.hpp"

namespace owl {
namespace fsm {
namespace {

std::vector<float> regspace(float a, float d, float b) {
    const int n = int(std::lround((b - a) / d)) + 1;
    std::vector<float> x(size_t(std::max(n, 0)));
    for (int k = 0; k < n; ++k) x[size_t(k)] = a + float(k) * d;
    return x;
}

std::vector<float> deriv(const std::vector<float>& x) {
    const size_t n = x.size();
    std::vector<float> d(n, 0.0f);
    if (n < 2) return d;
    d[0] = x[1] - x[0];
    d[n - 1] = x[n - 1] - x[n - 2];
    for (size_t i = 1; i + 1 < n; ++i) d[i] = 0.5f * (x[i + 1] - x[i - 1]);
    return d;
}

// segment k of an increasing x with x[k] <= v < x[k+1], clamped to the ends
int segment(const float* x, int n, float v) {
    if (v < x[1]) return 0;
    if (v >= x[n - 2]) return n - 2;
    int lo = 1, hi = n - 2;
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (x[mid] <= v)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

void rescale(const float* x, int n, float r1, float r2, float* y) {
    float lo = x[0], hi = x[0];
    for (int i = 1; i < n; ++i) {
        lo = std::min(lo, x[i]);
        hi = std::max(hi, x[i]);
    }
    for (int i = 0; i < n; ++i) y[i] = (x[i] - lo) / (hi - lo) * (r2 - r1) + r1;
}

}  // namespace

void ginterp(const float* x, const float* y, int n, const float* xi, int m, float* yi, bool cubic) {
    if (n == 1) {
        std::fill(yi, yi + m, y[0]);
        return;
    }
    if (!cubic || n < 3) {
        for (int k = 0; k < m; ++k) {
            const int s = segment(x, n, xi[k]);
            const float t = (xi[k] - x[s]) / (x[s + 1] - x[s]);
            yi[k] = y[s] + (y[s + 1] - y[s]) * t;
        }
        return;
    }
    // natural spline: second derivatives M from the tridiagonal system
    // (Thomas sweep), M_0 = M_{n-1} = 0
    std::vector<double> M(size_t(n), 0.0), cp(size_t(n), 0.0), dp(size_t(n), 0.0);
    for (int i = 1; i < n - 1; ++i) {
        const double h0 = double(x[i]) - x[i - 1], h1 = double(x[i + 1]) - x[i];
        const double a = h0, bb = 2.0 * (h0 + h1), c = h1;
        const double r = 6.0 * ((double(y[i + 1]) - y[i]) / h1 - (double(y[i]) - y[i - 1]) / h0);
        const double den = bb - a * cp[size_t(i - 1)];
        cp[size_t(i)] = c / den;
        dp[size_t(i)] = (r - a * dp[size_t(i - 1)]) / den;
    }
    for (int i = n - 2; i >= 1; --i) M[size_t(i)] = dp[size_t(i)] - cp[size_t(i)] * M[size_t(i + 1)];
    for (int k = 0; k < m; ++k) {
        const int s = segment(x, n, xi[k]);
        const double h = double(x[s + 1]) - x[s];
        const double A = (double(x[s + 1]) - xi[k]) / h, B = (double(xi[k]) - x[s]) / h;
        const double v =
            A * y[s] + B * y[s + 1] + ((A * A * A - A) * M[size_t(s)] + (B * B * B - B) * M[size_t(s + 1)]) * h * h / 6.0;
        yi[k] = float(v);
    }
}

Mesh build(int nx, int nz0, float dx, float dz0, float ox, float oz, int pml, const Surface& s, const float* c55,
           const float* rho, float f0min) {
    if (nx < 1 || nz0 < 3) fail(OWL_ERR_ARG, "the free surface needs at least 3 rows");
    if (!(s.refine > 0)) fail(OWL_ERR_ARG, "free_surface_dz_refine must be positive");
    if (!(s.dz_max > 0)) fail(OWL_ERR_ARG, "dz_max must be positive");
    if (!(f0min > 0)) fail(OWL_ERR_ARG, "source f0 must be positive");
    if (s.tx.size() != s.tz.size()) fail(OWL_ERR_ARG, "topography x and elevation differ in length");
    for (size_t k = 1; k < s.tx.size(); ++k)
        if (!(s.tx[k] > s.tx[k - 1])) fail(OWL_ERR_ARG, "topography x must increase");
    Mesh m;
    m.nx = nx;
    m.nz0 = nz0;
    m.pml = pml;
    m.dx = dx;
    m.dz0 = dz0;
    m.ox = ox;
    m.oz = oz;
    const int nci = nx + 2 * pml + 2;
    const std::vector<float> topox = regspace(ox - float(pml + 1) * dx, dx, ox + float(nx + pml) * dx);
    if (int(topox.size()) != nci) fail(OWL_ERR_ARG, "topography abscissae do not match the padded window");
    m.topo_i.assign(size_t(nci), 0.0f);
    m.slope_i.assign(size_t(nci), 0.0f);
    m.topo_h.assign(size_t(nci), 0.0f);
    m.slope_h.assign(size_t(nci), 0.0f);
    if (!s.tx.empty()) {
        const bool cubic = s.interp != 0;
        ginterp(s.tx.data(), s.tz.data(), int(s.tx.size()), topox.data(), nci, m.topo_i.data(), cubic);
        m.slope_i = deriv(m.topo_i);
        for (auto& v : m.slope_i) v = v / dx;
        std::vector<float> xh(static_cast<size_t>(nci));
        for (int k = 0; k < nci; ++k) xh[size_t(k)] = topox[size_t(k)] - 0.5f * dx;
        ginterp(topox.data(), m.topo_i.data(), nci, xh.data(), nci, m.topo_h.data(), cubic);
        m.slope_h = deriv(m.topo_h);
        for (auto& v : m.slope_h) v = v / dx;
    }
    const float tmax = *std::max_element(m.topo_i.begin(), m.topo_i.end());
    const float depth_max = float(nz0 - 1) * dz0 - tmax;
    if (!(depth_max > 0)) fail(OWL_ERR_ARG, "the topography rises above the window's depth range");
    float eta_max = depth_max;
    bool bad = false;
    for (int k = 0; k < nci; ++k)
        if ((m.topo_i[size_t(k)] + depth_max) / eta_max < std::fmax(1.0f, std::fabs(m.slope_i[size_t(k)]))) bad = true;
    if (bad) {
        float mn = INFINITY;
        for (int k = 0; k < nci; ++k)
            mn = std::fmin(mn, (m.topo_i[size_t(k)] + depth_max) / std::fmax(1.0f, std::fabs(m.slope_i[size_t(k)])));
        eta_max = 0.975f * mn;
    }
    if (!(eta_max > 0)) fail(OWL_ERR_ARG, "the topography leaves no depth below the free surface");
    const float dz = eta_max / float(nz0 - 1);
    // sqrt(mean(c55(:, 1)/rho(:, 1)))/minval(f0) over the padded first row
    float sm = 0.0f;
    for (int i = -pml; i <= nx + pml + 1; ++i) {
        const int ic = std::min(std::max(i, 1), nx);
        const size_t q = size_t(ic - 1) * size_t(nz0);
        sm = sm + c55[q] / rho[q];
    }
    const float rl = std::sqrt(sm / float(nci)) / f0min;
    const int refine_nz = int(std::ceil(0.5f * rl / dz));
    const float rr = s.refine, cap = s.dz_max * eta_max / (depth_max + tmax);
    float dzm = std::fmin(cap, dz / rr);
    if (!(dzm > 0) || refine_nz < 1) fail(OWL_ERR_ARG, "the free-surface mesh spacing is not positive");
    std::vector<float> eta = regspace(0.0f, dzm, float(refine_nz - 1) * dz);
    if (eta.empty()) fail(OWL_ERR_ARG, "empty refined layer");
    while (eta.back() < eta_max) {
        dzm = std::fmin(dzm * 1.025f, cap);
        eta.push_back(eta.back() + dzm);
        if (eta.size() > 1000000) fail(OWL_ERR_ARG, "the free-surface mesh does not reach eta_max");
    }
    const int nz = int(eta.size()), nrj = nz + 2 * pml + 2;
    m.nz = nz;
    m.dz = dz;
    m.eta_max = eta_max;
    m.depth_max = depth_max;
    m.topo_max = tmax;
    std::vector<float> ei(size_t(nrj), 0.0f), eh(size_t(nrj), 0.0f);
    auto I = [&](int j) { return size_t(j + pml); };
    for (int j = 1; j <= nz; ++j) ei[I(j)] = eta[size_t(j - 1)];
    eh[I(1)] = -0.5f * dz / rr;
    for (int j = 1; j <= nz - 1; ++j) eh[I(j + 1)] = 0.5f * (eta[size_t(j - 1)] + eta[size_t(j)]);
    for (int i = 1; i <= pml + 1; ++i) {
        ei[I(1 - i)] = ei[I(1)] - float(i) * dz / rr;
        ei[I(nz + i)] = ei[I(nz)] + float(i) * dzm;
        eh[I(1 - i)] = eh[I(1)] - float(i) * dz / rr;
        eh[I(nz + i)] = eh[I(nz)] + float(i) * dzm;
    }
    m.edz_i = deriv(ei);
    m.edz_h = deriv(eh);
    m.sc_i.resize(size_t(nrj));
    m.sc_h.resize(size_t(nrj));
    for (int k = 0; k < nrj; ++k) {
        m.sc_i[size_t(k)] = dz / m.edz_i[size_t(k)];
        m.sc_h[size_t(k)] = dz / m.edz_h[size_t(k)];
    }
    m.zz.assign(size_t(nci) * size_t(nrj), 0.0f);
    auto Z = [&](int i, int j) -> float& { return m.zz[size_t(i + pml) * size_t(nrj) + size_t(j + pml)]; };
    std::vector<float> col(static_cast<size_t>(nz));
    for (int i = -pml; i <= nx + pml + 1; ++i) {
        rescale(eta.data(), nz, -m.topo_i[size_t(i + pml)], depth_max, col.data());
        for (int j = 1; j <= nz; ++j) Z(i, j) = col[size_t(j - 1)];
    }
    float dsum = 0.0f;
    for (int i = -pml; i <= nx + pml + 1; ++i) dsum = dsum + (Z(i, nz) - Z(i, nz - 1));
    const float dzmean = dsum / float(nci);
    for (int i = -pml; i <= nx + pml + 1; ++i) {
        for (int k = 1; k <= pml + 1; ++k) Z(i, nz + k) = Z(i, nz) + float(k) * dzmean;
        const float d1 = Z(i, 2) - Z(i, 1);
        for (int k = 1; k <= pml + 1; ++k) Z(i, 1 - k) = Z(i, 1) - float(k) * d1;
    }
    const float clo = float(-pml) * dz;
    for (int k = 0; k < nrj; ++k) {
        ei[size_t(k)] = std::fmin(std::fmax(ei[size_t(k)], clo), eta_max);
        eh[size_t(k)] = std::fmin(std::fmax(eh[size_t(k)], clo), eta_max);
    }
    m.eta_i = std::move(ei);
    m.eta_h = std::move(eh);
    return m;
}

std::vector<float> medium(const Mesh& m, const float* v) {
    const int nx = m.nx, nz0 = m.nz0, pml = m.pml, nrj = m.nrj(), nr = nz0 + 2 * pml + 2;
    std::vector<float> out(size_t(m.nci()) * size_t(nrj));
    std::vector<float> zr(static_cast<size_t>(nr)), vr(static_cast<size_t>(nr)), zq(static_cast<size_t>(nrj));
    for (int k = -pml; k <= nz0 + pml + 1; ++k) zr[size_t(k + pml)] = float(k - 1) * m.dz0 - m.topo_max;
    for (int i = -pml; i <= nx + pml + 1; ++i) {
        const int ic = std::min(std::max(i, 1), nx);
        for (int k = -pml; k <= nz0 + pml + 1; ++k) {
            const int kc = std::min(std::max(k, 1), nz0);
            vr[size_t(k + pml)] = v[size_t(kc - 1) + size_t(ic - 1) * size_t(nz0)];
        }
        for (int j = -pml; j <= m.nz + pml + 1; ++j) zq[size_t(j + pml)] = m.zz_at(i, j);
        ginterp(zr.data(), vr.data(), nr, zq.data(), nrj, out.data() + size_t(i + pml) * size_t(nrj), false);
    }
    return out;
}

ZLoc locate_z(const Mesh& m, const Surface& s, float x, float z, bool source) {
    const int pml = m.pml, nci = m.nci();
    std::vector<float> topox = regspace(0.0f, m.dx, float(m.nx + 2 * pml + 1) * m.dx);
    if (int(topox.size()) != nci) fail(OWL_ERR_ARG, "topography abscissae do not match the padded window");
    std::vector<float> surf(static_cast<size_t>(nci));
    for (int k = 0; k < nci; ++k) {
        topox[size_t(k)] = topox[size_t(k)] - float(pml + 1) * m.dx;
        surf[size_t(k)] = m.zz_at(k - pml, 1);
    }
    float stp = 0.0f;
    ginterp(topox.data(), surf.data(), nci, &x, 1, &stp, s.interp != 0);
    if (source ? s.src_from_surface : s.rec_from_surface) z = z + stp;
    ZLoc L;
    L.zv = (z - stp) / (-stp + m.depth_max) * m.eta_max;
    if (source && L.zv < m.edz_i[size_t(1 + pml)]) {
        L.zv = m.edz_i[size_t(1 + pml)];
        L.amp_scale = 1.2f;
    }
    const float zs = L.zv + (L.zv == 0.0f ? FLT_EPSILON : 0.0f);
    float bi = INFINITY, bh = INFINITY;
    for (int j = -pml; j <= m.nz + pml + 1; ++j) {
        const float di = std::fabs(L.zv - m.oz - m.eta_i[size_t(j + pml)]);
        if (di < bi) {
            bi = di;
            L.gz = j;
        }
        const float dh = std::fabs(zs - m.oz - m.eta_h[size_t(j + pml)]);
        if (dh < bh) {
            bh = dh;
            L.hz = j;
        }
    }
    kaiser_sinc((L.zv - m.oz - m.eta_i[size_t(L.gz + pml)]) / m.edz_i[size_t(L.gz + pml)], L.iz);
    kaiser_sinc((L.zv - m.oz - m.eta_h[size_t(L.hz + pml)]) / m.edz_h[size_t(L.hz + pml)], L.hzw);
    return L;
}

void to_regular(const Mesh& m, const float* g, float* out) {
    const int nx = m.nx, nz = m.nz, nz0 = m.nz0;
    std::vector<float> x(static_cast<size_t>(nz)), y(static_cast<size_t>(nz)), xq(static_cast<size_t>(nz0)), yq(static_cast<size_t>(nz0));
    for (int j = 1; j <= nz0; ++j) xq[size_t(j - 1)] = float(j - 1) * m.dz0 - m.topo_max;
    for (int i = 1; i <= nx; ++i) {
        for (int j = 1; j <= nz; ++j) {
            x[size_t(j - 1)] = m.zz_at(i, j);
            y[size_t(j - 1)] = g[size_t(i - 1) + size_t(j - 1) * size_t(nx)];
        }
        ginterp(x.data(), y.data(), nz, xq.data(), nz0, yq.data(), true);
        for (int j = 1; j <= nz0; ++j) {
            const float zz = float(j - 1) * m.dz0;
            out[size_t(j - 1) + size_t(i - 1) * size_t(nz0)] =
                zz < m.topo_max - m.topo_i[size_t(i + m.pml)] ? 0.0f : yq[size_t(j - 1)];
        }
    }
}

SsgMesh build_ssg(int nx, int nz0, float dz, int pml, float refine, float dz_max, const float* c55, const float* rho,
                  float f0min) {
    if (nx < 1 || nz0 < 3) fail(OWL_ERR_ARG, "the free surface needs at least 3 rows");
    if (!(refine > 0)) fail(OWL_ERR_ARG, "free_surface_dz_refine must be positive");
    if (!(dz_max > 0)) fail(OWL_ERR_ARG, "dz_max must be positive");
    if (!(f0min > 0)) fail(OWL_ERR_ARG, "source f0 must be positive");
    float sm = 0.0f;
    for (int i = -pml; i <= nx + pml + 1; ++i) {
        const int ic = std::min(std::max(i, 1), nx);
        const size_t q = size_t(ic - 1) * size_t(nz0);
        sm = sm + c55[q] / rho[q];
    }
    const float rl = std::sqrt(sm / float(nx + 2 * pml + 2)) / f0min;
    const int rnz = int(std::ceil(0.5f * rl / dz));
    SsgMesh m;
    m.rr = refine;
    float dzm = dz / refine;
    std::vector<float> z = regspace(0.0f, dzm, float(rnz - 1) * dz);
    if (z.empty()) fail(OWL_ERR_ARG, "empty refined layer");
    while (z.back() < float(nz0 - 1) * dz) {
        dzm = std::fmin(dzm * 1.025f, dz_max);
        z.push_back(z.back() + dzm);
        if (z.size() > 10000000) fail(OWL_ERR_ARG, "the free-surface mesh does not reach the window's depth");
    }
    const int nz = int(z.size()), nrj = nz + 2 * pml + 2;
    std::vector<float> zh(static_cast<size_t>(nz));
    zh[0] = -0.5f * dz / refine;
    for (int i = 0; i + 1 < nz; ++i) zh[size_t(i + 1)] = 0.5f * (z[size_t(i)] + z[size_t(i + 1)]);
    const std::vector<float> di = deriv(z), dh = deriv(zh);
    m.nz = nz;
    m.pml = pml;
    m.dz = dz;
    m.dzm = dzm;
    for (auto* v : {&m.zi, &m.zh, &m.dzi, &m.dzh, &m.sci, &m.sch}) v->assign(size_t(nrj), 0.0f);
    for (int j = -pml; j <= nz + pml + 1; ++j) {
        const int jc = std::min(std::max(j, 1), nz);
        const size_t k = size_t(j + pml), c = size_t(jc - 1);
        m.dzi[k] = di[c];
        m.dzh[k] = dh[c];
        m.sci[k] = dz / di[c];
        m.sch[k] = dz / dh[c];
        if (j >= 1 && j <= nz) {
            m.zi[k] = z[c];
            m.zh[k] = zh[c];
        }
    }
    for (int i = 1; i <= pml + 1; ++i) {
        m.zi[size_t(1 - i + pml)] = m.zi[size_t(1 + pml)] - float(i) * dz / refine;
        m.zi[size_t(nz + i + pml)] = m.zi[size_t(nz + pml)] + float(i) * dzm;
        m.zh[size_t(1 - i + pml)] = m.zh[size_t(1 + pml)] - float(i) * dz / refine;
        m.zh[size_t(nz + i + pml)] = m.zh[size_t(nz + pml)] + float(i) * dzm;
    }
    return m;
}

void map_column_ssg(const SsgMesh& m, int nx, int nz0, const float* v, int i, float* out) {
    const int pml = m.pml, nr = nz0 + 2 * pml + 2, nq = m.nz + 2 * pml + 2;
    std::vector<float> zr(static_cast<size_t>(nr)), vr(static_cast<size_t>(nr));
    const int ic = std::min(std::max(i, 1), nx);
    for (int k = -pml; k <= nz0 + pml + 1; ++k) {
        const int kc = std::min(std::max(k, 1), nz0);
        zr[size_t(k + pml)] = float(k - 1) * m.dz;
        vr[size_t(k + pml)] = v[size_t(kc - 1) + size_t(ic - 1) * size_t(nz0)];
    }
    ginterp(zr.data(), vr.data(), nr, m.zi.data(), nq, out, false);
}

void locate_ssg(const SsgMesh& m, float z, float oz, int& gz, int& hz, float* iz, float* hzw) {
    const int pml = m.pml;
    const float zs = z + (z == 0.0f ? FLT_EPSILON : 0.0f);
    float bi = INFINITY, bh = INFINITY;
    for (int j = -pml; j <= m.nz + pml + 1; ++j) {
        const float di = std::fabs(z - oz - m.zi[size_t(j + pml)]);
        if (di < bi) {
            bi = di;
            gz = j;
        }
        const float dh = std::fabs(zs - oz - m.zh[size_t(j + pml)]);
        if (dh < bh) {
            bh = dh;
            hz = j;
        }
    }
    kaiser_sinc((z - oz - m.zi[size_t(gz + pml)]) / m.dzi[size_t(gz + pml)], iz);
    kaiser_sinc((z - oz - m.zh[size_t(hz + pml)]) / m.dzh[size_t(hz + pml)], hzw);
}

void column_to_regular_ssg(const SsgMesh& m, int nz0, const double* g, double* out) {
    spline_to_regular(m.zi.data() + m.pml + 1, g, m.nz, m.dz, nz0, out);
}

void spline_to_regular(const float* x, const double* g, int n, float dz, int nz0, double* out) {
    if (n < 3) fail(OWL_ERR_ARG, "too few refined rows");
    std::vector<double> M(size_t(n), 0.0), cp(size_t(n), 0.0), dp(size_t(n), 0.0);
    for (int i = 1; i < n - 1; ++i) {
        const double h0 = double(x[i]) - x[i - 1], h1 = double(x[i + 1]) - x[i];
        const double r = 6.0 * ((g[i + 1] - g[i]) / h1 - (g[i] - g[i - 1]) / h0);
        const double den = 2.0 * (h0 + h1) - h0 * cp[size_t(i - 1)];
        cp[size_t(i)] = h1 / den;
        dp[size_t(i)] = (r - h0 * dp[size_t(i - 1)]) / den;
    }
    for (int i = n - 2; i >= 1; --i) M[size_t(i)] = dp[size_t(i)] - cp[size_t(i)] * M[size_t(i + 1)];
    for (int k = 0; k < nz0; ++k) {
        const float v = float(k) * dz;
        const int s = segment(x, n, v);
        const double h = double(x[s + 1]) - x[s];
        const double A = (double(x[s + 1]) - v) / h, B = (double(v) - x[s]) / h;
        out[k] = A * g[s] + B * g[s + 1] + ((A * A * A - A) * M[size_t(s)] + (B * B * B - B) * M[size_t(s + 1)]) * h * h / 6.0;
    }
}

}  // namespace fsm
}  // namespace owl
