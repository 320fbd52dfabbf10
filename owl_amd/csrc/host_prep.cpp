// Host-side preparation for OWL's acoustic solvers.  See host_prep.hpp.
#include "host_prep.hpp"

#include <cmath>
#include <cstring>

namespace owl {

namespace {
// Optimised staggered coefficients, src/lib/module_fdcoef.f90:24-71.
const float kCoef[9][8] = {
    {0},
    {0},
    {0.1129042e+1f, -0.4301412e-1f},
    {0.1185991e+1f, -0.7249965e-1f, 0.6301572e-2f},
    {0.1217990e+1f, -0.9382142e-1f, 0.1507536e-1f, -0.1700324e-2f},
    {0.1236607e+1f, -0.1082265e+0f, 0.2343440e-1f, -0.5033546e-2f, 0.6817483e-3f},
    {0.1247662e+1f, -0.1175538e+0f, 0.2997970e-1f, -0.8719078e-2f, 0.2215897e-2f, -0.3462075e-3f},
    {0.1254799e+1f, -0.1238928e+0f, 0.3494371e-1f, -0.1208897e-1f, 0.4132531e-2f, -0.1197110e-2f,
     0.2122227e-3f},
    {0.1259312e+1f, -0.1280347e+0f, 0.3841945e-1f, -0.1473229e-1f, 0.5924913e-2f, -0.2248618e-2f,
     0.7179226e-3f, -0.1400855e-3f},
};

constexpr double kPi = 3.14159265358979323846264338327950288;

// Modified Bessel function of the first kind, order 0, from its series
// sum_k ((x/2)^k / k!)^2; terms are added until they no longer change the sum.
double bessel_i0(double x) {
    const double q = 0.25 * x * x;
    double term = 1.0, total = 1.0;
    for (int k = 1; k < 500; ++k) {
        term = term * (q / (double(k) * double(k)));
        total += term;
        if (term < 1e-18 * total) break;
    }
    return total;
}
}  // namespace

bool fd_coefficients(int R, float* c) {
    if (R < 2 || R > 8) return false;
    for (int m = 0; m < R; ++m) c[m] = kCoef[R][m];
    return true;
}

int fortran_nint(float x) { return int(std::round(x)); }

void kaiser_sinc(float x, float h[9]) {
    // kaiser_b0 is a double initialised from the default-real literal 6.31.
    const double beta = double(6.31f);
    for (int k = 0; k < 9; ++k) h[k] = 0.0f;
    if (x == 0.0f) {
        h[kNkw] = 1.0f;
        return;
    }
    const double norm = bessel_i0(beta);
    for (int i = -kNkw; i <= kNkw; ++i) {
        const float t = float(i) - x;
        const double arg = kPi * double(t);
        const float s = float(arg == 0.0 ? 1.0 : std::sin(arg) / arg);
        float w = 0.0f;
        if (t >= -float(kNkw) && t <= float(kNkw)) {
            const float u = t / float(kNkw);
            const float r = std::sqrt(1.0f - u * u);
            w = float(bessel_i0(beta * double(r)) / norm);
        }
        h[i + kNkw] = s * w;
    }
}

Located locate(int dims, float x, float y, float z, float ox, float oy, float oz, float dx,
               float dy, float dz) {
    Located L;
    L.gx = fortran_nint((x - ox) / dx) + 1;
    L.gz = fortran_nint((z - oz) / dz) + 1;
    L.hx = fortran_nint((x - ox + 0.5f * dx) / dx) + 1;
    L.hz = fortran_nint((z - oz + 0.5f * dz) / dz) + 1;
    kaiser_sinc((x - ox) / dx - float(L.gx - 1), L.ix);
    kaiser_sinc((z - oz) / dz - float(L.gz - 1), L.iz);
    kaiser_sinc((x - ox + 0.5f * dx) / dx - float(L.hx - 1), L.hxw);
    kaiser_sinc((z - oz + 0.5f * dz) / dz - float(L.hz - 1), L.hzw);
    if (dims == 3) {
        L.gy = fortran_nint((y - oy) / dy) + 1;
        L.hy = fortran_nint((y - oy + 0.5f * dy) / dy) + 1;
        kaiser_sinc((y - oy) / dy - float(L.gy - 1), L.iy);
        kaiser_sinc((y - oy + 0.5f * dy) / dy - float(L.hy - 1), L.hyw);
    } else {
        kaiser_sinc(0.0f, L.iy);
        kaiser_sinc(0.0f, L.hyw);
    }
    return L;
}

int stf_length(float f0, float dt, float f0factor) {
    return int(std::ceil(2.0f * f0factor / f0 / dt));
}

std::vector<float> source_time_function(int wavelet, float f0, float dt, float f0factor,
                                        int* hnt) {
    const int n = stf_length(f0, dt, f0factor);
    std::vector<float> w(n > 0 ? n : 0);
    float peak = 0.0f;
    for (int t = 1; t <= n; ++t) {
        const float tau = float(t - 1) * dt - f0factor / f0;
        // (const_pi*f0*t)**2 is formed in double precision and stored as real.
        const double a = kPi * double(f0) * double(tau);
        const float q = float(a * a);
        float v;
        switch (wavelet) {
            case kGaussian: v = std::exp(-q); break;
            case kGaussianDeriv: v = -tau * std::exp(-q); break;
            case kRickerDeriv: v = ((2.0f * q - 3.0f) * std::exp(-q)) * tau; break;
            default: v = (1.0f - 2.0f * q) * std::exp(-q); break;
        }
        w[t - 1] = v;
        peak = std::fabs(v) > peak ? std::fabs(v) : peak;
    }
    for (float& v : w) v = v / peak;
    if (hnt) *hnt = fortran_nint(float(n) / 2.0f);
    return w;
}

std::vector<float> stf_time_derivative(const std::vector<float>& stf, float dt, int nt) {
    const size_t n = stf.size();
    std::vector<float> d(n, 0.0f);
    if (n >= 2) {
        d[0] = stf[1] - stf[0];
        d[n - 1] = stf[n - 1] - stf[n - 2];
        for (size_t i = 1; i + 1 < n; ++i) d[i] = 0.5f * (stf[i + 1] - stf[i - 1]);
    }
    float peak = 0.0f;
    for (float& v : d) {
        v = v / dt;
        peak = std::fabs(v) > peak ? std::fabs(v) : peak;
    }
    if (n > 0 && peak > 0.0f && std::fabs(d[n - 1]) / peak > 0.1f && size_t(nt) > n) d.resize(size_t(nt), d[n - 1]);
    return d;
}

float dt_stable(int R, int dims, float vmax, float dx, float dy, float dz) {
    float csum = 0.0f;
    for (int m = 0; m < R; ++m) csum = csum + std::fabs(kCoef[R][m]);
    const float s = dims == 3 ? 1.0f / (dx * dx) + 1.0f / (dy * dy) + 1.0f / (dz * dz)
                              : 1.0f / (dx * dx) + 1.0f / (dz * dz);
    return 1.0f / (csum * vmax * std::sqrt(s));
}

int time_steps(float tmax, float dt) { return fortran_nint(tmax / dt + 1.0f); }

namespace {
// damp_coef, acoustic_iso_2d/module_cfspml.f90:37-56 (npower_a = 1,
// npower_d = 2, kmax = 1, R0 = 1e-5).
void damping(float pmlvp, float d, float dist, int pml, float dt, float alphamax, float* a,
             float* b) {
    const float r0 = 1.0e-5f;
    const float nd = dist / (float(pml) * d);
    float alpha = alphamax * (1.0f - nd);
    const float floor_alpha = 0.25f * alphamax;
    if (floor_alpha > alpha) alpha = floor_alpha;
    const float damp = pmlvp * (2.0f + 1.0f) * std::log(1.0f / r0) / (2.0f * float(pml) * d) * (nd * nd);
    const float kappa = 1.0f;
    const float adt = std::fabs(dt);
    *a = -(2.0f * adt * damp / kappa / (2.0f + adt * (alpha + damp / kappa)));
    *b = (2.0f - adt * (alpha + damp / kappa)) / (2.0f + adt * (alpha + damp / kappa));
}
}  // namespace

PmlAxis cfspml_axis(int n, int pml, float d, float dt, float pmlvp, float f0max) {
    // alphamax = 1.0*max(f0)*const_pi, with const_pi in double precision.
    const float alphamax = float(double(1.0f * f0max) * kPi);
    PmlAxis P;
    const int len = n + 2 * pml;
    P.ai.assign(len, 0.0f);
    P.bi.assign(len, 0.0f);
    P.ah.assign(len, 0.0f);
    P.bh.assign(len, 0.0f);
    for (int i = -pml + 1; i <= n + pml; ++i) {
        if (!(i <= 1 || i >= n + 1)) continue;
        float di, dh;
        if (i <= 1) {
            di = std::fabs(float(i - 1) * d);
            dh = std::fabs((float(i - 1) - 0.5f) * d);
        } else {
            di = std::fabs(float(i - n) * d);
            dh = std::fabs((float(i - n) - 0.5f) * d);
        }
        const int q = i + pml - 1;
        damping(pmlvp, d, di, pml, dt, alphamax, &P.ai[q], &P.bi[q]);
        damping(pmlvp, d, dh, pml, dt, alphamax, &P.ah[q], &P.bh[q]);
    }
    return P;
}

// ---------------------------------------------------------------------------
// M-PML damping ratio of a general 2D anisotropic medium
// (elastic_tti_2d/module_cfspml.f90:80-280, Meza-Fajardo and Papageorgiou
// 2008 extended to anisotropy).  A = trace and B = minus the determinant of
// the 2D Christoffel matrix; for a stable medium both squared eigenvalues
// -(A rho +- sqrt((A^2 + 4B) rho^2)) / (2 rho^2) are negative reals, so the
// eigenvalues are purely imaginary and the complex arithmetic of
// mpml_eigenvalue_derivative reduces to the real arithmetic below, in the
// same order.  Integer powers as gfortran expands them.
// ---------------------------------------------------------------------------
namespace {

inline float p2(float x) { return x * x; }
inline float p3(float x) { return (x * x) * x; }
inline float p4(float x) {
    const float y = x * x;
    return y * y;
}
inline float chr_a(float c11, float c15, float c33, float c35, float c55, float k1, float k3) {
    return (c11 + c55) * p2(k1) + 2 * (c15 + c35) * k1 * k3 + (c33 + c55) * p2(k3);
}
inline float chr_b(float c11, float c13, float c15, float c33, float c35, float c55, float k1, float k3) {
    return (p2(c15) - c11 * c55) * p4(k1) + 2 * (c13 * c15 - c11 * c35) * p3(k1) * k3 +
           (p2(c13) - c11 * c33 - 2 * c15 * c35 + 2 * c13 * c55) * p2(k1) * p2(k3) +
           2 * (-(c15 * c33) + c13 * c35) * k1 * p3(k3) + (p2(c35) - c33 * c55) * p4(k3);
}
// real part of eigval**2 for mode 1 or 2 (mpml_eigenvalue, :80-112)
float eigen_sq(const float* c, float rho, int mode, float k1, float k3) {
    const float A = chr_a(c[0], c[2], c[3], c[4], c[5], k1, k3);
    const float B = chr_b(c[0], c[1], c[2], c[3], c[4], c[5], k1, k3);
    const float s = std::sqrt((p2(A) + 4 * B) * p2(rho));
    const float negx = mode == 1 ? (A * rho + s) / p2(rho) : -((-(A * rho) + s) / p2(rho));
    const float b = std::sqrt(negx) / std::sqrt(2.0f);
    return -(b * b);
}
// real part of mpml_eigenvalue_derivative (:114-160) for scaled moduli c
float eigen_deriv(const float* c, float rho, bool dirx, float e2, float ratio, float k1, float k3) {
    const float c11 = c[0], c13 = c[1], c15 = c[2], c33 = c[3], c35 = c[4], c55 = c[5];
    const float e4 = e2 * e2;
    const float den = 2 * chr_b(c11, c13, c15, c33, c35, c55, k1, k3) -
                      3 * e2 * chr_a(c11, c15, c33, c35, c55, k1, k3) * rho - 4 * e4 * p2(rho);
    float num;
    if (dirx) {
        num = 2 * c15 * c35 * p2(k1) * p2(k3) + 3 * c15 * c33 * k1 * p3(k3) - 2 * p2(c35) * p4(k3) +
              2 * c33 * c55 * p4(k3) - 2 * p2(c15) * p4(k1) * ratio + 2 * c15 * c35 * p2(k1) * p2(k3) * ratio +
              c15 * c33 * k1 * p3(k3) * ratio - p2(c13) * p2(k1) * p2(k3) * (1 + ratio) -
              c13 * k1 * k3 * (c15 * p2(k1) * (1 + 3 * ratio) + k3 * (2 * c55 * k1 * (1 + ratio) + c35 * k3 * (3 + ratio))) +
              e2 * (k3 * (3 * c15 * k1 * (1 + ratio) + 3 * c35 * k1 * (1 + ratio) + c33 * k3 * (2 + ratio)) +
                    c55 * (p2(k3) * (2 + ratio) + p2(k1) * (1 + 2 * ratio))) * rho +
              2 * e4 * (1 + ratio) * p2(rho) +
              c11 * p2(k1) * (2 * c55 * p2(k1) * ratio + c33 * p2(k3) * (1 + ratio) + c35 * k1 * k3 * (1 + 3 * ratio) +
                              e2 * (1 + 2 * ratio) * rho);
    } else {
        num = -2 * p2(c15) * p4(k1) +
              p2(k3) * (-2 * (p2(c35) - c33 * c55) * p2(k3) * ratio - p2(c13) * p2(k1) * (1 + ratio) -
                        c13 * k1 * (2 * c55 * k1 * (1 + ratio) + c35 * k3 * (1 + 3 * ratio))) +
              e2 * (k3 * (3 * c35 * k1 * (1 + ratio) + c33 * k3 * (1 + 2 * ratio)) +
                    c55 * (p2(k1) * (2 + ratio) + p2(k3) * (1 + 2 * ratio))) * rho +
              2 * e4 * (1 + ratio) * p2(rho) +
              c15 * k1 * k3 * (-(c13 * p2(k1) * (3 + ratio)) + k3 * (2 * c35 * k1 * (1 + ratio) + c33 * k3 * (1 + 3 * ratio)) +
                               3 * e2 * (1 + ratio) * rho) +
              c11 * p2(k1) * (2 * c55 * p2(k1) + k3 * (c33 * k3 * (1 + ratio) + c35 * k1 * (3 + ratio)) +
                              e2 * (2 + ratio) * rho);
    }
    return num / den;
}

}  // namespace

float mpml_damp_ratio(const float* c, float rho, bool dirx) {
    const float dtheta = float(kPi / 180.0), alpha = -0.01f, dratio = 0.002f, sc = 1.0e-9f;
    const float d[6] = {c[0] * sc, c[1] * sc, c[2] * sc, c[3] * sc, c[4] * sc, c[5] * sc};
    const float dr = rho * sc;
    float ratio = 0.0f;
    for (int i = 1; i <= 180; ++i) {
        const float theta = (float(i) + 0.5f) * dtheta;
        const float kx = std::sin(theta), kz = std::cos(theta);
        const float e1 = eigen_sq(c, rho, 1, kx, kz), e2 = eigen_sq(c, rho, 2, kx, kz);
        float dg = std::fmax(eigen_deriv(d, dr, dirx, e1, ratio, kx, kz), eigen_deriv(d, dr, dirx, e2, ratio, kx, kz));
        while (dg >= alpha && ratio < 1.0f) {
            ratio = ratio + dratio;
            dg = std::fmax(eigen_deriv(d, dr, dirx, e1, ratio, kx, kz), eigen_deriv(d, dr, dirx, e2, ratio, kx, kz));
        }
    }
    return ratio;
}

}  // namespace owl
