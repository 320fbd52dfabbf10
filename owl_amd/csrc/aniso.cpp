// 2D TTI moduli from Thomsen or Alkhalifah-Tsvankin parameters, and the
// chain rule that turns the c_ij gradients of the FSG solver into gradients
// of those parameters (OWL's mod_anisotropy, src/lib/module_anisotropy.f90,
// and the parameter block of elastic_tti_2d/submodule_fwi_gradient.f90:
// 484-560).
//
// The moduli are float32 in OWL's operation order.  The rotation is the
// Bond transformation of a rotation about y, restricted to the Voigt indices
// 1, 3, 5 (x, z, xz) that a 2D medium uses; every other entry of OWL's 6x6
// products is zero, so the float32 sums keep their order.  The chain rule is
// evaluated in double from the derivatives of the unrotated moduli, rotated
// by the same Bond matrix (the rotation does not depend on the parameters).
#include <cmath>

#include "common.hpp"
#include "owl_amd.h"

namespace owl {
namespace {

// c11, c13, c15, c33, c35, c55 as a symmetric 3x3 over (x, z, xz)
template <class T>
struct Voigt3 {
    T m[3][3];
};

template <class T>
Voigt3<T> from6(const T c[6]) {
    return Voigt3<T>{{{c[0], c[1], c[2]}, {c[1], c[3], c[4]}, {c[2], c[4], c[5]}}};
}

template <class T>
void to6(const Voigt3<T>& v, T c[6]) {
    c[0] = v.m[0][0];
    c[1] = v.m[0][1];
    c[2] = v.m[0][2];
    c[3] = v.m[1][1];
    c[4] = v.m[1][2];
    c[5] = v.m[2][2];
}

// Bond matrix rows for (x, z, xz) of the rotation about y by t,
// R = [[cos t, 0, sin t], [0, 1, 0], [-sin t, 0, cos t]] (bond_matrix,
// module_anisotropy.f90:126-173).
template <class T>
Voigt3<T> bond(T t) {
    const T co = std::cos(t), si = std::sin(t), ms = -si;
    return Voigt3<T>{{{co * co, si * si, T(2) * co * si},
                      {ms * ms, co * co, T(2) * ms * co},
                      {co * ms, si * co, co * co + si * ms}}};
}

// B C B^T with the inner product first (matmul(a, matmul(c, transpose(a))),
// cij_rotate_2d, module_anisotropy.f90:184-240), sums in index order.
template <class T>
Voigt3<T> rotate(const Voigt3<T>& B, const Voigt3<T>& C) {
    Voigt3<T> P{}, R{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            T s = T(0);
            for (int k = 0; k < 3; ++k) s = s + C.m[i][k] * B.m[j][k];
            P.m[i][j] = s;
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            T s = T(0);
            for (int k = 0; k < 3; ++k) s = s + B.m[i][k] * P.m[k][j];
            R.m[i][j] = s;
        }
    return R;
}

// thomsen_to_cij2d (module_anisotropy.f90:305-330), float32
void thomsen_point(float vp, float vs, float rho, float eps, float del, float the, float c[6]) {
    const float c33 = vp * vp * rho;
    const float c55 = vs * vs * rho;
    const float c11 = c33 * (1.0f + 2.0f * eps);
    const float tmp = 2.0f * c33 * (c33 - c55) * del + (c33 - c55) * (c33 - c55);
    const float c13 = std::sqrt(std::fabs(tmp)) - c55;
    const float c0[6] = {c11, c13, 0.0f, c33, 0.0f, c55};
    if (the != 0.0f) {
        to6(rotate(bond(-the), from6(c0)), c);
    } else {
        for (int k = 0; k < 6; ++k) c[k] = c0[k];
    }
}

// alkhalifah_tsvankin_to_cij2d (module_anisotropy.f90:399-418), float32
void at_point(float vp, float vs, float rho, float eps, float eta, float the, float c[6]) {
    const float c11 = vp * vp * rho;
    const float c55 = vs * vs * rho;
    const float c33 = c11 / (1.0f + 2.0f * eps);
    const float c13 = rho * (std::sqrt(std::fabs((vp * vp / (1.0f + 2.0f * eps) - vs * vs) *
                                                 (vp * vp / (1.0f + 2.0f * eta) - vs * vs))) -
                             vs * vs);
    const float c0[6] = {c11, c13, 0.0f, c33, 0.0f, c55};
    if (the != 0.0f) {
        to6(rotate(bond(-the), from6(c0)), c);
    } else {
        for (int k = 0; k < 6; ++k) c[k] = c0[k];
    }
}

// Derivatives of the unrotated moduli (c11, c13, 0, c33, 0, c55) with
// respect to the four parameters, in double.
void thomsen_dc0(double vp, double vs, double rho, double eps, double del, double d[4][6]) {
    const double c33 = rho * vp * vp, c55 = rho * vs * vs;
    const double T = 2 * c33 * (c33 - c55) * del + (c33 - c55) * (c33 - c55);
    const double sq = std::sqrt(std::fabs(T)), sg = T < 0 ? -1.0 : 1.0;
    const double d13_dT = sg / (2 * sq);
    for (int p = 0; p < 4; ++p)
        for (int k = 0; k < 6; ++k) d[p][k] = 0;
    // vp
    const double dc33 = 2 * rho * vp;
    d[0][0] = dc33 * (1 + 2 * eps);
    d[0][3] = dc33;
    d[0][1] = d13_dT * (2 * (2 * c33 - c55) * del + 2 * (c33 - c55)) * dc33;
    // vs
    const double dc55 = 2 * rho * vs;
    d[1][5] = dc55;
    d[1][1] = d13_dT * (-2 * c33 * del - 2 * (c33 - c55)) * dc55 - dc55;
    // epsilon
    d[2][0] = 2 * c33;
    // delta
    d[3][1] = d13_dT * 2 * c33 * (c33 - c55);
}

void at_dc0(double vp, double vs, double rho, double eps, double eta, double d[4][6]) {
    const double fe = 1 + 2 * eps, fn = 1 + 2 * eta;
    const double a = vp * vp / fe - vs * vs, b = vp * vp / fn - vs * vs;
    const double U = a * b, sq = std::sqrt(std::fabs(U)), sg = U < 0 ? -1.0 : 1.0;
    const double d13_dU = rho * sg / (2 * sq);
    for (int p = 0; p < 4; ++p)
        for (int k = 0; k < 6; ++k) d[p][k] = 0;
    // vp
    d[0][0] = 2 * rho * vp;
    d[0][3] = 2 * rho * vp / fe;
    d[0][1] = d13_dU * (2 * vp / fe * b + a * 2 * vp / fn);
    // vs
    d[1][5] = 2 * rho * vs;
    d[1][1] = d13_dU * (-2 * vs * b - 2 * vs * a) - 2 * rho * vs;
    // epsilon
    d[2][3] = -2 * rho * vp * vp / (fe * fe);
    d[2][1] = d13_dU * (-2 * vp * vp / (fe * fe)) * b;
    // eta
    d[3][1] = d13_dU * a * (-2 * vp * vp / (fn * fn));
}

}  // namespace
}  // namespace owl

using namespace owl;

extern "C" {

owl_status owl_tti2d_cij(int param, long long n, const float* vp, const float* vs, const float* rho,
                         const float* p1, const float* p2, const float* theta, float* const* cij) {
    return guarded([&] {
        if (n < 0 || !vp || !vs || !rho || !cij) fail(OWL_ERR_ARG, "null argument");
        for (int k = 0; k < 6; ++k)
            if (!cij[k]) fail(OWL_ERR_ARG, "null c_ij output");
        if (param != OWL_ANISO_ISO && (!p1 || !p2 || !theta)) fail(OWL_ERR_ARG, "null anisotropy parameter");
        for (long long q = 0; q < n; ++q) {
            float c[6];
            if (param == OWL_ANISO_ISO) {
                // 'iso' branch of prepare_modeling (elastic_tti_2d/module_vars.f90:287-299)
                const float c33 = rho[q] * (vp[q] * vp[q]), c55 = rho[q] * (vs[q] * vs[q]);
                c[0] = c33;
                c[1] = c33 - 2.0f * c55;
                c[2] = 0.0f;
                c[3] = c33;
                c[4] = 0.0f;
                c[5] = c55;
            } else if (param == OWL_ANISO_THOMSEN) {
                thomsen_point(vp[q], vs[q], rho[q], p1[q], p2[q], theta[q], c);
            } else if (param == OWL_ANISO_AT) {
                at_point(vp[q], vs[q], rho[q], p1[q], p2[q], theta[q], c);
            } else {
                fail(OWL_ERR_ARG, "unknown parameterisation");
            }
            for (int k = 0; k < 6; ++k) cij[k][q] = c[k];
        }
    });
}

owl_status owl_tti2d_param_gradient(int param, long long n, const float* vp, const float* vs, const float* rho,
                                    const float* p1, const float* p2, const float* theta,
                                    const float* const* grad_cij, float* const* grads) {
    return guarded([&] {
        if (n < 0 || !vp || !vs || !rho || !grad_cij || !grads) fail(OWL_ERR_ARG, "null argument");
        for (int k = 0; k < 6; ++k)
            if (!grad_cij[k]) fail(OWL_ERR_ARG, "null c_ij gradient");
        const int np = param == OWL_ANISO_ISO ? 2 : 4;
        for (int k = 0; k < np; ++k)
            if (!grads[k]) fail(OWL_ERR_ARG, "null parameter gradient output");
        if (param != OWL_ANISO_ISO && (!p1 || !p2 || !theta)) fail(OWL_ERR_ARG, "null anisotropy parameter");
        if (param != OWL_ANISO_ISO && param != OWL_ANISO_THOMSEN && param != OWL_ANISO_AT)
            fail(OWL_ERR_ARG, "unknown parameterisation");
        for (long long q = 0; q < n; ++q) {
            // OWL rescales the velocities to km/s and the density to g/cm^3
            // before the chain rule (submodule_fwi_gradient.f90:476-480, 486-488)
            const double v = double(vp[q]) * 1.0e-3, s = double(vs[q]) * 1.0e-3, r = double(rho[q]) * 1.0e-3;
            double g[6];
            for (int k = 0; k < 6; ++k) g[k] = grad_cij[k][q];
            if (param == OWL_ANISO_ISO) {
                grads[0][q] = float(2 * r * v * (g[0] + g[1] + g[3]));
                grads[1][q] = float(2 * r * s * (g[5] - 2 * g[1]));
                continue;
            }
            double d[4][6];
            if (param == OWL_ANISO_THOMSEN)
                thomsen_dc0(v, s, r, p1[q], p2[q], d);
            else
                at_dc0(v, s, r, p1[q], p2[q], d);
            const Voigt3<double> B = bond(-double(theta[q]));
            for (int p = 0; p < 4; ++p) {
                double dr[6];
                to6(rotate(B, from6(d[p])), dr);
                double acc = 0;
                for (int k = 0; k < 6; ++k) acc += dr[k] * g[k];
                grads[p][q] = float(acc);
            }
        }
    });
}

}  // extern "C"
