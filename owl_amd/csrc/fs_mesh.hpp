// OWL's free-surface mesh for the 2D FSG solver: the depth-refined,
// topography-following mesh of one shot window (elastic_tti_2d/
// module_vars.f90:360-519), the window's medium on it
// (map_regular_to_irregular, :169-203), source and receiver rows on it
// (:639-680 and the depth-varying branch of prepare_geometry,
// src/lib/module_source_receiver.f90:556-760), and the refined-mesh
// gradient back on the window (map_irregular_to_regular, :208-249).
//
// OWL builds this with FLIT's regspace, deriv, rescale, ginterp and
// pad_array, which are not part of the OWL tree.  They are taken as:
//   regspace(a, d, b): nint((b - a)/d) + 1 values a + k d;
//   deriv(x, 'center'): (x(i+1) - x(i-1))/2 inside, one-sided at the ends;
//   rescale(x, [r1, r2]): (x - min)/(max - min)*(r2 - r1) + r1;
//   ginterp 'linear': y0 + (y1 - y0) t on the segment x0 <= v < x1, end
//     segments extended; 'cubic': natural cubic spline evaluated in double;
//   pad_array: edge replication; float_small = FLT_EPSILON.
// All in float32 in OWL's operation order.
#pragma once
#include <vector>

namespace owl {
namespace fsm {

struct Surface {
    std::vector<float> tx, tz;  // topography control points (x, elevation up); empty: flat at 0
    int interp = 1;             // 0 linear, 1 cubic (OWL's topo_interp, default 'cubic')
    float refine = 4.0f;        // free_surface_dz_refine
    float dz_max = 0.0f;        // the caller's dz_max (OWL's default 1.5 dz)
    bool src_from_surface = false, rec_from_surface = false;
};

// Mesh of one window: row tables over j = -pml .. nz+pml+1 at [j + pml],
// column tables over i = -pml .. nx+pml+1 at [i + pml].  eta_i and eta_h
// are clipped to [-pml dz, eta_max] as OWL clips them before stepping.
struct Mesh {
    int nx = 0, nz0 = 0, pml = 0, nz = 0;  // window columns, window rows, damping cells, refined rows
    float dx = 0, dz0 = 0, ox = 0, oz = 0;
    float dz = 0;  // eta_max / (nz0 - 1): the z spacing the update divides by
    float eta_max = 0, depth_max = 0, topo_max = 0;
    std::vector<float> topo_i, slope_i, topo_h, slope_h;
    std::vector<float> eta_i, eta_h, edz_i, edz_h, sc_i, sc_h;
    std::vector<float> zz;  // physical depth of integer row j in column i at [(i + pml) * nrj + j + pml]
    int nci() const { return nx + 2 * pml + 2; }
    int nrj() const { return nz + 2 * pml + 2; }
    float zz_at(int i, int j) const { return zz[size_t(i + pml) * size_t(nrj()) + size_t(j + pml)]; }
};

// FLIT's ginterp on increasing x: 'linear' (cubic = false) or the natural
// cubic spline evaluated in double (cubic = true), end segments extended.
void ginterp(const float* x, const float* y, int n, const float* xi, int m, float* yi, bool cubic);

// Builds the mesh.  c55, rho: the window (nz0, nx) z fastest; f0min: the
// smallest source frequency of the shot.  Throws owl::Error on input that
// cannot give a mesh.
Mesh build(int nx, int nz0, float dx, float dz0, float ox, float oz, int pml, const Surface& s, const float* c55,
           const float* rho, float f0min);

// The window's values v (nz0, nx) z fastest on the refined mesh, over
// i = -pml .. nx+pml+1, j = -pml .. nz+pml+1 at [(i + pml) * nrj + j + pml].
std::vector<float> medium(const Mesh& m, const float* v);

// A source or receiver on the mesh: virtual depth, nearest rows and
// Kaiser-sinc weights in z; x as on the regular mesh.  amp_scale is 1.2 for
// a source moved down to the first row below the surface, else 1.
struct ZLoc {
    float zv = 0;
    int gz = 1, hz = 1;
    float iz[9], hzw[9];
    float amp_scale = 1.0f;
};
ZLoc locate_z(const Mesh& m, const Surface& s, float x, float z, bool source);

// One refined-mesh gradient column set (nx, nz refined; x fastest) to the
// window (nz0, nx) z fastest: the column's cubic interpolant at the regular
// depths, zero above the free surface.
void to_regular(const Mesh& m, const float* g, float* out);

// The depth-refined mesh of OWL's 2D standard-staggered solver with a free
// surface (elastic_vhtiort_2d/module_vars.f90:310-371): rows dz/refine
// apart over half a Rayleigh wavelength (sqrt of the mean c55/rho of the
// padded first row over the smallest f0), then spacings min(1.025 dz_prev,
// dz_max) until (nz0 - 1) dz; half rows at the midpoints, the first at
// -dz/(2 refine); dz_i = deriv(z_i), dz_h = deriv(z_h) and the scalings dz/dz_i,
// dz/dz_h on the unpadded rows, then edge-padded; z_i, z_h extended by
// dz/refine above and by the last spacing below.  Tables over j = -pml ..
// nz+pml+1 at [j + pml].  c55, rho: the window (nz0, nx) z fastest.
struct SsgMesh {
    int nz = 0, pml = 0;
    float dz = 0, rr = 0, dzm = 0;
    std::vector<float> zi, zh, dzi, dzh, sci, sch;
};
SsgMesh build_ssg(int nx, int nz0, float dz, int pml, float refine, float dz_max, const float* c55, const float* rho,
                  float f0min);

// Column i of the window's values v (nz0, nx) z fastest, padded by edge
// replication over rows k = -pml .. nz0+pml+1 at (k - 1) dz, mapped onto
// z_i of the refined mesh by FLIT's linear ginterp (map_regular_to_irregular,
// module_vars.f90:139-170): out[j + pml], j = -pml .. nz+pml+1.
void map_column_ssg(const SsgMesh& m, int nx, int nz0, const float* v, int i, float* out);

// A point's refined rows (prepare_geometry's depth-varying branch,
// src/lib/module_source_receiver.f90:596-680): gz, hz the first rows nearest
// to z - oz on z_i and to (z + float_small at z = 0) - oz on z_h, weights
// Kaiser-sinc of the offsets over dz_i(gz), dz_h(hz).
void locate_ssg(const SsgMesh& m, float z, float oz, int& gz, int& hz, float* iz, float* hzw);

// One window column of refined-row values g[j - 1], j = 1 .. nz, to the
// regular rows 1 .. nz0 at (j - 1) dz by the natural cubic spline through
// z_i(1 .. nz) (map_irregular_to_regular, module_vars.f90:175-203).
void column_to_regular_ssg(const SsgMesh& m, int nz0, const double* g, double* out);

// The natural cubic spline through (x[k], g[k]), k < n, at (k - 1) dz for
// k = 1 .. nz0.
void spline_to_regular(const float* x, const double* g, int n, float dz, int nz0, double* out);

// Stable dt of the refined mesh (module_vars.f90:532-567): min over the
// mesh of 1 / (sum|c| vmax dz0) with OWL's metric-scaled spacing.
// vmax(i, j): the largest phase velocity at refined cell (i, j).
template <class V>
float dt_stable(const Mesh& m, float csum, V&& vmax);

}  // namespace fsm
}  // namespace owl

// ---------------------------------------------------------------------------
#include <cmath>

template <class V>
float owl::fsm::dt_stable(const Mesh& m, float csum, V&& vmax) {
    float dts = INFINITY;
    for (int j = 1; j <= m.nz + m.pml + 1; ++j)
        for (int i = -m.pml + 1; i <= m.nx + m.pml; ++i) {
            const float e = m.eta_i[size_t(j + m.pml)], ed = m.edz_i[size_t(j + m.pml)];
            const float a = 1.0f / m.dx + std::fmax(0.0f, m.eta_max - e) / m.eta_max * m.slope_i[size_t(i + m.pml)] *
                                              1.0f / ed;
            const float dz0 = std::sqrt(a * a + (1.0f / ed) * (1.0f / ed));
            dts = std::fmin(dts, 1.0f / (csum * vmax(i, j) * dz0));
        }
    return dts;
}
