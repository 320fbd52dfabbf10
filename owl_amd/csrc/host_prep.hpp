// Host-side preparation for OWL's acoustic solvers: FD coefficients,
// source/receiver grid location with Kaiser-windowed sinc weights, source
// time functions, the CFL bound and CFS-PML damping profiles.
//
// Everything here is evaluated in the reference's float32 arithmetic, in the
// order the Fortran writes it, so grid indices and PML masks come out
// bit-identical to OWL's and the weights/coefficients match to the last bit
// wherever OWL's FLIT helpers are the textbook functions.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace owl {

constexpr int kNkw = 4;  // half-width of the Kaiser-windowed sinc (module_source_receiver.f90:146)

// Staggered-grid coefficients c_1..c_R for R = fdhalf in 2..8
// (src/lib/module_fdcoef.f90:24-71).  Returns false for other R.
bool fd_coefficients(int R, float* c);

// Fortran nint: round half away from zero.
int fortran_nint(float x);

// Nine Kaiser-windowed sinc weights for taps -4..4 around a point at
// fractional offset x (module_source_receiver.f90:789-828).
void kaiser_sinc(float x, float h[9]);

// A source or receiver located on one shot window
// (prepare_geometry, module_source_receiver.f90:556-785, uniform mesh).
struct Located {
    int gx = 1, gy = 1, gz = 1;  // nearest integer nodes (1-based)
    int hx = 1, hy = 1, hz = 1;  // nearest half nodes (1-based)
    float ix[9], iy[9], iz[9];   // weights on integer nodes
    float hxw[9], hyw[9], hzw[9];  // weights on half nodes
};
Located locate(int dims, float x, float y, float z, float ox, float oy, float oz, float dx,
               float dy, float dz);

// Wavelet kinds accepted by prepare_stf (module_source_receiver.f90:1036-1061).
enum Wavelet { kRicker = 0, kGaussian = 1, kGaussianDeriv = 2, kRickerDeriv = 3 };

// Number of STF samples: ceiling(2*f0factor/f0/dt) in float32 (:1027).
int stf_length(float f0, float dt, float f0factor);
// Normalised source time function and hnt = nint(nt_stf/2.0) (:1034-1097).
std::vector<float> source_time_function(int wavelet, float f0, float dt, float f0factor, int* hnt);

// The wavelet of a stress-drop (explosion) source in the elastic solvers:
// OWL sets time_integration = -1 (elastic_vhtiort_2d/module_vars.f90:476,
// elastic_tti_2d/module_vars.f90:689) and prepare_stf then takes
// deriv(stf)/dt (module_source_receiver.f90:1111-1124).  FLIT's deriv is
// taken as centred inside, one-sided at the ends; when the last sample keeps
// more than a tenth of the peak the wavelet is extended to nt samples with
// its last value (pad_array, edge replication).
std::vector<float> stf_time_derivative(const std::vector<float>& stf, float dt, int nt);

// dt_stable = 1/(sum|c| * vmax * sqrt(sum 1/d^2)) (acoustic_iso_2d/module_vars.f90:121-125).
float dt_stable(int R, int dims, float vmax, float dx, float dy, float dz);
// nt = nint(tmax/dt + 1) (acoustic_iso_2d/module_vars.f90:134).
int time_steps(float tmax, float dt);

// CFS-PML profile of one axis of n interior nodes with pml cells each side,
// entries for index i = -pml+1 .. n+pml stored at [i + pml - 1]
// (acoustic_iso_2d/module_cfspml.f90:37-133).  kappa is identically 1 in
// OWL (kmax = 1), so only a and b are produced; outside the slabs a = b = 0.
struct PmlAxis {
    std::vector<float> ai, bi;  // integer nodes, slab where i <= 0 or i >= n+1
    std::vector<float> ah, bh;  // half nodes m, slab where m <= 1 or m >= n+1
};
PmlAxis cfspml_axis(int n, int pml, float d, float dt, float pmlvp, float f0max);

// Integer-node / half-node PML masks as used by the update sweeps
// (acoustic_iso_2d/module_wavefield.f90:58-65, 79-82, 96-99).
// M-PML damping ratio of a 2D anisotropic medium at one boundary cell
// (elastic_tti_2d/module_cfspml.f90:164-219): c = c11, c13, c15, c33, c35,
// c55; dirx selects the x (left/right) or z (top/bottom) direction.
float mpml_damp_ratio(const float* c, float rho, bool dirx);

inline bool pml_mask_int(int i, int n) { return i <= 0 || i >= n + 1; }
inline bool pml_mask_half(int m, int n) { return m <= 1 || m >= n + 1; }

}  // namespace owl
