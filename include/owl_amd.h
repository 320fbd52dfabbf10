/*
 * owl_amd.h -- C ABI of the MI355X implementation of OWL's acoustic
 * staggered-grid finite-difference solvers and their FWI gradient.
 *
 * Plain C types only: pointers to host memory, sizes, and opaque handles.
 * Every entry point returns OWL_OK (0) or a negative owl_status; the message
 * of the last failure on the calling thread is returned by owl_last_error().
 * Nothing here exits the process (OWL's solver calls `stop` instead).
 *
 * Array conventions are those of OWL's Fortran callers:
 *   models      vp(nz, nx) / rho(nz, nx), column-major, z fastest;
 *               3D: vp(nz, ny, nx), z fastest, then y, then x;
 *   traces      per shot [nr][nt] (receiver-major, as the SU traces of
 *               type(su) hold them: seis_p%trace(ir)%data(1:nt));
 *   gradients   (nz, nx) column-major, the array written to
 *               shot_<id>_grad_<name>.grd / grad_<name>.bin.
 * Reference file:line citations are relative to OWL's source tree.
 */
#ifndef OWL_AMD_H
#define OWL_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int owl_status;
#define OWL_OK 0
#define OWL_ERR_ARG (-1)       /* invalid argument or unsupported option */
#define OWL_ERR_CFL (-2)       /* dt above the stability bound (OWL resets it) */
#define OWL_ERR_DEVICE (-3)    /* HIP runtime failure */
#define OWL_ERR_NOMEM (-4)     /* device allocation failed */
#define OWL_ERR_STATE (-5)     /* call out of order (e.g. adjoint before forward) */
#define OWL_ERR_NONFINITE (-6) /* the wavefield became NaN or infinite (OWL stops with a
                                  NaN check in verbose mode, module_wavefield.f90) */

/* Source mechanisms and wavelets (module_source_receiver.f90:76-80, 1036-1061). */
#define OWL_EXPLOSION 0
#define OWL_FORCE 1
#define OWL_RICKER 0
#define OWL_GAUSSIAN 1
#define OWL_GAUSSIAN_DERIV 2
#define OWL_RICKER_DERIV 3

/* One source: type(source), module_source_receiver.f90:56-109.  polar and
 * azimuth are in radians (load_geometry converts degrees, :257-258). */
typedef struct owl_source {
    float x, y, z;
    float f0, amp, t0;
    float polar, azimuth;
    int mechanism;
    int wavelet;
} owl_source;

/* One receiver: type(receiver), module_source_receiver.f90:29-53. */
typedef struct owl_receiver {
    float x, y, z;
    float weight;
} owl_receiver;

/* One shot: type(source_receiver_geometry) (module_source_receiver.f90:
 * 111-142) plus the shot window that set_adaptive_range computes
 * (module_utility.f90:410-481): 1-based inclusive node ranges in the full
 * model.  A range of 0..0 means the full model along that axis. */
typedef struct owl_shot {
    int id;
    int ns;
    const owl_source *sources;
    int nr;
    const owl_receiver *receivers;
    int nxbeg, nxend;
    int nybeg, nyend;
    int nzbeg, nzend;
} owl_shot;

/* Solver configuration: the components of type(wave_solver_acoustic_iso_2d)
 * and _3d that are not arrays (acoustic_iso_2d/interface.f90:30-53). */
typedef struct owl_acoustic_config {
    int nx, ny, nz;          /* full model; ny = 1 in 2D */
    float dx, dy, dz;
    float ox, oy, oz;
    float dt, tmax;          /* nt = nint(tmax/dt + 1) */
    float data_dt, data_tmax;/* must equal dt, tmax: resample afterwards with owl_resample_traces */
    int pml;                 /* npml, default 15 (module_model.f90:174) */
    int fdhalf;              /* R: 2..8; OWL's -D_fdorder<2R>_ */
    int free_surface;        /* 2D only */
    float f0factor;          /* default 1.0 (module_source_receiver.f90:134) */
    int cc_step_interval;    /* imaging every k-th reverse step, default 1 */
    float trace_discard_threshold; /* module_parameters.f90:306, 570; OWL's default 1.0e-6.
                                      A trace is left out of the misfit and adjoint source when
                                      its norm is below this fraction of the largest one
                                      (module_inversion_adjoint_source.f90:1042-1054); 0 keeps
                                      every trace with a non-zero norm */
} owl_acoustic_config;

/* Gradient algorithms. */
#define OWL_GRAD_OWL 0    /* OWL's algorithm: boundary-saving reconstruction
                             with -dt (acoustic_iso_2d/submodule_fwi_gradient.f90:75-167) */
#define OWL_GRAD_EXACT 1  /* exact discrete adjoint of the forward scheme */

/* Forward-state storage for OWL_GRAD_EXACT. */
#define OWL_CKPT_FULL 0      /* keep every step's state */
#define OWL_CKPT_UNIFORM 1   /* snapshots every k steps, segment recompute */
#define OWL_CKPT_BINOMIAL 2  /* binomial (revolve) schedule with a snapshot budget */
#define OWL_STORE_F32 0
#define OWL_STORE_F16 1      /* with a per-snapshot scale */
#define OWL_STORE_BF16 2     /* with a per-snapshot scale */

typedef struct owl_acoustic2d owl_acoustic2d; /* opaque batch of 2D shots */

/* Version string of the library. */
const char *owl_version(void);
/* Message of the last failed call on this thread ("" if none). */
const char *owl_last_error(void);

/* Stable dt of the scheme (acoustic_iso_2d/module_vars.f90:121-125):
 * 1/(sum|c_m| * vmax * sqrt(1/dx^2 [+ 1/dy^2] + 1/dz^2)), float32. */
float owl_dt_stable(int fdhalf, int dims, float vmax, float dx, float dy, float dz);

/* Grid location and Kaiser-windowed sinc weights of one point, as
 * source_receiver_geometry%prepare_geometry computes them
 * (module_source_receiver.f90:556-785).  idx[6] = gx, gy, gz, hx, hy, hz
 * (1-based); w[6*9] = ix, iy, iz, hx, hy, hz weights for taps -4..4. */
owl_status owl_locate_point(int dims, float x, float y, float z, float ox, float oy, float oz,
                            float dx, float dy, float dz, int *idx, float *w);

/* Normalised source time function (prepare_stf, module_source_receiver.f90:
 * 1002-1097).  Returns the sample count; writes up to cap samples and hnt. */
int owl_source_time_function(int wavelet, float f0, float dt, float f0factor, float *stf,
                             int cap, int *hnt);

/* CFS-PML coefficients of one axis (compute_cfspml_damping_coef,
 * acoustic_iso_2d/module_cfspml.f90:61-133): four arrays of n + 2*pml
 * entries for i = -pml+1 .. n+pml (a, b on integer nodes; a, b on half
 * nodes).  kappa is 1 everywhere. */
owl_status owl_cfspml_axis(int n, int pml, float d, float dt, float pmlvp, float f0max, float *ai,
                           float *bi, float *ah, float *bh);

/* Parameters of the per-shot aperture window (module_parameters.f90:90-91,
 * 123-129, 419-420, 463-468; model extents module_model.f90:97-111).  2D:
 * ny = 1, dy = 1, yn_adpy = 0 and every y coordinate 0. */
typedef struct owl_window_params {
    int nx, ny, nz;                 /* full model */
    float dx, dy, dz;
    float ox, oy, oz;
    float xmin, xmax, ymin, ymax, zmin, zmax;   /* model extent (default ox .. ox+(nx-1)dx) */
    int yn_adpx, yn_adpy, yn_adpz;  /* adaptive range along each axis (default 0) */
    float adp_extrax, adp_extray, adp_extraz;   /* margin added around the shot (default 0) */
    float offset_min, offset_max;   /* absolute-offset limits (default 0, float_huge) */
} owl_window_params;

/* set_adaptive_range (module_utility.f90:410-510) together with the
 * absolute-offset rule of prepare_geometry (module_source_receiver.f90:
 * 421-436), float32 as OWL computes them: zeroes the weight of receivers
 * outside [offset_min, offset_max] and outside the window, and returns the
 * window as 1-based inclusive node ranges window[6] = nxbeg, nxend, nybeg,
 * nyend, nzbeg, nzend (for owl_shot) and, if range is not NULL, its
 * coordinates range[6] = xbeg, xend, ybeg, yend, zbeg, zend. */
owl_status owl_set_adaptive_range(const owl_window_params *p, int ns, const owl_source *src, int nr,
                                  owl_receiver *rec, int *window, float *range);

/* Trace resampling of su%resamp (src/lib/module_su.f90:425-468), which the
 * solvers call when data_dt/data_tmax differ from dt/tmax
 * (acoustic_iso_2d/submodule_forward_modeling.f90:124-125) and on the
 * adjoint source (submodule_fwi_gradient.f90:46): in[nr][nt] sampled at
 * k dt to out[nr][nnt] at oot + k ddt by FLIT's interp 'cubic' (taken as
 * the natural cubic spline, end cubics extended); all-zero traces stay
 * zero.  Host code; in and out may not alias. */
owl_status owl_resample_traces(int nr, int nt, float dt, const float *in, int nnt, float ddt, float oot,
                               float *out);

/* FLIT's ginterp(x, y, xi, 'linear' | 'cubic') on increasing x, as
 * load_geometry calls it on the topography to drop sources and receivers
 * above the free surface (module_source_receiver.f90:302-316). */
owl_status owl_ginterp(int n, const float *x, const float *y, int m, const float *xi, float *yi, int cubic);

/* ---- 2D acoustic: batched shots on one GPU --------------------------------
 * Replaces the per-shot loops of forward_modeling
 * (src/forward/2d/module_modeling.f90:66-95) and compute_gradient_shots
 * (src/inverse/2d/module_gradient.f90:93-388) for which_medium =
 * 'acoustic-iso'.  vp, rho: full model (nz, nx), z fastest.  Every shot runs
 * on its own window; all shots of the batch advance together on the GPU. */
owl_status owl_acoustic2d_create(const owl_acoustic_config *cfg, const float *vp,
                                 const float *rho, int nshot, const owl_shot *shots,
                                 owl_acoustic2d **out);
void owl_acoustic2d_destroy(owl_acoustic2d *b);
/* Number of steps nt and the per-shot receiver counts. */
int owl_acoustic2d_nt(const owl_acoustic2d *b);
/* Replace the model (same shape) without re-planning. */
owl_status owl_acoustic2d_set_model(owl_acoustic2d *b, const float *vp, const float *rho);

/* Forward modelling of every shot (compute_forward,
 * acoustic_iso_2d/submodule_forward_modeling.f90:31-132).  traces: shot
 * after shot, each [nr][nt].  save_boundary != 0 keeps the data OWL's
 * gradient needs (reconstruct = .true.). */
owl_status owl_acoustic2d_forward(owl_acoustic2d *b, float *traces, int save_boundary);

/* Forward modelling that allocates no device memory: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Same arguments as
 * owl_acoustic2d_create followed by owl_acoustic2d_forward(..., traces, 0),
 * and the same traces.  Every device buffer of the call is carved from a
 * static device array of 32 MiB per GPU (held until the process ends): a
 * run whose fields, model, tables and traces need more returns
 * OWL_ERR_NOMEM.  Host-side planning uses ordinary host memory.  Calls on one
 * GPU share the array and are serialised within a process; the array keeps
 * the last call's contents, and every call clears what must start at zero. */
owl_status owl_acoustic2d_forward_static(const owl_acoustic_config *cfg, const float *vp,
                                         const float *rho, int nshot, const owl_shot *shots,
                                         float *traces);

/* FWI gradient over all shots.  obs: observed traces, laid out like the
 * forward traces.  Computes the L2 misfit and the adjoint source
 * syn - obs (compute_adjsrc_waveform,
 * module_inversion_adjoint_source.f90:371-383, with the trace-discard rule
 * of :1042-1054), back-propagates it, and sums the per-shot window
 * gradients into the full model (module_gradient.f90:366-413).
 *   mode      OWL_GRAD_OWL | OWL_GRAD_EXACT
 *   ckpt      OWL_CKPT_* (EXACT only); ckpt_param = snapshot interval
 *             (UNIFORM) or snapshot budget (BINOMIAL)
 *   store     OWL_STORE_* for snapshots (EXACT only)
 *   misfit    per-shot misfit values [nshot] (FP64), may be NULL
 *   grad_vp, grad_rho: (nz, nx) z fastest, summed over shots (FP64).
 * traces_out (may be NULL): the synthetic traces of the same run. */
owl_status owl_acoustic2d_gradient(owl_acoustic2d *b, const float *obs, int mode, int ckpt,
                                   int ckpt_param, int store, double *misfit, double *grad_vp,
                                   double *grad_rho, float *traces_out);
/* energy_precond of OWL's gradient (on != 0): the OWL-algorithm gradient of
 * each shot window divided by sqrt(source energy x receiver energy) of its
 * imaging terms, each energy raised by 1e-3 of its largest value
 * (acoustic_iso_2d/submodule_fwi_gradient.f90:173-184, 223-265).  Applies to
 * OWL_GRAD_OWL and owl_acoustic2d_adjoint; default off. */
owl_status owl_acoustic2d_set_energy_precond(owl_acoustic2d *b, int on);
/* Misfit only: forward modelling of every shot and the per-shot L2 misfit
 * with the trace-discard rule, no adjoint (compute_gradient_shots with
 * yn_misfit_only = .true., src/inverse/2d/module_gradient.f90:93-388, as
 * the step-size trials of module_inversion_step_size.f90:562-587 call it).
 * obs may be NULL to keep the observed data already set. */
owl_status owl_acoustic2d_misfit(owl_acoustic2d *b, const float *obs, double *misfit, float *traces_out);

/* Differentiable path (exact adjoint): model every shot and keep the forward
 * states the adjoint needs under the given storage policy (OWL_CKPT_*,
 * ckpt_param, OWL_STORE_*); then owl_acoustic2d_backward applies the
 * transpose of the forward map to dtraces (dL/dtraces, laid out like the
 * traces) and returns dL/dvp, dL/drho on the full model (nz, nx), z fastest,
 * summed over shots.  No OWL counterpart (OWL's gradient is the
 * reconstruction algorithm of OWL_GRAD_OWL); the forward map is OWL's
 * compute_forward.  Explosion sources, no free surface. */
owl_status owl_acoustic2d_forward_store(owl_acoustic2d *b, float *traces, int ckpt, int ckpt_param,
                                        int store);
owl_status owl_acoustic2d_backward(owl_acoustic2d *b, const float *dtraces, double *grad_vp,
                                   double *grad_rho);
/* The differentiable path with device-resident tensors (no OWL
 * counterpart; for torch.autograd): the same calls with every array in
 * device memory of the batch's GPU, so nothing is staged through the host.
 * The caller completes its writes of vp, rho and dtraces before the call;
 * each call returns after its work is complete.  set_model_device replaces
 * the model as owl_acoustic2d_set_model does (vp, rho: (nz, nx) z fastest,
 * or x fastest when x_fastest != 0), with the CFL check and the PML
 * profiles of the new edge velocities, without planning the batch again;
 * explosion sources only.  forward_store_device writes the traces in the
 * trace layout; backward_device takes dL/dtraces in that layout and writes
 * dL/dvp, dL/drho as float (nz, nx) (z fastest, or x fastest). */
owl_status owl_acoustic2d_set_model_device(owl_acoustic2d *b, const float *vp, const float *rho, int x_fastest);
owl_status owl_acoustic2d_forward_store_device(owl_acoustic2d *b, float *traces, int ckpt, int ckpt_param,
                                               int store);
owl_status owl_acoustic2d_backward_device(owl_acoustic2d *b, const float *dtraces, float *grad_vp,
                                          float *grad_rho, int x_fastest);
/* Identifier of the forward states that the batch holds for
 * owl_acoustic2d_backward: a new positive value after every
 * owl_acoustic2d_forward_store, 0 when none are held (after set_model, a
 * gradient call, or before any forward_store).  A caller that keeps several
 * differentiable evaluations compares it before calling backward.  No OWL
 * counterpart. */
owl_status owl_acoustic2d_store_id(const owl_acoustic2d *b, long long *id);

/* Adjoint pass alone, OWL's algorithm, on data from a preceding
 * owl_acoustic2d_forward(..., save_boundary = 1): adjoint source traces adj
 * (laid out like the forward traces) in, per-shot window gradients out
 * (compute_adjoint, acoustic_iso_2d/submodule_fwi_gradient.f90:31-199).
 * grad_vp, grad_rho: shot after shot, each (nz_s, nx_s) of the shot's
 * window, z fastest: the arrays of shot_<id>_grad_vp.grd / _rho.grd. */
owl_status owl_acoustic2d_adjoint(owl_acoustic2d *b, const float *adj, float *grad_vp,
                                  float *grad_rho);

/* Window size of shot s: writes nx_s, nz_s (and the number of receivers). */
owl_status owl_acoustic2d_shot_dims(const owl_acoustic2d *b, int s, int *nx, int *nz, int *nr);

/* Keep observed traces resident on the device (laid out like the forward
 * traces); owl_acoustic2d_gradient(b, NULL, ...) then uses them.  This is
 * the data OWL's compute_adjoint_source reads from dir_field
 * (src/inverse/2d/module_gradient.f90:321-327). */
owl_status owl_acoustic2d_set_observed(owl_acoustic2d *b, const float *obs);
/* Model the batch's current model and keep the traces as observed data
 * (synthetic inversion tests generate d_obs this way). */
owl_status owl_acoustic2d_observe(owl_acoustic2d *b);

/* Device timing of the last forward/gradient call (ms) and the number of
 * stencil grid-point updates it performed. */
owl_status owl_acoustic2d_last_stats(const owl_acoustic2d *b, double *ms, double *updates);

/* Forward-step kernel of the exact-adjoint paths (no OWL counterpart;
 * results are bit-identical either way): mode 0 two sweeps (p, then v)
 * and the two-pass adjoint, 1 one fused z-marching kernel per forward step
 * and per adjoint step, 2 fused for batches of 4 shots or more
 * (default).  zrows: z rows per fused tile (default 16).  The fused kernel needs
 * distinct input and output states and no free surface; elsewhere the two
 * sweeps run. */
owl_status owl_acoustic2d_set_step_kernel(owl_acoustic2d *b, int mode, int zrows);
/* z rows per tile of the fused adjoint step (default 8; no OWL counterpart,
 * results are bit-identical for any value). */
owl_status owl_acoustic2d_set_adjoint_rows(owl_acoustic2d *b, int zrows);
/* Autotuner of the fused kernels (no OWL counterpart): on (default), the
 * first exact gradient times z-row candidates of the fused forward (8..32)
 * and adjoint (4..16) steps on the batch's buffers and keeps the fastest.
 * set_step_kernel and set_adjoint_rows turn it off; turning it on again
 * re-tunes at the next exact gradient. */
owl_status owl_acoustic2d_set_autotune(owl_acoustic2d *b, int on);
/* The z rows in use by the fused forward and adjoint steps. */
owl_status owl_acoustic2d_fused_rows(const owl_acoustic2d *b, int *zrows, int *adjoint_rows);

/* HIP graphs for owl_acoustic2d_gradient in OWL_GRAD_EXACT mode (no OWL
 * counterpart; results are bit-identical): on (default), the second call
 * with the same schedule and buffers captures the whole gradient (storage
 * schedule, steps, residual, fold) into a graph that later calls replay;
 * calls with profiling on run eagerly.  Turning it off drops the graph. */
owl_status owl_acoustic2d_set_graphs(owl_acoustic2d *b, int on);

/* Per-kernel timing with HIP events on the batch's stream.  No OWL
 * counterpart (OWL has no profiler); used by the benchmark.  kind:
 * 0 p-sweep and 1 v-sweep over one field set (forward), 2 source/receiver
 * step, 3 imaging condition, 4 boundary injection, 5 adjoint-source
 * injection, 6 misfit, 7 merge, 8 p-sweep and 9 v-sweep over two field sets
 * (reverse pass: reconstruction + adjoint), 10 exact-adjoint p-pass,
 * 11 exact-adjoint v-pass, 12 state storage, 13 fused forward step,
 * 14 fused exact-adjoint step, 0 <= kind < 16.
 * Totals since profiling was switched on: device time, launch count, and
 * algorithmic (compulsory) bytes of those launches (0 if not modelled). */
owl_status owl_acoustic2d_set_profiling(owl_acoustic2d *b, int on);
owl_status owl_acoustic2d_kernel_stats(const owl_acoustic2d *b, int kind, double *total_ms,
                                       long long *launches, double *bytes_total);

/* ---- 3D acoustic: batched shots on one GPU --------------------------------
 * OWL's acoustic_iso_3d solver without domain decomposition, for
 * which_medium = 'acoustic-iso' in src/forward/3d/module_modeling.f90 and
 * src/inverse/3d/module_gradient.f90.  vp, rho: (nz, ny, nx) z fastest.
 * Explosion and force sources; free surface at the top when
 * cfg->free_surface (every window then starts at z index 1).  Traces per
 * shot [nr][nt]. */
typedef struct owl_acoustic3d owl_acoustic3d;
owl_status owl_acoustic3d_create(const owl_acoustic_config *cfg, const float *vp,
                                 const float *rho, int nshot, const owl_shot *shots,
                                 owl_acoustic3d **out);
void owl_acoustic3d_destroy(owl_acoustic3d *b);
int owl_acoustic3d_nt(const owl_acoustic3d *b);
/* Forward modelling of every shot (compute_forward,
 * acoustic_iso_3d/submodule_forward_modeling.f90:31-183; update_wavefield
 * module_wavefield.f90:40-317, add_source :322-444).  traces may be NULL.
 * save_boundary != 0 keeps the boundary records and the final state that
 * OWL's gradient needs (reconstruct = .true.,
 * acoustic_iso_3d/module_boundary_saving.f90), in device memory. */
owl_status owl_acoustic3d_forward(owl_acoustic3d *b, float *traces, int save_boundary);
/* Adjoint pass alone, OWL's algorithm (compute_adjoint,
 * acoustic_iso_3d/submodule_fwi_gradient.f90:33-274, compute_gradient
 * :276-361, energy preconditioning off), after a forward with
 * save_boundary = 1.  adj: adjoint source traces laid out like the forward
 * traces.  grad_vp, grad_rho: shot after shot, each (nz_s, ny_s, nx_s) of the
 * shot's window, z fastest: the arrays of shot_<id>_grad_vp.grd / _rho.grd. */
owl_status owl_acoustic3d_adjoint(owl_acoustic3d *b, const float *adj, float *grad_vp,
                                  float *grad_rho);
/* FWI gradient over all shots: forward with boundary saving, L2 misfit and
 * adjoint source syn - obs (module_inversion_adjoint_source.f90:371-383,
 * trace-discard rule :1042-1054), OWL's reverse pass, and the sum of the
 * window gradients on the full model (nz, ny, nx), z fastest, FP64.
 * obs: observed traces (kept on the device; NULL reuses the last ones).
 * misfit[nshot], grad_vp, grad_rho and traces_out may be NULL. */
owl_status owl_acoustic3d_gradient(owl_acoustic3d *b, const float *obs, double *misfit,
                                   double *grad_vp, double *grad_rho, float *traces_out);
/* Launch shape of the sweeps (results do not depend on it; no OWL
 * counterpart): z planes one block marches over, and the kernel variant
 * (0: x-y planes staged in LDS; 1, 2, 3: barrier-free column segments with
 * 4, 8, 16 rows per block, 5-7 the same with an XCD-contiguous tile order;
 * 8: four-wide columns (float4 per thread), 24: the same with the
 * XCD-contiguous order; 9, 10, 11: software-pipelined columns with 4, 8, 16
 * rows per block, 13-15 the same with the XCD-contiguous order). */
owl_status owl_acoustic3d_set_launch(owl_acoustic3d *b, int planes, int variant);
/* HIP graph of the forward time loop (no OWL counterpart): with on != 0
 * (default), forward modelling without boundary records and with profiling
 * off runs the whole time loop as one captured graph, captured by the first
 * such call and replayed while the launch shape and buffers stay the same;
 * results are bit-identical to eager launches.  on = 0 drops the graph and
 * launches eagerly. */
owl_status owl_acoustic3d_set_graphs(owl_acoustic3d *b, int on);
/* Launch-shape autotuner (no OWL counterpart): times `steps` sweeps of the
 * batch under each candidate (variant, planes) and keeps the fastest;
 * writes the chosen shape (either pointer may be NULL).  Leaves the wave
 * state zeroed; results do not depend on the shape. */
owl_status owl_acoustic3d_autotune(owl_acoustic3d *b, int steps, int *planes, int *variant);
/* Timing as for the 2D batch; kinds: 0 p-sweep, 1 v-sweep, 2 source and
 * receivers, 3 imaging condition, 4 boundary records, 5 adjoint-source
 * injection. */
owl_status owl_acoustic3d_set_profiling(owl_acoustic3d *b, int on);
owl_status owl_acoustic3d_kernel_stats(const owl_acoustic3d *b, int kind, double *total_ms,
                                       long long *launches, double *bytes_total);
owl_status owl_acoustic3d_last_stats(const owl_acoustic3d *b, double *ms, double *updates);

/* ---- 3D isotropic elastic: batched shots on one GPU -----------------------
 * OWL's elastic_vhtiort_3d forward solver for anisotropy_type = 'iso'
 * (compute_forward, elastic_vhtiort_3d/submodule_forward_modeling.f90:31-140;
 * update_wavefield module_wavefield.f90:52-545, free surface :549-1090,
 * add_source :1095-1200; CFS-MPML module_cfspml.f90:428-625) for
 * which_medium = 'elastic-iso' in src/forward/3d/module_modeling.f90.
 * vp, vs, rho: (nz, ny, nx) z fastest; cfg as for the acoustic solvers
 * (free_surface: flat free surface at z index 1 on a uniform mesh).  Force
 * sources (polar, azimuth).  Traces: shot after shot, each three blocks
 * (vx, vy, vz) of [nr][nt]. */
typedef struct owl_elastic3d owl_elastic3d;
owl_status owl_elastic3d_create(const owl_acoustic_config *cfg, const float *vp, const float *vs,
                                const float *rho, int nshot, const owl_shot *shots,
                                owl_elastic3d **out);
void owl_elastic3d_destroy(owl_elastic3d *b);
int owl_elastic3d_nt(const owl_elastic3d *b);
/* save_boundary != 0 keeps the boundary records of every step and the
 * final state for owl_elastic3d_adjoint (OWL's forward part of
 * compute_gradient, elastic_vhtiort_3d/module_boundary_saving.f90:62-236). */
owl_status owl_elastic3d_forward(owl_elastic3d *b, float *traces, int save_boundary);
/* OWL's reverse pass for the isotropic elastic gradient (compute_adjoint and
 * compute_gradient, elastic_vhtiort_3d/submodule_fwi_gradient.f90:31-827,
 * chain rule :486-495): adj in the trace layout (three components) is
 * injected into the adjoint velocities below the free surface; the
 * wavefield runs backwards from the final state with the boundary records;
 * per-shot window gradients grad_vp, grad_vs, grad_rho (nz_s, ny_s, nx_s)
 * z fastest, shot after shot (the .grd arrays), NaN/Inf -> 0.  Needs a
 * preceding owl_elastic3d_forward(..., save_boundary = 1). */
owl_status owl_elastic3d_adjoint(owl_elastic3d *b, const float *adj, float *grad_vp, float *grad_vs,
                                 float *grad_rho);
/* Forward with boundary saving, L2 waveform adjoint source per component
 * (module_inversion_adjoint_source.f90:1020-1125; discard threshold 1e-6 of
 * each component's largest trace norm), reverse pass, and the sum of the
 * shots' gradients on the full model (module_gradient.f90:366-388), FP64.
 * misfit: one value per shot (sum over the three components). */
owl_status owl_elastic3d_gradient(owl_elastic3d *b, const float *obs, double *misfit, double *grad_vp,
                                  double *grad_vs, double *grad_rho, float *traces_out);
/* Launch shape (no OWL counterpart; results do not depend on it): z planes
 * per block and thread rows per block of 64 threads (1, 2 or 4). */
owl_status owl_elastic3d_set_launch(owl_elastic3d *b, int planes, int rows);
/* Times `steps` steps of the stress and velocity sweeps under each launch
 * shape (8 .. 256 z planes, 1, 2 or 4 y rows per tile) on this batch,
 * keeps the fastest and returns it (planes, rows may be NULL).  No
 * reference counterpart. */
owl_status owl_elastic3d_autotune(owl_elastic3d *b, int steps, int *planes, int *rows);
/* As owl_fsg2d_set_graphs, for this batch's forward loop. */
owl_status owl_elastic3d_set_graphs(owl_elastic3d *b, int on);
/* As owl_fsg2d_set_record_storage. */
owl_status owl_elastic3d_set_record_storage(owl_elastic3d *b, int kind);
/* Timing as for the acoustic batches; kinds: 0 stress sweep, 1 velocity
 * sweep, 2 sources and receivers, 3 imaging, 4 boundary records. */
owl_status owl_elastic3d_set_profiling(owl_elastic3d *b, int on);
owl_status owl_elastic3d_kernel_stats(const owl_elastic3d *b, int kind, double *total_ms,
                                      long long *launches, double *bytes_total);
owl_status owl_elastic3d_last_stats(const owl_elastic3d *b, double *ms, double *updates);

/* ---- 2D isotropic elastic: batched shots on one GPU -----------------------
 * OWL's elastic_vhtiort_2d solver for anisotropy_type = 'iso' (compute_forward,
 * elastic_vhtiort_2d/submodule_forward_modeling.f90:31-180; update_wavefield
 * module_wavefield.f90:42-134; CFS-MPML module_cfspml.f90:285-578) and its
 * gradient (compute_adjoint / compute_gradient, submodule_fwi_gradient.f90;
 * boundary records module_boundary_saving.f90:53-160), for which_medium =
 * 'elastic-iso' in src/forward/2d/module_modeling.f90 and
 * src/inverse/2d/module_gradient.f90.  vp, vs, rho: (nz, nx) z fastest; cfg
 * as for the acoustic solvers with ny = 1; free_surface = 1 runs OWL's
 * free surface on its depth-refined mesh (owl_elastic2d_create_surface).
 * Force (polar) and explosion sources.  Traces: shot after shot, two blocks
 * (vx, vz) of [nr][nt]. */
typedef struct owl_elastic2d owl_elastic2d;
owl_status owl_elastic2d_create(const owl_acoustic_config *cfg, const float *vp, const float *vs,
                                const float *rho, int nshot, const owl_shot *shots,
                                owl_elastic2d **out);
/* OWL's 2D elastic free surface (elastic_vhtiort_2d with free_surface:
 * module_vars.f90:310-371, module_wavefield.f90:143-289, module_cfspml.f90:
 * 398-480): each shot steps on its depth-refined mesh (rows dz/dz_refine
 * apart over half a Rayleigh wavelength, then growing by 2.5 % per row up
 * to dz_max) with its medium mapped onto it; traction-free surface row,
 * odd stress mirrors, even velocity images for the receivers, sources and
 * adjoint sources on rows >= 2, sources above the first row moved to it
 * with amplitude x 1.2, gradients mapped back to the regular window.  The
 * surface is flat at z = oz of each window. */
typedef struct owl_elastic_surface {
    float dz_refine;  /* free_surface_dz_refine, OWL's default 4.0 (<= 0: 4.0) */
    float dz_max;     /* largest z spacing; OWL's default 1.5 dz (<= 0: 1.5 dz) */
} owl_elastic_surface;
/* As owl_elastic2d_create with that free surface (cfg->free_surface = 1 in
 * owl_elastic2d_create takes OWL's defaults). */
owl_status owl_elastic2d_create_surface(const owl_acoustic_config *cfg, const float *vp, const float *vs,
                                        const float *rho, int nshot, const owl_shot *shots,
                                        const owl_elastic_surface *surface, owl_elastic2d **out);
/* Rows shot `shot` steps (its window's nz, or the refined mesh's). */
int owl_elastic2d_rows(const owl_elastic2d *b, int shot);
void owl_elastic2d_destroy(owl_elastic2d *b);
int owl_elastic2d_nt(const owl_elastic2d *b);
/* save_boundary != 0 keeps the boundary records of every step and the final
 * state for owl_elastic2d_adjoint. */
owl_status owl_elastic2d_forward(owl_elastic2d *b, float *traces, int save_boundary);
/* OWL's reverse pass: adj (trace layout, two components) into the adjoint
 * velocities; per-shot window gradients grad_vp, grad_vs, grad_rho
 * (nz_s, nx_s) z fastest, shot after shot (the .grd arrays; chain rule
 * submodule_fwi_gradient.f90:330-360), NaN/Inf -> 0.  Needs a preceding
 * owl_elastic2d_forward(..., save_boundary = 1). */
owl_status owl_elastic2d_adjoint(owl_elastic2d *b, const float *adj, float *grad_vp, float *grad_vs,
                                 float *grad_rho);
/* Forward with boundary saving, L2 waveform adjoint source per component
 * (module_inversion_adjoint_source.f90:1020-1125), reverse pass, sum of the
 * shots' gradients on the full model (module_gradient.f90:366-388), FP64.
 * misfit: one value per shot (sum over the two components). */
owl_status owl_elastic2d_gradient(owl_elastic2d *b, const float *obs, double *misfit, double *grad_vp,
                                  double *grad_vs, double *grad_rho, float *traces_out);
/* Launch shape (no OWL counterpart; results do not depend on it): z rows
 * per tile and 64-thread tiles per block (1, 2 or 4). */
owl_status owl_elastic2d_set_launch(owl_elastic2d *b, int planes, int rows);
/* As owl_fsg2d_autotune and owl_fsg2d_set_graphs, for this batch's stress
 * and velocity sweeps and forward loop.  No reference counterpart. */
owl_status owl_elastic2d_autotune(owl_elastic2d *b, int steps, int *planes, int *rows);
owl_status owl_elastic2d_set_graphs(owl_elastic2d *b, int on);
/* As owl_fsg2d_set_record_storage. */
owl_status owl_elastic2d_set_record_storage(owl_elastic2d *b, int kind);
/* Timing as for the acoustic batches; kinds: 0 stress sweep, 1 velocity
 * sweep, 2 sources and receivers, 3 imaging, 4 boundary records. */
owl_status owl_elastic2d_set_profiling(owl_elastic2d *b, int on);
owl_status owl_elastic2d_kernel_stats(const owl_elastic2d *b, int kind, double *total_ms,
                                      long long *launches, double *bytes_total);
owl_status owl_elastic2d_last_stats(const owl_elastic2d *b, double *ms, double *updates);

/* ---- 2D general-anisotropic elastic on the Lebedev fully staggered grid ----
 * OWL's elastic_tti_2d solver (compute_forward,
 * elastic_tti_2d/submodule_forward_modeling.f90:30-300; update_wavefield
 * module_wavefield.f90:35-252; add_source :593-736; CFS-MPML with damping
 * ratios module_cfspml.f90:80-576) and its gradient (compute_adjoint /
 * compute_gradient, submodule_fwi_gradient.f90:40-727; boundary records
 * module_boundary_saving.f90:100-200), for which_medium = 'elastic-tti' with
 * anisotropy_type = 'cij' in src/forward/2d/module_modeling.f90 and
 * src/inverse/2d/module_gradient.f90.  cij: six arrays c11, c13, c15, c33,
 * c35, c55 and rho, each (nz, nx) z fastest.  aniso = 0 sets the M-PML
 * damping ratios to 0 (OWL's anisotropy_type = 'iso', module_cfspml.f90:
 * 344-352), 1 takes them from the Christoffel scan.  Force and explosion
 * sources; no free surface (OWL's FSG free surface runs on its FLIT-built
 * topography-following mesh).  Traces: shot after shot, two blocks (vx, vz)
 * of [nr][nt], each the half-and-half mean of the two velocity grids. */
/* ---- Misfits and adjoint sources -------------------------------------------
 * misfit_type of OWL's FWI (module_inversion_adjoint_source.f90). */
#define OWL_MISFIT_WAVEFORM 0  /* sum (obs - syn)^2; adjoint syn - obs (:371-383) */
#define OWL_MISFIT_CORR 1      /* zero-lag normalised correlation (:391-413) */
#define OWL_MISFIT_ENVELOPE 2  /* envelope difference (:424-452) */
#define OWL_MISFIT_PHASE 3     /* instantaneous-phase difference (:460-481) */
#define OWL_MISFIT_ADAPTIVE 4  /* adaptive waveform inversion, Warner & Guasch (:490-537) */
#define OWL_MISFIT_LOCAL_ADAPTIVE 5 /* localised AWI, Yong et al. (:545-625) */
#define OWL_MISFIT_ADAPTIVE_SPACETIME 6 /* AWI over groups of neighbouring traces (:629-701, 1171-1215) */

/* Lag penalties of the adaptive misfit (xcorr_penalty, :321-363). */
#define OWL_PENALTY_LINEAR 0
#define OWL_PENALTY_POWER 1
#define OWL_PENALTY_GAUSSIAN 2
#define OWL_PENALTY_EXP 3

/* Parameters of the adaptive misfit (module_parameters.f90:262-265,
 * 551-557). */
typedef struct owl_misfit_params {
    float dt;            /* trace sampling (s) */
    float tlag_max;      /* largest penalised lag (s); OWL's default 0.1 tmax */
    int penalty;         /* OWL_PENALTY_*; OWL's default linear */
    float penalty_power; /* default 1 */
    float deconv_eps;    /* Wiener-filter stabilisation relative to max |D|^2, default 0.1 */
    float lawi_sigma;    /* local adaptive: Gabor window standard deviation (s), default 0.25 */
    int adaptive_hw;     /* space-time adaptive: half width of a trace group, default 3 */
} owl_misfit_params;

/* Misfit and adjoint source of the ntr traces of one shot component, on the
 * GPU (compute_adjoint_source, module_inversion_adjoint_source.f90:985-1160,
 * without resampling): obs, syn and adj are [ntr][nt]; weight (may be NULL)
 * holds the receiver weights; threshold is trace_discard_threshold.  A trace
 * with zero weight, an obs or syn norm below threshold x the largest of the
 * shot, or a zero norm gets a zero adjoint and misfit.  misfit[ntr]: per-trace
 * misfit in FP64 (OWL sums them into the shot's misfit).  Envelope and phase
 * use the analytic signal of a length-nt FFT (FLIT's hilbert, envelope and
 * instant_phase are not in the reference tree). */
owl_status owl_adjoint_source(int type, int ntr, int nt, const float *obs, const float *syn,
                              const float *weight, float threshold, float *adj, double *misfit);
/* As owl_adjoint_source, with the parameters that OWL_MISFIT_ADAPTIVE and
 * OWL_MISFIT_LOCAL_ADAPTIVE need (p may be NULL for the other types).  The adaptive misfit per trace is
 * 0.5 sum (w xc)^2 / sum xc^2 of the Wiener matching filter xc of syn onto
 * obs (both tapered over nint(0.025 nt) samples at each end and padded to the
 * 2-3-5 length >= 2 nt), w the lag penalty; its adjoint source is the
 * gradient of that misfit (tapered at the end).  FLIT's taper, fft and
 * rescale are taken as a Hann half-window, the unnormalised forward and
 * 1/n inverse DFT, and the linear [min, max] -> [0, 1] map.  The local
 * adaptive misfit sums v_i^2 over one matching filter per sample i under the
 * Gabor window exp(-(k - i)^2 / (2 (lawi_sigma/dt)^2)) (FLIT's gaussian,
 * taken unnormalised), v_i = sum (w xc_i)^2 / (sum xc_i^2 + eta).  The
 * space-time adaptive misfit of trace i averages the filters of traces
 * i - adaptive_hw .. i + adaptive_hw whose obs spectrum is not zero (no
 * traces beyond the gather), pads to the 2-3-5 length >= 4 nt and
 * back-projects through trace i's spectrum; as OWL's driver, it applies the
 * receiver weights but not the discard threshold. */
owl_status owl_adjoint_source_ex(int type, const owl_misfit_params *p, int ntr, int nt, const float *obs,
                                 const float *syn, const float *weight, float threshold, float *adj,
                                 double *misfit);

/* Parameterisations of a 2D TTI medium (anisotropy_type of OWL's
 * elastic_tti_2d solver, module_vars.f90:284-340). */
#define OWL_ANISO_ISO 0      /* vp, vs, rho */
#define OWL_ANISO_THOMSEN 1  /* vp, vs, rho, epsilon, delta, tilt theta (radians) */
#define OWL_ANISO_AT 2       /* vp, vs, rho, epsilon, eta, tilt theta (Alkhalifah-Tsvankin) */

/* c11, c13, c15, c33, c35, c55 of n points from the parameters (host only,
 * elementwise, float32): 'iso' as in prepare_modeling
 * (elastic_tti_2d/module_vars.f90:287-299); thomsen_to_cij2d and
 * alkhalifah_tsvankin_to_cij2d (src/lib/module_anisotropy.f90:305-330,
 * 399-418) with the tilt applied as cij_rotate_2d(..., -theta) (:184-240).
 * p1, p2: epsilon and delta (THOMSEN) or epsilon and eta (AT); p1, p2,
 * theta may be NULL for ISO.  cij[0..5]: outputs. */
owl_status owl_tti2d_cij(int param, long long n, const float *vp, const float *vs, const float *rho,
                         const float *p1, const float *p2, const float *theta, float *const *cij);
/* Chain rule from the c_ij gradients of owl_fsg2d_adjoint/gradient to the
 * parameter gradients (elastic_tti_2d/submodule_fwi_gradient.f90:476-560,
 * derivative macros src/lib/macro_thomsen_2d.f90): with vp, vs rescaled to
 * km/s and rho to g/cm^3 as OWL does; grads[0..1] = vp, vs (ISO), grads[0..3] = vp, vs, p1,
 * p2 (THOMSEN, AT).  The AT derivatives are the exact ones (OWL's macros
 * write the exponent 3/2 as an integer division). */
owl_status owl_tti2d_param_gradient(int param, long long n, const float *vp, const float *vs,
                                    const float *rho, const float *p1, const float *p2,
                                    const float *theta, const float *const *grad_cij, float *const *grads);

typedef struct owl_fsg2d owl_fsg2d;
owl_status owl_fsg2d_create(const owl_acoustic_config *cfg, const float *const *cij, const float *rho,
                            int aniso, int nshot, const owl_shot *shots, owl_fsg2d **out);
/* A new medium (c11 .. c55 and rho as for owl_fsg2d_create) for the same
 * shots: the batch is planned again (free-surface meshes, damping ratios,
 * stability check, sources' rho) and keeps its stream, launch shape and
 * storage settings; boundary records of an earlier forward are dropped.
 * OWL's iterations compute each gradient on the current model
 * (src/inverse/2d/module_gradient.f90:57-417).  A failed call (OWL_ERR_CFL,
 * OWL_ERR_ARG) leaves the batch unusable: destroy it. */
owl_status owl_fsg2d_set_model(owl_fsg2d *b, const float *const *cij, const float *rho);
void owl_fsg2d_destroy(owl_fsg2d *b);
int owl_fsg2d_nt(const owl_fsg2d *b);
/* The damping-ratio profiles of one shot (compute_damping_profile_ratio,
 * module_cfspml.f90:222-276): left and right over j = -pml .. nz+pml+1
 * (nz + 2 pml + 2 values), top and bottom over i = -pml .. nx+pml+1. */
owl_status owl_fsg2d_damping_ratios(const owl_fsg2d *b, int shot, float *left, float *right, float *top,
                                    float *bottom);
/* damp_ratio (module_cfspml.f90:164-219) of one medium, host only: c =
 * c11, c13, c15, c33, c35, c55; dirx != 0 for the x direction (left and
 * right edges), 0 for z (top and bottom). */
owl_status owl_fsg2d_damp_ratio(const float *c, float rho, int dirx, float *ratio);
/* OWL reads the final state back with vz_hxiz named twice
 * (module_boundary_saving.f90:85-90): the saved vz_ixhz lands in vz_hxiz and
 * vz_ixhz stays zero.  on = 0 (default) keeps that order; on = 1 restores
 * every field (no OWL counterpart). */
owl_status owl_fsg2d_set_restore_all(owl_fsg2d *b, int on);
/* save_boundary != 0 keeps the boundary records and the final state for
 * owl_fsg2d_adjoint. */
owl_status owl_fsg2d_forward(owl_fsg2d *b, float *traces, int save_boundary);
/* OWL's reverse pass: adj (trace layout, two components) into both adjoint
 * velocity grids; grads[0..6] = per-shot window gradients of c11, c13,
 * c15, c33, c35, c55 and rho (nz_s, nx_s) z fastest, shot after shot (the
 * .grd arrays of anisotropy_type 'cij'), NaN/Inf -> 0; null entries are
 * skipped.  Needs a preceding owl_fsg2d_forward(..., save_boundary = 1). */
owl_status owl_fsg2d_adjoint(owl_fsg2d *b, const float *adj, float *const *grads);
/* Forward with boundary saving, L2 waveform adjoint source per component,
 * reverse pass, sum of the shots' gradients on the full model, FP64.
 * grads[0..6] as above on (nz, nx); misfit: one value per shot. */
owl_status owl_fsg2d_gradient(owl_fsg2d *b, const float *obs, double *misfit, double *const *grads,
                              float *traces_out);
/* Launch shape (no OWL counterpart; results do not depend on it). */
owl_status owl_fsg2d_set_launch(owl_fsg2d *b, int planes, int rows);
/* Times `steps` steps of the two sweeps under each launch shape (rows per
 * tile column 4 .. 128, 1, 2 or 4 tiles per block) on this batch, keeps the
 * fastest and returns it (planes, rows may be NULL).  No reference
 * counterpart: OWL runs on the CPU. */
owl_status owl_fsg2d_autotune(owl_fsg2d *b, int steps, int *planes, int *rows);
/* on != 0 (the default): modelling without profiling or boundary records
 * replays the forward time loop as one captured HIP graph, re-captured when
 * a launch argument changes; 0 launches eagerly. */
owl_status owl_fsg2d_set_graphs(owl_fsg2d *b, int on);
/* Storage of the boundary records that OWL's gradient saves each forward
 * step (elastic_tti_2d/module_boundary_saving.f90:55-200): OWL_STORE_F32 (OWL's, the default),
 * OWL_STORE_F16 or OWL_STORE_BF16 with one scale per (shot, step) record,
 * halving their HBM; takes effect at the next forward with boundary
 * records. */
owl_status owl_fsg2d_set_record_storage(owl_fsg2d *b, int kind);
/* Timing; kinds: 0 stress sweep, 1 velocity sweep, 2 sources and
 * receivers, 3 imaging, 4 boundary records. */
owl_status owl_fsg2d_set_profiling(owl_fsg2d *b, int on);
owl_status owl_fsg2d_kernel_stats(const owl_fsg2d *b, int kind, double *total_ms, long long *launches,
                                  double *bytes_total);
owl_status owl_fsg2d_last_stats(const owl_fsg2d *b, double *ms, double *updates);

/* OWL's free surface for the FSG solver (elastic_tti_2d: free_surface,
 * file_topo, topo_interp, free_surface_dz_refine, dz_max,
 * measure_source/receiver_depth_from_surface; interface.f90:39-68 and
 * module_model.f90:176-178). */
typedef struct owl_fsg_surface {
    int ntopo;              /* topography control points; 0: flat at elevation 0 */
    const float *topo_x;    /* x of each point (increasing) */
    const float *topo_z;    /* elevation of each point, positive up (OWL's file_topo columns) */
    int topo_interp;        /* 0 'linear', 1 'cubic' (OWL's default) */
    float dz_refine;        /* free_surface_dz_refine, OWL's default 4.0 */
    float dz_max;           /* largest z spacing; OWL's default 1.5 dz */
    int src_depth_from_surface, rec_depth_from_surface;
} owl_fsg_surface;

/* As owl_fsg2d_create, with OWL's topography-following free surface on the
 * depth-refined mesh (elastic_tti_2d/module_vars.f90:360-519, 532-567,
 * 639-680; module_wavefield.f90:257-736; module_cfspml.f90:395-480): each
 * shot window gets its own refined mesh, the medium mapped onto it, source
 * and receiver rows on it, and the traction-free update at the surface.
 * Gradients from owl_fsg2d_adjoint/gradient are mapped back to the regular
 * window (map_irregular_to_regular, :208-249) and have the shapes above.
 * surface = NULL is owl_fsg2d_create. */
owl_status owl_fsg2d_create_surface(const owl_acoustic_config *cfg, const float *const *cij, const float *rho,
                                    int aniso, int nshot, const owl_shot *shots, const owl_fsg_surface *surface,
                                    owl_fsg2d **out);
/* The refined mesh of one window, host only (the same construction as
 * owl_fsg2d_create_surface): c55, rho (nz, nx) z fastest for the window of
 * cfg; f0min the shot's smallest f0.  *nz: refined rows; scal: dz (the
 * spacing the update divides by), eta_max, depth_max, topo_max.  Row tables
 * over j = -pml .. nz+pml+1 (cap entries available; OWL_ERR_ARG with *nz
 * set when too small): eta_i, eta_h (clipped), sc_i, sc_h (dz scaling),
 * edz_i; column tables over i = -pml .. nx+pml+1: topo_i, slope_i, topo_h,
 * slope_h. */
/* Rows of one shot's mesh (the refined row count with the free surface,
 * else the window's); 0 for a bad shot index. */
int owl_fsg2d_rows(const owl_fsg2d *b, int shot);
owl_status owl_fsg2d_surface_mesh(const owl_acoustic_config *cfg, const float *c55, const float *rho, float f0min,
                                  const owl_fsg_surface *surface, int cap, int *nz, float *scal, float *eta_i,
                                  float *eta_h, float *sc_i, float *sc_h, float *edz_i, float *topo_i,
                                  float *slope_i, float *topo_h, float *slope_h);

#ifdef __cplusplus
}
#endif
#endif /* OWL_AMD_H */
