/*
 * lmw.h -- C ABI of the MI355X (gfx950) dynamical core: the drop-in replacement of lmweather's weather_core
 * library (reference CMakeLists.txt:24: cpp/sphere.cpp + cpp/plane.cpp) below the wx.* operator registrations
 * (reference cpp/main.cpp:108-495).
 *
 * Conventions
 *   - Every function returns lmw_status.  LMW_EINVAL stands for std::invalid_argument, LMW_ERUNTIME for
 *     std::runtime_error; lmw_last_error() returns the message (thread-local), byte-identical to the reference's
 *     exception text.  LMW_EDEVICE reports a HIP failure.  No C++ exception crosses this boundary.
 *   - All arrays are HOST pointers in the reference's layouts: grids are nlat x nlon row-major, south to north
 *     (cpp/sphere.hpp:88); spectral arrays are (N+1)^2 complex numbers stored as interleaved (re, im) doubles at
 *     index m*(N+1)+n (cpp/sphere.hpp:107,110; entries with n < m are zero); plane h and u are ny x nx and v is
 *     (ny+1) x nx (cpp/plane.hpp:32-37).  The model state stays resident in device memory between calls.
 *   - Calls are synchronous: results are complete on return.  Handles are not thread-safe (one caller thread
 *     per handle, like the single-threaded reference).
 */
#ifndef LMW_H
#define LMW_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif
#if defined(__GNUC__) || defined(__clang__)
#pragma GCC visibility push(default) /* the library builds with hidden visibility: export exactly these */
#endif

typedef int lmw_status;
#define LMW_OK 0
#define LMW_EINVAL 1
#define LMW_ERUNTIME 2
#define LMW_EDEVICE 3

/* arithmetic of the time step's transforms.  LMW_FP64: everything in FP64.  The three mixed modes change the two
 * Legendre contractions (below) AND the grid-point kernel between them: its longitude FFTs, the nonlinear products
 * and the Bernoulli function run in FP32 on FP32 Fourier rows.  The state, the RK4 update, the Legendre tables
 * (built in FP64, then rounded or split), the Gauss nodes and weights, set-up, diagnostics and the conservation
 * sums stay FP64 in every mode. */
#define LMW_FP64 0
#define LMW_FP32 1   /* FP32 tables and operands, FP32 accumulation */
#define LMW_BF16X3 2 /* split BF16 hi+lo operands, three MFMA products, FP32 accumulation */
#define LMW_BF16X6 3 /* operands split in three BF16 parts (FP32-exact), six MFMA products, FP32 accumulation */

/* grid fields of lmw_sphere_get_grid (reference grid_height/grid_u/grid_v/grid_vorticity/grid_tracer,
 * cpp/sphere.hpp:91-95,55) */
#define LMW_GRID_H 0
#define LMW_GRID_U 1
#define LMW_GRID_V 2
#define LMW_GRID_VORTICITY 3
#define LMW_GRID_TRACER 4
#define LMW_GRID_OROGRAPHY 5 /* the orography grid as set (grid_orography, cpp/sphere.hpp:50) */

/* spectral fields (reference zeta_, delta_, phi_, q_: cpp/sphere.hpp:109) */
#define LMW_SPEC_ZETA 0
#define LMW_SPEC_DELTA 1
#define LMW_SPEC_PHI 2
#define LMW_SPEC_Q 3

/* scalar diagnostics of lmw_sphere_diag (reference cpp/sphere.hpp:70-88,55-62) */
#define LMW_DIAG_MEAN_HEIGHT 0
#define LMW_DIAG_ENERGY 1
#define LMW_DIAG_KINETIC 2
#define LMW_DIAG_POTENTIAL 3
#define LMW_DIAG_ENSTROPHY 4
#define LMW_DIAG_MAX_WIND 5
#define LMW_DIAG_ERROR_L2 6
#define LMW_DIAG_ERROR_LINF 7
#define LMW_DIAG_TIME 8
/* tracer_mean/tracer_mass/tracer_min/tracer_max/tracer_error_l2/tracer_error_linf: cpp/sphere.hpp:58-64 */
#define LMW_DIAG_TRACER_MEAN 9
#define LMW_DIAG_TRACER_MASS 10
#define LMW_DIAG_TRACER_MIN 11
#define LMW_DIAG_TRACER_MAX 12
#define LMW_DIAG_TRACER_ERROR_L2 13
#define LMW_DIAG_TRACER_ERROR_LINF 14

/* analytic initial states of lmw_sphere_setup / lmw_ensemble_setup: the test cases of Williamson et al. (1992,
 * J. Comput. Phys. 102:211-224) behind the Sphere.williamson1/2/5/6 operators (reference cpp/main.cpp:174-291) */
#define LMW_SETUP_WILLIAMSON1 1 /* cosine bell tracer in solid-body rotation about an axis tilted by alpha; advect-only */
#define LMW_SETUP_WILLIAMSON2 2 /* steady geostrophic solid-body rotation about an axis tilted by alpha */
#define LMW_SETUP_WILLIAMSON5 5 /* zonal flow over a conical mountain (sets the orography) */
#define LMW_SETUP_WILLIAMSON6 6 /* Rossby-Haurwitz wave of wavenumber 4 */

/* point evaluators of lmw_sphere_point (reference height_at/wind_at/vorticity_at: cpp/sphere.hpp:83-86) */
#define LMW_POINT_HEIGHT 0
#define LMW_POINT_U 1
#define LMW_POINT_V 2
#define LMW_POINT_VORTICITY 3
#define LMW_POINT_TRACER 4 /* tracer_at: cpp/sphere.hpp:60 / sphere.cpp:287-299 */

/* plane fields and diagnostics (reference cpp/plane.hpp:32-56) */
#define LMW_PLANE_H 0
#define LMW_PLANE_U 1
#define LMW_PLANE_V 2
#define LMW_PDIAG_MEAN_DEPTH 0
#define LMW_PDIAG_ENERGY 1
#define LMW_PDIAG_MAX_SPEED 2
#define LMW_PDIAG_PV_ANOMALY_RMS 3
#define LMW_PDIAG_TIME 4

/* reference SphereParams, cpp/sphere.hpp:20-25 */
typedef struct {
  double radius;    /* m */
  double omega;     /* rad/s */
  double gravity;   /* m/s2 */
  double diffusion; /* del^4 coefficient, m4/s (0 = none) */
} lmw_sphere_params;

/* reference PlaneParams, cpp/plane.hpp:9-18 */
typedef struct {
  int nx, ny;
  double lx, ly;
  double gravity;
  double depth;
  double f0, beta;
  double viscosity;
  double drag;
  int linear;
} lmw_plane_params;

typedef struct lmw_sphere lmw_sphere;
typedef struct lmw_plane lmw_plane;
typedef struct lmw_ensemble lmw_ensemble;
typedef struct lmw_corrector lmw_corrector;

/* ------------------------------------------------------------------------------------------- library */
const char *lmw_last_error(void);
/* device name, compute-unit count and device bytes held by live handles */
lmw_status lmw_device_info(char *name, size_t name_len, int *compute_units, size_t *bytes_in_use);

/* ------------------------------------------------------------------------------------------- free functions */
/* replaces fft(std::vector<cplx>&, int sign), cpp/sphere.hpp:14 / sphere.cpp:21-45: `count` independent
 * transforms of length n (a power of two, 2..4096), in place, unnormalized, sign -1 forward, +1 backward */
lmw_status lmw_fft(double *re_im, size_t n, int count, int sign);
/* replaces gauss_legendre(int, x, w), cpp/sphere.hpp:16 / sphere.cpp:47-64 (computed on the GPU) */
lmw_status lmw_gauss_legendre(int n, double *x, double *w);

/* ------------------------------------------------------------------------------------------- sphere */
/* replaces Sphere::Sphere(int truncation, const SphereParams&), cpp/sphere.hpp:29 / sphere.cpp:75-103 */
lmw_status lmw_sphere_create(int truncation, const lmw_sphere_params *params, lmw_sphere **out);
lmw_status lmw_sphere_destroy(lmw_sphere *s);
/* truncation(), nlat(), nlon(): cpp/sphere.hpp:30-32 */
lmw_status lmw_sphere_dims(const lmw_sphere *s, int *truncation, int *nlat, int *nlon);
/* arithmetic of the Legendre contractions (LMW_FP64 default) -- an addition, no reference counterpart */
lmw_status lmw_sphere_set_precision(lmw_sphere *s, int precision);
/* set_state_grid(h, u, v): cpp/sphere.hpp:39 / sphere.cpp:217-234 */
lmw_status lmw_sphere_set_state_grid(lmw_sphere *s, const double *h, const double *u, const double *v, size_t n);
/* set_orography_grid(hs): cpp/sphere.hpp:48 / sphere.cpp:105-109 */
lmw_status lmw_sphere_set_orography_grid(lmw_sphere *s, const double *hs, size_t n);
/* set_axis_tilt(alpha): cpp/sphere.hpp:40 / sphere.cpp:111-118 */
lmw_status lmw_sphere_set_axis_tilt(lmw_sphere *s, double alpha);
/* set_diffusion(nu4): cpp/sphere.hpp:41 */
lmw_status lmw_sphere_set_diffusion(lmw_sphere *s, double nu4);
/* mark_reference(): cpp/sphere.hpp:42 / sphere.cpp:236-239 */
lmw_status lmw_sphere_mark_reference(lmw_sphere *s);
/* add_mode(n, m, amplitude, mean_depth): cpp/sphere.hpp:43 / sphere.cpp:301-311 */
lmw_status lmw_sphere_add_mode(lmw_sphere *s, int n, int m, double amplitude, double mean_depth);
/* perturb(amplitude, seed, max_wave): cpp/sphere.hpp:44 / sphere.cpp:313-336 */
lmw_status lmw_sphere_perturb(lmw_sphere *s, double amplitude, uint64_t seed, int max_wave);
/* step(dt): cpp/sphere.hpp:63 / sphere.cpp:407-436 */
lmw_status lmw_sphere_step(lmw_sphere *s, double dt);
/* `count` steps of dt without the periodic checks of run() (the hot loop of sphere.cpp:471-472) */
lmw_status lmw_sphere_steps(lmw_sphere *s, double dt, int64_t count);
/* run(seconds, dt): cpp/sphere.hpp:64 / sphere.cpp:457-484 */
lmw_status lmw_sphere_run(lmw_sphere *s, double seconds, double dt);
/* stable_dt(safety): cpp/sphere.hpp:65 / sphere.cpp:438-447 */
lmw_status lmw_sphere_stable_dt(lmw_sphere *s, double safety, double *out);
/* grid_height/grid_u/grid_v/grid_vorticity/grid_tracer: cpp/sphere.hpp:91-94,55 / sphere.cpp:486-508 */
lmw_status lmw_sphere_get_grid(lmw_sphere *s, int field, double *out, size_t n);
/* the private spectral state zeta_, delta_, phi_ (cpp/sphere.hpp:109), reference layout */
lmw_status lmw_sphere_get_spectral(lmw_sphere *s, int field, double *re_im, size_t n_complex);
lmw_status lmw_sphere_set_spectral(lmw_sphere *s, int field, const double *re_im, size_t n_complex);
/* mean_height/energy/kinetic_energy/potential_energy/potential_enstrophy/max_wind/height_error_l2/linf/time:
 * cpp/sphere.hpp:70-81,34 / sphere.cpp:510-590 */
/* set_tracer_grid: cpp/sphere.hpp:53 / sphere.cpp:241-246 (single spheres; the tracer is stepped with the state by
 * lmw_sphere_step/steps/run, sphere.cpp:407-436); set_advect_only: sphere.hpp:55; has_tracer/advect_only: :54,56 */
lmw_status lmw_sphere_set_tracer_grid(lmw_sphere *s, const double *q, size_t n);
lmw_status lmw_sphere_set_advect_only(lmw_sphere *s, int on);
lmw_status lmw_sphere_tracer_flags(const lmw_sphere *s, int *has_tracer, int *advect_only);
lmw_status lmw_sphere_diag(lmw_sphere *s, int which, double *out);
/* spectrum(ke, pe): cpp/sphere.hpp:78 / sphere.cpp:536-551 (n = truncation + 1 entries each) */
lmw_status lmw_sphere_spectrum(lmw_sphere *s, double *ke, double *pe, size_t n);
/* height_at/wind_at/vorticity_at: cpp/sphere.hpp:83-86 / sphere.cpp:592-637 (lat, lon in radians) */
lmw_status lmw_sphere_point(lmw_sphere *s, int which, double lat, double lon, double *out);
/* mode_coefficient(n, m, imag), wave_crest(m, lat): cpp/sphere.hpp:87-88 / sphere.cpp:639-652 */
lmw_status lmw_sphere_mode_coefficient(lmw_sphere *s, int n, int m, int imag, double *out);
lmw_status lmw_sphere_wave_crest(lmw_sphere *s, int m, double lat, double *out);
/* set-up: one of the LMW_SETUP_* states, evaluated at the Gauss grid and analysed on the device (no host grids);
 * williamson1(alpha) / williamson2(alpha) set the axis tilt to alpha, williamson5() / williamson6() to 0, as the
 * reference's operators do (cpp/main.cpp:174-291).  LMW_EINVAL "angle must be a finite number" for a non-finite
 * alpha. */
lmw_status lmw_sphere_setup(lmw_sphere *s, int which, double alpha);
/* expression grids (the Sphere.fields / orography / tracer operators, cpp/main.cpp:158-172, :212-228, :252-268,
 * over the expression language of cpp/expr.hpp:16-166): NUL-terminated formulas in lon, lat (radians) and the
 * planet's a, omega, g, compiled on the host and evaluated at every Gauss grid point on the device.
 *   fields: h, u, v, then set_state_grid; orography: set_orography_grid; tracer: set_tracer_grid.
 * LMW_ERUNTIME with the reference's texts: 'expression "<text>": <what> at column <c>' for a string that does not
 * compile, "Sphere.fields: an expression is not finite on the grid" (h + u + v checked per point), "Sphere.orography:
 * the expression is not finite on the grid", "Sphere.tracer: the expression is not finite on the grid"; the state is
 * unchanged then. */
lmw_status lmw_sphere_fields(lmw_sphere *s, const char *h, const char *u, const char *v);
lmw_status lmw_sphere_orography(lmw_sphere *s, const char *hs);
lmw_status lmw_sphere_tracer(lmw_sphere *s, const char *q);
/* text output, "%.10g" per number (the Sphere.save_csv / save_spectrum / save_zonal_mean operators,
 * cpp/main.cpp:357-371, :328-335, :373-383): rows "lon lat h u v zeta q hs" in degrees with a blank line after each
 * latitude row; "n ke pe" for n = 1..T; "lat mean_h mean_u" per latitude.  LMW_ERUNTIME "<who>: cannot write
 * <path>" when the file cannot be written (main.cpp:86-92). */
lmw_status lmw_sphere_save_csv(lmw_sphere *s, const char *path);
lmw_status lmw_sphere_save_spectrum(lmw_sphere *s, const char *path);
lmw_status lmw_sphere_save_zonal_mean(lmw_sphere *s, const char *path);
/* the transforms exposed for tests: synthesize_scalar, analyze_scalar, analyze_divergence (cpp/sphere.hpp:97-100,
 * sphere.cpp:138-177) and winds_spectral_to_grid (sphere.cpp:179-206) */
lmw_status lmw_sphere_synthesize(lmw_sphere *s, const double *spec, double *grid);
lmw_status lmw_sphere_analyze(lmw_sphere *s, const double *grid, double *spec);
lmw_status lmw_sphere_analyze_divergence(lmw_sphere *s, const double *A, const double *B, double *spec);
lmw_status lmw_sphere_winds(lmw_sphere *s, const double *zeta, const double *delta, double *U, double *V);
/* one evaluation of tendencies(zeta, delta, phi) on the current state: cpp/sphere.cpp:338-390 */
lmw_status lmw_sphere_tendencies(lmw_sphere *s, double *dzeta, double *ddelta, double *dphi);
/* gauss_mu()/gauss_weights() (cpp/sphere.hpp:95-96) and the tables P_, H_ [lat][m][n-m] (cpp/sphere.hpp:108),
 * returned as nlat x (N+1) x (N+1) doubles */
lmw_status lmw_sphere_gauss(lmw_sphere *s, double *mu, double *w);
lmw_status lmw_sphere_get_tables(lmw_sphere *s, double *P, double *H);

/* ------------------------------------------------------------------------------------------- ensembles */
/* An addition with no single reference counterpart: `members` independent Sphere objects (as many independent
 * wx.sphere objects in one run, reference tests/determinism.lms:3-6) with possibly different truncations,
 * stepped together.  Members of one truncation share one set of tables and one kernel launch per stage. */
lmw_status lmw_ensemble_create(int members, const int *truncations, const lmw_sphere_params *params, int precision, lmw_ensemble **out);
lmw_status lmw_ensemble_destroy(lmw_ensemble *e);
lmw_status lmw_ensemble_member_dims(const lmw_ensemble *e, int member, int *truncation, int *nlat, int *nlon);
lmw_status lmw_ensemble_set_state_grid(lmw_ensemble *e, int member, const double *h, const double *u, const double *v, size_t n);
/* set_state_grid of members member0 .. member0 + count - 1 (one truncation) from grids [count][nlat][nlon] each, in
 * batched launches (no reference counterpart: the ensemble set-up path); n = count * nlat * nlon */
lmw_status lmw_ensemble_set_state_grids(lmw_ensemble *e, int member0, int count, const double *h, const double *u, const double *v, size_t n);
/* lmw_sphere_setup for members member0 .. member0 + count - 1 in batched launches; LMW_SETUP_WILLIAMSON1 needs a
 * tracer, which ensemble members do not carry (LMW_EINVAL) */
lmw_status lmw_ensemble_setup(lmw_ensemble *e, int member0, int count, int which, double alpha);
/* lmw_sphere_fields / lmw_sphere_orography for members member0 .. member0 + count - 1: each truncation group's grid
 * is evaluated once and analysed in batched launches; members given one orography share one device plane.  A
 * non-finite value in a later group leaves the earlier groups set. */
lmw_status lmw_ensemble_fields(lmw_ensemble *e, int member0, int count, const char *h, const char *u, const char *v);
lmw_status lmw_ensemble_orography(lmw_ensemble *e, int member0, int count, const char *hs);
lmw_status lmw_ensemble_set_orography_grid(lmw_ensemble *e, int member, const double *hs, size_t n);
lmw_status lmw_ensemble_set_axis_tilt(lmw_ensemble *e, int member, double alpha);
lmw_status lmw_ensemble_set_spectral(lmw_ensemble *e, int member, int field, const double *re_im, size_t n_complex);
lmw_status lmw_ensemble_get_spectral(lmw_ensemble *e, int member, int field, double *re_im, size_t n_complex);
lmw_status lmw_ensemble_get_grid(lmw_ensemble *e, int member, int field, double *out, size_t n);
lmw_status lmw_ensemble_diag(lmw_ensemble *e, int member, int which, double *out);
/* advance every member by `count` steps of its truncation's dt (dts[t] for the t-th distinct truncation in
 * ascending order; lmw_ensemble_groups lists them) */
lmw_status lmw_ensemble_groups(const lmw_ensemble *e, int *ngroups, int *truncations, int *members_per_group);
lmw_status lmw_ensemble_steps(lmw_ensemble *e, const double *dts, const int64_t *counts);
/* replay captured HIP graphs per group (8 RK4 steps, then single steps for the rest; hipGraph path); counts per group */
lmw_status lmw_ensemble_steps_graph(lmw_ensemble *e, const double *dts, const int64_t *counts);
/* `count` macro steps: in each, group g advances substeps[g] steps of dts[g] (e.g. 600 s = 1 x 600 s at T42,
 * 2 x 300 s at T85, 4 x 150 s at T170).  One macro step, and a run of 4 of them, are captured once as HIP graphs
 * whose branches are the groups' streams, so the truncations run concurrently; the host replays the 4-step graph,
 * then the single one for the rest (no reference counterpart) */
lmw_status lmw_ensemble_macro_steps(lmw_ensemble *e, const double *dts, const int *substeps, int64_t count);
/* persistent macro steps (mixed-precision ensembles of at most 4 truncation groups whose sub-step counts divide the
 * largest): in each stage, one persistent kernel walks a static work list of the inverse Legendre transforms of
 * every group due a step, the groups' grid-point kernels follow, then one persistent kernel for the forward
 * transforms; captured as one HIP graph per macro step like the default mode (no reference counterpart) */
lmw_status lmw_ensemble_set_persistent(lmw_ensemble *e, int on);
/* Tile shapes of the Legendre kernels, chosen per truncation group by timing every variant on the group's state
 * (restored afterwards) and cached in $LMW_TUNE_CACHE (default ~/.cache/lmweather_amd/tiles_v1.txt) keyed by
 * architecture, truncation, members and precision; force != 0 re-times.  tiles[4 g + 0..3] = variant indices of
 * the FP64 inverse, FP64 forward, mixed inverse and mixed forward transforms of group g; best_ms[2 g + 0..1] = the
 * chosen inverse and forward transform times (0 when taken from the cache).  Untuned groups use the mixed row chunk
 * 32 and latitude chunk 32 with 4 waves, and for FP64 the FP64 MFMA kernels from 32 members on, else the FMA
 * kernels with 4 members per block.  No reference counterpart. */
lmw_status lmw_ensemble_autotune(lmw_ensemble *e, int force, int *tiles, double *best_ms);
lmw_status lmw_sphere_autotune(lmw_sphere *s, int force, int *tiles, double *best_ms);
/* set the four tile variant indices of group g directly (tests, or a caller's own tuning) */
lmw_status lmw_ensemble_set_tiles(lmw_ensemble *e, int group, const int *tiles);
/* the shape of tile variant `index`: kind 0 = FP64 inverse members per block (a); 1 = mixed inverse row chunk (a);
 * 2 = mixed forward latitude chunk (a) and waves per block (b); 3 = FP64 forward: members per block of the FMA
 * kernel (a, b = 0) or the FP64 MFMA kernel with b waves per block (a = 0) */
lmw_status lmw_tile_variants(int kind, int index, int *a, int *b);
/* the same steps launched eagerly with a HIP event pair around every kernel, on the stream the kernels run on;
 * kernel_ms[3 g + 0..2] = total milliseconds of the inverse Legendre, grid-point (FFT + products) and forward
 * Legendre kernels of group g (measurement entry point, no reference counterpart) */
lmw_status lmw_ensemble_steps_timed(lmw_ensemble *e, const double *dts, const int64_t *counts, double *kernel_ms);

/* ------------------------------------------------------------------------------------------- learned corrector */
/* No reference counterpart: a spectral-space network (4 dense layers of width 8..64 over the packed coefficients
 * (m, n) of zeta, delta, phi, with per-degree mixing in the two inner layers; BF16 weights, FP32 accumulation)
 * that corrects a coarse large-step integration toward a high-resolution model truncated to the same triangle.
 * Spectra cross this boundary as float [sample][3][C][2] with C = (T+1)(T+2)/2 in the packed triangular order
 * m-major, index m (T+1) - m (m-1)/2 + n - m.  The plain spectral model stays the fallback. */
lmw_status lmw_corrector_create(int truncation, int width, uint64_t seed, lmw_corrector **out);
lmw_status lmw_corrector_destroy(lmw_corrector *c);
lmw_status lmw_corrector_dims(const lmw_corrector *c, int *truncation, int *width, size_t *nparams);
/* the FP32 master parameters: W1 [W][8], b1 [W], W2 [W][W], b2 [W], V2 [W][W], W3 [W][W], b3 [W], V3 [W][W],
 * W4 [6][W], b4 [6]; set_params also resets the Adam moments */
lmw_status lmw_corrector_get_params(lmw_corrector *c, float *out, size_t n);
lmw_status lmw_corrector_set_params(lmw_corrector *c, const float *in, size_t n);
/* per-degree scales [3][T+1] (zeta, delta, phi): inputs are divided by in_scale, outputs multiplied by out_scale */
lmw_status lmw_corrector_set_scales(lmw_corrector *c, const float *in_scale, const float *out_scale, size_t n);
/* mean squared error (trained entries: all but the (0,0) row and the imaginary parts at m = 0) of the network on
 * nsamples inputs against target corrections, and its gradient (grad may be NULL) */
lmw_status lmw_corrector_loss_grad(lmw_corrector *c, const float *x, const float *target, int nsamples, double *loss, float *grad);
/* the correction for nsamples inputs, and the rms of the network's outputs per sample (residual may be NULL) */
lmw_status lmw_corrector_forward(lmw_corrector *c, const float *x, int nsamples, float *correction, double *residual);
/* lmw_corrector_forward on DEVICE arrays, allocating nothing: no counterpart in the reference.  For callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  x and
 * correction (and residual, which may be NULL) are device pointers of the sizes lmw_corrector_forward takes; the
 * activations live in a static 32 MiB device workspace per GPU, held until the process ends, so nsamples is limited
 * to what fits (T42, width 64: 40 samples); LMW_EINVAL says how many bytes a larger call would need.  Synchronous;
 * calls that overlap on one GPU share the workspace and race. */
lmw_status lmw_corrector_forward_noalloc(lmw_corrector *c, const float *x, int nsamples, float *correction, double *residual);
/* the training set, resident on the device: inputs and target corrections */
lmw_status lmw_corrector_set_dataset(lmw_corrector *c, const float *x, const float *target, int nsamples);
/* `epochs` passes of Adam (lr, beta 0.9/0.999) over the training set in shuffled mini-batches; losses[e] = mean
 * loss of epoch e */
lmw_status lmw_corrector_train(lmw_corrector *c, int epochs, int batch, float lr, uint64_t seed, double *losses);
/* one corrected step of dt: coarse_substeps RK4 steps of dt / coarse_substeps, then the correction; a member whose
 * residual exceeds `threshold` (or is not finite) is stepped from its start state with fallback_substeps plain RK4
 * steps of dt / fallback_substeps instead.  used_fallback and residual (per member, may be NULL). */
lmw_status lmw_sphere_corrected_step(lmw_sphere *s, lmw_corrector *c, double dt, int coarse_substeps, int fallback_substeps, double threshold,
                                     int *used_fallback, double *residual);
lmw_status lmw_ensemble_corrected_step(lmw_ensemble *e, lmw_corrector *c, double dt, int coarse_substeps, int fallback_substeps,
                                       double threshold, int *used_fallback, double *residual);
/* the spectral state of every member of a single-truncation ensemble truncated to T (a coefficient copy: the
 * normalization does not depend on the truncation), float [member][3][C_T][2]; and setting it (zero above T) */
lmw_status lmw_ensemble_get_truncated(lmw_ensemble *e, int truncation, float *out, size_t n);
lmw_status lmw_ensemble_set_truncated(lmw_ensemble *e, int truncation, const float *in, size_t n);

/* ------------------------------------------------------------------------------------------- plane */
/* replaces Plane::Plane(const PlaneParams&), cpp/plane.hpp:22 / plane.cpp:20-25; `batch` independent members
 * share the parameters; precision LMW_FP64 or LMW_FP32 (state and arithmetic) */
lmw_status lmw_plane_create(const lmw_plane_params *params, int batch, int precision, lmw_plane **out);
lmw_status lmw_plane_destroy(lmw_plane *p);
/* params() setters (cpp/plane.hpp:24, the Plane.gravity/depth/coriolis/linear/viscosity/drag operators of
 * cpp/main.cpp:387-430); nx, ny, lx, ly are fixed at creation */
lmw_status lmw_plane_set_params(lmw_plane *p, const lmw_plane_params *params);
/* reset(): cpp/plane.hpp:39 / plane.cpp:27-32 (all members) */
lmw_status lmw_plane_reset(lmw_plane *p);
/* perturb(amplitude, seed): cpp/plane.hpp:40 / plane.cpp:34-57 (one member) */
lmw_status lmw_plane_perturb(lmw_plane *p, int member, double amplitude, uint64_t seed);
/* set_h/u/v_from_grid: cpp/plane.hpp:57-59 / plane.cpp:312-318 and the accessors h/u/v(i, j) */
lmw_status lmw_plane_set_field(lmw_plane *p, int member, int field, const double *g, size_t n);
/* Plane.fields(h, u, v) (cpp/main.cpp:410-431): formulas in x, y and lx, ly, H, g, f0, beta, evaluated on the
 * device: h = H + h(x, y) at the cell centres, u at the west faces, v at the interior south faces (the wall rows stay
 * 0).  LMW_ERUNTIME with the expression compiler's texts, or "Plane.fields: the height|u|v expression is not
 * finite" (state unchanged). */
lmw_status lmw_plane_fields(lmw_plane *p, int member, const char *h, const char *u, const char *v);
lmw_status lmw_plane_get_field(lmw_plane *p, int member, int field, double *out, size_t n);
/* step(dt): cpp/plane.hpp:41 / plane.cpp:143-158 (all members) */
lmw_status lmw_plane_step(lmw_plane *p, double dt);
lmw_status lmw_plane_steps(lmw_plane *p, double dt, int64_t count);
/* lmw_plane_steps, also reporting the time of the step kernels alone (HIP events on the plane's stream) in
 * *kernel_ms (no reference counterpart: instrumentation for the benchmark) */
lmw_status lmw_plane_steps_timed(lmw_plane *p, double dt, int64_t count, double *kernel_ms);
/* run(seconds, dt): cpp/plane.hpp:42 / plane.cpp:160-189 (all members; the limit is the smallest over members) */
lmw_status lmw_plane_run(lmw_plane *p, double seconds, double dt);
/* stable_dt(safety): cpp/plane.hpp:43 / plane.cpp:131-141 (the smallest over members) */
lmw_status lmw_plane_stable_dt(lmw_plane *p, double safety, double *out);
/* mean_depth/energy/max_speed/potential_vorticity_anomaly_rms/time: cpp/plane.hpp:45-48,25 / plane.cpp:191-243 */
lmw_status lmw_plane_diag(lmw_plane *p, int member, int which, double *out);
/* the plane's step kernel (no reference counterpart): the row march, one launch per RK4 step with each wavefront
 * walking a 56-column strip in registers, with rows per wavefront chosen from the occupancy (0, default) or
 * 1 = 32, 2 = 64, 3 = 128, 4 = 256 rows; or 5 = four stage launches; all bit-identical */
lmw_status lmw_plane_set_tile(lmw_plane *p, int tile);
/* one evaluation of tendency(h, u, v) for one member: cpp/plane.cpp:59-129 */
lmw_status lmw_plane_tendency(lmw_plane *p, int member, double *dh, double *du, double *dv);
/* the parameters (Plane::params(), cpp/plane.hpp:24), members and precision of a plane; any pointer may be NULL */
lmw_status lmw_plane_get_params(lmw_plane *p, lmw_plane_params *params, int *batch, int *precision);
/* zonal harmonics of one member (cpp/plane.hpp:52-55 / plane.cpp:264-310): the amplitude 2|c_m| and a crest
 * x in [0, lx/m) of harmonic m of h on row j, of v on v row j (1..ny-1), and the kinetic energy per unit mass in
 * harmonic m averaged over the channel; LMW_EINVAL "zonal harmonic or row out of range" etc. as the reference */
lmw_status lmw_plane_zonal_amplitude(lmw_plane *p, int member, int m, int j, double *out);
lmw_status lmw_plane_zonal_crest(lmw_plane *p, int member, int m, int j, double *out);
lmw_status lmw_plane_v_zonal_crest(lmw_plane *p, int member, int m, int j, double *out);
lmw_status lmw_plane_harmonic_energy(lmw_plane *p, int member, int m, double *out);
/* text output of one member, "%.10g" per number (the Plane.save_csv / save_row / save_column operators,
 * cpp/main.cpp:464-493): rows "x y h u v zeta" at the cell centres with a blank line after each row (u, v averaged
 * to the centre, zeta the mean of the four corners); "x h" along row `row`; "y h" along column `column` */
lmw_status lmw_plane_save_csv(lmw_plane *p, int member, const char *path);
lmw_status lmw_plane_save_row(lmw_plane *p, int member, int row, const char *path);
lmw_status lmw_plane_save_column(lmw_plane *p, int member, int column, const char *path);

#if defined(__GNUC__) || defined(__clang__)
#pragma GCC visibility pop
#endif
#ifdef __cplusplus
}
#endif
#endif
