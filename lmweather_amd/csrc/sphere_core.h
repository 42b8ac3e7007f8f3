// SphereCore: B members of one truncation resident on the GPU, stepped by the spectral transform method of the
// reference (cpp/sphere.cpp:338-436) with the kernels in sphere_core.hip.
//
// Device data layout (all FP64 unless stated):
//   spectral state      x0/xs/acc [b][field][S] double2, packed triangle: entry (m, n) at moff[m] + (n - m),
//                        moff[m] = m (N+1) - m (m-1)/2, S = (N+1)(N+2)/2; fields zeta, delta, phi.
//   Legendre tables     P, H for the NORTHERN latitudes only (P(-mu) = (-1)^(n-m) P(mu), H(-mu) = -(-1)^(n-m) H(mu)).
//                        Within each m the rows are permuted: even n-m first (kpe[m] rows, padded to 16), then odd
//                        (kpo[m] rows); row r of m is table row poff[m] + r.
//                        "synthesis" layout Ps/Hs [row][jh] (jh fastest, stride Jhp);
//                        "analysis"  layout Pa/Ha [m][jh][r]  (r fastest, stride kpe+kpo).
//   Fourier coefficients Fs [b][j][m][4] double2: U, V (= u cos(lat), v cos(lat)), zeta, phi per latitude j.
//   analysis operands    Gb [b][m][jh][16] double2: X+(4), X-(4), Y+(4), Y-(4) where X multiplies P and Y multiplies H
//                        and +/- are the sums/differences of the northern and southern latitude of the pair jh.
//   grids                f, hs, ref_phi [b][j][k]; diagnostic grids [b][4][j][k] = h, u, v, zeta.
#pragma once
#include <memory>

#include "common.h"
#include "expr.h"

namespace lmw {

struct Geo {
  int N, M1, J, K, logK, Jh, Jhp, B, S, Tmax;
  double a, g, omega;
  const int *moff, *poff, *kpe, *kpo;
  // element strides of the spectral state arrays x0/xs/acc: entry (member b, field f, packed q) at
  // b sb + f sf + q sq.  Member-major [b][f][q] (sb = 3S, sf = S, sq = 1) for a single member; member-minor [f][q][b]
  // (sb = 1, sf = S B, sq = B) for groups of several members in every precision: the MFMA Legendre kernels map lanes
  // to members and so read and write 8 or 16 consecutive members as one contiguous segment (the FP64 FMA kernels
  // run on either layout).
  long long sb, sf, sq;
  __host__ __device__ size_t at(int b, int f, size_t q) const {
    return static_cast<size_t>(b * sb + f * sf + static_cast<long long>(q) * sq);
  }
};

enum AnalMode { ANAL_TEND = 0, ANAL_STATE = 1, ANAL_RAW = 2, ANAL_TENDOUT = 3 };
// PREP_TRACER: in3 = U q, V q, q delta -> o1 = D(U q, V q), o4 = the scalar analysis of q delta (sphere.cpp:250-263)
enum PrepMode { PREP_STATE = 0, PREP_DIV = 1, PREP_SCALAR = 2, PREP_TRACER = 3 };
enum GridMode { GRID_DIAG = 0, GRID_RAW = 1 };

// Tile variants of the Legendre kernels chosen per truncation group by the autotuner (autotune.hip): FP64
// members per block; mixed-precision synthesis row chunk; mixed-precision analysis (latitude chunk, waves).
// tile variants: FP64 synthesis (members per block, then the FP64 MFMA kernel with row chunks of 32 or 64), FP64
// analysis (members per block, then the FP64 MFMA kernel with 4 or 8 waves), mixed synthesis (row chunk), mixed
// analysis (latitude chunk, waves); the first kFp64Tiles FP64 variants are the FMA kernels' members per block
constexpr int kFp64Tiles = 3, kFp64SynthTiles = 5, kFp64AnalTiles = 5, kSynthTiles = 2, kAnalTiles = 6;
constexpr int kSetupChunk = 32;  // members per launch of the batched set-up path (bounds the staging buffer in3)
constexpr int SETUP_TRACER = 100;  // k_setup: test case 1's tracer bell (internal; not an LMW_SETUP_* value)
constexpr int kMaxGroups = 4;  // truncation groups served by one multi-group (persistent) Legendre launch
extern const int kFp64MB[kFp64Tiles];
extern const int kSynthRC[kSynthTiles];
extern const int kAnalKC[kAnalTiles], kAnalNW[kAnalTiles];

struct SphereCore {
  // geometry
  int N, M1, J, K, logK, Jh, Jhp, S, B, R, Tmax;
  lmw_sphere_params prm;
  int precision = LMW_FP64;
  hipStream_t stream = nullptr;
  // member parts of a large group stepped as concurrent branches of the step (parts of whole 16-member tiles): the
  // second and later parts run on these streams, forked from and joined to `stream` with these events
  static constexpr int kMaxParts = 2;  // (C3, 741 -> 746 member-days/s against 4 parts; 8 parts: 680)
  hipStream_t part_stream[kMaxParts] = {};
  hipEvent_t part_fork = nullptr, part_join[kMaxParts] = {};
  int step_parts() const;
  std::vector<int> h_moff, h_poff, h_kpe, h_kpo;
  // per member host state
  std::vector<double> t;                      // model time
  std::vector<double> tilt;                   // axis tilt
  std::vector<char> has_orography;
  // orography planes: member b reads plane hs_slot[b] of hs.  Members given bit-identical orography share one
  // plane (set_orography compares contents; test case 5's cone is one plane per group), so a batch over one terrain
  // reads it once.  hs_refs counts the members of each plane; hs_kind: 0 unknown, 1 set from the host (hs_key its
  // hash), 2 test case 5's cone.
  std::vector<int> hs_slot, hs_refs;
  std::vector<char> hs_kind;
  std::vector<uint64_t> hs_key;
  // device
  DevArray<int> moff, poff, kpe, kpo;
  DevArray<double> mu, w;
  DevArray<double> Ps, Hs, Pa, Ha;             // FP64 tables
  // mixed-precision operands of the Legendre kernels, rounded/split once from the FP64 tables (precision.hip)
  bool mixed_ready = false;
  DevArray<int> d_mrow, d_soff, d_ke32, d_ko32;
  DevArray<float> A32, Ps32, Hs32;
  DevArray<double> A64;  // FP64 forward-transform operand of the FP64 MFMA variant (built on first use)
  DevArray<__bf16> Abh, Abl, Sph, Spl, Shh, Shl;  // BF16 parts 1, 2 (BF16X3)
  DevArray<float> Abf, Spf, Shf;                  // the same FP32 values in that layout, split in three on load (BF16X6)
  DevArray<double2> tw;                        // exp(-2 pi i q / K), q < K
  DevArray<float2> tw32;                       // tw rounded to FP32 (mixed-precision grid kernel)
  DevArray<double> knt;                        // -a^2 / (n (n + 1)), n = 0 .. N + 1 (0 at n = 0): psi, chi from zeta, delta
  DevArray<double2> x0, xs, acc;               // [b][3][S]
  DevArray<double2> Fs;                        // [b][J][M1][4]
  DevArray<double2> Gb;                        // [b][M1][Jh][16]
  DevArray<double> f, hs, ref_phi;             // [b][J][K]
  DevArray<double> grid;                       // [b][4][J][K]
  DevArray<double> in3;                        // [b][3][J][K] input grids for the analysis path
  DevArray<double2> spec_io;                   // [b][4][S] raw transform outputs / inputs
  DevArray<double> red;                        // [b][16] reductions
  DevArray<double> rowred;                     // [b][9][J] per-row partial reductions
  DevArray<int> mflags;                        // [b] bit 0 untilted axis, bit 1 no orography, bits 2.. hs plane
  DevArray<float2> Fs32, Gb32;                 // FP32 storage of Fs and Gb for the mixed-precision step
  // passive tracer q (single-member spheres only): state, RK4 stage state and accumulator [S], the spectral input
  // of its transforms [3][S], the delta grid scratch [4][J][K] and the reference tracer grid [J][K]
  bool has_tracer = false, advect_only = false;
  DevArray<double2> q0, qs, qacc, tq;
  DevArray<double> grid2, ref_q;
  hipGraphExec_t graph = nullptr;       // one step
  hipGraphExec_t graph_many = nullptr;  // kGraphSteps steps
  static constexpr int kGraphSteps = 8;
  double graph_dt = 0;
  // tile variant indices (defaults: FP64 4 members per block, or the FP64 MFMA kernels for 32 members and more (set
  // in the constructor); RC 32, KC 32 with 4 waves: the mixed shapes the autotuner picks at T42, T85 and T170, and
  // the only forward shape with LDS-buffered records)
  int tile_fp64_synth = 1, tile_fp64_anal = 1, tile_synth = 0, tile_anal = 0;
  bool member_minor = false;  // layout of x0/xs/acc (see Geo)

  SphereCore(int truncation, const lmw_sphere_params &p, int members, int precision);
  ~SphereCore();
  SphereCore(const SphereCore &) = delete;
  SphereCore &operator=(const SphereCore &) = delete;

  Geo geo() const;
  void set_precision(int p);
  void sync();

  // set-up (member b)
  void set_axis_tilt(int b, double alpha);
  void update_flags(int b);
  void set_orography(int b, const double *hs_host);
  // members b0 .. b0 + nb - 1 get the orography hs_host (host bits, for the content key); a new plane is filled from
  // the device copy dev_src when given, else uploaded
  void set_orography_run(int b0, int nb, const double *hs_host, const double *dev_src);
  void point_orography(int b, int plane);
  int own_orography(int b);
  void set_state_grid(int b, const double *h, const double *u, const double *v);
  void set_state_grids(int b0, int nb, const double *h, const double *u, const double *v);
  void setup(int b0, int nb, int which, double alpha);
  void ensure_in3(int nb);
  void state_from_in3(int b0, int nb);
  void mark_reference_range(int b0, int nb);
  void tracer_from_in3();
  // expression grids (expr.hip): Sphere.fields / orography / tracer of cpp/main.cpp:158-172,212-228,252-268
  ExprDevice expr;
  DevArray<double> lonlat;  // [K] longitudes, [J] latitudes
  void upload_lonlat();
  bool eval_exprs(const ExprProgram *const progs[], int count, int nb);
  void fields_expr(int b0, int nb, const std::string &h, const std::string &u, const std::string &v);
  void orography_expr(int b0, int nb, const std::string &text);
  void tracer_expr(const std::string &text);
  void set_spectral_packed(int b, int field, const double2 *packed);  // host packed triangle
  void get_spectral_packed(int b, int field, double2 *packed);
  void mark_reference(int b);
  // tracer (member 0 of a single-member core): set_tracer_grid (sphere.cpp:241-246), advect-only switch
  void set_tracer_grid(const double *q);
  void set_tracer_spectral(const double2 *packed);
  void set_advect_only(bool on);
  void tracer_grid(double *out);  // host grid of q
  double tracer_abs_sum();
  void drop_graph();

  // stepping (all members)
  void launch_stage(int stage, double dt, int b0, int nb);
  void enqueue_step(double dt);
  void enqueue_steps(double dt, int nsteps);  // member parts forked once for all the steps
  void steps(double dt, int64_t count);
  void steps_graph(double dt, int64_t count);
  // eager steps with a HIP event pair around every launch; ms[0..2] += total milliseconds of the inverse Legendre,
  // grid-point and forward Legendre kernels
  void steps_timed(double dt, int64_t count, double ms[3]);

  // diagnostics
  void compute_grids(int b0, int nb);                 // grid <- h, u, v, zeta of x0
  void reduce(int b0, int nb);                        // red <- sums over grid
  void download_grid(int b, int field, double *out);  // after compute_grids
  void download_red(int b, double *out16);
  double spectral_abs_sum(int b);

  // raw transforms on member 0's scratch (exposed for tests)
  void synthesize_raw(const double2 *z, const double2 *d, const double2 *phi, double *U, double *V, double *Z, double *PHI);
  void analyze_raw(int mode, const double *g0, const double *g1, const double *g2, double2 *out4);
  void tendencies_raw(int b, double2 *dz, double2 *dd, double2 *df);

  // launches
  // scratch: spec is a contiguous [member][3][S] buffer (spec_io, tq), not the state arrays
  void launch_synth(const double2 *spec, int b0, int nb, bool scratch = false);
  void launch_tend_grid(int b0, int nb, int mixed = 0);
  void launch_anal(int mode, int stage, double dt, int b0, int nb, double2 *raw_out);
  void launch_prep(int mode, int b0, int nb);
  void launch_to_grid(int mode, double *out, int b0, int nb);
  void launch_tracer_tendency();                     // spec_io[0] = D(U q, V q), spec_io[3] = analysis of q delta
  void launch_tracer_update(int stage, double dt);   // RK4 stage update of q from spec_io
  void launch_state_filter(double dt);               // del^4 filter of zeta, delta, phi (advect-only steps)
  void build_mixed_tables();
  void build_fp64_mfma_tables();
  void launch_anal_fp64_mfma(int nw, int mode, int stage, double dt, int b0, int nb, double2 *raw_out);
  void launch_synth_fp64_mfma(int rc, const double2 *spec, int b0, int nb, bool scratch);
  void launch_synth_mx(const double2 *spec, int b0, int nb);
  void launch_anal_mx(int mode, int stage, double dt, int b0, int nb, double2 *raw_out);

  size_t packed_index(int m, int n) const { return static_cast<size_t>(h_moff[m] + (n - m)); }
};

// Persistent multi-group Legendre launches (precision.hip): groups[i] are all the groups of an ensemble (same mixed
// precision, at most kMaxGroups); work items name them by index; dts[i] is group i's step.  Tiles: those of the last
// (largest) group.  Each launch has as many blocks as fit on the device at once; block x walks its own list of items,
// items synth/anal[start[x] .. start[x + 1]), built on the host (build_multi_work).
struct MultiWork {
  // (group, m, member tile start, first latitude tile) / (group, m, parity, member tile start)
  DevArray<int4> synth, anal;
  DevArray<int> synth_start, anal_start;
  int nsynth = 0, nanal = 0, synth_blocks = 0, anal_blocks = 0;
};
void build_multi_work(const std::vector<SphereCore *> &groups, unsigned active_mask, MultiWork &w);
void launch_synth_mx_multi(const std::vector<SphereCore *> &groups, const MultiWork &w, hipStream_t stream);
void launch_anal_mx_multi(const std::vector<SphereCore *> &groups, const std::vector<double> &dts, int stage, const MultiWork &w,
                          hipStream_t stream);

// free functions on the GPU: batched FFT (host arrays in/out) and Gauss-Legendre quadrature
void fft_batch(double2 *host, int n, int count, int sign);
void gauss_device(int n, double *x, double *w);

}  // namespace lmw
