// Autotuner of the Legendre kernels' tile shapes, per truncation group.  For each variant it times the kernel
// on the group's own state with HIP events on the group's stream (three trials of five launches; the fastest
// average wins), restores the state afterwards, and records the choice in a cache file keyed by
// (architecture, truncation, members, precision), so later runs start from the tuned shapes without timing.
//
// Variants: FP64 members per block {2, 4, 8} for the inverse and forward transforms; in the mixed modes the inverse
// transform's row chunk {32, 64} and the forward transform's (latitude chunk, waves) {32, 64, 128} x {4, 8}.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <map>
#include <sstream>
#include <string>
#include <sys/stat.h>

#include "sphere_core.h"

namespace lmw {

SphereCore &single_group(lmw_ensemble *e);
std::vector<SphereCore *> all_groups(lmw_ensemble *e);
SphereCore &core_of(lmw_sphere *s);
void drop_macro(lmw_ensemble *e);

namespace {

std::string cache_path() {
  if (const char *p = std::getenv("LMW_TUNE_CACHE")) return p;
  const char *home = std::getenv("HOME");
  std::string dir = std::string(home ? home : "/tmp") + "/.cache";
  ::mkdir(dir.c_str(), 0755);
  dir += "/lmweather_amd";
  ::mkdir(dir.c_str(), 0755);
  return dir + "/tiles_v1.txt";
}

std::string arch_name() {
  int dev = 0;
  LMW_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  LMW_HIP(hipGetDeviceProperties(&prop, dev));
  std::string a = prop.gcnArchName;
  const size_t colon = a.find(':');
  return colon == std::string::npos ? a : a.substr(0, colon);
}

std::string key_of(const SphereCore &c) {
  std::ostringstream k;
  k << arch_name() << ":T" << c.N << ":B" << c.B << ":P" << c.precision;
  return k.str();
}

std::map<std::string, std::array<int, 4>> read_cache() {
  std::map<std::string, std::array<int, 4>> out;
  std::ifstream f(cache_path());
  std::string key;
  std::array<int, 4> t;
  while (f >> key >> t[0] >> t[1] >> t[2] >> t[3]) out[key] = t;
  return out;
}

void append_cache(const std::string &key, const std::array<int, 4> &t) {
  std::ofstream f(cache_path(), std::ios::app);
  if (f) f << key << ' ' << t[0] << ' ' << t[1] << ' ' << t[2] << ' ' << t[3] << '\n';  // best effort
}

// average milliseconds per call of fn over `reps` calls, best of three trials
double time_it(SphereCore &c, const std::function<void()> &fn, int reps = 5) {
  hipEvent_t a, b;
  LMW_HIP(hipEventCreate(&a));
  LMW_HIP(hipEventCreate(&b));
  fn();  // warm-up (code object load, caches)
  double best = 1e30;
  for (int trial = 0; trial < 3; trial++) {
    LMW_HIP(hipEventRecord(a, c.stream));
    for (int i = 0; i < reps; i++) fn();
    LMW_HIP(hipEventRecord(b, c.stream));
    LMW_HIP(hipEventSynchronize(b));
    float ms = 0;
    LMW_HIP(hipEventElapsedTime(&ms, a, b));
    best = std::min(best, static_cast<double>(ms) / reps);
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return best;
}

}  // namespace

// Tunes (or loads from the cache) the tile variants of one group; tiles[4] = FP64 synth, FP64 anal, mixed synth,
// mixed anal variant indices; best_ms[2] = the chosen inverse and forward transform times (0 when loaded).
void autotune(SphereCore &c, bool force, int *tiles, double *best_ms) {
  const std::string key = key_of(c);
  std::array<int, 4> t{c.tile_fp64_synth, c.tile_fp64_anal, c.tile_synth, c.tile_anal};
  double ms[2] = {0, 0};
  auto cache = read_cache();
  auto it = cache.find(key);
  if (!force && it != cache.end()) {
    t = it->second;
  } else {
    const size_t n = c.x0.n;
    DevArray<double2> s0, s1, s2;
    s0.alloc(n);
    s1.alloc(n);
    s2.alloc(n);
    LMW_HIP(hipMemcpyAsync(s0.p, c.x0.p, n * sizeof(double2), hipMemcpyDeviceToDevice, c.stream));
    LMW_HIP(hipMemcpyAsync(s1.p, c.xs.p, n * sizeof(double2), hipMemcpyDeviceToDevice, c.stream));
    LMW_HIP(hipMemcpyAsync(s2.p, c.acc.p, n * sizeof(double2), hipMemcpyDeviceToDevice, c.stream));
    const bool mx = c.precision != LMW_FP64;
    const double dt = 1.0;  // any dt: only the launch time matters, the state is restored
    if (!mx) {
      double bs = 1e30, ba = 1e30;
      for (int v = 0; v < kFp64SynthTiles; v++) {
        c.tile_fp64_synth = v;
        const double x = time_it(c, [&] { c.launch_synth(c.xs.p, 0, c.B); });
        if (x < bs) bs = x, t[0] = v;
      }
      c.tile_fp64_synth = t[0];
      c.launch_synth(c.xs.p, 0, c.B);
      c.launch_tend_grid(0, c.B, 0);  // valid analysis operands
      c.build_fp64_mfma_tables();
      for (int v = 0; v < kFp64AnalTiles; v++) {
        c.tile_fp64_anal = v;
        const double x = time_it(c, [&] { c.launch_anal(ANAL_TEND, 2, dt, 0, c.B, nullptr); });
        if (x < ba) ba = x, t[1] = v;
      }
      ms[0] = bs;
      ms[1] = ba;
    } else {
      double bs = 1e30, ba = 1e30;
      for (int v = 0; v < kSynthTiles; v++) {
        c.tile_synth = v;
        const double x = time_it(c, [&] { c.launch_synth_mx(c.xs.p, 0, c.B); });
        if (x < bs) bs = x, t[2] = v;
      }
      c.tile_synth = t[2];
      c.launch_synth_mx(c.xs.p, 0, c.B);
      c.launch_tend_grid(0, c.B, 1);
      for (int v = 0; v < kAnalTiles; v++) {
        c.tile_anal = v;
        const double x = time_it(c, [&] { c.launch_anal_mx(ANAL_TEND, 2, dt, 0, c.B, nullptr); });
        if (x < ba) ba = x, t[3] = v;
      }
      ms[0] = bs;
      ms[1] = ba;
    }
    LMW_HIP(hipMemcpyAsync(c.x0.p, s0.p, n * sizeof(double2), hipMemcpyDeviceToDevice, c.stream));
    LMW_HIP(hipMemcpyAsync(c.xs.p, s1.p, n * sizeof(double2), hipMemcpyDeviceToDevice, c.stream));
    LMW_HIP(hipMemcpyAsync(c.acc.p, s2.p, n * sizeof(double2), hipMemcpyDeviceToDevice, c.stream));
    c.sync();
    append_cache(key, t);
  }
  for (int i = 0; i < 4; i++)
    if (t[i] < 0 || t[i] >= (i == 0 ? kFp64SynthTiles : i == 1 ? kFp64AnalTiles : i == 2 ? kSynthTiles : kAnalTiles))
      runtime("corrupt tile cache entry for " + key);
  c.tile_fp64_synth = t[0];
  c.tile_fp64_anal = t[1];
  c.tile_synth = t[2];
  c.tile_anal = t[3];
  if (c.tile_fp64_anal >= kFp64Tiles) c.build_fp64_mfma_tables();  // before any graph capture
  c.drop_graph();
  if (tiles)
    for (int i = 0; i < 4; i++) tiles[i] = t[i];
  if (best_ms) {
    best_ms[0] = ms[0];
    best_ms[1] = ms[1];
  }
}

}  // namespace lmw

using namespace lmw;

extern "C" {

lmw_status lmw_ensemble_autotune(lmw_ensemble *e, int force, int *tiles, double *best_ms) {
  return guard([&] {
    if (!e) invalid("null pointer argument");
    std::vector<SphereCore *> gs = all_groups(e);
    for (size_t g = 0; g < gs.size(); g++) autotune(*gs[g], force != 0, tiles ? tiles + 4 * g : nullptr, best_ms ? best_ms + 2 * g : nullptr);
    drop_macro(e);
  });
}
lmw_status lmw_sphere_autotune(lmw_sphere *s, int force, int *tiles, double *best_ms) {
  return guard([&] {
    if (!s) invalid("null pointer argument");
    autotune(core_of(s), force != 0, tiles, best_ms);
  });
}
lmw_status lmw_ensemble_set_tiles(lmw_ensemble *e, int group, const int *tiles) {
  return guard([&] {
    if (!e || !tiles) invalid("null pointer argument");
    std::vector<SphereCore *> gs = all_groups(e);
    if (group < 0 || group >= static_cast<int>(gs.size())) invalid("group index out of range");
    for (int i = 0; i < 4; i++)
      if (tiles[i] < 0 || tiles[i] >= (i == 0 ? kFp64SynthTiles : i == 1 ? kFp64AnalTiles : i == 2 ? kSynthTiles : kAnalTiles))
        invalid("tile index out of range");
    SphereCore &c = *gs[group];
    c.tile_fp64_synth = tiles[0];
    c.tile_fp64_anal = tiles[1];
    if (c.tile_fp64_anal >= kFp64Tiles) c.build_fp64_mfma_tables();
    c.tile_synth = tiles[2];
    c.tile_anal = tiles[3];
    c.drop_graph();
    drop_macro(e);
  });
}
lmw_status lmw_tile_variants(int kind, int index, int *a, int *b) {
  return guard([&] {
    if (!a || !b) invalid("null pointer argument");
    switch (kind) {
      case 0:  // FP64 synthesis: members per block of the FMA kernel (b = 0), or 0 and the row chunk of the MFMA kernel
        if (index < 0 || index >= kFp64SynthTiles) invalid("tile index out of range");
        *a = index < kFp64Tiles ? kFp64MB[index] : 0;
        *b = index < kFp64Tiles ? 0 : index == kFp64Tiles ? 32 : 64;
        break;
      case 1:
        if (index < 0 || index >= kSynthTiles) invalid("tile index out of range");
        *a = kSynthRC[index];
        *b = 0;
        break;
      case 2:
        if (index < 0 || index >= kAnalTiles) invalid("tile index out of range");
        *a = kAnalKC[index];
        *b = kAnalNW[index];
        break;
      case 3:  // FP64 analysis: members per block of the FMA kernel (b = 0), or 0 and the waves of the MFMA kernel
        if (index < 0 || index >= kFp64AnalTiles) invalid("tile index out of range");
        *a = index < kFp64Tiles ? kFp64MB[index] : 0;
        *b = index < kFp64Tiles ? 0 : index == kFp64Tiles ? 4 : 8;
        break;
      default: invalid("unknown tile kind");
    }
  });
}

}  // extern "C"
