// Shared host/device helpers of the gfx950 dynamical core: error translation for the C ABI, device buffers,
// complex arithmetic on double2.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "lmw.h"

namespace lmw {

// An error that crosses the C ABI as (status, message).
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] inline void invalid(const std::string &m) { throw Error(LMW_EINVAL, m); }
[[noreturn]] inline void runtime(const std::string &m) { throw Error(LMW_ERUNTIME, m); }

#define LMW_HIP(x)                                                                                          \
  do {                                                                                                      \
    hipError_t e_ = (x);                                                                                    \
    if (e_ != hipSuccess) throw ::lmw::Error(LMW_EDEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)
#define LMW_LAUNCH_CHECK() LMW_HIP(hipGetLastError())

void set_error(const std::string &msg);
void note_bytes(int64_t delta);

// Runs fn, mapping exceptions to a status and the thread-local message (no exception crosses the C ABI).
template <class F> lmw_status guard(F &&fn) {
  try {
    fn();
    return LMW_OK;
  } catch (const Error &e) {
    set_error(e.what());
    return e.code;
  } catch (const std::bad_alloc &) {
    set_error("out of host memory");
    return LMW_ERUNTIME;
  } catch (const std::exception &e) {
    set_error(e.what());
    return LMW_ERUNTIME;
  }
}

// std::to_string(double) formatting ("%f"), used in the reference's run() messages (sphere.cpp:465, plane.cpp:238).
inline std::string fstr(double x) {
  char b[64];
  std::snprintf(b, sizeof b, "%f", x);
  return b;
}

// Device array owned by a handle; zero-filled at allocation.
template <class T> struct DevArray {
  T *p = nullptr;
  size_t n = 0;
  DevArray() = default;
  DevArray(const DevArray &) = delete;
  DevArray &operator=(const DevArray &) = delete;
  ~DevArray() { release(); }
  void alloc(size_t count) {
    release();
    if (count == 0) return;
    LMW_HIP(hipMalloc(&p, count * sizeof(T)));
    n = count;
    note_bytes(static_cast<int64_t>(count * sizeof(T)));
    LMW_HIP(hipMemset(p, 0, count * sizeof(T)));
  }
  void release() {
    if (p) {
      (void)hipFree(p);
      note_bytes(-static_cast<int64_t>(n * sizeof(T)));
    }
    p = nullptr;
    n = 0;
  }
  size_t bytes() const { return n * sizeof(T); }
  void upload(const T *src, size_t count, size_t offset, hipStream_t s) {
    LMW_HIP(hipMemcpyAsync(p + offset, src, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void download(T *dst, size_t count, size_t offset, hipStream_t s) const {
    LMW_HIP(hipMemcpyAsync(dst, p + offset, count * sizeof(T), hipMemcpyDeviceToHost, s));
  }
};

// ---------------------------------------------------------------------------------- complex arithmetic
__host__ __device__ inline double2 cx(double r, double i) { return make_double2(r, i); }
__host__ __device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__host__ __device__ inline double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__host__ __device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// (i m) * a, as (I * double(m)) * a of the reference: (-m a.im, m a.re)
__host__ __device__ inline double2 cim(double m, double2 a) { return make_double2(-m * a.y, m * a.x); }
// storage conversions of complex values (FP64 compute, FP64 or FP32 storage)
__host__ __device__ inline double2 to_d2(double2 a) { return a; }
__host__ __device__ inline double2 to_d2(float2 a) { return make_double2(a.x, a.y); }
template <class T> __host__ __device__ inline T from_d2(double2 a);
template <> __host__ __device__ inline double2 from_d2<double2>(double2 a) { return a; }
template <> __host__ __device__ inline float2 from_d2<float2>(double2 a) { return make_float2(static_cast<float>(a.x), static_cast<float>(a.y)); }
// the same operations on float2 (FP32 grid-point work of the mixed-precision step)
__host__ __device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__host__ __device__ inline float2 cmul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__host__ __device__ inline float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__host__ __device__ inline float2 cim(float m, float2 a) { return make_float2(-m * a.y, m * a.x); }
// complex number of type C from two doubles; the real type of a complex type
template <class C> __host__ __device__ inline C mkc(double re, double im);
template <> __host__ __device__ inline double2 mkc<double2>(double re, double im) { return make_double2(re, im); }
template <> __host__ __device__ inline float2 mkc<float2>(double re, double im) { return make_float2(static_cast<float>(re), static_cast<float>(im)); }
template <class C> struct real_of;
template <> struct real_of<double2> { using type = double; };
template <> struct real_of<float2> { using type = float; };
// storage conversions between the complex types
template <class C> __host__ __device__ inline C to_c(double2 a);
template <class C> __host__ __device__ inline C to_c(float2 a);
template <> __host__ __device__ inline double2 to_c<double2>(double2 a) { return a; }
template <> __host__ __device__ inline double2 to_c<double2>(float2 a) { return make_double2(a.x, a.y); }
template <> __host__ __device__ inline float2 to_c<float2>(double2 a) { return make_float2(static_cast<float>(a.x), static_cast<float>(a.y)); }
template <> __host__ __device__ inline float2 to_c<float2>(float2 a) { return a; }

// acc += a * s (complex times real)
__device__ inline void cfma(double2 &acc, double2 a, double s) {
  acc.x = fma(a.x, s, acc.x);
  acc.y = fma(a.y, s, acc.y);
}

}  // namespace lmw
