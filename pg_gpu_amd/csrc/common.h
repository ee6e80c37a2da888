// Shared helpers for libpgld: status codes, the thread-local error string and
// launch checking.  Every C-ABI entry point returns an int status and never
// lets an exception or abort cross the boundary.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <string>

#define PGLD_OK 0
#define PGLD_EINVAL (-1)     // invalid argument (null pointer, bad size)
#define PGLD_EDOMAIN (-2)    // input value outside {-1,0,1} / {-1,0,1,2}
#define PGLD_ECAPACITY (-3)  // output capacity or fixed maximum exceeded
#define PGLD_EUNSUPPORTED (-4)

namespace pgld {

void set_error(const char* fmt, ...);
const char* get_error();

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Returns PGLD_OK or the positive hipError_t of the last launch.
inline int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: %s", what, hipGetErrorString(e));
    return static_cast<int>(e);
  }
  return PGLD_OK;
}

inline int check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    set_error("%s: %s", what, hipGetErrorString(e));
    return static_cast<int>(e);
  }
  return PGLD_OK;
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// p[0..n) = v with a kernel on stream s (used instead of hipMemsetAsync so that
// a captured graph holds only kernel nodes).
int fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s);

}  // namespace pgld

#define PGLD_REQUIRE(cond, ...)          \
  do {                                   \
    if (!(cond)) {                       \
      ::pgld::set_error(__VA_ARGS__);    \
      return PGLD_EINVAL;                \
    }                                    \
  } while (0)

#define PGLD_TRY_HIP(expr, what)                       \
  do {                                                 \
    int _rc = ::pgld::check_hip((expr), (what));       \
    if (_rc != PGLD_OK) return _rc;                    \
  } while (0)
