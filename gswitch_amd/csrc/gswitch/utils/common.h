// Enums, constants and error plumbing shared by the engine.
// Names and values follow src/utils/common.cuh:6-37 of the reference so that
// application code written against GSWITCH compiles unchanged.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#define ALL_ACTIVE -1
#define ALL_INACTIVE -2

enum Status { Inactive = 0, Active = 1, All = 2, Fixed = 3 };
enum GraphFmt { CSR, COO };
enum Centric { EC, VC };
enum Computation { Idem, Mono, ASSO };
enum QueueMode { Normal, Cached };
enum Direction { Push, Pull };
enum ASFmt { Queue, Bitmap };
enum LB { TM, WM, CM, STRICT, TWOD, ELB, TWC };
enum Fusion { Fused, Standalone };
enum CullMode { Interleave, Stride };

// Empty edge / vertex payload (src/utils/utils.cuh:208-218).
struct Empty {
  __host__ __device__ Empty() {}
  template <typename T> __host__ __device__ Empty(T) {}
};

namespace gsw {

constexpr int WAVE = 64;          // CDNA wavefront width
constexpr int LOG_WAVE = 6;
constexpr int BLOCK = 256;        // default workgroup: 4 waves
constexpr int MAX_GRID = 2048;    // 256 CUs x 8 resident workgroups

// Host-side error type; the C-ABI turns it into a status code.
struct error : std::runtime_error {
  int code;
  error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

enum : int {
  E_OK = 0,
  E_INVALID = 1,
  E_HIP = 2,
  E_ALLOC = 3,
  E_OVERFLOW = 4,
  E_UNSUPPORTED = 5,
  E_INTERNAL_CHECK = 6,
};

inline void hip_check(hipError_t e, const char *what, const char *file,
                      int line) {
  if (e != hipSuccess) {
    char buf[512];
    snprintf(buf, sizeof(buf), "%s failed at %s:%d: %s", what, file, line,
             hipGetErrorString(e));
    throw error(e == hipErrorOutOfMemory ? E_ALLOC : E_HIP, buf);
  }
}

}  // namespace gsw

#define H_ERR(x) ::gsw::hip_check((x), #x, __FILE__, __LINE__)
#define GSW_REQUIRE(cond, code, msg)                  \
  do {                                                \
    if (!(cond)) throw ::gsw::error((code), (msg));   \
  } while (0)

#ifndef MIN
#define MIN(a, b) ((a) < (b) ? (a) : (b))
#endif
#ifndef MAX
#define MAX(a, b) ((a) > (b) ? (a) : (b))
#endif
#define DIV(a, b) ((b) == 0 ? 0 : (((a) + .0) / (b)))
