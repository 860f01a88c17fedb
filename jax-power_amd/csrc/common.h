// Shared helpers for the MI355X mesh power-spectrum kernels: error reporting through the
// C ABI, per-kernel event timing, and small device utilities.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#define MP_EXPORT extern "C" __attribute__((visibility("default")))

namespace mp {

// Error codes returned by every C entry point.
enum : int {
  MP_OK = 0,
  MP_ERR_INVALID = -1,   // bad argument (shape, size, null pointer)
  MP_ERR_UNSUPPORTED = -2,  // valid for the reference but not implemented on this path
  MP_ERR_HIP = -3,       // HIP runtime error
  MP_ERR_WORKSPACE = -4, // workspace too small
};

// Thread-local last error message (read with mp_last_error()).
int set_error(int code, const char* fmt, ...);

// Kernel timing: when enabled, every launch made through MP_LAUNCH records a pair of HIP
// events on the launch stream; mp_timing_collect() synchronises them and accumulates the
// per-kernel-name totals.
bool timing_enabled();
void timing_begin(const char* name, hipStream_t stream, void** token);
void timing_end(void* token, hipStream_t stream);

}  // namespace mp

#define MP_CHECK_HIP(expr)                                                                \
  do {                                                                                    \
    hipError_t _e = (expr);                                                               \
    if (_e != hipSuccess)                                                                 \
      return ::mp::set_error(::mp::MP_ERR_HIP, "%s failed: %s (%s:%d)", #expr,             \
                             hipGetErrorString(_e), __FILE__, __LINE__);                  \
  } while (0)

// Launch a kernel on `stream`, with optional event timing under `name`, then check the
// launch error.  Usage: MP_LAUNCH("paint_tile", kernel, grid, block, shmem, stream, args...)
#define MP_LAUNCH(name, kernel, grid, block, shmem, stream, ...)                          \
  do {                                                                                    \
    void* _tok = nullptr;                                                                 \
    if (::mp::timing_enabled()) ::mp::timing_begin(name, stream, &_tok);                  \
    hipLaunchKernelGGL(kernel, grid, block, shmem, stream, __VA_ARGS__);                  \
    hipError_t _e = hipGetLastError();                                                    \
    if (_tok) ::mp::timing_end(_tok, stream);                                             \
    if (_e != hipSuccess)                                                                 \
      return ::mp::set_error(::mp::MP_ERR_HIP, "launch of %s failed: %s", name,           \
                             hipGetErrorString(_e));                                      \
  } while (0)

#define MP_REQUIRE(cond, ...)                                                             \
  do {                                                                                    \
    if (!(cond)) return ::mp::set_error(::mp::MP_ERR_INVALID, __VA_ARGS__);                \
  } while (0)

namespace mp {

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
__host__ inline bool is_pow2(int64_t n) { return n > 0 && (n & (n - 1)) == 0; }
__host__ inline size_t align_up(size_t n, size_t a) { return (n + a - 1) / a * a; }

// Simple bump allocator over a caller-provided workspace.
struct Workspace {
  char* base;
  size_t size;
  size_t used;
  Workspace(void* b, size_t s) : base(static_cast<char*>(b)), size(s), used(0) {}
  template <typename T>
  T* take(size_t count) {
    size_t off = align_up(used, 256);
    size_t bytes = align_up(count * sizeof(T), 256);
    if (!base || off + bytes > size) return nullptr;
    used = off + bytes;
    return reinterpret_cast<T*>(base + off);
  }
};

}  // namespace mp
