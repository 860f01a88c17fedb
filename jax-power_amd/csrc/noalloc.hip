// An entry point that allocates nothing: the fused x pass + FP64 binning of
// mp_bin_poles_partial with its workspace (combine tables and per-block partial histograms)
// taken from a fixed device array instead of a caller-provided buffer, for callers that cannot
// allocate device memory (code under stream capture, serving loops, harnesses without an
// allocator).
//
// The array is a file-scope __device__ symbol of NOALLOC_SCRATCH_BYTES.  Its name is only a
// host-side stand-in: each device holds its own copy at its own address, which is read once
// per device with hipGetSymbolAddress (on the first call made with that device current) and
// used for every launch.  Every call on one device shares that device's array, so calls that
// overlap on one device (two streams, two threads) race; a call leaves its contents behind, and
// nothing in it needs to start at zero (the tables and histograms are written before they are
// read).
#include "bin.h"
#include "common.h"
#include "fft.h"

#include <mutex>
#include <vector>

namespace mp {

constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)16 << 20;  // 16 MiB per device

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mutex;
static std::vector<void*> g_scratch_ptr;  // per device, the address of that device's copy

// Address of the current device's copy of `scratch`; null with the error set on failure.
static void* scratch_for_current_device() {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    set_error(MP_ERR_HIP, "hipGetDevice failed: %s", hipGetErrorString(e));
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_error(MP_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_ptr[dev] = ptr;
  }
  return g_scratch_ptr[dev];
}

}  // namespace mp

using namespace mp;

MP_EXPORT int64_t mp_noalloc_scratch_bytes(void) { return (int64_t)NOALLOC_SCRATCH_BYTES; }

MP_EXPORT int mp_bin_poles_partial_noalloc(void* stream, const void* partial, int64_t pitch, const int32_t meshsize[3],
                                           int order, int interlacing, int compensate, int zero_dc, const double* kx,
                                           const double* ky, const double* kz, const double* k2_thresholds, int nedges,
                                           const double los[3], int nprod, const int32_t* prod_ell,
                                           const int32_t* prod_remove_zero, double* out) {
  MP_REQUIRE(meshsize && kx && ky && kz && k2_thresholds, "mp_bin_poles_partial_noalloc: null pointer");
  MP_REQUIRE(order >= 1 && order <= 4, "resampler order %d not in 1..4", order);
  MP_REQUIRE(nedges >= 2, "mp_bin_poles_partial_noalloc: need at least 2 edges, got %d", nedges);
  MP_REQUIRE(nprod >= 1, "mp_bin_poles_partial_noalloc: %d products", nprod);
  const int ni = interlacing < 1 ? 1 : interlacing;
  // the tables and at least one partial histogram must fit in the fixed array
  const size_t need = xpass_tables_bytes(meshsize, ni) + 256 + align_up((size_t)nprod * (nedges - 1) * 2 * sizeof(double), 256);
  if (need > NOALLOC_SCRATCH_BYTES)
    return set_error(MP_ERR_UNSUPPORTED,
                     "mp_bin_poles_partial_noalloc: meshsize (%d, %d, %d) with %d products of %d bins needs %zu bytes of "
                     "scratch, more than the fixed %zu",
                     meshsize[0], meshsize[1], meshsize[2], nprod, nedges - 1, need, NOALLOC_SCRATCH_BYTES);
  void* ptr = scratch_for_current_device();
  if (!ptr) return MP_ERR_HIP;
  BinGeom g{};
  g.n0 = meshsize[0];
  g.n1 = meshsize[1];
  g.ncols = meshsize[2] / 2 + 1;
  g.nedges = nedges;
  g.kx = kx;
  g.ky = ky;
  g.kz = kz;
  g.thr = k2_thresholds;
  return bin_partial((hipStream_t)stream, (const float2*)partial, pitch, meshsize, order, ni, compensate, zero_dc, g, los, nprod,
                     prod_ell, prod_remove_zero, out, ptr, NOALLOC_SCRATCH_BYTES, 0, nullptr, true);
}
