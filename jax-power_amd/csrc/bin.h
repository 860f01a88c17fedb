// Shared pieces of the |k|-shell binning: geometry, exact k^2 in float64, bin search, and the
// deterministic reduction of per-block partial histograms.
#pragma once

#include "common.h"
#include "sph.h"

namespace mp {

constexpr int BIN_LDS_DOUBLES = 4096;  // 32 KiB histogram budget per block

struct BinGeom {
  int n0, n1, ncols;
  int nedges;
  const double* kx;
  const double* ky;
  const double* kz;
  const double* thr;  // thr[e] = min{t : fl(sqrt(t)) >= edge e}
  // Added to kx, ky, kz before the norm that picks the bin (mode oversampling); mu and the
  // hermitian weights use the unshifted wavenumbers.  Zero except in mp_bin_modes/mp_bin_poles.
  double shift[3];
};

// (kx^2 + ky^2) + kz^2 exactly as IEEE float64 (no fused multiply-add), the order in which the
// reference sums the squared components.
__device__ __forceinline__ double k2_sum(double a2, double kz) {
#pragma clang fp contract(off)
  return a2 + kz * kz;
}
__device__ __forceinline__ double k2_xy(double kx, double ky) {
#pragma clang fp contract(off)
  return kx * kx + ky * ky;
}

// Number of thresholds <= k2: the mode is in bin (result - 1) when 1 <= result < nedges,
// which is digitize(|k|, edges, right=False) of the reference.
__device__ __forceinline__ int first_bin(const double* thr, int nedges, double k2) {
  int lo = 0, hi = nedges;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (thr[mid] <= k2) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Move a bin position found for a nearby k2 to the one of k2 (either direction).
__device__ __forceinline__ int walk_bin(const double* thr, int nedges, int b, double k2) {
  while (b < nedges && thr[b] <= k2) ++b;
  while (b > 0 && thr[b - 1] > k2) --b;
  return b;
}

// Separation-shell binning of a real (configuration-space) mesh: geometry and the multipoles
// asked for (sum of value * P_ell(mu_s) per bin, the (0,0,0) cell skipped where rz is set).
constexpr int CB_MAX_ELL = 6;
constexpr int CB_MAX_BLOCKS = 1024;
static_assert(CB_MAX_ELL <= LEG_MAX, "Legendre table too short");
struct RealBinArgs {
  const float* mesh;  // (n0, n1, n2), z fastest; unused by the fused z pass
  int n0, n1, n2, nedges;
  const double* sx;
  const double* sy;
  const double* sz;
  const double* thr;  // thr[e] = min{t : fl(sqrt(t)) >= edge e}
  double shift[3];    // added to sx, sy, sz for the bin only (mode oversampling), mu unshifted
  int nell;
  int ell[CB_MAX_ELL], rz[CB_MAX_ELL];
  LegendreTable leg;
  float los[3];
  int need_mu;
};

// out[e] = sum over blocks of partial[block][e], in a fixed order.
int reduce_partials(hipStream_t s, const double* partial, int nblocks, int size, double* out, int nmock = 1);

// Products of stored half spectra binned with Legendre weights (bin.hip): product p is
// conj(mesh[pi[p]]) mesh[pj[p]] (pi = -1: the constant 1), weight P_ell[p](mu), the (0,0,0) mode
// skipped where rz[p]; out[p][b][re, im].
constexpr int BIN_MAX_PROD = 6;
constexpr int BIN_MAX_MESH = 4;
struct BinProducts {
  int nmesh, nprod;
  const float2* mesh[BIN_MAX_MESH];
  int64_t pitch[BIN_MAX_MESH];
  int pi[BIN_MAX_PROD], pj[BIN_MAX_PROD], ell[BIN_MAX_PROD], rz[BIN_MAX_PROD];
  float los[3];
  int need_mu;
};
size_t bin_workspace_bytes(int nprod, int nedges);
int bin_poles(hipStream_t s, const BinGeom& g, const BinProducts& P, double* out, void* ws, size_t ws_bytes);

}  // namespace mp
