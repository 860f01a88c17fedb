// Composite entry points: interlaced paint -> spectrum, paint -> real mesh, the local
// line-of-sight Y_lm accumulation, and FP64 reductions.
//
// mp_paint_spectrum replaces _paint(..., out='complex') (jaxpower/mesh.py:2673-2754): the
// n_i shifted meshes are painted in one pass, each is transformed along z and y, and one x
// pass transforms all n_i partial spectra, applies exp(+i s_j (kx+ky+kz)) / n_i (mesh.py:
// 2740-2744), sums them, and applies the compensation prod_a sinc(k_a/2pi)^-p
// (resamplers.py:16-26) before the single write of the result.
#include "bin.h"
#include "common.h"
#include "fft.h"
#include "paint.h"

#include <cmath>

namespace mp {

int fill_f32(hipStream_t s, float* p, int64_t n, float v);
int read_mesh(hipStream_t s, const float* mesh, const GridMap& g, const void* pos, int pos_f64, int64_t np, int order, float* out);

static int g_paint_path = 0;  // automatic: fused paint + z-FFT where supported
int paint_path_option() { return g_paint_path; }

static bool use_fused(const int32_t n[3], int64_t np, int order, int ni) {
  int opt = paint_path_option();
  if (opt == 1 || opt == 3) return false;
  return paint_zfft_supported(n, np, order, ni);
}

// Workspace of paint_partial beyond its output (the n_i partial spectra).
size_t paint_partial_workspace_bytes(const int32_t n[3], int64_t np, int order, int ni) {
  if (use_fused(n, np, order, ni)) return paint_zfft_workspace_bytes(n, np) + 512;
  const int64_t ncell = (int64_t)n[0] * n[1] * n[2];
  return align_up((size_t)ni * ncell * 4, 256) + paint_workspace_bytes(n, np, order, ni) + 512;
}

// The n_i shifted meshes painted and transformed along z and y: partial + j * n0 * n1 * pitch.
int paint_partial(hipStream_t s, const void* pos, int pos_f64, const float* w, int64_t np, const GridMap& g, int order, int ni,
                  float2* partial, int64_t pitch, void* ws, size_t ws_bytes) {
  const int32_t n[3] = {g.n[0], g.n[1], g.n[2]};
  MP_REQUIRE(partial, "paint_partial: null output");
  MP_REQUIRE(pitch >= n[2] / 2 + 1, "paint_partial: pitch too small");
  const int64_t ncell = (int64_t)n[0] * n[1] * n[2];
  const int64_t nspec = (int64_t)n[0] * n[1] * pitch;
  const bool fused = use_fused(n, np, order, ni);
  if (paint_path_option() == 2 && !fused) return set_error(MP_ERR_UNSUPPORTED, "paint: fused path not available");
  Workspace wsp(ws, ws_bytes);
  int rc = MP_OK;
  if (fused) {
    size_t pbytes = paint_zfft_workspace_bytes(n, np);
    void* pws = wsp.take<char>(pbytes);
    if (!pws)
      return set_error(MP_ERR_WORKSPACE, "paint: workspace of %zu bytes < %zu needed", ws_bytes,
                       paint_partial_workspace_bytes(n, np, order, ni));
    rc = paint_zfft(s, pos, pos_f64, w, np, g, order, ni, partial, pitch, nspec, pws, pbytes);
    if (rc) return rc;
    return ypass_meshes(s, partial, ni, nspec, n, pitch);
  }
  float* meshes = wsp.take<float>((size_t)ni * ncell);
  size_t pbytes = paint_workspace_bytes(n, np, order, ni);
  void* pws = wsp.take<char>(pbytes);
  if (!meshes || !pws)
    return set_error(MP_ERR_WORKSPACE, "paint: workspace of %zu bytes < %zu needed", ws_bytes,
                     paint_partial_workspace_bytes(n, np, order, ni));
  rc = paint_meshes(s, pos, pos_f64, w, np, g, order, ni, meshes, ncell, pws, pbytes, paint_path_option() == 3);
  if (rc) return rc;
  RealPrologue pro{};
  for (int j = 0; j < ni; ++j) {
    rc = rfft3d_zy(s, meshes + j * ncell, partial + j * nspec, n, pitch, pro);
    if (rc) return rc;
  }
  return MP_OK;
}

size_t paint_spectrum_workspace_bytes(const int32_t n[3], int64_t np, int order, int ni) {
  const int64_t nspec = (int64_t)n[0] * n[1] * spectrum_pitch(n[2]);
  return align_up((size_t)ni * nspec * 8, 256) + paint_partial_workspace_bytes(n, np, order, ni) +
         combine_partial_workspace_bytes(n, ni) + 1024;
}

int paint_spectrum(hipStream_t s, const void* pos, int pos_f64, const float* w, int64_t np, const GridMap& g, int order, int ni,
                   int compensate, float2* out, int64_t out_pitch, void* ws, size_t ws_bytes) {
  const int32_t n[3] = {g.n[0], g.n[1], g.n[2]};
  MP_REQUIRE(out_pitch >= n[2] / 2 + 1, "paint_spectrum: out_pitch too small");
  if (!fft_any_shape_ok(n)) return set_error(MP_ERR_UNSUPPORTED, "paint_spectrum: meshsize (%d, %d, %d) not in 1..2048", n[0], n[1], n[2]);
  const int64_t P = spectrum_pitch(n[2]);
  const int64_t nspec = (int64_t)n[0] * n[1] * P;
  Workspace wsp(ws, ws_bytes);
  float2* partial = wsp.take<float2>((size_t)ni * nspec);
  const size_t pbytes = paint_partial_workspace_bytes(n, np, order, ni), cbytes = combine_partial_workspace_bytes(n, ni);
  void* pws = wsp.take<char>(pbytes);
  void* cws = wsp.take<char>(cbytes);
  if (!partial || !pws || !cws)
    return set_error(MP_ERR_WORKSPACE, "paint_spectrum: workspace of %zu bytes < %zu needed", ws_bytes,
                     paint_spectrum_workspace_bytes(n, np, order, ni));
  int rc = paint_partial(s, pos, pos_f64, w, np, g, order, ni, partial, P, pws, pbytes);
  if (rc) return rc;
  return combine_partial(s, partial, P, n, order, ni, compensate, 0, out, out_pitch, cws, cbytes);
}

// ------------------------------------------------------------------ FP64 reductions

constexpr int RED_NT = 256;
constexpr int RED_BLOCKS = 1024;

__global__ void __launch_bounds__(RED_NT)
dot_partial_kernel(const float* __restrict__ a, const float* __restrict__ b, int64_t n, int square, double* __restrict__ partial) {
  __shared__ double sh[RED_NT];
  double s = 0.;
  auto term = [&](float u, float v) {
    const double x = (double)u;
    return b ? x * (double)v : (square ? x * x : x);
  };
  const int64_t first = (int64_t)blockIdx.x * RED_NT + threadIdx.x, stride = (int64_t)gridDim.x * RED_NT;
  int64_t done = 0;
  // 16-byte loads where both vectors allow them (weights of particle fields are), the rest one
  // element at a time; the grid size fixes the summation order, so sums are reproducible
  if (((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & 15) == 0) {
    const int64_t n4 = n / 4;
    const float4* a4 = reinterpret_cast<const float4*>(a);
    const float4* b4 = reinterpret_cast<const float4*>(b);
    for (int64_t i = first; i < n4; i += stride) {
      const float4 u = a4[i];
      const float4 v = b ? b4[i] : make_float4(0.f, 0.f, 0.f, 0.f);
      s += (term(u.x, v.x) + term(u.y, v.y)) + (term(u.z, v.z) + term(u.w, v.w));
    }
    done = n4 * 4;
  }
  for (int64_t i = done + first; i < n; i += stride) s += term(a[i], b ? b[i] : 0.f);
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int d = RED_NT / 2; d > 0; d >>= 1) {
    if (threadIdx.x < d) sh[threadIdx.x] += sh[threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// out[r] = sum of the nblocks partial sums of row r = blockIdx.x, in a fixed order.
__global__ void __launch_bounds__(RED_NT) final_sum_kernel(const double* __restrict__ partial, int nblocks, double* __restrict__ out) {
  __shared__ double sh[RED_NT];
  partial += (int64_t)blockIdx.x * nblocks;
  out += blockIdx.x;
  double s = 0.;
  for (int i = threadIdx.x; i < nblocks; i += RED_NT) s += partial[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int d = RED_NT / 2; d > 0; d >>= 1) {
    if (threadIdx.x < d) sh[threadIdx.x] += sh[threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

// Sum w and sum w^2 (FP64) of each catalogue of a batch, catalogue m = blockIdx.y holding
// w[moff[m] .. moff[m + 1]) (w null: unit weights); partial[(2 m + q) gridDim.x + b] for
// block b, q = 0 (w) and 1 (w^2), reduced by final_sum_kernel.  About 4096 blocks in all
// (at least 16 per catalogue): the loads are 4-byte and latency-bound, so the grid must fill
// the GPU for one catalogue as for many.
static int cat_blocks(int nmock) {
  const int b = (4096 + nmock - 1) / nmock;
  return b < 16 ? 16 : (b > 1024 ? 1024 : b);
}

__global__ void __launch_bounds__(RED_NT)
catalogue_sums_kernel(const float* __restrict__ w, const int64_t* __restrict__ moff, double* __restrict__ partial) {
  __shared__ double sh[2][RED_NT];
  const int m = blockIdx.y;
  const int64_t i0 = moff[m], i1 = moff[m + 1];
  double s1 = 0., s2 = 0.;
  if (w) {
    for (int64_t i = i0 + (int64_t)blockIdx.x * RED_NT + threadIdx.x; i < i1; i += (int64_t)gridDim.x * RED_NT) {
      const double x = (double)w[i];
      s1 += x;
      s2 += x * x;
    }
  } else if (blockIdx.x == 0 && threadIdx.x == 0) {
    s1 = s2 = (double)(i1 - i0);
  }
  sh[0][threadIdx.x] = s1;
  sh[1][threadIdx.x] = s2;
  __syncthreads();
  for (int d = RED_NT / 2; d > 0; d >>= 1) {
    if (threadIdx.x < d) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + d];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + d];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[(2 * (int64_t)m) * gridDim.x + blockIdx.x] = sh[0][0];
    partial[(2 * (int64_t)m + 1) * gridDim.x + blockIdx.x] = sh[1][0];
  }
}

}  // namespace mp

using namespace mp;

static GridMap make_map(const int32_t meshsize[3], const double boxsize[3], const double boxcenter[3]) {
  GridMap g;
  for (int a = 0; a < 3; ++a) {
    g.n[a] = meshsize[a];
    g.cell[a] = boxsize[a] / (double)meshsize[a];
    g.shift[a] = boxsize[a] / 2. - boxcenter[a];
  }
  return g;
}

MP_EXPORT int64_t mp_paint_spectrum_workspace_bytes(const int32_t meshsize[3], int64_t np, int order, int interlacing) {
  return (int64_t)paint_spectrum_workspace_bytes(meshsize, np, order, interlacing < 1 ? 1 : interlacing);
}

MP_EXPORT int mp_paint_spectrum(void* stream, const void* positions, int positions_f64, const float* weights, int64_t np,
                                const int32_t meshsize[3], const double boxsize[3], const double boxcenter[3], int order,
                                int interlacing, int compensate, void* out, int64_t out_pitch, void* workspace,
                                size_t workspace_bytes) {
  MP_REQUIRE(meshsize && boxsize && boxcenter && out, "mp_paint_spectrum: null pointer");
  MP_REQUIRE(order >= 1 && order <= 4, "resampler order %d not in 1..4", order);
  GridMap g = make_map(meshsize, boxsize, boxcenter);
  return paint_spectrum((hipStream_t)stream, positions, positions_f64, weights, np, g, order, interlacing < 1 ? 1 : interlacing,
                        compensate, (float2*)out, out_pitch, workspace, workspace_bytes);
}

MP_EXPORT int64_t mp_paint_partial_workspace_bytes(const int32_t meshsize[3], int64_t np, int order, int interlacing) {
  return (int64_t)paint_partial_workspace_bytes(meshsize, np, order, interlacing < 1 ? 1 : interlacing);
}

MP_EXPORT int mp_paint_partial(void* stream, const void* positions, int positions_f64, const float* weights, int64_t np,
                               const int32_t meshsize[3], const double boxsize[3], const double boxcenter[3], int order,
                               int interlacing, void* partial, int64_t pitch, void* workspace, size_t workspace_bytes) {
  MP_REQUIRE(meshsize && boxsize && boxcenter && partial, "mp_paint_partial: null pointer");
  MP_REQUIRE(order >= 1 && order <= 4, "resampler order %d not in 1..4", order);
  GridMap g = make_map(meshsize, boxsize, boxcenter);
  return paint_partial((hipStream_t)stream, positions, positions_f64, weights, np, g, order, interlacing < 1 ? 1 : interlacing,
                       (float2*)partial, pitch, workspace, workspace_bytes);
}

MP_EXPORT int64_t mp_combine_partial_workspace_bytes(const int32_t meshsize[3], int interlacing) {
  return (int64_t)combine_partial_workspace_bytes(meshsize, interlacing < 1 ? 1 : interlacing);
}

MP_EXPORT int mp_combine_partial(void* stream, const void* partial, int64_t pitch, const int32_t meshsize[3], int order,
                                 int interlacing, int compensate, int zero_dc, void* out, int64_t out_pitch, void* workspace,
                                 size_t workspace_bytes) {
  MP_REQUIRE(meshsize, "mp_combine_partial: null meshsize");
  MP_REQUIRE(order >= 1 && order <= 4, "resampler order %d not in 1..4", order);
  return combine_partial((hipStream_t)stream, (const float2*)partial, pitch, meshsize, order, interlacing < 1 ? 1 : interlacing,
                         compensate, zero_dc, (float2*)out, out_pitch, workspace, workspace_bytes);
}

MP_EXPORT int mp_combine_partial_power(void* stream, const void* partial, int64_t pitch, const int32_t meshsize[3], int order,
                                       int interlacing, int compensate, int zero_dc, int inverse_x, void* out, int64_t out_pitch,
                                       void* workspace, size_t workspace_bytes) {
  MP_REQUIRE(meshsize, "mp_combine_partial_power: null meshsize");
  MP_REQUIRE(order >= 1 && order <= 4, "resampler order %d not in 1..4", order);
  return combine_partial((hipStream_t)stream, (const float2*)partial, pitch, meshsize, order, interlacing < 1 ? 1 : interlacing,
                         compensate, zero_dc, (float2*)out, out_pitch, workspace, workspace_bytes, inverse_x ? 2 : 1);
}

MP_EXPORT int64_t mp_bin_partial_workspace_bytes(const int32_t meshsize[3], int interlacing, int nprod, int nedges) {
  return (int64_t)bin_partial_workspace_bytes(meshsize, interlacing < 1 ? 1 : interlacing, nprod, nedges);
}

MP_EXPORT int mp_bin_poles_partial(void* stream, const void* partial, int64_t pitch, const int32_t meshsize[3], int order,
                                   int interlacing, int compensate, int zero_dc, const double* kx, const double* ky,
                                   const double* kz, const double* k2_thresholds, int nedges, const double los[3], int nprod,
                                   const int32_t* prod_ell, const int32_t* prod_remove_zero, double* out, void* workspace,
                                   size_t workspace_bytes) {
  MP_REQUIRE(meshsize && kx && ky && kz && k2_thresholds, "mp_bin_poles_partial: null pointer");
  MP_REQUIRE(order >= 1 && order <= 4, "resampler order %d not in 1..4", order);
  MP_REQUIRE(nedges >= 2, "mp_bin_poles_partial: need at least 2 edges, got %d", nedges);
  BinGeom g{};
  g.n0 = meshsize[0];
  g.n1 = meshsize[1];
  g.ncols = meshsize[2] / 2 + 1;
  g.nedges = nedges;
  g.kx = kx;
  g.ky = ky;
  g.kz = kz;
  g.thr = k2_thresholds;
  return bin_partial((hipStream_t)stream, (const float2*)partial, pitch, meshsize, order, interlacing < 1 ? 1 : interlacing,
                     compensate, zero_dc, g, los, nprod, prod_ell, prod_remove_zero, out, workspace, workspace_bytes);
}

MP_EXPORT int64_t mp_ylm_workspace_bytes(const int32_t meshsize[3]) {
  return (int64_t)meshsize[0] * meshsize[1] * spectrum_pitch(meshsize[2]) * 8 + 1024;
}
// (meshes whose sizes are not all powers of two run the x pass in place on the same scratch)

// C_l(k) = sum_m r2c(rho * Y_lm(xhat))(k) * Y_lm(khat), jaxpower/mesh2.py:319-329.
MP_EXPORT int mp_ylm_accumulate(void* stream, const float* rmesh, const int32_t meshsize[3], const double boxsize[3],
                                const double boxcenter[3], int ell, void* out, int64_t out_pitch, void* workspace,
                                size_t workspace_bytes) {
  MP_REQUIRE(rmesh && meshsize && boxsize && boxcenter && out, "mp_ylm_accumulate: null pointer");
  MP_REQUIRE(ell >= 1 && ell <= MAX_ELL, "mp_ylm_accumulate: ell %d not in 1..%d", ell, MAX_ELL);
  MP_REQUIRE(out_pitch >= meshsize[2] / 2 + 1, "mp_ylm_accumulate: out_pitch too small");
  hipStream_t s = (hipStream_t)stream;
  const int64_t P = spectrum_pitch(meshsize[2]);
  Workspace wsp(workspace, workspace_bytes);
  float2* tmp = wsp.take<float2>((size_t)meshsize[0] * meshsize[1] * P);
  if (!tmp) return set_error(MP_ERR_WORKSPACE, "mp_ylm_accumulate: workspace too small");
  for (int m = -ell; m <= ell; ++m) {
    RealPrologue pro{};
    pro.kind = PRO_YLM;
    pro.ell = ell;
    pro.m = m;
    pro.ny = meshsize[1];
    pro.nz = meshsize[2];
    pro.x0 = boxcenter[0] - boxsize[0] / 2.;
    pro.y0 = boxcenter[1] - boxsize[1] / 2.;
    pro.z0 = boxcenter[2] - boxsize[2] / 2.;
    pro.dx = boxsize[0] / meshsize[0];
    pro.dy = boxsize[1] / meshsize[1];
    pro.dz = boxsize[2] / meshsize[2];
    int rc = rfft3d_zy(s, rmesh, tmp, meshsize, P, pro);
    if (rc) return rc;
    KPrologue kp{};
    kp.ell = ell;
    kp.m = m;
    kp.nx = meshsize[0];
    kp.ny = meshsize[1];
    kp.kfx = (float)(2. * M_PI / boxsize[0]);
    kp.kfy = (float)(2. * M_PI / boxsize[1]);
    kp.kfz = (float)(2. * M_PI / boxsize[2]);
    if (pow2_shape(meshsize)) {
      rc = ylm_acc_x(s, tmp, (float2*)out, meshsize, P, out_pitch, kp, m == -ell);
    } else {  // x pass in place on the scratch, then the Y_lm(khat) multiply-accumulate
      const int64_t ny = meshsize[1];
      rc = any_col_fft(s, tmp, tmp, meshsize[0], meshsize[2] / 2 + 1, ny, ny * P, P, ny * P, P, false, 1.f, "fft_xpass_any");
      if (!rc) rc = ylm_acc_any(s, tmp, P, (float2*)out, out_pitch, meshsize, kp, m == -ell);
    }
    if (rc) return rc;
  }
  return MP_OK;
}

MP_EXPORT int64_t mp_ylm_bin_workspace_bytes(const int32_t meshsize[3], int ell, int nedges) {
  return (int64_t)ylm_bin_workspace_bytes(meshsize, ell, nedges);
}

// The ell > 0 terms of the local line-of-sight estimator (jaxpower/mesh2.py:311-333) for one
// ell, fused: sum over bins of conj(C_ell) A0 w with C_ell = sum_m r2c(rho Y_lm(xhat)) Y_lm(khat),
// and with_ell0 != 0 also |A0|^2 w; C_ell is never stored.
MP_EXPORT int mp_ylm_bin(void* stream, const float* rmesh, const void* A0, int64_t a0_pitch, const int32_t meshsize[3],
                         const double boxsize[3], const double boxcenter[3], int ell, int with_ell0, const double* kx,
                         const double* ky, const double* kz, const double* k2_thresholds, int nedges, double* out, void* workspace,
                         size_t workspace_bytes) {
  MP_REQUIRE(meshsize && boxsize && boxcenter && kx && ky && kz && k2_thresholds, "mp_ylm_bin: null pointer");
  MP_REQUIRE(nedges >= 2, "mp_ylm_bin: need at least 2 edges, got %d", nedges);
  BinGeom g{};
  g.n0 = meshsize[0];
  g.n1 = meshsize[1];
  g.ncols = meshsize[2] / 2 + 1;
  g.nedges = nedges;
  g.kx = kx;
  g.ky = ky;
  g.kz = kz;
  g.thr = k2_thresholds;
  return ylm_bin((hipStream_t)stream, rmesh, (const float2*)A0, a0_pitch, meshsize, boxsize, boxcenter, ell, with_ell0, g, out,
                 workspace, workspace_bytes);
}

MP_EXPORT int64_t mp_reduce_workspace_bytes(void) { return RED_BLOCKS * sizeof(double) + 256; }

// out[0] = sum_i a_i * b_i (b given), sum_i a_i^2 (square), or sum_i a_i, accumulated in FP64.
MP_EXPORT int mp_sum_f64(void* stream, const float* a, const float* b, int64_t n, int square, double* out, void* workspace,
                         size_t workspace_bytes) {
  MP_REQUIRE(out && (a || n == 0), "mp_sum_f64: null pointer");
  hipStream_t s = (hipStream_t)stream;
  Workspace wsp(workspace, workspace_bytes);
  double* partial = wsp.take<double>(RED_BLOCKS);
  if (!partial) return set_error(MP_ERR_WORKSPACE, "mp_sum_f64: workspace too small");
  int blocks = (int)(ceil_div(n, RED_NT) < RED_BLOCKS ? ceil_div(n, RED_NT) : RED_BLOCKS);
  if (blocks < 1) blocks = 1;
  MP_LAUNCH("sum_partial", dot_partial_kernel, dim3(blocks), dim3(RED_NT), 0, s, a, b, n, square, partial);
  MP_LAUNCH("sum_final", final_sum_kernel, dim3(1), dim3(RED_NT), 0, s, (const double*)partial, blocks, out);
  return MP_OK;
}

// Selects the paint path of mp_paint_spectrum (0 automatic, 1 separate paint + z pass,
// 2 fused paint + z pass only, 3 per-particle atomics + z pass); returns the previous value,
// and a negative `path` only returns the current one.
MP_EXPORT int mp_set_paint_path(int path) {
  int old = g_paint_path;
  if (path >= 0) g_paint_path = path <= 3 ? path : 0;  // a negative value only queries
  return old;
}



MP_EXPORT int mp_fill_f32(void* stream, float* p, int64_t n, float value) { return fill_f32((hipStream_t)stream, p, n, value); }

// ------------------------------------------------------------------ batches of catalogues
//
// A batch of mocks on one mesh geometry: nmock catalogues stacked in one particle array,
// catalogue m holding particles moff[m] .. moff[m + 1] (moff on the device, int64, nmock + 1
// entries, moff[0] = 0).  Each phase of the per-mock spectrum (row sort, fused paint + z pass,
// y pass, x pass + binning, weight sums) runs as one launch over every catalogue.

MP_EXPORT int64_t mp_paint_partial_batch_workspace_bytes(const int32_t meshsize[3], int64_t np, int nmock, int order,
                                                         int interlacing) {
  if (!meshsize || nmock < 1) return 0;
  return (int64_t)paint_zfft_workspace_bytes(meshsize, np, nmock) + 512;
}

MP_EXPORT int mp_paint_partial_batch(void* stream, const void* positions, int positions_f64, const float* weights, int64_t np,
                                     const int64_t* moff, int nmock, const int32_t meshsize[3], const double boxsize[3],
                                     const double boxcenter[3], int order, int interlacing, void* partial, int64_t pitch,
                                     void* workspace, size_t workspace_bytes) {
  MP_REQUIRE(meshsize && boxsize && boxcenter && partial && moff, "mp_paint_partial_batch: null pointer");
  MP_REQUIRE(order >= 1 && order <= 4, "resampler order %d not in 1..4", order);
  MP_REQUIRE(nmock >= 1, "mp_paint_partial_batch: %d catalogues", nmock);
  MP_REQUIRE(np >= 0 && (np == 0 || positions), "mp_paint_partial_batch: %lld particles", (long long)np);
  MP_REQUIRE(pitch >= meshsize[2] / 2 + 1, "mp_paint_partial_batch: pitch too small");
  const int ni = interlacing < 1 ? 1 : interlacing;
  if (!paint_zfft_supported(meshsize, np, order, ni))
    return set_error(MP_ERR_UNSUPPORTED, "mp_paint_partial_batch: no fused paint for meshsize (%d, %d, %d), order %d, %lld particles",
                     meshsize[0], meshsize[1], meshsize[2], order, (long long)np);
  GridMap g = make_map(meshsize, boxsize, boxcenter);
  hipStream_t s = (hipStream_t)stream;
  const int64_t nspec = (int64_t)meshsize[0] * meshsize[1] * pitch;
  int rc = paint_zfft(s, positions, positions_f64, weights, np, g, order, ni, (float2*)partial, pitch, nspec, workspace,
                      workspace_bytes, moff, nmock);
  if (rc) return rc;
  return ypass_meshes(s, (float2*)partial, nmock * ni, nspec, meshsize, pitch);
}

MP_EXPORT int64_t mp_bin_partial_batch_workspace_bytes(const int32_t meshsize[3], int interlacing, int nprod, int nedges,
                                                       int nmock) {
  if (!meshsize || nmock < 1) return 0;
  return (int64_t)bin_partial_workspace_bytes(meshsize, interlacing < 1 ? 1 : interlacing, nprod, nedges, nmock);
}

MP_EXPORT int mp_bin_poles_partial_batch(void* stream, const void* partial, int64_t pitch, int nmock, const int32_t meshsize[3],
                                         int order, int interlacing, int compensate, int zero_dc, const double* kx,
                                         const double* ky, const double* kz, const double* k2_thresholds, int nedges,
                                         const double los[3], int nprod, const int32_t* prod_ell, const int32_t* prod_remove_zero,
                                         double* out, void* workspace, size_t workspace_bytes) {
  MP_REQUIRE(meshsize && kx && ky && kz && k2_thresholds, "mp_bin_poles_partial_batch: null pointer");
  MP_REQUIRE(order >= 1 && order <= 4, "resampler order %d not in 1..4", order);
  MP_REQUIRE(nedges >= 2, "mp_bin_poles_partial_batch: need at least 2 edges, got %d", nedges);
  BinGeom g{};
  g.n0 = meshsize[0];
  g.n1 = meshsize[1];
  g.ncols = meshsize[2] / 2 + 1;
  g.nedges = nedges;
  g.kx = kx;
  g.ky = ky;
  g.kz = kz;
  g.thr = k2_thresholds;
  return bin_partial((hipStream_t)stream, (const float2*)partial, pitch, meshsize, order, interlacing < 1 ? 1 : interlacing,
                     compensate, zero_dc, g, los, nprod, prod_ell, prod_remove_zero, out, workspace, workspace_bytes, 0, nullptr,
                     false, nmock);
}

MP_EXPORT int64_t mp_catalogue_sums_workspace_bytes(int nmock) {
  return nmock < 1 ? 0 : (int64_t)align_up((size_t)nmock * 2 * cat_blocks(nmock) * sizeof(double), 256) + 256;
}

MP_EXPORT int mp_catalogue_sums(void* stream, const float* weights, const int64_t* moff, int nmock, double* out, void* workspace,
                                size_t workspace_bytes) {
  MP_REQUIRE(moff && out, "mp_catalogue_sums: null pointer");
  MP_REQUIRE(nmock >= 1 && nmock <= 65535, "mp_catalogue_sums: %d catalogues not in 1..65535", nmock);
  hipStream_t s = (hipStream_t)stream;
  Workspace wsp(workspace, workspace_bytes);
  const int nb = cat_blocks(nmock);
  double* partial = wsp.take<double>((size_t)nmock * 2 * nb);
  if (!partial) return set_error(MP_ERR_WORKSPACE, "mp_catalogue_sums: workspace too small");
  MP_LAUNCH("catalogue_sums", catalogue_sums_kernel, dim3((unsigned)nb, (unsigned)nmock), dim3(RED_NT), 0, s, weights, moff, partial);
  MP_LAUNCH("sum_final", final_sum_kernel, dim3((unsigned)(2 * nmock)), dim3(RED_NT), 0, s, (const double*)partial, nb, out);
  return MP_OK;
}
