// Correlation-function multipoles: binning of a real mesh in separation shells, and the
// elementwise pieces of the estimator.
//
// Replaces BinMesh2CorrelationPoles (jaxpower/mesh2.py:120-184, edges from _make_edges2 with
// kind='real', mesh2.py:25-62) and the binning of compute_mesh2_correlation
// (jaxpower/mesh2.py:377-504):
//   * separations s_a = fftfreq(N_a) * cellsize_a * N_a (mesh.py:361-403), every cell weighted 1
//     (no hermitian weights on a real mesh),
//   * bin = digitize(|s|, edges, right=False) with |s| = fl(sqrt((sx^2 + sy^2) + sz^2)) in
//     float64, tested on s^2 against the thresholds t_e = min{t : fl(sqrt(t)) >= edge e},
//   * mu = s.los / |s| with |s| := 1 at s = 0 (mesh2.py:493), the (0,0,0) cell skipped for the
//     ells that ask for remove_zero (mesh2.py:190-192),
//   * mode oversampling (mesh2.py:48-61): the bin is that of s + shift for one integer shift per
//     call (mu from the unshifted s); the caller sums the calls over the shifts.
//
// Work decomposition: each thread sweeps CB_SEG consecutive z cells of one (x, y) row.  Along a
// row |s| rises to the middle and falls after it, so the bin index moves by small steps in
// either direction; FP64 running sums stay in registers until the bin changes, then go to an
// LDS histogram that each block writes once, and a second kernel sums the block partials in a
// fixed order.
#include "bin.h"
#include "fft.h"
#include "sph.h"

namespace mp {

int fill_i32_dev(hipStream_t s, int* p, int64_t n, int v);

constexpr int CB_SEG = 32;
constexpr int CB_NT = 256;

static int real_bin_blocks(const RealBinArgs& a) {
  const int64_t nseg = ceil_div(a.n2, CB_SEG);
  const int64_t blocks = ceil_div((int64_t)a.n0 * a.n1 * nseg, CB_NT);
  return (int)(blocks < CB_MAX_BLOCKS ? (blocks < 1 ? 1 : blocks) : CB_MAX_BLOCKS);
}

// Cell counts and sum |s| per separation bin (exact integer counts).
__global__ void __launch_bounds__(CB_NT)
bin_real_modes_kernel(RealBinArgs a, unsigned long long* __restrict__ nmodes, double* __restrict__ ssum) {
  const int nseg = (a.n2 + CB_SEG - 1) / CB_SEG;
  const int64_t nwork = (int64_t)a.n0 * a.n1 * nseg;
  for (int64_t q = (int64_t)blockIdx.x * CB_NT + threadIdx.x; q < nwork; q += (int64_t)gridDim.x * CB_NT) {
    const int64_t row = q / nseg;
    const int seg = (int)(q % nseg);
    const int ix = (int)(row / a.n1), iy = (int)(row % a.n1);
    const int z0 = seg * CB_SEG, z1 = min(z0 + CB_SEG, a.n2);
    const double a2 = k2_xy(a.sx[ix] + a.shift[0], a.sy[iy] + a.shift[1]);
    int b = first_bin(a.thr, a.nedges, k2_sum(a2, a.sz[z0] + a.shift[2]));
    unsigned long long cnt = 0;
    double ss = 0.;
    for (int iz = z0; iz < z1; ++iz) {
      const double s2 = k2_sum(a2, a.sz[iz] + a.shift[2]);
      const int nb = walk_bin(a.thr, a.nedges, b, s2);
      if (nb != b) {
        if (b >= 1 && b < a.nedges && cnt) {
          atomicAdd(&nmodes[b - 1], cnt);
          atomicAdd(&ssum[b - 1], ss);
        }
        cnt = 0;
        ss = 0.;
        b = nb;
      }
      if (b < 1 || b >= a.nedges) continue;
      cnt += 1;
      ss += sqrt(s2);
    }
    if (b >= 1 && b < a.nedges && cnt) {
      atomicAdd(&nmodes[b - 1], cnt);
      atomicAdd(&ssum[b - 1], ss);
    }
  }
}

// Per-block histogram of sum mesh * P_ell(mu), one double per (ell, bin).
template <int NE>
__global__ void __launch_bounds__(CB_NT)
bin_real_kernel(RealBinArgs a, double* __restrict__ partial) {
  extern __shared__ double hist[];
  const int nbins = a.nedges - 1;
  const int hsize = NE * nbins;
  double* thr = hist + hsize;  // thresholds and z separations, read per cell
  double* szt = thr + a.nedges;
  for (int i = threadIdx.x; i < hsize; i += CB_NT) hist[i] = 0.;
  for (int i = threadIdx.x; i < a.nedges; i += CB_NT) thr[i] = a.thr[i];
  for (int i = threadIdx.x; i < a.n2; i += CB_NT) szt[i] = a.sz[i];
  __syncthreads();
  const int nseg = (a.n2 + CB_SEG - 1) / CB_SEG;
  const int64_t nwork = (int64_t)a.n0 * a.n1 * nseg;
  for (int64_t q = (int64_t)blockIdx.x * CB_NT + threadIdx.x; q < nwork; q += (int64_t)gridDim.x * CB_NT) {
    const int64_t row = q / nseg;
    const int seg = (int)(q % nseg);
    const int ix = (int)(row / a.n1), iy = (int)(row % a.n1);
    const int z0 = seg * CB_SEG, z1 = min(z0 + CB_SEG, a.n2);
    const double sxv = a.sx[ix], syv = a.sy[iy];
    const double a2 = k2_xy(sxv + a.shift[0], syv + a.shift[1]);
    int b = first_bin(thr, a.nedges, k2_sum(a2, szt[z0] + a.shift[2]));
    double acc[NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) acc[e] = 0.;
    const float fsx = (float)sxv, fsy = (float)syv;
    const float lxy = fsx * a.los[0] + fsy * a.los[1];
    const float fa2 = fsx * fsx + fsy * fsy;
    const float* src = a.mesh + row * a.n2;
    for (int iz = z0; iz < z1; ++iz) {
      const double szv = szt[iz];
      const int nb = walk_bin(thr, a.nedges, b, k2_sum(a2, szv + a.shift[2]));
      if (nb != b) {
        if (b >= 1 && b < a.nedges) {
#pragma unroll
          for (int e = 0; e < NE; ++e) {
            if (acc[e] != 0.) atomicAdd(&hist[e * nbins + b - 1], acc[e]);
            acc[e] = 0.;
          }
        }
        b = nb;
      }
      if (b < 1 || b >= a.nedges) continue;
      const float v = src[iz];
      float mu = 0.f;
      if (a.need_mu) {
        const float fsz = (float)szv;
        const float s2 = fa2 + fsz * fsz;
        mu = s2 > 0.f ? (lxy + fsz * a.los[2]) * rsqrtf(s2) : 0.f;
      }
      const bool origin = (ix == 0) & (iy == 0) & (iz == 0);
      const float m2 = mu * mu;
#pragma unroll
      for (int e = 0; e < NE; ++e) {
        if (origin && a.rz[e]) continue;
        acc[e] += (double)(v * legendre_eval(a.leg, e, mu, m2));
      }
    }
    if (b >= 1 && b < a.nedges) {
#pragma unroll
      for (int e = 0; e < NE; ++e)
        if (acc[e] != 0.) atomicAdd(&hist[e * nbins + b - 1], acc[e]);
    }
  }
  __syncthreads();
  double* dst = partial + (int64_t)blockIdx.x * hsize;
  for (int i = threadIdx.x; i < hsize; i += CB_NT) dst[i] = hist[i];
}

// The same binning for more bins than an LDS histogram holds (e.g. the unique separation shells
// of a large mesh): thresholds read from global memory and the running sums added straight to
// out with FP64 atomics (out zeroed beforehand).  The order of the additions is not fixed, so the
// sums can differ run to run in their last bits.
template <int NE>
__global__ void __launch_bounds__(CB_NT)
bin_real_global_kernel(RealBinArgs a, double* __restrict__ out) {
  const int nbins = a.nedges - 1;
  const int nseg = (a.n2 + CB_SEG - 1) / CB_SEG;
  const int64_t nwork = (int64_t)a.n0 * a.n1 * nseg;
  for (int64_t q = (int64_t)blockIdx.x * CB_NT + threadIdx.x; q < nwork; q += (int64_t)gridDim.x * CB_NT) {
    const int64_t row = q / nseg;
    const int seg = (int)(q % nseg);
    const int ix = (int)(row / a.n1), iy = (int)(row % a.n1);
    const int z0 = seg * CB_SEG, z1 = min(z0 + CB_SEG, a.n2);
    const double sxv = a.sx[ix], syv = a.sy[iy];
    const double a2 = k2_xy(sxv + a.shift[0], syv + a.shift[1]);
    int b = first_bin(a.thr, a.nedges, k2_sum(a2, a.sz[z0] + a.shift[2]));
    double acc[NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) acc[e] = 0.;
    const float fsx = (float)sxv, fsy = (float)syv;
    const float lxy = fsx * a.los[0] + fsy * a.los[1];
    const float fa2 = fsx * fsx + fsy * fsy;
    const float* src = a.mesh + row * a.n2;
    for (int iz = z0; iz < z1; ++iz) {
      const double szv = a.sz[iz];
      const int nb = walk_bin(a.thr, a.nedges, b, k2_sum(a2, szv + a.shift[2]));
      if (nb != b) {
        if (b >= 1 && b < a.nedges) {
#pragma unroll
          for (int e = 0; e < NE; ++e) {
            if (acc[e] != 0.) atomicAdd(&out[e * nbins + b - 1], acc[e]);
            acc[e] = 0.;
          }
        }
        b = nb;
      }
      if (b < 1 || b >= a.nedges) continue;
      const float v = src[iz];
      float mu = 0.f;
      if (a.need_mu) {
        const float fsz = (float)szv;
        const float s2 = fa2 + fsz * fsz;
        mu = s2 > 0.f ? (lxy + fsz * a.los[2]) * rsqrtf(s2) : 0.f;
      }
      const bool origin = (ix == 0) & (iy == 0) & (iz == 0);
      const float m2 = mu * mu;
#pragma unroll
      for (int e = 0; e < NE; ++e) {
        if (origin && a.rz[e]) continue;
        acc[e] += (double)(v * legendre_eval(a.leg, e, mu, m2));
      }
    }
    if (b >= 1 && b < a.nedges) {
#pragma unroll
      for (int e = 0; e < NE; ++e)
        if (acc[e] != 0.) atomicAdd(&out[e * nbins + b - 1], acc[e]);
    }
  }
}

static bool real_bin_fits_lds(int nell, int nedges, int n2) {
  return ((size_t)nell * (nedges - 1) + nedges + n2) * sizeof(double) <= 96 * 1024;
}

// out = conj(a) * b on half spectra, or |a|^2 (imaginary part 0) when b is null; out may be a.
__global__ void spectrum_product_kernel(const float2* a, int64_t a_pitch, const float2* b, int64_t b_pitch,
                                        int64_t nrows, int ncols, float2* out, int64_t out_pitch) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows * ncols) return;
  const int64_t r = i / ncols;
  const int c = (int)(i % ncols);
  const float2 x = a[r * a_pitch + c];
  float2 v;
  if (b) {
    const float2 y = b[r * b_pitch + c];
    v = make_float2(x.x * y.x + x.y * y.y, x.x * y.y - x.y * y.x);
  } else {
    v = make_float2(x.x * x.x + x.y * x.y, 0.f);
  }
  out[r * out_pitch + c] = v;
}

// out (+)= in * Y_lm(cx[i], cy[j], cz[k]) on a real mesh (ylm_at_zero at the zero vector).
__global__ void mul_ylm_kernel(const float* in, float* out, int n0, int n1, int n2, const double* __restrict__ cx,
                               const double* __restrict__ cy, const double* __restrict__ cz, int ell, int m, int accumulate) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)n0 * n1 * n2;
  if (i >= total) return;
  const int iz = (int)(i % n2);
  const int64_t row = i / n2;
  const int iy = (int)(row % n1), ix = (int)(row / n1);
  const float y = ylm_of_vector(ell, m, (float)cx[ix], (float)cy[iy], (float)cz[iz]);
  const float v = in[i] * y;
  out[i] = accumulate ? out[i] + v : v;
}

static int real_args(RealBinArgs& a, const float* mesh, const int32_t n[3], const double* sx, const double* sy, const double* sz,
                     const double* thr, int nedges, const double* mode_shift = nullptr) {
  MP_REQUIRE(n && sx && sy && sz && thr, "correlation binning: null pointer");
  MP_REQUIRE(nedges >= 2, "correlation binning: need at least 2 edges, got %d", nedges);
  MP_REQUIRE(n[0] >= 1 && n[1] >= 1 && n[2] >= 1, "correlation binning: bad meshsize");
  a.mesh = mesh;
  a.n0 = n[0];
  a.n1 = n[1];
  a.n2 = n[2];
  a.nedges = nedges;
  a.sx = sx;
  a.sy = sy;
  a.sz = sz;
  a.thr = thr;
  for (int d = 0; d < 3; ++d) a.shift[d] = mode_shift ? mode_shift[d] : 0.;
  return MP_OK;
}

static int real_ells(RealBinArgs& a, const double los[3], int nell, const int32_t* ells, const int32_t* remove_zero) {
  MP_REQUIRE(ells && remove_zero, "correlation binning: null ells");
  MP_REQUIRE(nell >= 1 && nell <= CB_MAX_ELL, "correlation binning: %d ells not in 1..%d", nell, CB_MAX_ELL);
  a.nell = nell;
  a.need_mu = 0;
  for (int e = 0; e < nell; ++e) {
    MP_REQUIRE(ells[e] >= 0 && ells[e] <= 10, "correlation binning: ell %d not in 0..10", ells[e]);
    a.ell[e] = ells[e];
    a.rz[e] = remove_zero[e];
    legendre_table_set(a.leg, e, ells[e]);
    if (ells[e]) a.need_mu = 1;
  }
  for (int d = 0; d < 3; ++d) a.los[d] = los ? (float)los[d] : 0.f;
  return MP_OK;
}

template <int NE>
static int launch_real_bin(hipStream_t s, const RealBinArgs& a, double* out, void* ws, size_t ws_bytes) {
  const int hsize = NE * (a.nedges - 1);
  const int blocks = real_bin_blocks(a);
  if (!real_bin_fits_lds(NE, a.nedges, a.n2)) {
    int rc = fill_i32_dev(s, reinterpret_cast<int*>(out), 2 * (int64_t)hsize, 0);
    if (rc) return rc;
    MP_LAUNCH("bin_real", bin_real_global_kernel<NE>, dim3(blocks), dim3(CB_NT), 0, s, a, out);
    return MP_OK;
  }
  Workspace w(ws, ws_bytes);
  double* partial = w.take<double>((size_t)blocks * hsize);
  if (!partial) return set_error(MP_ERR_WORKSPACE, "correlation binning: workspace too small");
  const size_t lds = ((size_t)hsize + a.nedges + a.n2) * sizeof(double);
  if (lds > 65536) MP_CHECK_HIP(hipFuncSetAttribute((const void*)bin_real_kernel<NE>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  MP_LAUNCH("bin_real", bin_real_kernel<NE>, dim3(blocks), dim3(CB_NT), lds, s, a, partial);
  return reduce_partials(s, partial, blocks, hsize, out);
}

}  // namespace mp

using namespace mp;

MP_EXPORT int mp_bin_real_modes(void* stream, const int32_t meshsize[3], const double* sx, const double* sy, const double* sz,
                                const double* s2_thresholds, int nedges, const double mode_shift[3], int64_t* nmodes,
                                double* ssum) {
  RealBinArgs a{};
  int rc = real_args(a, nullptr, meshsize, sx, sy, sz, s2_thresholds, nedges, mode_shift);
  if (rc) return rc;
  MP_REQUIRE(nmodes && ssum, "mp_bin_real_modes: null output");
  hipStream_t s = (hipStream_t)stream;
  const int nbins = nedges - 1;
  rc = fill_i32_dev(s, reinterpret_cast<int*>(nmodes), 2 * (int64_t)nbins, 0);
  if (rc) return rc;
  rc = fill_i32_dev(s, reinterpret_cast<int*>(ssum), 2 * (int64_t)nbins, 0);
  if (rc) return rc;
  MP_LAUNCH("bin_real_modes", bin_real_modes_kernel, dim3(real_bin_blocks(a)), dim3(CB_NT), 0, s, a, (unsigned long long*)nmodes, ssum);
  return MP_OK;
}

MP_EXPORT int64_t mp_bin_real_workspace_bytes(const int32_t meshsize[3], int nell, int nedges) {
  RealBinArgs a{};
  a.n0 = meshsize[0];
  a.n1 = meshsize[1];
  a.n2 = meshsize[2];
  if (!real_bin_fits_lds(nell < 1 ? 1 : nell, nedges, a.n2)) return 256;  // global atomics: no partial histograms
  return (int64_t)(align_up((size_t)real_bin_blocks(a) * (nell < 1 ? 1 : nell) * (nedges - 1 > 0 ? nedges - 1 : 1) * sizeof(double), 256) + 256);
}

MP_EXPORT int mp_bin_real_poles(void* stream, const float* mesh, const int32_t meshsize[3], const double* sx, const double* sy,
                                const double* sz, const double* s2_thresholds, int nedges, const double mode_shift[3],
                                const double los[3], int nell, const int32_t* ells, const int32_t* remove_zero, double* out,
                                void* workspace, size_t workspace_bytes) {
  RealBinArgs a{};
  int rc = real_args(a, mesh, meshsize, sx, sy, sz, s2_thresholds, nedges, mode_shift);
  if (rc) return rc;
  MP_REQUIRE(mesh && out, "mp_bin_real_poles: null pointer");
  rc = real_ells(a, los, nell, ells, remove_zero);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  switch (nell) {
    case 1: return launch_real_bin<1>(s, a, out, workspace, workspace_bytes);
    case 2: return launch_real_bin<2>(s, a, out, workspace, workspace_bytes);
    case 3: return launch_real_bin<3>(s, a, out, workspace, workspace_bytes);
    case 4: return launch_real_bin<4>(s, a, out, workspace, workspace_bytes);
    case 5: return launch_real_bin<5>(s, a, out, workspace, workspace_bytes);
    default: return launch_real_bin<6>(s, a, out, workspace, workspace_bytes);
  }
}

MP_EXPORT int64_t mp_irfft3d_bin_real_workspace_bytes(const int32_t meshsize[3], int nell, int nedges) {
  return (int64_t)irfft3d_bin_real_workspace_bytes(meshsize, nell, nedges);
}

MP_EXPORT int mp_irfft3d_bin_real(void* stream, const void* in, int64_t in_pitch, int x_done, const int32_t meshsize[3],
                                  const double* sx, const double* sy, const double* sz, const double* s2_thresholds, int nedges,
                                  const double los[3], int nell, const int32_t* ells, const int32_t* remove_zero, double* out,
                                  void* workspace, size_t workspace_bytes) {
  RealBinArgs a{};
  int rc = real_args(a, nullptr, meshsize, sx, sy, sz, s2_thresholds, nedges);
  if (rc) return rc;
  rc = real_ells(a, los, nell, ells, remove_zero);
  if (rc) return rc;
  MP_REQUIRE(in && out, "mp_irfft3d_bin_real: null pointer");
  MP_REQUIRE(real_bin_fits_lds(nell, nedges, meshsize[2]), "mp_irfft3d_bin_real: too many bins for the LDS histogram");
  return irfft3d_bin_real((hipStream_t)stream, (const float2*)in, in_pitch, meshsize, a, out, workspace, workspace_bytes, x_done);
}

MP_EXPORT int mp_spectrum_product(void* stream, const void* a, int64_t a_pitch, const void* b, int64_t b_pitch,
                                  const int32_t meshsize[3], void* out, int64_t out_pitch) {
  MP_REQUIRE(a && out && meshsize, "mp_spectrum_product: null pointer");
  const int ncols = meshsize[2] / 2 + 1;
  MP_REQUIRE(a_pitch >= ncols && out_pitch >= ncols && (!b || b_pitch >= ncols), "mp_spectrum_product: pitch < Nz/2+1");
  const int64_t nrows = (int64_t)meshsize[0] * meshsize[1];
  const int64_t total = nrows * ncols;
  if (total == 0) return MP_OK;
  MP_LAUNCH("spectrum_product", spectrum_product_kernel, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, (hipStream_t)stream,
            (const float2*)a, a_pitch, (const float2*)b, b_pitch, nrows, ncols, (float2*)out, out_pitch);
  return MP_OK;
}

MP_EXPORT int mp_mul_ylm(void* stream, const float* in, float* out, const int32_t meshsize[3], const double* cx, const double* cy,
                         const double* cz, int ell, int m, int accumulate) {
  MP_REQUIRE(in && out && meshsize && cx && cy && cz, "mp_mul_ylm: null pointer");
  MP_REQUIRE(ell >= 0 && ell <= MAX_ELL && m >= -ell && m <= ell, "mp_mul_ylm: (ell, m) = (%d, %d) not supported", ell, m);
  const int64_t total = (int64_t)meshsize[0] * meshsize[1] * meshsize[2];
  if (total == 0) return MP_OK;
  MP_LAUNCH("mul_ylm", mul_ylm_kernel, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, (hipStream_t)stream, in, out, meshsize[0],
            meshsize[1], meshsize[2], cx, cy, cz, ell, m, accumulate);
  return MP_OK;
}
