// Binning of half-spectrum products into |k| shells, with Legendre weights (global line of
// sight) or plain products (local line of sight), accumulated in FP64.
//
// Replaces BinMesh2SpectrumPoles (jaxpower/mesh2.py:65-119) with _get_bin_attrs / _bincount
// (jaxpower/mesh.py:2831-2910) and the binning loops of compute_mesh2_spectrum
// (jaxpower/mesh2.py:304-333, 351-369):
//   * mode weight 1 + (kz > 0) (mesh.py:2812-2828),
//   * bin = digitize(|k|, edges, right=False): |k| = fl(sqrt((kx^2 + ky^2) + kz^2)) in float64;
//     the comparison is made on k^2 against per-edge thresholds t_e = min{t : fl(sqrt(t)) >= e},
//     which is the same test and needs no device square root,
//   * remove_zero zeroes the (0,0,0) mode for the products that ask for it (mesh2.py:117-118).
//
// Mode oversampling (mesh2.py:48-61): the bin of a mode is that of the shifted vector k + s for
// one integer shift s per call; the caller sums the calls over the shifts.
//
// Work decomposition: each thread sweeps SEG consecutive kz of one (kx, ky) row; along a row
// the bin index moves by a step or two (forward for s = 0, either way otherwise) and the thread
// keeps FP64 running sums in registers, flushing them into an LDS histogram when the bin changes.  Each block writes
// its histogram once; a second kernel sums the block partials in a fixed order.
#include "bin.h"
#include "sph.h"

namespace mp {

constexpr int BIN_SEG = 16;
constexpr int BIN_NT = 256;

// v[m] for a run-time m without indexing a private array (which would live in scratch memory).
__device__ __forceinline__ float2 pick_mesh(const float2 (&v)[BIN_MAX_MESH], int m) {
  float2 r = v[0];
#pragma unroll
  for (int q = 1; q < BIN_MAX_MESH; ++q)
    if (m == q) r = v[q];
  return r;
}

// Per-block histogram of Sigma w * product * L_ell, two doubles (re, im) per (product, bin).
template <int NP, bool USE_LDS>
__global__ void __launch_bounds__(BIN_NT)
bin_poles_kernel(BinGeom g, BinProducts P, double* __restrict__ partial, double* __restrict__ out_atomic) {
  extern __shared__ double hist[];
  const int nbins = g.nedges - 1;
  const int hsize = NP * nbins * 2;
  if constexpr (USE_LDS) {
    for (int i = threadIdx.x; i < hsize; i += BIN_NT) hist[i] = 0.;
    __syncthreads();
  }
  // the histogram lives in LDS (ds_add_f64) or, for very many bins, in global memory
  double* H = USE_LDS ? hist : out_atomic;
  const int nseg = (g.ncols + BIN_SEG - 1) / BIN_SEG;
  const int64_t nwork = (int64_t)g.n0 * g.n1 * nseg;
  for (int64_t q = (int64_t)blockIdx.x * BIN_NT + threadIdx.x; q < nwork; q += (int64_t)gridDim.x * BIN_NT) {
    const int64_t row = q / nseg;
    const int seg = (int)(q % nseg);
    const int ix = (int)(row / g.n1), iy = (int)(row % g.n1);
    const int z0 = seg * BIN_SEG, z1 = min(z0 + BIN_SEG, g.ncols);
    const double kx = g.kx[ix], ky = g.ky[iy];
    const double a2 = k2_xy(kx + g.shift[0], ky + g.shift[1]);
    int b = first_bin(g.thr, g.nedges, k2_sum(a2, g.kz[z0] + g.shift[2]));
    double accr[NP], acci[NP];
#pragma unroll
    for (int p = 0; p < NP; ++p) accr[p] = acci[p] = 0.;
    const float fkx = (float)kx, fky = (float)ky;
    const float lxy = fkx * P.los[0] + fky * P.los[1];
    const float fa2 = fkx * fkx + fky * fky;
    for (int iz = z0; iz < z1; ++iz) {
      const double kzv = g.kz[iz];
      const int nb = walk_bin(g.thr, g.nedges, b, k2_sum(a2, kzv + g.shift[2]));
      if (nb != b) {
        if (b >= 1 && b < g.nedges) {
#pragma unroll
          for (int p = 0; p < NP; ++p) {
            if (accr[p] != 0.) atomicAdd(&H[(p * nbins + b - 1) * 2], accr[p]);
            if (acci[p] != 0.) atomicAdd(&H[(p * nbins + b - 1) * 2 + 1], acci[p]);
            accr[p] = acci[p] = 0.;
          }
        }
        b = nb;
      }
      if (b < 1 || b >= g.nedges) continue;
      float2 v[BIN_MAX_MESH];
#pragma unroll
      for (int m = 0; m < BIN_MAX_MESH; ++m)
        v[m] = m < P.nmesh ? P.mesh[m][row * P.pitch[m] + iz] : make_float2(0.f, 0.f);
      const float w = iz > 0 ? 2.f : 1.f;
      float mu = 0.f;
      if (P.need_mu) {
        const float fkz = (float)kzv;
        float kn2 = fa2 + fkz * fkz;
        mu = kn2 > 0.f ? (lxy + fkz * P.los[2]) * rsqrtf(kn2) : 0.f;
      }
      const bool origin = (ix == 0) & (iy == 0) & (iz == 0);
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        if (origin && P.rz[p]) continue;
        float2 a = P.pi[p] < 0 ? make_float2(1.f, 0.f) : pick_mesh(v, P.pi[p]);
        float2 c = pick_mesh(v, P.pj[p]);
        float re = a.x * c.x + a.y * c.y;  // conj(a) * c
        float im = a.x * c.y - a.y * c.x;
        float f = w * (P.ell[p] ? legendre(P.ell[p], mu) : 1.f);
        accr[p] += (double)(re * f);
        acci[p] += (double)(im * f);
      }
    }
    if (b >= 1 && b < g.nedges) {
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        if (accr[p] != 0.) atomicAdd(&H[(p * nbins + b - 1) * 2], accr[p]);
        if (acci[p] != 0.) atomicAdd(&H[(p * nbins + b - 1) * 2 + 1], acci[p]);
      }
    }
  }
  if constexpr (USE_LDS) {
    __syncthreads();
    double* dst = partial + (int64_t)blockIdx.x * hsize;
    for (int i = threadIdx.x; i < hsize; i += BIN_NT) dst[i] = hist[i];
  }
}

// out[e] = sum_b partial[b][e]: one block per entry e, fixed summation order (deterministic).
constexpr int RED_NT = 256;
__global__ void __launch_bounds__(RED_NT)
reduce_partials_kernel(const double* __restrict__ partial, int nblocks, int size, double* __restrict__ out) {
  __shared__ double sh[RED_NT];
  const int e = blockIdx.x;
  partial += (int64_t)blockIdx.y * nblocks * size;  // catalogue blockIdx.y of a batch
  out += (int64_t)blockIdx.y * size;
  double s = 0.;
  for (int b = threadIdx.x; b < nblocks; b += RED_NT) s += partial[(int64_t)b * size + e];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int d = RED_NT / 2; d > 0; d >>= 1) {
    if (threadIdx.x < d) sh[threadIdx.x] += sh[threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[e] = sh[0];
}

int reduce_partials(hipStream_t s, const double* partial, int nblocks, int size, double* out, int nmock) {
  MP_LAUNCH("bin_reduce", reduce_partials_kernel, dim3((unsigned)size, (unsigned)nmock), dim3(RED_NT), 0, s, partial, nblocks, size,
            out);
  return MP_OK;
}

__global__ void fill_f64_kernel(double* p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.;
}

// Mode counts and Sigma w |k| per bin (mesh.py:2831-2863), exact integer counts.  With
// USE_LDS each block adds into its own LDS histogram and then once per bin into the outputs
// (the run-length flushes of every thread went to ~nbins global addresses: 24.6 ms at 512^3,
// contended atomics; now one global addition per (block, bin)).
template <bool USE_LDS>
__global__ void __launch_bounds__(BIN_NT)
bin_modes_kernel(BinGeom g, unsigned long long* __restrict__ nmodes_out, double* __restrict__ ksum_out) {
  extern __shared__ double lhist[];  // [nbins] counts (as unsigned long long), then [nbins] sums
  const int nbins = g.nedges - 1;
  unsigned long long* nmodes = nmodes_out;
  double* ksum = ksum_out;
  if constexpr (USE_LDS) {
    nmodes = reinterpret_cast<unsigned long long*>(lhist);
    ksum = lhist + nbins;
    for (int i = threadIdx.x; i < 2 * nbins; i += BIN_NT) lhist[i] = 0.;  // zero bits either way
    __syncthreads();
  }
  const int nseg = (g.ncols + BIN_SEG - 1) / BIN_SEG;
  const int64_t nwork = (int64_t)g.n0 * g.n1 * nseg;
  for (int64_t q = (int64_t)blockIdx.x * BIN_NT + threadIdx.x; q < nwork; q += (int64_t)gridDim.x * BIN_NT) {
    const int64_t row = q / nseg;
    const int seg = (int)(q % nseg);
    const int ix = (int)(row / g.n1), iy = (int)(row % g.n1);
    const int z0 = seg * BIN_SEG, z1 = min(z0 + BIN_SEG, g.ncols);
    const double a2 = k2_xy(g.kx[ix] + g.shift[0], g.ky[iy] + g.shift[1]);
    int b = first_bin(g.thr, g.nedges, k2_sum(a2, g.kz[z0] + g.shift[2]));
    unsigned long long cnt = 0;
    double ks = 0.;
    for (int iz = z0; iz < z1; ++iz) {
      const double k2 = k2_sum(a2, g.kz[iz] + g.shift[2]);
      const int nb = walk_bin(g.thr, g.nedges, b, k2);
      if (nb != b) {
        if (b >= 1 && b < g.nedges && cnt) {
          atomicAdd(&nmodes[b - 1], cnt);
          atomicAdd(&ksum[b - 1], ks);
        }
        cnt = 0;
        ks = 0.;
        b = nb;
      }
      if (b < 1 || b >= g.nedges) continue;
      const unsigned w = iz > 0 ? 2u : 1u;
      cnt += w;
      ks += (double)w * sqrt(k2);
    }
    if (b >= 1 && b < g.nedges && cnt) {
      atomicAdd(&nmodes[b - 1], cnt);
      atomicAdd(&ksum[b - 1], ks);
    }
  }
  if constexpr (USE_LDS) {
    __syncthreads();
    for (int i = threadIdx.x; i < nbins; i += BIN_NT) {
      if (nmodes[i]) {
        atomicAdd(&nmodes_out[i], nmodes[i]);
        atomicAdd(&ksum_out[i], ksum[i]);
      }
    }
  }
}

static int bin_blocks(const BinGeom& g) {
  int64_t nseg = (g.ncols + BIN_SEG - 1) / BIN_SEG;
  int64_t nwork = (int64_t)g.n0 * g.n1 * nseg;
  int64_t blocks = ceil_div(nwork, BIN_NT);
  return (int)(blocks < 1024 ? (blocks < 1 ? 1 : blocks) : 1024);
}

size_t bin_workspace_bytes(int nprod, int nedges) {
  int nbins = nedges - 1;
  return align_up((size_t)1024 * nprod * nbins * 2 * sizeof(double), 256) + 256;
}

template <int NP>
static int launch_poles(hipStream_t s, const BinGeom& g, const BinProducts& P, double* out, void* ws, size_t ws_bytes) {
  const int nbins = g.nedges - 1;
  const int hsize = NP * nbins * 2;
  const int blocks = bin_blocks(g);
  const bool use_lds = hsize <= BIN_LDS_DOUBLES;
  if (use_lds) {
    Workspace w(ws, ws_bytes);
    double* partial = w.take<double>((size_t)blocks * hsize);
    if (!partial) return set_error(MP_ERR_WORKSPACE, "bin: workspace too small");
    MP_LAUNCH("bin_poles", (bin_poles_kernel<NP, true>), dim3(blocks), dim3(BIN_NT), hsize * sizeof(double), s, g, P, partial,
              (double*)nullptr);
    return reduce_partials(s, partial, blocks, hsize, out);
  } else {
    MP_LAUNCH("bin_zero", fill_f64_kernel, dim3((unsigned)ceil_div(hsize, 256)), dim3(256), 0, s, out, (int64_t)hsize);
    MP_LAUNCH("bin_poles", (bin_poles_kernel<NP, false>), dim3(blocks), dim3(BIN_NT), 0, s, g, P, (double*)nullptr, out);
  }
  return MP_OK;
}

int bin_poles(hipStream_t s, const BinGeom& g, const BinProducts& P, double* out, void* ws, size_t ws_bytes) {
  switch (P.nprod) {
    case 1: return launch_poles<1>(s, g, P, out, ws, ws_bytes);
    case 2: return launch_poles<2>(s, g, P, out, ws, ws_bytes);
    case 3: return launch_poles<3>(s, g, P, out, ws, ws_bytes);
    case 4: return launch_poles<4>(s, g, P, out, ws, ws_bytes);
    case 5: return launch_poles<5>(s, g, P, out, ws, ws_bytes);
    case 6: return launch_poles<6>(s, g, P, out, ws, ws_bytes);
    default: return set_error(MP_ERR_INVALID, "bin: %d products not in 1..%d", P.nprod, BIN_MAX_PROD);
  }
}

}  // namespace mp

using namespace mp;

static int make_geom(BinGeom& g, const int32_t meshsize[3], const double* kx, const double* ky, const double* kz, const double* thr,
                     int nedges, const double* mode_shift) {
  MP_REQUIRE(meshsize && kx && ky && kz && thr, "bin: null pointer");
  MP_REQUIRE(nedges >= 2, "bin: need at least 2 edges, got %d", nedges);
  MP_REQUIRE(meshsize[0] >= 1 && meshsize[1] >= 1 && meshsize[2] >= 1, "bin: bad meshsize");
  g.n0 = meshsize[0];
  g.n1 = meshsize[1];
  g.ncols = meshsize[2] / 2 + 1;
  g.nedges = nedges;
  g.kx = kx;
  g.ky = ky;
  g.kz = kz;
  g.thr = thr;
  for (int a = 0; a < 3; ++a) g.shift[a] = mode_shift ? mode_shift[a] : 0.;
  return MP_OK;
}

MP_EXPORT int64_t mp_bin_workspace_bytes(int nprod, int nedges) { return (int64_t)bin_workspace_bytes(nprod, nedges); }

MP_EXPORT int mp_bin_modes(void* stream, const int32_t meshsize[3], const double* kx, const double* ky, const double* kz,
                           const double* k2_thresholds, int nedges, const double mode_shift[3], int64_t* nmodes,
                           double* ksum) {
  BinGeom g{};
  int rc = make_geom(g, meshsize, kx, ky, kz, k2_thresholds, nedges, mode_shift);
  if (rc) return rc;
  MP_REQUIRE(nmodes && ksum, "mp_bin_modes: null output");
  hipStream_t s = (hipStream_t)stream;
  const int nbins = nedges - 1;
  MP_LAUNCH("bin_zero", fill_f64_kernel, dim3((unsigned)ceil_div(nbins, 256)), dim3(256), 0, s, (double*)nmodes, (int64_t)nbins);
  MP_LAUNCH("bin_zero", fill_f64_kernel, dim3((unsigned)ceil_div(nbins, 256)), dim3(256), 0, s, ksum, (int64_t)nbins);
  const size_t lds = (size_t)nbins * 16;
  if (lds <= 65536) {
    MP_LAUNCH("bin_modes", bin_modes_kernel<true>, dim3(bin_blocks(g)), dim3(BIN_NT), lds, s, g, (unsigned long long*)nmodes, ksum);
  } else {
    MP_LAUNCH("bin_modes", bin_modes_kernel<false>, dim3(bin_blocks(g)), dim3(BIN_NT), 0, s, g, (unsigned long long*)nmodes, ksum);
  }
  return MP_OK;
}

MP_EXPORT int mp_bin_poles(void* stream, int nmesh, const void* const* meshes, const int64_t* pitches, const int32_t meshsize[3],
                           const double* kx, const double* ky, const double* kz, const double* k2_thresholds, int nedges,
                           const double mode_shift[3], const double los[3], int nprod, const int32_t* prod_i, const int32_t* prod_j,
                           const int32_t* prod_ell, const int32_t* prod_remove_zero, double* out, void* workspace,
                           size_t workspace_bytes) {
  BinGeom g{};
  int rc = make_geom(g, meshsize, kx, ky, kz, k2_thresholds, nedges, mode_shift);
  if (rc) return rc;
  MP_REQUIRE(nmesh >= 1 && nmesh <= BIN_MAX_MESH, "mp_bin_poles: nmesh %d not in 1..%d", nmesh, BIN_MAX_MESH);
  MP_REQUIRE(nprod >= 1 && nprod <= BIN_MAX_PROD, "mp_bin_poles: nprod %d not in 1..%d", nprod, BIN_MAX_PROD);
  MP_REQUIRE(meshes && pitches && prod_i && prod_j && prod_ell && prod_remove_zero && out, "mp_bin_poles: null pointer");
  BinProducts P{};
  P.nmesh = nmesh;
  P.nprod = nprod;
  for (int m = 0; m < nmesh; ++m) {
    MP_REQUIRE(meshes[m], "mp_bin_poles: null mesh %d", m);
    MP_REQUIRE(pitches[m] >= g.ncols, "mp_bin_poles: pitch of mesh %d < Nz/2+1", m);
    P.mesh[m] = (const float2*)meshes[m];
    P.pitch[m] = pitches[m];
  }
  P.need_mu = 0;
  for (int p = 0; p < nprod; ++p) {
    MP_REQUIRE(prod_i[p] >= -1 && prod_i[p] < nmesh && prod_j[p] >= 0 && prod_j[p] < nmesh, "mp_bin_poles: bad mesh index");
    MP_REQUIRE(prod_ell[p] >= 0 && prod_ell[p] <= 10, "mp_bin_poles: ell %d not in 0..10", prod_ell[p]);
    P.pi[p] = prod_i[p];
    P.pj[p] = prod_j[p];
    P.ell[p] = prod_ell[p];
    P.rz[p] = prod_remove_zero[p];
    if (prod_ell[p]) P.need_mu = 1;
  }
  for (int a = 0; a < 3; ++a) P.los[a] = los ? (float)los[a] : 0.f;
  return bin_poles((hipStream_t)stream, g, P, out, workspace, workspace_bytes);
}
