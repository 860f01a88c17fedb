// Real-to-complex and complex-to-real 3-D FFTs for power-of-two meshes on gfx950.
//
// Replaces MeshAttrs.r2c / c2r with fft_backend='jax' (jaxpower/mesh.py:1679-1719):
// forward unnormalised e^{-ikx} onto the half spectrum (Nx, Ny, Nz/2+1), inverse with 1/N^3.
//
// Layout in HBM: real meshes are C-order (x slowest, z fastest); half spectra are
// (Nx, Ny, pitch) complex64 rows with Nz/2+1 valid entries and pitch >= Nz/2+1 (our own
// buffers pad the pitch to a multiple of 16 so every row starts on a 128-byte boundary).
//
// Passes (each one HBM round trip):
//   z: rows of Nz reals viewed as Nz/2 complex -> Nz/2-point FFT in LDS -> even/odd split
//      -> Nz/2+1 outputs per row.
//   y, x: tiles of (L along the FFT axis) x (C contiguous columns) are staged in LDS with the
//      contiguous axis fastest (coalesced 128-byte row segments, conflict-free LDS access),
//      transformed by a radix-8/4/2 Stockham FFT, and written back.
//   The x pass of the interlaced paint (combine of the n_i shifted meshes) is in xpass.hip.
#include "bin.h"
#include "common.h"
#include "fft.h"
#include "fft_core.h"
#include "sph.h"
#include "tune.h"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

namespace mp {

// exp(-2 pi i e / TW_N), e in [0, TW_N): one table serves every power-of-two length <= TW_N.
__device__ float2 g_twiddle[TW_N];

static std::mutex g_tw_mutex;
static std::vector<int> g_tw_ready;  // per device

int ensure_twiddles() {
  int dev = 0;
  MP_CHECK_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_tw_mutex);
  if ((int)g_tw_ready.size() <= dev) g_tw_ready.resize(dev + 1, 0);
  if (g_tw_ready[dev]) return MP_OK;
  std::vector<float2> host(TW_N);
  for (int e = 0; e < TW_N; ++e) {
    // exact reduction of the angle to the first octant keeps the table symmetric
    double a = -2.0 * M_PI * (double)e / (double)TW_N;
    host[e] = make_float2((float)cos(a), (float)sin(a));
  }
  for (int e = 0; e < TW_N; e += TW_N / 4) {  // exact values on the axes
    int q = e / (TW_N / 4);
    const float re[4] = {1.f, 0.f, -1.f, 0.f}, im[4] = {0.f, -1.f, 0.f, 1.f};
    host[e] = make_float2(re[q], im[q]);
  }
  MP_CHECK_HIP(hipMemcpyToSymbol(HIP_SYMBOL(g_twiddle), host.data(), sizeof(float2) * TW_N));
  g_tw_ready[dev] = 1;
  return MP_OK;
}

// ------------------------------------------------------------------ launch geometry

// z passes: C rows of M = Nz/2 complex per block, rows padded by one element in LDS.
template <int M>
struct RowCfg {
  static constexpr int C = M <= 256 ? 16 : 4096 / M;
  static constexpr int LP = M + 1;
  static constexpr int NT0 = C * M / 8;
  static constexpr int NT = NT0 < 64 ? 64 : NT0;
  static constexpr size_t LDS = (size_t)C * LP * sizeof(float2);
};

// ------------------------------------------------------------------ z pass kernels

// Real rows (nrows x 2M floats) -> half-complex rows (nrows x pitch, M+1 valid).
// Optional real-space prologue: multiply each input by a real function of the node (used for
// the local line of sight), selected by `pro`.
template <int M>
__global__ void __launch_bounds__(RowCfg<M>::NT)
zpass_r2c_kernel(const float* __restrict__ in, float2* __restrict__ out, int64_t nrows, int64_t out_pitch,
                 RealPrologue pro) {
  using Cfg = RowCfg<M>;
  constexpr int C = Cfg::C, LP = Cfg::LP, NT = Cfg::NT;
  extern __shared__ float2 lds[];
  __shared__ float2 stw[2 * M];
  const int tid = threadIdx.x;
  const int64_t row0 = (int64_t)blockIdx.x * C;
  load_twiddles<2 * M, NT>(stw, g_twiddle);
  const float2* in2 = reinterpret_cast<const float2*>(in);
  for (int idx = tid; idx < C * M; idx += NT) {
    int c = idx / M, l = idx % M;
    int64_t row = row0 + c;
    float2 v = make_float2(0.f, 0.f);
    if (row < nrows) {
      v = in2 ? in2[row * M + l] : make_float2(1.f, 1.f);  // no input: the prologue alone
      if (pro.kind != PRO_NONE) {
        v.x *= pro.eval(row, 2 * l);
        v.y *= pro.eval(row, 2 * l + 1);
      }
    }
    lds[c * LP + l] = v;
  }
  __syncthreads();
  fft_lds<M, C, 1, LP, false, NT, false, 2 * M>(lds, stw);
  constexpr int NK = M / 2 + 1;
  for (int idx = tid; idx < C * NK; idx += NT) {
    int c = idx / NK, k = idx % NK;
    int64_t row = row0 + c;
    if (row >= nrows) continue;
    rfft_post<2 * M>(lds + c * LP, out + row * out_pitch, M, k, stw);
  }
}

// z pass of rho * Y_lm(xhat) for every m = -ell..ell of pro.ell: the real rows are read once
// and kept in registers; for each m they are multiplied by Y_lm of the node direction, staged
// in LDS, transformed, and written to out + (m + ell) * mesh_stride.
template <int M>
__global__ void __launch_bounds__(RowCfg<M>::NT)
zpass_r2c_ylm_kernel(const float* __restrict__ in, float2* __restrict__ out, int64_t nrows, int64_t out_pitch, int64_t mesh_stride,
                     RealPrologue pro) {
  using Cfg = RowCfg<M>;
  constexpr int C = Cfg::C, LP = Cfg::LP, NT = Cfg::NT;
  constexpr int EPT = (C * M + NT - 1) / NT;
  extern __shared__ float2 lds[];
  __shared__ float2 stw[2 * M];
  __shared__ float spz[2 * M];  // node z of every cell of a row
  const int tid = threadIdx.x;
  const int64_t row0 = (int64_t)blockIdx.x * C;
  load_twiddles<2 * M, NT>(stw, g_twiddle);
  for (int z = tid; z < 2 * M; z += NT) spz[z] = node_coord(pro.z0, pro.dz, z);
  const float2* in2 = reinterpret_cast<const float2*>(in);
  float2 reg[EPT];
  float px[EPT], py[EPT];  // node x, y of each element's row, computed once for every m
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const int idx = tid + e * NT;
    const int64_t row = row0 + idx / M;
    const bool ok = idx < C * M && row < nrows;
    reg[e] = ok ? in2[row * M + idx % M] : make_float2(0.f, 0.f);
    const int r = ok ? (int)row : 0;
    const int ix = r / pro.ny, iy = r - ix * pro.ny;
    px[e] = node_coord(pro.x0, pro.dx, ix);
    py[e] = node_coord(pro.y0, pro.dy, iy);
  }
  __syncthreads();
  constexpr int NK = M / 2 + 1;
  for (int m = -pro.ell; m <= pro.ell; ++m) {
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int idx = tid + e * NT;
      if (idx < C * M) {
        const int c = idx / M, l = idx % M;
        float2 v = reg[e];
        // opaque per m: otherwise the compiler hoists every loop-invariant term of every
        // Y_lm case out of the m loop and runs out of registers
        asm volatile("" : "+v"(px[e]), "+v"(py[e]));
        v.x *= ylm_of_vector_rsq(pro.ell, m, px[e], py[e], spz[2 * l]);
        v.y *= ylm_of_vector_rsq(pro.ell, m, px[e], py[e], spz[2 * l + 1]);
        lds[c * LP + l] = v;
      }
    }
    __syncthreads();
    fft_lds<M, C, 1, LP, false, NT, false, 2 * M>(lds, stw);
    float2* o = out + (int64_t)(m + pro.ell) * mesh_stride;
    for (int idx = tid; idx < C * NK; idx += NT) {
      const int c = idx / NK, k = idx % NK;
      const int64_t row = row0 + c;
      if (row < nrows) rfft_post<2 * M>(lds + c * LP, o + row * out_pitch, M, k, stw);
    }
    __syncthreads();
  }
}

// Half-complex rows -> real rows, times `scale`.  Imaginary parts of the k=0 and k=M entries
// are ignored, as numpy's irfftn does on its last axis.
template <int M>
__global__ void __launch_bounds__(RowCfg<M>::NT)
zpass_c2r_kernel(const float2* __restrict__ in, int64_t in_pitch, float* __restrict__ out, int64_t nrows, float scale) {
  using Cfg = RowCfg<M>;
  constexpr int C = Cfg::C, LP = Cfg::LP, NT = Cfg::NT;
  extern __shared__ float2 lds[];
  __shared__ float2 stw[2 * M];
  const int tid = threadIdx.x;
  const int64_t row0 = (int64_t)blockIdx.x * C;
  load_twiddles<2 * M, NT>(stw, g_twiddle);
  __syncthreads();
  constexpr int NK = M / 2 + 1;
  for (int idx = tid; idx < C * NK; idx += NT) {
    int c = idx / NK, k = idx % NK;
    int64_t row = row0 + c;
    float2 zk = make_float2(0.f, 0.f), zmk = make_float2(0.f, 0.f);
    if (row < nrows) {
      const float2* x = in + row * in_pitch;
      if (k == 0) {
        float x0 = x[0].x, xm = x[M].x;
        zk = make_float2(0.5f * (x0 + xm), 0.5f * (x0 - xm));
      } else {
        float2 xk = x[k], xmk = cconj(x[M - k]);
        float2 E = cscale(cadd(xk, xmk), 0.5f);
        float2 O = cmul(cscale(csub(xk, xmk), 0.5f), cconj(stw[k]));
        zk = make_float2(E.x - O.y, E.y + O.x);        // E + i O
        zmk = make_float2(E.x + O.y, -E.y + O.x);      // conj(E) + i conj(O)
      }
    }
    lds[c * LP + k] = zk;
    if (k != 0 && 2 * k != M) lds[c * LP + (M - k)] = zmk;
  }
  __syncthreads();
  fft_lds<M, C, 1, LP, false, NT, true, 2 * M>(lds, stw);
  float2* out2 = reinterpret_cast<float2*>(out);
  for (int idx = tid; idx < C * M; idx += NT) {
    int c = idx / M, l = idx % M;
    int64_t row = row0 + c;
    if (row < nrows) out2[row * M + l] = cscale(lds[c * LP + l], scale);
  }
}

// The z pass of a convolution by a real selection mesh, in place: each half-complex row
// (already through its inverse x and y passes) is taken back to real space in LDS (as in
// zpass_c2r_kernel), multiplied point by point by the matching real row of Q times `scale`,
// and transformed forward again (as in zpass_r2c_kernel).  The real row is never written.
template <int M>
__global__ void __launch_bounds__(RowCfg<M>::NT)
zpass_mul_real_kernel(float2* __restrict__ data, int64_t pitch, const float* __restrict__ Q, int64_t nrows, float scale) {
  using Cfg = RowCfg<M>;
  constexpr int C = Cfg::C, LP = Cfg::LP, NT = Cfg::NT;
  extern __shared__ float2 lds[];
  __shared__ float2 stw[2 * M];
  const int tid = threadIdx.x;
  const int64_t row0 = (int64_t)blockIdx.x * C;
  load_twiddles<2 * M, NT>(stw, g_twiddle);
  __syncthreads();
  constexpr int NK = M / 2 + 1;
  for (int idx = tid; idx < C * NK; idx += NT) {
    int c = idx / NK, k = idx % NK;
    int64_t row = row0 + c;
    float2 zk = make_float2(0.f, 0.f), zmk = make_float2(0.f, 0.f);
    if (row < nrows) {
      const float2* x = data + row * pitch;
      if (k == 0) {
        float x0 = x[0].x, xm = x[M].x;
        zk = make_float2(0.5f * (x0 + xm), 0.5f * (x0 - xm));
      } else {
        float2 xk = x[k], xmk = cconj(x[M - k]);
        float2 E = cscale(cadd(xk, xmk), 0.5f);
        float2 O = cmul(cscale(csub(xk, xmk), 0.5f), cconj(stw[k]));
        zk = make_float2(E.x - O.y, E.y + O.x);
        zmk = make_float2(E.x + O.y, -E.y + O.x);
      }
    }
    lds[c * LP + k] = zk;
    if (k != 0 && 2 * k != M) lds[c * LP + (M - k)] = zmk;
  }
  __syncthreads();
  fft_lds<M, C, 1, LP, false, NT, true, 2 * M>(lds, stw);
  const float2* q2 = reinterpret_cast<const float2*>(Q);
  for (int idx = tid; idx < C * M; idx += NT) {
    int c = idx / M, l = idx % M;
    int64_t row = row0 + c;
    const float2 qv = row < nrows ? q2[row * M + l] : make_float2(0.f, 0.f);
    const float2 v = lds[c * LP + l];
    lds[c * LP + l] = make_float2(v.x * (qv.x * scale), v.y * (qv.y * scale));
  }
  __syncthreads();
  fft_lds<M, C, 1, LP, false, NT, false, 2 * M>(lds, stw);
  for (int idx = tid; idx < C * NK; idx += NT) {
    int c = idx / NK, k = idx % NK;
    int64_t row = row0 + c;
    if (row >= nrows) continue;
    rfft_post<2 * M>(lds + c * LP, data + row * pitch, M, k, stw);
  }
}

// The z c2r pass fused with the separation-shell binning of the real rows it produces (the
// correlation function of a product spectrum, jaxpower/mesh2.py:471-498): the rows never
// leave LDS.  Each thread then sweeps RUN consecutive complex elements (2 RUN reals) of one
// row, keeping FP64 running sums per multipole until the bin changes (along a row |s| rises
// to the middle and falls after it, so the bin moves by small steps either way), then adds
// them into an LDS histogram.  Persistent blocks; each writes its histogram once.
template <int M, int NE>
__global__ void __launch_bounds__(RowCfg<M>::NT)
zpass_c2r_bin_kernel(const float2* __restrict__ in, int64_t in_pitch, int64_t nrows, float scale, RealBinArgs a,
                     double* __restrict__ partial) {
  using Cfg = RowCfg<M>;
  constexpr int C = Cfg::C, LP = Cfg::LP, NT = Cfg::NT;
  constexpr int TPR = NT / C;
  constexpr int RUN = (M + TPR - 1) / TPR;
  extern __shared__ float2 lds[];
  double* hist = reinterpret_cast<double*>(lds + C * LP);
  double* thr = hist + NE * (a.nedges - 1);  // thresholds and z separations, read per element
  double* szt = thr + a.nedges;
  __shared__ float2 stw[2 * M];
  const int tid = threadIdx.x;
  const int nbins = a.nedges - 1, hsize = NE * nbins;
  load_twiddles<2 * M, NT>(stw, g_twiddle);
  for (int i = tid; i < hsize; i += NT) hist[i] = 0.;
  for (int i = tid; i < a.nedges; i += NT) thr[i] = a.thr[i];
  for (int i = tid; i < 2 * M; i += NT) szt[i] = a.sz[i];
  __syncthreads();
  constexpr int NK = M / 2 + 1;
  const int64_t ngroups = (nrows + C - 1) / C;
  for (int64_t grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    const int64_t row0 = grp * C;
    for (int idx = tid; idx < C * NK; idx += NT) {
      int c = idx / NK, k = idx % NK;
      int64_t row = row0 + c;
      float2 zk = make_float2(0.f, 0.f), zmk = make_float2(0.f, 0.f);
      if (row < nrows) {
        const float2* x = in + row * in_pitch;
        if (k == 0) {
          float x0 = x[0].x, xm = x[M].x;
          zk = make_float2(0.5f * (x0 + xm), 0.5f * (x0 - xm));
        } else {
          float2 xk = x[k], xmk = cconj(x[M - k]);
          float2 E = cscale(cadd(xk, xmk), 0.5f);
          float2 O = cmul(cscale(csub(xk, xmk), 0.5f), cconj(stw[k]));
          zk = make_float2(E.x - O.y, E.y + O.x);
          zmk = make_float2(E.x + O.y, -E.y + O.x);
        }
      }
      lds[c * LP + k] = zk;
      if (k != 0 && 2 * k != M) lds[c * LP + (M - k)] = zmk;
    }
    __syncthreads();
    fft_lds<M, C, 1, LP, false, NT, true, 2 * M>(lds, stw);
    const int c = tid / TPR, l0 = (tid % TPR) * RUN;
    const int64_t row = row0 + c;
    if (c < C && row < nrows && l0 < M) {
      const int ix = (int)(row / a.n1), iy = (int)(row % a.n1);
      const double sxv = a.sx[ix], syv = a.sy[iy];
      const double a2 = k2_xy(sxv, syv);
      int b = first_bin(thr, a.nedges, k2_sum(a2, szt[2 * l0]));
      double acc[NE];
#pragma unroll
      for (int e = 0; e < NE; ++e) acc[e] = 0.;
      const float fsx = (float)sxv, fsy = (float)syv;
      const float lxy = fsx * a.los[0] + fsy * a.los[1];
      const float fa2 = fsx * fsx + fsy * fsy;
      const int l1 = min(l0 + RUN, M);
      for (int l = l0; l < l1; ++l) {
        const float2 v2 = lds[c * LP + l];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int iz = 2 * l + h;
          const double szv = szt[iz];
          const int nb = walk_bin(thr, a.nedges, b, k2_sum(a2, szv));
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
          const float v = (h ? v2.y : v2.x) * scale;
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
      }
      if (b >= 1 && b < a.nedges) {
#pragma unroll
        for (int e = 0; e < NE; ++e)
          if (acc[e] != 0.) atomicAdd(&hist[e * nbins + b - 1], acc[e]);
      }
    }
    __syncthreads();
  }
  double* dst = partial + (int64_t)blockIdx.x * hsize;
  for (int i = tid; i < hsize; i += NT) dst[i] = hist[i];
}

// ------------------------------------------------------------------ column passes

// Generic column FFT: block (chunk, outer) transforms columns c0..c0+C-1 (c < ncols) of
// in[outer * in_ostride + l * in_fstride + c] into out[...] (may alias in).
template <int L, bool INV>
__global__ void __launch_bounds__(ColCfg<L>::NT)
col_fft_kernel(const float2* in, float2* out, int ncols, int64_t in_fstride, int64_t in_ostride,
               int64_t out_fstride, int64_t out_ostride, float scale) {
  using Cfg = ColCfg<L>;
  constexpr int C = Cfg::C, NT = Cfg::NT;
  extern __shared__ float2 lds[];
  __shared__ float2 stw[L];
  const int tid = threadIdx.x;
  const int c0 = blockIdx.x * C;
  load_twiddles<L, NT>(stw, g_twiddle);
  const float2* src = in + (int64_t)blockIdx.y * in_ostride + c0;
  for (int idx = tid; idx < C * L; idx += NT) {
    int c = idx % C, l = idx / C;
    lds[l * C + c] = (c0 + c < ncols) ? src[l * in_fstride + c] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft_lds<L, C, C, 1, true, NT, INV, L>(lds, stw);
  float2* dst = out + (int64_t)blockIdx.y * out_ostride + c0;
  for (int idx = tid; idx < C * L; idx += NT) {
    int c = idx % C, l = idx / C;
    if (c0 + c < ncols) dst[l * out_fstride + c] = cscale(lds[l * C + c], scale);
  }
}

// Inverse FFT along x of a generated half spectrum (the window's |k| shell): as
// col_fft_kernel<L, true> with the input of column c, row l of slice blockIdx.y computed from
// the mode indices (l, blockIdx.y, c) instead of read.
template <int L>
__global__ void __launch_bounds__(ColCfg<L>::NT)
xpass_inv_shell_kernel(ShellGen g, float2* out, int ncols, int64_t out_fstride, int64_t out_ostride) {
  using Cfg = ColCfg<L>;
  constexpr int C = Cfg::C, NT = Cfg::NT;
  extern __shared__ float2 lds[];
  __shared__ float2 stw[L];
  const int tid = threadIdx.x;
  const int c0 = blockIdx.x * C, y = blockIdx.y;
  load_twiddles<L, NT>(stw, g_twiddle);
  for (int idx = tid; idx < C * L; idx += NT) {
    int c = idx % C, l = idx / C;
    lds[l * C + c] = make_float2((c0 + c < ncols) ? g.eval(l, y, c0 + c) : 0.f, 0.f);
  }
  __syncthreads();
  fft_lds<L, C, C, 1, true, NT, true, L>(lds, stw);
  float2* dst = out + (int64_t)y * out_ostride + c0;
  for (int idx = tid; idx < C * L; idx += NT) {
    int c = idx % C, l = idx / C;
    if (c0 + c < ncols) dst[l * out_fstride + c] = lds[l * C + c];
  }
}

// Pipelined column pass over several meshes: persistent blocks walk tiles (mesh, outer,
// chunk of C columns); the global loads of the next tile are issued before the FFT of the
// current one, and the tile is written back in the order it was read (in place allowed).
struct ColPassArgs {
  float2* data;
  int64_t mesh_stride, fstride, ostride;  // between meshes, FFT rows, outer slices
  int ncols, nouter, nmesh, nchunks, ntiles;
  const float2* tw;
};

template <int L, bool INV, int CC = XCfg<L>::C>
__global__ void __launch_bounds__((XCfgC<L, CC>::NT))
colpass_kernel(ColPassArgs a) {
  using Cfg = XCfgC<L, CC>;
  constexpr int C = Cfg::C, NT = Cfg::NT, EPT = Cfg::EPT;
  extern __shared__ float2 lds[];
  __shared__ float2 stw[L];
  const int tid = threadIdx.x;
  load_twiddles<L, NT>(stw, a.tw);
  float2 reg[EPT];
  auto base = [&](int tile) {
    const int chunk = tile % a.nchunks, rest = tile / a.nchunks;
    const int o = rest % a.nouter, m = rest / a.nouter;
    return a.data + m * a.mesh_stride + (int64_t)o * a.ostride + chunk * C;
  };
  auto valid = [&](int tile, int cc) { return (tile % a.nchunks) * C + cc < a.ncols; };
  auto load = [&](int tile) {
    const float2* src = base(tile);
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int idx = tid + e * NT, cc = idx % C, l = idx / C;
      if (!valid(tile, cc)) reg[e] = make_float2(0.f, 0.f);
      else if constexpr (C >= 16) reg[e] = load_stream(src + l * a.fstride + cc);
      else reg[e] = src[l * a.fstride + cc];
    }
  };
  // Tiles narrower than a 128-byte line share lines with their neighbours: at L >= 1024 those
  // take XCD-contiguous runs (1024^3 y pass 11.64 -> 11.41 ms); at L = 512 the runs made the
  // pass 0.61 -> 0.79 ms with 16 or 8 columns, so it keeps the round-robin order.
  int tile = (C < 16 && L >= 1024) ? xcd_first_tile() : (int)blockIdx.x;
  if (tile < a.ntiles) load(tile);
  while (tile < a.ntiles) {  // (the first sync below also publishes the twiddle table)
    const int next = tile + gridDim.x;
    float2* dst = base(tile);
    if constexpr (RegFFT<L, C, NT>::ok) {
      // first and last stages on registers; the middle stages in LDS
      fft_first_from_regs<L, C, NT, INV>(reg, lds);
      __syncthreads();
      if (next < a.ntiles) load(next);
      fft_stages_until<L, C, C, 1, true, NT, INV, RegFFT<L, C, NT>::R1, RegFFT<L, C, NT>::NSLAST, L>(lds, stw);
      float2 out[EPT];
      fft_last_to_regs<L, C, NT, INV, L>(lds, out, stw);
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const int idx = tid + e * NT, cc = idx % C, l = idx / C;
        if (!valid(tile, cc)) continue;
        if constexpr (C >= 16) store_stream(dst + l * a.fstride + cc, out[e]);
        else dst[l * a.fstride + cc] = out[e];
      }
    } else {
#pragma unroll
      for (int e = 0; e < EPT; ++e) lds[tid + e * NT] = reg[e];
      __syncthreads();
      if (next < a.ntiles) load(next);
      fft_lds<L, C, C, 1, true, NT, INV, L>(lds, stw);
#pragma unroll
      for (int e = 0; e < EPT; ++e) {
        const int idx = tid + e * NT, cc = idx % C, l = idx / C;
        if (!valid(tile, cc)) continue;
        if constexpr (C >= 16) store_stream(dst + l * a.fstride + cc, lds[idx]);
        else dst[l * a.fstride + cc] = lds[idx];
      }
    }
    __syncthreads();
    tile = next;
  }
}

// Local line of sight: acc(kx, ky, kz) += FFT_x(in)(kx) * Y_lm(k) for one (l, m).
template <int L>
__global__ void __launch_bounds__(ColCfg<L>::NT)
xpass_ylm_acc_kernel(const float2* __restrict__ in, float2* __restrict__ acc, int ncols, int64_t in_fstride,
                     int64_t in_ostride, int64_t acc_fstride, int64_t acc_ostride, KPrologue kp, int first) {
  using Cfg = ColCfg<L>;
  constexpr int C = Cfg::C, NT = Cfg::NT;
  extern __shared__ float2 lds[];
  __shared__ float2 stw[L];
  const int tid = threadIdx.x;
  const int c0 = blockIdx.x * C;
  const int y = blockIdx.y;
  load_twiddles<L, NT>(stw, g_twiddle);
  const float2* src = in + (int64_t)y * in_ostride + c0;
  for (int idx = tid; idx < C * L; idx += NT) {
    int c = idx % C, l = idx / C;
    lds[l * C + c] = (c0 + c < ncols) ? src[l * in_fstride + c] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft_lds<L, C, C, 1, true, NT, false, L>(lds, stw);
  float2* dst = acc + (int64_t)y * acc_ostride + c0;
  for (int idx = tid; idx < C * L; idx += NT) {
    int c = idx % C, l = idx / C;
    if (c0 + c < ncols) {
      float yk = kp.eval(l, y, c0 + c);
      float2 v = cscale(lds[l * C + c], yk);
      float2* d = dst + l * acc_fstride + c;
      *d = first ? v : cadd(*d, v);
    }
  }
}

// ------------------------------------------------------------------ host dispatch

template <int M>
static int launch_zpass_r2c(const float* in, float2* out, int64_t nrows, int64_t pitch, const RealPrologue& pro,
                            hipStream_t stream) {
  using Cfg = RowCfg<M>;
  dim3 grid((unsigned)ceil_div(nrows, Cfg::C));
  MP_LAUNCH("fft_zpass_r2c", zpass_r2c_kernel<M>, grid, dim3(Cfg::NT), Cfg::LDS, stream, in, out, nrows, pitch, pro);
  return MP_OK;
}

template <int M>
static int launch_zpass_r2c_ylm(const float* in, float2* out, int64_t nrows, int64_t pitch, int64_t mstride, const RealPrologue& pro,
                                hipStream_t stream) {
  using Cfg = RowCfg<M>;
  dim3 grid((unsigned)ceil_div(nrows, Cfg::C));
  MP_LAUNCH("fft_zpass_r2c_ylm", zpass_r2c_ylm_kernel<M>, grid, dim3(Cfg::NT), Cfg::LDS, stream, in, out, nrows, pitch, mstride, pro);
  return MP_OK;
}

template <int M>
static int launch_zpass_c2r(const float2* in, int64_t pitch, float* out, int64_t nrows, float scale, hipStream_t stream) {
  using Cfg = RowCfg<M>;
  dim3 grid((unsigned)ceil_div(nrows, Cfg::C));
  MP_LAUNCH("fft_zpass_c2r", zpass_c2r_kernel<M>, grid, dim3(Cfg::NT), Cfg::LDS, stream, in, pitch, out, nrows, scale);
  return MP_OK;
}

template <int L>
static int launch_xpass_inv_shell(const ShellGen& g, float2* out, int ncols, int ny, int64_t ofs, int64_t oos, hipStream_t stream) {
  using Cfg = ColCfg<L>;
  dim3 grid((unsigned)ceil_div(ncols, Cfg::C), (unsigned)ny);
  MP_LAUNCH("window_xpass_inv", xpass_inv_shell_kernel<L>, grid, dim3(Cfg::NT), Cfg::LDS, stream, g, out, ncols, ofs, oos);
  return MP_OK;
}

template <int M>
static int launch_zpass_mul_real(float2* data, int64_t pitch, const float* Q, int64_t nrows, float scale, hipStream_t stream) {
  using Cfg = RowCfg<M>;
  dim3 grid((unsigned)ceil_div(nrows, Cfg::C));
  MP_LAUNCH("window_zpass", zpass_mul_real_kernel<M>, grid, dim3(Cfg::NT), Cfg::LDS, stream, data, pitch, Q, nrows, scale);
  return MP_OK;
}

template <int L, bool INV>
static int launch_col(const char* name, const float2* in, float2* out, int ncols, int64_t nouter, int64_t ifs, int64_t ios,
                      int64_t ofs, int64_t oos, float scale, hipStream_t stream) {
  using Cfg = ColCfg<L>;
  dim3 grid((unsigned)ceil_div(ncols, Cfg::C), (unsigned)nouter);
  MP_LAUNCH(name, (col_fft_kernel<L, INV>), grid, dim3(Cfg::NT), Cfg::LDS, stream, in, out, ncols, ifs, ios, ofs, oos, scale);
  return MP_OK;
}

template <int L>
static int launch_ylm_acc(const float2* in, float2* acc, int ncols, int ny, int64_t ifs, int64_t ios, int64_t afs, int64_t aos,
                          const KPrologue& kp, int first, hipStream_t stream) {
  using Cfg = ColCfg<L>;
  dim3 grid((unsigned)ceil_div(ncols, Cfg::C), (unsigned)ny);
  MP_LAUNCH("fft_xpass_ylm_acc", xpass_ylm_acc_kernel<L>, grid, dim3(Cfg::NT), Cfg::LDS, stream, in, acc, ncols, ifs, ios, afs, aos,
            kp, first);
  return MP_OK;
}

#define MP_POW2_SWITCH(n, F, ...)         \
  switch (n) {                            \
    case 1: return F<1>(__VA_ARGS__);     \
    case 2: return F<2>(__VA_ARGS__);     \
    case 4: return F<4>(__VA_ARGS__);     \
    case 8: return F<8>(__VA_ARGS__);     \
    case 16: return F<16>(__VA_ARGS__);   \
    case 32: return F<32>(__VA_ARGS__);   \
    case 64: return F<64>(__VA_ARGS__);   \
    case 128: return F<128>(__VA_ARGS__); \
    case 256: return F<256>(__VA_ARGS__); \
    case 512: return F<512>(__VA_ARGS__); \
    case 1024: return F<1024>(__VA_ARGS__); \
    case 2048: return F<2048>(__VA_ARGS__); \
    default: return set_error(MP_ERR_UNSUPPORTED, "FFT length %d is not a power of two in [1, 2048]", (int)(n)); \
  }

static int zpass_r2c(int M, const float* in, float2* out, int64_t nrows, int64_t pitch, const RealPrologue& pro, hipStream_t s) {
  if (M > 1024) return set_error(MP_ERR_UNSUPPORTED, "z length %d too large", 2 * M);
  MP_POW2_SWITCH(M, launch_zpass_r2c, in, out, nrows, pitch, pro, s);
}

static int zpass_c2r(int M, const float2* in, int64_t pitch, float* out, int64_t nrows, float scale, hipStream_t s) {
  if (M > 1024) return set_error(MP_ERR_UNSUPPORTED, "z length %d too large", 2 * M);
  MP_POW2_SWITCH(M, launch_zpass_c2r, in, pitch, out, nrows, scale, s);
}

template <int L>
static int launch_col_fwd(const char* name, const float2* in, float2* out, int ncols, int64_t nouter, int64_t ifs, int64_t ios,
                          int64_t ofs, int64_t oos, float scale, hipStream_t s) {
  return launch_col<L, false>(name, in, out, ncols, nouter, ifs, ios, ofs, oos, scale, s);
}
template <int L>
static int launch_col_inv(const char* name, const float2* in, float2* out, int ncols, int64_t nouter, int64_t ifs, int64_t ios,
                          int64_t ofs, int64_t oos, float scale, hipStream_t s) {
  return launch_col<L, true>(name, in, out, ncols, nouter, ifs, ios, ofs, oos, scale, s);
}

static int col_fft(int L, bool inv, const char* name, const float2* in, float2* out, int ncols, int64_t nouter, int64_t ifs,
                   int64_t ios, int64_t ofs, int64_t oos, float scale, hipStream_t s) {
  if (L < 2) return set_error(MP_ERR_UNSUPPORTED, "column FFT length %d < 2", L);
  if (inv) {
    MP_POW2_SWITCH(L, launch_col_inv, name, in, out, ncols, nouter, ifs, ios, ofs, oos, scale, s);
  }
  MP_POW2_SWITCH(L, launch_col_fwd, name, in, out, ncols, nouter, ifs, ios, ofs, oos, scale, s);
}

template <int L, int CC = XCfg<L>::C>
static int launch_colpass(hipStream_t s, ColPassArgs& a) {
  using Cfg = XCfgC<L, CC>;
  auto kern = colpass_kernel<L, false, CC>;
  a.nchunks = (int)ceil_div(a.ncols, Cfg::C);
  a.ntiles = a.nchunks * a.nouter * a.nmesh;
  int per = 0;
  MP_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, Cfg::NT, Cfg::LDS));
  if (per < 1) per = 1;
  int dev = 0, ncu = 0;
  MP_CHECK_HIP(hipGetDevice(&dev));
  MP_CHECK_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  int blocks = ncu * per;
  if (blocks > a.ntiles) blocks = a.ntiles;
  MP_LAUNCH("fft_ypass", kern, dim3((unsigned)blocks), dim3(Cfg::NT), Cfg::LDS, s, a);
  return MP_OK;
}

// Tile widths the autotuner tries for the y pass (candidate 0 is the default): half of it while
// at least 4 columns remain, and twice it up to 64 KiB tiles (0: does not apply; lengths below 64
// keep the default).
template <int L>
struct YPTune {
  static constexpr int C0 = XCfg<L>::C;
  static constexpr int C1 = L >= 64 && C0 >= 8 ? C0 / 2 : 0;
  static constexpr int C2 = L >= 64 && 2 * C0 * L * 8 <= 65536 ? 2 * C0 : 0;
};
constexpr size_t YP_TUNE_BYTES = (size_t)512 << 20;  // larger than the last-level cache

// The y pass with its tile width chosen by the autotuner per (L, columns, x rows, meshes).  The
// pass works in place, so candidates are timed on a scratch copy of the first x rows of the
// data (at most 512 MiB, allocated for the measurement only: with a copy that fit the 256 MB
// last-level cache the candidates were ranked by their cache-resident speed, and an 8-column
// tile that took 0.79 ms on the whole 512^3 data against 0.61 for 16 columns came first).
template <int L>
static int launch_colpass_tuned(hipStream_t s, ColPassArgs& a) {
  using T = YPTune<L>;
  auto run = [&](int q, ColPassArgs& b) -> int {
    if constexpr (T::C1 > 0) {
      if (q == 1) return launch_colpass<L, T::C1>(s, b);
    }
    if constexpr (T::C2 > 0) {
      if (q == 2) return launch_colpass<L, T::C2>(s, b);
    }
    return launch_colpass<L, T::C0>(s, b);
  };
  const bool ok[3] = {true, T::C1 > 0, T::C2 > 0};
  const int forced = tune_forced(TK_YPASS_COLS);
  if (forced >= 0) return run(forced < 3 && ok[forced] ? forced : 0, a);
  char key[160];
  snprintf(key, sizeof(key), "ypass_cols L=%d ncols=%d nouter=%d meshes=%d", L, a.ncols, a.nouter, a.nmesh);
  const int q0 = tune_get(key);
  if (q0 >= 0) return run(q0 < 3 && ok[q0] ? q0 : 0, a);
  if (!(ok[1] || ok[2]) || !tune_measure_ok(s)) return run(0, a);
  const size_t row_bytes = (size_t)a.ostride * sizeof(float2);  // one x row of one mesh
  const int64_t rows = std::max<int64_t>(1, std::min<int64_t>(a.nouter, (int64_t)(YP_TUNE_BYTES / row_bytes)));
  float2* scratch = nullptr;
  MP_CHECK_HIP(hipMalloc(&scratch, (size_t)rows * row_bytes));
  ColPassArgs b = a;
  b.data = scratch;
  b.nouter = (int)rows;
  b.nmesh = 1;
  int rc = MP_OK;
  if (hipMemcpyAsync(scratch, a.data, (size_t)rows * row_bytes, hipMemcpyDeviceToDevice, s) != hipSuccess)
    rc = set_error(MP_ERR_HIP, "y pass tuning: copy to the scratch rows");
  float ms[3];
  const int best = rc ? -1 : tune_time(s, 3, ok, [&](int q) { return run(q, b); }, ms);
  if (hipStreamSynchronize(s) != hipSuccess && !rc) rc = set_error(MP_ERR_HIP, "y pass tuning: synchronize");
  (void)hipFree(scratch);
  if (rc) return rc;
  if (best < 0) return MP_ERR_HIP;
  tune_put(key, TK_YPASS_COLS, best, ms, 3);
  return run(best, a);
}

static int xpass_ylm_acc(int L, const float2* in, float2* acc, int ncols, int ny, int64_t ifs, int64_t ios, int64_t afs,
                         int64_t aos, const KPrologue& kp, int first, hipStream_t s) {
  if (L < 2) return set_error(MP_ERR_UNSUPPORTED, "x length %d < 2", L);
  MP_POW2_SWITCH(L, launch_ylm_acc, in, acc, ncols, ny, ifs, ios, afs, aos, kp, first, s);
}

// Check the LDS attribute once per kernel instantiation that needs more than 64 KiB (none
// with the current tile sizes: every configuration is <= 64 KiB).
static int check_shape(const int32_t n[3]) {
  for (int a = 0; a < 3; ++a) {
    if (!is_pow2(n[a])) return set_error(MP_ERR_UNSUPPORTED, "meshsize[%d] = %d is not a power of two", a, n[a]);
    if (n[a] > 2048) return set_error(MP_ERR_UNSUPPORTED, "meshsize[%d] = %d > 2048", a, n[a]);
  }
  if (n[0] < 2 || n[1] < 2 || n[2] < 2) return set_error(MP_ERR_UNSUPPORTED, "meshsize must be >= 2 per axis");
  return MP_OK;
}

bool pow2_shape(const int32_t n[3]) {
  for (int a = 0; a < 3; ++a)
    if (!is_pow2(n[a]) || n[a] < 2 || n[a] > 2048) return false;
  return true;
}

// Any-length r2c along z then complex passes along y (and x): the specialised passes apply only
// when every axis is a power of two.
static int rfft3d_any(hipStream_t s, const float* in, float2* out, const int32_t n[3], int64_t pitch, const RealPrologue& pro,
                      bool with_x) {
  if (!fft_any_shape_ok(n)) return set_error(MP_ERR_UNSUPPORTED, "meshsize (%d, %d, %d) not in 1..2048", n[0], n[1], n[2]);
  const int ncols = n[2] / 2 + 1;
  if (pitch < ncols) return set_error(MP_ERR_INVALID, "pitch %lld < Nz/2+1", (long long)pitch);
  const int64_t nx = n[0], ny = n[1];
  int rc = any_zpass_r2c(s, in, out, n[2], nx * ny, pitch, pro);
  if (rc) return rc;
  rc = any_col_fft(s, out, out, n[1], ncols, nx, pitch, ny * pitch, pitch, ny * pitch, false, 1.f, "fft_ypass_any");
  if (rc || !with_x) return rc;
  return any_col_fft(s, out, out, n[0], ncols, ny, ny * pitch, pitch, ny * pitch, pitch, false, 1.f, "fft_xpass_any");
}

// ------------------------------------------------------------------ internal API

int rfft3d(hipStream_t s, const float* in, float2* out, const int32_t n[3], int64_t pitch, const RealPrologue& pro) {
  if (!pow2_shape(n)) return rfft3d_any(s, in, out, n, pitch, pro, true);
  int rc = check_shape(n);
  if (rc) return rc;
  rc = ensure_twiddles();
  if (rc) return rc;
  const int M = n[2] / 2, ncols = M + 1;
  if (pitch < ncols) return set_error(MP_ERR_INVALID, "pitch %lld < Nz/2+1", (long long)pitch);
  const int64_t nx = n[0], ny = n[1];
  rc = zpass_r2c(M, in, out, nx * ny, pitch, pro, s);
  if (rc) return rc;
  rc = col_fft(n[1], false, "fft_ypass", out, out, ncols, nx, pitch, ny * pitch, pitch, ny * pitch, 1.f, s);
  if (rc) return rc;
  return col_fft(n[0], false, "fft_xpass", out, out, ncols, ny, ny * pitch, pitch, ny * pitch, pitch, 1.f, s);
}

int rfft3d_zy(hipStream_t s, const float* in, float2* out, const int32_t n[3], int64_t pitch, const RealPrologue& pro) {
  if (!pow2_shape(n)) return rfft3d_any(s, in, out, n, pitch, pro, false);
  int rc = check_shape(n);
  if (rc) return rc;
  rc = ensure_twiddles();
  if (rc) return rc;
  const int M = n[2] / 2, ncols = M + 1;
  const int64_t nx = n[0], ny = n[1];
  rc = zpass_r2c(M, in, out, nx * ny, pitch, pro, s);
  if (rc) return rc;
  return col_fft(n[1], false, "fft_ypass", out, out, ncols, nx, pitch, ny * pitch, pitch, ny * pitch, 1.f, s);
}

int ypass_meshes(hipStream_t s, float2* data, int nmesh, int64_t mesh_stride, const int32_t n[3], int64_t pitch) {
  if (!pow2_shape(n)) {
    if (!fft_any_shape_ok(n)) return set_error(MP_ERR_UNSUPPORTED, "meshsize (%d, %d, %d) not in 1..2048", n[0], n[1], n[2]);
    const int ncols = n[2] / 2 + 1;
    const int64_t ny = n[1];
    for (int m = 0; m < nmesh; ++m) {
      float2* d = data + m * mesh_stride;
      int rc = any_col_fft(s, d, d, n[1], ncols, n[0], pitch, ny * pitch, pitch, ny * pitch, false, 1.f, "fft_ypass_any");
      if (rc) return rc;
    }
    return MP_OK;
  }
  int rc = check_shape(n);
  if (rc) return rc;
  rc = ensure_twiddles();
  if (rc) return rc;
  ColPassArgs a{};
  a.data = data;
  a.mesh_stride = mesh_stride;
  a.fstride = pitch;
  a.ostride = (int64_t)n[1] * pitch;
  a.ncols = n[2] / 2 + 1;
  a.nouter = n[0];
  a.tw = twiddle_device_ptr();
  if (!a.tw) return set_error(MP_ERR_HIP, "ypass: twiddle table address");
  a.nmesh = nmesh;
  switch (n[1]) {
#define MP_YCASE(L) \
  case L:           \
    return launch_colpass_tuned<L>(s, a);
    MP_YCASE(2) MP_YCASE(4) MP_YCASE(8) MP_YCASE(16) MP_YCASE(32) MP_YCASE(64) MP_YCASE(128) MP_YCASE(256) MP_YCASE(512)
    MP_YCASE(1024) MP_YCASE(2048)
#undef MP_YCASE
    default: return set_error(MP_ERR_UNSUPPORTED, "ypass: length %d", n[1]);
  }
}

int rfft_zy_ylm(hipStream_t s, const float* in, float2* Z, int64_t mesh_stride, const int32_t n[3], int64_t pitch,
                const RealPrologue& pro) {
  int rc = check_shape(n);
  if (rc) return rc;
  rc = ensure_twiddles();
  if (rc) return rc;
  const int M = n[2] / 2;
  if (M > 1024) return set_error(MP_ERR_UNSUPPORTED, "z length %d too large", 2 * M);
  MP_REQUIRE(pro.ell >= 1 && pro.ell <= MAX_ELL, "rfft_zy_ylm: ell %d not in 1..%d", pro.ell, MAX_ELL);
  const int64_t nrows = (int64_t)n[0] * n[1];
  rc = [&]() -> int { MP_POW2_SWITCH(M, launch_zpass_r2c_ylm, in, Z, nrows, pitch, mesh_stride, pro, s); }();
  if (rc) return rc;
  return ypass_meshes(s, Z, 2 * pro.ell + 1, mesh_stride, n, pitch);
}

int ypass(hipStream_t s, float2* data, const int32_t n[3], int64_t pitch) {
  if (!pow2_shape(n)) return ypass_meshes(s, data, 1, 0, n, pitch);
  int rc = check_shape(n);
  if (rc) return rc;
  rc = ensure_twiddles();
  if (rc) return rc;
  const int ncols = n[2] / 2 + 1;
  const int64_t nx = n[0], ny = n[1];
  return col_fft(n[1], false, "fft_ypass", data, data, ncols, nx, pitch, ny * pitch, pitch, ny * pitch, 1.f, s);
}

int ylm_acc_x(hipStream_t s, const float2* partial, float2* acc, const int32_t n[3], int64_t in_pitch, int64_t acc_pitch,
              const KPrologue& kp, int first) {
  const int ncols = n[2] / 2 + 1;
  const int64_t ny = n[1];
  return xpass_ylm_acc(n[0], partial, acc, ncols, n[1], ny * in_pitch, in_pitch, ny * acc_pitch, acc_pitch, kp, first, s);
}

int irfft3d(hipStream_t s, const float2* in, int64_t in_pitch, float* out, const int32_t n[3], float2* scratch,
            int64_t scratch_pitch) {
  if (!pow2_shape(n)) {
    if (!fft_any_shape_ok(n)) return set_error(MP_ERR_UNSUPPORTED, "meshsize (%d, %d, %d) not in 1..2048", n[0], n[1], n[2]);
    const int ncols = n[2] / 2 + 1;
    const int64_t nx = n[0], ny = n[1], sp = scratch_pitch;
    int rc = any_col_fft(s, in, scratch, n[0], ncols, ny, ny * in_pitch, in_pitch, ny * sp, sp, true, 1.f, "fft_xpass_inv_any");
    if (rc) return rc;
    rc = any_col_fft(s, scratch, scratch, n[1], ncols, nx, sp, ny * sp, sp, ny * sp, true, 1.f, "fft_ypass_inv_any");
    if (rc) return rc;
    const float scale = (float)(1.0 / ((double)n[0] * (double)n[1] * (double)n[2]));
    return any_zpass_c2r(s, scratch, sp, out, n[2], nx * ny, scale);
  }
  int rc = check_shape(n);
  if (rc) return rc;
  rc = ensure_twiddles();
  if (rc) return rc;
  const int M = n[2] / 2, ncols = M + 1;
  const int64_t nx = n[0], ny = n[1];
  rc = col_fft(n[0], true, "fft_xpass_inv", in, scratch, ncols, ny, ny * in_pitch, in_pitch, ny * scratch_pitch, scratch_pitch,
               1.f, s);
  if (rc) return rc;
  rc = col_fft(n[1], true, "fft_ypass_inv", scratch, scratch, ncols, nx, scratch_pitch, ny * scratch_pitch, scratch_pitch,
               ny * scratch_pitch, 1.f, s);
  if (rc) return rc;
  // the z pass returns M * (x_even + i x_odd) for M = Nz/2 (unnormalised M-point inverse)
  const float scale = (float)(1.0 / ((double)n[0] * (double)n[1] * (double)M));
  return zpass_c2r(M, scratch, scratch_pitch, out, nx * ny, scale, s);
}

int convolve_shell_zy(hipStream_t s, const ShellGen& g, float2* data, int64_t pitch, const float* Q, const int32_t n[3], float scale) {
  int rc = check_shape(n);
  if (rc) return rc;
  rc = ensure_twiddles();
  if (rc) return rc;
  const int M = n[2] / 2, ncols = M + 1;
  if (M > 1024) return set_error(MP_ERR_UNSUPPORTED, "z length %d too large", 2 * M);
  if (pitch < ncols) return set_error(MP_ERR_INVALID, "pitch %lld < Nz/2+1", (long long)pitch);
  const int64_t nx = n[0], ny = n[1];
  rc = [&]() -> int { MP_POW2_SWITCH(n[0], launch_xpass_inv_shell, g, data, ncols, (int)ny, ny * pitch, pitch, s); }();
  if (rc) return rc;
  rc = col_fft(n[1], true, "window_ypass_inv", data, data, ncols, nx, pitch, ny * pitch, pitch, ny * pitch, 1.f, s);
  if (rc) return rc;
  const float zscale = (float)((double)scale / ((double)n[0] * (double)n[1] * (double)M));
  rc = [&]() -> int { MP_POW2_SWITCH(M, launch_zpass_mul_real, data, pitch, Q, nx * ny, zscale, s); }();
  if (rc) return rc;
  return col_fft(n[1], false, "window_ypass", data, data, ncols, nx, pitch, ny * pitch, pitch, ny * pitch, 1.f, s);
}

int irfft3d_shell(hipStream_t s, const ShellGen& g, float* out, const int32_t n[3], float2* scratch, int64_t sp) {
  int rc = check_shape(n);
  if (rc) return rc;
  rc = ensure_twiddles();
  if (rc) return rc;
  const int M = n[2] / 2, ncols = M + 1;
  if (M > 1024) return set_error(MP_ERR_UNSUPPORTED, "z length %d too large", 2 * M);
  if (sp < ncols) return set_error(MP_ERR_INVALID, "pitch %lld < Nz/2+1", (long long)sp);
  const int64_t nx = n[0], ny = n[1];
  rc = [&]() -> int { MP_POW2_SWITCH(n[0], launch_xpass_inv_shell, g, scratch, ncols, (int)ny, ny * sp, sp, s); }();
  if (rc) return rc;
  rc = col_fft(n[1], true, "window_ypass_inv", scratch, scratch, ncols, nx, sp, ny * sp, sp, ny * sp, 1.f, s);
  if (rc) return rc;
  const float scale = (float)(1.0 / ((double)n[0] * (double)n[1] * (double)M));
  return zpass_c2r(M, scratch, sp, out, nx * ny, scale, s);
}

int convolve_real(hipStream_t s, float2* data, int64_t pitch, const float* Q, const int32_t n[3], float scale) {
  int rc = check_shape(n);
  if (rc) return rc;
  rc = ensure_twiddles();
  if (rc) return rc;
  const int M = n[2] / 2, ncols = M + 1;
  if (M > 1024) return set_error(MP_ERR_UNSUPPORTED, "z length %d too large", 2 * M);
  if (pitch < ncols) return set_error(MP_ERR_INVALID, "pitch %lld < Nz/2+1", (long long)pitch);
  const int64_t nx = n[0], ny = n[1];
  rc = col_fft(n[0], true, "window_xpass_inv", data, data, ncols, ny, ny * pitch, pitch, ny * pitch, pitch, 1.f, s);
  if (rc) return rc;
  rc = col_fft(n[1], true, "window_ypass_inv", data, data, ncols, nx, pitch, ny * pitch, pitch, ny * pitch, 1.f, s);
  if (rc) return rc;
  // irfftn normalisation 1/(N0 N1 Nz) with the z pass returning M (x_even + i x_odd), M = Nz/2
  const float zscale = (float)((double)scale / ((double)n[0] * (double)n[1] * (double)M));
  rc = [&]() -> int { MP_POW2_SWITCH(M, launch_zpass_mul_real, data, pitch, Q, nx * ny, zscale, s); }();
  if (rc) return rc;
  rc = col_fft(n[1], false, "window_ypass", data, data, ncols, nx, pitch, ny * pitch, pitch, ny * pitch, 1.f, s);
  if (rc) return rc;
  return col_fft(n[0], false, "window_xpass", data, data, ncols, ny, ny * pitch, pitch, ny * pitch, pitch, 1.f, s);
}

template <int M, int NE>
static int launch_c2r_bin_ne(const float2* in, int64_t pitch, int64_t nrows, float scale, const RealBinArgs& a, double* partial,
                             int blocks, hipStream_t s) {
  using Cfg = RowCfg<M>;
  const size_t lds = Cfg::LDS + ((size_t)NE * (a.nedges - 1) + a.nedges + 2 * M) * sizeof(double);
  auto kern = zpass_c2r_bin_kernel<M, NE>;
  if (lds > 65536) MP_CHECK_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  MP_LAUNCH("fft_zpass_c2r_bin", kern, dim3((unsigned)blocks), dim3(Cfg::NT), lds, s, in, pitch, nrows, scale, a, partial);
  return MP_OK;
}

static int c2r_bin_blocks(int M, int64_t nrows) {
  const int C = M <= 256 ? 16 : 4096 / M;
  const int64_t g = ceil_div(nrows, C);
  return (int)(g < CB_MAX_BLOCKS ? (g < 1 ? 1 : g) : CB_MAX_BLOCKS);
}

template <int M>
static int launch_c2r_bin(const float2* in, int64_t pitch, int64_t nrows, float scale, const RealBinArgs& a, double* partial,
                          int blocks, hipStream_t s) {
  switch (a.nell) {
    case 1: return launch_c2r_bin_ne<M, 1>(in, pitch, nrows, scale, a, partial, blocks, s);
    case 2: return launch_c2r_bin_ne<M, 2>(in, pitch, nrows, scale, a, partial, blocks, s);
    case 3: return launch_c2r_bin_ne<M, 3>(in, pitch, nrows, scale, a, partial, blocks, s);
    case 4: return launch_c2r_bin_ne<M, 4>(in, pitch, nrows, scale, a, partial, blocks, s);
    case 5: return launch_c2r_bin_ne<M, 5>(in, pitch, nrows, scale, a, partial, blocks, s);
    default: return launch_c2r_bin_ne<M, 6>(in, pitch, nrows, scale, a, partial, blocks, s);
  }
}

size_t irfft3d_bin_real_workspace_bytes(const int32_t n[3], int nell, int nedges) {
  const size_t H = (size_t)n[0] * n[1] * spectrum_pitch(n[2]) * sizeof(float2);
  const int blocks = c2r_bin_blocks(n[2] / 2, (int64_t)n[0] * n[1]);
  return align_up(H, 256) + align_up((size_t)blocks * (nell < 1 ? 1 : nell) * (nedges > 1 ? nedges - 1 : 1) * sizeof(double), 256) +
         512;
}

// c2r (x and y inverse passes, then the z c2r pass fused with the separation binning of
// the real mesh): out[e * nbins + b] = sum over cells of bin b of irfftn(in) P_ell[e](mu_s).
int irfft3d_bin_real(hipStream_t s, const float2* in, int64_t in_pitch, const int32_t n[3], const RealBinArgs& a, double* out,
                     void* ws, size_t ws_bytes, int x_done) {
  int rc = check_shape(n);
  if (rc) return rc;
  rc = ensure_twiddles();
  if (rc) return rc;
  const int M = n[2] / 2, ncols = M + 1;
  if (M > 1024) return set_error(MP_ERR_UNSUPPORTED, "z length %d too large", 2 * M);
  MP_REQUIRE(in_pitch >= ncols, "irfft3d_bin_real: in_pitch too small");
  const int64_t nx = n[0], ny = n[1];
  const int64_t sp = spectrum_pitch(n[2]);
  const int blocks = c2r_bin_blocks(M, nx * ny);
  const int hsize = a.nell * (a.nedges - 1);
  Workspace w(ws, ws_bytes);
  float2* scratch = w.take<float2>((size_t)nx * ny * sp);
  double* partial = w.take<double>((size_t)blocks * hsize);
  if (!scratch || !partial)
    return set_error(MP_ERR_WORKSPACE, "irfft3d_bin_real: workspace of %zu bytes < %zu", ws_bytes,
                     irfft3d_bin_real_workspace_bytes(n, a.nell, a.nedges));
  if (x_done) {  // the input already carries its inverse FFT along x
    rc = col_fft(n[1], true, "fft_ypass_inv", in, scratch, ncols, nx, in_pitch, ny * in_pitch, sp, ny * sp, 1.f, s);
  } else {
    rc = col_fft(n[0], true, "fft_xpass_inv", in, scratch, ncols, ny, ny * in_pitch, in_pitch, ny * sp, sp, 1.f, s);
    if (rc) return rc;
    rc = col_fft(n[1], true, "fft_ypass_inv", scratch, scratch, ncols, nx, sp, ny * sp, sp, ny * sp, 1.f, s);
  }
  if (rc) return rc;
  const float scale = (float)(1.0 / ((double)n[0] * (double)n[1] * (double)M));
  rc = [&]() -> int { MP_POW2_SWITCH(M, launch_c2r_bin, scratch, sp, nx * ny, scale, a, partial, blocks, s); }();
  if (rc) return rc;
  return reduce_partials(s, partial, blocks, hsize, out);
}

int64_t spectrum_pitch(int nz) { return (int64_t)align_up((size_t)(nz / 2 + 1), 16); }

const float2* twiddle_device_ptr() {
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(g_twiddle)) != hipSuccess) return nullptr;
  return (const float2*)p;
}

}  // namespace mp

using namespace mp;

// ------------------------------------------------------------------ C ABI

MP_EXPORT int64_t mp_spectrum_pitch(int32_t nz) { return spectrum_pitch(nz); }

MP_EXPORT int mp_rfft3d(void* stream, const float* in, void* out, const int32_t meshsize[3], int64_t out_pitch) {
  MP_REQUIRE(in && out && meshsize, "mp_rfft3d: null pointer");
  RealPrologue pro{};
  return rfft3d((hipStream_t)stream, in, (float2*)out, meshsize, out_pitch, pro);
}

MP_EXPORT int64_t mp_irfft3d_workspace_bytes(const int32_t meshsize[3]) {
  return (int64_t)meshsize[0] * meshsize[1] * spectrum_pitch(meshsize[2]) * (int64_t)sizeof(float2) + 256;
}

MP_EXPORT int mp_irfft3d(void* stream, const void* in, int64_t in_pitch, float* out, const int32_t meshsize[3],
                         void* workspace, size_t workspace_bytes) {
  MP_REQUIRE(in && out && meshsize, "mp_irfft3d: null pointer");
  MP_REQUIRE(in_pitch >= meshsize[2] / 2 + 1, "mp_irfft3d: in_pitch too small");
  Workspace ws(workspace, workspace_bytes);
  int64_t sp = spectrum_pitch(meshsize[2]);
  float2* scratch = ws.take<float2>((size_t)meshsize[0] * meshsize[1] * sp);
  if (!scratch) return set_error(MP_ERR_WORKSPACE, "mp_irfft3d: workspace too small");
  return irfft3d((hipStream_t)stream, (const float2*)in, in_pitch, out, meshsize, scratch, sp);
}
