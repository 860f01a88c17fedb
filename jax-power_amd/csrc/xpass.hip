// The x pass of an interlaced paint, with the interlacing combine folded in:
//   A(kx, ky, kz) = comp(k) / n_i * sum_j exp(+i s_j (kx+ky+kz)_circ) FFT_x(partial_j)(kx, ky, kz)
// (jaxpower/mesh.py:2739-2751, compensation resamplers.py:16-26).  Two consumers:
//   * write: A is stored (the complex mesh of _paint(..., out='complex'));
//   * bin:   A is never stored; |A|^2 w L_ell(mu) is summed into |k| shells in FP64, which is
//            compute_mesh2_spectrum with a global line of sight on an auto spectrum
//            (jaxpower/mesh2.py:346-369, BinMesh2SpectrumPoles.__call__ mesh2.py:100-119).
//
// Work decomposition: persistent blocks walk units (tile, shift j); a tile is C contiguous kz
// columns of one ky row, for all L values of kx.  The global loads of the next unit are issued
// before the FFT of the current one, so HBM latency overlaps the LDS work.  After the FFT each
// thread owns one column and EPT consecutive kx of the tile for all shifts, so the combine
// stays in registers, and in bin mode |k| is monotone along the thread's kx run: the FP64
// shell sums stay in registers until the shell changes, then go to an LDS histogram that each
// block writes once at the end.
#include "bin.h"
#include "common.h"
#include "fft.h"
#include "fft_core.h"
#include "sph.h"
#include "tune.h"

#include <type_traits>

#include <algorithm>
#include <mutex>
#include <vector>

namespace mp {

constexpr int XB_MAX_PROD = 6;

static int multiprocessors();

struct XPassArgs {
  const float2* in;
  int64_t mesh_stride;  // between the partial spectra of consecutive shifts
  int64_t fstride;      // between consecutive kx rows of the input
  int64_t ostride;      // between consecutive ky rows of the input
  int ni, ncols, ny, nchunks, ntiles;
  CombineTables tab;
  float scale;  // 1 / n_i
  int zero_dc;  // the (0,0,0) mode of A is set to 0
  int power;    // write mode: 1 stores |A|^2 (imaginary part 0) instead of A, 2 its inverse FFT along x
  int linear;   // bin mode: bins Re(A) (a linear functional of the mesh) instead of |A|^2
  int ylm;      // bin mode: Y_lm(khat) weights instead of P_ell(mu)
  int ylm_m[XB_MAX_PROD];  // bin mode with YLM: the weight of product p is Y_{ell[p], ylm_m[p]}(khat)
  const float2* tw;
  // write mode
  float2* out;
  int64_t out_fstride, out_ostride;
  // bin mode
  BinGeom g;
  int ell[XB_MAX_PROD], rz[XB_MAX_PROD];
  LegendreTable leg;
  float los[3];
  int need_mu;
  double* partial;      // [gridDim.y][gridDim.x][nprod][nbins][2]
  int64_t mock_stride;  // bin mode: between the partial spectra of the catalogues (blockIdx.y) of a batch
};
static_assert(XB_MAX_PROD <= LEG_MAX, "Legendre table too short");

// Bytes of LDS the kernel needs beyond its tile: the per-kx and per-kz tables (TAB) and, in
// bin mode, the thresholds and the histogram.
__host__ __device__ inline size_t xpass_extra_lds(int L, int ncols, int ni, int nedges, int nprod, bool bin, bool tab) {
  auto r8 = [](size_t b) { return (b + 7) & ~(size_t)7; };
  size_t bytes = 0;
  if (tab) {
    bytes += r8((size_t)ni * L * 8) + r8((size_t)ni * ncols * 8) + r8((size_t)L * 4) + r8((size_t)ncols * 4);
    if (bin) bytes += (size_t)L * 8 + (size_t)ncols * 8;
  }
  if (bin) bytes += (size_t)(nedges + 4) * 8 + (size_t)nprod * (nedges - 1) * 8;  // thresholds + 2 sentinels a side
  return bytes;
}

// Half-size tiles from L = 512 up (32 KiB): two blocks share a CU, which hides more of the HBM
// latency than one block with twice the columns (measured on the 512^3 x pass).
template <int L>
using XPCfg = XCfgC<L, L < 512 ? (L >= 4 ? 16 : 64 / L) : 4096 / L>;
// Register budget of the local line-of-sight binning x pass: at least 4 waves per SIMD
// (<= 128 VGPRs), so that two 8-wave blocks share a CU; unconstrained it took 160 VGPRs at
// L = 1024 and ran 1.3x slower (measured on the 1024^3 FKP spectrum).  The combine kernel
// keeps its registers: capped it spills and slows down.
constexpr int XP_WAVES_PER_SIMD = 4;

template <int L, bool BIN, int NP, bool TAB, bool YLM = false, int CC = XPCfg<L>::C, bool LEAN = false>
__global__ void __launch_bounds__((XCfgC<L, CC>::NT))
xpass_kernel(XPassArgs a) {
  using Cfg = XCfgC<L, CC>;
  constexpr int C = Cfg::C, NT = Cfg::NT;
  constexpr int EPT = Cfg::EPT;
  static_assert(EPT >= 1 && EPT * NT == C * L && (L / 2) % EPT == 0, "tile must split into monotone runs");
  extern __shared__ float2 lds[];
  __shared__ float2 stw[L];
  const int tid = threadIdx.x;
  load_twiddles<L, NT>(stw, a.tw);
  const int c = tid % C;           // this thread's column in the tile
  const int l0 = (tid / C) * EPT;  // and its first kx
  const int nedges = a.g.nedges, nbins = nedges - 1;
  const bool phases = a.tab.phx != nullptr, comp = a.tab.cx != nullptr;
  // tables indexed by kx or kz: copied to LDS (TAB) or read where they are
  const float2* phx = a.tab.phx;
  const float2* phz = a.tab.phz;
  const float* cx = a.tab.cx;
  const float* cz = a.tab.cz;
  const double* kxt = a.g.kx;
  const double* kzt = a.g.kz;
  int phx_stride = a.tab.stride, phz_stride = a.tab.stride;
  char* cur = reinterpret_cast<char*>(lds + C * L);
  auto take = [&](size_t bytes) {
    char* p = cur;
    cur += (bytes + 7) & ~(size_t)7;
    return p;
  };
  if constexpr (TAB) {
    float2* sphx = reinterpret_cast<float2*>(take((size_t)a.ni * L * 8));
    float2* sphz = reinterpret_cast<float2*>(take((size_t)a.ni * a.ncols * 8));
    float* scx = reinterpret_cast<float*>(take((size_t)L * 4));
    float* scz = reinterpret_cast<float*>(take((size_t)a.ncols * 4));
    if (phases) {
      for (int i = tid; i < a.ni * L; i += NT) sphx[i] = a.tab.phx[(i / L) * a.tab.stride + i % L];
      for (int i = tid; i < a.ni * a.ncols; i += NT) sphz[i] = a.tab.phz[(i / a.ncols) * a.tab.stride + i % a.ncols];
    }
    // ones without compensation, so that kernels reading the tables here need no test of comp
    for (int i = tid; i < L; i += NT) scx[i] = comp ? a.tab.cx[i] : 1.f;
    for (int i = tid; i < a.ncols; i += NT) scz[i] = comp ? a.tab.cz[i] : 1.f;
    phx = sphx;
    phz = sphz;
    cx = scx;
    cz = scz;
    phx_stride = L;
    phz_stride = a.ncols;
    if constexpr (BIN) {
      double* skx = reinterpret_cast<double*>(take((size_t)L * 8));
      double* skz = reinterpret_cast<double*>(take((size_t)a.ncols * 8));
      for (int i = tid; i < L; i += NT) skx[i] = a.g.kx[i];
      for (int i = tid; i < a.ncols; i += NT) skz[i] = a.g.kz[i];
      kxt = skx;
      kzt = skz;
    }
  }
  // thresholds with two sentinels on each side: thr[-2], thr[-1] = -inf, thr[nedges], thr[nedges + 1] = +inf
  double* thr = reinterpret_cast<double*>(take(BIN ? (size_t)(nedges + 4) * 8 : 0)) + 2;
  double* hist = reinterpret_cast<double*>(take(BIN ? (size_t)NP * nbins * 8 : 0));
  __shared__ int swsteps;
  if constexpr (BIN) {
    for (int i = tid; i < nedges; i += NT) thr[i] = a.g.thr[i];
    if (tid < 4) thr[tid < 2 ? tid - 2 : nedges + tid - 2] = tid < 2 ? -INFINITY : INFINITY;
    for (int i = tid; i < NP * nbins; i += NT) hist[i] = 0.;
    // Thresholds one kx step can cross: |k| changes by at most the kx spacing between
    // neighbours of a run, so ceil-like (spacing / narrowest bin) + 1, with a 1 % margin; 0 (the
    // general walk) for narrower bins or empty ones.
    if (tid < 64) {
      double wmin = 1e300;
      for (int i = tid; i + 1 < nedges; i += 64) wmin = fmin(wmin, sqrt(a.g.thr[i + 1]) - sqrt(a.g.thr[i]));
#pragma unroll
      for (int d = 32; d > 0; d >>= 1) wmin = fmin(wmin, __shfl_xor(wmin, d, 64));
      if (tid == 0) {
        const double dk = L > 1 ? fabs(a.g.kx[1] - a.g.kx[0]) : 0.;
        const double r = wmin > 0. ? 1.01 * dk / wmin : 1e300;
        swsteps = r < 4. ? (int)r + 1 : 0;
      }
    }
  }
  float2 reg[EPT];
  const float2* in = a.in + blockIdx.y * a.mock_stride;  // this block's catalogue
  auto load = [&](int tile, int j) {
    const int c0 = (tile % a.nchunks) * C, y = tile / a.nchunks;
    const float2* src = in + j * a.mesh_stride + (int64_t)y * a.ostride + c0;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int idx = tid + e * NT, cc = idx % C, l = idx / C;
      reg[e] = (c0 + cc < a.ncols) ? src[l * a.fstride + cc] : make_float2(0.f, 0.f);
    }
  };
  float2 acc[EPT];
#pragma unroll
  for (int e = 0; e < EPT; ++e) acc[e] = make_float2(0.f, 0.f);
  int tile = xcd_first_tile(), j = 0;
  if (tile < a.ntiles) load(tile, 0);
  while (tile < a.ntiles) {
    if constexpr (RegFFT<L, C, NT>::ok) {
      fft_first_from_regs<L, C, NT, false>(reg, lds);  // the first stage on registers
    } else {
#pragma unroll
      for (int e = 0; e < EPT; ++e) lds[tid + e * NT] = reg[e];  // element (l, cc) at l * C + cc
    }
    __syncthreads();
    int ntile = tile, nj = j + 1;
    if (nj == a.ni) {
      nj = 0;
      ntile += gridDim.x;
    }
    if (ntile < a.ntiles) load(ntile, nj);
    if constexpr (RegFFT<L, C, NT>::ok) {
      fft_stages_until<L, C, C, 1, true, NT, false, RegFFT<L, C, NT>::R1, L, L>(lds, stw);
    } else {
      fft_lds<L, C, C, 1, true, NT, false, L>(lds, stw);
    }
    const int c0 = (tile % a.nchunks) * C, y = tile / a.nchunks;
    const int kz = min(c0 + c, a.ncols - 1);
    if (phases && j > 0) {  // shift 0: every phase is exactly 1
      const float2 pyz = cmul(a.tab.phy[j * a.tab.stride + y], phz[j * phz_stride + kz]);
      const float2* px = phx + j * phx_stride + l0;
#pragma unroll
      for (int e = 0; e < EPT; ++e) acc[e] = cadd(acc[e], cmul(lds[(l0 + e) * C + c], cmul(px[e], pyz)));
    } else {
#pragma unroll
      for (int e = 0; e < EPT; ++e) acc[e] = cadd(acc[e], lds[(l0 + e) * C + c]);
    }
    if (j == a.ni - 1) {
      bool done = false;
      if constexpr (!BIN) {
        if (a.power == 2) {  // |A|^2, then its inverse FFT along x in the same tile
          const bool col_ok = c0 + c < a.ncols;
          float s = a.scale;
          if (comp) s *= a.tab.cy[y] * cz[kz];
          __syncthreads();  // the combine above read the tile
#pragma unroll
          for (int e = 0; e < EPT; ++e) {
            const int l = l0 + e;
            float2 v = cscale(acc[e], comp ? s * cx[l] : s);
            if (a.zero_dc && l == 0 && y == 0 && kz == 0) v = make_float2(0.f, 0.f);
            lds[l * C + c] = make_float2(col_ok ? v.x * v.x + v.y * v.y : 0.f, 0.f);
          }
          __syncthreads();
          fft_lds<L, C, C, 1, true, NT, true, L>(lds, stw);
          if (col_ok) {
            float2* dst = a.out + (int64_t)y * a.out_ostride + kz;
#pragma unroll
            for (int e = 0; e < EPT; ++e) dst[(l0 + e) * a.out_fstride] = lds[(l0 + e) * C + c];
          }
          done = true;
        }
      }
      if (!done && c0 + c < a.ncols) {
        float s = a.scale;
        if (comp) s *= a.tab.cy[y] * cz[kz];
        if constexpr (!BIN) {
          float2* dst = a.out + (int64_t)y * a.out_ostride + kz;
#pragma unroll
          for (int e = 0; e < EPT; ++e) {
            const int l = l0 + e;
            float2 v = cscale(acc[e], comp ? s * cx[l] : s);
            if (a.zero_dc && l == 0 && y == 0 && kz == 0) v = make_float2(0.f, 0.f);
            if (a.power) v = make_float2(v.x * v.x + v.y * v.y, 0.f);
            dst[l * a.out_fstride] = v;
          }
        } else {
          const double kyv = a.g.ky[y], kzv = kzt[kz];
          const float w = kz > 0 ? 2.f : 1.f;
          const float fky = (float)kyv, fkz = (float)kzv;
          const float lyz = fky * a.los[1] + fkz * a.los[2];
          const float f2yz = fky * fky + fkz * fkz;
          double sum[NP];
#pragma unroll
          for (int p = 0; p < NP; ++p) sum[p] = 0.;
          // Bin position b (thresholds <= k2) along the thread's run of kx, on which k2 is
          // monotone: found once by bisection, then moved by at most `steps` thresholds per kx
          // step in the run's direction (a fixed number of reads, no data-dependent loop); runs
          // through bins narrower than the kx spacing take the general walk.
          const int steps = __builtin_amdgcn_readfirstlane(swsteps);  // uniform
          const int dir = l0 < L / 2 ? 1 : -1;  // k2 increases along the run below L/2
          if constexpr (LEAN && TAB && !YLM) {
            // |A|^2 w L_ell(mu) with even ell <= 4 only (chosen on the host): the general loop
            // below without its other cases.  With FS > 0 the bin moves by exactly FS conditional
            // threshold reads per kx step (steps beyond the needed ones do not move it), so a
            // run is one basic block around the flush branch; FS = 0 takes the general walk.
            const bool orow = (l0 == 0) & (y == 0) & (kz == 0);  // the origin is a run's first mode
            const int po = dir > 0 ? 0 : -2, pstep = dir > 0 ? 2 : 0;
            auto run = [&](auto fixed) {
              constexpr int FS = decltype(fixed)::value;
              int b = first_bin(thr, nedges, k2_sum(k2_xy(kxt[l0], kyv), kzv));
#pragma unroll
              for (int e = 0; e < EPT; ++e) {
                const int l = l0 + e;
                const double kxv = kxt[l];
                if (e > 0) {
                  const double k2 = k2_sum(k2_xy(kxv, kyv), kzv);
                  int nb = b;
                  if constexpr (FS > 0) {
                    // up to two thresholds crossed: the pair beyond b in the run's direction,
                    // one ds_read2 and no branch (the sentinels keep b in [0, nedges])
                    const double t0 = thr[nb + po], t1 = thr[nb + po + 1];
                    nb += pstep - ((t0 > k2 ? 1 : 0) + (t1 > k2 ? 1 : 0));
                  } else if (steps > 0) {
                    for (int t = 0; t < steps; ++t) {
                      if (dir > 0) nb += (nb < nedges && thr[min(nb, nedges - 1)] <= k2) ? 1 : 0;
                      else nb -= (nb > 0 && thr[max(nb - 1, 0)] > k2) ? 1 : 0;
                    }
                  } else {
                    nb = walk_bin(thr, nedges, b, k2);
                  }
                  if (nb != b) {
                    if (b >= 1 && b < nedges) {
#pragma unroll
                      for (int p = 0; p < NP; ++p) atomicAdd(&hist[p * nbins + b - 1], sum[p]);
                    }
#pragma unroll
                    for (int p = 0; p < NP; ++p) sum[p] = 0.;
                    b = nb;
                  }
                }
                const bool valid = b >= 1 && b < nedges;
                const bool origin = e == 0 && orow;
                float2 A = cscale(acc[e], s * cx[l]);  // cx: an LDS table, ones without compensation
                if (origin && a.zero_dc) A = make_float2(0.f, 0.f);
                // mu is finite (los is), and with ell = 0 only c1 = c2 = 0: no need_mu test
                const float fkx = (float)kxv;
                const float kn2 = fkx * fkx + f2yz;
                const float mu = kn2 > 0.f ? (fkx * a.los[0] + lyz) * __builtin_amdgcn_rsqf(kn2) : 0.f;
                const float m2 = mu * mu;
                const float wp2 = valid ? w * (A.x * A.x + A.y * A.y) : 0.f;  // w = 1 or 2: exact
#pragma unroll
                for (int p = 0; p < NP; ++p) {
                  const float q = fmaf(fmaf(a.leg.c2[p], m2, a.leg.c1[p]), m2, a.leg.c0[p]);
                  sum[p] += (double)(origin && a.rz[p] ? 0.f : wp2 * q);
                }
              }
              if (b >= 1 && b < nedges) {
#pragma unroll
                for (int p = 0; p < NP; ++p) atomicAdd(&hist[p * nbins + b - 1], sum[p]);
              }
            };
            if (steps == 1 || steps == 2) run(std::integral_constant<int, 2>{});
            else run(std::integral_constant<int, 0>{});
          } else {
          int b = -1;
#pragma unroll
          for (int e = 0; e < EPT; ++e) {
            const int l = l0 + e;
            const double kxv = kxt[l];
            const double k2 = k2_sum(k2_xy(kxv, kyv), kzv);
            int nb;
            if (e == 0) {
              nb = first_bin(thr, nedges, k2);
            } else if (steps > 0) {
              nb = b;
              for (int t = 0; t < steps; ++t) {
                if (dir > 0) nb += (nb < nedges && thr[min(nb, nedges - 1)] <= k2) ? 1 : 0;
                else nb -= (nb > 0 && thr[max(nb - 1, 0)] > k2) ? 1 : 0;
              }
            } else {
              nb = walk_bin(thr, nedges, b, k2);
            }
            if (nb != b) {
              if (b >= 1 && b < nedges) {
#pragma unroll
                for (int p = 0; p < NP; ++p) atomicAdd(&hist[p * nbins + b - 1], sum[p]);
              }
#pragma unroll
              for (int p = 0; p < NP; ++p) sum[p] = 0.;
              b = nb;
            }
            if (b < 1 || b >= nedges) continue;
            float2 A = cscale(acc[e], comp ? s * cx[l] : s);
            const bool origin = (l == 0) & (y == 0) & (kz == 0);
            if (origin && a.zero_dc) A = make_float2(0.f, 0.f);
            const float p2 = a.linear ? A.x : A.x * A.x + A.y * A.y;
            if constexpr (YLM) {  // real spherical harmonics of khat (local line of sight)
#pragma unroll
              for (int p = 0; p < NP; ++p) {
                if (origin && a.rz[p]) continue;
                const float f = w * ylm_of_vector_rsq(a.ell[p], a.ylm_m[p], (float)kxv, fky, fkz);
                sum[p] += (double)(p2 * f);
              }
            } else {
              float mu = 0.f;
              if (a.need_mu) {
                const float fkx = (float)kxv;
                const float kn2 = fkx * fkx + f2yz;
                mu = kn2 > 0.f ? (fkx * a.los[0] + lyz) * rsqrtf(kn2) : 0.f;
              }
              const float m2 = mu * mu;
#pragma unroll
              for (int p = 0; p < NP; ++p) {
                if (origin && a.rz[p]) continue;
                const float f = w * legendre_eval(a.leg, p, mu, m2);
                sum[p] += (double)(p2 * f);
              }
            }
          }
          if (b >= 1 && b < nedges) {
#pragma unroll
            for (int p = 0; p < NP; ++p)
              if (sum[p] != 0.) atomicAdd(&hist[p * nbins + b - 1], sum[p]);
          }
          }
        }
      }
#pragma unroll
      for (int e = 0; e < EPT; ++e) acc[e] = make_float2(0.f, 0.f);
    }
    __syncthreads();
    tile = ntile;
    j = nj;
  }
  if constexpr (BIN) {
    __syncthreads();
    double* dst = a.partial + ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * NP * nbins * 2;
    for (int i = tid; i < NP * nbins; i += NT) {
      dst[2 * i] = hist[i];
      dst[2 * i + 1] = 0.;
    }
  }
}

// Local line of sight, one ell: C(k) = sum_m FFT_x(Z_m)(k) Y_lm(khat) over the 2 ell + 1 partial
// spectra Z_m (rho Y_lm(xhat) after their z and y passes), never written: each tile's C is
// multiplied by the combined mesh A0 at the same modes and binned in FP64 as
// conj(C) A0 w (remove_zero at k = 0), and optionally |A0|^2 w (the ell = 0 term of an auto
// spectrum).  jaxpower/mesh2.py:304-333.
struct YlmBinArgs {
  const float2* Z;
  int64_t mesh_stride, fstride, ostride;  // of Z (and of A0 with a0_fstride, a0_ostride)
  const float2* A0;
  int64_t a0_fstride, a0_ostride;
  int nm, ell, ncols, ny, nchunks, ntiles;
  KPrologue kp;
  const float2* tw;
  BinGeom g;
  int with0;
  double* partial;  // [gridDim.x][NPP][nbins][2]
};

template <int L, int NPP>
__global__ void __launch_bounds__(XPCfg<L>::NT, XP_WAVES_PER_SIMD)
xpass_ylm_bin_kernel(YlmBinArgs a) {
  using Cfg = XPCfg<L>;
  constexpr int C = Cfg::C, NT = Cfg::NT;
  constexpr int EPT = Cfg::EPT;
  static_assert(EPT >= 1 && EPT * NT == C * L && (L / 2) % EPT == 0, "tile must split into monotone runs");
  extern __shared__ float2 lds[];
  __shared__ float2 stw[L];
  const int tid = threadIdx.x;
  load_twiddles<L, NT>(stw, a.tw);
  const int c = tid % C;
  const int l0 = (tid / C) * EPT;
  const int nedges = a.g.nedges, nbins = nedges - 1;
  double* thr = reinterpret_cast<double*>(lds + C * L);
  double* hist = thr + nedges;
  for (int i = tid; i < nedges; i += NT) thr[i] = a.g.thr[i];
  for (int i = tid; i < NPP * nbins * 2; i += NT) hist[i] = 0.;
  float2 reg[EPT];
  auto load = [&](int tile, int m) {
    const int c0 = (tile % a.nchunks) * C, y = tile / a.nchunks;
    const float2* src = a.Z + m * a.mesh_stride + (int64_t)y * a.ostride + c0;
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
      const int idx = tid + e * NT, cc = idx % C, l = idx / C;
      reg[e] = (c0 + cc < a.ncols) ? src[l * a.fstride + cc] : make_float2(0.f, 0.f);
    }
  };
  float2 acc[EPT];
#pragma unroll
  for (int e = 0; e < EPT; ++e) acc[e] = make_float2(0.f, 0.f);
  int tile = xcd_first_tile(), m = 0;
  if (tile < a.ntiles) load(tile, 0);
  while (tile < a.ntiles) {
    if constexpr (RegFFT<L, C, NT>::ok) {
      fft_first_from_regs<L, C, NT, false>(reg, lds);
    } else {
#pragma unroll
      for (int e = 0; e < EPT; ++e) lds[tid + e * NT] = reg[e];
    }
    __syncthreads();
    int ntile = tile, nm = m + 1;
    if (nm == a.nm) {
      nm = 0;
      ntile += gridDim.x;
    }
    if (ntile < a.ntiles) load(ntile, nm);
    if constexpr (RegFFT<L, C, NT>::ok) {
      fft_stages_until<L, C, C, 1, true, NT, false, RegFFT<L, C, NT>::R1, L, L>(lds, stw);
    } else {
      fft_lds<L, C, C, 1, true, NT, false, L>(lds, stw);
    }
    const int c0 = (tile % a.nchunks) * C, y = tile / a.nchunks;
    const int kz = min(c0 + c, a.ncols - 1);
    KPrologue kp = a.kp;
    kp.m = m - a.ell;
#pragma unroll
    for (int e = 0; e < EPT; ++e) acc[e] = cadd(acc[e], cscale(lds[(l0 + e) * C + c], kp.eval(l0 + e, y, kz)));
    if (m == a.nm - 1) {
      if (c0 + c < a.ncols) {
        const double kyv = a.g.ky[y], kzv = a.g.kz[kz];
        const float w = kz > 0 ? 2.f : 1.f;
        double sum[NPP][2];
#pragma unroll
        for (int p = 0; p < NPP; ++p) sum[p][0] = sum[p][1] = 0.;
        int b = -1;
        const float2* a0 = a.A0 + (int64_t)y * a.a0_ostride + kz;
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
          const int l = l0 + e;
          const double k2 = k2_sum(k2_xy(a.g.kx[l], kyv), kzv);
          const int nb = e == 0 ? first_bin(thr, nedges, k2) : walk_bin(thr, nedges, b, k2);
          if (nb != b) {
            if (b >= 1 && b < nedges) {
#pragma unroll
              for (int p = 0; p < NPP; ++p)
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                  if (sum[p][q] != 0.) atomicAdd(&hist[(p * nbins + b - 1) * 2 + q], sum[p][q]);
                  sum[p][q] = 0.;
                }
            }
            b = nb;
          }
          if (b < 1 || b >= nedges) continue;
          const float2 A = a0[l * a.a0_fstride];
          const float2 Cv = acc[e];
          const bool origin = (l == 0) & (y == 0) & (kz == 0);
          if (!origin) {  // remove_zero for the ell > 0 product
            sum[0][0] += (double)((Cv.x * A.x + Cv.y * A.y) * w);
            sum[0][1] += (double)((Cv.x * A.y - Cv.y * A.x) * w);
          }
          if (NPP > 1) sum[NPP - 1][0] += (double)((A.x * A.x + A.y * A.y) * w);
        }
        if (b >= 1 && b < nedges) {
#pragma unroll
          for (int p = 0; p < NPP; ++p)
#pragma unroll
            for (int q = 0; q < 2; ++q)
              if (sum[p][q] != 0.) atomicAdd(&hist[(p * nbins + b - 1) * 2 + q], sum[p][q]);
        }
      }
#pragma unroll
      for (int e = 0; e < EPT; ++e) acc[e] = make_float2(0.f, 0.f);
    }
    __syncthreads();
    tile = ntile;
    m = nm;
  }
  __syncthreads();
  double* dst = a.partial + (int64_t)blockIdx.x * NPP * nbins * 2;
  for (int i = tid; i < NPP * nbins * 2; i += NT) dst[i] = hist[i];
}

template <int L, int NPP>
static int launch_ylm_bin(hipStream_t s, YlmBinArgs& a, double* partial, size_t partial_doubles, int* nblocks_out) {
  using Cfg = XPCfg<L>;
  auto kern = xpass_ylm_bin_kernel<L, NPP>;
  a.nchunks = (int)ceil_div(a.ncols, Cfg::C);
  a.ntiles = a.nchunks * a.ny;
  const int nbins = a.g.nedges - 1;
  const size_t lds = Cfg::LDS + (size_t)a.g.nedges * 8 + (size_t)NPP * nbins * 2 * 8;
  if (lds > 150 * 1024) return set_error(MP_ERR_UNSUPPORTED, "ylm bin: %zu bytes of LDS needed", lds);
  if (lds > 65536) MP_CHECK_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  int per = 0;
  MP_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, Cfg::NT, lds));
  if (per < 1) per = 1;
  int blocks = multiprocessors() * per;
  if (blocks > a.ntiles) blocks = a.ntiles;
  const int cap = (int)(partial_doubles / ((size_t)NPP * nbins * 2));
  if (cap < 1) return set_error(MP_ERR_WORKSPACE, "ylm bin: no room for the partial histograms");
  if (blocks > cap) blocks = cap;
  a.partial = partial;
  *nblocks_out = blocks;
  MP_LAUNCH("fft_xpass_ylm_bin", kern, dim3((unsigned)blocks), dim3(Cfg::NT), lds, s, a);
  return MP_OK;
}

template <int L>
static int ylm_bin_L(hipStream_t s, YlmBinArgs& a, double* partial, size_t pd, int* nblocks) {
  return a.with0 ? launch_ylm_bin<L, 2>(s, a, partial, pd, nblocks) : launch_ylm_bin<L, 1>(s, a, partial, pd, nblocks);
}

// Per-axis interlacing phases exp(i s_j k_circ) and compensation sinc(k/2pi)^-order, in FP64
// then rounded to FP32.  k_circ = 2 pi fftfreq(n) (rfftfreq on the last axis), mesh.py:1605-1636.
__global__ void combine_tables_kernel(int n0, int n1, int n2, int ni, int order, float2* __restrict__ ph, float* __restrict__ comp) {
  const int ncols = n2 / 2 + 1;
  const int len = n0 + n1 + ncols;
  const int total = (ni + 1) * len;
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  int j = t / len, r = t % len;
  int axis = r < n0 ? 0 : (r < n0 + n1 ? 1 : 2);
  int m = axis == 0 ? r : (axis == 1 ? r - n0 : r - n0 - n1);
  int n = axis == 0 ? n0 : (axis == 1 ? n1 : n2);
  int f = (axis == 2 || m < (n + 1) / 2) ? m : m - n;
  double x = (double)f / (double)n;  // fftfreq value
  if (j < ni) {
    double s = (double)j / (double)ni;
    double ang = 2.0 * M_PI * x * s;
    ph[t] = make_float2((float)cos(ang), (float)sin(ang));
  } else {
    double cval = 1.0;
    if (f != 0 && order > 0) {
      double px = M_PI * x;
      cval = pow(sin(px) / px, -(double)order);
    }
    comp[r] = (float)cval;
  }
}

// ------------------------------------------------------------------ host side

static int multiprocessors() {
  static std::mutex mu;
  static std::vector<int> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  std::lock_guard<std::mutex> lock(mu);
  if ((int)cache.size() <= dev) cache.resize(dev + 1, 0);
  if (!cache[dev]) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n < 1) n = 256;
    cache[dev] = n;
  }
  return cache[dev];
}

size_t xpass_tables_bytes(const int32_t n[3], int ni) {
  const size_t len = (size_t)n[0] + n[1] + n[2] / 2 + 1;
  return align_up((size_t)(ni + 1) * len * sizeof(float2), 256) + align_up(len * sizeof(float), 256);
}

// Builds the phase and compensation tables in the workspace.
int combine_tables(hipStream_t s, const int32_t n[3], int order, int ni, int compensate, Workspace& wsp, CombineTables& tab) {
  const int len = n[0] + n[1] + n[2] / 2 + 1;
  float2* ph = wsp.take<float2>((size_t)(ni + 1) * len);
  float* comp = wsp.take<float>((size_t)len);
  if (!ph || !comp) return set_error(MP_ERR_WORKSPACE, "x pass: workspace too small for the tables");
  const int total = (ni + 1) * len;
  tab = CombineTables{};
  tab.stride = len;
  if (ni == 1 && !compensate) return MP_OK;  // no phase and no compensation: nothing to tabulate
  MP_LAUNCH("combine_tables", combine_tables_kernel, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, s, n[0], n[1], n[2], ni,
            compensate ? order : 0, ph, comp);
  if (ni > 1) {
    tab.phx = ph;
    tab.phy = ph + n[0];
    tab.phz = ph + n[0] + n[1];
  }
  if (compensate) {
    tab.cx = comp;
    tab.cy = comp + n[0];
    tab.cz = comp + n[0] + n[1];
  }
  return MP_OK;
}

template <int L, bool BIN, int NP, bool TAB, bool YLM = false, int CC = XPCfg<L>::C, bool LEAN = false>
static int launch_xpass_t(hipStream_t s, XPassArgs& a, int* nblocks_out, double* partial_ws, size_t partial_doubles, int nmock) {
  using Cfg = XCfgC<L, CC>;
  auto kern = xpass_kernel<L, BIN, NP, TAB, YLM, CC, LEAN>;
  a.nchunks = (int)ceil_div(a.ncols, Cfg::C);
  a.ntiles = a.nchunks * a.ny;
  const size_t lds = Cfg::LDS + xpass_extra_lds(L, a.ncols, a.ni, a.g.nedges, NP, BIN, TAB);
  if (lds > 160 * 1024) return set_error(MP_ERR_UNSUPPORTED, "x pass: %zu bytes of LDS needed", lds);
  if (lds > 65536) MP_CHECK_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  int per = 0;
  MP_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, Cfg::NT, lds));
  if (per < 1) per = 1;
  int blocks = multiprocessors() * per;
  if (nmock > 1) blocks = (int)ceil_div(ceil_div(blocks, nmock), 8) * 8;  // per catalogue, a multiple of the 8 XCDs
  if (blocks > a.ntiles) blocks = a.ntiles;
  if (BIN) {
    const size_t per_block = (size_t)NP * (a.g.nedges - 1) * 2;
    const int cap = (int)(partial_doubles / per_block / nmock);
    if (cap < 1) return set_error(MP_ERR_WORKSPACE, "x pass: no room for the partial histograms");
    if (blocks > cap) blocks = cap;
    a.partial = partial_ws;
  }
  *nblocks_out = blocks;
  MP_LAUNCH(BIN ? "fft_xpass_bin" : "fft_xpass_combine", kern, dim3((unsigned)blocks, (unsigned)nmock), dim3(Cfg::NT), lds, s, a);
  return MP_OK;
}

// The kx/kz tables go to LDS when they fit beside the tile.
template <int L, bool BIN, int NP, bool YLM = false, int CC = XPCfg<L>::C, bool LEAN = false>
static int launch_xpass(hipStream_t s, XPassArgs& a, int* nblocks_out, double* partial_ws, size_t partial_doubles, int nmock = 1) {
  const size_t with_tab = XCfgC<L, CC>::LDS + xpass_extra_lds(L, a.ncols, a.ni, a.g.nedges, NP, BIN, true);
  if (with_tab <= 160 * 1024)
    return launch_xpass_t<L, BIN, NP, true, YLM, CC, LEAN>(s, a, nblocks_out, partial_ws, partial_doubles, nmock);
  return launch_xpass_t<L, BIN, NP, false, YLM, CC>(s, a, nblocks_out, partial_ws, partial_doubles, nmock);
}

// Tile widths the autotuner tries for the binning x pass (candidate 0 is the default): half the
// default while at least 4 columns remain, and twice it up to 64 KiB tiles; 0 marks a candidate
// that does not apply.  Lengths below 64 keep the default.
template <int L>
struct XPTune {
  static constexpr int C0 = XPCfg<L>::C;
  static constexpr int C1 = L >= 64 && C0 >= 8 ? C0 / 2 : 0;
  static constexpr int C2 = L >= 64 && 2 * C0 * L * 8 <= 65536 ? 2 * C0 : 0;
};

// The binning x pass with a global line of sight (P_ell weights) and at most 3 products, its
// tile width chosen by the autotuner per (L, Ny, columns, n_i, products, catalogues).  Timing
// launches only rewrite the partial histograms; the last launch is the chosen one.
template <int L, int NP>
static int launch_xpass_bin_tuned(hipStream_t s, XPassArgs& a, int* nblocks, double* partial, size_t partial_doubles, int nmock) {
  using T = XPTune<L>;
  // the lean binning applies to |A|^2 with even multipoles <= 4 (a Legendre polynomial of mu^2)
  bool lean = !a.linear;
  for (int p = 0; p < NP; ++p) lean = lean && a.ell[p] % 2 == 0 && a.ell[p] <= 4;
  auto run = [&](int q) -> int {
    if constexpr (T::C1 > 0) {
      if (q == 1)
        return lean ? launch_xpass<L, true, NP, false, T::C1, true>(s, a, nblocks, partial, partial_doubles, nmock)
                    : launch_xpass<L, true, NP, false, T::C1>(s, a, nblocks, partial, partial_doubles, nmock);
    }
    if constexpr (T::C2 > 0) {
      if (q == 2)
        return lean ? launch_xpass<L, true, NP, false, T::C2, true>(s, a, nblocks, partial, partial_doubles, nmock)
                    : launch_xpass<L, true, NP, false, T::C2>(s, a, nblocks, partial, partial_doubles, nmock);
    }
    return lean ? launch_xpass<L, true, NP, false, T::C0, true>(s, a, nblocks, partial, partial_doubles, nmock)
                : launch_xpass<L, true, NP, false, T::C0>(s, a, nblocks, partial, partial_doubles, nmock);
  };
  const bool ok[3] = {true, T::C1 > 0, T::C2 > 0};
  const int forced = tune_forced(TK_XPASS_COLS);
  if (forced >= 0) return run(forced < 3 && ok[forced] ? forced : 0);
  char key[160];
  snprintf(key, sizeof(key), "xpass_cols L=%d ny=%d ncols=%d interlacing=%d products=%d catalogues=%d", L, a.ny, a.ncols, a.ni, NP,
           nmock);
  const int q0 = tune_get(key);
  if (q0 >= 0) return run(q0 < 3 && ok[q0] ? q0 : 0);
  if (!(ok[1] || ok[2]) || !tune_measure_ok(s)) return run(0);
  float ms[3];
  const int best = tune_time(s, 3, ok, run, ms);
  if (best < 0) return MP_ERR_HIP;
  tune_put(key, TK_XPASS_COLS, best, ms, 3);
  return run(best);
}

template <int L>
static int xpass_write_L(hipStream_t s, XPassArgs& a) {
  int nb = 0;
  return launch_xpass<L, false, 1>(s, a, &nb, nullptr, 0);
}

template <int L>
static int xpass_bin_L(hipStream_t s, XPassArgs& a, int nprod, double* partial, size_t partial_doubles, int* nblocks, int nmock) {
  if (a.ylm) {
    if (nprod != 1) return set_error(MP_ERR_INVALID, "x pass: Y_lm binning takes one product, got %d", nprod);
    return launch_xpass<L, true, 1, true>(s, a, nblocks, partial, partial_doubles, nmock);
  }
  switch (nprod) {
    case 1: return launch_xpass_bin_tuned<L, 1>(s, a, nblocks, partial, partial_doubles, nmock);
    case 2: return launch_xpass_bin_tuned<L, 2>(s, a, nblocks, partial, partial_doubles, nmock);
    case 3: return launch_xpass_bin_tuned<L, 3>(s, a, nblocks, partial, partial_doubles, nmock);
    case 4: return launch_xpass<L, true, 4>(s, a, nblocks, partial, partial_doubles, nmock);
    case 5: return launch_xpass<L, true, 5>(s, a, nblocks, partial, partial_doubles, nmock);
    case 6: return launch_xpass<L, true, 6>(s, a, nblocks, partial, partial_doubles, nmock);
    default: return set_error(MP_ERR_INVALID, "x pass: %d products not in 1..%d", nprod, XB_MAX_PROD);
  }
}

#define MP_XPASS_SWITCH2(n, F, ...)                                        \
  switch (n) {                                                             \
    case 2: rc = F<2>(__VA_ARGS__); break;                                 \
    case 4: rc = F<4>(__VA_ARGS__); break;                                 \
    case 8: rc = F<8>(__VA_ARGS__); break;                                 \
    case 16: rc = F<16>(__VA_ARGS__); break;                               \
    case 32: rc = F<32>(__VA_ARGS__); break;                               \
    case 64: rc = F<64>(__VA_ARGS__); break;                               \
    case 128: rc = F<128>(__VA_ARGS__); break;                             \
    case 256: rc = F<256>(__VA_ARGS__); break;                             \
    case 512: rc = F<512>(__VA_ARGS__); break;                             \
    case 1024: rc = F<1024>(__VA_ARGS__); break;                           \
    case 2048: rc = F<2048>(__VA_ARGS__); break;                           \
    default: return set_error(MP_ERR_UNSUPPORTED, "x length %d", (int)(n)); \
  }

#define MP_XPASS_SWITCH(n, F, ...)                                                                      \
  switch (n) {                                                                                          \
    case 2: return F<2>(__VA_ARGS__);                                                                   \
    case 4: return F<4>(__VA_ARGS__);                                                                   \
    case 8: return F<8>(__VA_ARGS__);                                                                   \
    case 16: return F<16>(__VA_ARGS__);                                                                 \
    case 32: return F<32>(__VA_ARGS__);                                                                 \
    case 64: return F<64>(__VA_ARGS__);                                                                 \
    case 128: return F<128>(__VA_ARGS__);                                                               \
    case 256: return F<256>(__VA_ARGS__);                                                               \
    case 512: return F<512>(__VA_ARGS__);                                                               \
    case 1024: return F<1024>(__VA_ARGS__);                                                             \
    case 2048: return F<2048>(__VA_ARGS__);                                                             \
    default: return set_error(MP_ERR_UNSUPPORTED, "x pass: length %d not a power of two in [2, 2048]", (int)(n)); \
  }

bool xpass_supported(const int32_t n[3]) {
  for (int q = 0; q < 3; ++q)
    if (!is_pow2(n[q]) || n[q] < 2 || n[q] > 2048) return false;
  return true;
}

// Meshes whose sizes are not all powers of two: the x FFT of each partial spectrum runs as an
// any-length pass into a scratch mesh, and this elementwise kernel forms
// acc (+)= exp(+i s_j (kx+ky+kz)) t / n_i, applying on the last shift the compensation, the
// zeroed (0,0,0) mode and, for power, |A|^2 (jaxpower/mesh.py:2739-2751).
__global__ void combine_any_kernel(const float2* __restrict__ t, int64_t tpitch, float2* __restrict__ acc, int64_t apitch, int n0,
                                   int n1, int ncols, CombineTables tab, int j, float scale, int first, int last, int zero_dc,
                                   int power) {
  const int64_t total = (int64_t)n0 * n1 * ncols;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int iz = (int)(i % ncols);
    const int64_t row = i / ncols;
    const int iy = (int)(row % n1), ix = (int)(row / n1);
    float2 v = t[row * tpitch + iz];
    if (tab.phx) {
      const int64_t o = (int64_t)j * tab.stride;
      v = cmul(v, cmul(tab.phx[o + ix], cmul(tab.phy[o + iy], tab.phz[o + iz])));
    }
    v = cscale(v, scale);
    float2* d = acc + row * apitch + iz;
    float2 a = first ? v : cadd(*d, v);
    if (last) {
      if (tab.cx) a = cscale(a, tab.cx[ix] * tab.cy[iy] * tab.cz[iz]);
      if (zero_dc && i == 0) a = make_float2(0.f, 0.f);
      if (power) a = make_float2(a.x * a.x + a.y * a.y, 0.f);
    }
    *d = a;
  }
}

// the any-length combine of the n_i partial spectra into acc (pitch apitch); tmp holds one mesh
static int combine_any(hipStream_t s, const float2* partial, int64_t pitch, const int32_t n[3], int ni, const CombineTables& tab,
                       int zero_dc, int power, float2* acc, int64_t apitch, float2* tmp, int64_t tpitch) {
  if (!fft_any_shape_ok(n)) return set_error(MP_ERR_UNSUPPORTED, "combine: meshsize (%d, %d, %d) not in 1..2048", n[0], n[1], n[2]);
  const int ncols = n[2] / 2 + 1;
  const int64_t ny = n[1], mstride = (int64_t)n[0] * n[1] * pitch;
  const int64_t total = (int64_t)n[0] * n[1] * ncols;
  const int64_t blocks = std::min<int64_t>(std::max<int64_t>(ceil_div(total, 256), 1), 65536);
  for (int j = 0; j < ni; ++j) {
    int rc = any_col_fft(s, partial + j * mstride, tmp, n[0], ncols, ny, ny * pitch, pitch, ny * tpitch, tpitch, false, 1.f,
                         "fft_xpass_any");
    if (rc) return rc;
    MP_LAUNCH("combine_any", combine_any_kernel, dim3((unsigned)blocks), dim3(256), 0, s, tmp, tpitch, acc, apitch, n[0], n[1],
              ncols, tab, j, 1.f / (float)ni, j == 0, j == ni - 1, zero_dc, power);
  }
  return MP_OK;
}

static void fill_common(XPassArgs& a, const float2* partial, int64_t pitch, const int32_t n[3], int ni, const CombineTables& tab,
                        int zero_dc) {
  a.in = partial;
  a.ncols = n[2] / 2 + 1;
  a.ny = n[1];
  a.ostride = pitch;
  a.fstride = (int64_t)n[1] * pitch;
  a.mesh_stride = (int64_t)n[0] * n[1] * pitch;
  a.ni = ni;
  a.tab = tab;
  a.scale = 1.f / (float)ni;
  a.zero_dc = zero_dc;
  a.tw = twiddle_device_ptr();
}

static size_t mesh_bytes(const int32_t n[3]) { return align_up((size_t)n[0] * n[1] * spectrum_pitch(n[2]) * sizeof(float2), 256); }

size_t combine_partial_workspace_bytes(const int32_t n[3], int ni) {
  return xpass_tables_bytes(n, ni) + (xpass_supported(n) ? 0 : mesh_bytes(n)) + 512;
}

int combine_partial(hipStream_t s, const float2* partial, int64_t pitch, const int32_t n[3], int order, int ni, int compensate,
                    int zero_dc, float2* out, int64_t out_pitch, void* ws, size_t ws_bytes, int power) {
  MP_REQUIRE(partial && out, "combine: null pointer");
  MP_REQUIRE(ni >= 1, "combine: interlacing %d < 1", ni);
  MP_REQUIRE(pitch >= n[2] / 2 + 1 && out_pitch >= n[2] / 2 + 1, "combine: pitch < Nz/2+1");
  const bool any = !xpass_supported(n);
  if (any && power == 2) return set_error(MP_ERR_UNSUPPORTED, "combine: meshsize (%d, %d, %d) not supported", n[0], n[1], n[2]);
  int rc = ensure_twiddles();
  if (rc) return rc;
  Workspace wsp(ws, ws_bytes);
  CombineTables tab;
  rc = combine_tables(s, n, order, ni, compensate, wsp, tab);
  if (rc) return rc;
  if (any) {
    const int64_t P = spectrum_pitch(n[2]);
    float2* tmp = wsp.take<float2>((size_t)n[0] * n[1] * P);
    if (!tmp) return set_error(MP_ERR_WORKSPACE, "combine: workspace of %zu bytes < %zu", ws_bytes, combine_partial_workspace_bytes(n, ni));
    return combine_any(s, partial, pitch, n, ni, tab, zero_dc, power, out, out_pitch, tmp, P);
  }
  XPassArgs a{};
  fill_common(a, partial, pitch, n, ni, tab, zero_dc);
  if (!a.tw) return set_error(MP_ERR_HIP, "combine: twiddle table address");
  a.out = out;
  a.out_ostride = out_pitch;
  a.out_fstride = (int64_t)n[1] * out_pitch;
  a.power = power;
  MP_XPASS_SWITCH(n[0], xpass_write_L, s, a);
}

constexpr int XB_MAX_BLOCKS = 1024;  // partial histograms: the bin-mode grid is capped here

size_t bin_partial_workspace_bytes(const int32_t n[3], int ni, int nprod, int nedges, int nmock) {
  if (!xpass_supported(n))  // combined mesh and one scratch mesh, then the plain binning
    return xpass_tables_bytes(n, ni) + 2 * mesh_bytes(n) + bin_workspace_bytes(nprod, nedges) + 1024;
  // a batch: per catalogue at most ceil(XB_MAX_BLOCKS / nmock) blocks rounded up to 8
  const size_t blocks = nmock > 1 ? (size_t)nmock * (ceil_div(ceil_div(XB_MAX_BLOCKS, nmock), 8) * 8) : XB_MAX_BLOCKS;
  return xpass_tables_bytes(n, ni) + align_up(blocks * nprod * (nedges - 1) * 2 * sizeof(double), 256) + 512;
}

int bin_partial(hipStream_t s, const float2* partial, int64_t pitch, const int32_t n[3], int order, int ni, int compensate,
                int zero_dc, const BinGeom& g, const double los[3], int nprod, const int32_t* ell, const int32_t* remove_zero,
                double* out, void* ws, size_t ws_bytes, int linear, const int32_t* ylm_m, bool fit_partials, int nmock) {
  MP_REQUIRE(partial && out && ell && remove_zero, "bin_partial: null pointer");
  MP_REQUIRE(ni >= 1, "bin_partial: interlacing %d < 1", ni);
  MP_REQUIRE(nmock >= 1 && nmock <= 65535, "bin_partial: %d catalogues not in 1..65535", nmock);
  MP_REQUIRE(nprod >= 1 && nprod <= XB_MAX_PROD, "bin_partial: %d products not in 1..%d", nprod, XB_MAX_PROD);
  const bool any = !xpass_supported(n);
  if (any && (linear || ylm_m || nmock > 1))
    return set_error(MP_ERR_UNSUPPORTED, "bin_partial: meshsize (%d, %d, %d) not supported", n[0], n[1], n[2]);
  if (nprod * (g.nedges - 1) > BIN_LDS_DOUBLES) return set_error(MP_ERR_UNSUPPORTED, "bin_partial: too many bins for the LDS histogram");
  MP_REQUIRE(pitch >= n[2] / 2 + 1, "bin_partial: pitch < Nz/2+1");
  int rc = ensure_twiddles();
  if (rc) return rc;
  Workspace wsp(ws, ws_bytes);
  CombineTables tab;
  rc = combine_tables(s, n, order, ni, compensate, wsp, tab);
  if (rc) return rc;
  if (any) {  // combine into a stored mesh, then bin |A|^2 L_ell(mu) as mp_bin_poles does
    const int64_t P = spectrum_pitch(n[2]);
    float2* acc = wsp.take<float2>((size_t)n[0] * n[1] * P);
    float2* tmp = wsp.take<float2>((size_t)n[0] * n[1] * P);
    const size_t bws = bin_workspace_bytes(nprod, g.nedges);
    void* bw = wsp.take<char>(bws);
    if (!acc || !tmp || !bw)
      return set_error(MP_ERR_WORKSPACE, "bin_partial: workspace of %zu bytes < %zu", ws_bytes,
                       bin_partial_workspace_bytes(n, ni, nprod, g.nedges));
    rc = combine_any(s, partial, pitch, n, ni, tab, zero_dc, 0, acc, P, tmp, P);
    if (rc) return rc;
    BinProducts prods{};
    prods.nmesh = 1;
    prods.nprod = nprod;
    prods.mesh[0] = acc;
    prods.pitch[0] = P;
    prods.need_mu = 0;
    for (int p = 0; p < nprod; ++p) {
      MP_REQUIRE(ell[p] >= 0 && ell[p] <= 10, "bin_partial: ell %d not in 0..10", ell[p]);
      prods.pi[p] = 0;
      prods.pj[p] = 0;
      prods.ell[p] = ell[p];
      prods.rz[p] = remove_zero[p];
      if (ell[p]) prods.need_mu = 1;
    }
    for (int q = 0; q < 3; ++q) prods.los[q] = los ? (float)los[q] : 0.f;
    return bin_poles(s, g, prods, out, bw, bws);
  }
  const size_t per_block = (size_t)nprod * (g.nedges - 1) * 2;
  size_t pd = (nmock > 1 ? (size_t)nmock * (ceil_div(ceil_div(XB_MAX_BLOCKS, nmock), 8) * 8) : (size_t)XB_MAX_BLOCKS) * per_block;
  if (fit_partials) {  // a fixed workspace: as many per-block histograms as it holds (fewer blocks)
    const size_t off = align_up(wsp.used, 256);
    const size_t room = ws_bytes > off ? (ws_bytes - off) / sizeof(double) : 0;
    if (room < pd) pd = room / per_block * per_block;
    if (pd < per_block)
      return set_error(MP_ERR_WORKSPACE, "bin_partial: %zu bytes hold the tables but not one partial histogram of %zu bytes",
                       ws_bytes, per_block * sizeof(double));
  }
  double* pbuf = wsp.take<double>(pd);
  if (!pbuf) return set_error(MP_ERR_WORKSPACE, "bin_partial: workspace too small");
  XPassArgs a{};
  fill_common(a, partial, pitch, n, ni, tab, zero_dc);
  if (!a.tw) return set_error(MP_ERR_HIP, "bin_partial: twiddle table address");
  a.mock_stride = ni * a.mesh_stride;
  a.g = g;
  a.need_mu = 0;
  a.linear = linear;
  a.ylm = ylm_m != nullptr;
  for (int p = 0; p < nprod; ++p) {
    MP_REQUIRE(ell[p] >= 0 && ell[p] <= 10, "bin_partial: ell %d not in 0..10", ell[p]);
    if (ylm_m) {
      MP_REQUIRE(ell[p] <= 4 && ylm_m[p] >= -ell[p] && ylm_m[p] <= ell[p], "bin_partial: (ell, m) = (%d, %d) not supported",
                 ell[p], ylm_m[p]);
      a.ylm_m[p] = ylm_m[p];
    }
    a.ell[p] = ell[p];
    a.rz[p] = remove_zero[p];
    legendre_table_set(a.leg, p, ell[p]);
    if (ell[p]) a.need_mu = 1;
  }
  for (int q = 0; q < 3; ++q) a.los[q] = los ? (float)los[q] : 0.f;
  int nblocks = 0;
  switch (n[0]) {
    case 2: rc = xpass_bin_L<2>(s, a, nprod, pbuf, pd, &nblocks, nmock); break;
    case 4: rc = xpass_bin_L<4>(s, a, nprod, pbuf, pd, &nblocks, nmock); break;
    case 8: rc = xpass_bin_L<8>(s, a, nprod, pbuf, pd, &nblocks, nmock); break;
    case 16: rc = xpass_bin_L<16>(s, a, nprod, pbuf, pd, &nblocks, nmock); break;
    case 32: rc = xpass_bin_L<32>(s, a, nprod, pbuf, pd, &nblocks, nmock); break;
    case 64: rc = xpass_bin_L<64>(s, a, nprod, pbuf, pd, &nblocks, nmock); break;
    case 128: rc = xpass_bin_L<128>(s, a, nprod, pbuf, pd, &nblocks, nmock); break;
    case 256: rc = xpass_bin_L<256>(s, a, nprod, pbuf, pd, &nblocks, nmock); break;
    case 512: rc = xpass_bin_L<512>(s, a, nprod, pbuf, pd, &nblocks, nmock); break;
    case 1024: rc = xpass_bin_L<1024>(s, a, nprod, pbuf, pd, &nblocks, nmock); break;
    case 2048: rc = xpass_bin_L<2048>(s, a, nprod, pbuf, pd, &nblocks, nmock); break;
    default: return set_error(MP_ERR_UNSUPPORTED, "bin_partial: x length %d", n[0]);
  }
  if (rc) return rc;
  return reduce_partials(s, pbuf, nblocks, nprod * (g.nedges - 1) * 2, out, nmock);
}

size_t ylm_bin_workspace_bytes(const int32_t n[3], int ell, int nedges) {
  const size_t nspec = (size_t)n[0] * n[1] * spectrum_pitch(n[2]);
  return align_up((size_t)(2 * ell + 1) * nspec * 8, 256) + align_up((size_t)XB_MAX_BLOCKS * 2 * (nedges - 1) * 2 * 8, 256) + 512;
}

int ylm_bin(hipStream_t s, const float* rmesh, const float2* A0, int64_t a0_pitch, const int32_t n[3], const double boxsize[3],
            const double boxcenter[3], int ell, int with0, const BinGeom& g, double* out, void* ws, size_t ws_bytes) {
  MP_REQUIRE(rmesh && A0 && out, "ylm_bin: null pointer");
  MP_REQUIRE(ell >= 1 && ell <= MAX_ELL, "ylm_bin: ell %d not in 1..%d", ell, MAX_ELL);
  if (!xpass_supported(n)) return set_error(MP_ERR_UNSUPPORTED, "ylm_bin: meshsize (%d, %d, %d) not supported", n[0], n[1], n[2]);
  if ((with0 ? 2 : 1) * (g.nedges - 1) * 2 > BIN_LDS_DOUBLES) return set_error(MP_ERR_UNSUPPORTED, "ylm_bin: too many bins");
  MP_REQUIRE(a0_pitch >= n[2] / 2 + 1, "ylm_bin: A0 pitch < Nz/2+1");
  const int64_t P = spectrum_pitch(n[2]);
  const int64_t nspec = (int64_t)n[0] * n[1] * P;
  const int nm = 2 * ell + 1;
  Workspace wsp(ws, ws_bytes);
  float2* Z = wsp.take<float2>((size_t)nm * nspec);
  const size_t pd = (size_t)XB_MAX_BLOCKS * 2 * (g.nedges - 1) * 2;
  double* pbuf = wsp.take<double>(pd);
  if (!Z || !pbuf) return set_error(MP_ERR_WORKSPACE, "ylm_bin: workspace of %zu bytes < %zu", ws_bytes, ylm_bin_workspace_bytes(n, ell, g.nedges));
  RealPrologue pro{};
  pro.kind = PRO_YLM;
  pro.ell = ell;
  pro.ny = n[1];
  pro.nz = n[2];
  pro.x0 = boxcenter[0] - boxsize[0] / 2.;
  pro.y0 = boxcenter[1] - boxsize[1] / 2.;
  pro.z0 = boxcenter[2] - boxsize[2] / 2.;
  pro.dx = boxsize[0] / n[0];
  pro.dy = boxsize[1] / n[1];
  pro.dz = boxsize[2] / n[2];
  int rc = rfft_zy_ylm(s, rmesh, Z, nspec, n, P, pro);
  if (rc) return rc;
  YlmBinArgs a{};
  a.Z = Z;
  a.mesh_stride = nspec;
  a.ostride = P;
  a.fstride = (int64_t)n[1] * P;
  a.A0 = A0;
  a.a0_ostride = a0_pitch;
  a.a0_fstride = (int64_t)n[1] * a0_pitch;
  a.nm = nm;
  a.ell = ell;
  a.ncols = n[2] / 2 + 1;
  a.ny = n[1];
  a.kp.ell = ell;
  a.kp.nx = n[0];
  a.kp.ny = n[1];
  a.kp.kfx = (float)(2. * M_PI / boxsize[0]);
  a.kp.kfy = (float)(2. * M_PI / boxsize[1]);
  a.kp.kfz = (float)(2. * M_PI / boxsize[2]);
  a.tw = twiddle_device_ptr();
  if (!a.tw) return set_error(MP_ERR_HIP, "ylm_bin: twiddle table address");
  a.g = g;
  a.with0 = with0 ? 1 : 0;
  int nblocks = 0;
  MP_XPASS_SWITCH2(n[0], ylm_bin_L, s, a, pbuf, pd, &nblocks);
  if (rc) return rc;
  return reduce_partials(s, pbuf, nblocks, (with0 ? 2 : 1) * (g.nedges - 1) * 2, out);
}

}  // namespace mp
