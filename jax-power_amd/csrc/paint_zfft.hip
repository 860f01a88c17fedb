// Fused paint + z-axis r2c for the interlaced spectrum path.
//
// The n_i shifted meshes of _paint (jaxpower/mesh.py:2707-2747, resamplers.py:73-103) are
// never written to HBM as real meshes.  Particles are counting-sorted by their (x, y) home
// cell; one workgroup owns a block of RX x RY full z-rows of one shifted mesh (up to 132 KiB
// of LDS, which gfx950 allows per workgroup), gathers every particle whose stencil reaches
// the block, accumulates the painted rows in LDS, runs the z-axis real-to-complex FFT on the
// rows in place, and writes the half-spectrum rows once.  The y and x passes follow as in
// fft.hip.
//
// Record layout of a sorted particle: float4(fx, fy, zbits, w): fx, fy in [0, 1) are the
// offsets from the home cell along x and y, zbits is z~ (wrapped into [0, Nz)) as unsigned
// fixed point with ZFRAC fractional bits, w the weight.
#include "common.h"
#include "fft.h"
#include "fft_core.h"
#include "paint.h"

namespace mp {

constexpr int ZFRAC = 20;                 // fractional bits of z in the record (Nz <= 2048)
constexpr int PZ_NT = 1024;
constexpr int PZ_MAX_SEG = 512;
constexpr size_t PZ_LDS_BUDGET = 132 * 1024 + 1024;

int exclusive_scan_i32(hipStream_t s, int* data, int64_t n, int* tmp, int* copy);
size_t scan_workspace_bytes(int64_t n);
int fill_i32_dev(hipStream_t s, int* p, int64_t n, int v);

struct RowGeom {
  int n[3];
  int rx, ry;          // rows per block along x and y
  int nbx, nby;        // blocks per axis
  int lo, hi;          // stencil reach relative to the home cell
  int ni;
};

template <int ORDER>
__host__ __device__ __forceinline__ float kern1(float s) {
  if constexpr (ORDER == 1) {
    return 1.f;
  } else if constexpr (ORDER == 2) {
    return 1.f - s;
  } else if constexpr (ORDER == 3) {
    return s <= 0.5f ? 0.75f - s * s : 0.5f * (1.5f - s) * (1.5f - s);
  } else {
    return s <= 1.f ? (4.f - 6.f * s * s + 3.f * s * s * s) * (1.f / 6.f) : (2.f - s) * (2.f - s) * (2.f - s) * (1.f / 6.f);
  }
}

__device__ __forceinline__ int pmod(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

template <typename T>
__device__ __forceinline__ void grid_coords(const T* __restrict__ pos, int64_t i, const GridMap& g, T xg[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) xg[a] = (pos[3 * i + a] + (T)g.shift[a]) / (T)g.cell[a];
}

template <typename T>
__global__ void row_count_kernel(const T* __restrict__ pos, int64_t np, GridMap g, int* __restrict__ counts) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  T xg[3];
  grid_coords(pos, i, g, xg);
  int hx = pmod((int)floor(xg[0]), g.n[0]), hy = pmod((int)floor(xg[1]), g.n[1]);
  atomicAdd(&counts[(int64_t)hx * g.n[1] + hy], 1);
}

template <typename T>
__global__ void row_scatter_kernel(const T* __restrict__ pos, const float* __restrict__ w, int64_t np, GridMap g,
                                   int* __restrict__ cursor, float4* __restrict__ sorted) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  T xg[3];
  grid_coords(pos, i, g, xg);
  T fx = floor(xg[0]), fy = floor(xg[1]);
  int hx = pmod((int)fx, g.n[0]), hy = pmod((int)fy, g.n[1]);
  float ox = fminf(fmaxf((float)(xg[0] - fx), 0.f), 1.f);
  float oy = fminf(fmaxf((float)(xg[1] - fy), 0.f), 1.f);
  // z wrapped into [0, Nz) as fixed point
  T z = xg[2] - floor(xg[2] / (T)g.n[2]) * (T)g.n[2];
  double zs = (double)z * (double)(1u << ZFRAC);
  double zmax = (double)g.n[2] * (double)(1u << ZFRAC) - 1.;
  unsigned zbits = (unsigned)fmin(fmax(zs, 0.), zmax);
  int slot = atomicAdd(&cursor[(int64_t)hx * g.n[1] + hy], 1);
  sorted[slot] = make_float4(ox, oy, __uint_as_float(zbits), w ? w[i] : 1.f);
}

// One workgroup: RX x RY rows of shifted mesh j, painted in LDS, then z-FFT'd (r2c) in LDS.
template <int ORDER, int M, int RX, int RY>
__global__ void __launch_bounds__(PZ_NT)
paint_zfft_kernel(const float4* __restrict__ sorted, const int* __restrict__ offsets, RowGeom rg, float2* __restrict__ out,
                  int64_t pitch, int64_t mesh_stride, const float2* __restrict__ tw) {
  constexpr int NZ = 2 * M;
  constexpr int LPC = M + 1;       // complex row pitch in LDS
  constexpr int NROW = RX * RY;
  constexpr int OMIN = -((ORDER - 1) / 2);
  extern __shared__ float2 lds2[];
  float* rows = reinterpret_cast<float*>(lds2);
  __shared__ int sstart[PZ_MAX_SEG];
  __shared__ int spre[PZ_MAX_SEG + 1];
  const int tid = threadIdx.x;

  // XCD-aware work index: blocks sharing an x strip (hence most of their particles) and the
  // shifts of one block run on the same XCD, i.e. share an L2.
  const int id = blockIdx.x;
  const int xcd = id & 7;
  const int k = id >> 3;
  const int j = k % rg.ni;
  const int rest = k / rg.ni;
  const int by = rest % rg.nby;
  const int bx = (rest / rg.nby) * 8 + xcd;
  if (bx >= rg.nbx) return;
  const int x0 = bx * RX, y0 = by * RY;
  const float s = (float)j / (float)rg.ni;

  for (int i = tid; i < NROW * LPC; i += PZ_NT) lds2[i] = make_float2(0.f, 0.f);

  // neighbourhood rows: homes in [x0 - hi, x0 + RX - 1 - lo]
  const int sx = RX + rg.hi - rg.lo, sy = RY + rg.hi - rg.lo;
  const int nseg = sx * sy;
  for (int i = tid; i < nseg; i += PZ_NT) {
    int ix = i / sy, iy = i % sy;
    int hx = pmod(x0 - rg.hi + ix, rg.n[0]), hy = pmod(y0 - rg.hi + iy, rg.n[1]);
    int64_t key = (int64_t)hx * rg.n[1] + hy;
    int st = offsets[key];
    sstart[i] = st;
    spre[i + 1] = offsets[key + 1] - st;
  }
  __syncthreads();
  if (tid < 64) {
    constexpr int PER = PZ_MAX_SEG / 64;
    int local[PER];
    int sum = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      int i = tid * PER + q;
      local[q] = i < nseg ? spre[i + 1] : 0;
      sum += local[q];
    }
    int incl = sum;
#pragma unroll
    for (int d = 1; d < 64; d *= 2) {
      int v = __shfl_up(incl, d, 64);
      if (tid >= d) incl += v;
    }
    int run = incl - sum;
    if (tid == 0) spre[0] = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      int i = tid * PER + q;
      run += local[q];
      if (i < nseg) spre[i + 1] = run;
    }
  }
  __syncthreads();

  const int total = spre[nseg];
  for (int p = tid; p < total; p += PZ_NT) {
    int lo_i = 0, hi_i = nseg - 1;
    while (lo_i < hi_i) {
      int mid = (lo_i + hi_i + 1) >> 1;
      if (spre[mid] <= p) lo_i = mid;
      else hi_i = mid - 1;
    }
    const int seg = lo_i;
    const float4 rec = sorted[sstart[seg] + (p - spre[seg])];
    const int hxr = seg / sy - rg.hi;   // home row relative to x0 (unwrapped)
    const int hyr = seg % sy - rg.hi;
    const unsigned zb = __float_as_uint(rec.z);
    const int zc = (int)(zb >> ZFRAC);
    const float fz = (float)(zb & ((1u << ZFRAC) - 1u)) * (1.f / (float)(1u << ZFRAC));
    int cx[ORDER], cy[ORDER], cz[ORDER];
    float kx[ORDER], ky[ORDER], kz[ORDER];
    {
      float u = rec.x + s;
      float r = (ORDER % 2) ? rintf(u) : floorf(u);
#pragma unroll
      for (int o = 0; o < ORDER; ++o) {
        cx[o] = hxr + (int)r + OMIN + o;
        kx[o] = (cx[o] >= 0 && cx[o] < RX) ? kern1<ORDER>(fabsf(r + (float)(OMIN + o) - u)) : 0.f;
      }
      u = rec.y + s;
      r = (ORDER % 2) ? rintf(u) : floorf(u);
#pragma unroll
      for (int o = 0; o < ORDER; ++o) {
        cy[o] = hyr + (int)r + OMIN + o;
        ky[o] = (cy[o] >= 0 && cy[o] < RY) ? kern1<ORDER>(fabsf(r + (float)(OMIN + o) - u)) : 0.f;
      }
      u = fz + s;
      r = (ORDER % 2) ? rintf(u) : floorf(u);
#pragma unroll
      for (int o = 0; o < ORDER; ++o) {
        cz[o] = (zc + (int)r + OMIN + o) & (NZ - 1);
        kz[o] = kern1<ORDER>(fabsf(r + (float)(OMIN + o) - u));
      }
    }
#pragma unroll
    for (int ox = 0; ox < ORDER; ++ox) {
      if (!(cx[ox] >= 0 && cx[ox] < RX)) continue;
#pragma unroll
      for (int oy = 0; oy < ORDER; ++oy) {
        if (!(cy[oy] >= 0 && cy[oy] < RY)) continue;
        float* row = rows + (cx[ox] * RY + cy[oy]) * (2 * LPC);
        const float kxy = rec.w * kx[ox] * ky[oy];
#pragma unroll
        for (int oz = 0; oz < ORDER; ++oz) atomicAdd(row + cz[oz], kxy * kz[oz]);
      }
    }
  }
  __syncthreads();

  // z-axis r2c of the NROW rows held in LDS (row r: complex elements lds2[r*LPC + l])
  fft_lds<M, NROW, 1, LPC, false, PZ_NT, false>(lds2, tw);
  constexpr int NK = M / 2 + 1;
  float2* dst = out + j * mesh_stride;
  for (int idx = tid; idx < NROW * NK; idx += PZ_NT) {
    int r = idx / NK, kk = idx % NK;
    int rx = r / RY, ry = r % RY;
    float2* o = dst + ((int64_t)(x0 + rx) * rg.n[1] + (y0 + ry)) * pitch;
    rfft_post(lds2 + r * LPC, o, M, kk, tw);
  }
}

// ------------------------------------------------------------------ host side

template <int M>
struct PZCfg;
template <> struct PZCfg<1024> { static constexpr int RX = 4, RY = 4; };
template <> struct PZCfg<512> { static constexpr int RX = 8, RY = 4; };
template <> struct PZCfg<256> { static constexpr int RX = 8, RY = 8; };
template <> struct PZCfg<128> { static constexpr int RX = 16, RY = 8; };
template <> struct PZCfg<64> { static constexpr int RX = 16, RY = 16; };
template <> struct PZCfg<32> { static constexpr int RX = 16, RY = 16; };
template <> struct PZCfg<16> { static constexpr int RX = 16, RY = 16; };
template <> struct PZCfg<8> { static constexpr int RX = 16, RY = 16; };

static void reach_of(int order, int ni, int& lo, int& hi) {
  int omin = -((order - 1) / 2), omax = order - 1 - (order - 1) / 2;
  int i0max = (order % 2) ? (ni > 1 ? 2 : 1) : (ni > 1 ? 1 : 0);
  lo = omin;
  hi = i0max + omax + 1;  // one extra cell of margin for offsets rounded up to 1
}

static bool pz_rows(int M, int& rx, int& ry) {
  switch (M) {
    case 1024: rx = PZCfg<1024>::RX; ry = PZCfg<1024>::RY; return true;
    case 512: rx = PZCfg<512>::RX; ry = PZCfg<512>::RY; return true;
    case 256: rx = PZCfg<256>::RX; ry = PZCfg<256>::RY; return true;
    case 128: rx = PZCfg<128>::RX; ry = PZCfg<128>::RY; return true;
    case 64: rx = PZCfg<64>::RX; ry = PZCfg<64>::RY; return true;
    case 32: rx = PZCfg<32>::RX; ry = PZCfg<32>::RY; return true;
    case 16: rx = PZCfg<16>::RX; ry = PZCfg<16>::RY; return true;
    case 8: rx = PZCfg<8>::RX; ry = PZCfg<8>::RY; return true;
    default: return false;
  }
}

bool paint_zfft_supported(const int32_t n[3], int64_t np, int order, int ni) {
  if (order < 1 || order > 4 || ni < 1 || np >= (int64_t)INT32_MAX) return false;
  if (!is_pow2(n[2])) return false;
  int rx, ry;
  if (!pz_rows(n[2] / 2, rx, ry)) return false;
  if (n[0] % rx || n[1] % ry) return false;
  int lo, hi;
  reach_of(order, ni, lo, hi);
  // neighbourhood rows must not alias through the periodic wrap, and fit the segment table
  if (rx + hi - lo > n[0] || ry + hi - lo > n[1]) return false;
  return (rx + hi - lo) * (ry + hi - lo) <= PZ_MAX_SEG;
}

size_t paint_zfft_workspace_bytes(const int32_t n[3], int64_t np) {
  int64_t nkeys = (int64_t)n[0] * n[1];
  size_t bytes = align_up((size_t)(nkeys + 1) * 4, 256) * 2 + scan_workspace_bytes(nkeys + 1);
  bytes += align_up((size_t)(np > 0 ? np : 1) * 16, 256);
  return bytes + 256;
}

template <int ORDER, int M>
static int launch_pz(hipStream_t s, const float4* sorted, const int* offsets, const RowGeom& rg, float2* out, int64_t pitch,
                     int64_t mesh_stride) {
  constexpr int RX = PZCfg<M>::RX, RY = PZCfg<M>::RY;
  constexpr size_t lds = (size_t)RX * RY * (M + 1) * sizeof(float2);
  static_assert(lds <= PZ_LDS_BUDGET, "LDS budget");
  auto kern = paint_zfft_kernel<ORDER, M, RX, RY>;
  if (lds > 65536) MP_CHECK_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  // ids are laid out as ((q * nby + by) * ni + j) * 8 + xcd with bx = q * 8 + xcd
  int64_t nq = ceil_div(rg.nbx, 8);
  unsigned grid = (unsigned)(nq * rg.nby * rg.ni * 8);
  const float2* tw = twiddle_device_ptr();
  if (!tw) return set_error(MP_ERR_HIP, "paint_zfft: no twiddle table");
  MP_LAUNCH("paint_zfft", kern, dim3(grid), dim3(PZ_NT), lds, s, sorted, offsets, rg, out, pitch, mesh_stride, tw);
  return MP_OK;
}

template <int ORDER>
static int dispatch_pz(int M, hipStream_t s, const float4* sorted, const int* offsets, const RowGeom& rg, float2* out, int64_t pitch,
                       int64_t mesh_stride) {
  switch (M) {
    case 1024: return launch_pz<ORDER, 1024>(s, sorted, offsets, rg, out, pitch, mesh_stride);
    case 512: return launch_pz<ORDER, 512>(s, sorted, offsets, rg, out, pitch, mesh_stride);
    case 256: return launch_pz<ORDER, 256>(s, sorted, offsets, rg, out, pitch, mesh_stride);
    case 128: return launch_pz<ORDER, 128>(s, sorted, offsets, rg, out, pitch, mesh_stride);
    case 64: return launch_pz<ORDER, 64>(s, sorted, offsets, rg, out, pitch, mesh_stride);
    case 32: return launch_pz<ORDER, 32>(s, sorted, offsets, rg, out, pitch, mesh_stride);
    case 16: return launch_pz<ORDER, 16>(s, sorted, offsets, rg, out, pitch, mesh_stride);
    case 8: return launch_pz<ORDER, 8>(s, sorted, offsets, rg, out, pitch, mesh_stride);
    default: return set_error(MP_ERR_UNSUPPORTED, "paint_zfft: Nz/2 = %d", M);
  }
}

template <typename T>
static int paint_zfft_typed(hipStream_t s, const T* pos, const float* w, int64_t np, const GridMap& g, int order, int ni,
                            float2* out, int64_t pitch, int64_t mesh_stride, void* ws, size_t ws_bytes) {
  const int32_t n[3] = {g.n[0], g.n[1], g.n[2]};
  const int64_t nkeys = (int64_t)n[0] * n[1];
  Workspace wsp(ws, ws_bytes);
  int* offsets = wsp.take<int>(nkeys + 1);
  int* cursor = wsp.take<int>(nkeys + 1);
  int* scan_tmp = (int*)wsp.take<char>(scan_workspace_bytes(nkeys + 1));
  float4* sorted = wsp.take<float4>(np > 0 ? np : 1);
  if (!offsets || !cursor || !scan_tmp || !sorted)
    return set_error(MP_ERR_WORKSPACE, "paint_zfft: workspace of %zu bytes < %zu", ws_bytes, paint_zfft_workspace_bytes(n, np));
  int rc = fill_i32_dev(s, offsets, nkeys + 1, 0);
  if (rc) return rc;
  if (np > 0) MP_LAUNCH("paint_row_count", row_count_kernel<T>, dim3((unsigned)ceil_div(np, 256)), dim3(256), 0, s, pos, np, g, offsets);
  rc = exclusive_scan_i32(s, offsets, nkeys + 1, scan_tmp, cursor);
  if (rc) return rc;
  if (np > 0)
    MP_LAUNCH("paint_row_scatter", row_scatter_kernel<T>, dim3((unsigned)ceil_div(np, 256)), dim3(256), 0, s, pos, w, np, g, cursor,
              sorted);
  RowGeom rg;
  for (int a = 0; a < 3; ++a) rg.n[a] = n[a];
  pz_rows(n[2] / 2, rg.rx, rg.ry);
  rg.nbx = n[0] / rg.rx;
  rg.nby = n[1] / rg.ry;
  reach_of(order, ni, rg.lo, rg.hi);
  rg.ni = ni;
  const int M = n[2] / 2;
  switch (order) {
    case 1: return dispatch_pz<1>(M, s, sorted, offsets, rg, out, pitch, mesh_stride);
    case 2: return dispatch_pz<2>(M, s, sorted, offsets, rg, out, pitch, mesh_stride);
    case 3: return dispatch_pz<3>(M, s, sorted, offsets, rg, out, pitch, mesh_stride);
    default: return dispatch_pz<4>(M, s, sorted, offsets, rg, out, pitch, mesh_stride);
  }
}

// ni z-transformed half-spectrum rows (pitch `pitch`, mesh j at out + j * mesh_stride).
int paint_zfft(hipStream_t s, const void* pos, int pos_f64, const float* w, int64_t np, const GridMap& g, int order, int ni,
               float2* out, int64_t pitch, int64_t mesh_stride, void* ws, size_t ws_bytes) {
  int rc = ensure_twiddles();
  if (rc) return rc;
  if (pos_f64) return paint_zfft_typed<double>(s, (const double*)pos, w, np, g, order, ni, out, pitch, mesh_stride, ws, ws_bytes);
  return paint_zfft_typed<float>(s, (const float*)pos, w, np, g, order, ni, out, pitch, mesh_stride, ws, ws_bytes);
}

}  // namespace mp
