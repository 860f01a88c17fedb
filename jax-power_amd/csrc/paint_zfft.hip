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
// fixed point with zfrac = 32 - ceil(log2 Nz) fractional bits (rounded to nearest), w the weight.
#include "common.h"
#include "fft.h"
#include "fft_core.h"
#include "paint.h"
#include "tune.h"

#include <cstring>
#include <map>
#include <mutex>
#include <string>

namespace mp {

// Fractional bits of z in a record: all the bits of a 32-bit word that the cell index leaves.
static int zfrac_of(int nz) {
  int b = 1;
  while ((1 << b) < nz) ++b;
  return 32 - b;
}
constexpr int PZ_NT = 1024;
#ifndef PZ_YROT
#define PZ_YROT 1  // lanes rotate their y stencil slots as well as z (see LaneRotation)
#endif
// Segment table (home rows of a block's neighbourhood): the reach of a stencil is at most
// p + 1 <= 5 rows per axis, so (RX + 5) (RY + 5) rounded to a wave.
__host__ __device__ constexpr int pz_seg_capacity(int rx, int ry) { return ((rx + 5) * (ry + 5) + 63) / 64 * 64; }
constexpr size_t PZ_LDS_BUDGET = 140 * 1024;  // dynamic part; the static tables add at most 20 KiB

int exclusive_scan_i32(hipStream_t s, int* data, int64_t n, int* tmp, int* copy);
size_t scan_workspace_bytes(int64_t n);
int fill_i32_dev(hipStream_t s, int* p, int64_t n, int v);

struct RowGeom {
  int n[3];
  int rx, ry;          // rows per block along x and y
  int nbx, nby;        // blocks per axis
  int lo, hi;          // stencil reach relative to the home cell
  int ni;
  int zfrac;           // fractional bits of z in the records
  int nmock;           // catalogues stacked along x in the row sort (1 for one catalogue)
  const unsigned* wmax_bits;  // per catalogue: bits of the largest |w| (a non-negative float)
  const int* negw;            // per catalogue: nonzero when some weight is negative
};

// Catalogues stacked in one particle array (a batch of mocks): catalogue m holds particles
// [moff[m], moff[m + 1]).  A launch cuts every catalogue into chunks of its own size CH that
// never straddle two catalogues; choff[m] is the number of chunks before catalogue m.  One
// catalogue has moff == nullptr and its chunks are those of [0, np).
struct Catalogues {
  const int64_t* moff;  // device, nmock + 1 entries
  const int* choff;     // device, nmock + 1 entries, for the launch's chunk size
  int nmock;
  int64_t np;
};

// Catalogue and particle range of chunk `chunk`; false past the last chunk.
__device__ __forceinline__ bool chunk_range(const Catalogues& c, int chunk, int CH, int& m, int64_t& i0, int64_t& i1) {
  if (!c.moff) {
    m = 0;
    i0 = (int64_t)chunk * CH;
    i1 = i0 + CH < c.np ? i0 + CH : c.np;
    return i0 < i1;
  }
  if (chunk >= c.choff[c.nmock]) return false;
  int lo = 0, hi = c.nmock - 1;  // the last catalogue with choff <= chunk (it has chunks)
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (c.choff[mid] <= chunk) lo = mid;
    else hi = mid - 1;
  }
  m = lo;
  i0 = c.moff[m] + (int64_t)(chunk - c.choff[m]) * CH;
  i1 = i0 + CH < c.moff[m + 1] ? i0 + CH : c.moff[m + 1];
  return i0 < i1;
}

// Home z cell and offset in [0, 1] of a record's z word (the offset rounds to float).
__device__ __forceinline__ void z_of_record(unsigned zb, int zfrac, int& zc, float& fz) {
  zc = (int)(zb >> zfrac);
  fz = (float)(zb << (32 - zfrac)) * 2.3283064365386963e-10f;  // 2^-32
}

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
// The same for the last kernel to read the positions (the row-sort scatter): non-temporal
// loads, which leave the L2 to the records being written (512^3, 1.6e7 particles: scatter and
// fine sort 0.171 + 0.171 -> 0.165 + 0.160 ms).
template <typename T>
__device__ __forceinline__ void grid_coords_last(const T* __restrict__ pos, int64_t i, const GridMap& g, T xg[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) xg[a] = (__builtin_nontemporal_load(pos + 3 * i + a) + (T)g.shift[a]) / (T)g.cell[a];
}

// ------------------------------------------------------------------ two-level row sort
//
// Level A buckets particles by home x (atomics on an LDS histogram per block, one global
// atomic per (block, bucket) to reserve space); level B sorts each x bucket by home y in one
// workgroup and writes the (x, y) row offsets directly.  No per-particle global atomics.

constexpr int RS_NT = 1024;
constexpr int RS_CHUNK = 16384;
constexpr int RS_MAX_KEYS = 2048;

template <typename T>
__device__ __forceinline__ void row_record(const T* __restrict__ pos, const float* __restrict__ w, int64_t i, const GridMap& g,
                                           int zfrac, int& hx, int& hy, float4& rec) {
  T xg[3];
  grid_coords_last(pos, i, g, xg);
  T fx = floor(xg[0]), fy = floor(xg[1]);
  hx = pmod((int)fx, g.n[0]);
  hy = pmod((int)fy, g.n[1]);
  // offsets strictly below 1, so that round/floor(offset + shift) takes two values at most
  float ox = fminf(fmaxf((float)(xg[0] - fx), 0.f), 0.99999994f);
  float oy = fminf(fmaxf((float)(xg[1] - fy), 0.f), 0.99999994f);
  // z wrapped into [0, Nz) as fixed point
  T z = xg[2] - floor(xg[2] / (T)g.n[2]) * (T)g.n[2];
  const double zone = (double)(1ull << zfrac);
  double zs = rint((double)z * zone);
  double zmax = (double)g.n[2] * zone - 1.;
  unsigned zbits = (unsigned)fmin(fmax(zs, 0.), zmax);
  rec = make_float4(ox, oy, __uint_as_float(zbits), w ? __builtin_nontemporal_load(w + i) : 1.f);
}

__device__ __forceinline__ float wave_max_f32(float v) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v = fmaxf(v, __shfl_xor(v, d, 64));
  return v;
}

template <typename T>
__global__ void __launch_bounds__(RS_NT)
rowsort_hist_kernel(const T* __restrict__ pos, const float* __restrict__ w, Catalogues cat, GridMap g, int* __restrict__ bucket_total,
                    unsigned* __restrict__ wmax_bits, int* __restrict__ negw) {
  __shared__ int hist[RS_MAX_KEYS];
  __shared__ float wmax[RS_NT / 64];
  int mk;
  int64_t i0, i1;
  if (!chunk_range(cat, blockIdx.x, RS_CHUNK, mk, i0, i1)) return;
  bucket_total += (int64_t)mk * g.n[0];  // the x buckets of catalogue mk
  for (int b = threadIdx.x; b < g.n[0]; b += RS_NT) hist[b] = 0;
  __syncthreads();
  float m = w ? 0.f : 1.f;
  bool neg = false;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += RS_NT) {
    T xg[3];
    grid_coords(pos, i, g, xg);
    atomicAdd(&hist[pmod((int)floor(xg[0]), g.n[0])], 1);
    if (w) {
      m = fmaxf(m, fabsf(w[i]));
      neg = neg || w[i] < 0.f;
    }
  }
  if (__any(neg) && (threadIdx.x & 63) == 0) atomicOr(negw + mk, 1);
  m = wave_max_f32(m);
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
  __syncthreads();
  for (int b = threadIdx.x; b < g.n[0]; b += RS_NT)
    if (hist[b]) atomicAdd(&bucket_total[b], hist[b]);
  if (threadIdx.x == 0) {
    float bm = 0.f;
    for (int q = 0; q < RS_NT / 64; ++q) bm = fmaxf(bm, wmax[q]);
    atomicMax(wmax_bits + mk, __float_as_uint(bm));  // non-negative floats order like their bits
  }
}

__device__ __forceinline__ int wave_scan_incl(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d *= 2) {
    int t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  return v;
}

// Non-temporal load of a record that is read once (the fine sort's input: 512^3 fine sort
// 0.227 -> 0.171 ms).
typedef float rec_f4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 load_rec_stream(const float4* p) {
  const rec_f4 v = __builtin_nontemporal_load(reinterpret_cast<const rec_f4*>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}

// Exclusive scan of data[0..n) in LDS, n <= 2 * RS_NT; returns the total.
__device__ int block_excl_scan(int* data, int n, int* wsum) {
  const int tid = threadIdx.x;
  int a = 2 * tid < n ? data[2 * tid] : 0;
  int b = 2 * tid + 1 < n ? data[2 * tid + 1] : 0;
  int local = a + b;
  int incl = wave_scan_incl(local);
  __syncthreads();
  if ((tid & 63) == 63) wsum[tid >> 6] = incl;
  __syncthreads();
  if (tid < 64) {
    int x = tid < RS_NT / 64 ? wsum[tid] : 0;
    int xi = wave_scan_incl(x);
    if (tid < RS_NT / 64) wsum[tid] = xi - x;
    if (tid == RS_NT / 64 - 1) wsum[RS_NT / 64] = xi;
  }
  __syncthreads();
  int run = wsum[tid >> 6] + incl - local;
  if (2 * tid < n) data[2 * tid] = run;
  if (2 * tid + 1 < n) data[2 * tid + 1] = run + a;
  __syncthreads();
  return wsum[RS_NT / 64];
}


// Level A scatter: each block builds the records of RS_SCHUNK particles once, orders them by x
// bucket in LDS, reserves one range per (block, bucket) in the bucket arrays, and writes each
// bucket's run with consecutive lanes, so the global writes are contiguous runs rather than
// scattered 16-byte records.
constexpr int RS_SCHUNK = 6144;

template <typename T, int SCH = RS_SCHUNK>
__global__ void __launch_bounds__(RS_NT)
rowsort_scatter_kernel(const T* __restrict__ pos, const float* __restrict__ w, Catalogues cat, GridMap g, int zfrac,
                       int* __restrict__ bucket_cursor, float4* __restrict__ tmp_rec, short* __restrict__ tmp_key) {
  __shared__ int hist[RS_MAX_KEYS];
  __shared__ int lstart[RS_MAX_KEYS];
  __shared__ int gbase[RS_MAX_KEYS];
  __shared__ int wsum[RS_NT / 64 + 1];
  extern __shared__ float4 srec[];  // SCH records, then their y keys and x buckets
  short* skey = reinterpret_cast<short*>(srec + SCH);
  short* sbkt = skey + SCH;
  const int n0 = g.n[0];
  int mk;
  int64_t i0, i1;
  if (!chunk_range(cat, blockIdx.x, SCH, mk, i0, i1)) return;
  bucket_cursor += (int64_t)mk * n0;  // the x buckets of catalogue mk
  for (int b = threadIdx.x; b < n0; b += RS_NT) hist[b] = 0;
  __syncthreads();
  const int cnt = (int)(i1 - i0);
  constexpr int PER = SCH / RS_NT;
  int rank[PER], hxq[PER], hyq[PER];
  float4 recq[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int t = threadIdx.x + q * RS_NT;
    if (t < cnt) {
      row_record(pos, w, i0 + t, g, zfrac, hxq[q], hyq[q], recq[q]);
      rank[q] = atomicAdd(&hist[hxq[q]], 1);
    }
  }
  __syncthreads();
  for (int b = threadIdx.x; b < n0; b += RS_NT) {
    lstart[b] = hist[b];
    if (hist[b]) gbase[b] = atomicAdd(&bucket_cursor[b], hist[b]);
  }
  __syncthreads();
  for (int b0 = 0; b0 < n0; b0 += 2 * RS_NT) {  // exclusive scan of the bucket counts
    const int m = min(2 * RS_NT, n0 - b0);
    const int carry = b0 ? lstart[b0 - 1] + hist[b0 - 1] : 0;
    block_excl_scan(lstart + b0, m, wsum);
    for (int b = threadIdx.x; b < m; b += RS_NT) lstart[b0 + b] += carry;
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int t = threadIdx.x + q * RS_NT;
    if (t < cnt) {
      const int slot = lstart[hxq[q]] + rank[q];
      srec[slot] = recq[q];
      skey[slot] = (short)hyq[q];
      sbkt[slot] = (short)hxq[q];
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < cnt; t += RS_NT) {
    const int b = sbkt[t];
    const int64_t dst = (int64_t)gbase[b] + (t - lstart[b]);
    tmp_rec[dst] = srec[t];
    tmp_key[dst] = skey[t];
  }
}

// Chunk counts before each catalogue for two chunk sizes (one workgroup; any number of
// catalogues, RS_NT * 2 per round).
__global__ void __launch_bounds__(RS_NT)
catalogue_chunks_kernel(const int64_t* __restrict__ moff, int nmock, int ch_a, int* __restrict__ choff_a, int ch_b,
                        int* __restrict__ choff_b) {
  __shared__ int ca[2 * RS_NT], cb[2 * RS_NT];
  __shared__ int wsum[RS_NT / 64 + 1];
  int carry_a = 0, carry_b = 0;
  for (int m0 = 0; m0 < nmock; m0 += 2 * RS_NT) {
    const int n = min(2 * RS_NT, nmock - m0);
    for (int q = threadIdx.x; q < n; q += RS_NT) {
      const int64_t cnt = moff[m0 + q + 1] - moff[m0 + q];
      ca[q] = (int)((cnt + ch_a - 1) / ch_a);
      cb[q] = (int)((cnt + ch_b - 1) / ch_b);
    }
    __syncthreads();
    const int ta = block_excl_scan(ca, n, wsum);
    const int tb = block_excl_scan(cb, n, wsum);
    for (int q = threadIdx.x; q < n; q += RS_NT) {
      choff_a[m0 + q] = carry_a + ca[q];
      choff_b[m0 + q] = carry_b + cb[q];
    }
    carry_a += ta;
    carry_b += tb;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    choff_a[nmock] = carry_a;
    choff_b[nmock] = carry_b;
  }
}

// One workgroup per x bucket (bucket m * nx + x for x cell x of catalogue m): counting sort by
// y into ysorted (the order the paint reads); writes the row offsets and a bound on the sum of
// |w| of every row.  The bucket is moved in
// sub-chunks of RS_SCHUNK records: each is ordered by y in LDS and every y run is written
// contiguously after the records of that row already placed.
__global__ void __launch_bounds__(RS_NT)
rowsort_fine_kernel(const int* __restrict__ bucket_start, const float4* __restrict__ tmp_rec, const short* __restrict__ tmp_key, int ny,
                    int nx, int64_t nkeys, const unsigned* __restrict__ wmax_bits, float4* __restrict__ ysorted,
                    int* __restrict__ offsets, float* __restrict__ row_wabs) {
  __shared__ int cnt[RS_MAX_KEYS];     // whole bucket: counts, then the next free slot of each row
  __shared__ int lcnt[RS_MAX_KEYS];    // sub-chunk counts
  __shared__ int lstart[RS_MAX_KEYS];  // sub-chunk start of each row in LDS
  __shared__ unsigned long long qsum[RS_MAX_KEYS];
  __shared__ int wsum[RS_NT / 64 + 1];
  extern __shared__ float4 srec[];  // RS_SCHUNK records, then their y keys
  short* skey = reinterpret_cast<short*>(srec + RS_SCHUNK);
  const int hx = blockIdx.x;
  const int start = bucket_start[hx], end = bucket_start[hx + 1];
  for (int y = threadIdx.x; y < ny; y += RS_NT) {
    cnt[y] = 0;
    qsum[y] = 0ull;
  }
  __syncthreads();
  for (int i = start + threadIdx.x; i < end; i += RS_NT) atomicAdd(&cnt[tmp_key[i]], 1);
  __syncthreads();
  block_excl_scan(cnt, ny, wsum);
  for (int y = threadIdx.x; y < ny; y += RS_NT) {
    offsets[(int64_t)hx * ny + y] = start + cnt[y];
    cnt[y] += start;
  }
  if (hx == 0 && threadIdx.x == 0) offsets[nkeys] = bucket_start[gridDim.x];
  // Per-row bound on sum |w|, which sets the fixed-point scale of the paint.  It must not
  // depend on the order of the row's records (which the atomics leave unspecified), so it is
  // an integer sum: q_i = ceil(|w_i| / m * 2^24) with m the largest |w| of the catalogue, and
  // the bound is m * sum q_i / 2^24 >= sum |w|.
  const float m = __uint_as_float(wmax_bits[hx / nx]);  // of this bucket's catalogue
  const float qs = m > 0.f ? 16777216.f / m : 0.f;
  constexpr int PER = RS_SCHUNK / RS_NT;
  for (int sub = start; sub < end; sub += RS_SCHUNK) {
    const int n = min(RS_SCHUNK, end - sub);
    for (int y = threadIdx.x; y < ny; y += RS_NT) lcnt[y] = 0;
    __syncthreads();
    int rank[PER], key[PER];
    float4 rec[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int t = threadIdx.x + q * RS_NT;
      if (t < n) {
        key[q] = tmp_key[sub + t];
        rec[q] = load_rec_stream(tmp_rec + sub + t);
        rank[q] = atomicAdd(&lcnt[key[q]], 1);
        atomicAdd(&qsum[key[q]], (unsigned long long)ceilf(fabsf(rec[q].w) * qs));
      }
    }
    __syncthreads();
    for (int y = threadIdx.x; y < ny; y += RS_NT) lstart[y] = lcnt[y];
    __syncthreads();
    block_excl_scan(lstart, ny, wsum);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int t = threadIdx.x + q * RS_NT;
      if (t < n) {
        const int slot = lstart[key[q]] + rank[q];
        srec[slot] = rec[q];
        skey[slot] = (short)key[q];
      }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < n; t += RS_NT) {
      const int y = skey[t];
      ysorted[cnt[y] + (t - lstart[y])] = srec[t];
    }
    __syncthreads();
    for (int y = threadIdx.x; y < ny; y += RS_NT) cnt[y] += lcnt[y];
  }
  __syncthreads();
  for (int y = threadIdx.x; y < ny; y += RS_NT)
    row_wabs[(int64_t)hx * ny + y] = (float)((double)qsum[y] * (double)m * (1.0 / 16777216.0));
}

// Stencil of coordinate u (relative to cell base): returns the first cell,
// base + round/floor(u) + OMIN, and sets k[o] = K(|first + o - base - u|), o in [0, ORDER), in
// closed form in t = u - round/floor(u) (resamplers.py:63-69 evaluate K piecewise in |d|; the
// polynomials are the same).  base is relative to an even origin, so that an exact half rounds
// to the even mesh cell as round-half-even on the mesh coordinate does (only NGP, whose kernel
// is discontinuous there, depends on it).
template <int ORDER>
__device__ __forceinline__ int stencil_1d(float u, int base, float* k) {
  if constexpr (ORDER == 1) {
    float r = rintf(u);
    if (fabsf(u - r) == 0.5f && ((base + (int)r) & 1)) r = 2.f * u - r;
    k[0] = 1.f;
    return base + (int)r;
  } else if constexpr (ORDER == 2) {
    const float r = floorf(u), t = u - r;
    k[0] = 1.f - t;
    k[1] = t;
    return base + (int)r;
  } else if constexpr (ORDER == 3) {
    const float r = rintf(u), t = u - r;  // t in [-1/2, 1/2]
    const float a = 0.5f - t, b = 0.5f + t;
    k[0] = 0.5f * a * a;
    k[1] = 0.75f - t * t;
    k[2] = 0.5f * b * b;
    return base + (int)r - 1;
  } else {
    const float r = floorf(u), t = u - r, t1 = 1.f - t;  // t in [0, 1)
    const float t2 = t * t, t12 = t1 * t1;
    k[0] = t12 * t1 * (1.f / 6.f);
    k[1] = (4.f - 6.f * t2 + 3.f * t2 * t) * (1.f / 6.f);
    k[2] = (4.f - 6.f * t12 + 3.f * t12 * t1) * (1.f / 6.f);
    k[3] = t2 * t * (1.f / 6.f);
    return base + (int)r - 1;
  }
}

// LDS-space pointers: held in one 32-bit register, so an add to them stays one instruction and
// a constant offset folds into the LDS instruction.
typedef __attribute__((address_space(3))) int lds_i32;
typedef __attribute__((address_space(3))) unsigned char lds_u8;

// Fixed-point addend: round to nearest (halves up), in one instruction.
__device__ __forceinline__ int to_fixed(float v) {
  int r;
  asm("v_cvt_rpi_i32_f32 %0, %1" : "=v"(r) : "v"(v));
  return r;
}

// Stencil weights reordered per lane, k'[q] = k[(q + rot) mod ORDER], with the cell offsets
// reordered alike by the caller.  Lanes of one wave whose particles share stencil cells (a dense
// clump) then add to different cells in the same LDS atomic instruction instead of serialising
// on one address: lane l rotates z by l mod ORDER and y by (l / ORDER) mod ORDER (measured on
// lognormal mocks, TSC: 256^3 with 2.25e6 particles 0.194 -> 0.126 ms, 512^3 with 1.6e7
// particles 1.16 -> 1.09 ms).  The sums are unchanged: the adds are exact integers.
template <int ORDER>
__device__ __forceinline__ void rotate_stencil(float (&k)[ORDER], int rot) {
  float r[ORDER];
#pragma unroll
  for (int q = 0; q < ORDER; ++q) {
    float v = k[q];
#pragma unroll
    for (int t = 1; t < ORDER; ++t)
      if (rot == t) v = k[(q + t) % ORDER];
    r[q] = v;
  }
#pragma unroll
  for (int q = 0; q < ORDER; ++q) k[q] = r[q];
}

template <int ORDER>
struct LaneRotation {
  int zo[ORDER], yo[ORDER];  // cell offset of stencil slot q along z and y
  int zrot, yrot;
  __device__ __forceinline__ explicit LaneRotation(int tid) {
    const int lane = tid & 63;
    zrot = lane % ORDER;
    yrot = (lane / ORDER) % ORDER;
#pragma unroll
    for (int q = 0; q < ORDER; ++q) {
      zo[q] = (q + zrot) % ORDER;
      yo[q] = (q + yrot) % ORDER;
    }
  }
};

// A tighter bound on |cell| for catalogues with weights of both signs.  Their painted cells
// nearly cancel, so the row bound (sum of |w| over the home rows that reach a row, over the
// whole z extent) leaves most of the fixed-point range unused and the rounding of each addend
// becomes visible in the spectrum.  Here |w| is rounded up to 1/4096 of the catalogue's largest
// |w| and histogrammed by (home x row, home z cell) over every particle that reaches the block
// (all home y rows of the reach merged); summed over the homes that reach each target (x row,
// z cell), its largest value bounds every cell of the block.  Integer sums: the bound does not
// depend on the order of the particles.  `hist` is the zeroed row LDS with room for
// (RX + hi - lo) * nz ints; it is zeroed again on return.  All threads of the block call this.
template <int NT>
__device__ float tight_row_bound(int* hist, int hist_cap, int RX, int nz, int lo, int hi, int zfrac, float wmax,
                                 const float4* __restrict__ sorted, const int* sstart, const int* scount, const int* shome,
                                 const int* spre, int nseg, int total, int chunk, unsigned long long* sred) {
  const int span = hi - lo, sx = RX + span;
  const int tid = threadIdx.x;
  if (sx * nz > hist_cap || !(wmax > 0.f)) return 3.4e38f;  // no room or no weight: keep the row bound
  const float qs = 4096.f / wmax;
  for (int c = tid; c < total; c += NT) {
    int lo_i = 0, hi_i = nseg - 1;
    while (lo_i < hi_i) {
      const int mid = (lo_i + hi_i + 1) >> 1;
      if (spre[mid] <= c) lo_i = mid;
      else hi_i = mid - 1;
    }
    const int p0 = sstart[lo_i] + (c - spre[lo_i]) * chunk;
    const int p1 = min(p0 + chunk, sstart[lo_i] + scount[lo_i]);
    const int ix = (shome[lo_i] >> 16) + hi;  // home x row relative to the reach, in [0, sx)
    for (int p = p0; p < p1; ++p) {
      const float4 rec = sorted[p];
      const int zc = (int)(__float_as_uint(rec.z) >> zfrac);
      atomicAdd(&hist[ix * nz + zc], (int)ceilf(fabsf(rec.w) * qs));
    }
  }
  if (tid == 0) *sred = 0ull;
  __syncthreads();
  // target x row t is reached from ix in [t, t + span]; target z from home z in [z - hi, z - lo]
  unsigned long long m = 0ull;
  for (int i = tid; i < RX * nz; i += NT) {
    const int t = i / nz, z = i % nz;
    unsigned long long b = 0ull;
    for (int ix = t; ix <= t + span; ++ix)
      for (int d = lo; d <= hi; ++d) b += (unsigned)hist[ix * nz + ((z - d) % nz + nz) % nz];
    m = b > m ? b : m;
  }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    const unsigned long long o = __shfl_xor(m, d, 64);
    m = o > m ? o : m;
  }
  if ((tid & 63) == 0) atomicMax(sred, m);
  __syncthreads();
  const float bound = (float)((double)*sred * (double)wmax * (1.0 / 4096.0) * (1.0 + 1e-6));
  for (int i = tid; i < sx * nz; i += NT) hist[i] = 0;
  __syncthreads();
  return bound;
}

// First Stockham stage (radix first_radix(M), no twiddles) of the z transform of NROW rows:
// reads the painted fixed-point cells (row r, cell z at ints[r * RS + z]; complex element e is
// the pair z = 2e, 2e + 1), scales them to float and writes the stage output in the column
// layout (element l of row r at buf[l * RP + r]).  Every read precedes the barrier and every
// write follows it, as the two layouts share the LDS; the caller syncs before the call.
template <int M, int NROW, int RP, int RS, int NT>
__device__ __forceinline__ void pz_first_stage(const int* ints, float2* buf, float scale) {
  constexpr int R = first_radix(M);
  constexpr int NB = NROW * (M / R);
  constexpr int BPT = (NB + NT - 1) / NT;
  float2 v[BPT][R];
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    const int b = tid + q * NT;
    if (NB % NT == 0 || b < NB) {
      const int c = b % NROW, jj = b / NROW;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int2 p = *reinterpret_cast<const int2*>(ints + c * RS + 2 * (jj + r * (M / R)));
        v[q][r] = make_float2((float)p.x * scale, (float)p.y * scale);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    const int b = tid + q * NT;
    if (NB % NT == 0 || b < NB) {
      const int c = b % NROW, jj = b / NROW;
      dft<R, false>(v[q]);
#pragma unroll
      for (int r = 0; r < R; ++r) buf[(jj * R + r) * RP + c] = v[q][r];
    }
  }
  __syncthreads();
}

// One workgroup: RX x RY rows of shifted mesh j, painted in LDS, then z-FFT'd (r2c) in LDS.
template <int ORDER, int M, int RX, int RY>
__global__ void __launch_bounds__(PZ_NT)
paint_zfft_kernel(const float4* __restrict__ sorted, const int* __restrict__ offsets, const float* __restrict__ row_wabs, RowGeom rg,
                  float2* __restrict__ out, int64_t pitch, int64_t mesh_stride, const float2* __restrict__ tw) {
  constexpr int NZ = 2 * M;
  // Two LDS layouts share the dynamic LDS.  While painting, row r holds its NZ int32 cells at
  // words r RS + z (RS = NZ + 2: a stencil cell is one add from its row, and the first FFT stage
  // reads the 8-byte pairs of consecutive rows from different banks).  The z transform then works
  // on complex element l (reals z = 2l, 2l+1) of row r at lds2[l * RP + r], rows fastest with
  // one pad column, so the column FFT is bank-conflict free.
  constexpr int RP = RX * RY + 1;
  constexpr int RS = NZ + 2;
  constexpr int NROW = RX * RY;
  constexpr int OMIN = -((ORDER - 1) / 2);
  constexpr int CHUNK = 4;  // particles per work item, loaded together before any is painted
  extern __shared__ float2 lds2[];
  float* rows = reinterpret_cast<float*>(lds2);
  constexpr int SEG = pz_seg_capacity(RX, RY);
  __shared__ int sstart[SEG];
  __shared__ int scount[SEG];
  __shared__ int shome[SEG];  // home row relative to the block, (x << 16) | (y & 0xffff)
  __shared__ int spre[SEG + 2];
  __shared__ float swabs[SEG];
  __shared__ float sbound[1];
  __shared__ float tbound[NROW];
  __shared__ float2 stw[NZ];
  __shared__ int sregion[17];  // first segment of each chunk region, and the last segment
  __shared__ unsigned long long sredu;
  const int tid = threadIdx.x;
  load_twiddles<NZ, PZ_NT>(stw, tw);

  // XCD-aware work index: blocks sharing an x strip (hence most of their particles) and the
  // shifts of one block run on the same XCD, i.e. share an L2.
  const int id = blockIdx.x;
  const int xcd = id & 7;
  const int k = id >> 3;
  const int j = k % rg.ni;
  const int rest = k / rg.ni;
  const int by = rest % rg.nby;
  const int nq = (rg.nbx + 7) >> 3;
  const int q = rest / rg.nby;
  const int mk = q / nq;  // catalogue: its rows follow those of the catalogues before it
  const int bx = (q % nq) * 8 + xcd;
  if (bx >= rg.nbx) return;
  const int x0 = bx * RX, y0 = by * RY;
  const int64_t key0 = (int64_t)mk * rg.n[0] * rg.n[1];
  const float s = (float)j / (float)rg.ni;

  for (int i = tid; i < NROW * RS / 2; i += PZ_NT) lds2[i] = make_float2(0.f, 0.f);

  // neighbourhood rows: homes in [x0 - hi, x0 + RX - 1 - lo]; each row's particles are
  // contiguous.  Work items are chunks of up to CHUNK consecutive particles of one row.
  // rows reached from home row h at this shift: h + [lo, hi].  The stencil origin of x~ + s,
  // x~ in [h, h+1), is h or h+1 for s < 1/2 with round, h+1 or h+2 for s > 1/2; at s = 1/2
  // float rounding of x~ + s can give any of the three.  With floor it is h, or h and h+1 for
  // s > 0.
  const int lo = OMIN + ((ORDER % 2) && 2 * j > rg.ni ? 1 : 0);
  const int hi = (ORDER - 1 + OMIN) + ((ORDER % 2) ? (2 * j >= rg.ni ? 2 : 1) : (j > 0 ? 1 : 0));
  const int sx = RX + hi - lo, sy = RY + hi - lo;
  const int nseg = sx * sy;
  for (int i = tid; i < nseg; i += PZ_NT) {
    int ix = i / sy, iy = i % sy;
    int hx = pmod(x0 - hi + ix, rg.n[0]), hy = pmod(y0 - hi + iy, rg.n[1]);
    int64_t key = key0 + (int64_t)hx * rg.n[1] + hy;
    int st = offsets[key];
    int c = offsets[key + 1] - st;
    sstart[i] = st;
    scount[i] = c;
    shome[i] = ((ix - hi) * 65536) | ((iy - hi) & 0xffff);
    spre[i + 1] = (c + CHUNK - 1) / CHUNK;
    swabs[i] = row_wabs[key];
  }
  __syncthreads();
  if (tid < 64) {
    constexpr int PER = SEG / 64;
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
  // bound on |cell value| for the fixed-point accumulation: for each row of the block, the sum
  // of |w| over the home rows that can reach it; the block uses the largest
  if (tid < NROW) {
    const int rx = tid / RY, ry = tid % RY, span = hi - lo;
    float sum = 0.f;
    for (int a = 0; a <= span; ++a)
      for (int bb = 0; bb <= span; ++bb) sum += swabs[(rx + a) * sy + (ry + bb)];
    tbound[tid] = sum;
  }
  __syncthreads();
  // Work items are taken from G regions of the chunk list in turn (see the loop below); the
  // first segment of each region bounds the segment search of its chunks.
  constexpr int G = 16;
  const int total = spre[nseg];  // number of chunks
  const int Rg = (total + G - 1) / G;
  if (tid < 64) {
    float m = 0.f;
    for (int i = tid; i < NROW; i += 64) m = fmaxf(m, tbound[i]);
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) m = fmaxf(m, __shfl_xor(m, d, 64));
    if (tid == 0) sbound[0] = m;
  } else if (tid < 64 + G + 1) {
    const int g = tid - 64, c = g * Rg;
    int lo_i = 0, hi_i = nseg - 1;
    if (g == G || c >= total) {
      lo_i = hi_i;
    } else {
      while (lo_i < hi_i) {
        const int mid = (lo_i + hi_i + 1) >> 1;
        if (spre[mid] <= c) lo_i = mid;
        else hi_i = mid - 1;
      }
    }
    sregion[g] = lo_i;
  }
  __syncthreads();
  float bound = sbound[0];
  if (rg.negw[mk])
    bound = fminf(bound, tight_row_bound<PZ_NT>(reinterpret_cast<int*>(rows), NROW * RS, RX, NZ, lo, hi, rg.zfrac,
                                                 __uint_as_float(rg.wmax_bits[mk]), sorted, sstart, scount, shome, spre, nseg,
                                                 total, CHUNK, &sredu));
  // Fixed-point accumulation: integer LDS atomics are exact, order independent (the painted
  // rows are bitwise reproducible) and much faster than float LDS atomics on this target.
  // |sum| <= 0.99 * 2^31 (each K <= 1), and each addend is rounded to 1/scale.
  const float scale = bound > 0.f ? 2126008811.f / bound : 1.f;
  const int* irows = reinterpret_cast<const int*>(rows);
  lds_u8* const lrows = (lds_u8*)(rows);  // the rows as an LDS-space byte pointer
  // Work items are taken from G regions of the chunk list in turn (lane l of a wave works in
  // region l mod G), so the lanes of one wave hold chunks of different home rows: the particles
  // of a dense clump, which share stencil cells, then go to different instructions instead of
  // serialising one LDS atomic over many lanes at the same address (measured on the 512^3
  // lognormal mock: 1.41 -> 1.19 ms; a coprime stride over single chunks, which also breaks the
  // coalescing of the record loads, gave 1.22 ms).
  const LaneRotation<ORDER> rot(tid);
  for (int ci = tid; ci < G * Rg; ci += PZ_NT) {
    const int g = ci % G, c = g * Rg + ci / G;
    if (c >= total) continue;
    int lo_i = sregion[g], hi_i = sregion[g + 1];  // the segment of chunk c lies in this range
    while (lo_i < hi_i) {
      int mid = (lo_i + hi_i + 1) >> 1;
      if (spre[mid] <= c) lo_i = mid;
      else hi_i = mid - 1;
    }
    const int seg = lo_i;
    const int p0 = sstart[seg] + (c - spre[seg]) * CHUNK;
    const int plast = sstart[seg] + scount[seg] - 1;  // >= p0: a chunk holds at least one record
    const int cnt = min(CHUNK, plast + 1 - p0);
    const int home = shome[seg];
    const int hxr = home >> 16, hyr = (int)(short)(home & 0xffff);  // home row relative to the block
    float4 recs[CHUNK];
#pragma unroll
    for (int q = 0; q < CHUNK; ++q) recs[q] = sorted[min(p0 + q, plast)];  // no branch per load
#pragma unroll
    for (int q = 0; q < CHUNK; ++q) {
      if (q >= cnt) break;
      const float4 rec = recs[q];
      int zc;
      float fz;
      z_of_record(__float_as_uint(rec.z), rg.zfrac, zc, fz);
      float kx[ORDER], ky[ORDER], kz[ORDER];
      const int cx0 = stencil_1d<ORDER>(rec.x + s, hxr, kx);
      const int cy0 = stencil_1d<ORDER>(rec.y + s, hyr, ky);
      const int cz0 = stencil_1d<ORDER>(fz + s, zc, kz);
      if constexpr (PZ_YROT) rotate_stencil<ORDER>(ky, rot.yrot);
      rotate_stencil<ORDER>(kz, rot.zrot);
      // Cell (cx0 + ox, cy0 + yo, z) is word ((cx0 + ox) RY + cy0 + yo) RS + z: one base per
      // (z slot, y slot), computed before the x loop, and the x part an immediate offset of the
      // LDS atomic.  The dynamic LDS follows the static tables, so a base below the first row
      // stays >= 0.
      const int rowb = 4 * RS * (cx0 * RY + cy0);
      int zb[ORDER];  // byte offsets into the rows
#pragma unroll
      for (int oz = 0; oz < ORDER; ++oz) zb[oz] = rowb + 4 * ((cz0 + rot.zo[oz]) & (NZ - 1));
      bool inx[ORDER];
#pragma unroll
      for (int ox = 0; ox < ORDER; ++ox) inx[ox] = (unsigned)(cx0 + ox) < (unsigned)RX;
      const float ws = rec.w * scale;
      float wx[ORDER];
#pragma unroll
      for (int ox = 0; ox < ORDER; ++ox) wx[ox] = ws * kx[ox];
#pragma unroll
      for (int oy = 0; oy < ORDER; ++oy) {
        const int yo = PZ_YROT ? rot.yo[oy] : oy;
        if ((unsigned)(cy0 + yo) >= (unsigned)RY) continue;
        lds_i32* base[ORDER];
#pragma unroll
        for (int oz = 0; oz < ORDER; ++oz) {
          base[oz] = reinterpret_cast<lds_i32*>(lrows + zb[oz] + (PZ_YROT ? 4 * RS * yo : 0));
          asm volatile("" : "+v"(base[oz]));  // one add per (z, y) slot; the x slots use immediates
        }
#pragma unroll
        for (int ox = 0; ox < ORDER; ++ox) {
          if (!inx[ox]) continue;
          const float kxy = wx[ox] * ky[oy];
#pragma unroll
          for (int oz = 0; oz < ORDER; ++oz)
            __hip_atomic_fetch_add(base[oz] + RS * RY * ox + (PZ_YROT ? 0 : RS * oy), to_fixed(kxy * kz[oz]), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      }
    }
  }
  __syncthreads();

  // z-axis r2c of the NROW rows: the first Stockham stage converts the fixed-point cells (times
  // 1/(2 scale): the post-processing below then needs no factor 1/2) and moves them into the
  // column layout; the other stages run in place.
  pz_first_stage<M, NROW, RP, RS, PZ_NT>(irows, lds2, 0.5f / scale);
  fft_stages<M, NROW, RP, 1, true, PZ_NT, false, first_radix(M), NZ>(lds2, stw);
  float2* dst = out + ((int64_t)mk * rg.ni + j) * mesh_stride;
  // Outputs k and M - k of the N = 2M point real transform for k in [1, M/2] (k = M/2 writes
  // one), from Z = FFT_M(x[2m] + i x[2m+1]) / 2; k = 0 (outputs 0 and M) apart.  Offsets within
  // the mesh fit 32 bits (< 2^31 complex elements).  Streaming stores: the rows are read back
  // only by the y pass (512^3 TSCx3: paint 1.04 -> 0.98 ms, and the y pass 0.64 -> 0.615 ms).
  constexpr int HM = M / 2;
  for (int idx = tid; idx < NROW * HM; idx += PZ_NT) {
    const int r = idx / HM, k = idx % HM + 1;
    float2* o = dst + (unsigned)((x0 + r / RY) * rg.n[1] + y0 + r % RY) * (unsigned)pitch;
    const float2 zk = lds2[k * RP + r];
    const float2 zmk = cconj(lds2[(M - k) * RP + r]);
    const float2 E = cadd(zk, zmk), D = csub(zk, zmk);
    const float2 t = cmul(stw[k], make_float2(D.y, -D.x));  // exp(-2 pi i k / N) (-i D)
    store_stream(o + k, cadd(E, t));
    if (k != HM) store_stream(o + M - k, cconj(csub(E, t)));
  }
  if (tid < NROW) {
    const int r = tid;
    float2* o = dst + (unsigned)((x0 + r / RY) * rg.n[1] + y0 + r % RY) * (unsigned)pitch;
    const float2 z0 = lds2[r];
    store_stream(o, make_float2(2.f * (z0.x + z0.y), 0.f));
    store_stream(o + M, make_float2(2.f * (z0.x - z0.y), 0.f));
  }
}

// Real-space paint from the row sort (any mesh shape up to 2048 per axis): one workgroup per
// (block of RX x RY rows, shift); the rows (full z extent, int32 fixed point) accumulate in
// LDS like the fused paint, then are written to the real mesh.  Row blocks at the upper mesh
// edge may be partial.
constexpr int PR_NT = 512;

template <int ORDER, int RX, int RY>
__global__ void __launch_bounds__(PR_NT)
paint_rows_kernel(const float4* __restrict__ sorted, const int* __restrict__ offsets, const float* __restrict__ row_wabs,
                  RowGeom rg, float* __restrict__ out, int64_t mesh_stride) {
  constexpr int NROW = RX * RY;
  constexpr int CHUNK = 4;
  constexpr int OMIN = -((ORDER - 1) / 2);
  constexpr int SEG = pz_seg_capacity(RX, RY);
  extern __shared__ int irows[];  // row r, cell z at irows[r * rs + z]
  __shared__ int sstart[SEG];
  __shared__ int scount[SEG];
  __shared__ int shome[SEG];  // home row relative to the block, (x << 16) | (y & 0xffff)
  __shared__ int spre[SEG + 2];
  __shared__ float swabs[SEG];
  __shared__ float tbound[NROW];
  __shared__ float sbound[1];
  __shared__ unsigned long long sredu;
  const int tid = threadIdx.x;
  const int n2 = rg.n[2], rs = n2 + 1;
  const int id = blockIdx.x;
  const int j = id % rg.ni;
  const int rest = id / rg.ni;
  const int by = rest % rg.nby, bx = rest / rg.nby;
  const int x0 = bx * RX, y0 = by * RY;
  const float s = (float)j / (float)rg.ni;
  for (int i = tid; i < NROW * rs; i += PR_NT) irows[i] = 0;
  const int lo = OMIN + ((ORDER % 2) && 2 * j > rg.ni ? 1 : 0);
  const int hi = (ORDER - 1 + OMIN) + ((ORDER % 2) ? (2 * j >= rg.ni ? 2 : 1) : (j > 0 ? 1 : 0));
  const int sx = RX + hi - lo, sy = RY + hi - lo;
  const int nseg = sx * sy;
  for (int i = tid; i < nseg; i += PR_NT) {
    const int ix = i / sy, iy = i % sy;
    const int hx = pmod(x0 - hi + ix, rg.n[0]), hy = pmod(y0 - hi + iy, rg.n[1]);
    const int64_t key = (int64_t)hx * rg.n[1] + hy;
    const int st = offsets[key];
    const int c = offsets[key + 1] - st;
    sstart[i] = st;
    scount[i] = c;
    shome[i] = ((ix - hi) * 65536) | ((iy - hi) & 0xffff);
    spre[i + 1] = (c + CHUNK - 1) / CHUNK;
    swabs[i] = row_wabs[key];
  }
  __syncthreads();
  if (tid < 64) {
    constexpr int PER = SEG / 64;
    int local[PER];
    int sum = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int i = tid * PER + q;
      local[q] = i < nseg ? spre[i + 1] : 0;
      sum += local[q];
    }
    int incl = sum;
#pragma unroll
    for (int d = 1; d < 64; d *= 2) {
      const int v = __shfl_up(incl, d, 64);
      if (tid >= d) incl += v;
    }
    int run = incl - sum;
    if (tid == 0) spre[0] = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int i = tid * PER + q;
      run += local[q];
      if (i < nseg) spre[i + 1] = run;
    }
  }
  if (tid < NROW) {
    const int rx = tid / RY, ry = tid % RY, span = hi - lo;
    float sum = 0.f;
    for (int a = 0; a <= span; ++a)
      for (int bb = 0; bb <= span; ++bb) sum += swabs[(rx + a) * sy + (ry + bb)];
    tbound[tid] = sum;
  }
  __syncthreads();
  if (tid < 64) {
    float m = 0.f;
    for (int i = tid; i < NROW; i += 64) m = fmaxf(m, tbound[i]);
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) m = fmaxf(m, __shfl_xor(m, d, 64));
    if (tid == 0) sbound[0] = m;
  }
  __syncthreads();
  const int total = spre[nseg];
  float bound = sbound[0];
  if (rg.negw[0])
    bound = fminf(bound, tight_row_bound<PR_NT>(irows, NROW * rs, RX, n2, lo, hi, rg.zfrac, __uint_as_float(rg.wmax_bits[0]),
                                                sorted, sstart, scount, shome, spre, nseg, total, CHUNK, &sredu));
  const float scale = bound > 0.f ? 2126008811.f / bound : 1.f;
  const int rxmax = min(RX, rg.n[0] - x0), rymax = min(RY, rg.n[1] - y0);  // partial blocks at the edge
  constexpr int G = 16;  // chunks taken from G regions in turn, as in the fused paint
  const int Rg = (total + G - 1) / G;
  const LaneRotation<ORDER> rot(tid);
  for (int ci = tid; ci < G * Rg; ci += PR_NT) {
    const int c = (ci % G) * Rg + ci / G;
    if (c >= total) continue;
    int lo_i = 0, hi_i = nseg - 1;
    while (lo_i < hi_i) {
      const int mid = (lo_i + hi_i + 1) >> 1;
      if (spre[mid] <= c) lo_i = mid;
      else hi_i = mid - 1;
    }
    const int seg = lo_i;
    const int p0 = sstart[seg] + (c - spre[seg]) * CHUNK;
    const int cnt = min(CHUNK, sstart[seg] + scount[seg] - p0);
    const int home = shome[seg];
    const int hxr = home >> 16, hyr = (int)(short)(home & 0xffff);  // home row relative to the block
    float4 recs[CHUNK];
#pragma unroll
    for (int q = 0; q < CHUNK; ++q) recs[q] = q < cnt ? sorted[p0 + q] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int q = 0; q < CHUNK; ++q) {
      if (q >= cnt) break;
      const float4 rec = recs[q];
      int zc;
      float fz;
      z_of_record(__float_as_uint(rec.z), rg.zfrac, zc, fz);
      float kx[ORDER], ky[ORDER], kz[ORDER];
      const int cx0 = stencil_1d<ORDER>(rec.x + s, hxr, kx);
      const int cy0 = stencil_1d<ORDER>(rec.y + s, hyr, ky);
      const int cz0 = stencil_1d<ORDER>(fz + s, zc, kz);
      rotate_stencil<ORDER>(ky, rot.yrot);
      rotate_stencil<ORDER>(kz, rot.zrot);
      int zoff[ORDER];
#pragma unroll
      for (int oz = 0; oz < ORDER; ++oz) {
        const int z = cz0 + rot.zo[oz];
        zoff[oz] = z < 0 ? z + n2 : (z >= n2 ? z - n2 : z);
      }
      const float ws = rec.w * scale;
#pragma unroll
      for (int ox = 0; ox < ORDER; ++ox) {
        const int cx = cx0 + ox;
        if (!(cx >= 0 && cx < rxmax)) continue;
        const float wx = ws * kx[ox];
#pragma unroll
        for (int oy = 0; oy < ORDER; ++oy) {
          const int cy = cy0 + rot.yo[oy];
          if (!(cy >= 0 && cy < rymax)) continue;
          int* row = irows + (cx * RY + cy) * rs;
          const float kxy = wx * ky[oy];
#pragma unroll
          for (int oz = 0; oz < ORDER; ++oz) atomicAdd(row + zoff[oz], to_fixed(kxy * kz[oz]));
        }
      }
    }
  }
  __syncthreads();
  const float inv = 1.f / scale;
  float* dst = out + j * mesh_stride;
  for (int r = 0; r < NROW; ++r) {
    const int rx = r / RY, ry = r % RY;
    if (rx >= rxmax || ry >= rymax) continue;
    float* o = dst + ((int64_t)(x0 + rx) * rg.n[1] + (y0 + ry)) * n2;
    for (int z = tid; z < n2; z += PR_NT) o[z] = (float)irows[r * rs + z] * inv;
  }
}

// ------------------------------------------------------------------ host side

// Row-block shapes (RX, RY) tried per z half-length M; the first is the default.  Each fits
// M (RX RY + 1) complex in at most 140 KiB of LDS; the half-size shapes let two blocks share
// a CU.
struct PzShape {
  int rx, ry;
};
static const PzShape kPz1024[] = {{4, 4}, {4, 2}};
static const PzShape kPz512[] = {{8, 4}, {4, 4}};
static const PzShape kPz256[] = {{8, 4}, {8, 8}, {4, 4}};
static const PzShape kPz128[] = {{8, 8}, {16, 8}, {8, 4}};
static const PzShape kPz64[] = {{16, 8}, {16, 16}, {8, 8}};
static const PzShape kPzSmall[] = {{16, 16}, {16, 8}, {8, 8}};

static const PzShape* pz_candidates(int M, int& count) {
  switch (M) {
    case 1024: count = 2; return kPz1024;
    case 512: count = 2; return kPz512;
    case 256: count = 3; return kPz256;
    case 128: count = 3; return kPz128;
    case 64: count = 3; return kPz64;
    case 32: case 16: case 8: count = 3; return kPzSmall;
    default: count = 0; return nullptr;
  }
}

static void reach_of(int order, int ni, int& lo, int& hi) {
  int omin = -((order - 1) / 2), omax = order - 1 - (order - 1) / 2;
  int i0max = (order % 2) ? (ni > 1 ? 2 : 1) : (ni > 1 ? 1 : 0);
  lo = omin;
  hi = i0max + omax;  // offsets are stored strictly below 1, so no margin is needed
}

static bool pz_shape_ok(const int32_t n[3], const PzShape& sh, int order, int ni) {
  if (n[0] % sh.rx || n[1] % sh.ry) return false;
  int lo, hi;
  reach_of(order, ni, lo, hi);
  // neighbourhood rows must not alias through the periodic wrap, and fit the segment table
  if (sh.rx + hi - lo > n[0] || sh.ry + hi - lo > n[1]) return false;
  return (sh.rx + hi - lo) * (sh.ry + hi - lo) <= pz_seg_capacity(sh.rx, sh.ry);
}

bool paint_zfft_supported(const int32_t n[3], int64_t np, int order, int ni) {
  if (order < 1 || order > 4 || ni < 1 || np >= (int64_t)INT32_MAX) return false;
  if (!is_pow2(n[2])) return false;
  int count = 0;
  const PzShape* c = pz_candidates(n[2] / 2, count);
  for (int q = 0; q < count; ++q)
    if (pz_shape_ok(n, c[q], order, ni)) return true;
  return false;
}

size_t paint_zfft_workspace_bytes(const int32_t n[3], int64_t np, int nmock) {
  const int64_t nb = (int64_t)nmock * n[0];  // x buckets
  const int64_t nkeys = nb * n[1];
  size_t npp = (size_t)(np > 0 ? np : 1);
  size_t bytes = align_up((size_t)(nkeys + 1) * 4, 256);          // row offsets
  bytes += 2 * align_up((size_t)(nb + 2) * 4, 256);               // bucket start / cursor
  bytes += 2 * align_up((size_t)nmock * 4, 256);                  // max |w| and a negative-weight flag per catalogue
  bytes += 2 * align_up((size_t)(nmock + 1) * 4, 256);            // chunk offsets of the catalogues
  bytes += scan_workspace_bytes(nb + 1);
  bytes += align_up(npp * 16, 256) * 2 + align_up(npp * 2, 256);  // tmp records, sorted records, tmp keys
  bytes += align_up((size_t)nkeys * 4, 256);                      // per-row sum of |w|
  return bytes + 256;
}

template <int ORDER, int M, int RX, int RY>
static int launch_pz(hipStream_t s, const float4* sorted, const int* offsets, const float* row_wabs, const RowGeom& rg, float2* out,
                     int64_t pitch, int64_t mesh_stride) {
  // the larger of the two layouts: complex columns M (RX RY + 1), painted rows RX RY (2M + 2) int32
  constexpr size_t lds_fft = (size_t)M * (RX * RY + 1) * sizeof(float2);
  constexpr size_t lds_rows = (size_t)RX * RY * (2 * M + 2) * sizeof(int);
  constexpr size_t lds = lds_fft > lds_rows ? lds_fft : lds_rows;
  static_assert(lds <= PZ_LDS_BUDGET, "LDS budget");
  auto kern = paint_zfft_kernel<ORDER, M, RX, RY>;
  if (lds > 65536) MP_CHECK_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  // ids are laid out as (((m * nq + q) * nby + by) * ni + j) * 8 + xcd with bx = q * 8 + xcd
  // for catalogue m
  int64_t nq = ceil_div(rg.nbx, 8);
  unsigned grid = (unsigned)(rg.nmock * nq * rg.nby * rg.ni * 8);
  const float2* tw = twiddle_device_ptr();
  if (!tw) return set_error(MP_ERR_HIP, "paint_zfft: no twiddle table");
  MP_LAUNCH("paint_zfft", kern, dim3(grid), dim3(PZ_NT), lds, s, sorted, offsets, row_wabs, rg, out, pitch, mesh_stride, tw);
  return MP_OK;
}

template <int ORDER, int M>
static int launch_pz_shape(const PzShape& sh, hipStream_t s, const float4* sorted, const int* offsets, const float* row_wabs,
                           const RowGeom& rg, float2* out, int64_t pitch, int64_t mesh_stride) {
  constexpr bool small = M <= 32;
#define MP_PZ(RX_, RY_)                                                                                    \
  if (sh.rx == RX_ && sh.ry == RY_) return launch_pz<ORDER, M, RX_, RY_>(s, sorted, offsets, row_wabs, rg, out, pitch, mesh_stride);
  if constexpr (M == 1024) { MP_PZ(4, 4) MP_PZ(4, 2) }
  if constexpr (M == 512) { MP_PZ(8, 4) MP_PZ(4, 4) }
  if constexpr (M == 256) { MP_PZ(8, 4) MP_PZ(8, 8) MP_PZ(4, 4) }
  if constexpr (M == 128) { MP_PZ(8, 8) MP_PZ(16, 8) MP_PZ(8, 4) }
  if constexpr (M == 64) { MP_PZ(16, 8) MP_PZ(16, 16) MP_PZ(8, 8) }
  if constexpr (small) { MP_PZ(16, 16) MP_PZ(16, 8) MP_PZ(8, 8) }
#undef MP_PZ
  return set_error(MP_ERR_UNSUPPORTED, "paint_zfft: no %dx%d row blocks for Nz/2 = %d", sh.rx, sh.ry, M);
}

template <int ORDER>
static int dispatch_pz(int M, const PzShape& sh, hipStream_t s, const float4* sorted, const int* offsets, const float* row_wabs,
                       const RowGeom& rg, float2* out, int64_t pitch, int64_t mesh_stride) {
  switch (M) {
    case 1024: return launch_pz_shape<ORDER, 1024>(sh, s, sorted, offsets, row_wabs, rg, out, pitch, mesh_stride);
    case 512: return launch_pz_shape<ORDER, 512>(sh, s, sorted, offsets, row_wabs, rg, out, pitch, mesh_stride);
    case 256: return launch_pz_shape<ORDER, 256>(sh, s, sorted, offsets, row_wabs, rg, out, pitch, mesh_stride);
    case 128: return launch_pz_shape<ORDER, 128>(sh, s, sorted, offsets, row_wabs, rg, out, pitch, mesh_stride);
    case 64: return launch_pz_shape<ORDER, 64>(sh, s, sorted, offsets, row_wabs, rg, out, pitch, mesh_stride);
    case 32: return launch_pz_shape<ORDER, 32>(sh, s, sorted, offsets, row_wabs, rg, out, pitch, mesh_stride);
    case 16: return launch_pz_shape<ORDER, 16>(sh, s, sorted, offsets, row_wabs, rg, out, pitch, mesh_stride);
    case 8: return launch_pz_shape<ORDER, 8>(sh, s, sorted, offsets, row_wabs, rg, out, pitch, mesh_stride);
    default: return set_error(MP_ERR_UNSUPPORTED, "paint_zfft: Nz/2 = %d", M);
  }
}

static int launch_pz_any(int order, int M, const PzShape& sh, hipStream_t s, const float4* sorted, const int* offsets,
                         const float* row_wabs, RowGeom rg, float2* out, int64_t pitch, int64_t mesh_stride) {
  rg.rx = sh.rx;
  rg.ry = sh.ry;
  rg.nbx = rg.n[0] / sh.rx;
  rg.nby = rg.n[1] / sh.ry;
  switch (order) {
    case 1: return dispatch_pz<1>(M, sh, s, sorted, offsets, row_wabs, rg, out, pitch, mesh_stride);
    case 2: return dispatch_pz<2>(M, sh, s, sorted, offsets, row_wabs, rg, out, pitch, mesh_stride);
    case 3: return dispatch_pz<3>(M, sh, s, sorted, offsets, row_wabs, rg, out, pitch, mesh_stride);
    default: return dispatch_pz<4>(M, sh, s, sorted, offsets, row_wabs, rg, out, pitch, mesh_stride);
  }
}

// ------------------------------------------------------------------ autotuner
// On the first paint of a (meshsize, order, interlacing, batch) case, every row-block shape
// that fits is timed on the actual sorted particles (tune.h) and the fastest is kept.  Small
// catalogues take the first shape that fits.

constexpr int64_t PZ_TUNE_MIN_NP = 1 << 18;

static int pz_choose(hipStream_t s, int order, int M, const int32_t n[3], int ni, int64_t np, const float4* sorted,
                     const int* offsets, const float* row_wabs, const RowGeom& rg, float2* out, int64_t pitch, int64_t mesh_stride,
                     PzShape& chosen) {
  int count = 0;
  const PzShape* c = pz_candidates(M, count);
  bool ok[TUNE_MAX_CAND] = {};
  int first = -1, nok = 0;
  for (int q = 0; q < count && q < TUNE_MAX_CAND; ++q)
    if ((ok[q] = pz_shape_ok(n, c[q], order, ni))) {
      if (first < 0) first = q;
      ++nok;
    }
  if (first < 0) return set_error(MP_ERR_UNSUPPORTED, "paint_zfft: no row-block shape fits (%d, %d, %d)", n[0], n[1], n[2]);
  chosen = c[first];
  const int forced = tune_forced(TK_PAINT_ROWS);
  if (forced >= 0) {  // tests of every shape
    if (forced < count && ok[forced]) chosen = c[forced];
    return MP_OK;
  }
  char key[160];
  snprintf(key, sizeof(key), "paint_rows n=%dx%dx%d order=%d interlacing=%d%s", n[0], n[1], n[2], order, ni,
           rg.nmock > 1 ? " batch" : "");
  const int q0 = tune_get(key);
  if (q0 >= 0) {
    if (q0 < count && ok[q0]) chosen = c[q0];
    return MP_OK;
  }
  if (nok < 2 || np < PZ_TUNE_MIN_NP || !tune_measure_ok(s)) return MP_OK;
  float ms[TUNE_MAX_CAND];
  const int best = tune_time(s, count, ok, [&](int q) {
    return launch_pz_any(order, M, c[q], s, sorted, offsets, row_wabs, rg, out, pitch, mesh_stride);
  }, ms);
  if (best < 0) return MP_ERR_HIP;
  chosen = c[best];
  tune_put(key, TK_PAINT_ROWS, best, ms, count);
  return MP_OK;
}

// Scatter workgroup sizes the autotuner tries besides RS_SCHUNK: smaller LDS staging lets two
// workgroups share a CU.
constexpr int RS_SCHUNK_ALT[2] = {3072, 4096};

template <typename T, int SCH>
static int launch_scatter(hipStream_t s, const T* pos, const float* w, int64_t np, const Catalogues& cb, int extra, const GridMap& g,
                          int* bcursor, float4* tmp_rec, short* tmp_key) {
  constexpr size_t slds = (size_t)SCH * (sizeof(float4) + 2 * sizeof(short));
  auto skern = rowsort_scatter_kernel<T, SCH>;
  MP_CHECK_HIP(hipFuncSetAttribute((const void*)skern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)slds));
  MP_LAUNCH("paint_rowsort_scatter", skern, dim3((unsigned)(ceil_div(np, SCH) + extra)), dim3(RS_NT), slds, s, pos, w, cb, g,
            zfrac_of(g.n[2]), bcursor, tmp_rec, tmp_key);
  return MP_OK;
}

// The row sort shared by the fused paint and the real-space row paint: records grouped by home
// (x, y) row, row offsets, and the per-row bound on sum |w|.
struct RowSorted {
  const float4* sorted;
  const int* offsets;
  const float* row_wabs;
  const unsigned* wmax_bits;
  const int* negw;
};

// moff: nullptr for one catalogue of np particles, else the device offsets of nmock catalogues
// stacked in pos/w (np in total), whose rows then follow each other: the row of home cell
// (x, y) of catalogue m has key (m n0 + x) n1 + y.
template <typename T>
static int row_sort(hipStream_t s, const T* pos, const float* w, int64_t np, const GridMap& g, void* ws, size_t ws_bytes,
                    RowSorted& out, const int64_t* moff = nullptr, int nmock = 1) {
  const int32_t n[3] = {g.n[0], g.n[1], g.n[2]};
  MP_REQUIRE(n[0] <= RS_MAX_KEYS && n[1] <= RS_MAX_KEYS, "row sort: meshsize beyond %d", RS_MAX_KEYS);
  MP_REQUIRE(nmock >= 1 && (nmock == 1 || moff), "row sort: %d catalogues without offsets", nmock);
  MP_REQUIRE(np < (int64_t)INT32_MAX, "row sort: %lld particles, the row offsets are 32-bit", (long long)np);
  const int64_t nb = (int64_t)nmock * n[0];
  const int64_t nkeys = nb * n[1];
  Workspace wsp(ws, ws_bytes);
  int* offsets = wsp.take<int>(nkeys + 1);
  int* bstart = wsp.take<int>(nb + 1);
  int* bcursor = wsp.take<int>(nb + 2);
  unsigned* wmax = wsp.take<unsigned>(nmock);
  int* negw = wsp.take<int>(nmock);
  int* choff_a = wsp.take<int>(nmock + 1);
  int* choff_b = wsp.take<int>(nmock + 1);
  int* scan_tmp = (int*)wsp.take<char>(scan_workspace_bytes(nb + 1));
  float4* tmp_rec = wsp.take<float4>(np > 0 ? np : 1);
  float4* sorted = wsp.take<float4>(np > 0 ? np : 1);
  short* tmp_key = wsp.take<short>(np > 0 ? np : 1);
  float* row_wabs = wsp.take<float>(nkeys);
  if (!offsets || !bstart || !bcursor || !wmax || !negw || !choff_a || !choff_b || !scan_tmp || !tmp_rec || !sorted || !tmp_key || !row_wabs)
    return set_error(MP_ERR_WORKSPACE, "row sort: workspace of %zu bytes < %zu", ws_bytes, paint_zfft_workspace_bytes(n, np, nmock));
  int rc = fill_i32_dev(s, bstart, nb + 1, 0);
  if (rc) return rc;
  rc = fill_i32_dev(s, reinterpret_cast<int*>(wmax), nmock, 0);
  if (rc) return rc;
  rc = fill_i32_dev(s, negw, nmock, 0);
  if (rc) return rc;
  Catalogues ca{moff, choff_a, nmock, np}, cb{moff, choff_b, nmock, np};
  if (moff) MP_LAUNCH("paint_catalogue_chunks", catalogue_chunks_kernel, dim3(1), dim3(RS_NT), 0, s, moff, nmock, RS_CHUNK, choff_a,
                      RS_SCHUNK, choff_b);
  // at most nmock partial chunks beyond np / CH: the extra blocks find no chunk and return
  const int extra = moff ? nmock : 0;
  if (np > 0)
    MP_LAUNCH("paint_rowsort_hist", rowsort_hist_kernel<T>, dim3((unsigned)(ceil_div(np, RS_CHUNK) + extra)), dim3(RS_NT), 0, s,
              pos, w, ca, g, bstart, wmax, negw);
  rc = exclusive_scan_i32(s, bstart, nb + 1, scan_tmp, bcursor);
  if (rc) return rc;
  if (np > 0) {
    auto run = [&](int q) -> int {
      switch (q) {
        case 1: return launch_scatter<T, RS_SCHUNK_ALT[0]>(s, pos, w, np, cb, extra, g, bcursor, tmp_rec, tmp_key);
        case 2: return launch_scatter<T, RS_SCHUNK_ALT[1]>(s, pos, w, np, cb, extra, g, bcursor, tmp_rec, tmp_key);
        default: return launch_scatter<T, RS_SCHUNK>(s, pos, w, np, cb, extra, g, bcursor, tmp_rec, tmp_key);
      }
    };
    // Particles per scatter workgroup: tuned per mesh for single catalogues (a batch's chunk
    // offsets are laid out for the default).  Each timing run restarts the bucket cursors.
    int q = 0;
    if (!moff) {
      const int forced = tune_forced(TK_SORT_CHUNK);
      char key[128];
      snprintf(key, sizeof(key), "sort_chunk nx=%d ny=%d positions=%s", n[0], n[1], sizeof(T) == 8 ? "f64" : "f32");
      if (forced >= 0) {
        q = forced < 3 ? forced : 0;
      } else if ((q = tune_get(key)) < 0) {
        q = 0;
        if (np >= PZ_TUNE_MIN_NP && tune_measure_ok(s)) {
          const bool ok[3] = {true, true, true};
          float ms[3];
          q = tune_time(s, 3, ok, [&](int c) -> int {
            if (hipMemcpyAsync(bcursor, bstart, (size_t)(nb + 1) * sizeof(int), hipMemcpyDeviceToDevice, s) != hipSuccess)
              return set_error(MP_ERR_HIP, "row sort tuning: cursor reset");
            return run(c);
          }, ms);
          if (q < 0) return MP_ERR_HIP;
          tune_put(key, TK_SORT_CHUNK, q, ms, 3);
          MP_CHECK_HIP(hipMemcpyAsync(bcursor, bstart, (size_t)(nb + 1) * sizeof(int), hipMemcpyDeviceToDevice, s));
        }
      }
    }
    rc = run(q);
    if (rc) return rc;
  }
  constexpr size_t flds = (size_t)RS_SCHUNK * (sizeof(float4) + sizeof(short));
  MP_CHECK_HIP(hipFuncSetAttribute((const void*)rowsort_fine_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)flds));
  MP_LAUNCH("paint_rowsort_fine", rowsort_fine_kernel, dim3((unsigned)nb), dim3(RS_NT), flds, s, (const int*)bstart, tmp_rec,
            (const short*)tmp_key, n[1], n[0], nkeys, (const unsigned*)wmax, sorted, offsets, row_wabs);
  out.sorted = sorted;
  out.offsets = offsets;
  out.row_wabs = row_wabs;
  out.wmax_bits = wmax;
  out.negw = negw;
  return MP_OK;
}

template <typename T>
static int paint_zfft_typed(hipStream_t s, const T* pos, const float* w, int64_t np, const GridMap& g, int order, int ni,
                            float2* out, int64_t pitch, int64_t mesh_stride, void* ws, size_t ws_bytes, const int64_t* moff,
                            int nmock) {
  const int32_t n[3] = {g.n[0], g.n[1], g.n[2]};
  RowSorted rs;
  int rc = row_sort<T>(s, pos, w, np, g, ws, ws_bytes, rs, moff, nmock);
  if (rc) return rc;
  RowGeom rg;
  for (int a = 0; a < 3; ++a) rg.n[a] = n[a];
  reach_of(order, ni, rg.lo, rg.hi);
  rg.ni = ni;
  rg.zfrac = zfrac_of(n[2]);
  rg.nmock = nmock;
  rg.wmax_bits = rs.wmax_bits;
  rg.negw = rs.negw;
  const int M = n[2] / 2;
  PzShape shape{0, 0};
  rc = pz_choose(s, order, M, n, ni, np, rs.sorted, rs.offsets, rs.row_wabs, rg, out, pitch, mesh_stride, shape);
  if (rc) return rc;
  return launch_pz_any(order, M, shape, s, rs.sorted, rs.offsets, rs.row_wabs, rg, out, pitch, mesh_stride);
}

template <int ORDER, int RX, int RY>
static int launch_rows(hipStream_t s, const RowSorted& rs, RowGeom rg, float* meshes, int64_t mesh_stride) {
  rg.rx = RX;
  rg.ry = RY;
  rg.nbx = (int)ceil_div(rg.n[0], RX);
  rg.nby = (int)ceil_div(rg.n[1], RY);
  const size_t lds = (size_t)RX * RY * (rg.n[2] + 1) * sizeof(int);
  auto kern = paint_rows_kernel<ORDER, RX, RY>;
  if (lds > 65536) MP_CHECK_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const unsigned grid = (unsigned)((int64_t)rg.nbx * rg.nby * rg.ni);
  MP_LAUNCH("paint_rows", kern, dim3(grid), dim3(PR_NT), lds, s, rs.sorted, rs.offsets, rs.row_wabs, rg, meshes, mesh_stride);
  return MP_OK;
}

// Rows per block of the row paint: 8x8 rows up to Nz = 256, 4x4 above (at most 131 KiB of LDS).
static bool rows_big(int nz) { return nz > 256; }

bool paint_rows_supported(const int32_t n[3], int64_t np, int order, int ni) {
  if (order < 1 || order > 4 || ni < 1 || np >= (int64_t)INT32_MAX) return false;
  if (n[0] > RS_MAX_KEYS || n[1] > RS_MAX_KEYS || n[2] > 2048 || n[2] < 4) return false;
  const int r = rows_big(n[2]) ? 4 : 8;
  int lo, hi;
  reach_of(order, ni, lo, hi);
  return r + hi - lo <= n[0] && r + hi - lo <= n[1];
}

template <typename T>
static int paint_rows_typed(hipStream_t s, const T* pos, const float* w, int64_t np, const GridMap& g, int order, int ni,
                            float* meshes, int64_t mesh_stride, void* ws, size_t ws_bytes) {
  RowSorted rs;
  int rc = row_sort<T>(s, pos, w, np, g, ws, ws_bytes, rs);
  if (rc) return rc;
  RowGeom rg;
  for (int a = 0; a < 3; ++a) rg.n[a] = g.n[a];
  reach_of(order, ni, rg.lo, rg.hi);
  rg.ni = ni;
  rg.zfrac = zfrac_of(g.n[2]);
  rg.nmock = 1;
  rg.wmax_bits = rs.wmax_bits;
  rg.negw = rs.negw;
  const bool big = rows_big(g.n[2]);
#define MP_ROWS(O)                                                                                   \
  case O:                                                                                            \
    return big ? launch_rows<O, 4, 4>(s, rs, rg, meshes, mesh_stride) : launch_rows<O, 8, 8>(s, rs, rg, meshes, mesh_stride);
  switch (order) {
    MP_ROWS(1) MP_ROWS(2) MP_ROWS(3) MP_ROWS(4)
    default: return set_error(MP_ERR_INVALID, "resampler order %d not in 1..4", order);
  }
#undef MP_ROWS
}

// ni real meshes (mesh j at meshes + j * mesh_stride) painted from the row sort.
int paint_rows(hipStream_t s, const void* pos, int pos_f64, const float* w, int64_t np, const GridMap& g, int order, int ni,
               float* meshes, int64_t mesh_stride, void* ws, size_t ws_bytes) {
  if (!paint_rows_supported(g.n, np, order, ni)) return set_error(MP_ERR_UNSUPPORTED, "row paint: shape not supported");
  if (pos_f64) return paint_rows_typed<double>(s, (const double*)pos, w, np, g, order, ni, meshes, mesh_stride, ws, ws_bytes);
  return paint_rows_typed<float>(s, (const float*)pos, w, np, g, order, ni, meshes, mesh_stride, ws, ws_bytes);
}

// ni z-transformed half-spectrum rows (pitch `pitch`, mesh j at out + j * mesh_stride).
int paint_zfft(hipStream_t s, const void* pos, int pos_f64, const float* w, int64_t np, const GridMap& g, int order, int ni,
               float2* out, int64_t pitch, int64_t mesh_stride, void* ws, size_t ws_bytes, const int64_t* moff, int nmock) {
  int rc = ensure_twiddles();
  if (rc) return rc;
  if (pos_f64)
    return paint_zfft_typed<double>(s, (const double*)pos, w, np, g, order, ni, out, pitch, mesh_stride, ws, ws_bytes, moff, nmock);
  return paint_zfft_typed<float>(s, (const float*)pos, w, np, g, order, ni, out, pitch, mesh_stride, ws, ws_bytes, moff, nmock);
}

}  // namespace mp
