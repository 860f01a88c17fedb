// Particle painting (NGP/CIC/TSC/PCS) with interlacing shifts, and its adjoint readout.
//
// Replaces resamplers.paint / resamplers.read (jaxpower/resamplers.py:73-131) as driven by
// _paint / _read (jaxpower/mesh.py:2673-2754, 2065-2115):
//   x~ = (x + L/2 - c) / cellsize; i0 = round (odd order, half to even) or floor (even order);
//   stencil offsets arange(p) - (p-1)//2 with periodic wrap; weight prod_a K_p(|i_a - x~_a|).
// Mesh j of an interlaced paint receives the particles shifted by +j/n_i grid units.
//
// Tiled path (no global atomics on the mesh):
//   1. count particles per 4^3-cell brick (atomics on a small counter array),
//   2. exclusive scan -> brick offsets,
//   3. scatter particles into brick order as float4(local x, local y, local z, weight), with
//      local coordinates relative to the brick origin (full float precision at any N),
//   4. one workgroup per 16^3-cell tile: gathers every brick whose particles can reach the
//      tile, accumulates all n_i shifted meshes in LDS, and writes each tile once with
//      coalesced stores.  Every mesh cell is written exactly once, so no zero fill is needed.
// Meshes whose axes do not tile (size not a multiple of 4, or between 17 and 31) use a
// per-particle kernel with float atomics on a zero-filled mesh.
#include "paint.h"
#include "resampler.h"

namespace mp {

constexpr int BRICK = 4;
constexpr int TILE = 16;
constexpr int TILE_NT = 256;
constexpr int MAX_TILE_BRICKS = 256;  // 6^3 = 216 needed for order 3/4 with interlacing
constexpr int MAX_TILED_NI = 3;
constexpr int SCAN_NT = 1024;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_BLOCK = SCAN_NT * SCAN_ITEMS;

__device__ __forceinline__ int gcd_u32(unsigned a, unsigned b) {
  while (b) {
    unsigned t = a % b;
    a = b;
    b = t;
  }
  return (int)a;
}

// An odd stride near 0.618 * total that is coprime with total (1 for total <= 2).
__device__ __forceinline__ int coprime_stride(int total) {
  if (total <= 2) return 1;
  unsigned s = ((unsigned)((double)total * 0.6180339887498949)) | 1u;
  while (gcd_u32(s, (unsigned)total) != 1) s += 2;
  return (int)(s % (unsigned)total);
}

__global__ void fill_f32_kernel(float* __restrict__ p, int64_t n, float v) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

__global__ void fill_i32_kernel(int* __restrict__ p, int64_t n, int v) {
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

int fill_f32(hipStream_t s, float* p, int64_t n, float v) {
  if (n <= 0) return MP_OK;
  int64_t blocks = ceil_div(n, 256);
  if (blocks > 65536) blocks = 65536;
  MP_LAUNCH("fill_f32", fill_f32_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
  return MP_OK;
}

int fill_i32_dev(hipStream_t s, int* p, int64_t n, int v) {
  if (n <= 0) return MP_OK;
  int64_t blocks = ceil_div(n, 256);
  if (blocks > 65536) blocks = 65536;
  MP_LAUNCH("fill_i32", fill_i32_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
  return MP_OK;
}

// ------------------------------------------------------------------ atomic path

template <int ORDER, typename T>
__global__ void paint_atomic_kernel(const T* __restrict__ pos, const float* __restrict__ w, int64_t np, GridMap g, int ni,
                                    float* __restrict__ meshes, int64_t mesh_stride) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  constexpr int OMIN = -((ORDER - 1) / 2);
  T xg[3];
  to_grid(pos, i, g, xg);
  const float wi = w ? w[i] : 1.f;
  for (int j = 0; j < ni; ++j) {
    const T s = (T)j / (T)ni;
    int idx[3][ORDER];
    float ker[3][ORDER];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      T u = xg[a] + s;
      T r = round_or_floor(u, ORDER % 2 == 1);
      int i0 = (int)r;
#pragma unroll
      for (int o = 0; o < ORDER; ++o) {
        idx[a][o] = pos_mod(i0 + OMIN + o, g.n[a]);
        ker[a][o] = kern1d<ORDER>((float)fabs(r + (T)(OMIN + o) - u));
      }
    }
    float* m = meshes + j * mesh_stride;
#pragma unroll
    for (int ox = 0; ox < ORDER; ++ox)
#pragma unroll
      for (int oy = 0; oy < ORDER; ++oy) {
        const int64_t rowoff = ((int64_t)idx[0][ox] * g.n[1] + idx[1][oy]) * g.n[2];
        const float kxy = wi * ker[0][ox] * ker[1][oy];
#pragma unroll
        for (int oz = 0; oz < ORDER; ++oz) atomicAdd(m + rowoff + idx[2][oz], kxy * ker[2][oz]);
      }
  }
}

// ------------------------------------------------------------------ tiled path

struct TileGeom {
  int n[3];    // meshsize
  int nb[3];   // bricks per axis
  int t[3];    // tile extent per axis
  int nt[3];   // tiles per axis
  int full[3]; // tile spans the whole (periodic) axis
  int blo, bhi;  // brick range relative to the tile's first brick on partial axes
};

template <typename T>
__global__ void brick_count_kernel(const T* __restrict__ pos, int64_t np, GridMap g, TileGeom tg, int* __restrict__ counts) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  T xg[3];
  to_grid(pos, i, g, xg);
  int b[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) b[a] = pos_mod((int)floor(xg[a] * (T)(1.0 / BRICK)), tg.nb[a]);
  atomicAdd(&counts[((int64_t)b[0] * tg.nb[1] + b[1]) * tg.nb[2] + b[2]], 1);
}

template <typename T>
__global__ void brick_scatter_kernel(const T* __restrict__ pos, const float* __restrict__ w, int64_t np, GridMap g, TileGeom tg,
                                     int* __restrict__ cursor, float4* __restrict__ sorted) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  T xg[3];
  to_grid(pos, i, g, xg);
  int b[3];
  float l[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    T q = floor(xg[a] * (T)(1.0 / BRICK));
    l[a] = fminf(fmaxf((float)(xg[a] - q * (T)BRICK), 0.f), (float)BRICK);
    b[a] = pos_mod((int)q, tg.nb[a]);
  }
  int slot = atomicAdd(&cursor[((int64_t)b[0] * tg.nb[1] + b[1]) * tg.nb[2] + b[2]], 1);
  sorted[slot] = make_float4(l[0], l[1], l[2], w ? w[i] : 1.f);
}

template <int ORDER, int NI>
__global__ void __launch_bounds__(TILE_NT)
paint_tile_kernel(const float4* __restrict__ sorted, const int* __restrict__ offsets, TileGeom tg, float* __restrict__ meshes,
                  int64_t mesh_stride) {
  constexpr int OMIN = -((ORDER - 1) / 2);
  constexpr int TCELLS = TILE * TILE * TILE;
  __shared__ float acc[NI][TCELLS];
  __shared__ int bstart[MAX_TILE_BRICKS];
  __shared__ int bpre[MAX_TILE_BRICKS + 2];
  const int tid = threadIdx.x;

  int tile = blockIdx.x;
  int tc[3];
  tc[2] = tile % tg.nt[2];
  tile /= tg.nt[2];
  tc[1] = tile % tg.nt[1];
  tc[0] = tile / tg.nt[1];
  int t0[3], lo[3], cnt[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    t0[a] = tc[a] * tg.t[a];
    if (tg.full[a]) {
      lo[a] = 0;
      cnt[a] = tg.nb[a];
    } else {
      lo[a] = t0[a] / BRICK + tg.blo;
      cnt[a] = tg.bhi - tg.blo + 1;
    }
  }
  const int ncell = tg.t[0] * tg.t[1] * tg.t[2];
  const int nbr = cnt[0] * cnt[1] * cnt[2];

  for (int idx = tid; idx < NI * TCELLS; idx += TILE_NT) (&acc[0][0])[idx] = 0.f;
  for (int i = tid; i < nbr; i += TILE_NT) {
    int i2 = i % cnt[2], i1 = (i / cnt[2]) % cnt[1], i0 = i / (cnt[2] * cnt[1]);
    int b0 = pos_mod(lo[0] + i0, tg.nb[0]), b1 = pos_mod(lo[1] + i1, tg.nb[1]), b2 = pos_mod(lo[2] + i2, tg.nb[2]);
    int64_t id = ((int64_t)b0 * tg.nb[1] + b1) * tg.nb[2] + b2;
    int s = offsets[id];
    bstart[i] = s;
    bpre[i + 1] = offsets[id + 1] - s;
  }
  __syncthreads();
  if (tid < 64) {  // prefix sum of the <= 256 brick counts by one wavefront
    int local[4];
    int sum = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int i = tid * 4 + q;
      local[q] = i < nbr ? bpre[i + 1] : 0;
      sum += local[q];
    }
    int incl = sum;
#pragma unroll
    for (int d = 1; d < 64; d *= 2) {
      int v = __shfl_up(incl, d, 64);
      if (tid >= d) incl += v;
    }
    int run = incl - sum;
    if (tid == 0) bpre[0] = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int i = tid * 4 + q;
      run += local[q];
      if (i < nbr) bpre[i + 1] = run;
    }
  }
  __syncthreads();

  const int total = bpre[nbr];
  // strided visiting order (q * S mod total, S coprime with total): consecutive lanes take
  // particles from different bricks, which keeps LDS atomic conflicts low for clustered data
  if (tid == 0) bpre[MAX_TILE_BRICKS + 1] = coprime_stride(total);
  __syncthreads();
  const uint64_t S = (uint64_t)bpre[MAX_TILE_BRICKS + 1];
  for (int q = tid; q < total; q += TILE_NT) {
    const int p = (int)(((uint64_t)q * S) % (uint64_t)total);
    int lo_i = 0, hi_i = nbr - 1;  // largest i with bpre[i] <= p
    while (lo_i < hi_i) {
      int mid = (lo_i + hi_i + 1) >> 1;
      if (bpre[mid] <= p) lo_i = mid;
      else hi_i = mid - 1;
    }
    const int i = lo_i;
    const float4 rec = sorted[bstart[i] + (p - bpre[i])];
    const int bi[3] = {i / (cnt[2] * cnt[1]), (i / cnt[2]) % cnt[1], i % cnt[2]};
    int base[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) base[a] = (lo[a] + bi[a]) * BRICK - (tg.full[a] ? 0 : t0[a]);
    const float loc[3] = {rec.x, rec.y, rec.z};
#pragma unroll
    for (int j = 0; j < NI; ++j) {
      const float s = (float)j / (float)NI;
      int cell[3][ORDER];
      float ker[3][ORDER];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        float u = loc[a] + s;
        float r = (ORDER % 2) ? rintf(u) : floorf(u);
        int i0 = (int)r;
#pragma unroll
        for (int o = 0; o < ORDER; ++o) {
          int c = base[a] + i0 + OMIN + o;
          if (tg.full[a]) c = pos_mod(c, tg.n[a]);
          bool ok = c >= 0 && c < tg.t[a];
          cell[a][o] = c;
          ker[a][o] = ok ? kern1d<ORDER>(fabsf(r + (float)(OMIN + o) - u)) : 0.f;
        }
      }
#pragma unroll
      for (int ox = 0; ox < ORDER; ++ox) {
        if (!(cell[0][ox] >= 0 && cell[0][ox] < tg.t[0])) continue;
#pragma unroll
        for (int oy = 0; oy < ORDER; ++oy) {
          if (!(cell[1][oy] >= 0 && cell[1][oy] < tg.t[1])) continue;
          const int rowoff = (cell[0][ox] * tg.t[1] + cell[1][oy]) * tg.t[2];
          const float kxy = rec.w * ker[0][ox] * ker[1][oy];
#pragma unroll
          for (int oz = 0; oz < ORDER; ++oz) {
            if (!(cell[2][oz] >= 0 && cell[2][oz] < tg.t[2])) continue;
            atomicAdd(&acc[j][rowoff + cell[2][oz]], kxy * ker[2][oz]);
          }
        }
      }
    }
  }
  __syncthreads();

  for (int idx = tid; idx < NI * ncell; idx += TILE_NT) {
    int j = idx / ncell, c = idx % ncell;
    int cz = c % tg.t[2], cy = (c / tg.t[2]) % tg.t[1], cx = c / (tg.t[2] * tg.t[1]);
    int64_t gidx = ((int64_t)(t0[0] + cx) * tg.n[1] + (t0[1] + cy)) * tg.n[2] + (t0[2] + cz);
    meshes[j * mesh_stride + gidx] = acc[j][c];
  }
}

// ------------------------------------------------------------------ exclusive scan (int32)

__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d *= 2) {
    int t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  return v;
}

// Block-local exclusive scan of SCAN_BLOCK items; block totals to sums.
__global__ void __launch_bounds__(SCAN_NT) scan_block_kernel(int* __restrict__ data, int64_t n, int* __restrict__ sums) {
  __shared__ int wsum[SCAN_NT / 64];
  const int tid = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * SCAN_BLOCK + (int64_t)tid * SCAN_ITEMS;
  int v[SCAN_ITEMS];
  int local = 0;
#pragma unroll
  for (int q = 0; q < SCAN_ITEMS; ++q) {
    v[q] = base + q < n ? data[base + q] : 0;
    local += v[q];
  }
  int incl = wave_incl_scan(local);
  if ((tid & 63) == 63) wsum[tid >> 6] = incl;
  __syncthreads();
  if (tid < 64) {
    int x = tid < SCAN_NT / 64 ? wsum[tid] : 0;
    int xi = wave_incl_scan(x);
    if (tid < SCAN_NT / 64) wsum[tid] = xi - x;
    if (tid == SCAN_NT / 64 - 1 && sums) sums[blockIdx.x] = xi;
  }
  __syncthreads();
  int run = wsum[tid >> 6] + incl - local;
#pragma unroll
  for (int q = 0; q < SCAN_ITEMS; ++q) {
    if (base + q < n) data[base + q] = run;
    run += v[q];
  }
}

__global__ void scan_add_kernel(int* __restrict__ data, int64_t n, const int* __restrict__ sums, int* __restrict__ copy) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int v = data[i] + sums[i / SCAN_BLOCK];
  data[i] = v;
  if (copy) copy[i] = v;
}

size_t scan_workspace_bytes(int64_t n);

static size_t scan_workspace_ints(int64_t n) {
  size_t tot = 0;
  while (n > SCAN_BLOCK) {
    n = ceil_div(n, SCAN_BLOCK);
    tot += align_up((size_t)n, 64);
  }
  return tot + 64;
}

size_t scan_workspace_bytes(int64_t n) { return align_up(scan_workspace_ints(n) * 4, 256); }

// In-place exclusive scan of data[0..n); copy (optional) receives the result too.
int exclusive_scan_i32(hipStream_t s, int* data, int64_t n, int* tmp, int* copy) {
  int64_t nblocks = ceil_div(n, SCAN_BLOCK);
  if (nblocks <= 1) {
    MP_LAUNCH("scan_block", scan_block_kernel, dim3(1), dim3(SCAN_NT), 0, s, data, n, (int*)nullptr);
    if (copy) MP_CHECK_HIP(hipMemcpyAsync(copy, data, n * sizeof(int), hipMemcpyDeviceToDevice, s));
    return MP_OK;
  }
  int* sums = tmp;
  MP_LAUNCH("scan_block", scan_block_kernel, dim3((unsigned)nblocks), dim3(SCAN_NT), 0, s, data, n, sums);
  int rc = exclusive_scan_i32(s, sums, nblocks, tmp + align_up((size_t)nblocks, 64), nullptr);
  if (rc) return rc;
  MP_LAUNCH("scan_add", scan_add_kernel, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s, data, n, (const int*)sums, copy);
  return MP_OK;
}

// ------------------------------------------------------------------ host side

struct Reach {
  int lo, hi;  // min/max cell offset, relative to a particle's home cell, that it can touch
};

static Reach reach(int order, int ni) {
  int omin = -((order - 1) / 2), omax = order - 1 - (order - 1) / 2;
  int i0max = (order % 2) ? (ni > 1 ? 2 : 1) : (ni > 1 ? 1 : 0);
  return Reach{omin, i0max + omax};
}

static bool make_tile_geom(const int32_t n[3], int order, int ni, TileGeom& tg) {
  Reach r = reach(order, ni);
  // homes h in [t0 - r.hi, t0 + T - 1 - r.lo] -> bricks [floor((t0 - r.hi)/B), floor((t0 + T - 1 - r.lo)/B)]
  // (one extra cell of margin on the high side covers local coordinates rounded up to BRICK)
  tg.blo = -(int)ceil_div(r.hi + 1, BRICK);
  tg.bhi = (TILE - 1 - r.lo) / BRICK;
  for (int a = 0; a < 3; ++a) {
    tg.n[a] = n[a];
    if (n[a] % BRICK) return false;
    tg.nb[a] = n[a] / BRICK;
    if (n[a] <= TILE) {
      tg.t[a] = n[a];
      tg.full[a] = 1;
    } else {
      if (n[a] % TILE) return false;
      tg.t[a] = TILE;
      tg.full[a] = 0;
      if (tg.bhi - tg.blo + 1 > tg.nb[a]) return false;  // neighbour bricks would alias
    }
    tg.nt[a] = n[a] / tg.t[a];
  }
  int nbr = 1;
  for (int a = 0; a < 3; ++a) nbr *= tg.full[a] ? tg.nb[a] : tg.bhi - tg.blo + 1;
  return nbr <= MAX_TILE_BRICKS;
}

static bool use_tiled(const int32_t n[3], int64_t np, int order, int ni, TileGeom& tg) {
  if (ni < 1 || ni > MAX_TILED_NI || order < 1 || order > 4) return false;
  if (np >= (int64_t)INT32_MAX) return false;
  return make_tile_geom(n, order, ni, tg);
}

// The row paint serves every shape it supports unless the path option asks for the tiled (1)
// or atomic (3) paint.
static bool use_rows(const int32_t n[3], int64_t np, int order, int ni) {
  const int opt = paint_path_option();
  return opt != 1 && opt != 3 && paint_rows_supported(n, np, order, ni);
}

size_t paint_workspace_bytes(const int32_t n[3], int64_t np, int order, int ni) {
  if (use_rows(n, np, order, ni)) return paint_zfft_workspace_bytes(n, np);
  TileGeom tg;
  if (!use_tiled(n, np, order, ni, tg)) return 256;
  int64_t nbt = (int64_t)tg.nb[0] * tg.nb[1] * tg.nb[2];
  size_t bytes = 0;
  bytes += align_up((size_t)(nbt + 1) * 4, 256);  // offsets
  bytes += align_up((size_t)(nbt + 1) * 4, 256);  // cursor (the scan also copies the total)
  bytes += align_up(scan_workspace_ints(nbt + 1) * 4, 256);
  bytes += align_up((size_t)(np > 0 ? np : 1) * 16, 256);  // sorted particles
  return bytes + 256;
}

template <int ORDER, typename T>
static int launch_atomic(hipStream_t s, const T* pos, const float* w, int64_t np, const GridMap& g, int ni, float* meshes,
                         int64_t stride) {
  if (np == 0) return MP_OK;
  MP_LAUNCH("paint_atomic", (paint_atomic_kernel<ORDER, T>), dim3((unsigned)ceil_div(np, 256)), dim3(256), 0, s, pos, w, np, g, ni,
            meshes, stride);
  return MP_OK;
}

template <int ORDER, int NI>
static int launch_tile(hipStream_t s, const float4* sorted, const int* offsets, const TileGeom& tg, float* meshes, int64_t stride) {
  unsigned ntiles = (unsigned)(tg.nt[0] * tg.nt[1] * tg.nt[2]);
  MP_LAUNCH("paint_tile", (paint_tile_kernel<ORDER, NI>), dim3(ntiles), dim3(TILE_NT), 0, s, sorted, offsets, tg, meshes, stride);
  return MP_OK;
}

template <int ORDER>
static int dispatch_tile(int ni, hipStream_t s, const float4* sorted, const int* offsets, const TileGeom& tg, float* meshes,
                         int64_t stride) {
  switch (ni) {
    case 1: return launch_tile<ORDER, 1>(s, sorted, offsets, tg, meshes, stride);
    case 2: return launch_tile<ORDER, 2>(s, sorted, offsets, tg, meshes, stride);
    default: return launch_tile<ORDER, 3>(s, sorted, offsets, tg, meshes, stride);
  }
}

template <typename T>
static int paint_typed(hipStream_t s, const T* pos, const float* w, int64_t np, const GridMap& g, int order, int ni,
                       float* meshes, int64_t stride, void* ws, size_t ws_bytes, int force_atomic) {
  TileGeom tg;
  const int64_t ncell = (int64_t)g.n[0] * g.n[1] * g.n[2];
  if (force_atomic || !use_tiled(g.n, np, order, ni, tg)) {
    for (int j = 0; j < ni; ++j) {
      int rc = fill_f32(s, meshes + j * stride, ncell, 0.f);
      if (rc) return rc;
    }
    switch (order) {
      case 1: return launch_atomic<1, T>(s, pos, w, np, g, ni, meshes, stride);
      case 2: return launch_atomic<2, T>(s, pos, w, np, g, ni, meshes, stride);
      case 3: return launch_atomic<3, T>(s, pos, w, np, g, ni, meshes, stride);
      case 4: return launch_atomic<4, T>(s, pos, w, np, g, ni, meshes, stride);
      default: return set_error(MP_ERR_INVALID, "resampler order %d not in 1..4", order);
    }
  }
  const int64_t nbt = (int64_t)tg.nb[0] * tg.nb[1] * tg.nb[2];
  Workspace wsp(ws, ws_bytes);
  int* offsets = wsp.take<int>(nbt + 1);
  int* cursor = wsp.take<int>(nbt + 1);
  int* scan_tmp = wsp.take<int>(scan_workspace_ints(nbt + 1));
  float4* sorted = wsp.take<float4>(np > 0 ? np : 1);
  if (!offsets || !cursor || !scan_tmp || !sorted)
    return set_error(MP_ERR_WORKSPACE, "paint: workspace of %zu bytes < %zu needed", ws_bytes, paint_workspace_bytes(g.n, np, order, ni));
  int rc = fill_i32_dev(s, offsets, nbt + 1, 0);
  if (rc) return rc;
  if (np > 0) {
    MP_LAUNCH("paint_brick_count", brick_count_kernel<T>, dim3((unsigned)ceil_div(np, 256)), dim3(256), 0, s, pos, np, g, tg, offsets);
  }
  rc = exclusive_scan_i32(s, offsets, nbt + 1, scan_tmp, cursor);
  if (rc) return rc;
  if (np > 0) {
    MP_LAUNCH("paint_brick_scatter", brick_scatter_kernel<T>, dim3((unsigned)ceil_div(np, 256)), dim3(256), 0, s, pos, w, np, g, tg,
              cursor, sorted);
  }
  switch (order) {
    case 1: return dispatch_tile<1>(ni, s, sorted, offsets, tg, meshes, stride);
    case 2: return dispatch_tile<2>(ni, s, sorted, offsets, tg, meshes, stride);
    case 3: return dispatch_tile<3>(ni, s, sorted, offsets, tg, meshes, stride);
    default: return dispatch_tile<4>(ni, s, sorted, offsets, tg, meshes, stride);
  }
}

int paint_meshes(hipStream_t s, const void* pos, int pos_f64, const float* w, int64_t np, const GridMap& g, int order, int ni,
                 float* meshes, int64_t mesh_stride, void* ws, size_t ws_bytes, int force_atomic) {
  MP_REQUIRE(order >= 1 && order <= 4, "resampler order %d not in 1..4", order);
  MP_REQUIRE(ni >= 1, "interlacing count %d < 1", ni);
  MP_REQUIRE(np >= 0, "negative particle count");
  MP_REQUIRE(np == 0 || pos, "null positions");
  MP_REQUIRE(meshes, "null meshes");
  for (int a = 0; a < 3; ++a) MP_REQUIRE(g.n[a] >= 1 && g.n[a] < 32768, "meshsize[%d] = %d out of range", a, g.n[a]);
  if (!force_atomic && use_rows(g.n, np, order, ni)) {
    if (ws_bytes < paint_zfft_workspace_bytes(g.n, np))
      return set_error(MP_ERR_WORKSPACE, "paint: workspace of %zu bytes < %zu", ws_bytes, paint_zfft_workspace_bytes(g.n, np));
    return paint_rows(s, pos, pos_f64, w, np, g, order, ni, meshes, mesh_stride, ws, ws_bytes);
  }
  if (pos_f64)
    return paint_typed<double>(s, (const double*)pos, w, np, g, order, ni, meshes, mesh_stride, ws, ws_bytes, force_atomic);
  return paint_typed<float>(s, (const float*)pos, w, np, g, order, ni, meshes, mesh_stride, ws, ws_bytes, force_atomic);
}

// ------------------------------------------------------------------ readout

template <int ORDER, typename T>
__global__ void read_kernel(const float* __restrict__ mesh, GridMap g, const T* __restrict__ pos, int64_t np,
                            float* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  constexpr int OMIN = -((ORDER - 1) / 2);
  T xg[3];
  to_grid(pos, i, g, xg);
  int idx[3][ORDER];
  float ker[3][ORDER];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    T r = round_or_floor(xg[a], ORDER % 2 == 1);
    int i0 = (int)r;
#pragma unroll
    for (int o = 0; o < ORDER; ++o) {
      idx[a][o] = pos_mod(i0 + OMIN + o, g.n[a]);
      ker[a][o] = kern1d<ORDER>((float)fabs(r + (T)(OMIN + o) - xg[a]));
    }
  }
  float acc = 0.f;
#pragma unroll
  for (int ox = 0; ox < ORDER; ++ox)
#pragma unroll
    for (int oy = 0; oy < ORDER; ++oy) {
      const float* row = mesh + ((int64_t)idx[0][ox] * g.n[1] + idx[1][oy]) * g.n[2];
      float kxy = ker[0][ox] * ker[1][oy];
#pragma unroll
      for (int oz = 0; oz < ORDER; ++oz) acc += row[idx[2][oz]] * kxy * ker[2][oz];
    }
  out[i] = acc;
}

template <typename T>
static int read_typed(hipStream_t s, const float* mesh, const GridMap& g, const T* pos, int64_t np, int order, float* out) {
  if (np == 0) return MP_OK;
  dim3 grid((unsigned)ceil_div(np, 256)), block(256);
  switch (order) {
    case 1: MP_LAUNCH("read", (read_kernel<1, T>), grid, block, 0, s, mesh, g, pos, np, out); break;
    case 2: MP_LAUNCH("read", (read_kernel<2, T>), grid, block, 0, s, mesh, g, pos, np, out); break;
    case 3: MP_LAUNCH("read", (read_kernel<3, T>), grid, block, 0, s, mesh, g, pos, np, out); break;
    case 4: MP_LAUNCH("read", (read_kernel<4, T>), grid, block, 0, s, mesh, g, pos, np, out); break;
    default: return set_error(MP_ERR_INVALID, "resampler order %d not in 1..4", order);
  }
  return MP_OK;
}

int read_mesh(hipStream_t s, const float* mesh, const GridMap& g, const void* pos, int pos_f64, int64_t np, int order, float* out) {
  if (pos_f64) return read_typed<double>(s, mesh, g, (const double*)pos, np, order, out);
  return read_typed<float>(s, mesh, g, (const float*)pos, np, order, out);
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

// ------------------------------------------------------------------ C ABI

MP_EXPORT int64_t mp_paint_workspace_bytes(const int32_t meshsize[3], int64_t np, int order, int interlacing) {
  return (int64_t)paint_workspace_bytes(meshsize, np, order, interlacing < 1 ? 1 : interlacing);
}

MP_EXPORT int mp_paint(void* stream, const void* positions, int positions_f64, const float* weights, int64_t np,
                       const int32_t meshsize[3], const double boxsize[3], const double boxcenter[3], int order,
                       int interlacing, float* meshes, void* workspace, size_t workspace_bytes) {
  MP_REQUIRE(meshsize && boxsize && boxcenter, "mp_paint: null geometry");
  GridMap g = make_map(meshsize, boxsize, boxcenter);
  int ni = interlacing < 1 ? 1 : interlacing;
  int64_t stride = (int64_t)meshsize[0] * meshsize[1] * meshsize[2];
  return paint_meshes((hipStream_t)stream, positions, positions_f64, weights, np, g, order, ni, meshes, stride, workspace,
                      workspace_bytes, 0);
}

MP_EXPORT int mp_paint_atomic(void* stream, const void* positions, int positions_f64, const float* weights, int64_t np,
                              const int32_t meshsize[3], const double boxsize[3], const double boxcenter[3], int order,
                              int interlacing, float* meshes) {
  MP_REQUIRE(meshsize && boxsize && boxcenter, "mp_paint_atomic: null geometry");
  GridMap g = make_map(meshsize, boxsize, boxcenter);
  int ni = interlacing < 1 ? 1 : interlacing;
  int64_t stride = (int64_t)meshsize[0] * meshsize[1] * meshsize[2];
  return paint_meshes((hipStream_t)stream, positions, positions_f64, weights, np, g, order, ni, meshes, stride, nullptr, 0, 1);
}

MP_EXPORT int mp_read(void* stream, const float* mesh, const int32_t meshsize[3], const double boxsize[3],
                      const double boxcenter[3], const void* positions, int positions_f64, int64_t np, int order, float* out) {
  MP_REQUIRE(mesh && meshsize && boxsize && boxcenter && out, "mp_read: null pointer");
  MP_REQUIRE(order >= 1 && order <= 4, "resampler order %d not in 1..4", order);
  GridMap g = make_map(meshsize, boxsize, boxcenter);
  return read_mesh((hipStream_t)stream, mesh, g, positions, positions_f64, np, order, out);
}
