// FFT passes of any length (1..2048) along one axis, for meshes whose sizes are not all powers
// of two: the reference sizes meshes with next_fft_size(n, primes=(2, 3, 5, 7))
// (jaxpower/mesh.py:2226-2262), a cellsize fixes them (get_mesh_attrs, mesh.py:2270-2381), and
// its own tests use sizes such as (128, 124, 122) (tests/test_mesh2.py:36), all transformed by
// jnp.fft.rfftn / irfftn (MeshAttrs.r2c / c2r, mesh.py:1679-1719).
//
// Each block holds a tile of C lines of length L in LDS and runs a self-sorting (Stockham)
// mixed-radix transform over the prime factors of L: the stage of radix R and span NS writes
//   out[(j / NS) NS R + k + q NS] = sum_r in[j + r L/R] w^(r (k + q NS) L/(NS R)),  k = j mod NS,
// one output element per work item, with the L twiddles w^e = exp(-+2 pi i e / L) computed once
// per block in float64.  Any prime factor is handled by the same loop, so lengths like 2 x 61
// need no special case; the cost per element is the sum of the prime factors of L.
#include "common.h"
#include "fft.h"
#include "fft_core.h"

#include <cmath>

namespace mp {

constexpr int ANY_NT = 256;
constexpr int ANY_MAX_STAGES = 24;
constexpr int ANY_MAX_L = 2048;

enum { ANY_COLS = 0, ANY_R2C = 1, ANY_C2R = 2 };

struct AnyArgs {
  int L, nst, C;
  int R[ANY_MAX_STAGES];
  // complex columns (ANY_COLS): line (c, o), c in [0, ncols) consecutive in memory, o in
  // [0, nouter) at o * ostride; element l at l * fstride (output: out_fstride, out_ostride)
  const float2* in;
  float2* out;
  int64_t fstride, ostride, out_fstride, out_ostride;
  int ncols, nchunks;
  int64_t nouter;
  // rows along z (ANY_R2C: real rows of L in, half rows of L/2+1 out at pitch; ANY_C2R back)
  const float* rin;
  float* rout;
  int64_t pitch, nrows;
  RealPrologue pro;
  float scale;
};

template <int MODE, bool INV>
__global__ void __launch_bounds__(ANY_NT) any_fft_kernel(AnyArgs a) {
  extern __shared__ float2 sh[];
  const int L = a.L, C = a.C, tid = threadIdx.x;
  float2* tw = sh;
  float2* src = sh + L;
  float2* dst = src + (size_t)C * L;
  for (int e = tid; e < L; e += ANY_NT) {
    double sv, cv;
    sincospi(2.0 * (double)e / (double)L, &sv, &cv);
    tw[e] = make_float2((float)cv, INV ? (float)sv : -(float)sv);
  }
  const int64_t tile = blockIdx.x;
  // load the tile: lines in LDS at src[c * L + l]
  if constexpr (MODE == ANY_COLS) {
    const int64_t o = tile / a.nchunks;
    const int c0 = (int)(tile % a.nchunks) * C;
    const float2* base = a.in + o * a.ostride + c0;
    for (int idx = tid; idx < C * L; idx += ANY_NT) {
      const int l = idx / C, cc = idx % C;
      src[cc * L + l] = c0 + cc < a.ncols ? base[l * a.fstride + cc] : make_float2(0.f, 0.f);
    }
  } else if constexpr (MODE == ANY_R2C) {
    const int64_t row0 = tile * C;
    for (int idx = tid; idx < C * L; idx += ANY_NT) {
      const int cc = idx / L, l = idx % L;
      const int64_t row = row0 + cc;
      float v = 0.f;
      if (row < a.nrows) {
        v = a.rin[row * L + l];
        if (a.pro.kind != PRO_NONE) v *= a.pro.eval(row, l);
      }
      src[cc * L + l] = make_float2(v, 0.f);
    }
  } else {  // ANY_C2R: the hermitian row, X[L - k] = conj(X[k])
    const int64_t row0 = tile * C;
    for (int idx = tid; idx < C * L; idx += ANY_NT) {
      const int cc = idx / L, l = idx % L;
      const int64_t row = row0 + cc;
      float2 v = make_float2(0.f, 0.f);
      if (row < a.nrows) v = l <= L / 2 ? a.in[row * a.pitch + l] : cconj(a.in[row * a.pitch + (L - l)]);
      src[cc * L + l] = v;
    }
  }
  __syncthreads();
  int NS = 1;
  for (int st = 0; st < a.nst; ++st) {
    const int R = a.R[st];
    const int span = L / R, step = L / (NS * R);
    for (int idx = tid; idx < C * L; idx += ANY_NT) {
      const int cc = idx / L, o = idx % L;
      const int q = (o / NS) % R, k = o % NS, j = (o / (NS * R)) * NS + k;
      const int m = (k + q * NS) * step;  // < L
      const float2* line = src + cc * L;
      float2 acc = line[j];
      int t = m;
      for (int r = 1; r < R; ++r) {
        acc = cadd(acc, cmul(line[j + r * span], tw[t]));
        t += m;
        if (t >= L) t -= L;
      }
      dst[cc * L + o] = acc;
    }
    __syncthreads();
    float2* tmp = src;
    src = dst;
    dst = tmp;
    NS *= R;
  }
  // store
  if constexpr (MODE == ANY_COLS) {
    const int64_t o = tile / a.nchunks;
    const int c0 = (int)(tile % a.nchunks) * C;
    float2* base = a.out + o * a.out_ostride + c0;
    for (int idx = tid; idx < C * L; idx += ANY_NT) {
      const int l = idx / C, cc = idx % C;
      if (c0 + cc < a.ncols) base[l * a.out_fstride + cc] = cscale(src[cc * L + l], a.scale);
    }
  } else if constexpr (MODE == ANY_R2C) {
    const int H = L / 2 + 1;
    const int64_t row0 = tile * C;
    for (int idx = tid; idx < C * H; idx += ANY_NT) {
      const int cc = idx / H, k = idx % H;
      const int64_t row = row0 + cc;
      if (row < a.nrows) a.out[row * a.pitch + k] = src[cc * L + k];
    }
  } else {
    const int64_t row0 = tile * C;
    for (int idx = tid; idx < C * L; idx += ANY_NT) {
      const int cc = idx / L, l = idx % L;
      const int64_t row = row0 + cc;
      if (row < a.nrows) a.rout[row * L + l] = src[cc * L + l].x * a.scale;
    }
  }
}

// Elementwise: acc (+)= Y_lm(khat) * t over a half spectrum (the Y_lm(khat) factor and the m sum
// of the local line of sight, jaxpower/mesh2.py:319-329).
__global__ void ylm_acc_any_kernel(const float2* __restrict__ t, int64_t tpitch, float2* __restrict__ acc, int64_t apitch, int n0,
                                   int n1, int ncols, KPrologue kp, int first) {
  const int64_t total = (int64_t)n0 * n1 * ncols;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int iz = (int)(i % ncols);
    const int64_t row = i / ncols;
    const int iy = (int)(row % n1), ix = (int)(row / n1);
    const float y = kp.eval(ix, iy, iz);
    const float2 v = cscale(t[row * tpitch + iz], y);
    float2* d = acc + row * apitch + iz;
    *d = first ? v : cadd(*d, v);
  }
}

// ------------------------------------------------------------------ host side

static int any_plan(int L, AnyArgs& a) {
  if (L < 1 || L > ANY_MAX_L) return set_error(MP_ERR_UNSUPPORTED, "FFT length %d not in 1..%d", L, ANY_MAX_L);
  a.L = L;
  a.nst = 0;
  int n = L;
  for (int p = 2; n > 1; p = (p == 2 ? 3 : p + 2)) {
    while (n % p == 0) {
      if (a.nst >= ANY_MAX_STAGES) return set_error(MP_ERR_UNSUPPORTED, "FFT length %d: too many factors", L);
      a.R[a.nst++] = p;
      n /= p;
    }
    if ((int64_t)p * p > n && n > 1) {  // what is left is prime
      a.R[a.nst++] = n;
      n = 1;
    }
  }
  a.C = L >= ANY_MAX_L ? 1 : ANY_MAX_L / L;
  if (a.C > 16) a.C = 16;
  return MP_OK;
}

static size_t any_lds(const AnyArgs& a) { return ((size_t)a.L + 2 * (size_t)a.C * a.L) * sizeof(float2); }

template <int MODE, bool INV>
static int any_launch(hipStream_t s, AnyArgs& a, int64_t ntiles, const char* name) {
  if (ntiles <= 0) return MP_OK;
  const size_t lds = any_lds(a);
  auto kern = any_fft_kernel<MODE, INV>;
  if (lds > 65536) MP_CHECK_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  if (ntiles > 0x7fffffff) return set_error(MP_ERR_UNSUPPORTED, "%s: %lld tiles", name, (long long)ntiles);
  MP_LAUNCH(name, kern, dim3((unsigned)ntiles), dim3(ANY_NT), lds, s, a);
  return MP_OK;
}

bool fft_any_shape_ok(const int32_t n[3]) {
  for (int q = 0; q < 3; ++q)
    if (n[q] < 1 || n[q] > ANY_MAX_L) return false;
  return true;
}

// complex FFT of length L along columns: ncols consecutive columns per line group, nouter groups
int any_col_fft(hipStream_t s, const float2* in, float2* out, int L, int ncols, int64_t nouter, int64_t fstride, int64_t ostride,
                int64_t out_fstride, int64_t out_ostride, bool inv, float scale, const char* name) {
  AnyArgs a{};
  int rc = any_plan(L, a);
  if (rc) return rc;
  a.in = in;
  a.out = out;
  a.fstride = fstride;
  a.ostride = ostride;
  a.out_fstride = out_fstride;
  a.out_ostride = out_ostride;
  a.ncols = ncols;
  a.nouter = nouter;
  a.nchunks = (int)ceil_div(ncols, a.C);
  a.scale = scale;
  const int64_t ntiles = (int64_t)a.nchunks * nouter;
  return inv ? any_launch<ANY_COLS, true>(s, a, ntiles, name) : any_launch<ANY_COLS, false>(s, a, ntiles, name);
}

int any_zpass_r2c(hipStream_t s, const float* in, float2* out, int L, int64_t nrows, int64_t pitch, const RealPrologue& pro) {
  AnyArgs a{};
  int rc = any_plan(L, a);
  if (rc) return rc;
  a.rin = in;
  a.out = out;
  a.pitch = pitch;
  a.nrows = nrows;
  a.pro = pro;
  a.scale = 1.f;
  return any_launch<ANY_R2C, false>(s, a, ceil_div(nrows, a.C), "fft_zpass_r2c_any");
}

int any_zpass_c2r(hipStream_t s, const float2* in, int64_t pitch, float* out, int L, int64_t nrows, float scale) {
  AnyArgs a{};
  int rc = any_plan(L, a);
  if (rc) return rc;
  a.in = in;
  a.rout = out;
  a.pitch = pitch;
  a.nrows = nrows;
  a.scale = scale;
  return any_launch<ANY_C2R, true>(s, a, ceil_div(nrows, a.C), "fft_zpass_c2r_any");
}

int ylm_acc_any(hipStream_t s, const float2* t, int64_t tpitch, float2* acc, int64_t apitch, const int32_t n[3], const KPrologue& kp,
                int first) {
  const int ncols = n[2] / 2 + 1;
  const int64_t total = (int64_t)n[0] * n[1] * ncols;
  const int64_t blocks = ceil_div(total, 256) < 65536 ? ceil_div(total, 256) : 65536;
  MP_LAUNCH("ylm_acc_any", ylm_acc_any_kernel, dim3((unsigned)(blocks < 1 ? 1 : blocks)), dim3(256), 0, s, t, tpitch, acc, apitch,
            n[0], n[1], ncols, kp, first);
  return MP_OK;
}

}  // namespace mp
