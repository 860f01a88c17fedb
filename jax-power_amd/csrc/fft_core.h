// Device building blocks of the FFT passes: complex helpers, radix-2/4/8 butterflies, an
// in-LDS Stockham FFT over C independent columns, and the r2c post-processing step.
// Twiddles come from a table tw[e] = exp(-2 pi i e / TW_N).
#pragma once

#include <hip/hip_runtime.h>

#include "fft.h"

namespace mp {

// ------------------------------------------------------------------ device helpers

// Streaming (non-temporal) access to mesh data that a pass reads or writes once, for tiles
// whose rows cover whole 128-byte lines (measured: the y pass of a 512^3 spectrum, 16 columns,
// runs 10 % faster; with 4 or 8 columns per row the other half of each line is lost before
// the neighbouring tile reads it, and the 1024^3 y and x passes ran 1.4x slower).
typedef float stream_f2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 load_stream(const float2* p) {
  const stream_f2 v = __builtin_nontemporal_load(reinterpret_cast<const stream_f2*>(p));
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ void store_stream(float2* p, float2 v) {
  const stream_f2 t = {v.x, v.y};
  __builtin_nontemporal_store(t, reinterpret_cast<stream_f2*>(p));
}

// First tile of a persistent block.  Blocks are dispatched to the 8 XCDs round-robin; here
// the blocks of one XCD take a contiguous run of tiles, so the two column chunks that share a
// 128-byte line of the input meet in that XCD's L2 instead of being fetched by two XCDs.
// Advancing by gridDim.x keeps the runs contiguous.
__device__ __forceinline__ int xcd_first_tile() {
  const int g = gridDim.x, b = blockIdx.x;
  return (g & 7) ? b : (b & 7) * (g >> 3) + (b >> 3);
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 rot90(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

template <bool INV>
__device__ __forceinline__ void dft2(float2& a, float2& b) {
  float2 t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <bool INV>
__device__ __forceinline__ void dft4(float2& v0, float2& v1, float2& v2, float2& v3) {
  float2 t0 = cadd(v0, v2), t1 = csub(v0, v2);
  float2 t2 = cadd(v1, v3), t3 = rot90<INV>(csub(v1, v3));
  v0 = cadd(t0, t2);
  v2 = csub(t0, t2);
  v1 = cadd(t1, t3);
  v3 = csub(t1, t3);
}

template <int R, bool INV>
__device__ __forceinline__ void dft(float2* v) {
  if constexpr (R == 2) {
    dft2<INV>(v[0], v[1]);
  } else if constexpr (R == 4) {
    dft4<INV>(v[0], v[1], v[2], v[3]);
  } else {  // R == 8: two radix-4 on even/odd, then W8 twiddles
    float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<INV>(e0, e1, e2, e3);
    dft4<INV>(o0, o1, o2, o3);
    const float h = 0.70710678118654752f;
    // W8^1 = h(1 -+ i), W8^2 = -+i, W8^3 = h(-1 -+ i)  (upper sign forward)
    float2 w1 = INV ? make_float2(h * (o1.x - o1.y), h * (o1.x + o1.y)) : make_float2(h * (o1.x + o1.y), h * (o1.y - o1.x));
    float2 w2 = rot90<INV>(o2);
    float2 w3 = INV ? make_float2(-h * (o3.x + o3.y), h * (o3.x - o3.y)) : make_float2(h * (o3.y - o3.x), -h * (o3.x + o3.y));
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, w1);
    v[5] = csub(e1, w1);
    v[2] = cadd(e2, w2);
    v[6] = csub(e2, w2);
    v[3] = cadd(e3, w3);
    v[7] = csub(e3, w3);
  }
}

// Column passes: C columns of length L per block (C*L*8 bytes of LDS = 64 KiB at L >= 512).
template <int L>
struct ColCfg {
  static constexpr int C = L <= 512 ? 16 : 8192 / L;
  static constexpr int NT0 = C * L / 8;
  static constexpr int NT = NT0 < 64 ? 64 : NT0;
  static constexpr size_t LDS = (size_t)C * L * sizeof(float2);
};

// Tiles of the pipelined column passes: C columns x L rows, EPT = C*L/NT elements per thread
// (a whole number for every power of two L in [2, 2048]).
template <int L, int CC>
struct XCfgC {
  static constexpr int C = CC;
  static constexpr int NT0 = C * L / 8;
  static constexpr int NT = NT0 < 64 ? 64 : NT0;
  static constexpr int EPT = C * L / NT;
  static constexpr size_t LDS = (size_t)C * L * sizeof(float2);
};
template <int L>
using XCfg = XCfgC<L, L <= 512 ? (L >= 4 ? 16 : 64 / L) : 8192 / L>;

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }
constexpr int first_radix(int L) { return ilog2(L) % 3 == 0 ? 8 : (ilog2(L) % 3 == 1 ? 2 : 4); }

// One Stockham stage over C independent length-L transforms held in LDS; element (l, c)
// lives at buf[l * LS + c * CS].  CFAST: consecutive threads take consecutive c.  tw is a
// table of TWN twiddles exp(-2 pi i e / TWN) (TWN a multiple of L).
template <int L, int C, int LS, int CS, bool CFAST, int NT, bool INV, int R, int NS, int TWN>
__device__ __forceinline__ void fft_stage(float2* buf, const float2* __restrict__ tw) {
  constexpr int NB = C * (L / R);
  constexpr int BPT = (NB + NT - 1) / NT;
  constexpr int TWSTEP = TWN / (NS * R);
  float2 v[BPT][R];
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    int b = tid + q * NT;
    if (NB % NT == 0 || b < NB) {
      int c = CFAST ? b % C : b / (L / R);
      int jj = CFAST ? b / C : b % (L / R);
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = buf[(jj + r * (L / R)) * LS + c * CS];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < BPT; ++q) {
    int b = tid + q * NT;
    if (NB % NT == 0 || b < NB) {
      int c = CFAST ? b % C : b / (L / R);
      int jj = CFAST ? b / C : b % (L / R);
      int k = jj % NS;
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          float2 w = tw[r * k * TWSTEP];
          if (INV) w.y = -w.y;
          v[q][r] = cmul(v[q][r], w);
        }
      }
      dft<R, INV>(v[q]);
      int out = (jj / NS) * NS * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[(out + r * NS) * LS + c * CS] = v[q][r];
    }
  }
  __syncthreads();
}

template <int L, int C, int LS, int CS, bool CFAST, int NT, bool INV, int NS, int TWN>
__device__ __forceinline__ void fft_stages(float2* buf, const float2* __restrict__ tw) {
  if constexpr (NS < L) {
    constexpr int R = NS == 1 ? first_radix(L) : 8;
    fft_stage<L, C, LS, CS, CFAST, NT, INV, R, NS, TWN>(buf, tw);
    fft_stages<L, C, LS, CS, CFAST, NT, INV, NS * R, TWN>(buf, tw);
  }
}

template <int L, int C, int LS, int CS, bool CFAST, int NT, bool INV, int NS, int NSEND, int TWN>
__device__ __forceinline__ void fft_stages_until(float2* buf, const float2* __restrict__ tw) {
  if constexpr (NS < NSEND) {
    constexpr int R = NS == 1 ? first_radix(L) : 8;
    fft_stage<L, C, LS, CS, CFAST, NT, INV, R, NS, TWN>(buf, tw);
    fft_stages_until<L, C, LS, CS, CFAST, NT, INV, NS * R, NSEND, TWN>(buf, tw);
  }
}

// The column passes hold their tile in registers between the global loads/stores and LDS:
// element e of thread t is (l = t / C + e * NT / C, c = t % C).  Those are whole butterflies of
// the first Stockham stage (stride L / R1 in l) and whole outputs of the last one (positions
// jj + r L / 8), so these two stages can run on registers, saving an LDS round trip each.
template <int L, int C, int NT>
struct RegFFT {
  static constexpr int EPT = C * L / NT, S = NT / C, R1 = first_radix(L);
  static constexpr int NSLAST = L / 8;  // the last stage (radix 8) when there are at least two
  static constexpr bool ok = L > R1 && NT % C == 0 && EPT % R1 == 0 && (L / R1) % S == 0 && EPT % 8 == 0 && (L / 8) % S == 0;
};

// First stage (no twiddles) from the thread's registers into buf (element (l, c) at l * C + c);
// the caller syncs before the next stage reads buf.
template <int L, int C, int NT, bool INV>
__device__ __forceinline__ void fft_first_from_regs(const float2* reg, float2* buf) {
  using F = RegFFT<L, C, NT>;
  constexpr int R = F::R1, NB = F::EPT / R;
  const int c = threadIdx.x % C, j = threadIdx.x / C;
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = reg[q + r * NB];  // element j + q S + r L / R
    dft<R, INV>(v);
    const int out = (j + q * F::S) * R;
#pragma unroll
    for (int r = 0; r < R; ++r) buf[(out + r) * C + c] = v[r];
  }
}

// Last stage (radix 8, NS = L / 8) from buf (synced by the caller) into the thread's registers
// in the load layout; buf is only read.
template <int L, int C, int NT, bool INV, int TWN>
__device__ __forceinline__ void fft_last_to_regs(const float2* buf, float2* reg, const float2* __restrict__ tw) {
  using F = RegFFT<L, C, NT>;
  constexpr int R = 8, NS = F::NSLAST, NB = F::EPT / R, TWSTEP = TWN / (NS * R);
  const int c = threadIdx.x % C, j = threadIdx.x / C;
#pragma unroll
  for (int q = 0; q < NB; ++q) {
    const int jj = j + q * F::S;  // < NS, so k = jj and the outputs are jj + r NS
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = buf[(jj + r * NS) * C + c];
#pragma unroll
    for (int r = 1; r < R; ++r) {
      float2 w = tw[r * jj * TWSTEP];
      if (INV) w.y = -w.y;
      v[r] = cmul(v[r], w);
    }
    dft<R, INV>(v);
#pragma unroll
    for (int r = 0; r < R; ++r) reg[q + r * NB] = v[r];
  }
}

// Full in-place FFT of C columns in LDS (callers sync before; this syncs after), with a
// TWN-entry twiddle table (normally an LDS copy made by load_twiddles).
template <int L, int C, int LS, int CS, bool CFAST, int NT, bool INV, int TWN = TW_N>
__device__ __forceinline__ void fft_lds(float2* buf, const float2* __restrict__ tw) {
  static_assert(TWN % L == 0, "twiddle table must hold the L-th roots of unity");
  fft_stages<L, C, LS, CS, CFAST, NT, INV, 1, TWN>(buf, tw);
}

// Copies the TWN-entry table exp(-2 pi i e / TWN) out of the global TW_N-entry table (the
// caller synchronizes before use).  Twiddles read from LDS keep global-memory latency out of
// every butterfly stage.
template <int TWN, int NT>
__device__ __forceinline__ void load_twiddles(float2* stw, const float2* __restrict__ gtw) {
  static_assert(TW_N % TWN == 0, "twiddle table length must divide TW_N");
  for (int i = threadIdx.x; i < TWN; i += NT) stw[i] = gtw[i * (TW_N / TWN)];
}

// Post-processing of the r2c z pass for output index k in [0, M/2] of one row: the row holds
// Z = FFT_M(x[2m] + i x[2m+1]); writes X[k] and X[M-k] of the N = 2M point real transform.
// tw: table of TWN twiddles, TWN a multiple of 2M.
// Element m of the row is row[m * stride].
template <int TWN = TW_N>
__device__ __forceinline__ void rfft_post(const float2* row, float2* o, int M, int k, const float2* __restrict__ tw,
                                          int stride = 1) {
  float2 zk = row[k * stride];
  if (k == 0) {
    o[0] = make_float2(zk.x + zk.y, 0.f);
    o[M] = make_float2(zk.x - zk.y, 0.f);
    return;
  }
  float2 zmk = cconj(row[(M - k) * stride]);
  float2 E = cscale(cadd(zk, zmk), 0.5f);
  float2 D = cscale(csub(zk, zmk), 0.5f);
  float2 O = make_float2(D.y, -D.x);                // -i * D
  float2 t = cmul(tw[k * (TWN / (2 * M))], O);     // exp(-2 pi i k / (2M)) * O
  o[k] = cadd(E, t);
  if (2 * k != M) o[M - k] = cconj(csub(E, t));
}


}  // namespace mp
