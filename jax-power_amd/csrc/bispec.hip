// Bispectrum multipoles in the Scoccimarro basis with a global line of sight
// (jaxpower/mesh3.py:180-313 BinMesh3SpectrumPoles, 509-629 compute_mesh3_spectrum):
//   * shells: for each |k| shell b of a half spectrum A, the real mesh
//       I_b(x) = irfftn(A(k) [bin(k) == b] P_ell(k.los/|k|))
//     (mesh3.py:300-305; P_ell of the third mesh, mesh3.py:585-590; A absent: the shell
//     indicator, whose triangle sums count the closed triangles, mesh3.py:229-236);
//   * triangles: S(b1, b2, b3) = sum_x I1_b1(x) I2_b2(x) I3_b3(x) (mesh3.py:307-311), in FP32
//     per cell and FP64 across cells.
//
// The shell meshes are made one at a time in a single complex scratch mesh (split, then an
// in-place inverse FFT), so the workspace is one half spectrum whatever the number of shells.
// The triangle pass is a persistent sweep over cell tiles: the block stages every shell of
// its tile in LDS, and each wave owns a run of triangles (sorted by (b1, b2), so the product
// I1_b1 I2_b2 of a lane's cells is reused along b3) with the lane's running sums in registers.
#include "bessel.h"
#include "bin.h"
#include "common.h"
#include "fft.h"
#include "sph.h"

namespace mp {

constexpr int SH_NT = 256;
constexpr int TR_NT = 512;                // 8 waves: 2 per SIMD, up to 256 VGPRs each
constexpr int TR_WAVES = TR_NT / 64;
constexpr int TR_TILE = 128;              // cells per tile: 2 per lane
constexpr int TR_PER_WAVE = 32;           // triangles per wave and launch
constexpr int TR_FLUSH = 32;              // tiles summed in FP32 (64 cells a lane) before the FP64 add
constexpr int TR_MAX_SHELLS = 96;         // LDS: (distinct stacks x shells) x TR_TILE floats
constexpr int TR_MAX_BLOCKS = 1024;

static_assert(TR_TILE == 2 * 64, "two cells per lane");
static_assert(TR_PER_WAVE <= 64, "lane k gathers the total of the wave's triangle k");

// out = A(k) [bin(k) == b] f(k) on the half spectrum (f = P_ell(mu) when ell_on, else 1).
__global__ void __launch_bounds__(SH_NT)
shell_split_kernel(const float2* __restrict__ A, int64_t pitch, int n0, int n1, int nc, const double* __restrict__ kx,
                   const double* __restrict__ ky, const double* __restrict__ kz, const double* __restrict__ thr, int nedges,
                   int b, float lx, float ly, float lz, LegendreTable leg, int ell_on, float2* __restrict__ out,
                   int64_t out_pitch) {
  const int64_t total = (int64_t)n0 * n1 * nc;
  for (int64_t i = (int64_t)blockIdx.x * SH_NT + threadIdx.x; i < total; i += (int64_t)gridDim.x * SH_NT) {
    const int iz = (int)(i % nc);
    const int64_t row = i / nc;
    const int iy = (int)(row % n1), ix = (int)(row / n1);
    const double kxv = kx[ix], kyv = ky[iy], kzv = kz[iz];
    const int nb = first_bin(thr, nedges, k2_sum(k2_xy(kxv, kyv), kzv)) - 1;
    float2 v = make_float2(0.f, 0.f);
    if (nb == b) {
      v = A ? A[row * pitch + iz] : make_float2(1.f, 0.f);
      if (ell_on) {
        const float fx = (float)kxv, fy = (float)kyv, fz = (float)kzv;
        const float kn2 = fx * fx + fy * fy + fz * fz;
        const float mu = kn2 > 0.f ? (fx * lx + fy * ly + fz * lz) * rsqrtf(kn2) : 0.f;
        const float f = legendre_eval(leg, 0, mu, mu * mu);
        v = make_float2(v.x * f, v.y * f);
      }
    }
    out[row * out_pitch + iz] = v;
  }
}

// out = A(k) [bin(k) == b] Y_lm(khat) (-i if l is odd) on the half spectrum (A null: the shell
// indicator).  Y_lm is the real harmonic (value at the zero vector as the reference's), so for
// even l the product is hermitian and for odd l the factor -i makes it so: the inverse real FFT
// then gives Q with ifftn(A [b] Y_lm) = Q (even l) or i Q (odd l) exactly.
__global__ void __launch_bounds__(SH_NT)
shell_split_ylm_kernel(const float2* __restrict__ A, int64_t pitch, int n0, int n1, int nc, const double* __restrict__ kx,
                       const double* __restrict__ ky, const double* __restrict__ kz, const double* __restrict__ thr, int nedges,
                       int b, int ell, int m, float2* __restrict__ out, int64_t out_pitch) {
  const int64_t total = (int64_t)n0 * n1 * nc;
  for (int64_t i = (int64_t)blockIdx.x * SH_NT + threadIdx.x; i < total; i += (int64_t)gridDim.x * SH_NT) {
    const int iz = (int)(i % nc);
    const int64_t row = i / nc;
    const int iy = (int)(row % n1), ix = (int)(row / n1);
    const double kxv = kx[ix], kyv = ky[iy], kzv = kz[iz];
    const int nb = first_bin(thr, nedges, k2_sum(k2_xy(kxv, kyv), kzv)) - 1;
    float2 v = make_float2(0.f, 0.f);
    if (nb == b) {
      const float2 a = A ? A[row * pitch + iz] : make_float2(1.f, 0.f);
      const float y = ylm_of_vector(ell, m, (float)kxv, (float)kyv, (float)kzv);
      v = (ell & 1) ? make_float2(a.y * y, -a.x * y) : make_float2(a.x * y, a.y * y);
    }
    out[row * out_pitch + iz] = v;
  }
}

// The Bessel-filtered meshes of BinMesh3CorrelationPoles in the Sugiyama basis
// (mesh3.py:383-395): B(k) = A(k) j_l(|k| s) [k0 <= |k| < k1] Y_lm(k_ref) on the full complex
// grid, where k_ref is the reference's wavenumber vector of the mode's index (fftfreq on every
// axis, so -Nyquist at index N/2).  With k' the partner index -k mod N, the hermitian part
// B_h = A j (Y(k_ref) + Y(k'_ref)) / 2 and the antihermitian part B_a = A j (Y(k_ref) - Y(k'_ref)) / 2
// give ifftn(B) = irfftn(B_h) + i irfftn(-i B_a) exactly, the Nyquist planes included (there
// k'_ref is not -k_ref).  part 0 writes B_h, part 1 writes -i B_a, on the half spectrum.
__global__ void __launch_bounds__(SH_NT)
bessel_split_kernel(const float2* __restrict__ A, int64_t pitch, int n0, int n1, int n2, double kf0, double kf1, double kf2, int ell,
                    int m, int ellj, double s, double klo, double khi, int part, float2* __restrict__ out, int64_t out_pitch) {
  const int nc = n2 / 2 + 1;
  const int64_t total = (int64_t)n0 * n1 * nc;
  for (int64_t i = (int64_t)blockIdx.x * SH_NT + threadIdx.x; i < total; i += (int64_t)gridDim.x * SH_NT) {
    const int iz = (int)(i % nc);
    const int64_t row = i / nc;
    const int iy = (int)(row % n1), ix = (int)(row / n1);
    // signed index of fftfreq: i for i < (n + 1) / 2, i - n above (so -n/2 at n/2 for even n)
    auto sgn = [](int q, int n) { return q < (n + 1) / 2 ? q : q - n; };
    const double kx = sgn(ix, n0) * kf0, ky = sgn(iy, n1) * kf1, kz = sgn(iz, n2) * kf2;
    const int jx = ix ? n0 - ix : 0, jy = iy ? n1 - iy : 0, jz = iz ? n2 - iz : 0;
    const double px = sgn(jx, n0) * kf0, py = sgn(jy, n1) * kf1, pz = sgn(jz, n2) * kf2;
    const double kn = sqrt(k2_sum(k2_xy(kx, ky), kz));
    float2 v = make_float2(0.f, 0.f);
    if (kn >= klo && kn < khi) {
      const float2 a = A[row * pitch + iz];
      const float yk = ell < 0 ? 1.f : ylm_of_vector(ell, m, (float)kx, (float)ky, (float)kz);
      const float yp = ell < 0 ? 1.f : ylm_of_vector(ell, m, (float)px, (float)py, (float)pz);
      const float w = (float)spherical_jn(ellj, kn * s) * (part == 0 ? 0.5f * (yk + yp) : 0.5f * (yk - yp));
      v = part == 0 ? make_float2(a.x * w, a.y * w) : make_float2(a.y * w, -a.x * w);
    }
    out[row * out_pitch + iz] = v;
  }
}

struct TriArgs {
  const float* stack[3];  // distinct shell stacks (up to three), shell q of stack s at stack[s] + q * shell_stride
  int nshell[3];
  int lds_base[3];        // first LDS row of each stack
  int role[3];            // stack of each triangle side
  int nstack;
  int64_t shell_stride, ncells;
  const int* tri;         // (b1, b2, b3) per triangle, sorted by (b1, b2)
  int ntri;
  double* partial;        // [gridDim.x][ntri]
};

__global__ void __launch_bounds__(TR_NT)
triangle_sum_kernel(TriArgs a) {
  extern __shared__ float srow[];  // row r (a shell of a stack) of the tile at srow[r * TR_TILE]
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int t0 = (blockIdx.y * TR_WAVES + wave) * TR_PER_WAVE;  // this wave's first triangle
  const int nrows = a.lds_base[a.nstack - 1] + a.nshell[a.nstack - 1];
  const float* s1 = srow + a.lds_base[a.role[0]] * TR_TILE;
  const float* s2 = srow + a.lds_base[a.role[1]] * TR_TILE;
  const float* s3 = srow + a.lds_base[a.role[2]] * TR_TILE;
  float acc[TR_PER_WAVE];  // this lane's FP32 sums, one per triangle of the wave
#pragma unroll
  for (int k = 0; k < TR_PER_WAVE; ++k) acc[k] = 0.f;
  double own = 0.;  // FP64 total of triangle t0 + lane, gathered at every flush
  auto flush = [&]() {
#pragma unroll
    for (int k = 0; k < TR_PER_WAVE; ++k) {
      double v = (double)acc[k];
#pragma unroll
      for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
      if (lane == k) own += v;
      acc[k] = 0.f;
    }
  };
  const int64_t ntiles = ceil_div(a.ncells, TR_TILE);
  int since_flush = 0;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t c0 = tile * TR_TILE;
    for (int i = tid; i < nrows * TR_TILE; i += TR_NT) {
      const int r = i / TR_TILE, c = i % TR_TILE;
      int s = 0;
      while (s + 1 < a.nstack && r >= a.lds_base[s + 1]) ++s;
      const int q = r - a.lds_base[s];
      srow[i] = c0 + c < a.ncells ? a.stack[s][q * a.shell_stride + c0 + c] : 0.f;
    }
    __syncthreads();
    int pb1 = -1, pb2 = -1;
    float p0 = 0.f, p1 = 0.f;
#pragma unroll
    for (int k = 0; k < TR_PER_WAVE; ++k) {
      const int t = t0 + k;
      if (t < a.ntri) {
        const int b1 = a.tri[3 * t], b2 = a.tri[3 * t + 1], b3 = a.tri[3 * t + 2];
        const bool ok = (unsigned)b1 < (unsigned)a.nshell[a.role[0]] && (unsigned)b2 < (unsigned)a.nshell[a.role[1]] &&
                        (unsigned)b3 < (unsigned)a.nshell[a.role[2]];  // a triangle with a shell out of range sums to 0
        if (ok) {
          if (b1 != pb1 || b2 != pb2) {
            const float2 u = *reinterpret_cast<const float2*>(s1 + b1 * TR_TILE + 2 * lane);
            const float2 v = *reinterpret_cast<const float2*>(s2 + b2 * TR_TILE + 2 * lane);
            p0 = u.x * v.x;
            p1 = u.y * v.y;
            pb1 = b1;
            pb2 = b2;
          }
          const float2 w = *reinterpret_cast<const float2*>(s3 + b3 * TR_TILE + 2 * lane);
          acc[k] = fmaf(p1, w.y, fmaf(p0, w.x, acc[k]));
        }
      }
    }
    if (++since_flush == TR_FLUSH) {
      flush();
      since_flush = 0;
    }
    __syncthreads();
  }
  flush();
  if (lane < TR_PER_WAVE && t0 + lane < a.ntri) a.partial[(int64_t)blockIdx.x * a.ntri + t0 + lane] = own;
}

// Spherical-Bessel pair sums of a configuration-space mesh (the last step of the Sugiyama
// shot noise, mesh3.py:875-903): out[p] = sum over the cells s of F(s) j_l1(|s| k1[b1])
// j_l2(|s| k2[b2]) for the pairs (b1, b2) = pairs[2p], pairs[2p + 1].  A block walks tiles of
// HK_TILE cells: it evaluates j_l1(|s| k) and j_l2(|s| k) once per cell and distinct k (float64,
// stored as float in LDS), then each thread adds the tile's products for its HK_PER pairs
// (float per tile, float64 across tiles); per-block partials are summed in a fixed order.
constexpr int HK_NT = 256;
constexpr int HK_TILE = 32;
constexpr int HK_PER = 4;
constexpr int HK_MAX_BLOCKS = 1024;

__global__ void __launch_bounds__(HK_NT)
hankel_pairs_kernel(const float* __restrict__ F, int n0, int n1, int n2, const double* __restrict__ sx, const double* __restrict__ sy,
                    const double* __restrict__ sz, int l1, const double* __restrict__ k1, int nk1, int l2, const double* __restrict__ k2,
                    int nk2, const int* __restrict__ pairs, int npairs, double* __restrict__ partial) {
  extern __shared__ float hk[];
  float* fv = hk;                       // [HK_TILE] field values
  float* j1 = fv + HK_TILE;             // [nk1][HK_TILE]
  float* j2 = j1 + (size_t)nk1 * HK_TILE;  // [nk2][HK_TILE]
  const int tid = threadIdx.x;
  const int64_t ncells = (int64_t)n0 * n1 * n2;
  const int64_t ntiles = (ncells + HK_TILE - 1) / HK_TILE;
  int b1[HK_PER], b2[HK_PER];
  double acc[HK_PER];
#pragma unroll
  for (int q = 0; q < HK_PER; ++q) {
    const int p = tid + q * HK_NT;
    b1[q] = p < npairs ? pairs[2 * p] : -1;
    b2[q] = p < npairs ? pairs[2 * p + 1] : -1;
    acc[q] = 0.;
  }
  const int nk = nk1 + nk2;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t c0 = tile * HK_TILE;
    for (int i = tid; i < HK_TILE; i += HK_NT) {
      const int64_t c = c0 + i;
      fv[i] = c < ncells ? F[c] : 0.f;
    }
    for (int i = tid; i < nk * HK_TILE; i += HK_NT) {
      const int kq = i / HK_TILE, ci = i % HK_TILE;
      const int64_t c = c0 + ci;
      double v = 0.;
      if (c < ncells) {
        const int iz = (int)(c % n2);
        const int64_t row = c / n2;
        const int iy = (int)(row % n1), ix = (int)(row / n1);
        const double r = sqrt(k2_sum(k2_xy(sx[ix], sy[iy]), sz[iz]));
        v = kq < nk1 ? spherical_jn(l1, r * k1[kq]) : spherical_jn(l2, r * k2[kq - nk1]);
      }
      j1[(size_t)kq * HK_TILE + ci] = (float)v;  // j2 follows j1 contiguously
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < HK_PER; ++q) {
      if (b1[q] < 0) continue;
      const float* a = j1 + (size_t)b1[q] * HK_TILE;
      const float* b = j2 + (size_t)b2[q] * HK_TILE;
      float t = 0.f;
      for (int ci = 0; ci < HK_TILE; ++ci) t = fmaf(fv[ci] * a[ci], b[ci], t);
      acc[q] += (double)t;
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < HK_PER; ++q) {
    const int p = tid + q * HK_NT;
    if (p < npairs) partial[(int64_t)blockIdx.x * npairs + p] = acc[q];
  }
}

static int multiprocessor_count(int* ncu) {
  int dev = 0;
  MP_CHECK_HIP(hipGetDevice(&dev));
  MP_CHECK_HIP(hipDeviceGetAttribute(ncu, hipDeviceAttributeMultiprocessorCount, dev));
  return MP_OK;
}

}  // namespace mp

using namespace mp;

MP_EXPORT size_t mp_bispec_shells_workspace_bytes(const int32_t meshsize[3]) {
  if (!meshsize) return 0;
  return align_up((size_t)meshsize[0] * meshsize[1] * spectrum_pitch(meshsize[2]) * sizeof(float2), 256) + 256;
}

// Real meshes of the shells b0 .. b0 + nshell - 1 (bins of the edges whose squared-|k|
// thresholds are thr): out + q * N0 N1 N2 = irfftn(A [bin == b0 + q] P_ell(k.los/|k|)); A
// null: the shell indicators.  los is used when ell > 0.
MP_EXPORT int mp_bispec_shells(void* stream, const void* A, int64_t pitch, const int32_t meshsize[3], const double* kx, const double* ky,
                               const double* kz, const double* thr, int nedges, int b0, int nshell, const double los[3], int ell,
                               float* out, void* workspace, size_t workspace_bytes) {
  MP_REQUIRE(meshsize && kx && ky && kz && thr && out, "mp_bispec_shells: null pointer");
  MP_REQUIRE(nedges >= 2 && b0 >= 0 && nshell >= 1 && b0 + nshell <= nedges - 1, "mp_bispec_shells: shells %d..%d outside the %d bins",
             b0, b0 + nshell - 1, nedges - 1);
  MP_REQUIRE(ell >= 0 && ell <= 10, "mp_bispec_shells: ell %d not in 0..10", ell);
  MP_REQUIRE(ell == 0 || los, "mp_bispec_shells: ell > 0 needs a line of sight");
  MP_REQUIRE(!A || pitch >= meshsize[2] / 2 + 1, "mp_bispec_shells: pitch too small");
  hipStream_t s = (hipStream_t)stream;
  const int64_t sp = spectrum_pitch(meshsize[2]);
  Workspace ws(workspace, workspace_bytes);
  float2* scratch = ws.take<float2>((size_t)meshsize[0] * meshsize[1] * sp);
  if (!scratch) return set_error(MP_ERR_WORKSPACE, "mp_bispec_shells: workspace too small");
  LegendreTable leg{};
  legendre_table_set(leg, 0, ell);
  const float lx = ell ? (float)los[0] : 0.f, ly = ell ? (float)los[1] : 0.f, lz = ell ? (float)los[2] : 0.f;
  const int nc = meshsize[2] / 2 + 1;
  const int64_t total = (int64_t)meshsize[0] * meshsize[1] * nc;
  int ncu = 0;
  int rc = multiprocessor_count(&ncu);
  if (rc) return rc;
  const int blocks = (int)std::min<int64_t>(ceil_div(total, SH_NT), (int64_t)ncu * 16);
  const int64_t cells = (int64_t)meshsize[0] * meshsize[1] * meshsize[2];
  for (int q = 0; q < nshell; ++q) {
    MP_LAUNCH("bispec_shell_split", shell_split_kernel, dim3((unsigned)blocks), dim3(SH_NT), 0, s, (const float2*)A, pitch, meshsize[0],
              meshsize[1], nc, kx, ky, kz, thr, nedges, b0 + q, lx, ly, lz, leg, ell > 0 ? 1 : 0, scratch, sp);
    rc = irfft3d(s, scratch, sp, out + q * cells, meshsize, scratch, sp);  // in place: each pass rewrites the tiles it read
    if (rc) return rc;
  }
  return MP_OK;
}

// Real meshes of the shells b0 .. b0 + nshell - 1 weighted by the real harmonic Y_lm(khat):
// out + q * N0 N1 N2 = Q_q with ifftn(A [bin == b0 + q] Y_lm(khat)) = Q_q for even l and i Q_q
// for odd l (A null: the shell indicators).  These are the components of the Sugiyama basis'
// shell meshes ifftn(A [b] Y*_lm(khat)) in the complex harmonics (mesh3.py:599-606).
MP_EXPORT int mp_bispec_shells_ylm(void* stream, const void* A, int64_t pitch, const int32_t meshsize[3], const double* kx,
                                   const double* ky, const double* kz, const double* thr, int nedges, int b0, int nshell, int ell,
                                   int m, float* out, void* workspace, size_t workspace_bytes) {
  MP_REQUIRE(meshsize && kx && ky && kz && thr && out, "mp_bispec_shells_ylm: null pointer");
  MP_REQUIRE(nedges >= 2 && b0 >= 0 && nshell >= 1 && b0 + nshell <= nedges - 1,
             "mp_bispec_shells_ylm: shells %d..%d outside the %d bins", b0, b0 + nshell - 1, nedges - 1);
  MP_REQUIRE(ell >= 0 && ell <= MAX_ELL && m >= -ell && m <= ell, "mp_bispec_shells_ylm: (ell, m) = (%d, %d) not supported", ell, m);
  MP_REQUIRE(!A || pitch >= meshsize[2] / 2 + 1, "mp_bispec_shells_ylm: pitch too small");
  hipStream_t s = (hipStream_t)stream;
  const int64_t sp = spectrum_pitch(meshsize[2]);
  Workspace ws(workspace, workspace_bytes);
  float2* scratch = ws.take<float2>((size_t)meshsize[0] * meshsize[1] * sp);
  if (!scratch) return set_error(MP_ERR_WORKSPACE, "mp_bispec_shells_ylm: workspace too small");
  const int nc = meshsize[2] / 2 + 1;
  const int64_t total = (int64_t)meshsize[0] * meshsize[1] * nc;
  int ncu = 0;
  int rc = multiprocessor_count(&ncu);
  if (rc) return rc;
  const int blocks = (int)std::min<int64_t>(ceil_div(total, SH_NT), (int64_t)ncu * 16);
  const int64_t cells = (int64_t)meshsize[0] * meshsize[1] * meshsize[2];
  for (int q = 0; q < nshell; ++q) {
    MP_LAUNCH("bispec_shell_split", shell_split_ylm_kernel, dim3((unsigned)blocks), dim3(SH_NT), 0, s, (const float2*)A, pitch,
              meshsize[0], meshsize[1], nc, kx, ky, kz, thr, nedges, b0 + q, ell, m, scratch, sp);
    rc = irfft3d(s, scratch, sp, out + q * cells, meshsize, scratch, sp);
    if (rc) return rc;
  }
  return MP_OK;
}

// Real meshes of the Bessel filters of the Sugiyama three-point correlation function, one per
// separation s[q], q < ns: ifftn(A(k) j_ellj(|k| s[q]) [klo <= |k| < khi] Y_lm(k_ref)) =
// out_h[q] + i out_a[q] on the full complex grid (mesh3.py:383-395), A a half spectrum.  s: host
// array; kfun: the fundamental wavenumbers; l <= 4 (l = -1: no harmonic), ellj <= 5.
MP_EXPORT int mp_bessel_shells_ylm(void* stream, const void* A, int64_t pitch, const int32_t meshsize[3], const double kfun[3],
                                   const double* s, int ns, int ell, int m, int ellj, double klo, double khi, float* out_h,
                                   float* out_a, void* workspace, size_t workspace_bytes) {
  MP_REQUIRE(A && meshsize && kfun && s && out_h && out_a, "mp_bessel_shells_ylm: null pointer");
  MP_REQUIRE(ns >= 1, "mp_bessel_shells_ylm: no separation");
  MP_REQUIRE(ell == -1 || (ell >= 0 && ell <= MAX_ELL && m >= -ell && m <= ell), "mp_bessel_shells_ylm: (ell, m) = (%d, %d) not supported",
             ell, m);
  MP_REQUIRE(ellj >= 0 && ellj <= 5, "mp_bessel_shells_ylm: Bessel order %d not in 0..5", ellj);
  MP_REQUIRE(pitch >= meshsize[2] / 2 + 1, "mp_bessel_shells_ylm: pitch too small");
  hipStream_t st = (hipStream_t)stream;
  const int64_t sp = spectrum_pitch(meshsize[2]);
  Workspace ws(workspace, workspace_bytes);
  float2* scratch = ws.take<float2>((size_t)meshsize[0] * meshsize[1] * sp);
  if (!scratch) return set_error(MP_ERR_WORKSPACE, "mp_bessel_shells_ylm: workspace too small");
  const int nc = meshsize[2] / 2 + 1;
  const int64_t total = (int64_t)meshsize[0] * meshsize[1] * nc;
  int ncu = 0;
  int rc = multiprocessor_count(&ncu);
  if (rc) return rc;
  const int blocks = (int)std::min<int64_t>(ceil_div(total, SH_NT), (int64_t)ncu * 16);
  const int64_t cells = (int64_t)meshsize[0] * meshsize[1] * meshsize[2];
  for (int q = 0; q < ns; ++q) {
    for (int part = 0; part < 2; ++part) {
      MP_LAUNCH("bessel_shell_split", bessel_split_kernel, dim3((unsigned)blocks), dim3(SH_NT), 0, st, (const float2*)A, pitch, meshsize[0],
                meshsize[1], meshsize[2], kfun[0], kfun[1], kfun[2], ell, m, ellj, s[q], klo, khi, part, scratch, sp);
      rc = irfft3d(st, scratch, sp, (part ? out_a : out_h) + q * cells, meshsize, scratch, sp);
      if (rc) return rc;
    }
  }
  return MP_OK;
}

MP_EXPORT size_t mp_hankel_pairs_workspace_bytes(int npairs) {
  return align_up((size_t)HK_MAX_BLOCKS * (npairs > 0 ? npairs : 1) * sizeof(double), 256) + 256;
}

// out[p] = sum over the N0 N1 N2 cells s of F(s) j_l1(|s| k1[b1]) j_l2(|s| k2[b2]), (b1, b2) =
// pairs[2p .. 2p+1] (device int32), |s| from the device float64 separation tables sx, sy, sz,
// k1 (nk1) and k2 (nk2) device float64; l1, l2 <= 5; FP64 sums in a fixed order.
MP_EXPORT int mp_hankel_pairs(void* stream, const float* F, const int32_t meshsize[3], const double* sx, const double* sy,
                              const double* sz, int l1, const double* k1, int nk1, int l2, const double* k2, int nk2,
                              const int32_t* pairs, int npairs, double* out, void* workspace, size_t workspace_bytes) {
  MP_REQUIRE(F && meshsize && sx && sy && sz && k1 && k2 && pairs && out, "mp_hankel_pairs: null pointer");
  MP_REQUIRE(l1 >= 0 && l1 <= 5 && l2 >= 0 && l2 <= 5, "mp_hankel_pairs: (l1, l2) = (%d, %d) not in 0..5", l1, l2);
  MP_REQUIRE(nk1 >= 1 && nk2 >= 1 && nk1 + nk2 <= 1024, "mp_hankel_pairs: %d + %d wavenumbers", nk1, nk2);
  MP_REQUIRE(npairs >= 1 && npairs <= HK_NT * HK_PER, "mp_hankel_pairs: %d pairs not in 1..%d", npairs, HK_NT * HK_PER);
  hipStream_t s = (hipStream_t)stream;
  const int64_t ncells = (int64_t)meshsize[0] * meshsize[1] * meshsize[2];
  int ncu = 0;
  int rc = multiprocessor_count(&ncu);
  if (rc) return rc;
  const int blocks = (int)std::min<int64_t>(std::min<int64_t>(ceil_div(ncells, HK_TILE), (int64_t)ncu * 4), HK_MAX_BLOCKS);
  Workspace ws(workspace, workspace_bytes);
  double* partial = ws.take<double>((size_t)blocks * npairs);
  if (!partial) return set_error(MP_ERR_WORKSPACE, "mp_hankel_pairs: workspace too small");
  const size_t lds = (size_t)HK_TILE * (1 + nk1 + nk2) * sizeof(float);
  if (lds > 65536) MP_CHECK_HIP(hipFuncSetAttribute((const void*)hankel_pairs_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  MP_LAUNCH("hankel_pairs", hankel_pairs_kernel, dim3((unsigned)blocks), dim3(HK_NT), lds, s, F, meshsize[0], meshsize[1], meshsize[2], sx,
            sy, sz, l1, k1, nk1, l2, k2, nk2, (const int*)pairs, npairs, partial);
  return reduce_partials(s, partial, blocks, npairs, out);
}

MP_EXPORT size_t mp_bispec_triangles_workspace_bytes(int ntri) {
  return align_up((size_t)TR_MAX_BLOCKS * (ntri > 0 ? ntri : 1) * sizeof(double), 256) + 256;
}

// S[t] = sum over the ncells cells of s1[b1](x) s2[b2](x) s3[b3](x), (b1, b2, b3) = tri[3t ..
// 3t + 2] (a device array, sorted by (b1, b2); a triangle naming a shell beyond its stack
// sums to 0); shell q of stack s at s + q * ncells.  Stacks given by equal pointers share
// their staging.
MP_EXPORT int mp_bispec_triangles(void* stream, const float* s1, const float* s2, const float* s3, int64_t ncells, int nshell1,
                                  int nshell2, int nshell3, const int32_t* tri, int ntri, double* out, void* workspace,
                                  size_t workspace_bytes) {
  MP_REQUIRE(s1 && s2 && s3 && tri && out, "mp_bispec_triangles: null pointer");
  MP_REQUIRE(ncells >= 1 && ntri >= 1, "mp_bispec_triangles: empty input");
  MP_REQUIRE(nshell1 >= 1 && nshell2 >= 1 && nshell3 >= 1, "mp_bispec_triangles: no shells");
  hipStream_t s = (hipStream_t)stream;
  TriArgs a{};
  const float* ptrs[3] = {s1, s2, s3};
  const int ns[3] = {nshell1, nshell2, nshell3};
  int rows = 0;
  for (int r = 0; r < 3; ++r) {
    int found = -1;
    for (int q = 0; q < a.nstack; ++q)
      if (a.stack[q] == ptrs[r]) found = q;
    if (found >= 0) {
      MP_REQUIRE(a.nshell[found] == ns[r], "mp_bispec_triangles: one stack given with two shell counts");
    } else {
      found = a.nstack++;
      a.stack[found] = ptrs[r];
      a.nshell[found] = ns[r];
      a.lds_base[found] = rows;
      rows += ns[r];
    }
    a.role[r] = found;
  }
  if (rows > TR_MAX_SHELLS) return set_error(MP_ERR_UNSUPPORTED, "mp_bispec_triangles: %d shell rows > %d", rows, TR_MAX_SHELLS);
  a.shell_stride = ncells;
  a.ncells = ncells;
  a.tri = tri;
  a.ntri = ntri;
  Workspace ws(workspace, workspace_bytes);
  int ncu = 0;
  int rc = multiprocessor_count(&ncu);
  if (rc) return rc;
  const int groups = (int)ceil_div(ntri, TR_WAVES * TR_PER_WAVE);
  int bx = (int)std::min<int64_t>(ceil_div(ncells, TR_TILE), (int64_t)std::max(1, 2 * ncu / groups));
  bx = std::min(bx, TR_MAX_BLOCKS);
  a.partial = ws.take<double>((size_t)bx * ntri);
  if (!a.partial) return set_error(MP_ERR_WORKSPACE, "mp_bispec_triangles: workspace too small");
  const size_t lds = (size_t)rows * TR_TILE * sizeof(float);
  if (lds > 65536) MP_CHECK_HIP(hipFuncSetAttribute((const void*)triangle_sum_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  MP_LAUNCH("bispec_triangles", triangle_sum_kernel, dim3((unsigned)bx, (unsigned)groups), dim3(TR_NT), lds, s, a);
  return reduce_partials(s, a.partial, bx, ntri, out);
}
