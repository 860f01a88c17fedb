// Internal interfaces of the FFT passes and their fused prologues/epilogues.
#pragma once

#include "common.h"
#include "sph.h"

namespace mp {

constexpr int TW_N = 4096;  // twiddle table length (largest supported transform length is 2048)

enum { PRO_NONE = 0, PRO_YLM = 1, PRO_TOPHAT = 2 };

// Node coordinate o + i d formed as numpy forms arange(n) * d + o in float64 (the product
// rounded, then the sum: no fused multiply-add), then rounded to float.  A node that should be
// the origin can carry a residual such as -5.7e-14 this way (1000/122 * 61 - 500), which fixes
// its direction as the reference's float64 coordinates (MeshAttrs.rcoords, mesh.py:1573-1603) do.
__host__ __device__ __forceinline__ float node_coord(double o, double d, int i) {
#pragma clang fp contract(off)
  return (float)((double)i * d + o);
}

// Real-space prologue of the z pass: multiplies each real input by Y_lm(xhat) of the node
// position (local line of sight, jaxpower/mesh2.py:319-323), by the configuration-space shell
// of the window matrix, ttab[|n|^2] P_ell(n.los/|n|) with n the separation index vector
// (fftfreq order, equal cellsizes; jaxpower/mesh2.py:1078-1092), or by 1.
struct RealPrologue {
  int kind;  // PRO_NONE, PRO_YLM or PRO_TOPHAT
  int ell, m;
  int ny, nz;
  double x0, y0, z0;  // position of node (0,0,0): boxcenter - boxsize/2
  double dx, dy, dz;  // cellsize
  // PRO_TOPHAT
  int nx;
  const float* ttab;
  float lx, ly, lz;
  LegendreTable leg;
  __device__ __forceinline__ float eval(int64_t row, int z) const {
    int ix = (int)(row / ny), iy = (int)(row % ny);
    if (kind == PRO_TOPHAT) {
      const int fx = ix < (nx + 1) / 2 ? ix : ix - nx, fy = iy < (ny + 1) / 2 ? iy : iy - ny, fz = z < (nz + 1) / 2 ? z : z - nz;
      const int n2 = fx * fx + fy * fy + fz * fz;
      float v = ttab[n2];
      if (ell) {
        const float mu = n2 ? ((float)fx * lx + (float)fy * ly + (float)fz * lz) * rsqrtf((float)n2) : 0.f;
        v *= legendre_eval(leg, 0, mu, mu * mu);
      }
      return v;
    }
    return ylm_of_vector_rsq(ell, m, node_coord(x0, dx, ix), node_coord(y0, dy, iy), node_coord(z0, dz, z));
  }
};

// Fourier-space factor of the x pass: Y_lm(khat) with k from the half-spectrum indices.
struct KPrologue {
  int ell, m;
  int nx, ny;
  float kfx, kfy, kfz;  // fundamental wavenumbers
  __device__ __forceinline__ float eval(int ix, int iy, int iz) const {
    int mx = ix < (nx + 1) / 2 ? ix : ix - nx;
    int my = iy < (ny + 1) / 2 ? iy : iy - ny;
    return ylm_of_vector_rsq(ell, m, kfx * mx, kfy * my, kfz * iz);
  }
};

// Tables for the interlacing combine: per-axis phases exp(i s_j k_circ) for each shift j
// (row-major [j][index]) and per-axis window compensation; null pointers disable each part.
struct CombineTables {
  int stride;  // distance between the phase rows of consecutive shifts
  const float2* phx;
  const float2* phy;
  const float2* phz;
  const float* cx;
  const float* cy;
  const float* cz;
};

// Any-length passes (fft_any.hip), for meshes whose sizes are not all powers of two.
bool fft_any_shape_ok(const int32_t n[3]);
int any_col_fft(hipStream_t s, const float2* in, float2* out, int L, int ncols, int64_t nouter, int64_t fstride, int64_t ostride,
                int64_t out_fstride, int64_t out_ostride, bool inv, float scale, const char* name);
int any_zpass_r2c(hipStream_t s, const float* in, float2* out, int L, int64_t nrows, int64_t pitch, const RealPrologue& pro);
int any_zpass_c2r(hipStream_t s, const float2* in, int64_t pitch, float* out, int L, int64_t nrows, float scale);
int ylm_acc_any(hipStream_t s, const float2* t, int64_t tpitch, float2* acc, int64_t apitch, const int32_t n[3], const KPrologue& kp,
                int first);
// true when every axis is a power of two in [2, 2048] (the specialised passes apply)
bool pow2_shape(const int32_t n[3]);

int ensure_twiddles();
const float2* twiddle_device_ptr();
int64_t spectrum_pitch(int nz);
int rfft3d(hipStream_t s, const float* in, float2* out, const int32_t n[3], int64_t pitch, const RealPrologue& pro);
// z and y passes only (the x pass is fused into a consumer).
int rfft3d_zy(hipStream_t s, const float* in, float2* out, const int32_t n[3], int64_t pitch, const RealPrologue& pro);
int ypass(hipStream_t s, float2* data, const int32_t n[3], int64_t pitch);
// z and y passes of rho * Y_lm(xhat) for every m of pro.ell (pro.kind = PRO_YLM): half spectrum
// of m at Z + (m + ell) * mesh_stride.
int rfft_zy_ylm(hipStream_t s, const float* in, float2* Z, int64_t mesh_stride, const int32_t n[3], int64_t pitch,
                const RealPrologue& pro);
// y pass of nmesh half spectra at data + m * mesh_stride, in one pipelined launch.
int ypass_meshes(hipStream_t s, float2* data, int nmesh, int64_t mesh_stride, const int32_t n[3], int64_t pitch);
// x pass of the interlaced paint (xpass.hip): combine of the n_i partial spectra after their z
// and y passes (mesh j at partial + j * n0 * n1 * pitch), written out or binned.
struct BinGeom;
bool xpass_supported(const int32_t n[3]);
size_t combine_partial_workspace_bytes(const int32_t n[3], int ni);
int combine_partial(hipStream_t s, const float2* partial, int64_t pitch, const int32_t n[3], int order, int ni, int compensate,
                    int zero_dc, float2* out, int64_t out_pitch, void* ws, size_t ws_bytes, int power = 0);
size_t xpass_tables_bytes(const int32_t n[3], int ni);  // the phase and compensation tables of one combine
// With nmock > 1, partial holds the ni partial spectra of nmock catalogues one after the other
// and out their nmock sums, catalogue-major (one launch, one grid row per catalogue).
size_t bin_partial_workspace_bytes(const int32_t n[3], int ni, int nprod, int nedges, int nmock = 1);
int bin_partial(hipStream_t s, const float2* partial, int64_t pitch, const int32_t n[3], int order, int ni, int compensate,
                int zero_dc, const BinGeom& g, const double los[3], int nprod, const int32_t* ell, const int32_t* remove_zero,
                double* out, void* ws, size_t ws_bytes, int linear = 0, const int32_t* ylm_m = nullptr,
                bool fit_partials = false, int nmock = 1);
// Local line of sight for one ell, fused (xpass.hip): z/y passes of rho Y_lm(xhat) for all m,
// then the x pass forms C = sum_m FFT(.) Y_lm(khat) per tile and bins conj(C) A0 w (and |A0|^2 w
// when with0) without writing C; out[p][b][re, im] as mp_bin_poles writes it.
size_t ylm_bin_workspace_bytes(const int32_t n[3], int ell, int nedges);
int ylm_bin(hipStream_t s, const float* rmesh, const float2* A0, int64_t a0_pitch, const int32_t n[3], const double boxsize[3],
            const double boxcenter[3], int ell, int with0, const BinGeom& g, double* out, void* ws, size_t ws_bytes);
int ylm_acc_x(hipStream_t s, const float2* partial, float2* acc, const int32_t n[3], int64_t in_pitch, int64_t acc_pitch,
              const KPrologue& kp, int first);
int irfft3d(hipStream_t s, const float2* in, int64_t in_pitch, float* out, const int32_t n[3], float2* scratch,
            int64_t scratch_pitch);
// |k| shell of the window matrix (jaxpower/mesh2.py:1097-1098): the half-spectrum value
// [t_lo <= k^2 < t_hi] P_ell(k.los/|k|) scale at mode (ix, iy, iz), k^2 summed in float64 as the
// binning does (so shell membership is digitize of the same edges), mu = 0 at k = 0.
struct ShellGen {
  const double* kx;
  const double* ky;
  const double* kz;
  double t_lo, t_hi;
  double los[3];
  LegendreTable leg;
  int ell_on;
  float scale;
  int ylm_on, ylm_ell, ylm_m;  // weight Y_lm(khat) instead of P_ell(mu) (local line of sight)
  int ylm2_on, ylm2_ell, ylm2_m;  // and a second factor Y_l2,m2(-khat) ('local' theory line of sight)
  // tabulated radial theory instead of a shell (tn > 0): tvals[#(tkeys <= k^2) - 1] when that
  // index is in [0, tn), else 0 (digitize on the squared-edge thresholds, or an exact lookup in
  // the sorted distinct k^2 of the mesh)
  const double* tkeys;
  const float* tvals;
  int tnkeys, tn;
  __device__ __forceinline__ float eval(int ix, int iy, int iz) const {
    const double kxv = kx[ix], kyv = ky[iy], kzv = kz[iz];
    double k2;
    {
#pragma clang fp contract(off)
      k2 = (kxv * kxv + kyv * kyv) + kzv * kzv;
    }
    float v = scale;
    if (tn > 0) {
      int lo = 0, hi = tnkeys;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (tkeys[mid] <= k2) lo = mid + 1;
        else hi = mid;
      }
      if (lo < 1 || lo > tn) return 0.f;
      v *= tvals[lo - 1];
    } else if (!(k2 >= t_lo && k2 < t_hi)) {
      return 0.f;
    }
    if (ylm_on) {
      v *= ylm_of_vector(ylm_ell, ylm_m, (float)kxv, (float)kyv, (float)kzv);
      if (ylm2_on) v *= ylm_of_vector(ylm2_ell, ylm2_m, -(float)kxv, -(float)kyv, -(float)kzv);
    }
    if (ell_on) {
      const double mu = k2 > 0. ? (kxv * los[0] + kyv * los[1] + kzv * los[2]) / sqrt(k2) : 0.;
      const float m = (float)mu;
      v *= legendre_eval(leg, 0, m, m * m);
    }
    return v;
  }
};
// data <- the z and y passes of r2c(Q * scale * irfftn(S)), S the shell of g, with the inverse
// x pass generating S instead of reading it; the forward x pass is left to a consumer.
int convolve_shell_zy(hipStream_t s, const ShellGen& g, float2* data, int64_t pitch, const float* Q, const int32_t n[3], float scale);
// out = irfftn(S), S the shell of g, with the inverse x pass generating S (scratch: a half
// spectrum of row pitch scratch_pitch).
int irfft3d_shell(hipStream_t s, const ShellGen& g, float* out, const int32_t n[3], float2* scratch, int64_t scratch_pitch);
// In place on a half spectrum: data <- r2c(Q * scale * irfftn(data)), Q a real (N0, N1, N2)
// mesh; the real mesh is formed and multiplied inside the z pass only.
int convolve_real(hipStream_t s, float2* data, int64_t pitch, const float* Q, const int32_t n[3], float scale);
struct RealBinArgs;
size_t irfft3d_bin_real_workspace_bytes(const int32_t n[3], int nell, int nedges);
int irfft3d_bin_real(hipStream_t s, const float2* in, int64_t in_pitch, const int32_t n[3], const RealBinArgs& a, double* out,
                     void* ws, size_t ws_bytes, int x_done = 0);

}  // namespace mp
