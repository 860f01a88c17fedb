// Window matrix of the power-spectrum multipoles with a global line of sight
// (jaxpower/mesh2.py:921-1118, compute_mesh2_spectrum_window without flags): for each input
// multipole ellin and input |k| shell [e0, e1), the mesh-defined theory
//   A(k) = [e0 <= |k| < e1] P_ellin(k.los/|k|) scale                    (mesh2.py:1097-1098)
// is, when a selection Q is given, replaced by B = r2c(Q irfftn(A)) (mesh2.py:1099, with
// Q = c2r(A1 conj(A2)) / N^3 from mesh2.py:1022-1029), and binned with the output Legendre
// weights (mesh2.py:1037-1045).
//
// mp_window_bin does one column in four launches: the inverse x pass generates A from the mode
// indices instead of reading it, the inverse y pass runs in place, one z pass goes back to real
// space, multiplies by the row of Q and transforms forward again without writing the real row,
// the forward y pass runs in place, and the forward x pass bins Re(B) P_ell in FP64 without
// writing B.  mp_window_shell writes A (periodic box) or B as a half spectrum instead.
//
// With an infinite-resolution theory (flags=('infinite',), mesh2.py:1078-1092) each column is
// the real mesh xi_ellin(|s|) P_ellin(mu_s) V (times Q), transformed and binned:
// mp_window_bin_tophat tabulates xi per integer |n|^2 (equal cellsizes: |s| = cell |n|), the
// z pass generates the real rows from the table and multiplies them by Q as it loads them, then
// the y pass and the binning x pass follow.
#include "bessel.h"
#include "bin.h"
#include "common.h"
#include "fft.h"
#include "sph.h"

namespace mp {

constexpr int WS_NT = 256;

__global__ void __launch_bounds__(WS_NT)
window_shell_kernel(ShellGen g, int n0, int n1, int nc, float2* __restrict__ out, int64_t pitch) {
  const int64_t total = (int64_t)n0 * n1 * nc;
  for (int64_t i = (int64_t)blockIdx.x * WS_NT + threadIdx.x; i < total; i += (int64_t)gridDim.x * WS_NT) {
    const int iz = (int)(i % nc);
    const int64_t row = i / nc;
    const int iy = (int)(row % n1), ix = (int)(row / n1);
    out[row * pitch + iz] = make_float2(g.eval(ix, iy, iz), 0.f);
  }
}

static ShellGen make_shell(const double* kx, const double* ky, const double* kz, double k2_lo, double k2_hi, int ellin,
                           const double los[3], float scale) {
  ShellGen g{};
  g.kx = kx;
  g.ky = ky;
  g.kz = kz;
  g.t_lo = k2_lo;
  g.t_hi = k2_hi;
  for (int q = 0; q < 3; ++q) g.los[q] = ellin ? los[q] : 0.;
  legendre_table_set(g.leg, 0, ellin);
  g.ell_on = ellin > 0;
  g.scale = scale;
  return g;
}

}  // namespace mp

using namespace mp;

// One column of the window matrix before binning: out = A (no selection, Q null) or
// r2c(Q irfftn(A)) with A = [t_lo <= k^2 < t_hi] P_ellin(k.los/|k|) scale.
MP_EXPORT int mp_window_shell(void* stream, const float* Q, const int32_t meshsize[3], const double* kx, const double* ky,
                              const double* kz, double k2_lo, double k2_hi, int ellin, const double los[3], float scale,
                              void* out, int64_t out_pitch) {
  MP_REQUIRE(meshsize && kx && ky && kz && out, "mp_window_shell: null pointer");
  MP_REQUIRE(ellin >= 0 && ellin <= 5, "mp_window_shell: ellin %d not in 0..5", ellin);
  MP_REQUIRE(ellin == 0 || los, "mp_window_shell: ellin > 0 needs a line of sight");
  MP_REQUIRE(meshsize[0] >= 1 && meshsize[1] >= 1 && meshsize[2] >= 2, "mp_window_shell: bad meshsize");
  const int nc = meshsize[2] / 2 + 1;
  MP_REQUIRE(out_pitch >= nc, "mp_window_shell: out_pitch %lld < Nz/2+1", (long long)out_pitch);
  hipStream_t s = (hipStream_t)stream;
  const ShellGen g = make_shell(kx, ky, kz, k2_lo, k2_hi, ellin, los, scale);
  const int64_t total = (int64_t)meshsize[0] * meshsize[1] * nc;
  int dev = 0, ncu = 0;
  MP_CHECK_HIP(hipGetDevice(&dev));
  MP_CHECK_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const int blocks = (int)std::min<int64_t>(ceil_div(total, WS_NT), (int64_t)ncu * 16);
  float2* o = static_cast<float2*>(out);
  MP_LAUNCH("window_shell", window_shell_kernel, dim3((unsigned)blocks), dim3(WS_NT), 0, s, g, meshsize[0], meshsize[1], nc, o,
            out_pitch);
  if (!Q) return MP_OK;
  return convolve_real(s, o, out_pitch, Q, meshsize, 1.f);
}

namespace mp {

constexpr int TT_NT = 256;

// ttab[n2] = xi_ell(cell sqrt(n2)) scale for n2 = 0 .. n2max (the separation shell of the
// input bin [k0, k1)).
__global__ void __launch_bounds__(TT_NT)
tophat_table_kernel(int ell, double k0, double k1, double cell, double scale, int n2max, float* __restrict__ ttab) {
  const int i = blockIdx.x * TT_NT + threadIdx.x;
  if (i > n2max) return;
  ttab[i] = (float)(tophat_shell_xi(ell, k0, k1, cell * sqrt((double)i)) * scale);
}

static int64_t tophat_n2max(const int32_t n[3]) {
  int64_t m = 0;
  for (int a = 0; a < 3; ++a) m += (int64_t)(n[a] / 2) * (n[a] / 2);
  return m;
}

}  // namespace mp

// Host evaluation of the configuration-space shell xi_ell(s) of [k0, k1) used by
// mp_window_bin_tophat (the BesselIntegral weights of jaxpower/utils.py:470-501).
MP_EXPORT double mp_tophat_shell_xi(int ell, double k0, double k1, double s) {
  if (ell < 0 || ell > 5) return NAN;
  return tophat_shell_xi(ell, k0, k1, s);
}

// Host arrays: out[i] = xi_ell(s[i]) of the shell [k0, k1) (as mp_tophat_shell_xi), and
// out[i] = j_ell(x[i]) (jaxpower/utils.py:418-450), ell in 0..5; the Hankel sums of the smooth
// window (jaxpower/mesh2.py:1057-1068) are formed from these on the host.
MP_EXPORT int mp_tophat_shell_xi_array(int ell, double k0, double k1, const double* s, int64_t n, double* out) {
  MP_REQUIRE(ell >= 0 && ell <= 5, "mp_tophat_shell_xi_array: ell %d not in 0..5", ell);
  MP_REQUIRE(n >= 0 && (n == 0 || (s && out)), "mp_tophat_shell_xi_array: null pointer");
  for (int64_t i = 0; i < n; ++i) out[i] = tophat_shell_xi(ell, k0, k1, s[i]);
  return MP_OK;
}

MP_EXPORT int mp_spherical_jn_array(int ell, const double* x, int64_t n, double* out) {
  MP_REQUIRE(ell >= 0 && ell <= 5, "mp_spherical_jn_array: ell %d not in 0..5", ell);
  MP_REQUIRE(n >= 0 && (n == 0 || (x && out)), "mp_spherical_jn_array: null pointer");
  for (int64_t i = 0; i < n; ++i) out[i] = spherical_jn(ell, x[i]);
  return MP_OK;
}

MP_EXPORT int64_t mp_window_tophat_workspace_bytes(const int32_t meshsize[3], int nell, int nedges) {
  if (!meshsize || nell < 1 || nedges < 2) return 0;
  const size_t spec = align_up((size_t)meshsize[0] * meshsize[1] * spectrum_pitch(meshsize[2]) * sizeof(float2), 256);
  const size_t tab = align_up((size_t)(tophat_n2max(meshsize) + 1) * sizeof(float), 256);
  return (int64_t)(spec + tab + bin_partial_workspace_bytes(meshsize, 1, nell, nedges) + 512);
}

// Columns of the window matrix with an infinite-resolution theory: column c is the input shell
// [k0[c], k1[c]) of multipole ellin[c] (three HOST arrays; edges, not thresholds), taken to
// configuration space as xi_ellin(|s|) P_ellin(s.los/|s|) scale (times the real mesh Q when not
// NULL), then r2c and binned as in mp_window_bin: out[c][p][b][re, im].  Equal cellsizes on the
// three axes, power-of-two meshsizes.
MP_EXPORT int mp_window_bin_tophat(void* stream, const float* Q, const int32_t meshsize[3], double cellsize, const double* kx,
                                   const double* ky, const double* kz, const double* k2_thresholds, int nedges, int ncol,
                                   const double* k0, const double* k1, const int32_t* ellin, const double los[3], double scale,
                                   int nell, const int32_t* ells, double* out, void* workspace, size_t workspace_bytes) {
  MP_REQUIRE(meshsize && kx && ky && kz && k2_thresholds && ells && out && los && k0 && k1 && ellin, "mp_window_bin_tophat: null pointer");
  MP_REQUIRE(ncol >= 0 && nell >= 1 && nell <= 6 && nedges >= 2, "mp_window_bin_tophat: bad sizes");
  MP_REQUIRE(cellsize > 0., "mp_window_bin_tophat: cellsize %g <= 0", cellsize);
  for (int c = 0; c < ncol; ++c) MP_REQUIRE(ellin[c] >= 0 && ellin[c] <= 5, "mp_window_bin_tophat: ellin %d not in 0..5", ellin[c]);
  if (!xpass_supported(meshsize))
    return set_error(MP_ERR_UNSUPPORTED, "mp_window_bin_tophat: meshsize (%d, %d, %d) not supported", meshsize[0], meshsize[1],
                     meshsize[2]);
  hipStream_t s = (hipStream_t)stream;
  const int64_t sp = spectrum_pitch(meshsize[2]);
  const int64_t n2max = tophat_n2max(meshsize);
  Workspace ws(workspace, workspace_bytes);
  float2* spec = ws.take<float2>((size_t)meshsize[0] * meshsize[1] * sp);
  float* ttab = ws.take<float>((size_t)n2max + 1);
  const size_t rest = workspace_bytes > ws.used ? workspace_bytes - ws.used : 0;
  if (!spec || !ttab || rest < bin_partial_workspace_bytes(meshsize, 1, nell, nedges))
    return set_error(MP_ERR_WORKSPACE, "mp_window_bin_tophat: workspace of %zu bytes < %lld", workspace_bytes,
                     (long long)mp_window_tophat_workspace_bytes(meshsize, nell, nedges));
  BinGeom geo{};
  geo.n0 = meshsize[0];
  geo.n1 = meshsize[1];
  geo.ncols = meshsize[2] / 2 + 1;
  geo.nedges = nedges;
  geo.kx = kx;
  geo.ky = ky;
  geo.kz = kz;
  geo.thr = k2_thresholds;
  int32_t rz[6] = {1, 1, 1, 1, 1, 1};
  char* base = static_cast<char*>(workspace) + ws.used;
  const int64_t per_col = (int64_t)nell * (nedges - 1) * 2;
  for (int c = 0; c < ncol; ++c) {
    MP_LAUNCH("window_tophat_table", tophat_table_kernel, dim3((unsigned)ceil_div(n2max + 1, TT_NT)), dim3(TT_NT), 0, s, ellin[c],
              k0[c], k1[c], cellsize, scale, (int)n2max, ttab);
    RealPrologue pro{};
    pro.kind = PRO_TOPHAT;
    pro.ell = ellin[c];
    pro.nx = meshsize[0];
    pro.ny = meshsize[1];
    pro.nz = meshsize[2];
    pro.ttab = ttab;
    pro.lx = (float)los[0];
    pro.ly = (float)los[1];
    pro.lz = (float)los[2];
    legendre_table_set(pro.leg, 0, ellin[c]);
    int rc = rfft3d_zy(s, Q, spec, meshsize, sp, pro);
    if (rc) return rc;
    rc = bin_partial(s, spec, sp, meshsize, 0, 1, 0, 0, geo, los, nell, ells, rz, out + c * per_col, base, rest, 1);
    if (rc) return rc;
  }
  return MP_OK;
}

MP_EXPORT int64_t mp_window_bin_workspace_bytes(const int32_t meshsize[3], int nell, int nedges) {
  if (!meshsize || nell < 1 || nedges < 2) return 0;
  const size_t spec = align_up((size_t)meshsize[0] * meshsize[1] * spectrum_pitch(meshsize[2]) * sizeof(float2), 256);
  return (int64_t)(spec + bin_partial_workspace_bytes(meshsize, 1, nell, nedges) + 256);
}

// Columns of the window matrix, binned: for column c (input shell k2_lo[c] <= k^2 < k2_hi[c] of
// multipole ellin[c], three host arrays), out[c][p][b][re, im] = sum over the modes of bin b
// (mode k = 0 excluded) of w Re(B) P_ell[p](k.los/|k|), B = r2c(Q irfftn(A)) and A the shell of
// mp_window_shell.  Per column the shell is generated inside the inverse x pass, and the
// forward x pass bins B without writing it (the imaginary slots are 0).
MP_EXPORT int mp_window_bin(void* stream, const float* Q, const int32_t meshsize[3], const double* kx, const double* ky,
                            const double* kz, const double* k2_thresholds, int nedges, int ncol, const double* k2_lo,
                            const double* k2_hi, const int32_t* ellin, const double los[3], float scale, int nell,
                            const int32_t* ells, double* out, void* workspace, size_t workspace_bytes) {
  MP_REQUIRE(Q && meshsize && kx && ky && kz && k2_thresholds && ells && out && los && k2_lo && k2_hi && ellin,
             "mp_window_bin: null pointer");
  MP_REQUIRE(ncol >= 0, "mp_window_bin: %d columns", ncol);
  MP_REQUIRE(nell >= 1 && nell <= 6, "mp_window_bin: %d ells not in 1..6", nell);
  MP_REQUIRE(nedges >= 2, "mp_window_bin: need at least two edges");
  for (int c = 0; c < ncol; ++c) MP_REQUIRE(ellin[c] >= 0 && ellin[c] <= 5, "mp_window_bin: ellin %d not in 0..5", ellin[c]);
  if (!xpass_supported(meshsize))
    return set_error(MP_ERR_UNSUPPORTED, "mp_window_bin: meshsize (%d, %d, %d) not supported", meshsize[0], meshsize[1], meshsize[2]);
  hipStream_t s = (hipStream_t)stream;
  const int64_t sp = spectrum_pitch(meshsize[2]);
  Workspace ws(workspace, workspace_bytes);
  float2* spec = ws.take<float2>((size_t)meshsize[0] * meshsize[1] * sp);
  const size_t rest = workspace_bytes > ws.used ? workspace_bytes - ws.used : 0;
  if (!spec || rest < bin_partial_workspace_bytes(meshsize, 1, nell, nedges))
    return set_error(MP_ERR_WORKSPACE, "mp_window_bin: workspace of %zu bytes < %lld", workspace_bytes,
                     (long long)mp_window_bin_workspace_bytes(meshsize, nell, nedges));
  BinGeom geo{};
  geo.n0 = meshsize[0];
  geo.n1 = meshsize[1];
  geo.ncols = meshsize[2] / 2 + 1;
  geo.nedges = nedges;
  geo.kx = kx;
  geo.ky = ky;
  geo.kz = kz;
  geo.thr = k2_thresholds;
  int32_t rz[6] = {1, 1, 1, 1, 1, 1};
  char* base = static_cast<char*>(workspace) + ws.used;
  const int64_t per_col = (int64_t)nell * (nedges - 1) * 2;
  for (int c = 0; c < ncol; ++c) {
    const ShellGen g = make_shell(kx, ky, kz, k2_lo[c], k2_hi[c], ellin[c], los, scale);
    int rc = convolve_shell_zy(s, g, spec, sp, Q, meshsize, 1.f);
    if (rc) return rc;
    rc = bin_partial(s, spec, sp, meshsize, 0, 1, 0, 0, geo, los, nell, ells, rz, out + c * per_col, base, rest, 1);
    if (rc) return rc;
  }
  return MP_OK;
}

// Local (firstpoint) line of sight with an infinite-resolution theory (jaxpower/mesh2.py:1235-1260,
// wide-angle order 0), one output (ell, m) and one input multipole ellin: for column c (input
// shell [k0[c], k1[c]), HOST arrays), out[c][b][re, im] = sum over the modes of bin b (mode
// k = 0 excluded when remove_zero) of w Re(B) Y_ell,m(khat), B = r2c(xi_ellin(|s|) scale Qs),
// Qs the real (N0, N1, N2) selection term of this (ell, m) and ellin.  Equal cellsizes,
// power-of-two meshsizes, ell <= 4.
MP_EXPORT int mp_window_bin_tophat_local(void* stream, const float* Qs, const int32_t meshsize[3], double cellsize, const double* kx,
                                         const double* ky, const double* kz, const double* k2_thresholds, int nedges, int ncol,
                                         const double* k0, const double* k1, int ellin, int ell, int m, int remove_zero,
                                         double scale, double* out, void* workspace, size_t workspace_bytes) {
  MP_REQUIRE(Qs && meshsize && kx && ky && kz && k2_thresholds && out && k0 && k1, "mp_window_bin_tophat_local: null pointer");
  MP_REQUIRE(ncol >= 0 && nedges >= 2, "mp_window_bin_tophat_local: bad sizes");
  MP_REQUIRE(cellsize > 0., "mp_window_bin_tophat_local: cellsize %g <= 0", cellsize);
  MP_REQUIRE(ellin >= 0 && ellin <= 5, "mp_window_bin_tophat_local: ellin %d not in 0..5", ellin);
  MP_REQUIRE(ell >= 0 && ell <= 4 && m >= -ell && m <= ell, "mp_window_bin_tophat_local: (ell, m) = (%d, %d) not supported", ell, m);
  if (!xpass_supported(meshsize))
    return set_error(MP_ERR_UNSUPPORTED, "mp_window_bin_tophat_local: meshsize (%d, %d, %d) not supported", meshsize[0], meshsize[1],
                     meshsize[2]);
  hipStream_t s = (hipStream_t)stream;
  const int64_t sp = spectrum_pitch(meshsize[2]);
  const int64_t n2max = tophat_n2max(meshsize);
  Workspace ws(workspace, workspace_bytes);
  float2* spec = ws.take<float2>((size_t)meshsize[0] * meshsize[1] * sp);
  float* ttab = ws.take<float>((size_t)n2max + 1);
  const size_t rest = workspace_bytes > ws.used ? workspace_bytes - ws.used : 0;
  if (!spec || !ttab || rest < bin_partial_workspace_bytes(meshsize, 1, 1, nedges))
    return set_error(MP_ERR_WORKSPACE, "mp_window_bin_tophat_local: workspace of %zu bytes < %lld", workspace_bytes,
                     (long long)mp_window_tophat_workspace_bytes(meshsize, 1, nedges));
  BinGeom geo{};
  geo.n0 = meshsize[0];
  geo.n1 = meshsize[1];
  geo.ncols = meshsize[2] / 2 + 1;
  geo.nedges = nedges;
  geo.kx = kx;
  geo.ky = ky;
  geo.kz = kz;
  geo.thr = k2_thresholds;
  const int32_t ells[1] = {ell}, ms[1] = {m}, rz[1] = {remove_zero ? 1 : 0};
  const double zero_los[3] = {0., 0., 0.};
  char* base = static_cast<char*>(workspace) + ws.used;
  const int64_t per_col = (int64_t)(nedges - 1) * 2;
  for (int c = 0; c < ncol; ++c) {
    MP_LAUNCH("window_tophat_table", tophat_table_kernel, dim3((unsigned)ceil_div(n2max + 1, TT_NT)), dim3(TT_NT), 0, s, ellin, k0[c],
              k1[c], cellsize, scale, (int)n2max, ttab);
    RealPrologue pro{};
    pro.kind = PRO_TOPHAT;
    pro.ell = 0;  // isotropic: the angular part sits in Qs
    pro.nx = meshsize[0];
    pro.ny = meshsize[1];
    pro.nz = meshsize[2];
    pro.ttab = ttab;
    legendre_table_set(pro.leg, 0, 0);
    int rc = rfft3d_zy(s, Qs, spec, meshsize, sp, pro);
    if (rc) return rc;
    rc = bin_partial(s, spec, sp, meshsize, 0, 1, 0, 0, geo, zero_los, 1, ells, rz, out + c * per_col, base, rest, 1, ms);
    if (rc) return rc;
  }
  return MP_OK;
}

namespace mp {

constexpr int MA_NT = 256;

__global__ void __launch_bounds__(MA_NT)
mul_acc_kernel(const float4* __restrict__ a, const float4* __restrict__ b, float4* __restrict__ out, int64_t n4, int accumulate) {
  for (int64_t i = (int64_t)blockIdx.x * MA_NT + threadIdx.x; i < n4; i += (int64_t)gridDim.x * MA_NT) {
    const float4 x = a[i], y = b[i];
    float4 v = make_float4(x.x * y.x, x.y * y.y, x.z * y.z, x.w * y.w);
    if (accumulate) {
      const float4 o = out[i];
      v = make_float4(o.x + v.x, o.y + v.y, o.z + v.z, o.w + v.w);
    }
    out[i] = v;
  }
}

__global__ void __launch_bounds__(MA_NT)
mul_acc_tail_kernel(const float* __restrict__ a, const float* __restrict__ b, float* __restrict__ out, int64_t start, int64_t n,
                    int accumulate) {
  const int64_t i = start + (int64_t)blockIdx.x * MA_NT + threadIdx.x;
  if (i < n) out[i] = accumulate ? out[i] + a[i] * b[i] : a[i] * b[i];
}

}  // namespace mp

// The real-space products of the local window with the theory on the mesh
// (jaxpower/mesh2.py:1277-1281, Aell += xi * Q): out = a * b, or out += a * b when accumulate.
// Device arrays of n floats, 16-byte aligned.
MP_EXPORT int mp_mul_acc(void* stream, const float* a, const float* b, float* out, int64_t n, int accumulate) {
  MP_REQUIRE(a && b && out && n >= 0, "mp_mul_acc: null pointer");
  MP_REQUIRE(((uintptr_t)a | (uintptr_t)b | (uintptr_t)out) % 16 == 0, "mp_mul_acc: arrays must be 16-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  const int64_t n4 = n / 4;
  int dev = 0, ncu = 0;
  MP_CHECK_HIP(hipGetDevice(&dev));
  MP_CHECK_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  if (n4 > 0) {
    const int blocks = (int)std::min<int64_t>(ceil_div(n4, MA_NT), (int64_t)ncu * 8);
    MP_LAUNCH("window_mul_acc", mul_acc_kernel, dim3((unsigned)blocks), dim3(MA_NT), 0, s, (const float4*)a, (const float4*)b,
              (float4*)out, n4, accumulate ? 1 : 0);
  }
  if (n > 4 * n4)
    MP_LAUNCH("window_mul_acc", mul_acc_tail_kernel, dim3(1), dim3(MA_NT), 0, s, a, b, out, 4 * n4, n, accumulate ? 1 : 0);
  return MP_OK;
}

MP_EXPORT int64_t mp_window_shell_ylm_workspace_bytes(const int32_t meshsize[3]) {
  if (!meshsize) return 0;
  return (int64_t)(align_up((size_t)meshsize[0] * meshsize[1] * spectrum_pitch(meshsize[2]) * sizeof(float2), 256) + 256);
}

// The configuration-space shell of the local window with the theory on the mesh
// (jaxpower/mesh2.py:1276): out = c2r([k2_lo <= k^2 < k2_hi] Y_ell,m(khat) scale), a real
// (N0, N1, N2) mesh; the shell is generated inside the inverse x pass.  ell <= 4.
MP_EXPORT int mp_window_shell_ylm(void* stream, const int32_t meshsize[3], const double* kx, const double* ky, const double* kz,
                                  double k2_lo, double k2_hi, int ell, int m, float scale, float* out, void* workspace,
                                  size_t workspace_bytes) {
  MP_REQUIRE(meshsize && kx && ky && kz && out, "mp_window_shell_ylm: null pointer");
  MP_REQUIRE(ell >= 0 && ell <= 4 && m >= -ell && m <= ell, "mp_window_shell_ylm: (ell, m) = (%d, %d) not supported", ell, m);
  const int64_t sp = spectrum_pitch(meshsize[2]);
  Workspace ws(workspace, workspace_bytes);
  float2* scratch = ws.take<float2>((size_t)meshsize[0] * meshsize[1] * sp);
  if (!scratch) return set_error(MP_ERR_WORKSPACE, "mp_window_shell_ylm: workspace too small");
  ShellGen g = make_shell(kx, ky, kz, k2_lo, k2_hi, 0, nullptr, scale);
  g.ylm_on = 1;
  g.ylm_ell = ell;
  g.ylm_m = m;
  return irfft3d_shell((hipStream_t)stream, g, out, meshsize, scratch, sp);
}

MP_EXPORT int64_t mp_bin_ylm_real_workspace_bytes(const int32_t meshsize[3], int nedges) {
  if (!meshsize || nedges < 2) return 0;
  const size_t spec = align_up((size_t)meshsize[0] * meshsize[1] * spectrum_pitch(meshsize[2]) * sizeof(float2), 256);
  return (int64_t)(spec + bin_partial_workspace_bytes(meshsize, 1, 1, nedges) + 256);
}

// out[b][re, im] = sum over the modes of bin b (k = 0 excluded when remove_zero) of
// w Re(r2c(mesh)) Y_ell,m(khat): one output (ell, m) term of the local window with the theory on
// the mesh (jaxpower/mesh2.py:1283-1285); the half spectrum is binned in its x pass, not written.
MP_EXPORT int mp_bin_ylm_real(void* stream, const float* mesh, const int32_t meshsize[3], const double* kx, const double* ky,
                              const double* kz, const double* k2_thresholds, int nedges, int ell, int m, int remove_zero, double* out,
                              void* workspace, size_t workspace_bytes) {
  MP_REQUIRE(mesh && meshsize && kx && ky && kz && k2_thresholds && out, "mp_bin_ylm_real: null pointer");
  MP_REQUIRE(ell >= 0 && ell <= 4 && m >= -ell && m <= ell, "mp_bin_ylm_real: (ell, m) = (%d, %d) not supported", ell, m);
  MP_REQUIRE(nedges >= 2, "mp_bin_ylm_real: need at least two edges");
  if (!xpass_supported(meshsize))
    return set_error(MP_ERR_UNSUPPORTED, "mp_bin_ylm_real: meshsize (%d, %d, %d) not supported", meshsize[0], meshsize[1], meshsize[2]);
  hipStream_t s = (hipStream_t)stream;
  const int64_t sp = spectrum_pitch(meshsize[2]);
  Workspace ws(workspace, workspace_bytes);
  float2* spec = ws.take<float2>((size_t)meshsize[0] * meshsize[1] * sp);
  const size_t rest = workspace_bytes > ws.used ? workspace_bytes - ws.used : 0;
  if (!spec || rest < bin_partial_workspace_bytes(meshsize, 1, 1, nedges))
    return set_error(MP_ERR_WORKSPACE, "mp_bin_ylm_real: workspace of %zu bytes < %lld", workspace_bytes,
                     (long long)mp_bin_ylm_real_workspace_bytes(meshsize, nedges));
  RealPrologue pro{};
  pro.kind = PRO_NONE;
  int rc = rfft3d_zy(s, mesh, spec, meshsize, sp, pro);
  if (rc) return rc;
  BinGeom geo{};
  geo.n0 = meshsize[0];
  geo.n1 = meshsize[1];
  geo.ncols = meshsize[2] / 2 + 1;
  geo.nedges = nedges;
  geo.kx = kx;
  geo.ky = ky;
  geo.kz = kz;
  geo.thr = k2_thresholds;
  const int32_t ells[1] = {ell}, ms[1] = {m}, rz[1] = {remove_zero ? 1 : 0};
  const double zero_los[3] = {0., 0., 0.};
  char* base = static_cast<char*>(workspace) + ws.used;
  return bin_partial(s, spec, sp, meshsize, 0, 1, 0, 0, geo, zero_los, 1, ells, rz, out, base, rest, 1, ms);
}

// The mean spectrum of a theory through a selection (jaxpower/mesh2.py:1485-1680), global line of
// sight, one input multipole: the theory P(k) = tvals[#(tkeys <= k^2) - 1] (0 outside
// [0, ntab)) times P_ellin(k.los/|k|) scale is convolved as r2c(Q irfftn(.)) and binned:
// out[p][b][re, im] = sum over the modes of bin b (k = 0 included) of w Re(B) P_ell[p](mu).
// tkeys (nkeys float64) and tvals (ntab float32) are device arrays; Q a real mesh.
MP_EXPORT int mp_spectrum_mean_bin(void* stream, const float* Q, const int32_t meshsize[3], const double* kx, const double* ky,
                                   const double* kz, const double* k2_thresholds, int nedges, int nkeys, const double* tkeys, int ntab,
                                   const float* tvals, int ellin, const double los[3], float scale, int nell, const int32_t* ells,
                                   double* out, void* workspace, size_t workspace_bytes) {
  MP_REQUIRE(Q && meshsize && kx && ky && kz && k2_thresholds && tkeys && tvals && ells && out && los, "mp_spectrum_mean_bin: null pointer");
  MP_REQUIRE(nkeys >= 1 && ntab >= 1 && ntab <= nkeys, "mp_spectrum_mean_bin: %d values for %d keys", ntab, nkeys);
  MP_REQUIRE(ellin >= 0 && ellin <= 5 && nell >= 1 && nell <= 6 && nedges >= 2, "mp_spectrum_mean_bin: bad multipoles or bins");
  if (!xpass_supported(meshsize))
    return set_error(MP_ERR_UNSUPPORTED, "mp_spectrum_mean_bin: meshsize (%d, %d, %d) not supported", meshsize[0], meshsize[1], meshsize[2]);
  hipStream_t s = (hipStream_t)stream;
  const int64_t sp = spectrum_pitch(meshsize[2]);
  Workspace ws(workspace, workspace_bytes);
  float2* spec = ws.take<float2>((size_t)meshsize[0] * meshsize[1] * sp);
  const size_t rest = workspace_bytes > ws.used ? workspace_bytes - ws.used : 0;
  if (!spec || rest < bin_partial_workspace_bytes(meshsize, 1, nell, nedges))
    return set_error(MP_ERR_WORKSPACE, "mp_spectrum_mean_bin: workspace of %zu bytes < %lld", workspace_bytes,
                     (long long)mp_window_bin_workspace_bytes(meshsize, nell, nedges));
  ShellGen g = make_shell(kx, ky, kz, 0., 0., ellin, los, scale);
  g.tkeys = tkeys;
  g.tvals = tvals;
  g.tnkeys = nkeys;
  g.tn = ntab;
  int rc = convolve_shell_zy(s, g, spec, sp, Q, meshsize, 1.f);
  if (rc) return rc;
  BinGeom geo{};
  geo.n0 = meshsize[0];
  geo.n1 = meshsize[1];
  geo.ncols = meshsize[2] / 2 + 1;
  geo.nedges = nedges;
  geo.kx = kx;
  geo.ky = ky;
  geo.kz = kz;
  geo.thr = k2_thresholds;
  int32_t rz[6] = {0, 0, 0, 0, 0, 0};
  char* base = static_cast<char*>(workspace) + ws.used;
  return bin_partial(s, spec, sp, meshsize, 0, 1, 0, 0, geo, los, nell, ells, rz, out, base, rest, 1);
}

// out = c2r(P(k) Y_ell,m(khat) scale) with the tabulated theory P of mp_spectrum_mean_bin: the
// xi of the local line-of-sight mean (jaxpower/mesh2.py:1640-1645); workspace as
// mp_window_shell_ylm.
MP_EXPORT int mp_theory_ylm_real(void* stream, const int32_t meshsize[3], const double* kx, const double* ky, const double* kz,
                                 int nkeys, const double* tkeys, int ntab, const float* tvals, int ell, int m, float scale, float* out,
                                 void* workspace, size_t workspace_bytes) {
  MP_REQUIRE(meshsize && kx && ky && kz && tkeys && tvals && out, "mp_theory_ylm_real: null pointer");
  MP_REQUIRE(nkeys >= 1 && ntab >= 1 && ntab <= nkeys, "mp_theory_ylm_real: %d values for %d keys", ntab, nkeys);
  MP_REQUIRE(ell >= 0 && ell <= 4 && m >= -ell && m <= ell, "mp_theory_ylm_real: (ell, m) = (%d, %d) not supported", ell, m);
  const int64_t sp = spectrum_pitch(meshsize[2]);
  Workspace ws(workspace, workspace_bytes);
  float2* scratch = ws.take<float2>((size_t)meshsize[0] * meshsize[1] * sp);
  if (!scratch) return set_error(MP_ERR_WORKSPACE, "mp_theory_ylm_real: workspace too small");
  ShellGen g = make_shell(kx, ky, kz, 0., 0., 0, nullptr, scale);
  g.ylm_on = 1;
  g.ylm_ell = ell;
  g.ylm_m = m;
  g.tkeys = tkeys;
  g.tvals = tvals;
  g.tnkeys = nkeys;
  g.tn = ntab;
  return irfft3d_shell((hipStream_t)stream, g, out, meshsize, scratch, sp);
}

namespace mp {

// Gauss-Legendre nodes and weights on [-1, 1] (Newton iteration on P_n, the textbook method).
static void gauss_legendre(int n, double* x, double* w) {
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double z = cos(3.14159265358979323846 * (i + 0.75) / (n + 0.5)), pp = 0.;
    for (int it = 0; it < 100; ++it) {
      double p1 = 1., p2 = 0.;
      for (int j = 1; j <= n; ++j) {
        const double p3 = p2;
        p2 = p1;
        p1 = ((2. * j - 1.) * z * p2 - (j - 1.) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.);
      const double dz = p1 / pp;
      z -= dz;
      if (fabs(dz) < 1e-15) break;
    }
    x[i] = -z;
    x[n - 1 - i] = z;
    w[i] = w[n - 1 - i] = 2. / ((1. - z * z) * pp * pp);
  }
}

}  // namespace mp

// Real Gaunt coefficient int Y_l1,m1 Y_l2,m2 Y_l3,m3 dOmega with this library's real spherical
// harmonics (the ones every Y_lm kernel evaluates), by Gauss-Legendre quadrature in cos(theta)
// times a uniform rule in phi, exact for l1 + l2 + l3 <= 12; l <= 4.  The real Gaunt products of
// jaxpower/mesh2.py:1365-1372 (jaxpower/utils.py:540-600) pair each Y with the same Y elsewhere in
// the estimator, so computing them with the library's own Y keeps the result independent of the
// per-m sign convention.
MP_EXPORT double mp_real_gaunt(int l1, int m1, int l2, int m2, int l3, int m3) {
  const int ls[3] = {l1, l2, l3}, ms[3] = {m1, m2, m3};
  for (int i = 0; i < 3; ++i)
    if (ls[i] < 0 || ls[i] > 4 || ms[i] < -ls[i] || ms[i] > ls[i]) return NAN;
  constexpr int NT = 16, NP = 32;
  double x[NT], w[NT];
  gauss_legendre(NT, x, w);
  double sum = 0.;
  for (int i = 0; i < NT; ++i) {
    const double ct = x[i], st = sqrt(1. - ct * ct);
    for (int j = 0; j < NP; ++j) {
      const double phi = 2. * 3.14159265358979323846 * j / NP;
      const float ux = (float)(st * cos(phi)), uy = (float)(st * sin(phi)), uz = (float)ct;
      double f = 1.;
      for (int q = 0; q < 3; ++q) f *= (double)ylm_unit(ls[q], ms[q], ux, uy, uz);
      sum += w[i] * f;
    }
  }
  return sum * 2. * 3.14159265358979323846 / NP;
}

namespace mp {

__global__ void __launch_bounds__(MA_NT)
axpy_kernel(float alpha, const float* __restrict__ x, float* __restrict__ y, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * MA_NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * MA_NT) y[i] += alpha * x[i];
}

}  // namespace mp

// y += alpha x on real meshes of n floats: the Gaunt-weighted sums of the local-theory window
// (jaxpower/mesh2.py:1398-1402, Qs[p] += rg Y_p,mp(shat) Q).
MP_EXPORT int mp_axpy(void* stream, float alpha, const float* x, float* y, int64_t n) {
  MP_REQUIRE(x && y && n >= 0, "mp_axpy: null pointer");
  if (n == 0) return MP_OK;
  int dev = 0, ncu = 0;
  MP_CHECK_HIP(hipGetDevice(&dev));
  MP_CHECK_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const int blocks = (int)std::min<int64_t>(ceil_div(n, MA_NT), (int64_t)ncu * 8);
  MP_LAUNCH("window_axpy", axpy_kernel, dim3((unsigned)blocks), dim3(MA_NT), 0, (hipStream_t)stream, alpha, x, y, n);
  return MP_OK;
}

// The configuration-space shell of the window with the 'local' theory line of sight and the
// theory on the mesh (jaxpower/mesh2.py:1436): out = c2r([k2_lo <= k^2 < k2_hi] Y_l1,m1(khat)
// Y_l2,m2(-khat) scale), a real mesh; workspace as mp_window_shell_ylm.  l1, l2 <= 4.
MP_EXPORT int mp_window_shell_ylm2(void* stream, const int32_t meshsize[3], const double* kx, const double* ky, const double* kz,
                                   double k2_lo, double k2_hi, int l1, int m1, int l2, int m2, float scale, float* out,
                                   void* workspace, size_t workspace_bytes) {
  MP_REQUIRE(meshsize && kx && ky && kz && out, "mp_window_shell_ylm2: null pointer");
  MP_REQUIRE(l1 >= 0 && l1 <= 4 && m1 >= -l1 && m1 <= l1 && l2 >= 0 && l2 <= 4 && m2 >= -l2 && m2 <= l2,
             "mp_window_shell_ylm2: (l1, m1, l2, m2) = (%d, %d, %d, %d) not supported", l1, m1, l2, m2);
  const int64_t sp = spectrum_pitch(meshsize[2]);
  Workspace ws(workspace, workspace_bytes);
  float2* scratch = ws.take<float2>((size_t)meshsize[0] * meshsize[1] * sp);
  if (!scratch) return set_error(MP_ERR_WORKSPACE, "mp_window_shell_ylm2: workspace too small");
  ShellGen g = make_shell(kx, ky, kz, k2_lo, k2_hi, 0, nullptr, scale);
  g.ylm_on = 1;
  g.ylm_ell = l1;
  g.ylm_m = m1;
  g.ylm2_on = 1;
  g.ylm2_ell = l2;
  g.ylm2_m = m2;
  return irfft3d_shell((hipStream_t)stream, g, out, meshsize, scratch, sp);
}

namespace mp {

__global__ void __launch_bounds__(MA_NT)
radial_power_kernel(const float* __restrict__ in, float* __restrict__ out, int n0, int n1, int n2, const double* __restrict__ cx,
                    const double* __restrict__ cy, const double* __restrict__ cz, double power) {
  const int64_t total = (int64_t)n0 * n1 * n2;
  for (int64_t i = (int64_t)blockIdx.x * MA_NT + threadIdx.x; i < total; i += (int64_t)gridDim.x * MA_NT) {
    const int iz = (int)(i % n2);
    const int64_t row = i / n2;
    const int iy = (int)(row % n1), ix = (int)(row / n1);
    const double r = sqrt(cx[ix] * cx[ix] + cy[iy] * cy[iy] + cz[iz] * cz[iz]);
    out[i] = in[i] * (float)pow(r, power);
  }
}

}  // namespace mp

// out = in |c|^power on a real mesh, c = (cx[i], cy[j], cz[k]) (positions or separations; 0^p
// as pow gives it): the wide-angle factors |x|^-wa and |s|^wa of jaxpower/mesh2.py:1244, 1269,
// 1642.  In place allowed.
MP_EXPORT int mp_mul_radial_power(void* stream, const float* in, float* out, const int32_t meshsize[3], const double* cx,
                                  const double* cy, const double* cz, double power) {
  MP_REQUIRE(in && out && meshsize && cx && cy && cz, "mp_mul_radial_power: null pointer");
  const int64_t total = (int64_t)meshsize[0] * meshsize[1] * meshsize[2];
  int dev = 0, ncu = 0;
  MP_CHECK_HIP(hipGetDevice(&dev));
  MP_CHECK_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const int blocks = (int)std::min<int64_t>(ceil_div(total, MA_NT), (int64_t)ncu * 8);
  MP_LAUNCH("window_radial_power", radial_power_kernel, dim3((unsigned)blocks), dim3(MA_NT), 0, (hipStream_t)stream, in, out,
            meshsize[0], meshsize[1], meshsize[2], cx, cy, cz, power);
  return MP_OK;
}

// out = c2r(P(k) Y_l1,m1(khat) Y_l2,m2(-khat) scale) with the tabulated theory P of
// mp_spectrum_mean_bin: the xi of the mean spectrum with the 'local' theory line of sight
// (jaxpower/mesh2.py:1659); workspace as mp_window_shell_ylm.  l1, l2 <= 4.
MP_EXPORT int mp_theory_ylm2_real(void* stream, const int32_t meshsize[3], const double* kx, const double* ky, const double* kz,
                                  int nkeys, const double* tkeys, int ntab, const float* tvals, int l1, int m1, int l2, int m2,
                                  float scale, float* out, void* workspace, size_t workspace_bytes) {
  MP_REQUIRE(meshsize && kx && ky && kz && tkeys && tvals && out, "mp_theory_ylm2_real: null pointer");
  MP_REQUIRE(nkeys >= 1 && ntab >= 1 && ntab <= nkeys, "mp_theory_ylm2_real: %d values for %d keys", ntab, nkeys);
  MP_REQUIRE(l1 >= 0 && l1 <= 4 && m1 >= -l1 && m1 <= l1 && l2 >= 0 && l2 <= 4 && m2 >= -l2 && m2 <= l2,
             "mp_theory_ylm2_real: (l1, m1, l2, m2) = (%d, %d, %d, %d) not supported", l1, m1, l2, m2);
  const int64_t sp = spectrum_pitch(meshsize[2]);
  Workspace ws(workspace, workspace_bytes);
  float2* scratch = ws.take<float2>((size_t)meshsize[0] * meshsize[1] * sp);
  if (!scratch) return set_error(MP_ERR_WORKSPACE, "mp_theory_ylm2_real: workspace too small");
  ShellGen g = make_shell(kx, ky, kz, 0., 0., 0, nullptr, scale);
  g.ylm_on = 1;
  g.ylm_ell = l1;
  g.ylm_m = m1;
  g.ylm2_on = 1;
  g.ylm2_ell = l2;
  g.ylm2_m = m2;
  g.tkeys = tkeys;
  g.tvals = tvals;
  g.tnkeys = nkeys;
  g.tn = ntab;
  return irfft3d_shell((hipStream_t)stream, g, out, meshsize, scratch, sp);
}

// Weights of the bispectrum shot noise with a global line of sight (jaxpower/mesh3.py:820-839):
// out = Y_ell,m(khat) P_leg(k.los/|k|) on the half spectrum (imaginary part 0); ell < 0 drops
// the Y factor, leg = 0 the Legendre one.  Y = 0 and mu = 0 at k = 0.
MP_EXPORT int mp_ylm_legendre_spectrum(void* stream, const int32_t meshsize[3], const double* kx, const double* ky, const double* kz,
                                       int ell, int m, int leg, const double los[3], void* out, int64_t out_pitch) {
  MP_REQUIRE(meshsize && kx && ky && kz && out, "mp_ylm_legendre_spectrum: null pointer");
  MP_REQUIRE(ell < 0 || (ell <= 4 && m >= -ell && m <= ell), "mp_ylm_legendre_spectrum: (ell, m) = (%d, %d) not supported", ell, m);
  MP_REQUIRE(leg >= 0 && leg <= 5 && (leg == 0 || los), "mp_ylm_legendre_spectrum: Legendre order %d", leg);
  const int nc = meshsize[2] / 2 + 1;
  MP_REQUIRE(out_pitch >= nc, "mp_ylm_legendre_spectrum: pitch too small");
  ShellGen g = make_shell(kx, ky, kz, -1., INFINITY, leg, los, 1.f);
  if (ell >= 0) {
    g.ylm_on = 1;
    g.ylm_ell = ell;
    g.ylm_m = m;
  }
  const int64_t total = (int64_t)meshsize[0] * meshsize[1] * nc;
  int dev = 0, ncu = 0;
  MP_CHECK_HIP(hipGetDevice(&dev));
  MP_CHECK_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const int blocks = (int)std::min<int64_t>(ceil_div(total, WS_NT), (int64_t)ncu * 16);
  MP_LAUNCH("ylm_weights", window_shell_kernel, dim3((unsigned)blocks), dim3(WS_NT), 0, (hipStream_t)stream, g, meshsize[0], meshsize[1],
            nc, static_cast<float2*>(out), out_pitch);
  return MP_OK;
}
