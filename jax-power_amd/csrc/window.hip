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
