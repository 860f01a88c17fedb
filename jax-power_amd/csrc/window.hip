// Window matrix of the power-spectrum multipoles with a global line of sight
// (jaxpower/mesh2.py:921-1118, compute_mesh2_spectrum_window without flags): for each input
// multipole ellin and input |k| shell [e0, e1), the mesh-defined theory
//   A(k) = [e0 <= |k| < e1] P_ellin(k.los/|k|) scale                    (mesh2.py:1097-1098)
// is written as a half spectrum and, when a selection Q is given, replaced in place by
// r2c(Q irfftn(A)) (mesh2.py:1099, with Q = c2r(A1 conj(A2)) / N^3 from mesh2.py:1022-1029).
// The output multipoles are then binned by mp_bin_poles.
//
// The convolution runs the inverse x and y column passes in place, one z pass that goes back
// to real space, multiplies by the row of Q and transforms forward again without writing the
// real row, then the forward y and x passes: five passes over the half spectrum and one read
// of Q per shell.
#include "bin.h"
#include "common.h"
#include "fft.h"
#include "sph.h"

namespace mp {

constexpr int WS_NT = 256;

// out = [t_lo <= k^2 < t_hi] P_ell(mu) scale on the half spectrum (P_0 = 1), k^2 summed in
// float64 as the binning does, so shell membership matches digitize of the same edges.
__global__ void __launch_bounds__(WS_NT)
window_shell_kernel(int n0, int n1, int nc, const double* __restrict__ kx, const double* __restrict__ ky,
                    const double* __restrict__ kz, double t_lo, double t_hi, double lx, double ly, double lz, LegendreTable leg,
                    int ell_on, float scale, float2* __restrict__ out, int64_t pitch) {
  const int64_t total = (int64_t)n0 * n1 * nc;
  for (int64_t i = (int64_t)blockIdx.x * WS_NT + threadIdx.x; i < total; i += (int64_t)gridDim.x * WS_NT) {
    const int iz = (int)(i % nc);
    const int64_t row = i / nc;
    const int iy = (int)(row % n1), ix = (int)(row / n1);
    const double kxv = kx[ix], kyv = ky[iy], kzv = kz[iz];
    const double k2 = k2_sum(k2_xy(kxv, kyv), kzv);
    float v = 0.f;
    if (k2 >= t_lo && k2 < t_hi) {
      v = scale;
      if (ell_on) {
        // mu = k.los / |k|, 0 at k = 0 (mesh2.py:1032)
        const double mu = k2 > 0. ? (kxv * lx + kyv * ly + kzv * lz) / sqrt(k2) : 0.;
        const float m = (float)mu;
        v *= legendre_eval(leg, 0, m, m * m);
      }
    }
    out[row * pitch + iz] = make_float2(v, 0.f);
  }
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
  LegendreTable leg{};
  legendre_table_set(leg, 0, ellin);
  const double lx = ellin ? los[0] : 0., ly = ellin ? los[1] : 0., lz = ellin ? los[2] : 0.;
  const int64_t total = (int64_t)meshsize[0] * meshsize[1] * nc;
  int dev = 0, ncu = 0;
  MP_CHECK_HIP(hipGetDevice(&dev));
  MP_CHECK_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const int blocks = (int)std::min<int64_t>(ceil_div(total, WS_NT), (int64_t)ncu * 16);
  float2* o = static_cast<float2*>(out);
  MP_LAUNCH("window_shell", window_shell_kernel, dim3((unsigned)blocks), dim3(WS_NT), 0, s, meshsize[0], meshsize[1], nc, kx, ky,
            kz, k2_lo, k2_hi, lx, ly, lz, leg, ellin > 0 ? 1 : 0, scale, o, out_pitch);
  if (!Q) return MP_OK;
  return convolve_real(s, o, out_pitch, Q, meshsize, 1.f);
}
