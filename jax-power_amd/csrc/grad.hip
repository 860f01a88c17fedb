// Hand-written backward of the global-line-of-sight power spectrum of an interlaced paint,
// d loss / d (positions, weights) given d loss / d num_raw.  The reference obtains these by
// jax.grad through resamplers.paint (jaxpower/resamplers.py:73-103), the interlacing scan
// (jaxpower/mesh.py:2738-2751) and the binning of compute_mesh2_spectrum
// (jaxpower/mesh2.py:346-369); this file restates the same derivatives by hand.
//
// Forward: delta_j = paint(x + s_j) (real N^3), F_j = rfft(delta_j),
//   A = c(k)/n_i sum_j phi_j(k) F_j  (phi_j = exp(+i s_j sum k_circ), c = compensation),
//   num[p][b] = (2 ell_p + 1)/nmodes_b sum_{m in b} w_m |A_m|^2 L_ell_p(mu_m).
// Backward, with G[p][b] = d loss / d num[p][b] and H_m = w_m sum_p coef[p][b_m] L_ell_p(mu_m),
// coef = G (2 ell + 1)/nmodes:
//   d loss / d delta_j = N^3 irfftn(Y_j),  Y_j(m) = f_m H_m A_m c_m conj(phi_j,m) / n_i,
//   f_m = 2 on the kz = 0 and kz = N/2 planes (where the half spectrum holds both k and -k) and
//   1 elsewhere, which is 2 Re sum_half B e^{ikx} written as a hermitian inverse transform;
//   d loss / d w_i = sum_j sum_stencil D_j K,  d loss / d x~_i = sum_j w_i sum_stencil D_j dK/dx~,
//   with D_j = d loss / d delta_j and d(round/floor)/dx = 0, as in the reference.
#include "bin.h"
#include "fft.h"
#include "fft_core.h"
#include "resampler.h"
#include "sph.h"

namespace mp {

constexpr int ADJ_NT = 256;
constexpr int ADJ_MAX_PROD = 6;
constexpr int ADJ_MAX_NI = 8;

struct AdjArgs {
  const float2* A;  // combined half spectrum
  int64_t pitch;
  float2* Y;        // n_i half spectra (same pitch), mesh j at Y + j * mesh_stride
  int64_t mesh_stride;
  int ni, nz;
  BinGeom g{};
  CombineTables tab;
  float scale;      // N^3 / n_i
  int zero_dc;
  int nprod;
  int ell[ADJ_MAX_PROD], rz[ADJ_MAX_PROD];
  float los[3];
  const double* coef;  // [nprod][nbins]
};

__global__ void __launch_bounds__(ADJ_NT) bin_adjoint_kernel(AdjArgs a) {
  extern __shared__ double thr[];
  const int nedges = a.g.nedges, nbins = nedges - 1;
  for (int i = threadIdx.x; i < nedges; i += ADJ_NT) thr[i] = a.g.thr[i];
  __syncthreads();
  const int64_t nmodes = (int64_t)a.g.n0 * a.g.n1 * a.g.ncols;
  for (int64_t q = (int64_t)blockIdx.x * ADJ_NT + threadIdx.x; q < nmodes; q += (int64_t)gridDim.x * ADJ_NT) {
    const int iz = (int)(q % a.g.ncols);
    const int64_t row = q / a.g.ncols;
    const int iy = (int)(row % a.g.n1), ix = (int)(row / a.g.n1);
    const double kxv = a.g.kx[ix], kyv = a.g.ky[iy], kzv = a.g.kz[iz];
    const double k2 = k2_sum(k2_xy(kxv, kyv), kzv);
    const int b = first_bin(thr, nedges, k2);
    float2 base = make_float2(0.f, 0.f);
    const bool origin = (ix == 0) & (iy == 0) & (iz == 0);
    if (b >= 1 && b < nedges && !(origin && a.zero_dc)) {
      const float fkx = (float)kxv, fky = (float)kyv, fkz = (float)kzv;
      const float kn2 = fkx * fkx + fky * fky + fkz * fkz;
      const float mu = kn2 > 0.f ? (fkx * a.los[0] + fky * a.los[1] + fkz * a.los[2]) * rsqrtf(kn2) : 0.f;
      double h = 0.;
      for (int p = 0; p < a.nprod; ++p) {
        if (origin && a.rz[p]) continue;
        h += a.coef[p * nbins + b - 1] * (double)(a.ell[p] ? legendre(a.ell[p], mu) : 1.f);
      }
      const float w = iz > 0 ? 2.f : 1.f;
      const float f = (iz == 0 || 2 * iz == a.nz) ? 2.f : 1.f;
      float s = (float)h * w * f * a.scale;
      if (a.tab.cx) s *= a.tab.cx[ix] * a.tab.cy[iy] * a.tab.cz[iz];
      base = cscale(a.A[row * a.pitch + iz], s);
    }
    for (int j = 0; j < a.ni; ++j) {
      float2 v = base;
      if (a.tab.phx) {
        const int o = j * a.tab.stride;
        const float2 ph = cmul(cmul(a.tab.phx[o + ix], a.tab.phy[o + iy]), a.tab.phz[o + iz]);
        v = cmul(v, cconj(ph));
      }
      a.Y[j * a.mesh_stride + row * a.pitch + iz] = v;
    }
  }
}

// Adjoint readout with kernel derivatives: for each particle, gathers D_j over its stencil at
// every shift j; writes d loss / d w and d loss / d x (physical units).
template <int ORDER, typename T>
__global__ void __launch_bounds__(256)
read_grad_kernel(const float* __restrict__ D, int64_t mesh_stride, int ni, GridMap g, const T* __restrict__ pos,
                 const float* __restrict__ w, int64_t np, float* __restrict__ gw, float* __restrict__ gpos) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  constexpr int OMIN = -((ORDER - 1) / 2);
  T xg0[3];
  to_grid(pos, i, g, xg0);
  const float wi = w ? w[i] : 1.f;
  float accw = 0.f, accx = 0.f, accy = 0.f, accz = 0.f;
  for (int j = 0; j < ni; ++j) {
    const T s = (T)j / (T)ni;
    int idx[3][ORDER];
    float ker[3][ORDER], dker[3][ORDER];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      const T u = xg0[ax] + s;
      const T r = round_or_floor(u, ORDER % 2 == 1);
      const int i0 = (int)r;
#pragma unroll
      for (int o = 0; o < ORDER; ++o) {
        const T d = u - (r + (T)(OMIN + o));  // x~ - i
        const float sd = (float)fabs(d);
        idx[ax][o] = pos_mod(i0 + OMIN + o, g.n[ax]);
        ker[ax][o] = kern1d<ORDER>(sd);
        const float sign = d > 0 ? 1.f : (d < 0 ? -1.f : 0.f);
        dker[ax][o] = kern1d_ds<ORDER>(sd) * sign;  // dK/dx~ = K'(s) ds/dx~
      }
    }
    const float* mesh = D + j * mesh_stride;
#pragma unroll
    for (int ox = 0; ox < ORDER; ++ox)
#pragma unroll
      for (int oy = 0; oy < ORDER; ++oy) {
        const float* rowp = mesh + ((int64_t)idx[0][ox] * g.n[1] + idx[1][oy]) * g.n[2];
#pragma unroll
        for (int oz = 0; oz < ORDER; ++oz) {
          const float v = rowp[idx[2][oz]];
          const float kx = ker[0][ox], ky = ker[1][oy], kz = ker[2][oz];
          accw += v * kx * ky * kz;
          accx += v * dker[0][ox] * ky * kz;
          accy += v * kx * dker[1][oy] * kz;
          accz += v * kx * ky * dker[2][oz];
        }
      }
  }
  gw[i] = accw;
  gpos[3 * i + 0] = wi * accx / (float)g.cell[0];
  gpos[3 * i + 1] = wi * accy / (float)g.cell[1];
  gpos[3 * i + 2] = wi * accz / (float)g.cell[2];
}

template <typename T>
static int read_grad_typed(hipStream_t s, const float* D, int64_t mstride, int ni, const GridMap& g, const T* pos, const float* w,
                           int64_t np, int order, float* gw, float* gpos) {
  if (np == 0) return MP_OK;
  dim3 grid((unsigned)ceil_div(np, 256)), block(256);
  switch (order) {
    case 1: MP_LAUNCH("read_grad", (read_grad_kernel<1, T>), grid, block, 0, s, D, mstride, ni, g, pos, w, np, gw, gpos); break;
    case 2: MP_LAUNCH("read_grad", (read_grad_kernel<2, T>), grid, block, 0, s, D, mstride, ni, g, pos, w, np, gw, gpos); break;
    case 3: MP_LAUNCH("read_grad", (read_grad_kernel<3, T>), grid, block, 0, s, D, mstride, ni, g, pos, w, np, gw, gpos); break;
    case 4: MP_LAUNCH("read_grad", (read_grad_kernel<4, T>), grid, block, 0, s, D, mstride, ni, g, pos, w, np, gw, gpos); break;
    default: return set_error(MP_ERR_INVALID, "resampler order %d not in 1..4", order);
  }
  return MP_OK;
}

size_t xpass_tables_bytes(const int32_t n[3], int ni);
int combine_tables(hipStream_t s, const int32_t n[3], int order, int ni, int compensate, Workspace& wsp, CombineTables& tab);

}  // namespace mp

using namespace mp;

MP_EXPORT int64_t mp_bin_adjoint_workspace_bytes(const int32_t meshsize[3], int interlacing) {
  return (int64_t)(xpass_tables_bytes(meshsize, interlacing < 1 ? 1 : interlacing) + 512);
}

MP_EXPORT int mp_bin_adjoint(void* stream, const void* A, int64_t pitch, const int32_t meshsize[3], int order, int interlacing,
                             int compensate, int zero_dc, const double* kx, const double* ky, const double* kz,
                             const double* k2_thresholds, int nedges, const double los[3], int nprod, const int32_t* prod_ell,
                             const int32_t* prod_remove_zero, const double* coef, void* Y, void* workspace, size_t workspace_bytes) {
  MP_REQUIRE(meshsize && A && Y && kx && ky && kz && k2_thresholds && coef && prod_ell && prod_remove_zero,
             "mp_bin_adjoint: null pointer");
  MP_REQUIRE(order >= 1 && order <= 4, "resampler order %d not in 1..4", order);
  MP_REQUIRE(nedges >= 2 && nedges <= 8192, "mp_bin_adjoint: nedges %d not in 2..8192", nedges);
  MP_REQUIRE(nprod >= 1 && nprod <= ADJ_MAX_PROD, "mp_bin_adjoint: %d products not in 1..%d", nprod, ADJ_MAX_PROD);
  const int ni = interlacing < 1 ? 1 : interlacing;
  MP_REQUIRE(ni <= ADJ_MAX_NI, "mp_bin_adjoint: interlacing %d > %d", ni, ADJ_MAX_NI);
  MP_REQUIRE(pitch >= meshsize[2] / 2 + 1, "mp_bin_adjoint: pitch < Nz/2+1");
  hipStream_t s = (hipStream_t)stream;
  Workspace wsp(workspace, workspace_bytes);
  AdjArgs a{};
  int rc = combine_tables(s, meshsize, order, ni, compensate, wsp, a.tab);
  if (rc) return rc;
  a.A = (const float2*)A;
  a.pitch = pitch;
  a.Y = (float2*)Y;
  a.mesh_stride = (int64_t)meshsize[0] * meshsize[1] * pitch;
  a.ni = ni;
  a.nz = meshsize[2];
  a.g.n0 = meshsize[0];
  a.g.n1 = meshsize[1];
  a.g.ncols = meshsize[2] / 2 + 1;
  a.g.nedges = nedges;
  a.g.kx = kx;
  a.g.ky = ky;
  a.g.kz = kz;
  a.g.thr = k2_thresholds;
  a.scale = (float)((double)meshsize[0] * meshsize[1] * meshsize[2] / ni);
  a.zero_dc = zero_dc;
  a.nprod = nprod;
  for (int p = 0; p < nprod; ++p) {
    MP_REQUIRE(prod_ell[p] >= 0 && prod_ell[p] <= 10, "mp_bin_adjoint: ell %d not in 0..10", prod_ell[p]);
    a.ell[p] = prod_ell[p];
    a.rz[p] = prod_remove_zero[p];
  }
  for (int q = 0; q < 3; ++q) a.los[q] = los ? (float)los[q] : 0.f;
  a.coef = coef;
  const int64_t nmodes = (int64_t)meshsize[0] * meshsize[1] * (meshsize[2] / 2 + 1);
  int64_t blocks = ceil_div(nmodes, ADJ_NT);
  if (blocks > 4096) blocks = 4096;
  MP_LAUNCH("bin_adjoint", bin_adjoint_kernel, dim3((unsigned)blocks), dim3(ADJ_NT), (size_t)nedges * sizeof(double), s, a);
  return MP_OK;
}

MP_EXPORT int mp_read_grad(void* stream, const float* meshes, int interlacing, const int32_t meshsize[3], const double boxsize[3],
                           const double boxcenter[3], const void* positions, int positions_f64, const float* weights, int64_t np,
                           int order, float* grad_weights, float* grad_positions) {
  MP_REQUIRE(meshes && meshsize && boxsize && boxcenter && grad_weights && grad_positions, "mp_read_grad: null pointer");
  MP_REQUIRE(np == 0 || positions, "mp_read_grad: null positions");
  MP_REQUIRE(order >= 1 && order <= 4, "resampler order %d not in 1..4", order);
  const int ni = interlacing < 1 ? 1 : interlacing;
  GridMap g;
  for (int q = 0; q < 3; ++q) {
    g.n[q] = meshsize[q];
    g.cell[q] = boxsize[q] / (double)meshsize[q];
    g.shift[q] = boxsize[q] / 2. - boxcenter[q];
  }
  const int64_t mstride = (int64_t)meshsize[0] * meshsize[1] * meshsize[2];
  hipStream_t s = (hipStream_t)stream;
  if (positions_f64)
    return read_grad_typed<double>(s, meshes, mstride, ni, g, (const double*)positions, weights, np, order, grad_weights, grad_positions);
  return read_grad_typed<float>(s, meshes, mstride, ni, g, (const float*)positions, weights, np, order, grad_weights, grad_positions);
}
