/*
 * C ABI of libjaxpower_amd.so: the MI355X (gfx950) mesh power-spectrum path.
 *
 * Conventions
 *   - every pointer to array data is a device pointer unless stated otherwise; geometry
 *     arrays (meshsize, boxsize, boxcenter, los) and product descriptors are host pointers;
 *   - `stream` is a hipStream_t (NULL = default stream); calls are asynchronous on it;
 *   - real meshes are C-order float32 (N0, N1, N2); half spectra are complex64
 *     (N0, N1, pitch) with N2/2+1 valid entries per row and pitch >= N2/2+1;
 *   - every function returns 0 on success or a negative code (-1 invalid argument,
 *     -2 unsupported, -3 HIP error, -4 workspace too small); mp_last_error() explains it;
 *   - workspace arguments are caller-owned device buffers of at least the size returned by
 *     the matching *_workspace_bytes query; the library itself never allocates device memory
 *     except a 32 KiB twiddle table per device, set up on the first FFT call.
 *
 * Each entry point names the jax-power function it replaces (jaxpower/<file>:<line>).
 */
#ifndef JAXPOWER_AMD_H
#define JAXPOWER_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- errors, version, timing ------------------------------------------------------ */
const char* mp_last_error(void);
int mp_version(void);
/* Per-kernel HIP-event timing of every launch made by the library (for roofline figures). */
int mp_timing_enable(int enable);
int mp_timing_collect(void);
int mp_timing_reset(void);
/* Writes "kernel total_ms launches\n" lines into buf; returns the number of kernels. */
int mp_timing_report(char* buf, int64_t buflen);

/* ---- painting --------------------------------------------------------------------- */
/* resamplers.paint driven by _paint with out='real' and no compensation
 * (jaxpower/resamplers.py:73-103, jaxpower/mesh.py:2673-2736): paints np particles into
 * max(interlacing,1) meshes, mesh j shifted by +j/interlacing grid units, into
 * meshes[j * N0*N1*N2 ...].  positions: (np, 3) float32 (positions_f64 = 0) or float64;
 * weights: (np,) float32 or NULL for unit weights; order 1..4 = ngp, cic, tsc, pcs. */
int64_t mp_paint_workspace_bytes(const int32_t meshsize[3], int64_t np, int order, int interlacing);
int mp_paint(void* stream, const void* positions, int positions_f64, const float* weights, int64_t np,
             const int32_t meshsize[3], const double boxsize[3], const double boxcenter[3], int order,
             int interlacing, float* meshes, void* workspace, size_t workspace_bytes);
/* Same result with per-particle float atomics (any mesh shape; no workspace). */
int mp_paint_atomic(void* stream, const void* positions, int positions_f64, const float* weights, int64_t np,
                    const int32_t meshsize[3], const double boxsize[3], const double boxcenter[3], int order,
                    int interlacing, float* meshes);

/* _paint(..., out='complex') (jaxpower/mesh.py:2673-2754): interlaced paint, r2c of every
 * shifted mesh, phase exp(+i s_j (kx+ky+kz)) / n_i and sum, then (compensate != 0) the
 * window compensation prod sinc(k/2pi)^-order (jaxpower/resamplers.py:16-26). */
int64_t mp_paint_spectrum_workspace_bytes(const int32_t meshsize[3], int64_t np, int order, int interlacing);
int mp_paint_spectrum(void* stream, const void* positions, int positions_f64, const float* weights, int64_t np,
                      const int32_t meshsize[3], const double boxsize[3], const double boxcenter[3], int order,
                      int interlacing, int compensate, void* out, int64_t out_pitch, void* workspace,
                      size_t workspace_bytes);
/* The two halves of mp_paint_spectrum, so that a consumer can take the interlaced meshes
 * before their combine.  mp_paint_partial: the scan over interlacing shifts of _paint
 * (jaxpower/mesh.py:2738-2747) up to, not including, the x axis of each r2c: writes
 * `interlacing` partial spectra, mesh j at partial + j * N0 * N1 * pitch, each transformed
 * along z and y only. */
int64_t mp_paint_partial_workspace_bytes(const int32_t meshsize[3], int64_t np, int order, int interlacing);
int mp_paint_partial(void* stream, const void* positions, int positions_f64, const float* weights, int64_t np,
                     const int32_t meshsize[3], const double boxsize[3], const double boxcenter[3], int order,
                     int interlacing, void* partial, int64_t pitch, void* workspace, size_t workspace_bytes);
/* mp_combine_partial: the x axis of the r2c of each partial spectrum, the phases
 * exp(+i s_j (kx+ky+kz)) / n_i, the sum over shifts and the compensation
 * (jaxpower/mesh.py:2739-2751, jaxpower/resamplers.py:16-26); zero_dc sets mode (0,0,0) to 0,
 * as mesh.at[0, 0, 0].set(0.) does. */
int64_t mp_combine_partial_workspace_bytes(const int32_t meshsize[3], int interlacing);
int mp_combine_partial(void* stream, const void* partial, int64_t pitch, const int32_t meshsize[3], int order,
                       int interlacing, int compensate, int zero_dc, void* out, int64_t out_pitch, void* workspace,
                       size_t workspace_bytes);
/* mp_bin_poles_partial: compute_mesh2_spectrum with a global line of sight on the auto
 * spectrum of the combined mesh (jaxpower/mesh2.py:346-369, BinMesh2SpectrumPoles.__call__
 * mesh2.py:100-119), fused with mp_combine_partial so the combined mesh is never written:
 * out[p][b][re, im] = sum over modes of bin b of w |A|^2 L_ell[p](mu), im = 0, with the same
 * tables, thresholds and remove_zero rule as mp_bin_poles. */
int64_t mp_bin_partial_workspace_bytes(const int32_t meshsize[3], int interlacing, int nprod, int nedges);
int mp_bin_poles_partial(void* stream, const void* partial, int64_t pitch, const int32_t meshsize[3], int order,
                         int interlacing, int compensate, int zero_dc, const double* kx, const double* ky,
                         const double* kz, const double* k2_thresholds, int nedges, const double los[3], int nprod,
                         const int32_t* prod_ell, const int32_t* prod_remove_zero, double* out, void* workspace,
                         size_t workspace_bytes);
/* Batches of mocks on one mesh geometry (the per-mock loop of the reference's measurement
 * scripts, scripts/abacus_hf.py:357-373, which calls _paint and compute_mesh2_spectrum once per
 * mock) as one launch per phase: nmock catalogues stacked in one particle array, catalogue m
 * holding particles moff[m] .. moff[m + 1] (moff: device int64, nmock + 1 entries, moff[0] = 0,
 * np = moff[nmock] < 2^31).  No counterpart in the reference's API.
 * mp_paint_partial_batch: mp_paint_partial for every catalogue (fused paint only; -2 where it
 * does not apply), the nmock * interlacing partial spectra catalogue-major at
 * partial + (m * interlacing + j) * N0 * N1 * pitch.
 * mp_bin_poles_partial_batch: mp_bin_poles_partial for each catalogue of such a partial array,
 * out[m][p][b][re, im].
 * mp_catalogue_sums: out[m][0] = sum w, out[m][1] = sum w^2 over catalogue m in FP64 (weights
 * null: unit weights), for the box normalization and shot noise of each mock. */
int64_t mp_paint_partial_batch_workspace_bytes(const int32_t meshsize[3], int64_t np, int nmock, int order, int interlacing);
int mp_paint_partial_batch(void* stream, const void* positions, int positions_f64, const float* weights, int64_t np,
                           const int64_t* moff, int nmock, const int32_t meshsize[3], const double boxsize[3],
                           const double boxcenter[3], int order, int interlacing, void* partial, int64_t pitch, void* workspace,
                           size_t workspace_bytes);
int64_t mp_bin_partial_batch_workspace_bytes(const int32_t meshsize[3], int interlacing, int nprod, int nedges, int nmock);
int mp_bin_poles_partial_batch(void* stream, const void* partial, int64_t pitch, int nmock, const int32_t meshsize[3], int order,
                               int interlacing, int compensate, int zero_dc, const double* kx, const double* ky, const double* kz,
                               const double* k2_thresholds, int nedges, const double los[3], int nprod, const int32_t* prod_ell,
                               const int32_t* prod_remove_zero, double* out, void* workspace, size_t workspace_bytes);
int64_t mp_catalogue_sums_workspace_bytes(int nmock);
int mp_catalogue_sums(void* stream, const float* weights, const int64_t* moff, int nmock, double* out, void* workspace,
                      size_t workspace_bytes);
/* mp_bin_poles_partial without a workspace argument, for callers that cannot allocate device
 * memory (code under stream capture, serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  The combine tables and per-block partial histograms live in a
 * fixed device array of mp_noalloc_scratch_bytes() (16 MiB) per device, which the library holds
 * until the process ends; sizes whose tables and one partial histogram exceed it return -2.
 * Every call on one device shares that array: calls that overlap on one device race.  Same
 * arguments, kernels and results as mp_bin_poles_partial. */
int64_t mp_noalloc_scratch_bytes(void);
int mp_bin_poles_partial_noalloc(void* stream, const void* partial, int64_t pitch, const int32_t meshsize[3], int order,
                                 int interlacing, int compensate, int zero_dc, const double* kx, const double* ky,
                                 const double* kz, const double* k2_thresholds, int nedges, const double los[3], int nprod,
                                 const int32_t* prod_ell, const int32_t* prod_remove_zero, double* out);
/* ---- Backward (d loss / d positions, weights) ---------------------------------------- */
/* jax.grad of num_raw through the binning of compute_mesh2_spectrum (jaxpower/mesh2.py:
 * 346-369) and the interlacing combine (jaxpower/mesh.py:2739-2751), transposed by hand:
 * given coef[p][b] = d loss / d sums[p][b] (sums as written by mp_bin_poles_partial), writes
 * the `interlacing` half spectra Y_j whose c2r (mp_irfft3d) is d loss / d delta_j, the
 * derivative with respect to the j-th shifted painted mesh.  A is the combined spectrum
 * (mp_combine_partial with the same order, interlacing, compensate and zero_dc). */
int64_t mp_bin_adjoint_workspace_bytes(const int32_t meshsize[3], int interlacing);
int mp_bin_adjoint(void* stream, const void* A, int64_t pitch, const int32_t meshsize[3], int order, int interlacing,
                   int compensate, int zero_dc, const double* kx, const double* ky, const double* kz,
                   const double* k2_thresholds, int nedges, const double los[3], int nprod, const int32_t* prod_ell,
                   const int32_t* prod_remove_zero, const double* coef, void* Y, void* workspace, size_t workspace_bytes);
/* Transpose of resamplers.paint (jaxpower/resamplers.py:73-103) with respect to weights and
 * positions: grad_weights[i] = sum_j sum_stencil D_j K, grad_positions[i][a] =
 * w_i sum_j sum_stencil D_j dK/dx_a (physical units; d round/floor = 0).  meshes holds the
 * `interlacing` real meshes D_j contiguously. */
int mp_read_grad(void* stream, const float* meshes, int interlacing, const int32_t meshsize[3], const double boxsize[3],
                 const double boxcenter[3], const void* positions, int positions_f64, const float* weights, int64_t np,
                 int order, float* grad_weights, float* grad_positions);

/* Paint path of mp_paint_spectrum: 0 automatic (fused paint + z-FFT in LDS when the shape
 * allows), 1 tiled paint then z pass, 2 fused only, 3 per-particle atomics then z pass.
 * Returns the previous setting; a negative path only queries.  No counterpart in the
 * reference (an implementation choice). */
int mp_set_paint_path(int path);
/* Autotuner of the fused paint (no counterpart in the reference): on the first paint of each
 * (meshsize, order, interlacing) case, the row-block shapes that fit are timed on the actual
 * particles and the fastest is kept for the process (skipped under graph capture and below
 * 2^18 particles).  enable >= 0 sets it (default 1; 0 takes the first shape that fits;
 * 100 + q forces candidate q where it fits, for tests); returns the previous setting.  The
 * report writes one line per tuned case ("... rows=RXxRY ms: ...") and returns their count. */
int mp_set_paint_tuning(int enable);
int mp_paint_tuning_report(char* buf, int64_t buflen);
/* The autotuner's table (no counterpart in the reference, which leaves launch configuration to
 * XLA).  Knobs: 0 the fused paint's row-block shape, 1 the tile width of the binning x pass,
 * 2 the tile width of the y pass, 3 the particles per row-sort scatter workgroup.  On the first
 * launch of a case every candidate is timed on the actual data and the fastest kept (not under
 * graph capture or per-kernel timing).  mp_tune_export writes the table as a JSON object
 * {"<case key>": {"knob": name, "choice": q, "ms": [...], "measured": bool}} and returns the
 * length of the full text; mp_tune_set imports one entry; mp_tune_force(knob, q) forces
 * candidate q of a knob (-1 frees it) and returns the previous forced candidate;
 * mp_tune_enable(0) turns measurement off (untuned cases take candidate 0), a negative value
 * only queries; mp_tune_clear empties the table. */
int mp_tune_export(char* buf, int64_t buflen);
int mp_tune_set(const char* key, int knob, int choice);
int mp_tune_force(int knob, int choice);
int mp_tune_enable(int enable);
int mp_tune_clear(void);

/* resamplers.read via RealMeshField.read (jaxpower/resamplers.py:106-131,
 * jaxpower/mesh.py:2065-2115): out[i] = sum_stencil mesh * K. */
int mp_read(void* stream, const float* mesh, const int32_t meshsize[3], const double boxsize[3],
            const double boxcenter[3], const void* positions, int positions_f64, int64_t np, int order, float* out);

/* ---- FFT -------------------------------------------------------------------------- */
/* Row pitch (complex elements) the library uses for its own half spectra. */
int64_t mp_spectrum_pitch(int32_t nz);
/* MeshAttrs.r2c, fft_backend='jax' (jaxpower/mesh.py:1679-1697): rfftn, unnormalised. */
int mp_rfft3d(void* stream, const float* in, void* out, const int32_t meshsize[3], int64_t out_pitch);
/* MeshAttrs.c2r (jaxpower/mesh.py:1699-1719): irfftn(s=meshsize), 1/N^3 normalisation. */
int64_t mp_irfft3d_workspace_bytes(const int32_t meshsize[3]);
int mp_irfft3d(void* stream, const void* in, int64_t in_pitch, float* out, const int32_t meshsize[3],
               void* workspace, size_t workspace_bytes);

/* ---- binning ---------------------------------------------------------------------- */
/* _get_bin_attrs with hermitian weights (jaxpower/mesh.py:2812-2863): nmodes[b] (exact
 * integers) and ksum[b] = sum w |k + shift| for bins b = 0..nedges-2.  kx, ky, kz: device
 * float64 wavenumbers per index; k2_thresholds[e] = min{t : fl(sqrt(t)) >= edge[e]};
 * mode_shift: host (sx, sy, sz) added to the wavenumbers before the norm that picks the bin,
 * one shift of the mode-oversampling loop of _make_edges2 (jaxpower/mesh2.py:48-61), or NULL
 * for none.  The caller sums the calls over the shifts. */
int mp_bin_modes(void* stream, const int32_t meshsize[3], const double* kx, const double* ky, const double* kz,
                 const double* k2_thresholds, int nedges, const double mode_shift[3], int64_t* nmodes, double* ksum);
/* BinMesh2SpectrumPoles.__call__ over the products of compute_mesh2_spectrum
 * (jaxpower/mesh2.py:100-119, 304-333, 351-369): for product p,
 *   out[(p*nbins + b)*2 + {0,1}] = {Re, Im} sum_{modes in b} w conj(M_i) M_j P_ell(mu)
 * with mu = k.los/|k|, P_0 = 1, the (0,0,0) mode skipped when prod_remove_zero[p];
 * prod_i[p] = -1 stands for the constant 1 (plain binning of M_j).  mode_shift as for
 * mp_bin_modes: the bin of a mode is that of k + shift, mu and w those of k (one term of
 * _bincount over the ibin list, jaxpower/mesh.py:2874-2910). */
int64_t mp_bin_workspace_bytes(int nprod, int nedges);
int mp_bin_poles(void* stream, int nmesh, const void* const* meshes, const int64_t* pitches,
                 const int32_t meshsize[3], const double* kx, const double* ky, const double* kz,
                 const double* k2_thresholds, int nedges, const double mode_shift[3], const double los[3], int nprod,
                 const int32_t* prod_i,
                 const int32_t* prod_j, const int32_t* prod_ell, const int32_t* prod_remove_zero, double* out,
                 void* workspace, size_t workspace_bytes);

/* ---- correlation function ---------------------------------------------------------- */
/* _make_edges2(kind='real') -> _get_bin_attrs without weights (jaxpower/mesh2.py:25-62,
 * 120-158; jaxpower/mesh.py:2831-2863): nmodes[b] (exact cell counts) and ssum[b] = sum |s + shift|
 * for bins b = 0..nedges-2 over the full real mesh.  sx, sy, sz: device float64 separations
 * fftfreq(N) * cellsize * N per index; s2_thresholds[e] = min{t : fl(sqrt(t)) >= edge[e]};
 * mode_shift: host shift of the mode-oversampling loop (jaxpower/mesh2.py:48-61) or NULL. */
int mp_bin_real_modes(void* stream, const int32_t meshsize[3], const double* sx, const double* sy, const double* sz,
                      const double* s2_thresholds, int nedges, const double mode_shift[3], int64_t* nmodes, double* ssum);
/* BinMesh2CorrelationPoles.__call__ with the Legendre weights of compute_mesh2_correlation
 * (jaxpower/mesh2.py:175-184, 490-498): out[e * nbins + b] = sum over the cells of bin b of
 * mesh * P_ell[e](mu_s), mu_s = s.los/|s| (0 at s = 0), the (0,0,0) cell skipped when
 * remove_zero[e]; FP64 sums, at most 6 ells.  mode_shift as for mp_bin_real_modes (the bin
 * of s + shift, mu of s). */
int64_t mp_bin_real_workspace_bytes(const int32_t meshsize[3], int nell, int nedges);
int mp_bin_real_poles(void* stream, const float* mesh, const int32_t meshsize[3], const double* sx, const double* sy,
                      const double* sz, const double* s2_thresholds, int nedges, const double mode_shift[3],
                      const double los[3], int nell,
                      const int32_t* ells, const int32_t* remove_zero, double* out, void* workspace, size_t workspace_bytes);
/* mp_irfft3d then mp_bin_real_poles in one go (jaxpower/mesh2.py:477-498: c2r of the product
 * spectrum, then the multipole binning), the last inverse pass fused with the binning so the
 * real mesh is never stored: out[e * nbins + b] = sum over bin b of irfftn(in) P_ell[e](mu_s).
 * x_done != 0: `in` already carries its (unnormalised) inverse FFT along x. */
int64_t mp_irfft3d_bin_real_workspace_bytes(const int32_t meshsize[3], int nell, int nedges);
int mp_irfft3d_bin_real(void* stream, const void* in, int64_t in_pitch, int x_done, const int32_t meshsize[3], const double* sx,
                        const double* sy, const double* sz, const double* s2_thresholds, int nedges, const double los[3],
                        int nell, const int32_t* ells, const int32_t* remove_zero, double* out, void* workspace,
                        size_t workspace_bytes);
/* The spectrum products of compute_mesh2_correlation (jaxpower/mesh2.py:429-431, 447, 476-478):
 * out = conj(a) * b on half spectra, or |a|^2 (imaginary part 0) when b is NULL; out may be a. */
int mp_spectrum_product(void* stream, const void* a, int64_t a_pitch, const void* b, int64_t b_pitch,
                        const int32_t meshsize[3], void* out, int64_t out_pitch);
/* Real-space Y_lm weights of the local line-of-sight correlation (jaxpower/mesh2.py:447-450,
 * jaxpower/utils.py:190-334): out (+)= in * Y_lm(c) with c = (cx[i], cy[j], cz[k]) (positions
 * for xhat, separations for shat), Y = 0 at the zero vector; ell <= 4. */
int mp_mul_ylm(void* stream, const float* in, float* out, const int32_t meshsize[3], const double* cx, const double* cy,
               const double* cz, int ell, int m, int accumulate);
/* mp_combine_partial writing |A|^2 (imaginary part 0) instead of A: the auto-correlation
 * input Aell = |A|^2 of jaxpower/mesh2.py:476, without storing A; with inverse_x != 0 the
 * (unnormalised) inverse FFT along x of |A|^2 is written instead (the first pass of its c2r). */
int mp_combine_partial_power(void* stream, const void* partial, int64_t pitch, const int32_t meshsize[3], int order,
                             int interlacing, int compensate, int zero_dc, int inverse_x, void* out, int64_t out_pitch,
                             void* workspace, size_t workspace_bytes);

/* ---- local line of sight ---------------------------------------------------------- */
/* One term of the firstpoint estimator (jaxpower/mesh2.py:319-329) and the third side of the
 * local line-of-sight bispectrum (jaxpower/mesh3.py:563-576):
 * out = sum_m r2c(rmesh * Y_lm(xhat)) * Y_lm(khat), ell in 1..4, x measured from the origin. */
int64_t mp_ylm_workspace_bytes(const int32_t meshsize[3]);
int mp_ylm_accumulate(void* stream, const float* rmesh, const int32_t meshsize[3], const double boxsize[3],
                      const double boxcenter[3], int ell, void* out, int64_t out_pitch, void* workspace,
                      size_t workspace_bytes);
/* The ell > 0 terms of the local line-of-sight estimator for one ell, fused
 * (jaxpower/mesh2.py:311-333): out[p][b][re, im] = sum over modes of bin b of conj(C_ell) A0 w
 * (p = 0, mode k = 0 excluded) and, when with_ell0 != 0, |A0|^2 w (p = 1), with
 * C_ell = sum_m r2c(rmesh Y_lm(xhat)) Y_lm(khat) never stored.  rmesh: the real mesh (N0, N1, N2);
 * A0: its half spectrum (or the second field's for a cross spectrum). */
int64_t mp_ylm_bin_workspace_bytes(const int32_t meshsize[3], int ell, int nedges);
int mp_ylm_bin(void* stream, const float* rmesh, const void* A0, int64_t a0_pitch, const int32_t meshsize[3],
               const double boxsize[3], const double boxcenter[3], int ell, int with_ell0, const double* kx,
               const double* ky, const double* kz, const double* k2_thresholds, int nedges, double* out, void* workspace,
               size_t workspace_bytes);

/* ---- bispectrum, Scoccimarro basis, global line of sight ------------------------------ */
/* The shell meshes of BinMesh3SpectrumPoles.__call__ (jaxpower/mesh3.py:300-305), with the
 * Legendre factor of the third mesh of compute_mesh3_spectrum (mesh3.py:585-590):
 * out + q N0 N1 N2 = irfftn(A(k) [bin(k) == b0 + q] P_ell(k.los/|k|)) for q < nshell, float32
 * (N0, N1, N2) meshes; A == NULL gives the shell indicators whose triangle sums are the mode
 * counts (mesh3.py:229-236).  Bins as in mp_bin_poles (k tables and squared-|k| thresholds). */
size_t mp_bispec_shells_workspace_bytes(const int32_t meshsize[3]);
int mp_bispec_shells(void* stream, const void* A, int64_t pitch, const int32_t meshsize[3], const double* kx, const double* ky,
                     const double* kz, const double* k2_thresholds, int nedges, int b0, int nshell, const double los[3], int ell,
                     float* out, void* workspace, size_t workspace_bytes);
/* The Sugiyama basis' shell meshes (jaxpower/mesh3.py:595-621: bin(A Y*_l1m1(khat), ...) with
 * complex harmonics, mesh3.py:599-606), by components: out + q * N0 N1 N2 = Q_q, a real mesh
 * with ifftn(A [bin == b0 + q] Y_lm(khat)) = Q_q (even l) or i Q_q (odd l), Y_lm the real
 * harmonic of jaxpower/utils.py:191-334 (l <= 4); A null: the shell indicators.  Workspace as
 * mp_bispec_shells. */
int mp_bispec_shells_ylm(void* stream, const void* A, int64_t pitch, const int32_t meshsize[3], const double* kx, const double* ky,
                         const double* kz, const double* thr, int nedges, int b0, int nshell, int ell, int m, float* out,
                         void* workspace, size_t workspace_bytes);
/* The Bessel-filtered meshes of BinMesh3CorrelationPoles.__call__, Sugiyama basis
 * (jaxpower/mesh3.py:383-395), weighted by a real harmonic for compute_mesh3_correlation
 * (mesh3.py:690-716): for each separation s[q] (host array, q < ns),
 * ifftn(A(k) j_ellj(|k| s[q]) [klo <= |k| < khi] Y_lm(k)) = out_h[q] + i out_a[q] on the full
 * complex grid (k the fftfreq wavenumber vector of each index, -Nyquist at N/2), A a half
 * spectrum of a real mesh; kfun the fundamental wavenumbers.  l <= 4 (l = -1: no harmonic,
 * the filters of the bin call itself), ellj <= 5; workspace as mp_bispec_shells. */
int mp_bessel_shells_ylm(void* stream, const void* A, int64_t pitch, const int32_t meshsize[3], const double kfun[3], const double* s,
                         int ns, int ell, int m, int ellj, double klo, double khi, float* out_h, float* out_a, void* workspace,
                         size_t workspace_bytes);
/* The spherical-Bessel sums of the Sugiyama shot noise's last term (jaxpower/mesh3.py:875-903,
 * get_spherical_jn of jaxpower/utils.py:418-450): out[p] = sum over the cells s of F(s)
 * j_l1(|s| k1[b1]) j_l2(|s| k2[b2]) with (b1, b2) = pairs[2p .. 2p+1]; F a float32 real mesh,
 * |s| from the float64 separation tables; l1, l2 <= 5; FP64 sums. */
size_t mp_hankel_pairs_workspace_bytes(int npairs);
int mp_hankel_pairs(void* stream, const float* F, const int32_t meshsize[3], const double* sx, const double* sy, const double* sz,
                    int l1, const double* k1, int nk1, int l2, const double* k2, int nk2, const int32_t* pairs, int npairs,
                    double* out, void* workspace, size_t workspace_bytes);
/* The closed-triangle reduction of BinMesh3SpectrumPoles.bin_reduce_meshs with
 * reduce = prod(meshes).sum() (jaxpower/mesh3.py:238-275, 307-311): out[t] = sum over the ncells
 * cells of s1[b1] s2[b2] s3[b3] for (b1, b2, b3) = tri[3t .. 3t+2] (a device int32 array sorted by
 * (b1, b2)); shell q of a stack at stack + q * ncells; FP64 across cells. */
size_t mp_bispec_triangles_workspace_bytes(int ntri);
int mp_bispec_triangles(void* stream, const float* s1, const float* s2, const float* s3, int64_t ncells, int nshell1,
                        int nshell2, int nshell3, const int32_t* tri, int ntri, double* out, void* workspace,
                        size_t workspace_bytes);

/* ---- window matrix, global line of sight ----------------------------------------------- */
/* One input column of compute_mesh2_spectrum_window without flags (jaxpower/mesh2.py:1097-1099,
 * selection Q of mesh2.py:1022-1029): the half spectrum out = A, or r2c(Q irfftn(A)) when Q is
 * not NULL, with A = [k2_lo <= k^2 < k2_hi] P_ellin(k.los/|k|) scale (mu = 0 at k = 0), k^2 as in
 * mp_bin_poles (thresholds from the input edges e0, e1 the same way).  Q: real (N0, N1, N2)
 * float32 mesh; the output multipoles are binned with mp_bin_poles (prod_i = -1). */
int mp_window_shell(void* stream, const float* Q, const int32_t meshsize[3], const double* kx, const double* ky,
                    const double* kz, double k2_lo, double k2_hi, int ellin, const double los[3], float scale, void* out,
                    int64_t out_pitch);

/* mp_window_shell with a selection, then the output binning, for ncol columns in one call
 * (jaxpower/mesh2.py:1037-1045, 1099): column c is the input shell k2_lo[c] <= k^2 < k2_hi[c] of
 * multipole ellin[c] (k2_lo, k2_hi, ellin: HOST arrays of ncol entries) and
 * out[c][p][b][re, im] = sum over the modes of bin b, mode k = 0 excluded, of
 * w Re(B) P_ell[p](k.los/|k|) (imaginary slots 0), B = r2c(Q irfftn(A)); neither A nor B is
 * written.  Power-of-two meshsizes, nell <= 6; bins as in mp_bin_poles. */
int64_t mp_window_bin_workspace_bytes(const int32_t meshsize[3], int nell, int nedges);
int mp_window_bin(void* stream, const float* Q, const int32_t meshsize[3], const double* kx, const double* ky,
                  const double* kz, const double* k2_thresholds, int nedges, int ncol, const double* k2_lo,
                  const double* k2_hi, const int32_t* ellin, const double los[3], float scale, int nell, const int32_t* ells,
                  double* out, void* workspace, size_t workspace_bytes);

/* compute_mesh2_spectrum_window(flags=('infinite',)) with a global line of sight
 * (jaxpower/mesh2.py:1078-1092): column c takes the input shell [k0[c], k1[c]) of multipole
 * ellin[c] (HOST arrays; edges, not thresholds) to configuration space,
 * A(s) = xi_ellin(|s|) P_ellin(s.los/|s|) scale (times the real mesh Q unless NULL), with
 * xi_ell = BesselIntegral(..., method='exact', mode='backward').w (jaxpower/utils.py:470-501),
 * then bins r2c(A) as mp_window_bin does.  Equal cellsizes, power-of-two meshsizes, nell <= 6. */
int64_t mp_window_tophat_workspace_bytes(const int32_t meshsize[3], int nell, int nedges);
int mp_window_bin_tophat(void* stream, const float* Q, const int32_t meshsize[3], double cellsize, const double* kx,
                         const double* ky, const double* kz, const double* k2_thresholds, int nedges, int ncol,
                         const double* k0, const double* k1, const int32_t* ellin, const double los[3], double scale,
                         int nell, const int32_t* ells, double* out, void* workspace, size_t workspace_bytes);
/* Host-side value of xi_ell(s) for the shell [k0, k1), ell in 0..5 (NaN otherwise): the
 * configuration-space weights above, for checking them without a device. */
double mp_tophat_shell_xi(int ell, double k0, double k1, double s);

/* compute_mesh2_spectrum_window(flags=('infinite',)) with a local (firstpoint) line of sight,
 * theory line of sight 'firstpoint', wide-angle order 0 (jaxpower/mesh2.py:1235-1260), for one
 * output (ell, m) and input multipole ellin: column c (input shell [k0[c], k1[c]), HOST arrays)
 * gives out[c][b][re, im] = sum over bin b (k = 0 excluded when remove_zero) of
 * w Re(B) Y_ell,m(khat), B = r2c(xi_ellin(|s|) scale Qs), with Qs the real selection term
 * V_cell sum_m1 Y_ellin,m1(shat) c2r(conj(r2c(rho Y_ell,m Y_ellin,m1)) A0) the caller forms with
 * mp_mul_ylm, mp_rfft3d, mp_spectrum_product and mp_irfft3d.  Workspace:
 * mp_window_tophat_workspace_bytes(meshsize, 1, nedges).  Equal cellsizes, power-of-two
 * meshsizes, ell <= 4. */
int mp_window_bin_tophat_local(void* stream, const float* Qs, const int32_t meshsize[3], double cellsize, const double* kx,
                               const double* ky, const double* kz, const double* k2_thresholds, int nedges, int ncol,
                               const double* k0, const double* k1, int ellin, int ell, int m, int remove_zero, double scale,
                               double* out, void* workspace, size_t workspace_bytes);

/* compute_mesh2_spectrum_window without flags, local (firstpoint) line of sight, theory line of
 * sight 'firstpoint', wide-angle order 0 (jaxpower/mesh2.py:1262-1305), in three pieces the caller
 * chains per input shell and output (ell, m):
 *   mp_window_shell_ylm: out = c2r([k2_lo <= k^2 < k2_hi] Y_ell,m(khat) scale), a real mesh (the
 *     xi of mesh2.py:1276; the closed right edge |k| <= e1 is k^2 < t(nextafter(e1)));
 *   mp_mul_acc: out (+)= a * b on real meshes (Aell += xi * Q, mesh2.py:1277-1281);
 *   mp_bin_ylm_real: out[b][re, im] = sum over bin b (k = 0 excluded when remove_zero) of
 *     w Re(r2c(mesh)) Y_ell,m(khat) (mesh2.py:1283-1285), the spectrum binned in its x pass.
 * ell <= 4; mp_bin_ylm_real needs power-of-two meshsizes. */
int64_t mp_window_shell_ylm_workspace_bytes(const int32_t meshsize[3]);
int mp_window_shell_ylm(void* stream, const int32_t meshsize[3], const double* kx, const double* ky, const double* kz,
                        double k2_lo, double k2_hi, int ell, int m, float scale, float* out, void* workspace,
                        size_t workspace_bytes);
int mp_mul_acc(void* stream, const float* a, const float* b, float* out, int64_t n, int accumulate);
int64_t mp_bin_ylm_real_workspace_bytes(const int32_t meshsize[3], int nedges);
int mp_bin_ylm_real(void* stream, const float* mesh, const int32_t meshsize[3], const double* kx, const double* ky,
                    const double* kz, const double* k2_thresholds, int nedges, int ell, int m, int remove_zero, double* out,
                    void* workspace, size_t workspace_bytes);

/* compute_mesh2_spectrum_mean (jaxpower/mesh2.py:1485-1680) for a theory tabulated per multipole:
 * P(k) = tvals[#(tkeys <= k^2) - 1] when that index is in [0, ntab), else 0 (tkeys: nkeys
 * float64 squared-edge thresholds, or the sorted distinct k^2 of the mesh for a theory given as
 * a function of |k|; device arrays).
 *   mp_spectrum_mean_bin, global line of sight, one input multipole: out[p][b][re, im] = sum over
 *     bin b (k = 0 included) of w Re(r2c(Q irfftn(P P_ellin(mu) scale))) P_ell[p](mu); Q a real
 *     mesh; workspace mp_window_bin_workspace_bytes(meshsize, nell, nedges);
 *   mp_theory_ylm_real: out = c2r(P Y_ell,m(khat) scale), the xi of the local line of sight
 *     (mesh2.py:1640-1645), finished with mp_mul_acc and mp_bin_ylm_real; workspace
 *     mp_window_shell_ylm_workspace_bytes(meshsize). */
int mp_spectrum_mean_bin(void* stream, const float* Q, const int32_t meshsize[3], const double* kx, const double* ky,
                         const double* kz, const double* k2_thresholds, int nedges, int nkeys, const double* tkeys, int ntab,
                         const float* tvals, int ellin, const double los[3], float scale, int nell, const int32_t* ells,
                         double* out, void* workspace, size_t workspace_bytes);
int mp_theory_ylm_real(void* stream, const int32_t meshsize[3], const double* kx, const double* ky, const double* kz, int nkeys,
                       const double* tkeys, int ntab, const float* tvals, int ell, int m, float scale, float* out, void* workspace,
                       size_t workspace_bytes);

/* Host: real Gaunt coefficient int Y_l1,m1 Y_l2,m2 Y_l3,m3 dOmega with the library's own real
 * spherical harmonics (the coefficients of jaxpower/mesh2.py:1365-1372, jaxpower/utils.py:540-600,
 * in this library's per-m sign convention, which cancels in the estimator); l <= 4, NaN otherwise. */
double mp_real_gaunt(int l1, int m1, int l2, int m2, int l3, int m3);

/* The shell of the window with the 'local' theory line of sight and the theory on the mesh
 * (jaxpower/mesh2.py:1436): out = c2r([k2_lo <= k^2 < k2_hi] Y_l1,m1(khat) Y_l2,m2(-khat) scale),
 * a real mesh; workspace mp_window_shell_ylm_workspace_bytes(meshsize); l1, l2 <= 4. */
int mp_window_shell_ylm2(void* stream, const int32_t meshsize[3], const double* kx, const double* ky, const double* kz,
                         double k2_lo, double k2_hi, int l1, int m1, int l2, int m2, float scale, float* out, void* workspace,
                         size_t workspace_bytes);

/* out = in |c|^power on a real mesh with c = (cx[i], cy[j], cz[k]) (positions or separations): the
 * wide-angle factors |x|^-wa and |s|^wa of the firstpoint theory (jaxpower/mesh2.py:1244, 1269,
 * 1642); in place allowed. */
int mp_mul_radial_power(void* stream, const float* in, float* out, const int32_t meshsize[3], const double* cx,
                        const double* cy, const double* cz, double power);

/* out = c2r(P(k) Y_l1,m1(khat) Y_l2,m2(-khat) scale) with the tabulated theory of
 * mp_spectrum_mean_bin: the xi of compute_mesh2_spectrum_mean with the 'local' theory line of
 * sight (jaxpower/mesh2.py:1647-1662); workspace mp_window_shell_ylm_workspace_bytes(meshsize). */
int mp_theory_ylm2_real(void* stream, const int32_t meshsize[3], const double* kx, const double* ky, const double* kz, int nkeys,
                        const double* tkeys, int ntab, const float* tvals, int l1, int m1, int l2, int m2, float scale, float* out,
                        void* workspace, size_t workspace_bytes);

/* y += alpha x on real meshes (device arrays of n floats): the Gaunt-weighted sums of the window
 * with the 'local' theory line of sight (jaxpower/mesh2.py:1398-1402). */
int mp_axpy(void* stream, float alpha, const float* x, float* y, int64_t n);

/* Host arrays, for the smooth window (compute_mesh2_spectrum_window(flags=('smooth',)),
 * jaxpower/mesh2.py:1045-1076): out[i] = xi_ell(s[i]) of the shell [k0, k1) as above, and
 * out[i] = j_ell(x[i]) by the closed forms / series of jaxpower/utils.py:418-450; ell in 0..5. */
int mp_tophat_shell_xi_array(int ell, double k0, double k1, const double* s, int64_t n, double* out);
int mp_spherical_jn_array(int ell, const double* x, int64_t n, double* out);

/* Weights of compute_fkp3_shotnoise, Scoccimarro basis, global line of sight
 * (jaxpower/mesh3.py:786-852): out = Y_ell,m(khat) P_leg(k.los/|k|) as a half spectrum (imaginary
 * part 0; ell < 0 drops the Y factor, leg = 0 the Legendre one), binned against the
 * paint(w) conj(paint(w^2)) product with mp_bin_poles. */
int mp_ylm_legendre_spectrum(void* stream, const int32_t meshsize[3], const double* kx, const double* ky, const double* kz,
                             int ell, int m, int leg, const double los[3], void* out, int64_t out_pitch);

/* ---- reductions (normalisation and shot noise, jaxpower/mesh.py:3092-3129,
 *      jaxpower/mesh2.py:516-621) ------------------------------------------------------ */
int64_t mp_reduce_workspace_bytes(void);
/* out[0] = sum a*b (b != NULL), sum a^2 (square != 0) or sum a, accumulated in float64. */
int mp_sum_f64(void* stream, const float* a, const float* b, int64_t n, int square, double* out, void* workspace,
               size_t workspace_bytes);
int mp_fill_f32(void* stream, float* p, int64_t n, float value);

#ifdef __cplusplus
}
#endif

#endif /* JAXPOWER_AMD_H */
