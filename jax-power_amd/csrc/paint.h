// Internal interface of the particle painting and readout kernels.
#pragma once

#include "common.h"

namespace mp {

struct GridMap {
  int n[3];          // meshsize
  double shift[3];   // boxsize/2 - boxcenter
  double cell[3];    // cellsize
};

// Workspace bytes needed by paint_meshes (tiled path; 0 for the atomic path).
size_t paint_workspace_bytes(const int32_t n[3], int64_t np, int order, int ni);

// Paint np particles into ni real meshes (mesh j shifted by j/ni grid units), all zeroed first.
int paint_meshes(hipStream_t s, const void* pos, int pos_f64, const float* w, int64_t np, const GridMap& g, int order,
                 int ni, float* meshes, int64_t mesh_stride, void* ws, size_t ws_bytes, int force_atomic);

// Fused paint + z-axis r2c (paint_zfft.hip): ni half-spectrum meshes after the z pass.  With
// moff (device, nmock + 1 offsets), pos/w hold nmock catalogues one after the other and out
// their nmock * ni meshes, catalogue-major.
bool paint_zfft_supported(const int32_t n[3], int64_t np, int order, int ni);
size_t paint_zfft_workspace_bytes(const int32_t n[3], int64_t np, int nmock = 1);
int paint_zfft(hipStream_t s, const void* pos, int pos_f64, const float* w, int64_t np, const GridMap& g, int order, int ni,
               float2* out, int64_t pitch, int64_t mesh_stride, void* ws, size_t ws_bytes, const int64_t* moff = nullptr,
               int nmock = 1);

// Real-space paint from the row sort (paint_zfft.hip): any mesh up to 2048 per axis; its
// workspace is paint_zfft_workspace_bytes.
bool paint_rows_supported(const int32_t n[3], int64_t np, int order, int ni);
int paint_rows(hipStream_t s, const void* pos, int pos_f64, const float* w, int64_t np, const GridMap& g, int order, int ni,
               float* meshes, int64_t mesh_stride, void* ws, size_t ws_bytes);

// Path used by the interlaced spectrum paint: 0 automatic (fused when supported), 1 separate
// paint + z pass, 2 fused only (error when unsupported), 3 per-particle atomics + z pass.
int paint_path_option();

// Autotuner of the fused paint's row-block shape (paint_zfft.hip): enable >= 0 sets it,
// returns the previous setting; the report lists each tuned case with the timings.

}  // namespace mp
