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

}  // namespace mp
