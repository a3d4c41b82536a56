// ljg_cells.h -- position -> LinkedCells cell on the device (bit-exact restatement of
// CellBlock3D::get3DIndexOfPosition).  Include it only from translation units compiled with
// -ffp-contract=off: the index arithmetic must round exactly as the reference's FP64 expression.
#pragma once

#include "ljg_internal.h"

namespace ljg {

// CellBlock3D.h:320-349 and threeToOneD (ThreeDimensionalMapping.h:29-31).
__device__ inline int32_t cell_index_1d(const CellGeom &g, double px, double py, double pz) {
  const double p[3] = {px, py, pz};
  int32_t c3[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const long value = (long)floor((p[d] - g.box_min[d]) * g.recip[d]) + g.cpil;
    const long nonneg = value > 0 ? value : 0;
    const long upper = g.nh[d] - 1;
    long c = nonneg < upper ? nonneg : upper;
    if (p[d] >= g.box_max[d]) {
      const long first_upper_halo = g.nh[d] - g.cpil;
      c = c > first_upper_halo ? c : first_upper_halo;
    } else if (p[d] < g.box_min[d] && c == g.cpil) {
      --c;
    } else if (p[d] < g.box_max[d] && c == g.nh[d] - g.cpil) {
      --c;
    }
    c3[d] = (int32_t)c;
  }
  return (c3[2] * g.nh[1] + c3[1]) * g.nh[0] + c3[0];
}

}  // namespace ljg
