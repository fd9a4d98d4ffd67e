// bgk_strip.h -- row-strip tiling shared by the moment kernel
// (bgk_fused_mom.hip) and the FP32-population kernel (bgk_fused32.hip).
//
// A workgroup owns W consecutive columns of a strip of consecutive rows and
// walks the strip row by row, keeping a few rows of phi (and, for moment
// storage, of the moments) in an LDS ring, so every row of those fields is
// read from HBM once per step plus the strip's two boundary rows.  gfx950
// dispatches workgroup b to XCD b % 8 and each XCD has its own L2: all
// column segments of a strip go to one XCD (the cache line a segment reads
// past its ends is then an L2 hit) and the eight XCDs work on eight
// neighbouring strips.  The boundary-adjacent nodes (i in {1, Nx} or j in
// {1, Ny}) run one per thread in extra workgroups of the same launch, after
// the strip workgroups, where BC_f and BC_g are folded into the pull.
#ifndef BGK_STRIP_H
#define BGK_STRIP_H

#include <algorithm>

#include "bgk_internal.h"

namespace bgk {

struct StripGrid {
  int nseg;    // column segments of W
  int rows;    // rows per strip
  int nstrip;  // ceil(ny / rows)
  int nmain;   // workgroups of the strips (padded to a multiple of 8 strips)
};

// Strip and column segment of strip workgroup blk; false for the padding
// workgroups past the last strip.
__device__ __forceinline__ bool strip_of(const StripGrid& sg, int blk, int* strip, int* seg) {
  const int kq = blk / NXCD;
  *strip = (kq / sg.nseg) * NXCD + blk % NXCD;
  *seg = kq % sg.nseg;
  return *strip < sg.nstrip;
}

// Node (i, j) of the boundary-adjacent ring i in {1, Nx} or j in {1, Ny}
// for edge index e: row 1, row Ny, then columns 1 and Nx of rows 2..Ny-1.
__device__ __forceinline__ bool edge_node(int e, int nx, int ny, int* i, int* j) {
  const int nrow = ny > 1 ? 2 : 1;
  if (e < nrow * nx) {
    *j = (e < nx) ? 1 : ny;
    *i = e % nx + 1;
    return true;
  }
  e -= nrow * nx;
  const int ncol = nx > 1 ? 2 : 1;
  if (ny <= 2 || e >= ncol * (ny - 2)) return false;
  *j = e / ncol + 2;
  *i = (e % ncol == 0) ? 1 : nx;
  return true;
}

// Number of boundary-adjacent nodes (the range of edge_node's e).
inline int edge_count(int nx, int ny) {
  return nx * (ny > 1 ? 2 : 1) + (ny > 2 ? (ny - 2) * (nx > 1 ? 2 : 1) : 0);
}

// Row strips of about `target` rows, and at least one workgroup per slot
// of the device (slots = CUs x resident workgroups per CU) on small
// lattices, down to 2 rows per strip.  The strip count is chosen, where one
// exists within a factor two, as a multiple of 8 (the XCDs) whose workgroups
// fill whole rounds of the device: measured at 4096^2 with the moment
// kernel, 64-row strips (512 workgroups) run at 57 GLUPS and 96-row strips
// (43 strips, 1.5 rounds) at 43.  rows_override > 0 fixes the strip height.
inline StripGrid make_strip_grid(int nx, int ny, int w, int target, int slots,
                                 int rows_override) {
  StripGrid s;
  s.nseg = (nx + w - 1) / w;
  if (rows_override > 0) {
    s.rows = rows_override;
  } else {
    int want = std::max(1, (ny + target - 1) / target);
    if (slots > 0) want = std::max(want, (slots + s.nseg - 1) / s.nseg);
    want = std::min(want, std::max(1, (ny + 1) / 2));
    int pick = (want + NXCD - 1) / NXCD * NXCD;
    for (int n = pick; n <= 2 * want && n <= ny; n += NXCD)
      if (slots > 0 && ((int64_t)n * s.nseg) % slots == 0) {
        pick = n;
        break;
      }
    s.rows = std::max(1, (ny + pick - 1) / pick);
  }
  s.nstrip = (ny + s.rows - 1) / s.rows;
  s.nmain = ((s.nstrip + NXCD - 1) / NXCD) * NXCD * s.nseg;
  return s;
}

}  // namespace bgk

#endif  // BGK_STRIP_H
