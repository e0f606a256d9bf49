// fish.hpp -- AreaSegment block tests shared by the obstacle list builder and
// the swimmer kernels.
#pragma once

#include "device_util.hpp"

namespace cup2d {

// AreaSegment record (FishLibrary.h:149-168), FISH_SEGD doubles: safe
// distance, lab-frame box objBoxLabFr (x0 x1 y0 y1), objBoxObjFr (x0 x1 y0 y1),
// |normalI| (2), |normalJ| (2).
enum { SG_SAFE = 0, SG_LAB = 1, SG_OBJ = 5, SG_NI = 9, SG_NJ = 11 };

// AreaSegment::isIntersectingWithAABB (FishLibrary.cpp:344-401); start and
// end are the centres of the block's first and last cells.
__host__ __device__ inline bool seg_intersects(const double* S, const double start[2],
                                               const double end[2]) {
  const double safe = S[SG_SAFE];
  const double AABB_w[2] = {(end[0] - start[0]) / 2 + safe, (end[1] - start[1]) / 2 + safe};
  const double AABB_c[2] = {(end[0] + start[0]) / 2, (end[1] + start[1]) / 2};
  const double box[2][2] = {{AABB_c[0] - AABB_w[0], AABB_c[0] + AABB_w[0]},
                            {AABB_c[1] - AABB_w[1], AABB_c[1] + AABB_w[1]}};
  const double* lab = S + SG_LAB;
  const double* obj = S + SG_OBJ;
  const double* nI = S + SG_NI;
  const double* nJ = S + SG_NJ;
  const double iL00 = dmax(lab[0], box[0][0]), iL01 = dmin(lab[1], box[0][1]);
  const double iL10 = dmax(lab[2], box[1][0]), iL11 = dmin(lab[3], box[1][1]);
  if (iL01 - iL00 < 0 || iL11 - iL10 < 0) return false;
  const double widthXbox[2] = {AABB_w[0] * nI[0], AABB_w[0] * nJ[0]};
  const double widthYbox[2] = {AABB_w[1] * nI[1], AABB_w[1] * nJ[1]};
  const double bb00 = AABB_c[0] - widthXbox[0] - widthYbox[0];
  const double bb01 = AABB_c[0] + widthXbox[0] + widthYbox[0];
  const double bb10 = AABB_c[1] - widthXbox[1] - widthYbox[1];
  const double bb11 = AABB_c[1] + widthXbox[1] + widthYbox[1];
  const double iF00 = dmax(bb00, obj[0]), iF01 = dmin(bb01, obj[1]);
  const double iF10 = dmax(bb10, obj[2]), iF11 = dmin(bb11, obj[3]);
  if (iF01 - iF00 < 0 || iF11 - iF10 < 0) return false;
  return true;
}

__device__ inline void block_corners(const Grid& g, int bx, int by, double start[2], double end[2]) {
  start[0] = cell_pos(bx, 0, g.hb, g.h);
  start[1] = cell_pos(by, 0, g.hb, g.h);
  end[0] = cell_pos(bx, BS - 1, g.hb, g.h);
  end[1] = cell_pos(by, BS - 1, g.hb, g.h);
}

// Fish::create steps 6-7 (Fish.cpp:158-185): does any segment of swimmer F
// meet block (bx, by)?
__device__ inline bool fish_touching(const Grid& g, const FishDev& F, int bx, int by) {
  double start[2], end[2];
  block_corners(g, bx, by, start, end);
  for (int i = 0; i < F.Nseg; ++i)
    if (seg_intersects(F.seg + (size_t)i * FISH_SEGD, start, end)) return true;
  return false;
}

}  // namespace cup2d
