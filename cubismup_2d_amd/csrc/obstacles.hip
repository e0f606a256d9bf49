// obstacles.hip -- obstacle-block work of the CubismUP_2D step on gfx950.
//
// The reference allocates an ObstacleBlock for every (shape, block) pair whose
// block touches the shape's bounding box (ShapesSimple.cpp:14-35).  Here those
// pairs are "slots" in a fixed-capacity device pool, rebuilt each step by one
// thread from the bounding box (no host round trip).  Kernels run either one
// workgroup per slot (work private to one shape) or one workgroup per distinct
// obstacle block (work that combines shapes on shared cells, done in shape
// order like the reference's `for blocks: for shapes` loops).
//
// Sums over a shape are per-slot partials (each a fixed-order tree inside the
// workgroup) added in slot order, i.e. blockID order as in
// Shape::computeForces (Shape.cpp:260-271).  Every workgroup that needs a
// shape total recomputes it from the partials in the same order, so all of
// them see bit-identical values.
#include "device_util.hpp"
#include "fish.hpp"

namespace cup2d {

#define RC_OB(x)              \
  do {                        \
    const int rc_ = (x);      \
    if (rc_) return rc_;      \
  } while (0)

constexpr double EPSD = 2.220446049250313e-16;  // std::numeric_limits<double>::epsilon()
constexpr int NT = 256;
constexpr int LDS_RED = (NT / 64) * 20;  // block_sum_multi of up to 20 values

// the BS*BS cells of a block, NT per iteration, with all iterations unrolled
#define CUP2D_CELL_PRAGMA _Pragma("unroll")
#define CELL_LOOP(i) \
  CUP2D_CELL_PRAGMA for (int i##_it = 0, i = threadIdx.x; i##_it < BS * BS / NT; ++i##_it, i += NT)

// partial-sum offsets inside a slot's NPART doubles
enum { P_COM = 0, P_MOM = 3, P_RHS = 10, P_PEN = 13, P_FRC = 20, P_NSURF = 38, P_NPART_USED = 39 };
static_assert(P_NPART_USED <= NPART + 16, "partials");
constexpr int SLOT_PART = 40;

struct SlotView {
  double *chi, *dist, *rho, *udx, *udy, *sd, *sx, *sy, *part;
};
__device__ inline SlotView slot(const Obstacles& ob, int s) {
  SlotView v;
  const size_t o = (size_t)s * BS * BS;
  v.chi = ob.chi + o; v.dist = ob.dist + o; v.rho = ob.rho + o;
  v.udx = ob.udef + 2 * o; v.udy = ob.udef + 2 * o + BS * BS;
  v.sd = ob.sdelta + o; v.sx = ob.sdx + o; v.sy = ob.sdy + o;
  v.part = ob.part + (size_t)s * SLOT_PART;
  return v;
}

// value of a dense field at global cell (gx, gy) with BlockLabOpen ghosts
// (zero-gradient at the four domain faces, Definitions.h:125-157)
template <typename RT>
__device__ inline double lab_at(const RT* f, const Grid& g, int gx, int gy) {
  gx = gx < 0 ? 0 : (gx >= g.NX ? g.NX - 1 : gx);
  gy = gy < 0 ? 0 : (gy >= g.NY ? g.NY - 1 : gy);
  return f[(size_t)gy * g.NX + gx];
}

// ------------------------------------------------------------ begin step
// PutObjectsOnGrid.cpp:200-220: U_inf from the fixed shapes
// (Shape::updateLabVelocity, Shape.cpp:61-65), Shape::updatePosition
// (Shape.cpp:67-88), domain check; then the obstacle-block lists
// (Disk::create + FillBlocks_Cylinder::is_touching, ShapeLibrary.h:24-57).
__device__ inline bool touching(const Grid& g, int bx, int by, const double bbox[2][2]) {
  const double mn0 = cell_pos(bx, 0, g.hb, g.h), mn1 = cell_pos(by, 0, g.hb, g.h);
  const double mx0 = cell_pos(bx, BS - 1, g.hb, g.h), mx1 = cell_pos(by, BS - 1, g.hb, g.h);
  const double i00 = dmax(mn0, bbox[0][0]), i01 = dmin(mx0, bbox[0][1]);
  const double i10 = dmax(mn1, bbox[1][0]), i11 = dmin(mx1, bbox[1][1]);
  return i01 - i00 > 0 && i11 - i10 > 0;
}

constexpr int LCAP = OBST_LIST_CAP;  // candidate blocks / slots handled by k_obst_begin

// swimmer segments per chunk of the list builder's prefilter, and the
// (candidate, chunk) pairs it lists per swimmer and window
constexpr int SEG_CHUNK = 8;
constexpr int SEG_LIST = 4096;

// Does the box of a block (start, end: centres of its first and last cells)
// grown by a chunk's safe distance, as AreaSegment::isIntersectingWithAABB
// grows it, meet the union c[0..3] of the chunk's lab-frame boxes?  (Its
// first test, with the union in place of one segment's box: min and max are
// exact and the rounding of the differences monotone, so a miss here is a
// miss of that test for every segment of the chunk.)
__device__ inline bool box_meets_chunk(const double* c, const double start[2],
                                       const double end[2]) {
  const double safe = c[4];
  const double w0 = (end[0] - start[0]) / 2 + safe, w1 = (end[1] - start[1]) / 2 + safe;
  const double m0 = (end[0] + start[0]) / 2, m1 = (end[1] + start[1]) / 2;
  const double iL00 = dmax(c[0], m0 - w0), iL01 = dmin(c[1], m0 + w0);
  const double iL10 = dmax(c[2], m1 - w1), iL11 = dmin(c[3], m1 + w1);
  return !(iL01 - iL00 < 0 || iL11 - iL10 < 0);
}

// exclusive scan of a[0..n) in LDS (n <= LCAP), returns the total
__device__ int block_exscan(int* a, int n, int* tmp) {
  constexpr int CH = LCAP / NT;
  const int t = threadIdx.x;
  const int i0 = t * CH, i1 = (t + 1) * CH < n ? (t + 1) * CH : n;
  int s = 0;
  for (int i = i0; i < i1; ++i) s += a[i];
  __syncthreads();
  tmp[t] = s;
  __syncthreads();
  for (int off = 1; off < NT; off <<= 1) {
    const int v = t >= off ? tmp[t - off] : 0;
    __syncthreads();
    tmp[t] += v;
    __syncthreads();
  }
  int base = t == 0 ? 0 : tmp[t - 1];
  const int total = tmp[NT - 1];
  for (int i = i0; i < i1; ++i) {
    const int v = a[i];
    a[i] = base;
    base += v;
  }
  __syncthreads();
  return total;
}

// One workgroup.  (1) In the device loop: dt from the findMaxU partials left
// by the projection and penalization kernels of the previous step
// (Simulation::calcMaxTimestep, Simulation.cpp:254-277).  (2) U_inf
// (Shape::updateLabVelocity), body motion (Shape::updatePosition) and the
// domain check.  (3) Obstacle slots: blocks whose cell-centre box meets a
// shape's bounding box (FillBlocks_Cylinder::is_touching), in shape order
// and blockID order, then the distinct blocks with their slots in shape order.
//
// phase 0 does everything; with swimmers present phase 1 (dt, motion) runs
// first, then the swimmers' midline kernel, then phase 2 (lists), whose
// candidate test for a swimmer is AreaSegment::isIntersectingWithAABB over
// its segments (Fish.cpp:158-173).
__global__ __launch_bounds__(NT) void k_obst_begin(Grid g, const ShapeParams* __restrict__ sp,
                                                   cup2d_shape_state* __restrict__ ss,
                                                   StepState* __restrict__ st, Obstacles ob,
                                                   Obstacles pv,
                                                   unsigned char* __restrict__ obmap,
                                                   int* __restrict__ ubidx,
                                                   const FishDev* __restrict__ fd,
                                                   const double* __restrict__ maxpart,
                                                   int nmaxpart,
                                                   const double* __restrict__ ramp,
                                                   int compute_dt, int step_idx, int phase,
                                                   int end_prev) {
  __shared__ int s_a[LCAP], s_b[LCAP], s_c[LCAP], s_d[LCAP];
  __shared__ int s_tmp[NT];
  __shared__ int s_cstart[MAX_SHAPES + 1];
  __shared__ int s_first[MAX_SHAPES + 1];  // first slot of each shape
  __shared__ int s_rng[MAX_SHAPES][3];
  __shared__ double s_box[MAX_SHAPES][4];
  __shared__ double lds[LDS_RED];
  __shared__ double s_seg[(FISH_MAX_NM / 8) * FISH_SEGD];
  __shared__ double s_chunk[((FISH_MAX_NM / 8 + SEG_CHUNK - 1) / SEG_CHUNK) * 5];
  __shared__ int s_list[SEG_LIST];
  __shared__ int s_nlist;
  const int t = threadIdx.x;
  const int ns = g.nshapes;
  const int nub_prev = pv.cnt[1];
  __shared__ double s_uv[MAX_SHAPES][2];
  __shared__ int s_fix[MAX_SHAPES][2];
  __shared__ double s_step[3];  // dt, U_inf
  // Every global value of the dt / motion part is loaded first, so the
  // kernel waits for one round trip instead of one per dependent step.
  double mA = 0;  // the max of the projection's partials (count known upfront)
  if (phase != 2 && compute_dt) mA = max_array<NT>(maxpart, nmaxpart);
  double Tu = 0, Tv = 0, Tom = 0, Tcm0 = 0, Tcm1 = 0, Tl0 = 0, Tl1 = 0, Tor = 0, Tg0 = 0, Tg1 = 0;
  // the bounding-box inputs of the list part, one lane per shape
  int kind = 0;
  double R = 0, Cx = 0, Cy = 0, fbox[4] = {0, 0, 0, 0};
  if (phase != 1 && t < ns) {
    kind = sp[t].kind;
    if (kind == 1) {
      const double* B = fd[sp[t].fish].box;
      for (int q = 0; q < 4; ++q) fbox[q] = B[q];
    } else {
      R = sp[t].radius;
      if (phase == 2) {  // phase 0 moves the shape below, on this lane
        Cx = ss[t].center[0];
        Cy = ss[t].center[1];
      }
    }
  }
  if (phase != 2 && t < ns) {
    const cup2d_shape_state& T = ss[t];
    Tu = T.u; Tv = T.v; Tom = T.omega;
    Tcm0 = T.centerOfMass[0]; Tcm1 = T.centerOfMass[1];
    Tl0 = T.labCenterOfMass[0]; Tl1 = T.labCenterOfMass[1];
    Tor = T.orientation; Tg0 = T.d_gm[0]; Tg1 = T.d_gm[1];
    s_fix[t][0] = sp[t].bFixedx;
    s_fix[t][1] = sp[t].bFixedy;
  }
  double sTime = 0, sDt = 0, sU0 = 0, sU1 = 0, sRamp = 0;
  // the step's index (the CFL ramp, the end-time slot): the host's count, or
  // with step_idx < 0 (launches replayed from a captured graph) the device's
  int sIdx = step_idx;
  if (t == 0) {
    sTime = st->time;
    sDt = st->dt;
    sU0 = st->uinf[0];
    sU1 = st->uinf[1];
    if (sIdx < 0) sIdx = st->step + end_prev;
    if (phase != 2 && compute_dt && sIdx < 100) sRamp = ramp[sIdx];
    if (end_prev) {  // the previous whole step ends (Simulation.cpp:289-290)
      sTime += st->dt_done;
      st->time = sTime;
      st->step += 1;
    }
  }
  if (phase != 2) {
    if (t < ns) {
      s_uv[t][0] = Tu;
      s_uv[t][1] = Tv;
    }
    // the partials of the previous step's obstacle blocks follow the
    // projection's (max does not depend on the order)
    if (compute_dt)
      for (int i = nmaxpart + t; i < nmaxpart + nub_prev; i += NT) mA = nmax(mA, maxpart[i]);
    __syncthreads();
    if (compute_dt) {
      const double maxU = block_max<NT>(mA, lds);
      if (t == 0) {
        const double h = g.h;
        const double dtFourier = h * h / g.nu, dtCFL = maxU < 2.2e-16 ? 1 : h / maxU;
        double maxUb = 0;
        for (int k = 0; k < ns; ++k)
          maxUb = dmax(maxUb, sqrt(s_uv[k][0] * s_uv[k][0] + s_uv[k][1] * s_uv[k][1]));
        const double dtBody = maxUb < 2.2e-16 ? 1 : h / maxUb;
        const double mn = dmin(dmin(dtCFL, dtFourier), dtBody);
        double dt = g.CFL * mn;
        if (sIdx < 100) dt = sRamp * mn;
        st->dt = dt;
        st->maxU = maxU;
        if (!finite_max(maxU) && st->error == 0) st->error = CUP2D_ERR_NONFINITE;
        s_step[0] = dt;
      }
    } else if (t == 0) {
      s_step[0] = sDt;
    }
    if (t == 0) {  // U_inf, in shape order
      int nSum[2] = {0, 0};
      double uSum[2] = {0, 0};
      for (int k = 0; k < ns; ++k) {
        if (s_fix[k][0]) { nSum[0]++; uSum[0] -= s_uv[k][0]; }
        if (s_fix[k][1]) { nSum[1]++; uSum[1] -= s_uv[k][1]; }
      }
      double U0 = sU0, U1 = sU1;
      if (nSum[0] > 0) U0 = uSum[0] / nSum[0];
      if (nSum[1] > 0) U1 = uSum[1] / nSum[1];
      st->uinf[0] = U0;
      st->uinf[1] = U1;
      s_step[1] = U0;
      s_step[2] = U1;
      st->tnext[sIdx & 1] = sTime + s_step[0];  // this step's end time
    }
    __syncthreads();
    if (t < ns) {  // each shape moves on its own lane
      const double dt = s_step[0], U0 = s_step[1], U1 = s_step[2];
      cup2d_shape_state& T = ss[t];
      const double cm0 = Tcm0 + dt * (Tu + U0);
      const double cm1 = Tcm1 + dt * (Tv + U1);
      T.centerOfMass[0] = cm0;
      T.centerOfMass[1] = cm1;
      T.labCenterOfMass[0] = Tl0 + dt * Tu;
      T.labCenterOfMass[1] = Tl1 + dt * Tv;
      double o = Tor + dt * Tom;
      o = o > M_PI ? o - 2 * M_PI : o;
      o = o < -M_PI ? o + 2 * M_PI : o;
      T.orientation = o;
      double sinang, cosang;
      det_sincos(o, &sinang, &cosang);
      const double c0 = cm0 + cosang * Tg0 - sinang * Tg1;
      const double c1 = cm1 + sinang * Tg0 + cosang * Tg1;
      T.center[0] = c0;
      T.center[1] = c1;
      Cx = c0;
      Cy = c1;
      if (c0 < 0 || c0 > g.extent0 || c1 < 0 || c1 > g.extent1) st->error = CUP2D_ERR_BODY_OUT;
    }
  }
  if (phase != 2) {
    // the early row pass of the pressure solve runs when the previous step's
    // obstacle blocks lie in at most a quarter of the block rows (its patch
    // transforms the rows that hold obstacle blocks); ub_block is ascending
    __shared__ int s_rows;
    if (t == 0) s_rows = 0;
    __syncthreads();
    int nr = 0;
    for (int u = t; u < nub_prev; u += NT)
      nr += u == 0 || pv.ub_block[u - 1] / g.bpdx != pv.ub_block[u] / g.bpdx ? 1 : 0;
    if (nr) atomicAdd(&s_rows, nr);
    __syncthreads();
    if (t == 0) st->rows_early_on = 4 * s_rows <= g.bpdy ? 1 : 0;
  }
  if (phase != 1)
    for (int u = t; u < nub_prev; u += NT) {
      obmap[pv.ub_block[u]] = 0;
      ubidx[pv.ub_block[u]] = -1;
    }
  __syncthreads();
  if (phase == 1) return;
  if (t < ns) {  // each shape's candidate block range on its own lane
    const int k = t;
    if (kind == 1) {  // union of the swimmer's segment boxes
      for (int q = 0; q < 4; ++q) s_box[k][q] = fbox[q];
    } else {
      const double safety = 2 * g.h;
      s_box[k][0] = Cx - R - safety; s_box[k][1] = Cx + R + safety;
      s_box[k][2] = Cy - R - safety; s_box[k][3] = Cy + R + safety;
    }
    // a box that is not finite (a body moved by a non-finite velocity or dt;
    // CUP2D_ERR_NONFINITE is raised by the dt part) selects no block: the
    // block range is clamped in floating point before any integer conversion
    bool fin = true;
    double bb[4];
    for (int q = 0; q < 4; ++q) {
      const double x = s_box[k][q] / g.hb;
      fin = fin && x == x && fabs(x) < 1e9;
      bb[q] = fin ? x : 0.0;
    }
    int bx0 = (int)floor(bb[0]) - 1, bx1 = (int)floor(bb[1]) + 1;
    int by0 = (int)floor(bb[2]) - 1, by1 = (int)floor(bb[3]) + 1;
    bx0 = bx0 < 0 ? 0 : bx0; by0 = by0 < 0 ? 0 : by0;
    bx1 = bx1 >= g.bpdx ? g.bpdx - 1 : bx1; by1 = by1 >= g.bpdy ? g.bpdy - 1 : by1;
    const int w = fin && bx1 >= bx0 ? bx1 - bx0 + 1 : 0;
    const int hgt = fin && by1 >= by0 ? by1 - by0 + 1 : 0;
    s_rng[k][0] = bx0; s_rng[k][1] = by0; s_rng[k][2] = w;
    s_tmp[k] = w * hgt;
  }
  __syncthreads();
  if (t == 0) {
    int nc = 0;
    for (int k = 0; k < ns; ++k) {
      s_cstart[k] = nc;
      nc += s_tmp[k];
    }
    s_cstart[ns] = nc;
  }
  __syncthreads();
  const int NC = s_cstart[ns];
  // The candidates are taken in windows of at most LCAP, in shape and block
  // order; each window's selected blocks are appended to the slot list.
  const int cap = ob.cap < LCAP ? ob.cap : LCAP;
  int nob = 0;  // slots so far (the same in every thread)
  bool over = false;
  for (int w0 = 0; w0 < NC; w0 += LCAP) {
    const int nw = NC - w0 < LCAP ? NC - w0 : LCAP;
    // candidates: s_a = touching flag, s_b = block, s_c = shape
    for (int c = t; c < nw; c += NT) {
      const int cidx = w0 + c;
      int k = 0;
      while (k + 1 < ns && s_cstart[k + 1] <= cidx) ++k;
      const int local = cidx - s_cstart[k], w = s_rng[k][2];
      const int bx = s_rng[k][0] + local % w, by = s_rng[k][1] + local / w;
      const double bbox[2][2] = {{s_box[k][0], s_box[k][1]}, {s_box[k][2], s_box[k][3]}};
      s_a[c] = sp[k].kind == 0 && touching(g, bx, by, bbox) ? 1 : 0;
      s_b[c] = bx + g.bpdx * by;
      s_c[c] = k;
    }
    // swimmers: any of the segments (staged in LDS) meets the block.  The
    // segments are taken in chunks of SEG_CHUNK consecutive ones: a
    // (candidate, chunk) pair whose union of lab-frame boxes misses the
    // block's box (grown by the chunk's largest safe distance) cannot pass
    // the first test of any of its segments, so only the pairs that meet it
    // are listed, and their (candidate, segment) tests run in parallel.
    for (int k = 0; k < ns; ++k) {
      if (sp[k].kind != 1) continue;
      const int c0 = s_cstart[k] > w0 ? s_cstart[k] - w0 : 0;
      const int c1 = s_cstart[k + 1] - w0 < nw ? s_cstart[k + 1] - w0 : nw;
      if (c1 <= c0) continue;
      const FishDev& F = fd[sp[k].fish];
      const int nseg = F.Nseg;
      const int nch = (nseg + SEG_CHUNK - 1) / SEG_CHUNK;
      __syncthreads();
      stage_copy<NT, ((FISH_MAX_NM / 8) * FISH_SEGD + NT - 1) / NT>(s_seg, F.seg,
                                                                   nseg * FISH_SEGD);
      if (t == 0) s_nlist = 0;
      __syncthreads();
      for (int ch = t; ch < nch; ch += NT) {
        double u[5] = {1e300, -1e300, 1e300, -1e300, 0.0};  // x0 x1 y0 y1 safe
        const int i1 = (ch + 1) * SEG_CHUNK < nseg ? (ch + 1) * SEG_CHUNK : nseg;
        for (int i = ch * SEG_CHUNK; i < i1; ++i) {
          const double* S = s_seg + i * FISH_SEGD;
          u[0] = dmin(u[0], S[SG_LAB]);
          u[1] = dmax(u[1], S[SG_LAB + 1]);
          u[2] = dmin(u[2], S[SG_LAB + 2]);
          u[3] = dmax(u[3], S[SG_LAB + 3]);
          u[4] = dmax(u[4], S[SG_SAFE]);
        }
        for (int q = 0; q < 5; ++q) s_chunk[ch * 5 + q] = u[q];
      }
      __syncthreads();
      const int nck = c1 - c0;
      for (int pr = t; pr < nck * nch; pr += NT) {
        const int c = c0 + pr / nch, ch = pr % nch;
        const int b = s_b[c];
        double start[2], end[2];
        block_corners(g, b % g.bpdx, b / g.bpdx, start, end);
        if (box_meets_chunk(s_chunk + ch * 5, start, end)) {
          const int j = atomicAdd(&s_nlist, 1);
          if (j < SEG_LIST) s_list[j] = pr;
        }
      }
      __syncthreads();
      const int nl = s_nlist;
      if (nl <= SEG_LIST) {
        for (int q = t; q < nl * SEG_CHUNK; q += NT) {
          const int pr = s_list[q / SEG_CHUNK];
          const int i = (pr % nch) * SEG_CHUNK + q % SEG_CHUNK;
          if (i >= nseg) continue;
          const int c = c0 + pr / nch;
          const int b = s_b[c];
          double start[2], end[2];
          block_corners(g, b % g.bpdx, b / g.bpdx, start, end);
          if (seg_intersects(s_seg + i * FISH_SEGD, start, end)) s_a[c] = 1;
        }
      } else {  // more meeting pairs than the list holds: every pair
        for (int pr = t; pr < nck * nseg; pr += NT) {
          const int c = c0 + pr / nseg, i = pr % nseg;
          const int b = s_b[c];
          double start[2], end[2];
          block_corners(g, b % g.bpdx, b / g.bpdx, start, end);
          if (seg_intersects(s_seg + i * FISH_SEGD, start, end)) s_a[c] = 1;
        }
      }
    }
    __syncthreads();  // the window's flags are complete
    const int nsel = block_exscan(s_a, nw, s_tmp);  // s_a now holds exclusive prefix
    // each shape's first slot: the slots before its first candidate
    if (t <= ns) {
      const int cs = s_cstart[t];
      if (cs >= w0 && cs < w0 + nw) s_first[t] = nob + s_a[cs - w0];
    }
    if (nob + nsel > cap) over = true;
    // compacted slots (s_d = block of slot)
    for (int c = t; c < nw; c += NT) {
      const int pos = s_a[c];
      const int next = c + 1 < nw ? s_a[c + 1] : nsel;
      if (next > pos && nob + pos < cap) {
        ob.slot_shape[nob + pos] = s_c[c];
        ob.slot_block[nob + pos] = s_b[c];
        s_d[nob + pos] = s_b[c];
      }
    }
    nob += nsel;
    __syncthreads();  // s_a, s_b, s_c are refilled by the next window
  }
  if (over) {  // more obstacle blocks than the slot lists hold
    if (t == 0) st->error = CUP2D_ERR_CAPACITY;
    nob = 0;
    if (t <= ns) s_first[t] = 0;
    __syncthreads();
  }
  if (t <= ns) {
    const int v = s_cstart[t] >= NC ? nob : s_first[t];
    ob.slot_of_shape_start[t] = v;
    s_first[t] = v;
  }
  __syncthreads();
  // distinct blocks.  Each shape's slots are in increasing block order, so the
  // other shapes holding the block of slot s are found by a range test and a
  // binary search in their runs.  s_a = first-occurrence flag, s_b = rank
  // among the block's slots (shape order), s_c = the block's first slot,
  // s_cnt = the block's slot count (at first occurrences), s_d keeps the block.
  __shared__ int s_cnt[LCAP];
  __shared__ int s_n[LCAP];
  for (int s = t; s < nob; s += NT) {
    const int b = s_d[s];
    int first = s, rank = 0, cnt = 1;
    for (int k = 0; k < ns; ++k) {
      int lo = s_first[k], hi = s_first[k + 1];
      if (lo >= hi || (s >= lo && s < hi)) continue;  // no slots, or the own shape
      if (b < s_d[lo] || b > s_d[hi - 1]) continue;
      while (hi - lo > 1) {  // last slot of the run with block <= b
        const int mid = (lo + hi) >> 1;
        if (s_d[mid] <= b) lo = mid; else hi = mid;
      }
      if (s_d[lo] != b) continue;
      ++cnt;
      if (lo < s) {
        if (rank == 0) first = lo;
        ++rank;
      }
    }
    s_a[s] = rank == 0 ? 1 : 0;
    s_b[s] = rank;
    s_c[s] = first;
    s_cnt[s] = rank == 0 ? cnt : 0;
    s_n[s] = cnt;
  }
  __syncthreads();
  const int nub = block_exscan(s_a, nob, s_tmp);    // s_a: ub index of first occurrences
  block_exscan(s_cnt, nob, s_tmp);  // s_cnt: ub_start at first occurrences
  for (int s = t; s < nob; s += NT) {
    const int f = s_c[s];
    const int u = s_a[f];
    ob.ub_slots[s_cnt[f] + s_b[s]] = s;
    if (f == s) {
      ob.ub_block[u] = s_d[s];
      ob.ub_start[u] = s_cnt[s];
      ob.ub_count[u] = s_n[s];
      obmap[s_d[s]] = 1;
      ubidx[s_d[s]] = u;
    }
  }
  if (t == 0) {
    ob.cnt[0] = nob;
    ob.cnt[1] = nub;
  }
}

// Totals of NQ consecutive partials q0.. over the slots of shape k, reduced
// by each wave on its own (every lane gets the totals; no barrier).  The
// order of the additions is fixed, so every wave of every workgroup gets
// bit-identical totals, and the waves of one workgroup can take different
// shapes at once.
template <int NQ>
__device__ inline void shape_sums(const Obstacles& ob, int k, int q0, double* out) {
  const int s0 = ob.slot_of_shape_start[k], n = ob.slot_of_shape_start[k + 1] - s0;
  wave_sum_table<NQ>(ob.part + (size_t)s0 * SLOT_PART + q0, n, SLOT_PART, out);
}

// PutObjectsOnGrid.cpp:191-197 (chi=0, uDef=0, invRho=1) for one block
__device__ inline void clear_block(const Grid& g, int b, double* chi, double* udx, double* udy,
                                   double* invrho) {
  const int bx = b % g.bpdx, by = b / g.bpdx;
  CELL_LOOP(k) {
    const size_t d = (size_t)(by * BS + k / BS) * g.NX + bx * BS + k % BS;
    chi[d] = 0; udx[d] = 0; udy[d] = 0; invrho[d] = 1;
  }
}

// The shapes' kind, radius and centre, staged in LDS once per workgroup.
struct DiskLds {
  int kind[MAX_SHAPES];
  double R[MAX_SHAPES], c0[MAX_SHAPES], c1[MAX_SHAPES];
};

// The grid SDF the reference assembles in sim.tmp (PutObjectsOnGrid.cpp:
// 190-200 clears it to -1, Disk::create takes the max with each shape's
// distance R - |x - c| (ShapeLibrary.cpp:21-34) on that shape's obstacle
// blocks) at cell (gx, gy), clamped into the domain as the BlockLab does: max
// over the shapes whose bounding box touches the cell's block of their
// distance, -1 if none.
//
// Swimmers contribute the distance PutFishOnBlocks stored in their slots of
// that block (signedDistanceSqrt, FishLibrary.cpp:431).
// nbs[0..nnb) are the swimmer slots of the block that holds the (clamped)
// cell.
__device__ inline double grid_sdf(const Grid& g, const DiskLds& D, const Obstacles& ob,
                                  const int* nbs, int nnb, int gx, int gy) {
  gx = gx < 0 ? 0 : (gx >= g.NX ? g.NX - 1 : gx);
  gy = gy < 0 ? 0 : (gy >= g.NY ? g.NY - 1 : gy);
  double v = -1;
  for (int q = 0; q < nnb; ++q)
    v = dmax(v, ob.dist[(size_t)nbs[q] * BS * BS + (gy % BS) * BS + gx % BS]);
  for (int k = 0; k < g.nshapes; ++k) {
    if (D.kind[k] != 0) continue;
    const double safety = 2 * g.h, R = D.R[k];
    const double c0 = D.c0[k], c1 = D.c1[k];
    const double bbox[2][2] = {{c0 - R - safety, c0 + R + safety}, {c1 - R - safety, c1 + R + safety}};
    if (touching(g, gx / BS, gy / BS, bbox)) {
      const double p0 = cell_pos(gx / BS, gx % BS, g.hb, g.h) - c0;
      const double p1 = cell_pos(gy / BS, gy % BS, g.hb, g.h) - c1;
      v = dmax(v, R - sqrt(p0 * p0 + p1 * p1));
    }
  }
  return v;
}

// PutObjectsOnGrid::operator() for the obstacle blocks (PutObjectsOnGrid.cpp:
// 186-241): workgroups [0, nub), one per obstacle block of this step, clear
// the block (PutObjectsOnGrid.cpp:191-197), then for its shapes in order:
// Disk::create (ShapeLibrary.cpp:21-34: dist = R - |x - c|, rho = rhoS) and
// putChiOnGrid (PutObjectsOnGrid.cpp:17-116): Towers chi from the shape's own
// distance inside the block and the grid SDF across block edges, surface
// delta + normals (ObstacleBlock::write, ObstacleBlock.h:95-105),
// centre-of-mass partials, grid CHI = max and the invRho blend
// (PutObjectsOnGrid.cpp:69-73).  Distances are evaluated where they are
// needed, so no grid-wide SDF pass is required.  The next workgroups, one per
// obstacle block of the previous step (list pv), clear those that are not
// obstacle blocks now (every other cell already holds the cleared values).
//
// The shapes' data is staged in LDS once, and the grid CHI and invRho of the
// block stay in LDS across the block's slots: the block is cleared and
// written once, with no read back of what was just stored.
__global__ __launch_bounds__(NT) void k_obst_place(Grid g, const ShapeParams* __restrict__ sp,
                                                   const cup2d_shape_state* __restrict__ ss,
                                                   Obstacles ob, Obstacles pv,
                                                   const unsigned char* __restrict__ obmap,
                                                   const int* __restrict__ ubidx,
                                                   double* __restrict__ chi,
                                                   double* __restrict__ udx,
                                                   double* __restrict__ udy,
                                                   double* __restrict__ invrho) {
  __shared__ double lds[LDS_RED];
  __shared__ int s_nbs[4][MAX_FISH];  // swimmer slots of the E, W, N, S neighbour blocks
  __shared__ int s_nbn[4];
  __shared__ DiskLds D;
  const int nub = ob.cnt[1], w = blockIdx.x;
  if (w >= nub) {
    if (w >= nub + pv.cnt[1]) return;
    const int b = pv.ub_block[w - nub];
    if (!obmap[b]) clear_block(g, b, chi, udx, udy, invrho);
    return;
  }
  const int b = ob.ub_block[w];
  const int bx = b % g.bpdx, by = b / g.bpdx;
  for (int k = threadIdx.x; k < g.nshapes; k += NT) {
    D.kind[k] = sp[k].kind;
    D.R[k] = sp[k].radius;
    D.c0[k] = ss[k].center[0];
    D.c1[k] = ss[k].center[1];
  }
  if (threadIdx.x < 4) {
    const int d = threadIdx.x;
    int nx = bx + (d == 0) - (d == 1), ny = by + (d == 2) - (d == 3);
    nx = nx < 0 ? 0 : (nx >= g.bpdx ? g.bpdx - 1 : nx);  // clamped like the cells
    ny = ny < 0 ? 0 : (ny >= g.bpdy ? g.bpdy - 1 : ny);
    int n = 0;
    const int u = g.nfish > 0 ? ubidx[ny * g.bpdx + nx] : -1;
    if (u >= 0)
      for (int q = 0; q < ob.ub_count[u]; ++q) {
        const int s = ob.ub_slots[ob.ub_start[u] + q];
        if (sp[ob.slot_shape[s]].kind == 1 && n < MAX_FISH) s_nbs[d][n++] = s;
      }
    s_nbn[d] = n;
  }
  __syncthreads();
  // PutObjectsOnGrid.cpp:191-197: chi = 0, invRho = 1 (uDef = 0 is stored at
  // the end); each thread only touches its own cells of s_chi / s_inv
  __shared__ double s_chi[BS * BS], s_inv[BS * BS];
  for (int i = threadIdx.x; i < BS * BS; i += NT) { s_chi[i] = 0; s_inv[i] = 1; }
  const int s0 = ob.ub_start[w], cnt = ob.ub_count[w];
  const double h = g.h, i2h = 0.5 / h, fac = 0.5 * h;
  for (int q = 0; q < cnt; ++q) {
    const int s = ob.ub_slots[s0 + q], k = ob.slot_shape[s];
    SlotView o = slot(ob, s);
    // swimmers: dist, udef and rho = 1 come from PutFishOnBlocks
    const bool fish = D.kind[k] == 1;
    const double rhoS = fish ? 1.0 : sp[k].rhoS;
    const double R = D.R[k], Sc0 = D.c0[k], Sc1 = D.c1[k];
    const double C0 = ss[k].centerOfMass[0], C1 = ss[k].centerOfMass[1];
    double _x = 0, _y = 0, _m = 0, _ns = 0;
#pragma unroll 1
    for (int i = threadIdx.x; i < BS * BS; i += NT) {
      const int ix = i % BS, iy = i / BS;
      const int gx = bx * BS + ix, gy = by * BS + iy;
      // the shape's own distance inside the block, the grid SDF across its
      // edges (PutObjectsOnGrid.cpp:53-56)
      double sd, distPx, distMx, distPy, distMy;
#define GSDF(d, x, y) grid_sdf(g, D, ob, s_nbs[d], s_nbn[d], x, y)
#define DDIST(x, y)                                                              \
  (R - sqrt((cell_pos((x) / BS, (x) % BS, g.hb, g.h) - Sc0) *                   \
                (cell_pos((x) / BS, (x) % BS, g.hb, g.h) - Sc0) +                \
            (cell_pos((y) / BS, (y) % BS, g.hb, g.h) - Sc1) *                   \
                (cell_pos((y) / BS, (y) % BS, g.hb, g.h) - Sc1)))
      if (fish) {
        sd = o.dist[i];
        distPx = ix + 1 == BS ? GSDF(0, gx + 1, gy) : o.dist[i + 1];
        distMx = ix == 0 ? GSDF(1, gx - 1, gy) : o.dist[i - 1];
        distPy = iy + 1 == BS ? GSDF(2, gx, gy + 1) : o.dist[i + BS];
        distMy = iy == 0 ? GSDF(3, gx, gy - 1) : o.dist[i - BS];
      } else {
        sd = DDIST(gx, gy);
        distPx = ix + 1 == BS ? GSDF(0, gx + 1, gy) : DDIST(gx + 1, gy);
        distMx = ix == 0 ? GSDF(1, gx - 1, gy) : DDIST(gx - 1, gy);
        distPy = iy + 1 == BS ? GSDF(2, gx, gy + 1) : DDIST(gx, gy + 1);
        distMy = iy == 0 ? GSDF(3, gx, gy - 1) : DDIST(gx, gy - 1);
        o.dist[i] = sd;
        o.rho[i] = rhoS;
        o.udx[i] = 0; o.udy[i] = 0;
      }
#undef DDIST
#undef GSDF
      double X;
      if (sd > +2 * h || sd < -2 * h) {
        X = sd > 0 ? 1 : 0;
      } else {
        const double IplusX = distPx < 0 ? 0 : distPx, IminuX = distMx < 0 ? 0 : distMx;
        const double IplusY = distPy < 0 ? 0 : distPy, IminuY = distMy < 0 ? 0 : distMy;
        const double gradIX = i2h * (IplusX - IminuX), gradIY = i2h * (IplusY - IminuY);
        const double gradUX = i2h * (distPx - distMx), gradUY = i2h * (distPy - distMy);
        const double gradUSq = gradUX * gradUX + gradUY * gradUY;
        const double denum = 1 / (gradUSq < EPSD ? EPSD : gradUSq);
        X = (gradIX * gradUX + gradIY * gradUY) * denum;
      }
      o.chi[i] = X;
      // grid CHI = max and the invRho blend (PutObjectsOnGrid.cpp:69-73)
      if (X >= s_chi[i]) {
        s_chi[i] = X;
        s_inv[i] = X / rhoS + (1 - X) * s_inv[i];
      }
      if (X > 0) {
        const double p0 = cell_pos(bx, ix, g.hb, h), p1 = cell_pos(by, iy, g.hb, h);
        _x += rhoS * X * h * h * (p0 - C0);
        _y += rhoS * X * h * h * (p1 - C1);
        _m += rhoS * X * h * h;
      }
      double Dd = 0, nx = 0, ny = 0;
      if (!(sd > +2 * h || sd < -2 * h)) {
        const double HplusX = fabs(distPx) < EPSD ? 0.5 : (distPx < 0 ? 0 : 1);
        const double HminuX = fabs(distMx) < EPSD ? 0.5 : (distMx < 0 ? 0 : 1);
        const double HplusY = fabs(distPy) < EPSD ? 0.5 : (distPy < 0 ? 0 : 1);
        const double HminuY = fabs(distMy) < EPSD ? 0.5 : (distMy < 0 ? 0 : 1);
        const double gradUX = i2h * (distPx - distMx), gradUY = i2h * (distPy - distMy);
        const double gradHX = (HplusX - HminuX), gradHY = (HplusY - HminuY);
        const double gradUSq = gradUX * gradUX + gradUY * gradUY;
        const double denum = 1 / (gradUSq < EPSD ? EPSD : gradUSq);
        const double Dc = fac * (gradHX * gradUX + gradHY * gradUY) * denum;
        if (Dc > 0) {
          Dd = Dc;
          nx = -Dc * gradUX;
          ny = -Dc * gradUY;
          _ns += 1;
        }
      }
      o.sd[i] = Dd;
      o.sx[i] = nx;
      o.sy[i] = ny;
    }
    double a4[4] = {_x, _y, _m, _ns};
    block_sum_multi<NT, 4>(a4, lds);
    if (threadIdx.x == 0) {
      o.part[P_COM] = a4[0]; o.part[P_COM + 1] = a4[1]; o.part[P_COM + 2] = a4[2];
      o.part[P_NSURF] = a4[3];
    }
  }
  CELL_LOOP(i) {
    const size_t d = (size_t)(by * BS + i / BS) * g.NX + bx * BS + i % BS;
    chi[d] = s_chi[i];
    invrho[d] = s_inv[i];
    udx[d] = 0;
    udy[d] = 0;
  }
}

// centre of mass after putChiOnGrid (PutObjectsOnGrid.cpp:118-122) from the
// totals t[0..3) of the COM partials, without touching the stored state
__device__ inline void com_from_sums(const cup2d_shape_state* ss, int k, const double* t,
                                     double C[2], double* mass, bool* shifted) {
  const double _x = t[0], _y = t[1], _m = t[2];
  C[0] = ss[k].centerOfMass[0];
  C[1] = ss[k].centerOfMass[1];
  *shifted = _m > EPSD;
  *mass = _m;
  if (*shifted) {
    C[0] += _x / _m;
    C[1] += _y / _m;
  }
}

__device__ inline void shape_com(const Obstacles& ob, const cup2d_shape_state* ss, int k,
                                 double C[2], double* mass, bool* shifted) {
  double t[3];
  shape_sums<3>(ob, k, P_COM, t);
  com_from_sums(ss, k, t, C, mass, shifted);
}

// Shape::integrateObstBlock (Shape.cpp:117-155): per-slot partials about the
// updated centre of mass.
__device__ void obst_moments_wg(const Grid& g, const cup2d_shape_state* __restrict__ ss,
                                const Obstacles& ob, double* lds);

__global__ __launch_bounds__(NT) void k_obst_moments(Grid g, const cup2d_shape_state* __restrict__ ss,
                                                     const StepState* __restrict__ st,
                                                     Obstacles ob) {
  __shared__ double lds[LDS_RED];
  if ((int)blockIdx.x < ob.cnt[0]) obst_moments_wg(g, ss, ob, lds);
}

__device__ void obst_moments_wg(const Grid& g, const cup2d_shape_state* __restrict__ ss,
                                const Obstacles& ob, double* lds) {
  const int s = blockIdx.x, k = ob.slot_shape[s], b = ob.slot_block[s];
  const int bx = b % g.bpdx, by = b / g.bpdx;
  SlotView o = slot(ob, s);
  const double hsq = g.h * g.h;
  // the slot's cells are loaded before the centre of mass is reduced, so
  // the two memory latencies overlap
  constexpr int NC = BS * BS / NT;
  double cX[NC], cR[NC], cU[NC], cV[NC];
  CELL_LOOP(i) {
    cX[i_it] = o.chi[i];
    cR[i_it] = o.rho[i];
    cU[i_it] = o.udx[i];
    cV[i_it] = o.udy[i];
  }
  double C[2], mass;
  bool shifted;
  shape_com(ob, ss, k, C, &mass, &shifted);
  double a[7] = {0, 0, 0, 0, 0, 0, 0};
  CELL_LOOP(i) {
    if (cX[i_it] <= 0) continue;
    const int ix = i % BS, iy = i / BS;
    const double rhochi = cX[i_it] * cR[i_it] * hsq;
    const double p0 = cell_pos(bx, ix, g.hb, g.h) - C[0], p1 = cell_pos(by, iy, g.hb, g.h) - C[1];
    a[0] += rhochi * p0;
    a[1] += rhochi * p1;
    a[2] += rhochi;
    a[3] += rhochi * (p0 * p0 + p1 * p1);
    a[4] += rhochi * cU[i_it];
    a[5] += rhochi * cV[i_it];
    a[6] += rhochi * (p0 * cV[i_it] - p1 * cU[i_it]);
  }
  block_sum_multi<NT, 7>(a, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int j = 0; j < 7; ++j) o.part[P_MOM + j] = a[j];
}

// Shape::removeMoments (Shape.cpp:157-212) + putObjectVelOnGrid
// (PutObjectsOnGrid.cpp:149-184), per obstacle block, shapes in order.
__global__ __launch_bounds__(NT) void k_obst_udef(Grid g, cup2d_shape_state* __restrict__ ss,
                                                  const StepState* __restrict__ st, Obstacles ob,
                                                  const double* __restrict__ chi,
                                                  double* __restrict__ udx,
                                                  double* __restrict__ udy) {
  const int nub = ob.cnt[1];
  if ((int)blockIdx.x < nub) {
    const int b = ob.ub_block[blockIdx.x];
    const int bx = b % g.bpdx, by = b / g.bpdx;
    const int s0 = ob.ub_start[blockIdx.x], cnt = ob.ub_count[blockIdx.x];
    for (int q = 0; q < cnt; ++q) {
      const int s = ob.ub_slots[s0 + q], k = ob.slot_shape[s];
      SlotView o = slot(ob, s);
      // the block's cells are loaded before the shape's sums are reduced,
      // so the two memory latencies overlap
      constexpr int NC = BS * BS / NT;
      double cU[NC], cV[NC], cO[NC], cX[NC], dU[NC], dV[NC];
      CELL_LOOP(i) {
        const size_t d = (size_t)(by * BS + i / BS) * g.NX + bx * BS + i % BS;
        cU[i_it] = o.udx[i];
        cV[i_it] = o.udy[i];
        cO[i_it] = o.chi[i];
        cX[i_it] = chi[d];
        dU[i_it] = udx[d];
        dV[i_it] = udy[d];
      }
      static_assert(P_COM == 0 && P_MOM == 3, "COM and moment partials are adjacent");
      double t[10];
      shape_sums<10>(ob, k, P_COM, t);
      const double* I = t + P_MOM;
      const double Iu = I[4] / I[2], Iv = I[5] / I[2], Ia = I[6] / I[3];
      double C[2], mass;
      bool shifted;
      com_from_sums(ss, k, t, C, &mass, &shifted);
      CELL_LOOP(i) {
        const int ix = i % BS, iy = i / BS;
        const double p0 = cell_pos(bx, ix, g.hb, g.h) - C[0], p1 = cell_pos(by, iy, g.hb, g.h) - C[1];
        const double nu = cU[i_it] - (Iu - Ia * p1), nv = cV[i_it] - (Iv + Ia * p0);
        o.udx[i] = nu;
        o.udy[i] = nv;
        const size_t d = (size_t)(by * BS + iy) * g.NX + bx * BS + ix;
        if (cO[i_it] < cX[i_it] || cO[i_it] <= 0) continue;
        udx[d] = dU[i_it] + nu;
        udy[d] = dV[i_it] + nv;
      }
    }
  }
}


// one workgroup, after k_obst_udef: commit centre of mass (PutObjectsOnGrid.cpp:
// 118-122), M, J and d_gm (Shape.cpp:165-172) once every block has used the
// previous values
__device__ void obst_commit_wg(const Grid& g, cup2d_shape_state* __restrict__ ss,
                               const Obstacles& ob) {
  for (int k = threadIdx.x / 64; k < g.nshapes; k += NT / 64) {  // a shape per wave
    double C[2], mass;
    bool shifted;
    double t[10];
    shape_sums<10>(ob, k, P_COM, t);
    com_from_sums(ss, k, t, C, &mass, &shifted);
    const double m = t[P_MOM + 2], j = t[P_MOM + 3];
    double nsurf;
    shape_sums<1>(ob, k, P_NSURF, &nsurf);
    if (threadIdx.x % 64 == 0) {
      cup2d_shape_state& T = ss[k];
      T.centerOfMass[0] = C[0];
      T.centerOfMass[1] = C[1];
      T.M = m;
      T.J = j;
      T.n_surface_points = nsurf;
      T.n_obstacle_blocks = ob.slot_of_shape_start[k + 1] - ob.slot_of_shape_start[k];
      const double dCx = T.center[0] - T.centerOfMass[0], dCy = T.center[1] - T.centerOfMass[1];
      double sn, cs;
      det_sincos(T.orientation, &sn, &cs);
      T.d_gm[0] = dCx * cs + dCy * sn;
      T.d_gm[1] = -dCx * sn + dCy * cs;
    }
  }
}

__global__ __launch_bounds__(NT) void k_obst_commit(Grid g, cup2d_shape_state* __restrict__ ss,
                                                    Obstacles ob) {
  obst_commit_wg(g, ss, ob);
}

static int grid_cap(cup2d_ctx* c) { return c->ob.cap > 0 ? c->ob.cap : 1; }

// PutObjectsOnGrid::operator() (PutObjectsOnGrid.cpp:186-241).
// overlap: run everything after the list builder on the side stream; the
// caller joins with join_obstacles() before the next reader of obstacle data.
static int launch_forces(cup2d_ctx* c, hipStream_t s, const Obstacles& lists, int end_step,
                         const double* u, const double* v, int deferred);

// How much of the placement chain runs on the main stream before advDiff when
// the chain runs beside it on the side stream (stages of enqueue_put_objects:
// midline + list builder, PutFishOnBlocks, placement, moments, removeMoments,
// commit): with swimmers the midline and list builder, with disks nothing.
// Beside advDiff, which saturates HBM, these latency-bound kernels run 2-3x
// slower than alone; the split balances that against the main stream's wait.
constexpr int SPLIT_FISH = 1, SPLIT_DISKS = 0;

int enqueue_put_objects(cup2d_ctx* c, int compute_dt, int overlap) {
  const int G = grid_cap(c);
  // overlap 2: the chain on the main stream and advDiff beside it
  const bool chain_main = overlap == 2;
  // the previous whole step's ComputeForces runs after this step's obstacle
  // chain on the side stream, its step end in this step's list builder
  const bool defer_forces = overlap && c->forces_pending;
  if (!defer_forces) RC_OB(flush_forces(c));
  const int end_prev = c->end_pending ? 1 : 0;
  c->forces_pending = false;
  c->end_pending = false;
  RC_OB(join_forces(c));  // its lists are about to be rebuilt
  std::swap(c->ob, c->ob_prev);
  const bool fish = !c->hfish.empty();
  {
    TimerScope ts(c, T_PUT_OBJECTS);
    hipLaunchKernelGGL(k_obst_begin, dim3(1), dim3(NT), 0, c->stream, c->g, c->sp, c->ss, c->st,
                       c->ob, c->ob_prev, c->obmap, c->ubidx, c->fishd, c->maxpart, c->nwg,
                       c->ramp, compute_dt, c->graph_capture ? -1 : c->host_step, fish ? 1 : 0,
                       end_prev);
  }
  // overlap 2: advDiff needs only dt and U_inf from the first phase of the
  // list builder; it runs on the side stream beside the whole chain, which
  // stays on the main stream, so that no cross-stream hop lies on the chain.
  // The previous step's velocity stays with the deferred ComputeForces.
  const double* u_prev = c->u;
  const double* v_prev = c->v;
  auto launch_side = [&]() -> int {
    CUP2D_HIP_CHECK(c, hipEventRecord(c->ev_begin, c->stream));
    CUP2D_HIP_CHECK(c, hipStreamWaitEvent(c->aux, c->ev_begin, 0));
    RC_OB(enqueue_advdiff_on(c, 1, c->aux));
    RC_OB(enqueue_fade_sums_early(c, c->aux));
    // the forward row FFT of the RHS without the obstacle correction, which
    // needs no obstacle data: the pressure solve adds the correction's part
    if (c->g.nshapes > 0) RC_OB(enqueue_rows_early(c, c->aux));
    CUP2D_HIP_CHECK(c, hipEventRecord(c->ev_obst, c->aux));  // joined before the RHS
    c->obst_pending = true;
    return CUP2D_OK;
  };
  if (chain_main) RC_OB(launch_side());
  // The chain's stages in order; the first `split` of them run on the main
  // stream before advDiff, the rest on the side stream beside it.
  enum { ST_MIDLINE, ST_FISH_PLACE, ST_PLACE, ST_MOMENTS, ST_UDEF, ST_COMMIT, ST_N };
  const int split = !overlap || chain_main ? ST_N : (fish ? SPLIT_FISH : SPLIT_DISKS);
  bool on_aux = false;
  auto stream_for = [&](int stage) -> hipStream_t {
    if (stage < split) return c->stream;
    if (!on_aux) {  // first side-stream stage: the side stream joins here
      on_aux = true;
      if (hipEventRecord(c->ev_begin, c->stream) != hipSuccess ||
          hipStreamWaitEvent(c->aux, c->ev_begin, 0) != hipSuccess)
        return nullptr;
    }
    return c->aux;
  };
  hipStream_t s;
  // this step's time-only midline part runs here (on the main stream) unless
  // the previous step left it ready
  const bool mid_a_here = !c->midA_ready;
  if (fish) {
    if (!(s = stream_for(ST_MIDLINE))) return CUP2D_ERR_HIP;
    RC_OB(enqueue_fish_midline(c, s));
    hipLaunchKernelGGL(k_obst_begin, dim3(1), dim3(NT), 0, s, c->g, c->sp, c->ss, c->st, c->ob,
                       c->ob_prev, c->obmap, c->ubidx, c->fishd, c->maxpart, c->nwg, c->ramp, 0,
                       c->graph_capture ? -1 : c->host_step, 2, 0);
    if (chain_main && mid_a_here) CUP2D_HIP_CHECK(c, hipEventRecord(c->ev_lists, c->stream));
    if (!(s = stream_for(ST_FISH_PLACE))) return CUP2D_ERR_HIP;
    RC_OB(enqueue_fish_place(c, s));
  }
  if (c->g.nshapes > 0) {
    if (!(s = stream_for(ST_PLACE))) return CUP2D_ERR_HIP;
    hipLaunchKernelGGL(k_obst_place, dim3(2 * G), dim3(NT), 0, s, c->g, c->sp, c->ss, c->ob,
                       c->ob_prev, c->obmap, c->ubidx, c->chi, c->udx, c->udy, c->invrho);
    // Disks carry no deformation velocity (Disk::create leaves udef = 0), so
    // removeMoments and putObjectVelOnGrid would only add zeros: skipped
    // unless a shape kind with a deformation velocity is present.
    bool deforming = false;
    for (const ShapeParams& p : c->hsp) deforming |= p.kind != 0;
    if (!(s = stream_for(ST_MOMENTS))) return CUP2D_ERR_HIP;
    hipLaunchKernelGGL(k_obst_moments, dim3(G), dim3(NT), 0, s, c->g, c->ss, c->st, c->ob);
    if (deforming) {
      if (!(s = stream_for(ST_UDEF))) return CUP2D_ERR_HIP;
      hipLaunchKernelGGL(k_obst_udef, dim3(G), dim3(NT), 0, s, c->g, c->ss, c->st, c->ob, c->chi,
                         c->udx, c->udy);
    }
    if (!(s = stream_for(ST_COMMIT))) return CUP2D_ERR_HIP;
    // chain on the main stream: the commit rides in the obstacle RHS launch,
    // the next kernel that reads nothing it writes
    if (chain_main)
      c->commit_pending = true;
    else
      hipLaunchKernelGGL(k_obst_commit, dim3(1), dim3(NT), 0, s, c->g, c->ss, c->ob);
  }
  if (chain_main) {
    // The next step's midline part, then the previous step's ComputeForces,
    // on the midline stream; ev_frc, which the main stream waits for before
    // the pressure is overwritten, orders both before the rest of the step.
    // The midline part shares work arrays only with this step's own (ev_lists,
    // when that ran here); ComputeForces reads the previous velocity (which
    // advDiff only reads), pressure and lists, and writes only force fields
    // of the shape state besides the two block/surface counts, which the
    // commit also sets and the step's own ComputeForces sets last.
    RC_OB(enqueue_fish_mid_next(c, c->host_step & 1, mid_a_here ? c->ev_lists : c->ev_begin,
                                false));
    if (defer_forces) RC_OB(launch_forces(c, c->mids, c->ob_prev, 0, u_prev, v_prev, 1));
    CUP2D_HIP_CHECK(c, hipEventRecord(c->ev_frc, c->mids));
    c->frc_join = true;
    return launch_check(c, "put_objects");
  }
  // the side stream always joins (the deferred ComputeForces and the next
  // step's midline are ordered after this step's list builder through it)
  if (overlap && !stream_for(ST_N)) return CUP2D_ERR_HIP;
  if (overlap) {
    CUP2D_HIP_CHECK(c, hipEventRecord(c->ev_obst, c->aux));
    c->obst_pending = true;
    // the swimmers' time-only midline part of the next step on the midline
    // stream, joined here by the side stream, so that ev_frc (which the main
    // stream waits for before the pressure is overwritten) also orders the
    // next step's midline kernel after it
    RC_OB(enqueue_fish_mid_next(c, c->host_step & 1, c->ev_begin, true));
    if (defer_forces)  // the previous step's ComputeForces, on its own lists
      RC_OB(launch_forces(c, c->aux, c->ob_prev, 0, u_prev, v_prev, 1));
    if (defer_forces || !c->hfish.empty()) {
      CUP2D_HIP_CHECK(c, hipEventRecord(c->ev_frc, c->aux));
      c->frc_join = true;
    }
  }
  return launch_check(c, "put_objects");
}

// a placement commit still waiting for its obstacle RHS launch, on its own
int flush_commit(cup2d_ctx* c) {
  if (!c->commit_pending) return CUP2D_OK;
  c->commit_pending = false;
  hipLaunchKernelGGL(k_obst_commit, dim3(1), dim3(NT), 0, c->stream, c->g, c->ss, c->ob);
  return launch_check(c, "placement commit");
}

int join_obstacles(cup2d_ctx* c) {
  if (!c->obst_pending) return CUP2D_OK;
  CUP2D_HIP_CHECK(c, hipStreamWaitEvent(c->stream, c->ev_obst, 0));
  c->obst_pending = false;
  return CUP2D_OK;
}

// A block's cells for the RHS obstacle correction, every load issued at once
// (one memory round trip instead of one per cell the branch takes): the
// dense chi and its four neighbours (BlockLabOpen ghosts), the slot's chi and
// the RHS.
struct CellStage {
  static constexpr int NC = BS * BS / NT;
  double X[NC], oX[NC], xe[NC], xw[NC], xn[NC], xs[NC], r[NC];
  template <typename RT>
  __device__ void load(const Grid& g, int bx, int by, const double* __restrict__ chi,
                       const double* __restrict__ ochi, const RT* r_) {
    CELL_LOOP(i) {
      const int gx = bx * BS + i % BS, gy = by * BS + i / BS;
      const size_t d = (size_t)gy * g.NX + gx;
      X[i_it] = chi[d];
      oX[i_it] = ochi[i];
      xe[i_it] = lab_at(chi, g, gx + 1, gy);
      xw[i_it] = lab_at(chi, g, gx - 1, gy);
      xn[i_it] = lab_at(chi, g, gx, gy + 1);
      xs[i_it] = lab_at(chi, g, gx, gy - 1);
      r[i_it] = (double)r_[d];
    }
  }
  // the cell lies on the perimeter (a nonzero chi gradient)
  __device__ bool perim(int c) const {
    const double gradXx = xe[c] - xw[c], gradXy = xn[c] - xs[c];
    return gradXx * gradXx + gradXy * gradXy > 0;
  }
};

// ------------------------------------------------ RHS obstacle correction
// PressureSingle.cpp:130-168 (source term + perimeter sums) and :173-212
// (sign-selective rescale), per distinct block, shapes in order.
// With u != nullptr the workgroup first forms the RHS of its block from the
// advDiff output (rhs_cell: PressureSingle.cpp:112-128 with the deferred
// boundary-flux correction), the cells the whole-step row FFT takes from the
// RHS plane.
// RT: the type of the RHS and velocity planes (double, or float at FP32
// precision); the corrections are computed in double and stored rounded.
template <typename RT>
__global__ __launch_bounds__(NT) void k_rhs_obst1(Grid g, StepState* __restrict__ st,
                                                  Obstacles ob, const double* __restrict__ chi,
                                                  RT* __restrict__ r,
                                                  const RT* __restrict__ u,
                                                  const RT* __restrict__ v,
                                                  cup2d_shape_state* __restrict__ commit_ss) {
  __shared__ double lds[LDS_RED];
  // k_rhs_obst2 sets it again when an obstacle block lies in the fade band
  if (blockIdx.x == 0 && threadIdx.x == 0) st->rows_redo = 0;
  if (commit_ss && blockIdx.x == gridDim.x - 1) {  // the placement commit (k_obst_commit)
    obst_commit_wg(g, commit_ss, ob);
    return;
  }
  if ((int)blockIdx.x >= ob.cnt[1]) return;
  const int b = ob.ub_block[blockIdx.x];
  const int bx = b % g.bpdx, by = b / g.bpdx;
  if (u) {
    const RT corr = (RT)st->flux_corr;  // deferred advDiff flux correction
    const RT facDiv = (RT)(0.5 * g.h / st->dt);
    CELL_LOOP(i) {
      const int x = bx * BS + i % BS, y = by * BS + i / BS;
      r[(size_t)y * g.NX + x] = rhs_cell<RT>(u, v, g.NX, g.NY, x, y, corr, facDiv);
    }
  }
  const int s0 = ob.ub_start[blockIdx.x], cnt = ob.ub_count[blockIdx.x];
  for (int q = 0; q < cnt; ++q) {
    const int s = ob.ub_slots[s0 + q];
    SlotView o = slot(ob, s);
    double sum = 0, pos = 0, neg = 0;
    CellStage cs;
    cs.load(g, bx, by, chi, o.chi, r);
    CELL_LOOP(i) {
      const size_t d = (size_t)(by * BS + i / BS) * g.NX + bx * BS + i % BS;
      if (cs.X[i_it] > cs.oX[i_it]) continue;
      const double rd = cs.r[i_it];
      const double srcBulk = -cs.oX[i_it] * rd;
      sum += srcBulk;
      const RT rn = (RT)(rd + srcBulk);
      r[d] = rn;
      const double isPerim = cs.perim(i_it) ? 1.0 : 0.0;
      const double T = (double)rn;
      pos += isPerim * T * (double)(T > 0);
      neg -= isPerim * T * (double)(T < 0);
    }
    double a3[3] = {sum, pos, neg};
    block_sum_multi<NT, 3>(a3, lds);
    if (threadIdx.x == 0) {
      o.part[P_RHS] = a3[0]; o.part[P_RHS + 1] = a3[1]; o.part[P_RHS + 2] = a3[2];
    }
  }
}

// band_index != nullptr: the fade sums were formed from the advDiff output
// before the obstacle RHS; the partial of an obstacle block in the fade band
// is formed again here from its final RHS.
template <typename RT>
__global__ __launch_bounds__(NT) void k_rhs_obst2(Grid g, StepState* __restrict__ st,
                                                  Obstacles ob, const double* __restrict__ chi,
                                                  RT* __restrict__ r,
                                                  const int* __restrict__ band_index,
                                                  double* __restrict__ fadepart) {
  static_assert(NT == 256, "fade_block_sums is a 256-thread reduction");
  __shared__ double lds[12];
  if ((int)blockIdx.x >= ob.cnt[1]) return;
  const int b = ob.ub_block[blockIdx.x];
  const int bx = b % g.bpdx, by = b / g.bpdx;
  const int s0 = ob.ub_start[blockIdx.x], cnt = ob.ub_count[blockIdx.x];
  for (int q = 0; q < cnt; ++q) {
    const int s = ob.ub_slots[s0 + q], k = ob.slot_shape[s];
    SlotView o = slot(ob, s);
    // the block's cells are loaded before the shape's sums are reduced, so
    // the two memory latencies overlap
    CellStage cs;
    cs.load(g, bx, by, chi, o.chi, r);
    double t[3];
    shape_sums<3>(ob, k, P_RHS, t);
    const double sum = t[0], pos = t[1], neg = t[2];
    const double corrDenom = sum > 0 ? pos : neg;
    const double corr = sum / dmax(corrDenom, EPSD);
    CELL_LOOP(i) {
      const size_t d = (size_t)(by * BS + i / BS) * g.NX + bx * BS + i % BS;
      if (cs.X[i_it] > cs.oX[i_it]) continue;
      const bool isPerim = cs.perim(i_it);
      const double T = cs.r[i_it];
      if (isPerim && T > 0 && corr > 0) r[d] = (RT)(T - corr * T);
      else if (isPerim && T < 0 && corr < 0) r[d] = (RT)(T + corr * T);
    }
  }
  const int j = band_index ? band_index[b] : -1;
  if (j < 0) return;
  if (threadIdx.x == 0) st->rows_redo = 1;  // an early row pass is redone
  __syncthreads();  // the block's RHS is final
  double a[3];
  fade_block_sums<RT>(g, b, r, nullptr, nullptr, true, (RT)0, (RT)0, a, lds);
  if (threadIdx.x == 0) {
    fadepart[3 * j] = a[0];
    fadepart[3 * j + 1] = a[1];
    fadepart[3 * j + 2] = a[2];
  }
}

int enqueue_rhs_obstacles(cup2d_ctx* c, int compute) {
  if (c->g.nshapes == 0) return CUP2D_OK;
  const int G = grid_cap(c);
  TimerScope ts(c, T_RHS_OBST);
  // a pending placement commit takes one more workgroup
  const int G1 = c->commit_pending ? G + 1 : G;
  CUP2D_REAL_DISPATCH(c, hipLaunchKernelGGL(k_rhs_obst1<T>, dim3(G1), dim3(NT), 0, c->stream,
                                            c->g, c->st, c->ob, c->chi, tp<T>(c->rhs),
                                            compute ? tp<T>(c->u) : nullptr, tp<T>(c->v),
                                            c->commit_pending ? c->ss : nullptr));
  c->commit_pending = false;
  CUP2D_REAL_DISPATCH(c, hipLaunchKernelGGL(k_rhs_obst2<T>, dim3(G), dim3(NT), 0, c->stream,
                                            c->g, c->st, c->ob, c->chi, tp<T>(c->rhs),
                                            c->fade_early ? c->band_index : nullptr, c->red));
  return launch_check(c, "rhs_obstacles");
}

// ------------------------------------------------------------ body update
// 3x3 LU with partial pivoting (restates gsl_linalg_LU_decomp/LU_solve as
// called in Shape::updateVelocity, Shape.cpp:48-54).
__device__ void lu3_solve(double A[3][3], const double b[3], double x[3]) {
  int perm[3] = {0, 1, 2};
  for (int j = 0; j < 2; ++j) {
    double amax = fabs(A[j][j]);
    int ip = j;
    for (int i = j + 1; i < 3; ++i)
      if (fabs(A[i][j]) > amax) { amax = fabs(A[i][j]); ip = i; }
    if (ip != j) {
      for (int k = 0; k < 3; ++k) { const double t = A[j][k]; A[j][k] = A[ip][k]; A[ip][k] = t; }
      const int t = perm[j]; perm[j] = perm[ip]; perm[ip] = t;
    }
    const double ajj = A[j][j];
    if (ajj != 0.0)
      for (int i = j + 1; i < 3; ++i) {
        const double aij = A[i][j] / ajj;
        A[i][j] = aij;
        for (int k = j + 1; k < 3; ++k) A[i][k] -= aij * A[j][k];
      }
  }
  double y[3];
  for (int i = 0; i < 3; ++i) {
    double v = b[perm[i]];
    for (int k = 0; k < i; ++k) v -= A[i][k] * y[k];
    y[i] = v;
  }
  for (int i = 2; i >= 0; --i) {
    double v = y[i];
    for (int k = i + 1; k < 3; ++k) v -= A[i][k] * x[k];
    x[i] = v / A[i][i];
  }
}

// Shape::updateVelocity (Shape.cpp:25-59), Disk::updateVelocity
// (ShapesSimple.cpp:83-87), preventCollidingObstacles (PressureSingle.cpp:243-344).
// Shape::updateVelocity (Shape.cpp:25-59) from the penalization sums
// S = penalM penalJ penalDX penalDY fluidMomX fluidMomY fluidAngMom; uvw holds
// the current u v omega on entry (kept where forced) and the new ones on exit.
__device__ inline void shape_velocity(const ShapeParams& P, const double S[7], double dt,
                                      double time, double uvw[3]) {
  const double penalM = S[0], penalJ = S[1], penalDX = S[2], penalDY = S[3];
  double A[3][3] = {{penalM, 0, -penalDY}, {0, penalM, penalDX}, {-penalDY, penalDX, penalJ}};
  double bb[3] = {S[4] + dt * P.appliedForceX, S[5] + dt * P.appliedForceY,
                  S[6] + dt * P.appliedTorque};
  const double tf = P.timeForced;
  if (P.bForcedx && time < tf) { A[0][1] = 0; A[0][2] = 0; bb[0] = penalM * P.forcedu; }
  if (P.bForcedy && time < tf) { A[1][0] = 0; A[1][2] = 0; bb[1] = penalM * P.forcedv; }
  if (P.bBlockAng && time < tf) { A[2][0] = 0; A[2][1] = 0; bb[2] = penalJ * P.forcedomega; }
  double x[3];
  lu3_solve(A, bb, x);
  if (!P.bForcedy || time > tf) uvw[0] = x[0];  // Shape.cpp:56 tests bForcedy for u
  if (!P.bForcedy || time > tf) uvw[1] = x[1];
  if (!P.bBlockAng || time > tf) uvw[2] = x[2];
  if (P.tAccel > 0) {
    if (P.bForcedx && time < P.tAccel) uvw[0] = (time / P.tAccel) * P.forcedu;
    if (P.bForcedy && time < P.tAccel) uvw[1] = (time / P.tAccel) * P.forcedv;
  }
}

__device__ void body_solve_wg(const Grid& g, const ShapeParams* __restrict__ sp,
                              cup2d_shape_state* __restrict__ ss,
                              StepState* __restrict__ st, const Obstacles& ob,
                              const int* __restrict__ ubidx, ForceSnap* __restrict__ snap,
                              int collide, const SurfVel& surf) {
  __shared__ double lds[LDS_RED];
  __shared__ double coll[MAX_SHAPES][10];
  const double dt = st->dt, time = st->time;
  const int ns = g.nshapes;
  for (int k = threadIdx.x / 64; k < ns; k += NT / 64) {  // a shape per wave
    double S[7];
    shape_sums<7>(ob, k, P_PEN, S);
    if (threadIdx.x % 64 == 0) {
      cup2d_shape_state& T = ss[k];
      double uvw[3] = {T.u, T.v, T.omega};
      shape_velocity(sp[k], S, dt, time, uvw);
      T.penalM = S[0]; T.penalJ = S[1]; T.penalDX = S[2]; T.penalDY = S[3];
      T.fluidMomX = S[4]; T.fluidMomY = S[5]; T.fluidAngMom = S[6];
      T.u = uvw[0]; T.v = uvw[1]; T.omega = uvw[2];
    }
  }
  if (ns < 2 || !collide) return;
  __syncthreads();
  // collision sums: coll[i] accumulates over every j != i (PressureSingle.cpp:255-298),
  // over the blocks the two share.  No block holding two slots means no
  // overlap at all; otherwise the slots of shape i on a shared block are
  // found NT at a time (a ballot per wave) and each one finds the other
  // shapes' slots of its block through the distinct-block lists.
  if (ob.cnt[0] == ob.cnt[1]) return;
  __shared__ unsigned long long s_mask[NT / 64];
  for (int i = 0; i < ns; ++i) {
    double acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const cup2d_shape_state& si = ss[i];
    const int a0 = ob.slot_of_shape_start[i], a1 = ob.slot_of_shape_start[i + 1];
    for (int base = a0; base < a1; base += NT) {
      const int mine = base + threadIdx.x;
      const bool shared = mine < a1 && ob.ub_count[ubidx[ob.slot_block[mine]]] > 1;
      const unsigned long long m = __ballot(shared);
      __syncthreads();  // the previous chunk's masks are read
      if (threadIdx.x % 64 == 0) s_mask[threadIdx.x / 64] = m;
      __syncthreads();
      for (int w = 0; w < NT / 64; ++w)
        for (unsigned long long mw = s_mask[w]; mw != 0; mw &= mw - 1) {
          const int a = base + 64 * w + __builtin_ctzll(mw);
          const int blk = ob.slot_block[a], bx = blk % g.bpdx, by = blk / g.bpdx;
          const int u = ubidx[blk], q0 = ob.ub_start[u], nq = ob.ub_count[u];
          SlotView oi = slot(ob, a);
          for (int q = 0; q < nq; ++q) {
            const int bsl = ob.ub_slots[q0 + q], j = ob.slot_shape[bsl];
            if (j == i) continue;
            const cup2d_shape_state& sj = ss[j];
            SlotView oj = slot(ob, bsl);
            CELL_LOOP(c) {
              if (oi.chi[c] <= 0 || oj.chi[c] <= 0) continue;
              const double p0 = cell_pos(bx, c % BS, g.hb, g.h), p1 = cell_pos(by, c / BS, g.hb, g.h);
              const double iUr = -si.omega * (p1 - si.centerOfMass[1]), iVr = si.omega * (p0 - si.centerOfMass[0]);
              const double jUr = -sj.omega * (p1 - sj.centerOfMass[1]), jVr = sj.omega * (p0 - sj.centerOfMass[0]);
              const double X = oi.chi[c], Y = oj.chi[c];
              double iUd = oi.udx[c], iVd = oi.udy[c], jUd = oj.udx[c], jVd = oj.udy[c];
              if (i == surf.shape || j == surf.shape) {
                double sx, sy;
                if (i == surf.shape) {
                  surf_vel(surf, p0 - si.centerOfMass[0], p1 - si.centerOfMass[1], &sx, &sy);
                  iUd += sx; iVd += sy;
                } else {
                  surf_vel(surf, p0 - sj.centerOfMass[0], p1 - sj.centerOfMass[1], &sx, &sy);
                  jUd += sx; jVd += sy;
                }
              }
              acc[0] += X; acc[1] += X * p0; acc[2] += X * p1;
              acc[3] += X * (si.u + iUr + iUd);
              acc[4] += X * (si.v + iVr + iVd);
              acc[5] += Y; acc[6] += Y * p0; acc[7] += Y * p1;
              acc[8] += Y * (sj.u + jUr + jUd);
              acc[9] += Y * (sj.v + jVr + jVd);
            }
          }
        }
    }
    block_sum_multi<NT, 10>(acc, lds);
    if (threadIdx.x == 0)
      for (int q = 0; q < 10; ++q) coll[i][q] = acc[q];
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int i = 0; i < ns; ++i)
    for (int j = 0; j < ns; ++j) {
      if (i == j) continue;
      const double* c = coll[i];
      if (c[0] < 1 || c[5] < 1) continue;
      const double iPX = c[1] / c[0], iPY = c[2] / c[0], iUX = c[3] / c[0], iUY = c[4] / c[0];
      const double jPX = c[6] / c[5], jPY = c[7] / c[5], jUX = c[8] / c[5], jUY = c[9] / c[5];
      const double CX = (iPX + jPX) / 2, CY = (iPY + jPY) / 2;
      const double dirX = iPX - jPX, dirY = iPY - jPY;
      const double hitVelX = jUX - iUX, hitVelY = jUY - iUY;
      const double normF = dmax(sqrt(dirX * dirX + dirY * dirY), EPSD);
      const double NXn = dirX / normF, NYn = dirY / normF;
      const double projVel = hitVelX * NXn + hitVelY * NYn;
      if (projVel <= 0) continue;
      const ShapeParams &Pi = sp[i], &Pj = sp[j];
      const bool iForcedX = Pi.bForcedx && time < Pi.timeForced;
      const bool iForcedY = Pi.bForcedy && time < Pi.timeForced;
      const bool iForcedA = Pi.bBlockAng && time < Pi.timeForced;
      const bool jForcedX = Pj.bForcedx && time < Pj.timeForced;
      const bool jForcedY = Pj.bForcedy && time < Pj.timeForced;
      const double iInvMassX = iForcedX ? 0 : 1 / ss[i].M, iInvMassY = iForcedY ? 0 : 1 / ss[i].M;
      const double jInvMassX = jForcedX ? 0 : 1 / ss[j].M, jInvMassY = jForcedY ? 0 : 1 / ss[j].M;
      const double iInvJ = iForcedA ? 0 : 1 / ss[i].J;
      const double meanMassX = 2 / dmax(iInvMassX + jInvMassX, EPSD);
      const double meanMassY = 2 / dmax(iInvMassY + jInvMassY, EPSD);
      const double FXdt = meanMassX * projVel * NXn, FYdt = meanMassY * projVel * NYn;
      ss[i].u += FXdt * iInvMassX;
      ss[i].v += FYdt * iInvMassY;
      const double iCx = ss[i].centerOfMass[0], iCy = ss[i].centerOfMass[1];
      const double RcrossF = (CX - iCx) * FYdt - (CY - iCy) * FXdt;
      ss[i].omega += iInvJ * RcrossF;
    }
}

// PressureSingle::integrateMomenta (PressureSingle.cpp:23-67), per slot.
template <typename RT>
__device__ void momenta_wg(const Grid& g, const cup2d_shape_state* __restrict__ ss,
                           const StepState* __restrict__ st, const Obstacles& ob,
                           const RT* __restrict__ u, const RT* __restrict__ v, double* lds,
                           const SurfVel& sv) {
  const int s = blockIdx.x, k = ob.slot_shape[s], b = ob.slot_block[s];
  const int bx = b % g.bpdx, by = b / g.bpdx;
  SlotView o = slot(ob, s);
  const double hsq = g.h * g.h;
  const double lambdt = g.lambda * st->dt;
  const double Cx = ss[k].centerOfMass[0], Cy = ss[k].centerOfMass[1];
  double a[7] = {0, 0, 0, 0, 0, 0, 0};
  // every load of the block first (one memory round trip), then the sums
  // over the cells inside the body
  constexpr int NC = BS * BS / NT;
  double cX[NC], cU[NC], cV[NC], cR[NC];
  CELL_LOOP(i) {
    const size_t d = (size_t)(by * BS + i / BS) * g.NX + bx * BS + i % BS;
    cX[i_it] = o.chi[i];
    cU[i_it] = (double)u[d] - o.udx[i];
    cV[i_it] = (double)v[d] - o.udy[i];
    cR[i_it] = o.rho[i];
  }
  CELL_LOOP(i) {
    if (cX[i_it] <= 0) continue;
    const int ix = i % BS, iy = i / BS;
    double ud0 = cU[i_it], ud1 = cV[i_it];
    if (k == sv.shape) {
      double sx, sy;
      surf_vel(sv, cell_pos(bx, ix, g.hb, g.h) - Cx, cell_pos(by, iy, g.hb, g.h) - Cy, &sx, &sy);
      ud0 -= sx;
      ud1 -= sy;
    }
    double F;
    if (!g.implicit_penal) {
      F = hsq * cR[i_it] * cX[i_it];
    } else {
      const double Xlamdt = cX[i_it] * lambdt;
      F = hsq * cR[i_it] * Xlamdt / (1 + Xlamdt);
    }
    const double p0 = cell_pos(bx, ix, g.hb, g.h) - Cx, p1 = cell_pos(by, iy, g.hb, g.h) - Cy;
    a[0] += F;
    a[1] += F * (p0 * p0 + p1 * p1);
    a[2] += F * p0;
    a[3] += F * p1;
    a[4] += F * ud0;
    a[5] += F * ud1;
    a[6] += F * (p0 * ud1 - p1 * ud0);
  }
  block_sum_multi<NT, 7>(a, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int j = 0; j < 7; ++j) o.part[P_PEN + j] = a[j];
}

// collide: also preventCollidingObstacles (PressureSingle only).
__device__ void body_update_tail(const Grid& g, const ShapeParams* __restrict__ sp,
                                 cup2d_shape_state* __restrict__ ss, StepState* __restrict__ st,
                                 const Obstacles& ob, const int* __restrict__ ubidx,
                                 ForceSnap* __restrict__ snap, int collide,
                                 const SurfVel& surf) {
  body_solve_wg(g, sp, ss, st, ob, ubidx, snap, collide, surf);
  __syncthreads();
  // state ComputeForces uses (it may run after the next step moved the bodies)
  if (threadIdx.x < g.nshapes) {
    const cup2d_shape_state& T = ss[threadIdx.x];
    ForceSnap f;
    f.C[0] = T.centerOfMass[0]; f.C[1] = T.centerOfMass[1];
    f.u = T.u; f.v = T.v; f.omega = T.omega;
    f.time = st->time; f.step = st->step; f.pad_ = 0;
    snap[threadIdx.x] = f;
  }
  if (threadIdx.x == 0) st->dt_done = st->dt;
}

__global__ __launch_bounds__(NT) void k_body_solve(Grid g, const ShapeParams* __restrict__ sp,
                                                   cup2d_shape_state* __restrict__ ss,
                                                   StepState* __restrict__ st, Obstacles ob,
                                                   const int* __restrict__ ubidx,
                                                   ForceSnap* __restrict__ snap, int collide,
                                                   SurfVel surf) {
  body_update_tail(g, sp, ss, st, ob, ubidx, snap, collide, surf);
}

// PressureSingle::integrateMomenta (PressureSingle.cpp:23-67) per slot.
template <typename RT>
__global__ __launch_bounds__(NT) void k_momenta(Grid g, const ShapeParams* __restrict__ sp,
                                                cup2d_shape_state* __restrict__ ss,
                                                StepState* __restrict__ st, Obstacles ob,
                                                const RT* __restrict__ u,
                                                const RT* __restrict__ v, SurfVel sv) {
  __shared__ double lds[LDS_RED];
  if ((int)blockIdx.x < ob.cnt[0]) momenta_wg(g, ss, st, ob, u, v, lds, sv);
}

int enqueue_body_update(cup2d_ctx* c) {
  if (c->g.nshapes == 0) return CUP2D_OK;
  const int G = grid_cap(c);
  TimerScope ts(c, T_BODY);
  CUP2D_REAL_DISPATCH(c, hipLaunchKernelGGL(k_momenta<T>, dim3(G), dim3(NT), 0, c->stream, c->g,
                                            c->sp, c->ss, c->st, c->ob, tp<T>(c->u),
                                            tp<T>(c->v), surf_of(c)));
  hipLaunchKernelGGL(k_body_solve, dim3(1), dim3(NT), 0, c->stream, c->g, c->sp, c->ss, c->st,
                     c->ob, c->ubidx, c->fsnap, 1, surf_of(c));
  return launch_check(c, "body_update");
}

// PressureSingle::penalize (PressureSingle.cpp:69-108), per distinct block,
// shapes in order.  The block's velocities stay in registers across its
// slots and are stored once; the findMaxU partial of the block comes from
// the same registers.
//
// SOLVE: the body update's solve (Shape::updateVelocity) is done here by
// every workgroup, from the momenta partials in the same fixed order, so the
// separate single-workgroup solve kernel and its launch are saved (used with
// fewer than two shapes, where no collision step follows the solve).
// Workgroup 0 stores the shape state, the ComputeForces snapshot and
// dt_done; where a shape is forced it rewrites the value every workgroup read.
template <bool SOLVE, typename RT>
__global__ __launch_bounds__(NT) void k_penalize(Grid g, const ShapeParams* __restrict__ sp,
                                                 cup2d_shape_state* __restrict__ ss,
                                                 StepState* __restrict__ st, Obstacles ob,
                                                 const double* __restrict__ chi,
                                                 RT* __restrict__ u, RT* __restrict__ v,
                                                 double* __restrict__ maxpart,
                                                 ForceSnap* __restrict__ snap, SurfVel sv) {
  __shared__ double lds[LDS_RED];
  __shared__ double s_uvw[MAX_SHAPES][3];
  const int nub = ob.cnt[1];
  if ((int)blockIdx.x >= nub && !(SOLVE && blockIdx.x == 0)) return;
  const double dt = st->dt;
  // the block's velocities and chi are loaded first, so that their memory
  // latency overlaps the solve's reduction of the partials
  const bool mine = (int)blockIdx.x < nub;
  const int b = mine ? ob.ub_block[blockIdx.x] : 0;
  const int bx = b % g.bpdx, by = b / g.bpdx;
  constexpr int NC = BS * BS / NT;
  double uR[NC], vR[NC], cR[NC];
  // and the block's first slot (the only one unless shapes overlap)
  const int s0 = mine ? ob.ub_start[blockIdx.x] : 0, cnt = mine ? ob.ub_count[blockIdx.x] : 0;
  const int sf = mine ? ob.ub_slots[s0] : 0;
  double fX[NC], fU[NC], fV[NC];
  if (mine) {
    const SlotView o = slot(ob, sf);
    CELL_LOOP(i) {
      const size_t d = (size_t)(by * BS + i / BS) * g.NX + bx * BS + i % BS;
      uR[i_it] = u[d];
      vR[i_it] = v[d];
      cR[i_it] = chi[d];
      fX[i_it] = o.chi[i];
      fU[i_it] = o.udx[i];
      fV[i_it] = o.udy[i];
    }
  }
  if (SOLVE) {
    const double time = st->time;
    for (int k = 0; k < g.nshapes; ++k) {
      double S[7];
      shape_sums<7>(ob, k, P_PEN, S);
      if (threadIdx.x == 0) {
        cup2d_shape_state& T = ss[k];
        double uvw[3] = {T.u, T.v, T.omega};
        shape_velocity(sp[k], S, dt, time, uvw);
        s_uvw[k][0] = uvw[0]; s_uvw[k][1] = uvw[1]; s_uvw[k][2] = uvw[2];
        if (blockIdx.x == 0) {
          T.penalM = S[0]; T.penalJ = S[1]; T.penalDX = S[2]; T.penalDY = S[3];
          T.fluidMomX = S[4]; T.fluidMomY = S[5]; T.fluidAngMom = S[6];
          T.u = uvw[0]; T.v = uvw[1]; T.omega = uvw[2];
        }
      }
    }
    __syncthreads();
    if (blockIdx.x == 0) {
      // state ComputeForces uses (it may run after the next step moved the bodies)
      if ((int)threadIdx.x < g.nshapes) {
        const int k = threadIdx.x;
        ForceSnap f;
        f.C[0] = ss[k].centerOfMass[0]; f.C[1] = ss[k].centerOfMass[1];
        f.u = s_uvw[k][0]; f.v = s_uvw[k][1]; f.omega = s_uvw[k][2];
        f.time = st->time; f.step = st->step; f.pad_ = 0;
        snap[k] = f;
      }
      if (threadIdx.x == 0) st->dt_done = dt;
      if (nub == 0) return;
    }
  }
  for (int q = 0; q < cnt; ++q) {
    const int s = q == 0 ? sf : ob.ub_slots[s0 + q], k = ob.slot_shape[s];
    SlotView o = slot(ob, s);
    const double u_s = SOLVE ? s_uvw[k][0] : ss[k].u;
    const double v_s = SOLVE ? s_uvw[k][1] : ss[k].v;
    const double omega_s = SOLVE ? s_uvw[k][2] : ss[k].omega;
    const double Cx = ss[k].centerOfMass[0], Cy = ss[k].centerOfMass[1];
    // the slot's loads of all cells first (one memory round trip; the first
    // slot's were issued before the solve)
    double sX[NC], sU[NC], sV[NC];
    CELL_LOOP(i) {
      sX[i_it] = q == 0 ? fX[i_it] : o.chi[i];
      sU[i_it] = q == 0 ? fU[i_it] : o.udx[i];
      sV[i_it] = q == 0 ? fV[i_it] : o.udy[i];
    }
    CELL_LOOP(i) {
      const int ix = i % BS, iy = i / BS;
      const double X = sX[i_it];
      if (cR[i_it] > X) continue;
      if (X <= 0) continue;
      const double p0 = cell_pos(bx, ix, g.hb, g.h) - Cx, p1 = cell_pos(by, iy, g.hb, g.h) - Cy;
      const double alpha = g.implicit_penal ? 1 / (1 + g.lambda * dt * X) : 1 - X;
      double US = u_s - omega_s * p1 + sU[i_it];
      double VS = v_s + omega_s * p0 + sV[i_it];
      if (k == sv.shape) {
        double sx, sy;
        surf_vel(sv, p0, p1, &sx, &sy);
        US += sx;
        VS += sy;
      }
      uR[i_it] = alpha * uR[i_it] + (1 - alpha) * US;
      vR[i_it] = alpha * vR[i_it] + (1 - alpha) * VS;
    }
  }
  CELL_LOOP(i) {
    const size_t d = (size_t)(by * BS + i / BS) * g.NX + bx * BS + i % BS;
    u[d] = (RT)uR[i_it];
    v[d] = (RT)vR[i_it];
  }
  if (!maxpart) return;
  // findMaxU (Helpers.cpp:160-175) over this block's final velocities; the
  // projection kernel skips obstacle blocks.
  const double UINF = st->uinf[0], VINF = st->uinf[1];
  double m = 0;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    // the stored values
    const double a = (double)(RT)uR[c], w = (double)(RT)vR[c];
    m = nmax(m, fabs(a + UINF));
    m = nmax(m, fabs(w + VINF));
    m = nmax(m, fabs(a));
    m = nmax(m, fabs(w));
  }
  m = block_max<NT>(m, lds);
  if (threadIdx.x == 0) maxpart[blockIdx.x] = m;
}

int enqueue_penalize(cup2d_ctx* c, double* maxpart) {
  if (c->g.nshapes == 0) return CUP2D_OK;
  TimerScope ts(c, T_PENALIZE);
  CUP2D_REAL_DISPATCH(c, hipLaunchKernelGGL((k_penalize<false, T>), dim3(grid_cap(c)), dim3(NT), 0,
                                            c->stream, c->g, c->sp, c->ss, c->st, c->ob, c->chi,
                                            tp<T>(c->u), tp<T>(c->v), maxpart, c->fsnap,
                                            surf_of(c)));
  return launch_check(c, "penalize");
}

// PressureSingle's body update + penalization (PressureSingle.cpp:360-368):
// with fewer than two shapes the solve runs inside the penalization kernel.
int enqueue_body_penalize(cup2d_ctx* c, double* maxpart) {
  if (c->g.nshapes == 0) return CUP2D_OK;
  if (c->g.nshapes > 1) {
    RC_OB(enqueue_body_update(c));
    return enqueue_penalize(c, maxpart);
  }
  const int G = grid_cap(c);
  {
    TimerScope ts(c, T_BODY);
    CUP2D_REAL_DISPATCH(c, hipLaunchKernelGGL(k_momenta<T>, dim3(G), dim3(NT), 0, c->stream,
                                              c->g, c->sp, c->ss, c->st, c->ob, tp<T>(c->u),
                                              tp<T>(c->v), surf_of(c)));
  }
  TimerScope ts(c, T_PENALIZE);
  CUP2D_REAL_DISPATCH(c, hipLaunchKernelGGL((k_penalize<true, T>), dim3(G), dim3(NT), 0,
                                            c->stream, c->g, c->sp, c->ss, c->st, c->ob, c->chi,
                                            tp<T>(c->u), tp<T>(c->v), maxpart, c->fsnap,
                                            surf_of(c)));
  return launch_check(c, "body update + penalize");
}

// ------------------------------------------------- PressureIterator_unif
// integrateMomenta (PressureIterator_unif.cpp:77-126) with EXPL_INTEGRATE_MOM:
// F = h^2 chi on the velocity vFluid; penalDX = penalDY = 0.
__global__ __launch_bounds__(NT) void k_iter_momenta(Grid g, const cup2d_shape_state* __restrict__ ss,
                                                     Obstacles ob,
                                                     const double* __restrict__ fu,
                                                     const double* __restrict__ fv) {
  __shared__ double lds[LDS_RED];
  if ((int)blockIdx.x >= ob.cnt[0]) return;
  const int s = blockIdx.x, k = ob.slot_shape[s], b = ob.slot_block[s];
  const int bx = b % g.bpdx, by = b / g.bpdx;
  SlotView o = slot(ob, s);
  const double hsq = g.h * g.h;
  const double Cx = ss[k].centerOfMass[0], Cy = ss[k].centerOfMass[1];
  double a[7] = {0, 0, 0, 0, 0, 0, 0};
  // every load of the block first (one memory round trip), then the sums
  constexpr int NC = BS * BS / NT;
  double cX[NC], cU[NC], cV[NC];
  CELL_LOOP(i) {
    const size_t d = (size_t)(by * BS + i / BS) * g.NX + bx * BS + i % BS;
    cX[i_it] = o.chi[i];
    cU[i_it] = fu[d];
    cV[i_it] = fv[d];
  }
  CELL_LOOP(i) {
    if (cX[i_it] <= 0) continue;
    const int ix = i % BS, iy = i / BS;
    const double F = hsq * cX[i_it];
    const double ud0 = cU[i_it], ud1 = cV[i_it];
    const double p0 = cell_pos(bx, ix, g.hb, g.h) - Cx, p1 = cell_pos(by, iy, g.hb, g.h) - Cy;
    a[0] += F;
    a[1] += F * (p0 * p0 + p1 * p1);
    a[4] += F * ud0;
    a[5] += F * ud1;
    a[6] += F * (p0 * ud1 - p1 * ud0);
  }
  block_sum_multi<NT, 7>(a, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int j = 0; j < 7; ++j) o.part[P_PEN + j] = a[j];
}

int enqueue_iter_body(cup2d_ctx* c) {
  if (c->g.nshapes == 0) return CUP2D_OK;
  const int G = grid_cap(c);
  TimerScope ts(c, T_BODY);
  hipLaunchKernelGGL(k_iter_momenta, dim3(G), dim3(NT), 0, c->stream, c->g, c->ss, c->ob, c->vfu,
                     c->vfv);
  hipLaunchKernelGGL(k_body_solve, dim3(1), dim3(NT), 0, c->stream, c->g, c->sp, c->ss, c->st,
                     c->ob, c->ubidx, c->fsnap, 0, surf_of(c));
  return launch_check(c, "iterator body update");
}

// penalize (PressureIterator_unif.cpp:128-184), per distinct block, shapes in
// order: vel = tmpV + chi (U_s - vFluid) written as vel += DP; per-block sums
// of vel^2 and DP^2 (part[4 ub + 0..3]: MX, MY, DMX, DMY).
__global__ __launch_bounds__(NT) void k_iter_penalize(Grid g, const cup2d_shape_state* __restrict__ ss,
                                                      Obstacles ob, const double* __restrict__ chi,
                                                      const double* __restrict__ tu,
                                                      const double* __restrict__ tv,
                                                      const double* __restrict__ fu,
                                                      const double* __restrict__ fv,
                                                      double* __restrict__ u,
                                                      double* __restrict__ v,
                                                      double* __restrict__ part) {
  __shared__ double lds[LDS_RED];
  if ((int)blockIdx.x >= ob.cnt[1]) return;
  const int b = ob.ub_block[blockIdx.x];
  const int bx = b % g.bpdx, by = b / g.bpdx;
  const int s0 = ob.ub_start[blockIdx.x], cnt = ob.ub_count[blockIdx.x];
  double a[4] = {0, 0, 0, 0};
  for (int q = 0; q < cnt; ++q) {
    const int s = ob.ub_slots[s0 + q], k = ob.slot_shape[s];
    SlotView o = slot(ob, s);
    const double u_s = ss[k].u, v_s = ss[k].v, omega_s = ss[k].omega;
    const double Cx = ss[k].centerOfMass[0], Cy = ss[k].centerOfMass[1];
    CELL_LOOP(i) {
      const int ix = i % BS, iy = i / BS;
      const size_t d = (size_t)(by * BS + iy) * g.NX + bx * BS + ix;
      const double X = o.chi[i];
      if (chi[d] > X) continue;
      if (X <= 0) continue;
      const double p0 = cell_pos(bx, ix, g.hb, g.h) - Cx, p1 = cell_pos(by, iy, g.hb, g.h) - Cy;
      const double US = u_s - omega_s * p1 + o.udx[i];
      const double VS = v_s + omega_s * p0 + o.udy[i];
      const double penalFac = X;
      const double DFX = penalFac * (US - fu[d]);
      const double DFY = penalFac * (VS - fv[d]);
      const double DPX = tu[d] + DFX - u[d];
      const double DPY = tv[d] + DFY - v[d];
      const double u1 = u[d] + DPX, v1 = v[d] + DPY;
      u[d] = u1;
      v[d] = v1;
      a[0] += u1 * u1;
      a[1] += v1 * v1;
      a[2] += DPX * DPX;
      a[3] += DPY * DPY;
    }
  }
  block_sum_multi<NT, 4>(a, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int j = 0; j < 4; ++j) part[4 * blockIdx.x + j] = a[j];
}

// relDF = sqrt((DMX + DMY) / (eps + MX + MY)) (PressureIterator_unif.cpp:183)
__global__ __launch_bounds__(NT) void k_iter_reldf(Obstacles ob, const double* __restrict__ part,
                                                   double* __restrict__ out) {
  __shared__ double lds[LDS_RED];
  double t[4];
  block_sum_table<NT, 4>(part, ob.cnt[1], 4, t, lds);
  if (threadIdx.x == 0) out[0] = sqrt((t[2] + t[3]) / (EPSD + t[0] + t[1]));
}

int enqueue_iter_penalize(cup2d_ctx* c, double* reldf) {
  double* part = c->red + 4096;
  if (c->g.nshapes == 0) {  // no body cells: relDF = 0
    hipLaunchKernelGGL(k_iter_reldf, dim3(1), dim3(NT), 0, c->stream, c->ob, part, reldf);
    return launch_check(c, "iterator penalize");
  }
  TimerScope ts(c, T_PENALIZE);
  hipLaunchKernelGGL(k_iter_penalize, dim3(grid_cap(c)), dim3(NT), 0, c->stream, c->g, c->ss,
                     c->ob, c->chi, c->tu, c->tv, c->vfu, c->vfv, c->u, c->v, part);
  hipLaunchKernelGGL(k_iter_reldf, dim3(1), dim3(NT), 0, c->stream, c->ob, part, reldf);
  return launch_check(c, "iterator penalize");
}

// ----------------------------------------------------------- ComputeForces
// Shape::computeForces (Shape.cpp:252-281): per-block partials added in
// block order; optionally ends the step (time += dt, step++; Simulation.cpp:289-290).
// With a force log, each shape's record of this ComputeForces goes to the
// ring (the reference appends it to its .dat logs at every step); deferred:
// the record's time and step are those of the snapshot, since the next
// step's list builder may already have ended this step.
struct ForceLog {
  double* rec;               // [cap][nshapes][FLOG_REC], nullptr: off
  unsigned long long* cnt;   // records written so far
  int cap;
};
__device__ void forces_fin_wg(const Grid& g, cup2d_shape_state* __restrict__ ss,
                              const Obstacles& ob, StepState* __restrict__ st, int end_step,
                              const ForceSnap* __restrict__ snap, ForceLog flog, int deferred) {
  const unsigned long long nrec = flog.rec ? *flog.cnt : 0ull;
  for (int k = threadIdx.x / 64; k < g.nshapes; k += NT / 64) {  // a shape per wave
    double a[18];
    shape_sums<18>(ob, k, P_FRC, a);
    if (threadIdx.x % 64 != 0) continue;
    cup2d_shape_state& T = ss[k];
    T.perimeter = a[0]; T.circulation = a[1];
    T.forcex = a[2]; T.forcey = a[3]; T.forcex_V = a[4]; T.forcey_V = a[5];
    T.forcex_P = a[6]; T.forcey_P = a[7];
    T.torque = a[8]; T.torque_P = a[9]; T.torque_V = a[10];
    T.thrust = a[11]; T.drag = a[12];
    T.Pout = a[13]; T.defPower = a[14]; T.PoutBnd = a[15]; T.defPowerBnd = a[16];
    T.n_surface_points = a[17];
    T.n_obstacle_blocks = ob.slot_of_shape_start[k + 1] - ob.slot_of_shape_start[k];
    const double su = snap[k].u, sv = snap[k].v;
    T.Pthrust = T.thrust * sqrt(su * su + sv * sv);
    T.Pdrag = T.drag * sqrt(su * su + sv * sv);
    const double denUnb = T.Pthrust - dmin(T.defPower, 0.0);
    const double demBnd = T.Pthrust - T.defPowerBnd;
    T.EffPDef = T.Pthrust / dmax(denUnb, EPSD);
    T.EffPDefBnd = T.Pthrust / dmax(demBnd, EPSD);
    if (flog.rec) {
      double* r = flog.rec + ((nrec % flog.cap) * g.nshapes + k) * FLOG_REC;
      r[0] = deferred ? snap[k].time : st->time;
      r[1] = deferred ? snap[k].step : st->step;
      r[2] = T.forcex; r[3] = T.forcey; r[4] = T.forcex_P; r[5] = T.forcey_P;
      r[6] = T.forcex_V; r[7] = T.forcey_V; r[8] = T.torque; r[9] = T.torque_P;
      r[10] = T.torque_V; r[11] = T.drag; r[12] = T.thrust; r[13] = T.perimeter;
      r[14] = T.circulation; r[15] = T.Pthrust; r[16] = T.Pdrag; r[17] = T.PoutBnd;
      r[18] = T.Pout; r[19] = T.defPowerBnd; r[20] = T.defPower; r[21] = T.EffPDefBnd;
      r[22] = T.EffPDef; r[23] = T.n_surface_points;
    }
  }
  __syncthreads();  // every wave has read the count
  if (flog.rec && threadIdx.x == 0) *flog.cnt = nrec + 1;
  if (end_step && threadIdx.x == 0) {
    st->time += st->dt_done;
    st->step += 1;
  }
}

// ComputeForces.cpp:16-107, per slot; surface points are the cells with D>0.
template <typename RT>
__device__ void forces_wg(const Grid& g, const ForceSnap* __restrict__ snap, const Obstacles& ob,
                          const RT* __restrict__ u, const RT* __restrict__ v,
                          const double* __restrict__ p, double* lds, const SurfVel& surf) {
  const int s = blockIdx.x, k = ob.slot_shape[s], b = ob.slot_block[s];
  const int bx = b % g.bpdx, by = b / g.bpdx;
  SlotView o = slot(ob, s);
  const double NUoH = g.nu / g.h;
  const double Cx = snap[k].C[0], Cy = snap[k].C[1];
  const double su = snap[k].u, sv = snap[k].v;
  const double vel_norm = sqrt(su * su + sv * sv);
  const double vu0 = vel_norm > 0 ? su / vel_norm : 0, vu1 = vel_norm > 0 ? sv / vel_norm : 0;
  double a[18];
#pragma unroll
  for (int j = 0; j < 18; ++j) a[j] = 0;
  CELL_LOOP(i) {
    if (!(o.sd[i] > 0)) continue;
    const int ix = i % BS, iy = i / BS, gx = bx * BS + ix, gy = by * BS + iy;
    const double p0 = cell_pos(bx, ix, g.hb, g.h), p1 = cell_pos(by, iy, g.hb, g.h);
    const double D11 = NUoH * (lab_at(u, g, gx + 1, gy) - lab_at(u, g, gx - 1, gy));
    const double D22 = NUoH * (lab_at(v, g, gx, gy + 1) - lab_at(v, g, gx, gy - 1));
    const double D12 = NUoH / 2 *
                       (lab_at(u, g, gx, gy + 1) - lab_at(u, g, gx, gy - 1) +
                        lab_at(v, g, gx + 1, gy) - lab_at(v, g, gx - 1, gy));
    const double normX = o.sx[i], normY = o.sy[i];
    const size_t d = (size_t)gy * g.NX + gx;
    const double Pc = p[d];
    const double fXV = D11 * normX + D12 * normY, fXP = -Pc * normX;
    const double fYV = D12 * normX + D22 * normY, fYP = -Pc * normY;
    const double fXT = fXV + fXP, fYT = fYV + fYP;
    const double vx = (double)u[d], vy = (double)v[d];
    double vxDef = o.udx[i], vyDef = o.udy[i];
    if (k == surf.shape) {
      double sx, sy;
      surf_vel(surf, p0 - Cx, p1 - Cy, &sx, &sy);
      vxDef += sx;
      vyDef += sy;
    }
    a[0] += sqrt(normX * normX + normY * normY);
    a[1] += normX * vy - normY * vx;
    a[2] += fXT; a[3] += fYT;
    a[4] += fXV; a[5] += fYV;
    a[6] += fXP; a[7] += fYP;
    a[8] += (p0 - Cx) * fYT - (p1 - Cy) * fXT;
    a[9] += (p0 - Cx) * fYP - (p1 - Cy) * fXP;
    a[10] += (p0 - Cx) * fYV - (p1 - Cy) * fXV;
    const double forcePar = fXT * vu0 + fYT * vu1;
    a[11] += .5 * (forcePar + fabs(forcePar));
    a[12] -= .5 * (forcePar - fabs(forcePar));
    const double powOut = fXT * vx + fYT * vy;
    const double powDef = fXT * vxDef + fYT * vyDef;
    a[13] += powOut;
    a[14] += powDef;
    a[15] += dmin(0.0, powOut);
    a[16] += dmin(0.0, powDef);
    a[17] += 1;
  }
  block_sum_multi<NT, 18>(a, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int j = 0; j < 18; ++j) o.part[P_FRC + j] = a[j];
}

// the step end of a whole step (Simulation.cpp:289-290) with the dt its body
// update used
__global__ void k_step_end_done(StepState* st) {
  st->time += st->dt_done;
  st->step += 1;
}

// ComputeForces per slot (ComputeForces.cpp:16-107)
template <typename RT>
__global__ __launch_bounds__(NT) void k_forces(Grid g, const ForceSnap* __restrict__ snap,
                                               Obstacles ob, const RT* __restrict__ u,
                                               const RT* __restrict__ v,
                                               const double* __restrict__ p, SurfVel sv) {
  __shared__ double lds[LDS_RED];
  if ((int)blockIdx.x < ob.cnt[0]) forces_wg(g, snap, ob, u, v, p, lds, sv);
}

// shape totals (Shape::computeForces, Shape.cpp:252-300) and, with end_step,
// the step end
__global__ __launch_bounds__(NT) void k_forces_fin(Grid g, cup2d_shape_state* __restrict__ ss,
                                                   Obstacles ob, StepState* __restrict__ st,
                                                   int end_step, const ForceSnap* __restrict__ snap,
                                                   ForceLog flog, int deferred) {
  forces_fin_wg(g, ss, ob, st, end_step, snap, flog, deferred);
}

static int launch_forces(cup2d_ctx* c, hipStream_t s, const Obstacles& lists, int end_step,
                         const double* u, const double* v, int deferred) {
  TimerScope ts(c, T_FORCES, s);
  CUP2D_REAL_DISPATCH(c, hipLaunchKernelGGL(k_forces<T>, dim3(grid_cap(c)), dim3(NT), 0, s, c->g,
                                            c->fsnap, lists, tp<T>(u), tp<T>(v), c->pres,
                                            surf_of(c)));
  const ForceLog flog{c->flog, c->flog_cnt, c->flog_cap};
  hipLaunchKernelGGL(k_forces_fin, dim3(1), dim3(NT), 0, s, c->g, c->ss, lists, c->st, end_step,
                     c->fsnap, flog, deferred);
  return launch_check(c, "forces");
}

// ComputeForces::operator() + the step end.  on_aux: defer it to the side
// stream of the next step's PutObjectsOnGrid, where it overlaps advDiff: it
// reads the body snapshot of the body update, this step's obstacle lists
// (the next step builds the other list buffer) and this step's velocity and
// pressure, none of which is overwritten before the next step joins the side
// stream.  The step end adds dt_done (set by the body update), not dt.
int enqueue_forces(cup2d_ctx* c, int end_step, int on_aux) {
  RC_OB(flush_forces(c));
  if (c->g.nshapes == 0) {
    if (end_step) return enqueue_step_end(c);
    return CUP2D_OK;
  }
  if (on_aux && end_step) {
    c->forces_pending = true;
    c->end_pending = true;
    return CUP2D_OK;
  }
  return launch_forces(c, c->stream, c->ob, end_step, c->u, c->v, 0);
}

int flush_forces(cup2d_ctx* c) {
  if (c->forces_pending) {
    c->forces_pending = false;
    RC_OB(launch_forces(c, c->stream, c->ob, 0, c->u, c->v, 1));
  }
  if (c->end_pending) {
    c->end_pending = false;
    hipLaunchKernelGGL(k_step_end_done, dim3(1), dim3(1), 0, c->stream, c->st);
    return launch_check(c, "step end");
  }
  return CUP2D_OK;
}

int join_forces(cup2d_ctx* c) {
  if (!c->frc_join) return CUP2D_OK;
  CUP2D_HIP_CHECK(c, hipStreamWaitEvent(c->stream, c->ev_frc, 0));
  c->frc_join = false;
  return CUP2D_OK;
}

// Shape::resetAll (Shape.h:64-88).
// Fish::resetAll (Fish.cpp:187-194) clears the swimmers' internal frame.
__global__ void k_shapes_reset(Grid g, const ShapeParams* __restrict__ sp,
                               cup2d_shape_state* __restrict__ ss, const FishDev* __restrict__ fd) {
  const int k = threadIdx.x;
  if (k < g.nfish)
    for (int q = 0; q < 8; ++q) fd[k].st[q] = 0;
  if (k >= g.nshapes) return;
  cup2d_shape_state T;
  double* z = (double*)&T;
  for (size_t i = 0; i < sizeof(T) / sizeof(double); ++i) z[i] = 0;
  T.center[0] = sp[k].origC[0]; T.center[1] = sp[k].origC[1];
  T.centerOfMass[0] = sp[k].origC[0]; T.centerOfMass[1] = sp[k].origC[1];
  T.orientation = sp[k].origAng;
  T.u = sp[k].forcedu; T.v = sp[k].forcedv; T.omega = sp[k].forcedomega;
  ss[k] = T;
}

int enqueue_shapes_reset(cup2d_ctx* c) {
  if (c->g.nshapes == 0) return CUP2D_OK;
  hipLaunchKernelGGL(k_shapes_reset, dim3(1), dim3(MAX_SHAPES), 0, c->stream, c->g, c->sp, c->ss,
                     c->fishd);
  return launch_check(c, "shapes_reset");
}

// ------------------------------------------------------------ allocation
static int obstacles_alloc_one(cup2d_ctx* c, Obstacles& o, int cap) {
  o.cap = cap;
  const size_t cells = (size_t)cap * BS * BS;
  const size_t ni = (size_t)6 * cap + 2 + MAX_SHAPES + 1;
  int* ibuf = nullptr;
  CUP2D_HIP_CHECK(c, hipMalloc(&ibuf, sizeof(int) * ni));
  CUP2D_HIP_CHECK(c, hipMemset(ibuf, 0, sizeof(int) * ni));
  o.slot_shape = ibuf;
  o.slot_block = ibuf + cap;
  o.ub_block = ibuf + 2 * cap;
  o.ub_start = ibuf + 3 * cap;
  o.ub_count = ibuf + 4 * cap;
  o.ub_slots = ibuf + 5 * cap;
  o.cnt = ibuf + 6 * cap;
  o.slot_of_shape_start = ibuf + 6 * cap + 2;
  double* dbuf = nullptr;
  const size_t nd = cells * 9 + (size_t)cap * SLOT_PART;
  CUP2D_HIP_CHECK(c, hipMalloc(&dbuf, sizeof(double) * nd));
  CUP2D_HIP_CHECK(c, hipMemset(dbuf, 0, sizeof(double) * nd));
  o.chi = dbuf;
  o.dist = dbuf + cells;
  o.rho = dbuf + 2 * cells;
  o.udef = dbuf + 3 * cells;
  o.sdelta = dbuf + 5 * cells;
  o.sdx = dbuf + 6 * cells;
  o.sdy = dbuf + 7 * cells;
  o.part = dbuf + 9 * cells;
  return CUP2D_OK;
}

int obstacles_alloc(cup2d_ctx* c) {
  // capacity: blocks a disk's bounding box can touch, per shape
  int cap = 0;
  for (const ShapeParams& p : c->hsp) {
    const double w = 2 * p.radius + 4 * c->g.h;
    int n1 = (int)ceil(w / c->g.hb) + 2;
    cap += n1 * n1;
  }
  const int nblocks = c->g.bpdx * c->g.bpdy;
  cap = cap > 0 ? cap : 1;
  cap = cap > nblocks * (int)(c->hsp.size() ? c->hsp.size() : 1)
            ? nblocks * (int)(c->hsp.size() ? c->hsp.size() : 1)
            : cap;
  // the worst case (every shape's bound fully covered) can exceed what the
  // list builder holds while the actual obstacle blocks fit: the list
  // builder reports CUP2D_ERR_CAPACITY when they do not
  cap = cap > OBST_LIST_CAP ? OBST_LIST_CAP : cap;
  int rc = obstacles_alloc_one(c, c->ob, cap);
  if (rc) return rc;
  return obstacles_alloc_one(c, c->ob_prev, cap);
}

void obstacles_free(cup2d_ctx* c) {
  for (Obstacles* o : {&c->ob, &c->ob_prev}) {
    if (o->slot_shape) (void)hipFree(o->slot_shape);
    if (o->chi) (void)hipFree(o->chi);
    *o = Obstacles{};
  }
}

// Both list buffers empty (Simulation::reset).
__global__ void k_obst_reset(int* cnt0, int* cnt1) {
  if (threadIdx.x < 2) { cnt0[threadIdx.x] = 0; cnt1[threadIdx.x] = 0; }
}
int enqueue_obst_reset(cup2d_ctx* c) {
  if (!c->ob.cnt) return CUP2D_OK;
  hipLaunchKernelGGL(k_obst_reset, dim3(1), dim3(64), 0, c->stream, c->ob.cnt, c->ob_prev.cnt);
  return launch_check(c, "obstacle reset");
}

}  // namespace cup2d
