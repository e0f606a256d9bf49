/*
 * cup2d_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * FP64 C/OpenMP restatement of the CubismUP_2D per-timestep hot path (the
 * checker for the HIP library, and the timed CPU baseline).  Every function
 * cites the reference file:line it follows.  Deliberate, documented choices
 * where the reference is not deterministic or not available:
 *   - OpenMP `reduction(+:...)` clauses (advDiff.cpp:94, PressureSingle.cpp:30,
 *     136, PoissonSolver.cpp:81, PutObjectsOnGrid.cpp:23, Shape.cpp:121) are
 *     replaced by per-block partial sums accumulated sequentially in the
 *     reference's loop order inside the block, then summed in blockID order.
 *   - std::pow(x,2) is written x*x; GSL LU (Shape.cpp:48-54) is restated as
 *     Doolittle LU with partial pivoting; FFTW is the oracle's own FFT
 *     (cup2d_fft.c).
 *   - Cubism (absent) semantics: blockID = bx + bpdx*by, block spacing
 *     hb = extent/max(bpdx,bpdy), h = hb/32, origin = (bx*hb, by*hb),
 *     pos(ix,iy) = origin + h*(i+0.5); labs copy the block plus a one-cell
 *     halo from neighbour blocks, domain faces filled by BlockLabOpen's
 *     zero-gradient rule (Definitions.h:125-157).
 */
#include "cup2d_oracle.h"
#include "cup2d_fft.h"
#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define BS ORC_BS
#define BC_KILL_FAC 100 /* Definitions.h:38 */
#define LW (BS + 2)

static const double EPS = DBL_EPSILON;

typedef struct {
  double chi[BS][BS], dist[BS][BS], rho[BS][BS], udef[BS][BS][2];
  int nsurf;
  int sx[BS * BS], sy[BS * BS];
  double dchidx[BS * BS], dchidy[BS * BS], delta[BS * BS];
  /* additive per-block quantities (ObstacleBlock.h:44-47) */
  double perimeter, forcex, forcey, forcex_P, forcey_P, forcex_V, forcey_V;
  double torque, torque_P, torque_V, drag, thrust, Pout, PoutBnd, defPower,
      defPowerBnd, circulation;
} oblock;

/* AreaSegment (FishLibrary.h:149-168): a group of 8 midline intervals with
 * its bounding box in the fish frame and the lab-frame boxes of
 * changeToComputationalFrame. */
typedef struct {
  int s0, s1;
  double safe, w[2], c[2], nI[2], nJ[2], lab[2][2], obj[2][2];
} fseg_t;

/* FishData (FishLibrary.h:41-147) with the Fish internal-frame state
 * (Fish.h:19-25). */
typedef struct {
  int Nm, Nseg, envelope, profile;
  double length, Tperiod, phaseShift, h, env_a0, env_a2, waveLength, naca_t;
  double *rS, *rX, *rY, *vX, *vY, *norX, *norY, *vNorX, *vNorY, *width;
  double area_internal, J_internal, CoM_internal[2], vCoM_internal[2];
  double theta_internal, angvel_internal, angvel_internal_prev;
  double Su; /* scale of the fixed-point deformation-velocity sums */
  fseg_t* seg;
} fish_t;

typedef struct {
  orc_shape_desc d;
  double origC[2], origAng, forcedu, forcedv, forcedomega;
  double appliedForceX, appliedForceY, appliedTorque;
  orc_shape_state st;
  oblock** ob; /* [nblocks], NULL where the shape does not touch the block */
  fish_t* fish; /* swimmers only */
} shape_t;

struct orc_sim {
  orc_cfg cfg;
  int bpdx, bpdy, nb, NX, NY;
  double hb, h, extents[2], fadeLenX, fadeLenY;
  double* f[ORC_NFIELDS];
  double* pRHS; /* scanned by the Checker only */
  double time, dt;
  int step;
  double uinfx, uinfy;
  shape_t* shapes;
  int nshapes;
  int* bnd;
  int nbnd;
  /* Poisson */
  int MX, MY, MXh;
  size_t stride;
  double* buffer;
  double* kernel;
  fft2d fft;
  double t_total, t_fft;
  /* PressureIterator_unif state (PressureIterator_unif.h:17-18) */
  int iter_old_nsteps;
  double iter_target;
  int iter_last;
  double iter_relDF;
  int no_collide; /* test hook: skip preventCollidingObstacles */
  /* imposed surface velocity of one disk (the library's
   * cup2d_set_surface_velocity; the reference's Disk keeps udef = 0) */
  int surf_shape, surf_M;
  double surf_coef[4 * 32];
};

static void surf_add(const orc_sim* s, int k, double p0, double p1, double* ux, double* uy);
static fish_t* fish_new(const orc_shape_desc* d, double h);
static void fish_free(fish_t* f);

static const int NCOMP[ORC_NFIELDS] = {2, 1, 1, 1, 2, 2, 1, 2};
int orc_ncomp(int field) { return (field >= 0 && field < ORC_NFIELDS) ? NCOMP[field] : 0; }
int orc_num_threads(void) { return omp_get_max_threads(); }

static inline double sq(double x) { return x * x; }
static inline double dmax(double a, double b) { return (a < b) ? b : a; } /* std::max */
static inline double dmin(double a, double b) { return (b < a) ? b : a; } /* std::min */

/* sin and cos on the body path: Cody-Waite reduction by pi/2 (three-part
 * constant, the first two parts with 33 significant bits so k*P1 and k*P2
 * are exact for |k| < 2^20) and Taylor polynomials on [-pi/4, pi/4] (sine to
 * r^17, cosine to r^18; truncation below 1e-17).  The device library runs the
 * same operation sequence, so both sides produce identical bits; accuracy is
 * about one ulp. */
/* Reference arithmetic (orc_set_reference_arithmetic): on the body path the
 * sines and cosines come from libm (the reference calls std::sin/std::cos,
 * CarlingFish.cpp:21,26, Fish.cpp:118-123), the midline's running sums add
 * in index order (CarlingFish.cpp:82), the midline integrals add in index
 * order (the reference's OpenMP reductions, run on one thread) and
 * PutFishOnBlocks's hat P2M adds every contribution onto chi and udef in
 * loop order (FishLibrary.cpp:619-631).  Off (the default) the oracle runs
 * the device library's orders and sine/cosine, bit for bit.  The mode
 * measures how far the device's summation orders and sine/cosine move the
 * results from the reference's own arithmetic. */
static int g_ref_arith = 0;
void orc_set_reference_arithmetic(int on) { g_ref_arith = on ? 1 : 0; }

void orc_det_sincos(double x, double* sn, double* cs)
{
  if (g_ref_arith) {
    *sn = sin(x);
    *cs = cos(x);
    return;
  }
  const double P1 = 0x1.921fb544p+0, P2 = 0x1.0b4611a6p-34, P3 = 0x1.3198a2e037073p-69;
  const double k = floor(x * 0x1.45f306dc9c883p-1 + 0.5);
  const double r = ((x - k * P1) - k * P2) - k * P3;
  const double z = r * r;
  const double sp = r + (r * z) * (-1.0 / 6 + z * (1.0 / 120 + z * (-1.0 / 5040 + z * (1.0 / 362880 +
                    z * (-1.0 / 39916800 + z * (1.0 / 6227020800.0 + z * (-1.0 / 1307674368000.0 +
                    z * (1.0 / 355687428096000.0))))))));
  const double cp = 1 + z * (-0.5 + z * (1.0 / 24 + z * (-1.0 / 720 + z * (1.0 / 40320 +
                    z * (-1.0 / 3628800 + z * (1.0 / 479001600 + z * (-1.0 / 87178291200.0 +
                    z * (1.0 / 20922789888000.0 + z * (-1.0 / 6402373705728000.0)))))))));
  const int q = (int)(k - 4 * floor(k * 0.25));
  switch (q) {
    case 0: *sn = sp; *cs = cp; break;
    case 1: *sn = cp; *cs = -sp; break;
    case 2: *sn = -sp; *cs = -cp; break;
    default: *sn = -cp; *cs = sp; break;
  }
}
static inline double det_sin(double x) { double a, b; orc_det_sincos(x, &a, &b); return a; }
static inline double det_cos(double x) { double a, b; orc_det_sincos(x, &a, &b); return b; }

/* Sum of n values in a fixed order: 256 lanes, lane l adds v[l], v[l+256],
 * ... in turn, then the lanes are combined pairwise (l += l + 128, 64, ..., 1).
 * The reference sums these midline integrals with OpenMP reductions (no fixed
 * order); the device library uses this order, so both sides agree bit for
 * bit. */
double orc_lane_sum(const double* v, int n)
{
  if (g_ref_arith) {
    double a = 0;
    for (int i = 0; i < n; ++i) a += v[i];
    return a;
  }
  double p[256];
  for (int l = 0; l < 256; ++l) {
    double a = 0;
    for (int i = l; i < n; i += 256) a += v[i];
    p[l] = a;
  }
  for (int st = 128; st >= 1; st >>= 1)
    for (int l = 0; l < st; ++l) p[l] = p[l] + p[l + st];
  return p[0];
}

static inline int BX(const orc_sim* s, int b) { return b % s->bpdx; }
static inline int BY(const orc_sim* s, int b) { return b / s->bpdx; }

/* Cubism BlockInfo::pos: origin + h_gridpoint*(i+0.5) */
static inline void bpos(const orc_sim* s, int b, int ix, int iy, double p[2])
{
  const double ox = (double)BX(s, b) * s->hb, oy = (double)BY(s, b) * s->hb;
  p[0] = ox + s->h * (ix + 0.5);
  p[1] = oy + s->h * (iy + 0.5);
}

static inline double* blk(const orc_sim* s, int field, int b)
{
  return s->f[field] + (size_t)b * BS * BS * NCOMP[field];
}

/* ------------------------------------------------------------------ labs */
/* lab(ix,iy,c) for ix,iy in [-1,BS]; Cubism BlockLab::load + _apply_bc */
#define LAB(L, nc, ix, iy, c) (L)[(((iy) + 1) * LW + ((ix) + 1)) * (nc) + (c)]

static void lab_load(const orc_sim* s, const double* field, int nc, int b, double* L)
{
  const int bx = BX(s, b), by = BY(s, b);
  const size_t bsz = (size_t)BS * BS * nc;
  const double* B = field + (size_t)b * bsz;
  for (int iy = 0; iy < BS; ++iy)
    for (int ix = 0; ix < BS; ++ix)
      for (int c = 0; c < nc; ++c) LAB(L, nc, ix, iy, c) = B[(iy * BS + ix) * nc + c];
  /* neighbour faces and corners */
  for (int dy = -1; dy <= 1; ++dy)
    for (int dx = -1; dx <= 1; ++dx) {
      if (!dx && !dy) continue;
      const int nx = bx + dx, ny = by + dy;
      if (nx < 0 || ny < 0 || nx >= s->bpdx || ny >= s->bpdy) continue;
      const double* N = field + (size_t)(nx + s->bpdx * ny) * bsz;
      const int x0 = dx < 0 ? -1 : (dx > 0 ? BS : 0), x1 = dx < 0 ? -1 : (dx > 0 ? BS : BS - 1);
      const int y0 = dy < 0 ? -1 : (dy > 0 ? BS : 0), y1 = dy < 0 ? -1 : (dy > 0 ? BS : BS - 1);
      for (int iy = y0; iy <= y1; ++iy)
        for (int ix = x0; ix <= x1; ++ix) {
          const int sx = ix - dx * BS, sy = iy - dy * BS;
          for (int c = 0; c < nc; ++c) LAB(L, nc, ix, iy, c) = N[(sy * BS + sx) * nc + c];
        }
    }
  /* BlockLabOpen::_apply_bc, Definitions.h:151-157 (x faces, then y faces) */
  if (bx == 0)
    for (int iy = -1; iy <= BS; ++iy)
      for (int c = 0; c < nc; ++c) LAB(L, nc, -1, iy, c) = LAB(L, nc, 0, iy, c);
  if (bx == s->bpdx - 1)
    for (int iy = -1; iy <= BS; ++iy)
      for (int c = 0; c < nc; ++c) LAB(L, nc, BS, iy, c) = LAB(L, nc, BS - 1, iy, c);
  if (by == 0)
    for (int ix = -1; ix <= BS; ++ix)
      for (int c = 0; c < nc; ++c) LAB(L, nc, ix, -1, c) = LAB(L, nc, ix, 0, c);
  if (by == s->bpdy - 1)
    for (int ix = -1; ix <= BS; ++ix)
      for (int c = 0; c < nc; ++c) LAB(L, nc, ix, BS, c) = LAB(L, nc, ix, BS - 1, c);
}

/* -------------------------------------------------------------- Checker */
/* Checker::run, Helpers.cpp:178-288 (NaN/Inf scan of 10 reals per cell) */
static int checker(orc_sim* s)
{
  if (!s->cfg.checker) return 0;
  const size_t ncell = (size_t)s->nb * BS * BS;
  int bad = 0;
  #pragma omp parallel for schedule(static) reduction(| : bad)
  for (size_t i = 0; i < ncell; ++i) {
    const double a[10] = {s->f[ORC_VEL][2 * i], s->f[ORC_VEL][2 * i + 1], s->f[ORC_UDEF][2 * i],
                          s->f[ORC_UDEF][2 * i + 1], s->f[ORC_TMPV][2 * i], s->f[ORC_TMPV][2 * i + 1],
                          s->f[ORC_CHI][i], s->f[ORC_PRES][i], s->f[ORC_TMP][i], s->pRHS[i]};
    for (int k = 0; k < 10; ++k) bad |= (isnan(a[k]) || isinf(a[k]));
  }
  if (bad) fprintf(stderr, "oracle Checker: NaN/Inf found\n");
  return bad;
}

/* ------------------------------------------------------------ creation */
orc_sim* orc_create(const orc_cfg* cfg)
{
  if (cfg->bpdx <= 0 || cfg->bpdy <= 0) return NULL;
  orc_sim* s = (orc_sim*)calloc(1, sizeof(orc_sim));
  s->cfg = *cfg;
  s->bpdx = cfg->bpdx; s->bpdy = cfg->bpdy; s->nb = s->bpdx * s->bpdy;
  s->NX = BS * s->bpdx; s->NY = BS * s->bpdy;
  /* Cubism Grid(bpdx,bpdy,1,extent): block spacing extent/max(blocks) */
  s->hb = cfg->extent / (double)(s->bpdx > s->bpdy ? s->bpdx : s->bpdy);
  s->h = s->hb / BS;
  /* SimulationData.cpp:34-35 */
  s->extents[0] = (double)s->bpdx * s->h * BS;
  s->extents[1] = (double)s->bpdy * s->h * BS;
  /* Simulation.cpp:97-98 */
  s->fadeLenX = cfg->fadeLen * cfg->extent;
  s->fadeLenY = cfg->fadeLen * cfg->extent;
  s->iter_old_nsteps = 10000; /* PressureIterator_unif.h:17-18 */
  s->iter_target = 1e-4;
  const size_t ncell = (size_t)s->nb * BS * BS;
  for (int f = 0; f < ORC_NFIELDS; ++f) s->f[f] = (double*)calloc(ncell * NCOMP[f], sizeof(double));
  s->pRHS = (double*)calloc(ncell, sizeof(double));
  /* SimulationData.cpp:38-46 */
  s->bnd = (int*)malloc(sizeof(int) * (size_t)s->nb);
  for (int b = 0; b < s->nb; ++b)
    if (BX(s, b) == 0 || BX(s, b) == s->bpdx - 1 || BY(s, b) == 0 || BY(s, b) == s->bpdy - 1)
      s->bnd[s->nbnd++] = b;

  if (cfg->poisson == 2) { /* FFTW_dirichlet.h:25,70: MX x MY reals, stride MX */
    s->MX = s->NX; s->MY = s->NY;
    s->MXh = s->MX / 2 + 1;
    s->stride = (size_t)s->MX;
    s->buffer = (double*)calloc(s->stride * (size_t)s->MY, sizeof(double));
    return s;
  }
  if (cfg->poisson == 0) { /* FFTW_periodic.h:25,96-97 */
    s->MX = s->NX; s->MY = s->NY;
  } else { /* FFTW_freespace.h:25-27 (2N-1), or 2N */
    s->MX = cfg->fs_pad2 ? 2 * s->NX : 2 * s->NX - 1;
    s->MY = cfg->fs_pad2 ? 2 * s->NY : 2 * s->NY - 1;
  }
  s->MXh = s->MX / 2 + 1;
  s->stride = 2 * (size_t)s->MXh;
  s->buffer = (double*)calloc(s->stride * (size_t)s->MY, sizeof(double));
  fft2d_init(&s->fft, s->MX, s->MY, s->stride);
  if (cfg->poisson != 0) {
    /* FFTW_freespace.cpp:100-137 */
    const double h = s->h, fac = 1 / (2 * M_PI);
    const int totNx = s->NX, totNy = s->NY, MX = s->MX, MY = s->MY;
    #pragma omp parallel for schedule(static)
    for (int j = 0; j < MY; ++j)
      for (int i = 0; i < MX; ++i) {
        const double yi = j >= totNy ? MY - j : j;
        const double xi = i >= totNx ? MX - i : i;
        const double r = sqrt(xi * xi + yi * yi);
        s->buffer[i + s->stride * j] = fac * log(h * r);
      }
    s->buffer[0] = fac / 2 * (2 * log(h / sqrt(M_PI)) - 1);
    s->kernel = (double*)malloc(sizeof(double) * (size_t)s->MY * s->MXh);
    fft2d_exec(&s->fft, s->buffer, -1, MY);
    const double norm_factor = 1.0 / ((double)MY * MX);
    for (int j = 0; j < MY; ++j)
      for (int i = 0; i < s->MXh; ++i)
        s->kernel[i + (size_t)s->MXh * j] = s->buffer[2 * (i + (size_t)s->MXh * j)] * norm_factor;
    memset(s->buffer, 0, sizeof(double) * s->stride * (size_t)s->MY);
  }
  return s;
}

void orc_destroy(orc_sim* s)
{
  if (!s) return;
  for (int f = 0; f < ORC_NFIELDS; ++f) free(s->f[f]);
  free(s->pRHS);
  for (int k = 0; k < s->nshapes; ++k) {
    for (int b = 0; b < s->nb; ++b) free(s->shapes[k].ob[b]);
    free(s->shapes[k].ob);
    fish_free(s->shapes[k].fish);
  }
  free(s->shapes); free(s->bnd); free(s->buffer); free(s->kernel);
  fft2d_free(&s->fft);
  free(s);
}

static void shape_reset(orc_sim* s, shape_t* sh)
{ /* Shape::resetAll, Shape.h:64-88 */
  orc_shape_state* t = &sh->st;
  memset(t, 0, sizeof(*t));
  t->center[0] = sh->origC[0]; t->center[1] = sh->origC[1];
  t->centerOfMass[0] = sh->origC[0]; t->centerOfMass[1] = sh->origC[1];
  t->orientation = sh->origAng;
  t->u = sh->forcedu; t->v = sh->forcedv; t->omega = sh->forcedomega;
  sh->appliedForceX = sh->appliedForceY = sh->appliedTorque = 0;
  for (int b = 0; b < s->nb; ++b) { free(sh->ob[b]); sh->ob[b] = NULL; }
  if (sh->fish) { /* Fish::resetAll, Fish.cpp:187-194 */
    fish_t* f = sh->fish;
    f->CoM_internal[0] = f->CoM_internal[1] = 0;
    f->vCoM_internal[0] = f->vCoM_internal[1] = 0;
    f->theta_internal = f->angvel_internal = f->angvel_internal_prev = 0;
  }
}

int orc_add_shape(orc_sim* s, const orc_shape_desc* d)
{
  if (d->kind != 0 && d->kind != 1) return -1;
  fish_t* fish = NULL;
  if (d->kind == 1) {
    if (!(d->length > 0) || !(d->Tperiod > 0)) return -1;
    if (d->envelope == 1 && !(d->waveLength > 0)) return -1;
    fish = fish_new(d, s->h);
    if (!fish) return -1;
  }
  s->shapes = (shape_t*)realloc(s->shapes, sizeof(shape_t) * (size_t)(s->nshapes + 1));
  shape_t* sh = &s->shapes[s->nshapes];
  memset(sh, 0, sizeof(*sh));
  sh->d = *d;
  /* Shape::Shape, Shape.cpp:331-345 */
  sh->origC[0] = d->xpos; sh->origC[1] = d->ypos;
  sh->origAng = d->angle_deg * M_PI / 180;
  sh->forcedu = -d->xvel; sh->forcedv = -d->yvel; sh->forcedomega = -d->angvel;
  sh->ob = (oblock**)calloc((size_t)s->nb, sizeof(oblock*));
  sh->fish = fish;
  shape_reset(s, sh);
  return s->nshapes++;
}

int orc_set_field(orc_sim* s, int field, const double* src)
{
  if (field < 0 || field >= ORC_NFIELDS) return -1;
  memcpy(s->f[field], src, sizeof(double) * (size_t)s->nb * BS * BS * NCOMP[field]);
  return 0;
}
int orc_get_field(const orc_sim* s, int field, double* dst)
{
  if (field < 0 || field >= ORC_NFIELDS) return -1;
  memcpy(dst, s->f[field], sizeof(double) * (size_t)s->nb * BS * BS * NCOMP[field]);
  return 0;
}
void orc_get_scalars(const orc_sim* s, double* time, double* dt, int* step, double uinf[2], double* h)
{
  *time = s->time; *dt = s->dt; *step = s->step; uinf[0] = s->uinfx; uinf[1] = s->uinfy; *h = s->h;
}
void orc_set_scalars(orc_sim* s, double time, int step, double uinfx, double uinfy)
{
  s->time = time; s->step = step; s->uinfx = uinfx; s->uinfy = uinfy;
}
int orc_get_shape(const orc_sim* s, int i, orc_shape_state* out)
{
  if (i < 0 || i >= s->nshapes) return -1;
  *out = s->shapes[i].st;
  int nob = 0, ns = 0;
  for (int b = 0; b < s->nb; ++b)
    if (s->shapes[i].ob[b]) { ++nob; ns += s->shapes[i].ob[b]->nsurf; }
  out->n_obstacle_blocks = nob;
  out->n_surface_points = ns;
  return 0;
}
int orc_fish_info(const orc_sim* s, int k, double* out)
{
  if (k < 0 || k >= s->nshapes || !s->shapes[k].fish) return -1;
  const fish_t* f = s->shapes[k].fish;
  const int Nm = f->Nm;
  if (out) {
    const double* arr[10] = {f->rS, f->rX, f->rY, f->vX, f->vY, f->norX, f->norY,
                             f->vNorX, f->vNorY, f->width};
    for (int a = 0; a < 10; ++a) memcpy(out + (size_t)a * Nm, arr[a], sizeof(double) * (size_t)Nm);
    double* q = out + (size_t)10 * Nm;
    q[0] = f->area_internal; q[1] = f->J_internal;
    q[2] = f->CoM_internal[0]; q[3] = f->CoM_internal[1];
    q[4] = f->vCoM_internal[0]; q[5] = f->vCoM_internal[1];
    q[6] = f->theta_internal; q[7] = f->angvel_internal;
  }
  return Nm;
}

void orc_get_timing(const orc_sim* s, double* total, double* fft) { *total = s->t_total; *fft = s->t_fft; }

/* --------------------------------------------------------------- helpers */
/* IC::operator(), Helpers.cpp:38-82 */
static void ic(orc_sim* s)
{
  const size_t ncell = (size_t)s->nb * BS * BS;
  for (int f = 0; f < ORC_NFIELDS; ++f) memset(s->f[f], 0, sizeof(double) * ncell * NCOMP[f]);
  memset(s->pRHS, 0, sizeof(double) * ncell);
  for (size_t i = 0; i < ncell; ++i) s->f[ORC_INVRHO][i] = 1;
}

/* ApplyObjVel, Helpers.cpp:290-307 */
void orc_apply_obj_vel(orc_sim* s)
{
  const size_t ncell = (size_t)s->nb * BS * BS;
  double* U = s->f[ORC_VEL];
  const double *US = s->f[ORC_UDEF], *X = s->f[ORC_CHI];
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < ncell; ++i) {
    U[2 * i] = U[2 * i] * (1 - X[i]) + US[2 * i] * X[i];
    U[2 * i + 1] = U[2 * i + 1] * (1 - X[i]) + US[2 * i + 1] * X[i];
  }
}

/* findMaxU::run, Helpers.cpp:136-176 */
double orc_find_max_u(orc_sim* s)
{
  checker(s);
  const double UINF = s->uinfx, VINF = s->uinfy;
  const size_t ncell = (size_t)s->nb * BS * BS;
  const double* V = s->f[ORC_VEL];
  double U = 0, Vm = 0, u = 0, v = 0;
  #pragma omp parallel for schedule(static) reduction(max : U, Vm, u, v)
  for (size_t i = 0; i < ncell; ++i) {
    U = dmax(U, fabs(V[2 * i] + UINF));
    Vm = dmax(Vm, fabs(V[2 * i + 1] + VINF));
    u = dmax(u, fabs(V[2 * i]));
    v = dmax(v, fabs(V[2 * i + 1]));
  }
  return dmax(dmax(U, Vm), dmax(u, v));
}

static double max_rel_speed(const orc_sim* s)
{ /* SimulationData::maxRelSpeed + Shape::getMaxVel */
  double m = 0;
  for (int k = 0; k < s->nshapes; ++k) {
    const orc_shape_state* t = &s->shapes[k].st;
    m = dmax(m, sqrt(t->u * t->u + t->v * t->v));
  }
  return m;
}

/* Simulation::calcMaxTimestep, Simulation.cpp:254-277 */
double orc_calc_max_timestep(orc_sim* s)
{
  const double maxU = orc_find_max_u(s);
  const double h = s->h;
  const double dtFourier = h * h / s->cfg.nu, dtCFL = maxU < 2.2e-16 ? 1 : h / maxU;
  const double maxUb = max_rel_speed(s), dtBody = maxUb < 2.2e-16 ? 1 : h / maxUb;
  const double mn = dmin(dmin(dtCFL, dtFourier), dtBody);
  s->dt = s->cfg.CFL * mn;
  if (s->step < 100) {
    const double x = (s->step + 1.0) / 100;
    const double rampCFL = exp(log(1e-3) * (1 - x) + log(s->cfg.CFL) * x);
    s->dt = rampCFL * mn;
  }
  return s->dt;
}

/* ------------------------------------------------------- PutObjectsOnGrid */
static int disk_touching(const orc_sim* s, int b, const double bbox[2][2])
{ /* _is_touching, ShapeLibrary.h:24-36 */
  double MINP[2], MAXP[2];
  bpos(s, b, 0, 0, MINP);
  bpos(s, b, BS - 1, BS - 1, MAXP);
  const double i00 = dmax(MINP[0], bbox[0][0]), i01 = dmin(MAXP[0], bbox[0][1]);
  const double i10 = dmax(MINP[1], bbox[1][0]), i11 = dmin(MAXP[1], bbox[1][1]);
  return i01 - i00 > 0 && i11 - i10 > 0;
}

/* Disk::create + FillBlocks_Cylinder, ShapesSimple.cpp:14-35,
 * ShapeLibrary.h:38-57, ShapeLibrary.cpp:21-34 */
static void disk_create(orc_sim* s, shape_t* sh)
{
  const double h = s->h, radius = sh->d.radius, safety = 2 * h;
  const double pos[2] = {sh->st.center[0], sh->st.center[1]};
  const double bbox[2][2] = {{pos[0] - radius - safety, pos[0] + radius + safety},
                             {pos[1] - radius - safety, pos[1] + radius + safety}};
  for (int b = 0; b < s->nb; ++b) { free(sh->ob[b]); sh->ob[b] = NULL; }
  double* TMP = s->f[ORC_TMP];
  #pragma omp parallel for schedule(dynamic, 1)
  for (int b = 0; b < s->nb; ++b) {
    if (!disk_touching(s, b, bbox)) continue;
    oblock* o = (oblock*)calloc(1, sizeof(oblock));
    sh->ob[b] = o;
    double* B = TMP + (size_t)b * BS * BS;
    for (int iy = 0; iy < BS; iy++)
      for (int ix = 0; ix < BS; ix++) {
        double p[2];
        bpos(s, b, ix, iy, p);
        p[0] -= pos[0]; p[1] -= pos[1];
        o->dist[iy][ix] = radius - sqrt(p[0] * p[0] + p[1] * p[1]);
        B[iy * BS + ix] = dmax(B[iy * BS + ix], o->dist[iy][ix]);
        o->rho[iy][ix] = sh->d.rhoS;
      }
  }
}

/* ---------------------------------------------------------------- swimmers */
/* AmplitudeFish::_width (CarlingFish.cpp:41-48) or a NACA 4-digit
 * symmetric half-thickness with closed trailing edge,
 * 5 t L (0.2969 sqrt(x) - 0.1260 x - 0.3516 x^2 + 0.2843 x^3 - 0.1036 x^4). */
static double fish_width(const fish_t* f, double s, double L)
{
  if (s < 0 || s > L) return 0;
  if (f->profile == 0) {
    const double sb = .04 * f->length, st = .95 * f->length, wt = .01 * f->length,
                 wh = .04 * f->length;
    return s < sb ? sqrt(2 * wh * s - s * s)
                  : (s < st ? wh - (wh - wt) * ((s - sb) / (st - sb)) : (wt * (L - s) / (L - st)));
  }
  const double x = s / L;
  const double w = 5 * f->naca_t * L *
                   (0.2969 * sqrt(x) - 0.1260 * x - 0.3516 * x * x + 0.2843 * x * x * x -
                    0.1036 * x * x * x * x);
  return w < 0 ? 0 : w;
}

/* FishData::FishData (FishLibrary.cpp:14-38, FishLibrary.h:44-60) and
 * AmplitudeFish's width table (CarlingFish.cpp:37-38). */
static fish_t* fish_new(const orc_shape_desc* d, double h)
{
  const double length = d->length;
  const double fracRefined = 0.1, fracMid = 1 - 2 * fracRefined;
  const double dSmid_tgt = h / sqrt(2);
  const double dSrefine_tgt = 0.125 * h;
  const int Nmid = (int)ceil(length * fracMid / dSmid_tgt / 8) * 8;
  const double dSmid = length * fracMid / Nmid;
  const int Nend = (int)ceil(fracRefined * length * 2 / (dSmid + dSrefine_tgt) / 4) * 4;
  const double dSref = fracRefined * length * 2 / Nend - dSmid;
  const int Nm = Nmid + 2 * Nend + 1;
  if (!(dSref > 0) || Nend < 2 || Nm < 17) return NULL;
  fish_t* f = (fish_t*)calloc(1, sizeof(fish_t));
  f->Nm = Nm;
  f->Nseg = (Nm - 1) / 8;
  f->length = length; f->Tperiod = d->Tperiod; f->phaseShift = d->phaseShift; f->h = h;
  f->envelope = d->envelope; f->profile = d->profile;
  f->env_a0 = d->env_a0; f->env_a2 = d->env_a2; f->waveLength = d->waveLength;
  f->naca_t = d->naca_t;
  double** arr[10] = {&f->rS, &f->rX, &f->rY, &f->vX, &f->vY, &f->norX, &f->norY,
                      &f->vNorX, &f->vNorY, &f->width};
  for (int a = 0; a < 10; ++a) *arr[a] = (double*)calloc((size_t)Nm, sizeof(double));
  f->seg = (fseg_t*)calloc((size_t)f->Nseg, sizeof(fseg_t));
  double* rS = f->rS;
  rS[0] = 0;
  int k = 0;
  for (int i = 0; i < Nend; ++i, k++)
    rS[k + 1] = rS[k] + dSref + (dSmid - dSref) * i / ((double)Nend - 1.);
  for (int i = 0; i < Nmid; ++i, k++) rS[k + 1] = rS[k] + dSmid;
  for (int i = 0; i < Nend; ++i, k++)
    rS[k + 1] = rS[k] + dSref + (dSmid - dSref) * (Nend - i - 1) / ((double)Nend - 1.);
  rS[k] = dmin(rS[k], length);
  for (int i = 0; i < Nm; ++i) f->width[i] = fish_width(f, rS[i], length);
  return f;
}

static void fish_free(fish_t* f)
{
  if (!f) return;
  free(f->rS); free(f->rX); free(f->rY); free(f->vX); free(f->vY);
  free(f->norX); free(f->norY); free(f->vNorX); free(f->vNorY); free(f->width);
  free(f->seg); free(f);
}

/* AmplitudeFish::midlineLatPos / midlineLatVel (CarlingFish.cpp:19-27), or
 * the envelope L (a0 + a2 (s/L)^2) with wavelength waveLength*L. */
static inline double lat_pos(const fish_t* f, double s, double t)
{
  if (f->envelope == 0) {
    const double arg = 2 * M_PI * (s / f->length - t / f->Tperiod + f->phaseShift);
    return 4. / 33. * (s + 0.03125 * f->length) * det_sin(arg);
  }
  const double x = s / f->length;
  const double arg = 2 * M_PI * (x / f->waveLength - t / f->Tperiod + f->phaseShift);
  return f->length * (f->env_a0 + f->env_a2 * x * x) * det_sin(arg);
}
static inline double lat_vel(const fish_t* f, double s, double t)
{
  if (f->envelope == 0) {
    const double arg = 2 * M_PI * (s / f->length - t / f->Tperiod + f->phaseShift);
    return -4. / 33. * (s + 0.03125 * f->length) * (2 * M_PI / f->Tperiod) * det_cos(arg);
  }
  const double x = s / f->length;
  const double arg = 2 * M_PI * (x / f->waveLength - t / f->Tperiod + f->phaseShift);
  return -f->length * (f->env_a0 + f->env_a2 * x * x) * (2 * M_PI / f->Tperiod) * det_cos(arg);
}

/* AmplitudeFish::computeMidline (CarlingFish.cpp:51-88) */
static void fish_midline(fish_t* f, double t)
{
  const int Nm = f->Nm;
  const double T = f->Tperiod;
  const double rampFac = t < T ? det_sin(0.5 * M_PI * t / T) : 1.0;
  const double rampFacVel = t < T ? 0.5 * M_PI / T * det_cos(0.5 * M_PI * t / T) : 0.0;
  double *rS = f->rS, *rX = f->rX, *rY = f->rY, *vX = f->vX, *vY = f->vY;
  rX[0] = 0.0;
  rY[0] = rampFac * lat_pos(f, rS[0], t);
  vX[0] = 0.0;
  vY[0] = rampFac * lat_vel(f, rS[0], t) + rampFacVel * lat_pos(f, rS[0], t);
  for (int i = 1; i < Nm; ++i) {
    rY[i] = rampFac * lat_pos(f, rS[i], t);
    vY[i] = rampFac * lat_vel(f, rS[i], t) + rampFacVel * lat_pos(f, rS[i], t);
  }
  for (int i = 1; i < Nm; ++i) {
    const double dy = rY[i] - rY[i - 1], ds = rS[i] - rS[i - 1];
    const double dx = sqrt(ds * ds - dy * dy);
    const double dVy = vY[i] - vY[i - 1];
    const double dVx = -dy / dx * dVy;
    rX[i] = dx;
    vX[i] = dVx;
    f->norX[i - 1] = -dy / ds;
    f->norY[i - 1] = dx / ds;
    f->vNorX[i - 1] = -dVy / ds;
    f->vNorY[i - 1] = dVx / ds;
  }
  /* running sums rX[i] += rX[i-1]: the reference adds in index order; here
   * (as on the device) in a fixed two-level order: 64 chunks of
   * c = ceil((Nm-1)/64) increments, each summed in index order, chunk offsets
   * summed in chunk order, element = chunk-local sum + offset */
  if (g_ref_arith) {
    for (int i = 1; i < Nm; ++i) { rX[i] += rX[i - 1]; vX[i] += vX[i - 1]; }
  } else {
    const int c = (Nm - 1 + 63) / 64;
    double* arr[2] = {rX, vX};
    for (int q = 0; q < 2; ++q) {
      double* a = arr[q];
      double off = 0;
      for (int l = 0; l < 64; ++l) {
        const int i0 = 1 + l * c, i1 = i0 + c < Nm ? i0 + c : Nm;
        double p = 0;
        for (int i = i0; i < i1; ++i) { p = a[i] + p; a[i] = p; }
        const double total = i0 < i1 ? p : 0.0;
        if (l > 0)
          for (int i = i0; i < i1; ++i) a[i] = a[i] + off;
        off = total + off;
      }
    }
  }
  f->norX[Nm - 1] = f->norX[Nm - 2];
  f->norY[Nm - 1] = f->norY[Nm - 2];
  f->vNorX[Nm - 1] = f->vNorX[Nm - 2];
  f->vNorY[Nm - 1] = f->vNorY[Nm - 2];
}

/* FishData::_d_ds (FishLibrary.h:102-107) */
static inline double fish_d_ds(const fish_t* f, int idx, const double* vals)
{
  const double* rS = f->rS;
  const int maxidx = f->Nm;
  if (idx == 0) return (vals[idx + 1] - vals[idx]) / (rS[idx + 1] - rS[idx]);
  else if (idx == maxidx - 1) return (vals[idx] - vals[idx - 1]) / (rS[idx] - rS[idx - 1]);
  else return ((vals[idx + 1] - vals[idx]) / (rS[idx + 1] - rS[idx]) +
               (vals[idx] - vals[idx - 1]) / (rS[idx] - rS[idx - 1])) / 2;
}

/* trapezoid weight and _integrationFac1..3 (FishLibrary.h:108-118); w^3 is
 * written w*w*w */
static inline void fish_facs(const fish_t* f, int i, double* ds, double* fac1, double* fac2,
                             double* fac3)
{
  const int Nm = f->Nm;
  const double* rS = f->rS;
  *ds = (i == 0) ? rS[1] - rS[0] : ((i == Nm - 1) ? rS[Nm - 1] - rS[Nm - 2] : rS[i + 1] - rS[i - 1]);
  const double w = f->width[i], w3 = w * w * w;
  const double dnorXi = fish_d_ds(f, i, f->norX), dnorYi = fish_d_ds(f, i, f->norY);
  *fac1 = 2 * w;
  *fac2 = 2 * w3 * (dnorXi * f->norY[i] - dnorYi * f->norX[i]) / 3;
  *fac3 = 2 * w3 / 3;
}

/* FishData::_computeMidlineNormals (FishLibrary.cpp:66-84) */
static void fish_normals(fish_t* f)
{
  const int Nm = f->Nm;
  for (int i = 0; i < Nm - 1; i++) {
    const double ds = f->rS[i + 1] - f->rS[i];
    const double tX = f->rX[i + 1] - f->rX[i], tY = f->rY[i + 1] - f->rY[i];
    const double tVX = f->vX[i + 1] - f->vX[i], tVY = f->vY[i + 1] - f->vY[i];
    f->norX[i] = -tY / ds;
    f->norY[i] = tX / ds;
    f->vNorX[i] = -tVY / ds;
    f->vNorY[i] = tVX / ds;
  }
  f->norX[Nm - 1] = f->norX[Nm - 2];
  f->norY[Nm - 1] = f->norY[Nm - 2];
  f->vNorX[Nm - 1] = f->vNorX[Nm - 2];
  f->vNorY[Nm - 1] = f->vNorY[Nm - 2];
}

/* Fish::create steps 1-3 (Fish.cpp:78-95): midline, integrateLinearMomentum
 * (FishLibrary.cpp:86-117), changeToCoMFrameLinear (:152-159),
 * integrateAngularMomentum (:119-150), changeToCoMFrameAngular (:161-176). */
static void fish_frame(fish_t* f, double t)
{
  const int Nm = f->Nm;
  fish_midline(f, t);
  double* tmp = (double*)malloc(sizeof(double) * (size_t)Nm * 5);
  for (int i = 0; i < Nm; ++i) {
    double ds, fac1, fac2, fac3;
    fish_facs(f, i, &ds, &fac1, &fac2, &fac3);
    tmp[i] = fac1 * ds / 2;
    tmp[Nm + i] = (f->rX[i] * fac1 + f->norX[i] * fac2) * ds / 2;
    tmp[2 * Nm + i] = (f->rY[i] * fac1 + f->norY[i] * fac2) * ds / 2;
    tmp[3 * Nm + i] = (f->vX[i] * fac1 + f->vNorX[i] * fac2) * ds / 2;
    tmp[4 * Nm + i] = (f->vY[i] * fac1 + f->vNorY[i] * fac2) * ds / 2;
  }
  const double area = orc_lane_sum(tmp, Nm);
  const double cmx = orc_lane_sum(tmp + Nm, Nm), cmy = orc_lane_sum(tmp + 2 * Nm, Nm);
  const double lmx = orc_lane_sum(tmp + 3 * Nm, Nm), lmy = orc_lane_sum(tmp + 4 * Nm, Nm);
  f->area_internal = area;
  f->CoM_internal[0] = cmx / area; f->CoM_internal[1] = cmy / area;
  f->vCoM_internal[0] = lmx / area; f->vCoM_internal[1] = lmy / area;
  for (int i = 0; i < Nm; ++i) {
    f->rX[i] -= f->CoM_internal[0]; f->rY[i] -= f->CoM_internal[1];
    f->vX[i] -= f->vCoM_internal[0]; f->vY[i] -= f->vCoM_internal[1];
  }
  f->angvel_internal_prev = f->angvel_internal;
  for (int i = 0; i < Nm; ++i) {
    double ds, fac1, fac2, fac3;
    fish_facs(f, i, &ds, &fac1, &fac2, &fac3);
    const double rX = f->rX[i], rY = f->rY[i], vX = f->vX[i], vY = f->vY[i];
    const double norX = f->norX[i], norY = f->norY[i], vNorX = f->vNorX[i], vNorY = f->vNorY[i];
    const double tmp_M = (rX * vY - rY * vX) * fac1 +
                         (rX * vNorY - rY * vNorX + vY * norX - vX * norY) * fac2 +
                         (norX * vNorY - norY * vNorX) * fac3;
    const double tmp_J = (rX * rX + rY * rY) * fac1 + 2 * (rX * norX + rY * norY) * fac2 + fac3;
    tmp[i] = tmp_M * ds / 2;
    tmp[Nm + i] = tmp_J * ds / 2;
  }
  const double am = orc_lane_sum(tmp, Nm), J = orc_lane_sum(tmp + Nm, Nm);
  free(tmp);
  f->J_internal = J;
  f->angvel_internal = am / J;
  double sn, cs;
  orc_det_sincos(f->theta_internal, &sn, &cs);
  const double R[2][2] = {{cs, -sn}, {sn, cs}};
  const double vAin = f->angvel_internal;
  for (int i = 0; i < Nm; ++i) {
    f->vX[i] += vAin * f->rY[i];
    f->vY[i] -= vAin * f->rX[i];
    const double px = f->rX[i], py = f->rY[i];
    f->rX[i] = R[0][0] * px + R[0][1] * py;
    f->rY[i] = R[1][0] * px + R[1][1] * py;
    const double qx = f->vX[i], qy = f->vY[i];
    f->vX[i] = R[0][0] * qx + R[0][1] * qy;
    f->vY[i] = R[1][0] * qx + R[1][1] * qy;
  }
  fish_normals(f);
  /* bound on any component of a deformation velocity of PutFishOnBlocks,
   * |R (v + o vNor)| <= |vX| + |vY| + w (|vNorX| + |vNorY|): the internal
   * P2M sums below use the power-of-two scale Su with |w ud Su| < 2^52 */
  double vmax = 0;
  for (int i = 0; i < Nm; ++i)
    vmax = dmax(vmax, fabs(f->vX[i]) + fabs(f->vY[i]) +
                          f->width[i] * (fabs(f->vNorX[i]) + fabs(f->vNorY[i])));
  int e = 0;
  (void)frexp(vmax, &e);
  f->Su = ldexp(1.0, 52 - e);
}

/* Fish::create step 4-5 (Fish.cpp:118-150): segment boxes, AreaSegment
 * (FishLibrary.h:160-164) and AreaSegment::changeToComputationalFrame
 * (FishLibrary.cpp:295-342) */
static void fish_segments(fish_t* f, const double pos[2], double angle)
{
  const int Nm = f->Nm, Nseg = f->Nseg;
  double sn, cs;
  orc_det_sincos(angle, &sn, &cs);
  const double R[2][2] = {{cs, -sn}, {sn, cs}};
  for (int i = 0; i < Nseg; ++i) {
    fseg_t* S = &f->seg[i];
    const int next_idx = (i + 1) * (Nm - 1) / Nseg, idx = i * (Nm - 1) / Nseg;
    double bbox[2][2] = {{1e9, -1e9}, {1e9, -1e9}};
    for (int ss = idx; ss <= next_idx; ++ss) {
      const double w = f->width[ss];
      const double xBnd[2] = {f->rX[ss] - f->norX[ss] * w, f->rX[ss] + f->norX[ss] * w};
      const double yBnd[2] = {f->rY[ss] - f->norY[ss] * w, f->rY[ss] + f->norY[ss] * w};
      const double maxX = dmax(xBnd[0], xBnd[1]), minX = dmin(xBnd[0], xBnd[1]);
      const double maxY = dmax(yBnd[0], yBnd[1]), minY = dmin(yBnd[0], yBnd[1]);
      bbox[0][0] = dmin(bbox[0][0], minX);
      bbox[0][1] = dmax(bbox[0][1], maxX);
      bbox[1][0] = dmin(bbox[1][0], minY);
      bbox[1][1] = dmax(bbox[1][1], maxY);
    }
    S->s0 = idx; S->s1 = next_idx;
    S->safe = 2 * f->h;
    S->w[0] = (bbox[0][1] - bbox[0][0]) / 2 + S->safe;
    S->w[1] = (bbox[1][1] - bbox[1][0]) / 2 + S->safe;
    S->c[0] = (bbox[0][1] + bbox[0][0]) / 2;
    S->c[1] = (bbox[1][1] + bbox[1][0]) / 2;
    const double p[2] = {S->c[0], S->c[1]};
    const double nx[2] = {1, 0}, ny[2] = {0, 1};
    for (int d = 0; d < 2; ++d) {
      S->c[d] = R[d][0] * p[0] + R[d][1] * p[1];
      S->nI[d] = R[d][0] * nx[0] + R[d][1] * nx[1];
      S->nJ[d] = R[d][0] * ny[0] + R[d][1] * ny[1];
    }
    S->c[0] += pos[0];
    S->c[1] += pos[1];
    const double magI = sqrt(S->nI[0] * S->nI[0] + S->nI[1] * S->nI[1]);
    const double magJ = sqrt(S->nJ[0] * S->nJ[0] + S->nJ[1] * S->nJ[1]);
    const double invMagI = 1 / magI, invMagJ = 1 / magJ;
    for (int d = 0; d < 2; ++d) {
      S->nI[d] = fabs(S->nI[d]) * invMagI;
      S->nJ[d] = fabs(S->nJ[d]) * invMagJ;
    }
    const double wX[2] = {S->w[0] * S->nI[0], S->w[0] * S->nI[1]};
    const double wY[2] = {S->w[1] * S->nJ[0], S->w[1] * S->nJ[1]};
    for (int d = 0; d < 2; ++d) {
      S->lab[d][0] = S->c[d] - wX[d] - wY[d];
      S->lab[d][1] = S->c[d] + wX[d] + wY[d];
      S->obj[d][0] = S->c[d] - S->w[d];
      S->obj[d][1] = S->c[d] + S->w[d];
    }
  }
}

/* AreaSegment::isIntersectingWithAABB (FishLibrary.cpp:344-401) */
static int seg_intersects(const fseg_t* S, const double start[2], const double end[2])
{
  const double AABB_w[2] = {(end[0] - start[0]) / 2 + S->safe, (end[1] - start[1]) / 2 + S->safe};
  const double AABB_c[2] = {(end[0] + start[0]) / 2, (end[1] + start[1]) / 2};
  const double AABB_box[2][2] = {{AABB_c[0] - AABB_w[0], AABB_c[0] + AABB_w[0]},
                                 {AABB_c[1] - AABB_w[1], AABB_c[1] + AABB_w[1]}};
  const double iL[2][2] = {{dmax(S->lab[0][0], AABB_box[0][0]), dmin(S->lab[0][1], AABB_box[0][1])},
                           {dmax(S->lab[1][0], AABB_box[1][0]), dmin(S->lab[1][1], AABB_box[1][1])}};
  if (iL[0][1] - iL[0][0] < 0 || iL[1][1] - iL[1][0] < 0) return 0;
  const double widthXbox[2] = {AABB_w[0] * S->nI[0], AABB_w[0] * S->nJ[0]};
  const double widthYbox[2] = {AABB_w[1] * S->nI[1], AABB_w[1] * S->nJ[1]};
  const double boxBox[2][2] = {
      {AABB_c[0] - widthXbox[0] - widthYbox[0], AABB_c[0] + widthXbox[0] + widthYbox[0]},
      {AABB_c[1] - widthXbox[1] - widthYbox[1], AABB_c[1] + widthXbox[1] + widthYbox[1]}};
  const double iF[2][2] = {{dmax(boxBox[0][0], S->obj[0][0]), dmin(boxBox[0][1], S->obj[0][1])},
                           {dmax(boxBox[1][0], S->obj[1][0]), dmin(boxBox[1][1], S->obj[1][1])}};
  if (iF[0][1] - iF[0][0] < 0 || iF[1][1] - iF[1][0] < 0) return 0;
  return 1;
}

/* PutFishOnBlocks (FishLibrary.h:170-214): the frame of the fish at
 * (position, angle) */
typedef struct {
  double pos[2], R[2][2];
} fframe_t;

static inline void to_comp(const fframe_t* F, double x[2])
{ /* changeToComputationalFrame */
  const double p[2] = {x[0], x[1]};
  x[0] = F->R[0][0] * p[0] + F->R[0][1] * p[1];
  x[1] = F->R[1][0] * p[0] + F->R[1][1] * p[1];
  x[0] += F->pos[0];
  x[1] += F->pos[1];
}
static inline void vel_to_comp(const fframe_t* F, double x[2])
{ /* changeVelocityToComputationalFrame */
  const double p[2] = {x[0], x[1]};
  x[0] = F->R[0][0] * p[0] + F->R[0][1] * p[1];
  x[1] = F->R[1][0] * p[0] + F->R[1][1] * p[1];
}
static inline void from_comp(const fframe_t* F, double x[2])
{ /* changeFromComputationalFrame */
  const double p[2] = {x[0] - F->pos[0], x[1] - F->pos[1]};
  x[0] = F->R[0][0] * p[0] + F->R[1][0] * p[1];
  x[1] = F->R[0][1] * p[0] + F->R[1][1] * p[1];
}
static inline double dist_sq(const double a[2], const double b[2])
{
  return (a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]);
}

/* PutFishOnBlocks::constructSurface, constructInternl and signedDistanceSqrt
 * (FishLibrary.cpp:419-638) for block b and the segments sl[0..ns) that
 * intersect it, in segment order; B is the block of the grid SDF (sim.tmp). */
static void put_fish_on_block(const orc_sim* s, const fish_t* f, const fframe_t* F, int b,
                              const int* sl, int ns, oblock* o, double* B)
{
  double org[2];
  bpos(s, b, 0, 0, org);
  const double h = s->h, invh = 1.0 / h;
  const int Nm = f->Nm;
  const double *rX = f->rX, *rY = f->rY, *norX = f->norX, *norY = f->norY;
  const double *vX = f->vX, *vY = f->vY, *vNorX = f->vNorX, *vNorY = f->vNorY;
  const double* width = f->width;
  for (int iy = 0; iy < BS; ++iy)
    for (int ix = 0; ix < BS; ++ix) { o->dist[iy][ix] = -1; o->chi[iy][ix] = 0; }
  /* constructSurface */
  for (int i = 0; i < ns; ++i) {
    const fseg_t* S = &f->seg[sl[i]];
    const int firstSegm = S->s0 > 1 ? S->s0 : 1;
    const int lastSegm = S->s1 < Nm - 2 ? S->s1 : Nm - 2;
    for (int ss = firstSegm; ss <= lastSegm; ++ss) {
      for (int signp = -1; signp <= 1; signp += 2) {
        double myP[2] = {rX[ss + 0] + width[ss + 0] * signp * norX[ss + 0],
                         rY[ss + 0] + width[ss + 0] * signp * norY[ss + 0]};
        to_comp(F, myP);
        const int iap[2] = {(int)floor((myP[0] - org[0]) * invh), (int)floor((myP[1] - org[1]) * invh)};
        if (iap[0] + 3 <= 0 || iap[0] - 1 >= BS) continue;
        if (iap[1] + 3 <= 0 || iap[1] - 1 >= BS) continue;
        double pP[2] = {rX[ss + 1] + width[ss + 1] * signp * norX[ss + 1],
                        rY[ss + 1] + width[ss + 1] * signp * norY[ss + 1]};
        to_comp(F, pP);
        double pM[2] = {rX[ss - 1] + width[ss - 1] * signp * norX[ss - 1],
                        rY[ss - 1] + width[ss - 1] * signp * norY[ss - 1]};
        to_comp(F, pM);
        double udef[2] = {vX[ss + 0] + width[ss + 0] * signp * vNorX[ss + 0],
                          vY[ss + 0] + width[ss + 0] * signp * vNorY[ss + 0]};
        vel_to_comp(F, udef);
        const int sy0 = iap[1] - 1 > 0 ? iap[1] - 1 : 0, sy1 = iap[1] + 3 < BS ? iap[1] + 3 : BS;
        const int sx0 = iap[0] - 1 > 0 ? iap[0] - 1 : 0, sx1 = iap[0] + 3 < BS ? iap[0] + 3 : BS;
        for (int sy = sy0; sy < sy1; ++sy)
          for (int sx = sx0; sx < sx1; ++sx) {
            double p[2];
            bpos(s, b, sx, sy, p);
            const double dist0 = dist_sq(p, myP);
            const double distP = dist_sq(p, pP);
            const double distM = dist_sq(p, pM);
            if (fabs(o->dist[sy][sx]) < dmin(dmin(dist0, distP), distM)) continue;
            from_comp(F, p);
            int close_s = ss, secnd_s = ss + (distP < distM ? 1 : -1);
            double dist1 = dist0, dist2 = distP < distM ? distP : distM;
            if (distP < dist0 || distM < dist0) {
              dist1 = dist2; dist2 = dist0;
              close_s = secnd_s; secnd_s = ss;
            }
            const double dSsq = (rX[close_s] - rX[secnd_s]) * (rX[close_s] - rX[secnd_s]) +
                                (rY[close_s] - rY[secnd_s]) * (rY[close_s] - rY[secnd_s]);
            const double cnt2ML = width[close_s] * width[close_s];
            const double nxt2ML = width[secnd_s] * width[secnd_s];
            const double safeW = dmax(width[close_s], width[secnd_s]) + 2 * h;
            const double xMidl[2] = {rX[close_s], rY[close_s]};
            const double grd2ML = dist_sq(p, xMidl);
            const double diffH = fabs(width[close_s] - width[secnd_s]);
            double sign2d = 0;
            if (dSsq > diffH * diffH || grd2ML > safeW * safeW) {
              sign2d = grd2ML > cnt2ML ? -1 : 1;
            } else {
              const double corr = 2 * sqrt(cnt2ML * nxt2ML);
              const double Rsq = (cnt2ML + nxt2ML - corr + dSsq) * (cnt2ML + nxt2ML + corr + dSsq) / 4 / dSsq;
              const double maxAx = dmax(cnt2ML, nxt2ML);
              const int idAx1 = cnt2ML > nxt2ML ? close_s : secnd_s;
              const int idAx2 = idAx1 == close_s ? secnd_s : close_s;
              const double d = sqrt((Rsq - maxAx) / dSsq);
              const double xCentr[2] = {rX[idAx1] + (rX[idAx1] - rX[idAx2]) * d,
                                        rY[idAx1] + (rY[idAx1] - rY[idAx2]) * d};
              const double grd2Core = dist_sq(p, xCentr);
              sign2d = grd2Core > Rsq ? -1 : 1;
            }
            if (fabs(o->dist[sy][sx]) > dist1) {
              const double W = 1 - sqrt(dist1) * (invh / 3);
              o->udef[sy][sx][0] = W * udef[0];
              o->udef[sy][sx][1] = W * udef[1];
              o->dist[sy][sx] = sign2d * dist1;
              o->chi[sy][sx] = W;
              o->rho[sy][sx] = 1;
            }
          }
      }
    }
  }
  /* constructInternl.  The reference adds each contribution to udef and chi
   * in loop order; here they are summed in 64-bit fixed point (weights
   * x 2^52, weighted velocities x Su, each rounded to the nearest integer)
   * and added to the surface values at the end, so the result does not depend
   * on the summation order (the device sums in parallel).  This differs from
   * the reference's sequential double sums by rounding only. */
  static long long qw[BS][BS], qu[BS][BS], qv[BS][BS];
  static char touched[BS][BS];
  #pragma omp threadprivate(qw, qu, qv, touched)
  memset(qw, 0, sizeof(qw)); memset(qu, 0, sizeof(qu)); memset(qv, 0, sizeof(qv));
  memset(touched, 0, sizeof(touched));
  for (int i = 0; i < ns; ++i) {
    const fseg_t* S = &f->seg[sl[i]];
    const int firstSegm = S->s0 > 1 ? S->s0 : 1;
    const int lastSegm = S->s1 < Nm - 2 ? S->s1 : Nm - 2;
    for (int ss = firstSegm; ss <= lastSegm; ++ss) {
      const double myWidth = width[ss];
      const int Nw = (int)floor(myWidth / h);
      for (int iw = -Nw + 1; iw < Nw; ++iw) {
        const double offsetW = iw * h;
        double xp[2] = {rX[ss] + offsetW * norX[ss], rY[ss] + offsetW * norY[ss]};
        to_comp(F, xp);
        xp[0] = (xp[0] - org[0]) * invh;
        xp[1] = (xp[1] - org[1]) * invh;
        const double ap[2] = {floor(xp[0]), floor(xp[1])};
        const int iap[2] = {(int)ap[0], (int)ap[1]};
        if (iap[0] + 2 <= 0 || iap[0] >= BS) continue;
        if (iap[1] + 2 <= 0 || iap[1] >= BS) continue;
        double udef[2] = {vX[ss] + offsetW * vNorX[ss], vY[ss] + offsetW * vNorY[ss]};
        vel_to_comp(F, udef);
        double wghts[2][2];
        for (int c = 0; c < 2; ++c) {
          const double t[2] = {fabs(xp[c] - ap[c]), fabs(xp[c] - (ap[c] + 1))};
          wghts[c][0] = 1 - t[0];
          wghts[c][1] = 1 - t[1];
        }
        const int y0 = iap[1] > 0 ? iap[1] : 0, y1 = iap[1] + 2 < BS ? iap[1] + 2 : BS;
        const int x0 = iap[0] > 0 ? iap[0] : 0, x1 = iap[0] + 2 < BS ? iap[0] + 2 : BS;
        for (int idy = y0; idy < y1; ++idy)
          for (int idx = x0; idx < x1; ++idx) {
            const int sx = idx - iap[0], sy = idy - iap[1];
            const double wxwy = wghts[1][sy] * wghts[0][sx];
            if (g_ref_arith) {
              o->udef[idy][idx][0] += wxwy * udef[0];
              o->udef[idy][idx][1] += wxwy * udef[1];
              o->chi[idy][idx] += wxwy;
            } else {
              qu[idy][idx] += llrint(wxwy * udef[0] * f->Su);
              qv[idy][idx] += llrint(wxwy * udef[1] * f->Su);
              qw[idy][idx] += llrint(wxwy * 0x1p52);
              touched[idy][idx] = 1;
            }
            o->rho[idy][idx] = 1;
            if (fabs(o->dist[idy][idx] + 1) < EPS) o->dist[idy][idx] = 1;
          }
      }
    }
  }
  for (int iy = 0; iy < BS; iy++)
    for (int ix = 0; ix < BS; ix++) {
      if (!touched[iy][ix]) continue;
      o->chi[iy][ix] += (double)qw[iy][ix] * 0x1p-52;
      o->udef[iy][ix][0] += (double)qu[iy][ix] / f->Su;
      o->udef[iy][ix][1] += (double)qv[iy][ix] / f->Su;
    }
  /* signedDistanceSqrt */
  for (int iy = 0; iy < BS; iy++)
    for (int ix = 0; ix < BS; ix++) {
      const double normfac = o->chi[iy][ix] > EPS ? o->chi[iy][ix] : 1;
      o->udef[iy][ix][0] /= normfac;
      o->udef[iy][ix][1] /= normfac;
      o->dist[iy][ix] = o->dist[iy][ix] >= 0 ? sqrt(o->dist[iy][ix]) : -sqrt(-o->dist[iy][ix]);
      B[iy * BS + ix] = dmax(B[iy * BS + ix], o->dist[iy][ix]);
    }
  for (int iy = 0; iy < BS; iy++)
    for (int ix = 0; ix < BS; ix++) { o->chi[iy][ix] = 0; o->rho[iy][ix] = 1; }
}

/* Fish::create (Fish.cpp:60-218) */
static void fish_create(orc_sim* s, shape_t* sh)
{
  fish_t* f = sh->fish;
  fish_frame(f, s->time);
  fish_segments(f, sh->st.center, sh->st.orientation);
  for (int b = 0; b < s->nb; ++b) { free(sh->ob[b]); sh->ob[b] = NULL; }
  fframe_t F;
  F.pos[0] = sh->st.center[0];
  F.pos[1] = sh->st.center[1];
  double sn, cs;
  orc_det_sincos(sh->st.orientation, &sn, &cs);
  F.R[0][0] = cs; F.R[0][1] = -sn; F.R[1][0] = sn; F.R[1][1] = cs;
  double* TMP = s->f[ORC_TMP];
  #pragma omp parallel
  {
    int* sl = (int*)malloc(sizeof(int) * (size_t)f->Nseg);
    #pragma omp for schedule(dynamic, 1)
    for (int b = 0; b < s->nb; ++b) {
      double pStart[2], pEnd[2];
      bpos(s, b, 0, 0, pStart);
      bpos(s, b, BS - 1, BS - 1, pEnd);
      int ns = 0;
      for (int q = 0; q < f->Nseg; ++q)
        if (seg_intersects(&f->seg[q], pStart, pEnd)) sl[ns++] = q;
      if (!ns) continue;
      oblock* o = (oblock*)calloc(1, sizeof(oblock));
      sh->ob[b] = o;
      put_fish_on_block(s, f, &F, b, sl, ns, o, TMP + (size_t)b * BS * BS);
    }
    free(sl);
  }
}

/* PutObjectsOnGrid::putChiOnGrid, PutObjectsOnGrid.cpp:17-147 */
static void put_chi_on_grid(orc_sim* s, shape_t* sh)
{
  const double h = s->h, i2h = 0.5 / h, fac = 0.5 * h;
  double* px = (double*)calloc((size_t)s->nb * 3, sizeof(double));
  #pragma omp parallel
  {
    double* L = (double*)malloc(sizeof(double) * LW * LW);
    #pragma omp for schedule(dynamic, 1)
    for (int b = 0; b < s->nb; ++b) {
      oblock* o = sh->ob[b];
      if (!o) continue;
      lab_load(s, s->f[ORC_TMP], 1, b, L);
      double* CHI = blk(s, ORC_CHI, b);
      double* IRHO = blk(s, ORC_INVRHO, b);
      double _x = 0, _y = 0, _m = 0;
      for (int iy = 0; iy < BS; iy++)
        for (int ix = 0; ix < BS; ix++) {
          const double sdf = o->dist[iy][ix];
          double X;
          if (sdf > +2 * h || sdf < -2 * h) {
            X = sdf > 0 ? 1 : 0;
          } else {
            const double distPx = ix + 1 == BS ? LAB(L, 1, ix + 1, iy, 0) : o->dist[iy][ix + 1];
            const double distMx = ix == 0 ? LAB(L, 1, ix - 1, iy, 0) : o->dist[iy][ix - 1];
            const double distPy = iy + 1 == BS ? LAB(L, 1, ix, iy + 1, 0) : o->dist[iy + 1][ix];
            const double distMy = iy == 0 ? LAB(L, 1, ix, iy - 1, 0) : o->dist[iy - 1][ix];
            const double IplusX = distPx < 0 ? 0 : distPx, IminuX = distMx < 0 ? 0 : distMx;
            const double IplusY = distPy < 0 ? 0 : distPy, IminuY = distMy < 0 ? 0 : distMy;
            const double gradIX = i2h * (IplusX - IminuX), gradIY = i2h * (IplusY - IminuY);
            const double gradUX = i2h * (distPx - distMx), gradUY = i2h * (distPy - distMy);
            const double gradUSq = gradUX * gradUX + gradUY * gradUY;
            const double denum = 1 / (gradUSq < EPS ? EPS : gradUSq);
            X = (gradIX * gradUX + gradIY * gradUY) * denum;
          }
          o->chi[iy][ix] = X;
          if (X >= CHI[iy * BS + ix]) {
            CHI[iy * BS + ix] = X;
            IRHO[iy * BS + ix] = X / o->rho[iy][ix] + (1 - X) * IRHO[iy * BS + ix];
          }
          if (X > 0) {
            double p[2];
            bpos(s, b, ix, iy, p);
            _x += o->rho[iy][ix] * X * h * h * (p[0] - sh->st.centerOfMass[0]);
            _y += o->rho[iy][ix] * X * h * h * (p[1] - sh->st.centerOfMass[1]);
            _m += o->rho[iy][ix] * X * h * h;
          }
          if (sdf > +2 * h || sdf < -2 * h) continue;
          {
            const double distPx = ix + 1 == BS ? LAB(L, 1, ix + 1, iy, 0) : o->dist[iy][ix + 1];
            const double distMx = ix == 0 ? LAB(L, 1, ix - 1, iy, 0) : o->dist[iy][ix - 1];
            const double distPy = iy + 1 == BS ? LAB(L, 1, ix, iy + 1, 0) : o->dist[iy + 1][ix];
            const double distMy = iy == 0 ? LAB(L, 1, ix, iy - 1, 0) : o->dist[iy - 1][ix];
            const double HplusX = fabs(distPx) < EPS ? 0.5 : (distPx < 0 ? 0 : 1);
            const double HminuX = fabs(distMx) < EPS ? 0.5 : (distMx < 0 ? 0 : 1);
            const double HplusY = fabs(distPy) < EPS ? 0.5 : (distPy < 0 ? 0 : 1);
            const double HminuY = fabs(distMy) < EPS ? 0.5 : (distMy < 0 ? 0 : 1);
            const double gradUX = i2h * (distPx - distMx), gradUY = i2h * (distPy - distMy);
            const double gradHX = (HplusX - HminuX), gradHY = (HplusY - HminuY);
            const double gradUSq = gradUX * gradUX + gradUY * gradUY;
            const double denum = 1 / (gradUSq < EPS ? EPS : gradUSq);
            const double D = fac * (gradHX * gradUX + gradHY * gradUY) * denum;
            if (D > 0) { /* ObstacleBlock::write, ObstacleBlock.h:95-105 */
              const int k = o->nsurf++;
              o->sx[k] = ix; o->sy[k] = iy;
              o->dchidx[k] = -D * gradUX; o->dchidy[k] = -D * gradUY; o->delta[k] = D;
            }
          }
        }
      px[3 * b] = _x; px[3 * b + 1] = _y; px[3 * b + 2] = _m;
    }
    free(L);
  }
  double _x = 0, _y = 0, _m = 0;
  for (int b = 0; b < s->nb; ++b) { _x += px[3 * b]; _y += px[3 * b + 1]; _m += px[3 * b + 2]; }
  free(px);
  if (_m > EPS) {
    sh->st.centerOfMass[0] += _x / _m;
    sh->st.centerOfMass[1] += _y / _m;
    sh->st.M = _m;
  }
}

/* Shape::integrateObstBlock + removeMoments, Shape.cpp:117-212 */
static void remove_moments(orc_sim* s, shape_t* sh)
{
  orc_shape_state* t = &sh->st;
  const double hsq = s->h * s->h;
  double* part = (double*)calloc((size_t)s->nb * 7, sizeof(double));
  #pragma omp parallel for schedule(dynamic, 1)
  for (int b = 0; b < s->nb; ++b) {
    const oblock* o = sh->ob[b];
    if (!o) continue;
    double _x = 0, _y = 0, _m = 0, _j = 0, _u = 0, _v = 0, _a = 0;
    for (int iy = 0; iy < BS; ++iy)
      for (int ix = 0; ix < BS; ++ix) {
        if (o->chi[iy][ix] <= 0) continue;
        double p[2];
        bpos(s, b, ix, iy, p);
        const double rhochi = o->chi[iy][ix] * o->rho[iy][ix] * hsq;
        p[0] -= t->centerOfMass[0];
        p[1] -= t->centerOfMass[1];
        _x += rhochi * p[0];
        _y += rhochi * p[1];
        _m += rhochi;
        _j += rhochi * (p[0] * p[0] + p[1] * p[1]);
        _u += rhochi * o->udef[iy][ix][0];
        _v += rhochi * o->udef[iy][ix][1];
        _a += rhochi * (p[0] * o->udef[iy][ix][1] - p[1] * o->udef[iy][ix][0]);
      }
    double* q = part + 7 * (size_t)b;
    q[0] = _x; q[1] = _y; q[2] = _m; q[3] = _j; q[4] = _u; q[5] = _v; q[6] = _a;
  }
  double I[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int b = 0; b < s->nb; ++b)
    for (int k = 0; k < 7; ++k) I[k] += part[7 * (size_t)b + k];
  free(part);
  I[4] /= I[2]; I[5] /= I[2]; I[6] /= I[3];
  t->M = I[2]; t->J = I[3];
  const double dCx = t->center[0] - t->centerOfMass[0];
  const double dCy = t->center[1] - t->centerOfMass[1];
  double sn, cs;
  orc_det_sincos(t->orientation, &sn, &cs);
  t->d_gm[0] = dCx * cs + dCy * sn;
  t->d_gm[1] = -dCx * sn + dCy * cs;
  #pragma omp parallel for schedule(dynamic)
  for (int b = 0; b < s->nb; ++b) {
    oblock* o = sh->ob[b];
    if (!o) continue;
    for (int iy = 0; iy < BS; ++iy)
      for (int ix = 0; ix < BS; ++ix) {
        double p[2];
        bpos(s, b, ix, iy, p);
        p[0] -= t->centerOfMass[0];
        p[1] -= t->centerOfMass[1];
        o->udef[iy][ix][0] -= I[4] - I[6] * p[1];
        o->udef[iy][ix][1] -= I[5] + I[6] * p[0];
      }
  }
}

/* PutObjectsOnGrid::putObjectVelOnGrid, PutObjectsOnGrid.cpp:149-184 */
static void put_object_vel_on_grid(orc_sim* s, shape_t* sh)
{
  #pragma omp parallel for schedule(dynamic)
  for (int b = 0; b < s->nb; ++b) {
    const oblock* o = sh->ob[b];
    if (!o) continue;
    double* UDEF = blk(s, ORC_UDEF, b);
    const double* CHI = blk(s, ORC_CHI, b);
    for (int iy = 0; iy < BS; iy++)
      for (int ix = 0; ix < BS; ix++) {
        if (o->chi[iy][ix] < CHI[iy * BS + ix] || o->chi[iy][ix] <= 0) continue;
        UDEF[2 * (iy * BS + ix)] += o->udef[iy][ix][0];
        UDEF[2 * (iy * BS + ix) + 1] += o->udef[iy][ix][1];
      }
  }
}

/* Shape::updatePosition, Shape.cpp:67-88 */
static void update_position(orc_sim* s, shape_t* sh, double dt)
{
  orc_shape_state* t = &sh->st;
  t->centerOfMass[0] += dt * (t->u + s->uinfx);
  t->centerOfMass[1] += dt * (t->v + s->uinfy);
  t->labCenterOfMass[0] += dt * t->u;
  t->labCenterOfMass[1] += dt * t->v;
  t->orientation += dt * t->omega;
  t->orientation = t->orientation > M_PI ? t->orientation - 2 * M_PI : t->orientation;
  t->orientation = t->orientation < -M_PI ? t->orientation + 2 * M_PI : t->orientation;
  double sinang, cosang;
  orc_det_sincos(t->orientation, &sinang, &cosang);
  t->center[0] = t->centerOfMass[0] + cosang * t->d_gm[0] - sinang * t->d_gm[1];
  t->center[1] = t->centerOfMass[1] + sinang * t->d_gm[0] + cosang * t->d_gm[1];
  /* Fish::updatePosition, Fish.cpp:180-185 */
  if (sh->fish) sh->fish->theta_internal -= dt * sh->fish->angvel_internal;
}

/* PutObjectsOnGrid::operator(), PutObjectsOnGrid.cpp:186-241 */
int orc_put_objects(orc_sim* s, double dt)
{
  checker(s);
  const size_t ncell = (size_t)s->nb * BS * BS;
  #pragma omp parallel for schedule(static)
  for (size_t i = 0; i < ncell; ++i) {
    s->f[ORC_CHI][i] = 0;
    s->f[ORC_TMP][i] = -1;
    s->f[ORC_UDEF][2 * i] = 0;
    s->f[ORC_UDEF][2 * i + 1] = 0;
    s->f[ORC_INVRHO][i] = 1;
  }
  checker(s);
  int nSum[2] = {0, 0};
  double uSum[2] = {0, 0};
  for (int k = 0; k < s->nshapes; ++k) { /* Shape::updateLabVelocity, Shape.cpp:61-65 */
    shape_t* sh = &s->shapes[k];
    if (sh->d.bFixedx) { nSum[0]++; uSum[0] -= sh->st.u; }
    if (sh->d.bFixedy) { nSum[1]++; uSum[1] -= sh->st.v; }
  }
  if (nSum[0] > 0) s->uinfx = uSum[0] / nSum[0];
  if (nSum[1] > 0) s->uinfy = uSum[1] / nSum[1];
  int err = 0;
  for (int k = 0; k < s->nshapes; ++k) {
    shape_t* sh = &s->shapes[k];
    update_position(s, sh, dt);
    const double* p = sh->st.center;
    if (p[0] < 0 || p[0] > s->extents[0] || p[1] < 0 || p[1] > s->extents[1]) err = -2;
  }
  if (err) return err; /* reference: printf + exit(0), PutObjectsOnGrid.cpp:213-218 */
  checker(s);
  for (int k = 0; k < s->nshapes; ++k) {
    if (s->shapes[k].fish) fish_create(s, &s->shapes[k]);
    else disk_create(s, &s->shapes[k]);
  }
  checker(s);
  for (int k = 0; k < s->nshapes; ++k) put_chi_on_grid(s, &s->shapes[k]);
  checker(s);
  for (int k = 0; k < s->nshapes; ++k) {
    remove_moments(s, &s->shapes[k]);
    put_object_vel_on_grid(s, &s->shapes[k]);
  }
  return 0;
}

/* ---------------------------------------------------------------- advDiff */
static inline double dU_adv_dif(const double* V, const double uinf[2], double advF, double difF,
                                int ix, int iy)
{ /* advDiff.cpp:14-23 */
  const double upx = LAB(V, 2, ix + 1, iy, 0), upy = LAB(V, 2, ix, iy + 1, 0);
  const double ucc = LAB(V, 2, ix, iy, 0), vcc = LAB(V, 2, ix, iy, 1);
  const double ulx = LAB(V, 2, ix - 1, iy, 0), uly = LAB(V, 2, ix, iy - 1, 0);
  const double dUadv = (ucc + uinf[0]) * (upx - ulx) + (vcc + uinf[1]) * (upy - uly);
  const double dUdif = upx + upy + ulx + uly - 4 * ucc;
  return advF * dUadv + difF * dUdif;
}
static inline double dV_adv_dif(const double* V, const double uinf[2], double advF, double difF,
                                int ix, int iy)
{ /* advDiff.cpp:25-34 */
  const double vpx = LAB(V, 2, ix + 1, iy, 1), vpy = LAB(V, 2, ix, iy + 1, 1);
  const double ucc = LAB(V, 2, ix, iy, 0), vcc = LAB(V, 2, ix, iy, 1);
  const double vlx = LAB(V, 2, ix - 1, iy, 1), vly = LAB(V, 2, ix, iy - 1, 1);
  const double dVadv = (ucc + uinf[0]) * (vpx - vlx) + (vcc + uinf[1]) * (vpy - vly);
  const double dVdif = vpx + vpy + vlx + vly - 4 * vcc;
  return advF * dVadv + difF * dVdif;
}

/* advDiff::operator(), advDiff.cpp:36-118 */
void orc_advdiff(orc_sim* s, double dt)
{
  checker(s);
  const double UINF[2] = {s->uinfx, s->uinfy}, h = s->h;
  const double dfac = (s->cfg.nu / h) * (dt / h), afac = -0.5 * dt / h;
  const double norUinf = dmax(dmax(fabs(UINF[0]), fabs(UINF[1])), EPS);
  const double fadeW = 1 - sq(dmax(UINF[0], 0) / norUinf) / BC_KILL_FAC;
  const double fadeS = 1 - sq(dmax(UINF[1], 0) / norUinf) / BC_KILL_FAC;
  const double fadeE = 1 - sq(dmin(UINF[0], 0) / norUinf) / BC_KILL_FAC;
  const double fadeN = 1 - sq(dmin(UINF[1], 0) / norUinf) / BC_KILL_FAC;
  #pragma omp parallel
  {
    double* V = (double*)malloc(sizeof(double) * LW * LW * 2);
    #pragma omp for schedule(static)
    for (int b = 0; b < s->nb; ++b) {
      lab_load(s, s->f[ORC_VEL], 2, b, V);
      double* TMP = blk(s, ORC_TMPV, b);
      if (BX(s, b) == 0)
        for (int iy = -1; iy <= BS; ++iy) { LAB(V, 2, -1, iy, 0) *= fadeW; LAB(V, 2, -1, iy, 1) *= fadeW; }
      if (BY(s, b) == 0)
        for (int ix = -1; ix <= BS; ++ix) { LAB(V, 2, ix, -1, 0) *= fadeS; LAB(V, 2, ix, -1, 1) *= fadeS; }
      if (BX(s, b) == s->bpdx - 1)
        for (int iy = -1; iy <= BS; ++iy) { LAB(V, 2, BS, iy, 0) *= fadeE; LAB(V, 2, BS, iy, 1) *= fadeE; }
      if (BY(s, b) == s->bpdy - 1)
        for (int ix = -1; ix <= BS; ++ix) { LAB(V, 2, ix, BS, 0) *= fadeN; LAB(V, 2, ix, BS, 1) *= fadeN; }
      for (int iy = 0; iy < BS; ++iy)
        for (int ix = 0; ix < BS; ++ix) {
          TMP[2 * (iy * BS + ix)] = LAB(V, 2, ix, iy, 0) + dU_adv_dif(V, UINF, afac, dfac, ix, iy);
          TMP[2 * (iy * BS + ix) + 1] = LAB(V, 2, ix, iy, 1) + dV_adv_dif(V, UINF, afac, dfac, ix, iy);
        }
    }
    free(V);
  }
  memcpy(s->f[ORC_VEL], s->f[ORC_TMPV], sizeof(double) * (size_t)s->nb * BS * BS * 2);
  /* boundary flux correction, advDiff.cpp:89-116 */
  double* part = (double*)calloc((size_t)s->nbnd, sizeof(double));
  #pragma omp parallel for schedule(static)
  for (int k = 0; k < s->nbnd; ++k) {
    const int b = s->bnd[k];
    const double* V = blk(s, ORC_VEL, b);
    double IF = 0;
    if (BX(s, b) == 0) for (int iy = 0; iy < BS; ++iy) IF -= V[2 * (iy * BS + 0)];
    if (BX(s, b) == s->bpdx - 1) for (int iy = 0; iy < BS; ++iy) IF += V[2 * (iy * BS + BS - 1)];
    if (BY(s, b) == 0) for (int ix = 0; ix < BS; ++ix) IF -= V[2 * (0 * BS + ix) + 1];
    if (BY(s, b) == s->bpdy - 1) for (int ix = 0; ix < BS; ++ix) IF += V[2 * ((BS - 1) * BS + ix) + 1];
    part[k] = IF;
  }
  double IF = 0;
  for (int k = 0; k < s->nbnd; ++k) IF += part[k];
  free(part);
  const double corr = IF / (2 * (BS * s->bpdy - 0) + 2 * (BS * s->bpdx - 0));
  #pragma omp parallel for schedule(static)
  for (int k = 0; k < s->nbnd; ++k) {
    const int b = s->bnd[k];
    double* V = blk(s, ORC_VEL, b);
    if (BX(s, b) == 0) for (int iy = 0; iy < BS; ++iy) V[2 * (iy * BS + 0)] += corr;
    if (BX(s, b) == s->bpdx - 1) for (int iy = 0; iy < BS; ++iy) V[2 * (iy * BS + BS - 1)] -= corr;
    if (BY(s, b) == 0) for (int ix = 0; ix < BS; ++ix) V[2 * (0 * BS + ix) + 1] += corr;
    if (BY(s, b) == s->bpdy - 1) for (int ix = 0; ix < BS; ++ix) V[2 * ((BS - 1) * BS + ix) + 1] -= corr;
  }
}

/* --------------------------------------------------------- PressureSingle */
/* PressureSingle::updatePressureRHS, PressureSingle.cpp:110-214 */
void orc_pressure_rhs(orc_sim* s, double dt)
{
  const double h = s->h, facDiv = 0.5 * h / dt;
  #pragma omp parallel
  {
    double* V = (double*)malloc(sizeof(double) * LW * LW * 2);
    #pragma omp for schedule(static)
    for (int b = 0; b < s->nb; ++b) {
      lab_load(s, s->f[ORC_VEL], 2, b, V);
      double* TMP = blk(s, ORC_TMP, b);
      for (int iy = 0; iy < BS; ++iy)
        for (int ix = 0; ix < BS; ++ix) {
          const double divVx = LAB(V, 2, ix + 1, iy, 0) - LAB(V, 2, ix - 1, iy, 0);
          const double divVy = LAB(V, 2, ix, iy + 1, 1) - LAB(V, 2, ix, iy - 1, 1);
          TMP[iy * BS + ix] = facDiv * (divVx + divVy);
        }
    }
    free(V);
  }
  const int ns = s->nshapes;
  if (ns == 0) return;
  double* part = (double*)calloc((size_t)s->nb * ns * 3, sizeof(double));
  #pragma omp parallel
  {
    double* C = (double*)malloc(sizeof(double) * LW * LW);
    #pragma omp for schedule(static)
    for (int b = 0; b < s->nb; ++b)
      for (int j = 0; j < ns; ++j) {
        const oblock* o = s->shapes[j].ob[b];
        if (!o) continue;
        lab_load(s, s->f[ORC_CHI], 1, b, C);
        double* TMP = blk(s, ORC_TMP, b);
        double sum = 0, pos = 0, neg = 0;
        for (int iy = 0; iy < BS; ++iy)
          for (int ix = 0; ix < BS; ++ix) {
            if (LAB(C, 1, ix, iy, 0) > o->chi[iy][ix]) continue;
            const double gradXx = LAB(C, 1, ix + 1, iy, 0) - LAB(C, 1, ix - 1, iy, 0);
            const double gradXy = LAB(C, 1, ix, iy + 1, 0) - LAB(C, 1, ix, iy - 1, 0);
            const double srcBulk = -o->chi[iy][ix] * TMP[iy * BS + ix];
            sum += srcBulk;
            TMP[iy * BS + ix] += srcBulk;
            const double isPerim = (gradXx * gradXx + gradXy * gradXy > 0) ? 1.0 : 0.0;
            const double T = TMP[iy * BS + ix];
            pos += isPerim * T * (double)(T > 0);
            neg -= isPerim * T * (double)(T < 0);
          }
        double* q = part + 3 * ((size_t)b * ns + j);
        q[0] = sum; q[1] = pos; q[2] = neg;
      }
    free(C);
  }
  double* corr = (double*)calloc((size_t)ns, sizeof(double));
  for (int j = 0; j < ns; ++j) {
    double sum = 0, pos = 0, neg = 0;
    for (int b = 0; b < s->nb; ++b) {
      const double* q = part + 3 * ((size_t)b * ns + j);
      sum += q[0]; pos += q[1]; neg += q[2];
    }
    const double corrDenom = sum > 0 ? pos : neg;
    corr[j] = sum / dmax(corrDenom, EPS);
  }
  free(part);
  #pragma omp parallel
  {
    double* C = (double*)malloc(sizeof(double) * LW * LW);
    #pragma omp for schedule(static)
    for (int b = 0; b < s->nb; ++b)
      for (int j = 0; j < ns; ++j) {
        const oblock* o = s->shapes[j].ob[b];
        if (!o) continue;
        double* TMP = blk(s, ORC_TMP, b);
        lab_load(s, s->f[ORC_CHI], 1, b, C);
        for (int iy = 0; iy < BS; ++iy)
          for (int ix = 0; ix < BS; ++ix) {
            if (LAB(C, 1, ix, iy, 0) > o->chi[iy][ix]) continue;
            const double gradXx = LAB(C, 1, ix + 1, iy, 0) - LAB(C, 1, ix - 1, iy, 0);
            const double gradXy = LAB(C, 1, ix, iy + 1, 0) - LAB(C, 1, ix, iy - 1, 0);
            const int isPerim = gradXx * gradXx + gradXy * gradXy > 0;
            double* T = &TMP[iy * BS + ix];
            if (isPerim && *T > 0 && corr[j] > 0) *T -= corr[j] * *T;
            else if (isPerim && *T < 0 && corr[j] < 0) *T += corr[j] * *T;
          }
      }
    free(C);
  }
  free(corr);
}

/* PoissonSolver::cub2rhs, PoissonSolver.cpp:61-141 */
static inline double fade_arg(const orc_sim* s, const double p[2])
{
  const double invFadeX = 1 / dmax(s->fadeLenX, EPS), invFadeY = 1 / dmax(s->fadeLenY, EPS);
  const double yt = invFadeY * dmax(0, s->fadeLenY - s->extents[1] + p[1]);
  const double yb = invFadeY * dmax(0, s->fadeLenY - p[1]);
  const double xt = invFadeX * dmax(0, s->fadeLenX - s->extents[0] + p[0]);
  const double xb = invFadeX * dmax(0, s->fadeLenX - p[0]);
  return dmin(dmax(dmax(dmax(yt, yb), xt), xb), 1);
}
static int fade_touching(const orc_sim* s, int b)
{
  double mn[2], mx[2];
  bpos(s, b, 0, 0, mn);
  bpos(s, b, BS - 1, BS - 1, mx);
  return s->fadeLenX >= mn[0] || s->fadeLenX >= s->extents[0] - mx[0] ||
         s->fadeLenY >= mn[1] || s->fadeLenY >= s->extents[1] - mx[1];
}

void orc_cub2rhs(orc_sim* s)
{
  double* part = (double*)calloc((size_t)s->nb * 3, sizeof(double));
  #pragma omp parallel for schedule(dynamic)
  for (int b = 0; b < s->nb; ++b) {
    if (!fade_touching(s, b)) continue;
    const double* B = blk(s, ORC_TMP, b);
    double sumRHS = 0, sumPOS = 0, sumNEG = 0;
    for (int iy = 0; iy < BS; ++iy)
      for (int ix = 0; ix < BS; ++ix) {
        double p[2];
        bpos(s, b, ix, iy, p);
        const double fa = fade_arg(s, p);
        const double fadedRHS = B[iy * BS + ix] * (fa * fa);
        if (fadedRHS < 0) sumNEG -= fadedRHS;
        if (fadedRHS > 0) sumPOS += fadedRHS;
        sumRHS += fadedRHS;
      }
    part[3 * b] = sumRHS; part[3 * b + 1] = sumPOS; part[3 * b + 2] = sumNEG;
  }
  double sumRHS = 0, sumPOS = 0, sumNEG = 0;
  for (int b = 0; b < s->nb; ++b) { sumRHS += part[3 * b]; sumPOS += part[3 * b + 1]; sumNEG += part[3 * b + 2]; }
  free(part);
  sumNEG = dmax(sumNEG, EPS);
  sumPOS = dmax(sumPOS, EPS);
  const double corr = sumRHS > 0 ? sumRHS / sumPOS : sumRHS / sumNEG;
  #pragma omp parallel for schedule(dynamic)
  for (int b = 0; b < s->nb; ++b) {
    if (!fade_touching(s, b)) continue;
    double* B = blk(s, ORC_TMP, b);
    for (int iy = 0; iy < BS; ++iy)
      for (int ix = 0; ix < BS; ++ix) {
        double p[2];
        bpos(s, b, ix, iy, p);
        const double fa = fade_arg(s, p);
        const double fadeRHS = B[iy * BS + ix] * (fa * fa);
        const double fadeCorrection = corr * (corr > 0 ? dmax(fadeRHS, 0) : -dmin(fadeRHS, 0));
        B[iy * BS + ix] = B[iy * BS + ix] * (1 - fa * fa) + fadeCorrection;
      }
  }
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < s->nb; ++b) {
    const size_t blockStart = (size_t)BS * BX(s, b) + s->stride * (size_t)(BS * BY(s, b));
    const double* B = blk(s, ORC_TMP, b);
    for (int iy = 0; iy < BS; iy++)
      for (int ix = 0; ix < BS; ix++) s->buffer[blockStart + ix + s->stride * iy] = B[iy * BS + ix];
  }
}

/* PoissonSolver::sol2cub, PoissonSolver.cpp:143-168 */
static void sol2cub(orc_sim* s)
{
  #pragma omp parallel for schedule(static)
  for (int b = 0; b < s->nb; ++b) {
    const size_t blockStart = (size_t)BS * BX(s, b) + s->stride * (size_t)(BS * BY(s, b));
    double* B = blk(s, ORC_PRES, b);
    for (int iy = 0; iy < BS; iy++)
      for (int ix = 0; ix < BS; ix++) B[iy * BS + ix] = s->buffer[blockStart + ix + s->stride * iy];
  }
}

/* spectral step: FFTW_periodic::_solve_spectral (FFTW_periodic.h:77-92) or
 * the Green's function product (FFTW_freespace.cpp:31-42) */
static void spectral(orc_sim* s)
{
  double* io = s->buffer;
  const int MX = s->MX, MY = s->MY, MXh = s->MXh;
  if (s->cfg.poisson == 0) {
    const double waveFactX = 2 * M_PI / MX, waveFactY = 2 * M_PI / MY;
    const double norm_factor = 1.0 / ((double)MX * MY);
    #pragma omp parallel for schedule(static)
    for (int j = 0; j < MY; ++j)
      for (int i = 0; i < MXh; ++i) {
        const int kx = i <= MX / 2 ? i : MX - i, ky = j <= MY / 2 ? j : MY - j;
        const double rkx = kx * waveFactX, rky = ky * waveFactY;
        const double kinv = -1 / (rkx * rkx + rky * rky);
        io[2 * ((size_t)j * MXh + i)] *= kinv * norm_factor;
        io[2 * ((size_t)j * MXh + i) + 1] *= kinv * norm_factor;
      }
    io[0] = 0; io[1] = 0;
  } else {
    #pragma omp parallel for schedule(static)
    for (int j = 0; j < MY; ++j)
      for (int i = 0; i < MXh; ++i) {
        const size_t li = (size_t)j * MXh + i;
        io[2 * li] *= s->kernel[li];
        io[2 * li + 1] *= s->kernel[li];
      }
  }
}

/* FFTW's REDFT10 (DCT-II, Y_k = 2 sum_j x_j cos(pi (j+1/2) k / n)) or REDFT01
 * (DCT-III, y_j = X_0 + 2 sum_{k>=1} X_k cos(pi k (j+1/2) / n)) of n values at
 * stride `stride`, by the direct sums (the FFTW documentation's definitions). */
static void dct_direct(double* x, int n, size_t stride, int type3, const double* cs, double* tmp)
{
  for (int k = 0; k < n; ++k) {
    double acc = 0;
    if (!type3) {
      for (int j = 0; j < n; ++j) acc += x[(size_t)j * stride] * cs[(size_t)k * n + j];
      tmp[k] = 2 * acc;
    } else {
      for (int m = 1; m < n; ++m) acc += x[(size_t)m * stride] * cs[(size_t)m * n + k];
      tmp[k] = x[0] + 2 * acc;
    }
  }
  for (int k = 0; k < n; ++k) x[(size_t)k * stride] = tmp[k];
}

/* cs[k n + j] = cos(pi (j + 1/2) k / n) */
static double* dct_table(int n)
{
  double* cs = (double*)malloc(sizeof(double) * (size_t)n * n);
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j) cs[(size_t)k * n + j] = cos(M_PI * (j + 0.5) * k / n);
  return cs;
}

/* 2D r2r transform of the MY x MX buffer (rows, then columns) */
static void dct2d(orc_sim* s, int type3)
{
  const int MX = s->MX, MY = s->MY;
  double* csx = dct_table(MX);
  double* csy = dct_table(MY);
  #pragma omp parallel
  {
    double* tmp = (double*)malloc(sizeof(double) * (size_t)(MX > MY ? MX : MY));
    #pragma omp for schedule(static)
    for (int j = 0; j < MY; ++j) dct_direct(s->buffer + (size_t)j * MX, MX, 1, type3, csx, tmp);
    #pragma omp for schedule(static)
    for (int i = 0; i < MX; ++i) dct_direct(s->buffer + i, MY, (size_t)MX, type3, csy, tmp);
    free(tmp);
  }
  free(csx);
  free(csy);
}

/* FFTW_dirichlet::_solve (FFTW_dirichlet.h:45-61).  Line 51 declares the
 * blend weight as `static constexpr tol = 0.1;`, which has no type; it is
 * taken as the double 0.1 the initializer states.  COScoef are floats
 * (FFTW_dirichlet.h:26-27,65-68). */
static void cosine_spectral(orc_sim* s)
{
  const int MX = s->MX, MY = s->MY;
  const double waveFactX = M_PI / MX, waveFactY = M_PI / MY;
  const double tol = 0.1;
  const double norm_factor = 0.25 / ((double)MX * MY);
  double* io = s->buffer;
  #pragma omp parallel for schedule(static)
  for (int j = 0; j < MY; ++j)
    for (int i = 0; i < MX; ++i) {
      const float cx = (float)cos(M_PI / MX * 2.0 * i), cy = (float)cos(M_PI / MY * 2.0 * j);
      const double rkx = (i + 0.5) * waveFactX, rky = (j + 0.5) * waveFactY;
      const double denomFD = 1 - (double)cx / 2 - (double)cy / 2;
      const double denomSP = rkx * rkx + rky * rky;
      io[(size_t)j * MX + i] *= -norm_factor / ((1 - tol) * denomFD + tol * denomSP);
    }
  io[0] = 0;
}

static void fft_solve_buffer(orc_sim* s)
{
  if (s->cfg.poisson == 2) {
    const double t0 = omp_get_wtime();
    dct2d(s, 0);
    cosine_spectral(s);
    dct2d(s, 1);
    s->t_fft += omp_get_wtime() - t0;
    return;
  }
  const double t0 = omp_get_wtime();
  fft2d_exec(&s->fft, s->buffer, -1, s->cfg.poisson == 0 ? s->MY : s->NY);
  spectral(s);
  fft2d_exec(&s->fft, s->buffer, +1, -1);
  s->t_fft += omp_get_wtime() - t0;
}

/* PoissonSolver::solve (FFTW_periodic.h:135-166, FFTW_freespace.cpp:14-56) */
void orc_poisson_solve(orc_sim* s)
{
  if (s->cfg.poisson == 1) memset(s->buffer, 0, sizeof(double) * s->stride * (size_t)s->MY);
  checker(s);
  orc_cub2rhs(s);
  checker(s);
  fft_solve_buffer(s);
  checker(s); checker(s); checker(s);
  sol2cub(s);
}

int orc_poisson_dense(orc_sim* s, const double* rhs, double* sol)
{
  memset(s->buffer, 0, sizeof(double) * s->stride * (size_t)s->MY);
  for (int j = 0; j < s->NY; ++j)
    memcpy(s->buffer + s->stride * (size_t)j, rhs + (size_t)j * s->NX, sizeof(double) * s->NX);
  fft_solve_buffer(s);
  for (int j = 0; j < s->NY; ++j)
    memcpy(sol + (size_t)j * s->NX, s->buffer + s->stride * (size_t)j, sizeof(double) * s->NX);
  return 0;
}

int orc_fft1d(double* data, int n, int sign)
{
  fft1d p;
  fft1d_init(&p, n);
  double* work = (double*)malloc(sizeof(double) * (fft1d_work_size(&p) + 2));
  fft1d_exec(&p, data, sign, work);
  free(work);
  fft1d_free(&p);
  return 0;
}

/* PressureSingle::pressureCorrection, PressureSingle.cpp:216-241 */
void orc_pressure_correction(orc_sim* s, double dt)
{
  const double h = s->h, pFac = -0.5 * dt / h;
  #pragma omp parallel
  {
    double* P = (double*)malloc(sizeof(double) * LW * LW);
    #pragma omp for schedule(static)
    for (int b = 0; b < s->nb; ++b) {
      lab_load(s, s->f[ORC_PRES], 1, b, P);
      double* V = blk(s, ORC_VEL, b);
      for (int iy = 0; iy < BS; ++iy)
        for (int ix = 0; ix < BS; ++ix) {
          V[2 * (iy * BS + ix)] = V[2 * (iy * BS + ix)] + pFac * (LAB(P, 1, ix + 1, iy, 0) - LAB(P, 1, ix - 1, iy, 0));
          V[2 * (iy * BS + ix) + 1] = V[2 * (iy * BS + ix) + 1] + pFac * (LAB(P, 1, ix, iy + 1, 0) - LAB(P, 1, ix, iy - 1, 0));
        }
    }
    free(P);
  }
}

/* 3x3 solve: Doolittle LU with partial pivoting (restates gsl_linalg_LU_decomp/solve) */
static void lu3_solve(double A[3][3], const double b[3], double x[3])
{
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

/* PressureSingle::integrateMomenta (PressureSingle.cpp:23-67), Shape::updateVelocity
 * (Shape.cpp:25-59), Disk::updateVelocity (ShapesSimple.cpp:83-87),
 * PressureSingle::preventCollidingObstacles (PressureSingle.cpp:243-344) */
static void update_velocity(orc_sim* s, shape_t* sh, double dt);
static void prevent_colliding_obstacles(orc_sim* s);

void orc_body_update(orc_sim* s, double dt)
{
  const double hsq = s->h * s->h;
  const int impl = s->cfg.implicit_penal;
  double* part = (double*)malloc(sizeof(double) * (size_t)s->nb * 7);
  for (int k = 0; k < s->nshapes; ++k) {
    shape_t* sh = &s->shapes[k];
    orc_shape_state* t = &sh->st;
    const double Cx = t->centerOfMass[0], Cy = t->centerOfMass[1];
    memset(part, 0, sizeof(double) * (size_t)s->nb * 7);
    #pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < s->nb; ++b) {
      const oblock* o = sh->ob[b];
      if (!o) continue;
      const double* VEL = blk(s, ORC_VEL, b);
      const double lambdt = s->cfg.lambda * dt;
      double PM = 0, PJ = 0, PX = 0, PY = 0, UM = 0, VM = 0, AM = 0;
      for (int iy = 0; iy < BS; ++iy)
        for (int ix = 0; ix < BS; ++ix) {
          if (o->chi[iy][ix] <= 0) continue;
          double udiff[2] = {VEL[2 * (iy * BS + ix)] - o->udef[iy][ix][0],
                             VEL[2 * (iy * BS + ix) + 1] - o->udef[iy][ix][1]};
          double F;
          if (!impl) F = hsq * o->rho[iy][ix] * o->chi[iy][ix];
          else {
            const double Xlamdt = o->chi[iy][ix] * lambdt;
            F = hsq * o->rho[iy][ix] * Xlamdt / (1 + Xlamdt);
          }
          double p[2];
          bpos(s, b, ix, iy, p);
          p[0] -= Cx; p[1] -= Cy;
          {
            double su = 0, sv = 0;
            surf_add(s, (int)(sh - s->shapes), p[0], p[1], &su, &sv);
            udiff[0] -= su;
            udiff[1] -= sv;
          }
          PM += F;
          PJ += F * (p[0] * p[0] + p[1] * p[1]);
          PX += F * p[0]; PY += F * p[1];
          UM += F * udiff[0]; VM += F * udiff[1];
          AM += F * (p[0] * udiff[1] - p[1] * udiff[0]);
        }
      double* q = part + 7 * (size_t)b;
      q[0] = PM; q[1] = PJ; q[2] = PX; q[3] = PY; q[4] = UM; q[5] = VM; q[6] = AM;
    }
    double S[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int b = 0; b < s->nb; ++b)
      for (int i = 0; i < 7; ++i) S[i] += part[7 * (size_t)b + i];
    t->penalM = S[0]; t->penalJ = S[1]; t->penalDX = S[2]; t->penalDY = S[3];
    t->fluidMomX = S[4]; t->fluidMomY = S[5]; t->fluidAngMom = S[6];
    update_velocity(s, sh, dt);
  }
  free(part);
  if (!s->no_collide) prevent_colliding_obstacles(s);
}

void orc_set_collisions(orc_sim* s, int enabled) { s->no_collide = !enabled; }

/* Fourier-mode surface velocity of disk `shape`, added wherever udef is read
 * (integrateMomenta, collisions, penalize, ComputeForces' deformation
 * power): with z = e^{i phi} = p/|p| about the centre of mass,
 * u_n = sum_m a_m Re z^m + b_m Im z^m along the normal z and
 * u_t = sum_m c_m Re z^m + d_m Im z^m along the tangent i z. */
int orc_set_surface_velocity(orc_sim* s, int shape, int M, const double* coef)
{
  if (M < 0 || M > 32 || (M > 0 && (shape < 0 || shape >= s->nshapes))) return -1;
  s->surf_shape = M > 0 ? shape : -1;
  s->surf_M = M;
  for (int i = 0; i < 4 * M; ++i) s->surf_coef[i] = coef[i];
  return 0;
}
static void surf_add(const orc_sim* s, int k, double p0, double p1, double* ux, double* uy)
{
  if (k != s->surf_shape || s->surf_M <= 0) return;
  const double r = sqrt(p0 * p0 + p1 * p1);
  const double zr = r > 0 ? p0 / r : 1.0, zi = r > 0 ? p1 / r : 0.0;
  const int M = s->surf_M;
  const double* c = s->surf_coef;
  double wr = 1, wi = 0, un = 0, ut = 0;
  for (int m = 0; m < M; ++m) {
    const double tr = wr * zr - wi * zi;
    wi = wr * zi + wi * zr;
    wr = tr;
    un += c[m] * wr + c[M + m] * wi;
    ut += c[2 * M + m] * wr + c[3 * M + m] * wi;
  }
  *ux += un * zr - ut * zi;
  *uy += un * zi + ut * zr;
}

/* Shape::updateVelocity (Shape.cpp:25-59) from the momenta in the state */
static void update_velocity(orc_sim* s, shape_t* sh, double dt)
{
  {
    orc_shape_state* t = &sh->st;
    double A[3][3] = {{t->penalM, 0, -t->penalDY}, {0, t->penalM, t->penalDX},
                      {-t->penalDY, t->penalDX, t->penalJ}};
    double bb[3] = {t->fluidMomX + dt * sh->appliedForceX, t->fluidMomY + dt * sh->appliedForceY,
                    t->fluidAngMom + dt * sh->appliedTorque};
    const double tf = sh->d.timeForced;
    if (sh->d.bForcedx && s->time < tf) { A[0][1] = 0; A[0][2] = 0; bb[0] = t->penalM * sh->forcedu; }
    if (sh->d.bForcedy && s->time < tf) { A[1][0] = 0; A[1][2] = 0; bb[1] = t->penalM * sh->forcedv; }
    if (sh->d.bBlockAng && s->time < tf) { A[2][0] = 0; A[2][1] = 0; bb[2] = t->penalJ * sh->forcedomega; }
    double x[3];
    lu3_solve(A, bb, x);
    if (!sh->d.bForcedy || s->time > tf) t->u = x[0]; /* quirk: Shape.cpp:56 */
    if (!sh->d.bForcedy || s->time > tf) t->v = x[1];
    if (!sh->d.bBlockAng || s->time > tf) t->omega = x[2];
    if (sh->d.tAccel > 0) {
      if (sh->d.bForcedx && s->time < sh->d.tAccel) t->u = (s->time / sh->d.tAccel) * sh->forcedu;
      if (sh->d.bForcedy && s->time < sh->d.tAccel) t->v = (s->time / sh->d.tAccel) * sh->forcedv;
    }
  }
}

/* PressureSingle::preventCollidingObstacles (PressureSingle.cpp:243-344) */
static void prevent_colliding_obstacles(orc_sim* s)
{
  const int N = s->nshapes;
  if (N < 2) return;
  double* coll = (double*)calloc((size_t)N * 10, sizeof(double));
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      if (i == j) continue;
      double* c = coll + 10 * (size_t)i;
      const orc_shape_state *si = &s->shapes[i].st, *sj = &s->shapes[j].st;
      for (int b = 0; b < s->nb; ++b) {
        const oblock *oi = s->shapes[i].ob[b], *oj = s->shapes[j].ob[b];
        if (!oi || !oj) continue;
        for (int iy = 0; iy < BS; ++iy)
          for (int ix = 0; ix < BS; ++ix) {
            if (oi->chi[iy][ix] <= 0 || oj->chi[iy][ix] <= 0) continue;
            double pos[2];
            bpos(s, b, ix, iy, pos);
            const double iUr = -si->omega * (pos[1] - si->centerOfMass[1]), iVr = si->omega * (pos[0] - si->centerOfMass[0]);
            const double jUr = -sj->omega * (pos[1] - sj->centerOfMass[1]), jVr = sj->omega * (pos[0] - sj->centerOfMass[0]);
            const double X = oi->chi[iy][ix], Y = oj->chi[iy][ix];
            double iUd = oi->udef[iy][ix][0], iVd = oi->udef[iy][ix][1];
            double jUd = oj->udef[iy][ix][0], jVd = oj->udef[iy][ix][1];
            surf_add(s, i, pos[0] - si->centerOfMass[0], pos[1] - si->centerOfMass[1], &iUd, &iVd);
            surf_add(s, j, pos[0] - sj->centerOfMass[0], pos[1] - sj->centerOfMass[1], &jUd, &jVd);
            c[0] += X; c[1] += X * pos[0]; c[2] += X * pos[1];
            c[3] += X * (si->u + iUr + iUd);
            c[4] += X * (si->v + iVr + iVd);
            c[5] += Y; c[6] += Y * pos[0]; c[7] += Y * pos[1];
            c[8] += Y * (sj->u + jUr + jUd);
            c[9] += Y * (sj->v + jVr + jVd);
          }
      }
    }
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) {
      if (i == j) continue;
      const double* c = coll + 10 * (size_t)i;
      if (c[0] < 1 || c[5] < 1) continue;
      const double iPX = c[1] / c[0], iPY = c[2] / c[0], iUX = c[3] / c[0], iUY = c[4] / c[0];
      const double jPX = c[6] / c[5], jPY = c[7] / c[5], jUX = c[8] / c[5], jUY = c[9] / c[5];
      const double CX = (iPX + jPX) / 2, CY = (iPY + jPY) / 2;
      const double dirX = iPX - jPX, dirY = iPY - jPY;
      const double hitVelX = jUX - iUX, hitVelY = jUY - iUY;
      const double normF = dmax(sqrt(dirX * dirX + dirY * dirY), EPS);
      const double NX = dirX / normF, NY = dirY / normF;
      const double projVel = hitVelX * NX + hitVelY * NY;
      if (projVel <= 0) continue;
      shape_t *shi = &s->shapes[i], *shj = &s->shapes[j];
      const int iForcedX = shi->d.bForcedx && s->time < shi->d.timeForced;
      const int iForcedY = shi->d.bForcedy && s->time < shi->d.timeForced;
      const int iForcedA = shi->d.bBlockAng && s->time < shi->d.timeForced;
      const int jForcedX = shj->d.bForcedx && s->time < shj->d.timeForced;
      const int jForcedY = shj->d.bForcedy && s->time < shj->d.timeForced;
      const double iInvMassX = iForcedX ? 0 : 1 / shi->st.M, iInvMassY = iForcedY ? 0 : 1 / shi->st.M;
      const double jInvMassX = jForcedX ? 0 : 1 / shj->st.M, jInvMassY = jForcedY ? 0 : 1 / shj->st.M;
      const double iInvJ = iForcedA ? 0 : 1 / shi->st.J;
      const double meanMassX = 2 / dmax(iInvMassX + jInvMassX, EPS);
      const double meanMassY = 2 / dmax(iInvMassY + jInvMassY, EPS);
      const double FXdt = meanMassX * projVel * NX, FYdt = meanMassY * projVel * NY;
      shi->st.u += FXdt * iInvMassX;
      shi->st.v += FYdt * iInvMassY;
      const double iCx = shi->st.centerOfMass[0], iCy = shi->st.centerOfMass[1];
      const double RcrossF = (CX - iCx) * FYdt - (CY - iCy) * FXdt;
      shi->st.omega += iInvJ * RcrossF;
    }
  free(coll);
}

/* PressureSingle::penalize, PressureSingle.cpp:69-108 */
void orc_penalize(orc_sim* s, double dt)
{
  #pragma omp parallel for schedule(dynamic, 1)
  for (int b = 0; b < s->nb; ++b)
    for (int k = 0; k < s->nshapes; ++k) {
      const shape_t* sh = &s->shapes[k];
      const oblock* o = sh->ob[b];
      if (!o) continue;
      const double u_s = sh->st.u, v_s = sh->st.v, omega_s = sh->st.omega;
      const double Cx = sh->st.centerOfMass[0], Cy = sh->st.centerOfMass[1];
      const double* CHI = blk(s, ORC_CHI, b);
      double* V = blk(s, ORC_VEL, b);
      for (int iy = 0; iy < BS; ++iy)
        for (int ix = 0; ix < BS; ++ix) {
          if (CHI[iy * BS + ix] > o->chi[iy][ix]) continue;
          if (o->chi[iy][ix] <= 0) continue;
          double p[2];
          bpos(s, b, ix, iy, p);
          p[0] -= Cx; p[1] -= Cy;
          const double alpha = s->cfg.implicit_penal ? 1 / (1 + s->cfg.lambda * dt * o->chi[iy][ix])
                                                     : 1 - o->chi[iy][ix];
          double US = u_s - omega_s * p[1] + o->udef[iy][ix][0];
          double VS = v_s + omega_s * p[0] + o->udef[iy][ix][1];
          surf_add(s, k, p[0], p[1], &US, &VS);
          V[2 * (iy * BS + ix)] = alpha * V[2 * (iy * BS + ix)] + (1 - alpha) * US;
          V[2 * (iy * BS + ix) + 1] = alpha * V[2 * (iy * BS + ix) + 1] + (1 - alpha) * VS;
        }
    }
}

/* PressureSingle::operator(), PressureSingle.cpp:346-373 */
void orc_pressure_single(orc_sim* s, double dt)
{
  checker(s);
  orc_pressure_rhs(s, dt);
  orc_poisson_solve(s);
  checker(s);
  orc_pressure_correction(s, dt);
  checker(s);
  orc_body_update(s, dt);
  checker(s);
  orc_penalize(s, dt);
}

/* ------------------------------------------------ PressureIterator_unif */
/* pressureCorrection (PressureIterator_unif.cpp:51-75): vFluid = tmpV + pFac grad p */
static void iter_pressure_correction(orc_sim* s, double dt)
{
  const double h = s->h, pFac = -0.5 * dt / h;
  #pragma omp parallel
  {
    double* P = (double*)malloc(sizeof(double) * LW * LW);
    #pragma omp for schedule(static)
    for (int b = 0; b < s->nb; ++b) {
      lab_load(s, s->f[ORC_PRES], 1, b, P);
      double* V = blk(s, ORC_VFLUID, b);
      const double* T = blk(s, ORC_TMPV, b);
      for (int iy = 0; iy < BS; ++iy)
        for (int ix = 0; ix < BS; ++ix) {
          const int c = 2 * (iy * BS + ix);
          V[c] = T[c] + pFac * (LAB(P, 1, ix + 1, iy, 0) - LAB(P, 1, ix - 1, iy, 0));
          V[c + 1] = T[c + 1] + pFac * (LAB(P, 1, ix, iy + 1, 0) - LAB(P, 1, ix, iy - 1, 0));
        }
    }
    free(P);
  }
}

/* integrateMomenta (PressureIterator_unif.cpp:77-126) with EXPL_INTEGRATE_MOM
 * (F = h^2 chi, velocity vFluid, penalDX = penalDY = 0), then
 * Shape::updateVelocity */
static void iter_integrate_momenta(orc_sim* s, shape_t* sh, double dt)
{
  const double hsq = s->h * s->h;
  orc_shape_state* t = &sh->st;
  const double Cx = t->centerOfMass[0], Cy = t->centerOfMass[1];
  double* part = (double*)calloc((size_t)s->nb * 7, sizeof(double));
  #pragma omp parallel for schedule(dynamic, 1)
  for (int b = 0; b < s->nb; ++b) {
    const oblock* o = sh->ob[b];
    if (!o) continue;
    const double* V = blk(s, ORC_VFLUID, b);
    double PM = 0, PJ = 0, PX = 0, PY = 0, UM = 0, VM = 0, AM = 0;
    for (int iy = 0; iy < BS; ++iy)
      for (int ix = 0; ix < BS; ++ix) {
        if (o->chi[iy][ix] <= 0) continue;
        const double F = hsq * o->chi[iy][ix];
        const double udiff[2] = {V[2 * (iy * BS + ix)], V[2 * (iy * BS + ix) + 1]};
        double p[2];
        bpos(s, b, ix, iy, p);
        p[0] -= Cx; p[1] -= Cy;
        PM += F;
        PJ += F * (p[0] * p[0] + p[1] * p[1]);
        PX += F * p[0]; PY += F * p[1];
        UM += F * udiff[0]; VM += F * udiff[1];
        AM += F * (p[0] * udiff[1] - p[1] * udiff[0]);
      }
    double* q = part + 7 * (size_t)b;
    q[0] = PM; q[1] = PJ; q[2] = PX; q[3] = PY; q[4] = UM; q[5] = VM; q[6] = AM;
  }
  double S[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int b = 0; b < s->nb; ++b)
    for (int i = 0; i < 7; ++i) S[i] += part[7 * (size_t)b + i];
  free(part);
  t->fluidAngMom = S[6]; t->fluidMomX = S[4]; t->fluidMomY = S[5];
  t->penalDX = 0; t->penalDY = 0; t->penalM = S[0]; t->penalJ = S[1];
  update_velocity(s, sh, dt);
}

/* penalize (PressureIterator_unif.cpp:128-184): vel = tmpV + chi (U_s - vFluid)
 * on the bodies' cells; returns the relative L2 change of the momentum */
static double iter_penalize(orc_sim* s, double dt)
{
  (void)dt;
  double* part = (double*)calloc((size_t)s->nb * 4, sizeof(double));
  #pragma omp parallel for schedule(dynamic, 1)
  for (int b = 0; b < s->nb; ++b) {
    double MX = 0, MY = 0, DMX = 0, DMY = 0;
    for (int k = 0; k < s->nshapes; ++k) {
      const shape_t* sh = &s->shapes[k];
      const oblock* o = sh->ob[b];
      if (!o) continue;
      const double u_s = sh->st.u, v_s = sh->st.v, omega_s = sh->st.omega;
      const double Cx = sh->st.centerOfMass[0], Cy = sh->st.centerOfMass[1];
      const double* CHI = blk(s, ORC_CHI, b);
      const double* T = blk(s, ORC_TMPV, b);
      const double* UF = blk(s, ORC_VFLUID, b);
      double* V = blk(s, ORC_VEL, b);
      for (int iy = 0; iy < BS; ++iy)
        for (int ix = 0; ix < BS; ++ix) {
          const double X = o->chi[iy][ix];
          if (CHI[iy * BS + ix] > X) continue;
          if (X <= 0) continue;
          double p[2];
          bpos(s, b, ix, iy, p);
          p[0] -= Cx; p[1] -= Cy;
          const double US = u_s - omega_s * p[1] + o->udef[iy][ix][0];
          const double VS = v_s + omega_s * p[0] + o->udef[iy][ix][1];
          const double penalFac = X;
          const int c = 2 * (iy * BS + ix);
          const double DFX = penalFac * (US - UF[c]);
          const double DFY = penalFac * (VS - UF[c + 1]);
          const double DPX = T[c] + DFX - V[c];
          const double DPY = T[c + 1] + DFY - V[c + 1];
          V[c] += DPX;
          V[c + 1] += DPY;
          MX += V[c] * V[c]; DMX += DPX * DPX;
          MY += V[c + 1] * V[c + 1]; DMY += DPY * DPY;
        }
    }
    double* q = part + 4 * (size_t)b;
    q[0] = MX; q[1] = MY; q[2] = DMX; q[3] = DMY;
  }
  double MX = 0, MY = 0, DMX = 0, DMY = 0;
  for (int b = 0; b < s->nb; ++b) {
    MX += part[4 * (size_t)b]; MY += part[4 * (size_t)b + 1];
    DMX += part[4 * (size_t)b + 2]; DMY += part[4 * (size_t)b + 3];
  }
  free(part);
  return sqrt((DMX + DMY) / (EPS + MX + MY));
}

/* updatePressureRHS (PressureIterator_unif.cpp:22-49):
 * r = (h/2dt) (div vel - CHI div uDef) */
static void iter_pressure_rhs(orc_sim* s, double dt)
{
  const double facDiv = 0.5 * s->h / dt;
  #pragma omp parallel
  {
    double* V = (double*)malloc(sizeof(double) * LW * LW * 2);
    double* U = (double*)malloc(sizeof(double) * LW * LW * 2);
    #pragma omp for schedule(static)
    for (int b = 0; b < s->nb; ++b) {
      lab_load(s, s->f[ORC_VEL], 2, b, V);
      lab_load(s, s->f[ORC_UDEF], 2, b, U);
      const double* CHI = blk(s, ORC_CHI, b);
      double* T = blk(s, ORC_TMP, b);
      for (int iy = 0; iy < BS; ++iy)
        for (int ix = 0; ix < BS; ++ix) {
          const double divVx = LAB(V, 2, ix + 1, iy, 0) - LAB(V, 2, ix - 1, iy, 0);
          const double divVy = LAB(V, 2, ix, iy + 1, 1) - LAB(V, 2, ix, iy - 1, 1);
          const double divUSx = LAB(U, 2, ix + 1, iy, 0) - LAB(U, 2, ix - 1, iy, 0);
          const double divUSy = LAB(U, 2, ix, iy + 1, 1) - LAB(U, 2, ix, iy - 1, 1);
          T[iy * BS + ix] = facDiv * (divVx + divVy - CHI[iy * BS + ix] * (divUSx + divUSy));
        }
    }
    free(V);
    free(U);
  }
}

int orc_pressure_iterator(orc_sim* s, double dt)
{
  const size_t n2 = (size_t)s->nb * BS * BS * 2;
  memcpy(s->f[ORC_TMPV], s->f[ORC_VEL], sizeof(double) * n2);
  int iter;
  double relDF = 1e3;
  for (iter = 0; iter < 1000; ++iter) {
    iter_pressure_correction(s, dt);
    for (int k = 0; k < s->nshapes; ++k) iter_integrate_momenta(s, &s->shapes[k], dt);
    relDF = iter_penalize(s, dt);
    iter_pressure_rhs(s, dt);
    orc_poisson_solve(s);
    const int conv = relDF < s->iter_target || iter > 100 || iter > 2 * s->iter_old_nsteps;
    if (conv) {
      /* finalizePressure (PressureIterator_unif.cpp:186-212) */
      orc_pressure_correction(s, dt);
      break;
    }
  }
  s->iter_old_nsteps = iter + 1;
  if (s->iter_old_nsteps > 30) s->iter_target = dmax(relDF, s->iter_target);
  if (s->iter_old_nsteps > 10) s->iter_target *= 1.01;
  if (s->iter_old_nsteps <= 2) s->iter_target *= 0.99;
  s->iter_target = dmin(1e-3, dmax(1e-5, s->iter_target));
  s->iter_last = iter + 1;
  s->iter_relDF = relDF;
  return iter + 1;
}

void orc_iter_stats(const orc_sim* s, int* iters, double* relDF, double* target)
{
  if (iters) *iters = s->iter_last;
  if (relDF) *relDF = s->iter_relDF;
  if (target) *target = s->iter_target;
}

/* ComputeForces::operator() (ComputeForces.cpp:16-111), Shape::computeForces
 * (Shape.cpp:252-281) */
void orc_compute_forces(orc_sim* s)
{
  const double NUoH = s->cfg.nu / s->h;
  for (int k = 0; k < s->nshapes; ++k) {
    shape_t* sh = &s->shapes[k];
    orc_shape_state* t = &sh->st;
    const double Cx = t->centerOfMass[0], Cy = t->centerOfMass[1];
    const double vel_norm = sqrt(t->u * t->u + t->v * t->v);
    const double vel_unit[2] = {vel_norm > 0 ? t->u / vel_norm : 0, vel_norm > 0 ? t->v / vel_norm : 0};
    #pragma omp parallel
    {
      double* V = (double*)malloc(sizeof(double) * LW * LW * 2);
      double* P = (double*)malloc(sizeof(double) * LW * LW);
      #pragma omp for schedule(static)
      for (int b = 0; b < s->nb; ++b) {
        oblock* O = sh->ob[b];
        if (!O) continue;
        lab_load(s, s->f[ORC_VEL], 2, b, V);
        lab_load(s, s->f[ORC_PRES], 1, b, P);
        O->perimeter = O->forcex = O->forcey = O->forcex_P = O->forcey_P = 0;
        O->forcex_V = O->forcey_V = O->torque = O->torque_P = O->torque_V = 0;
        O->drag = O->thrust = O->Pout = O->PoutBnd = O->defPower = O->defPowerBnd = 0;
        O->circulation = 0;
        for (int q = 0; q < O->nsurf; ++q) {
          const int ix = O->sx[q], iy = O->sy[q];
          double p[2];
          bpos(s, b, ix, iy, p);
          const double D11 = NUoH * (LAB(V, 2, ix + 1, iy, 0) - LAB(V, 2, ix - 1, iy, 0));
          const double D22 = NUoH * (LAB(V, 2, ix, iy + 1, 1) - LAB(V, 2, ix, iy - 1, 1));
          const double D12 = NUoH / 2 * (LAB(V, 2, ix, iy + 1, 0) - LAB(V, 2, ix, iy - 1, 0) +
                                         LAB(V, 2, ix + 1, iy, 1) - LAB(V, 2, ix - 1, iy, 1));
          const double normX = O->dchidx[q], normY = O->dchidy[q];
          const double Pc = LAB(P, 1, ix, iy, 0);
          const double fXV = D11 * normX + D12 * normY, fXP = -Pc * normX;
          const double fYV = D12 * normX + D22 * normY, fYP = -Pc * normY;
          const double fXT = fXV + fXP, fYT = fYV + fYP;
          const double vx = LAB(V, 2, ix, iy, 0), vy = LAB(V, 2, ix, iy, 1);
          double vxDef = O->udef[iy][ix][0], vyDef = O->udef[iy][ix][1];
          surf_add(s, (int)(sh - s->shapes), p[0] - Cx, p[1] - Cy, &vxDef, &vyDef);
          O->perimeter += sqrt(normX * normX + normY * normY);
          O->circulation += normX * vy - normY * vx;
          O->forcex += fXT; O->forcey += fYT;
          O->forcex_V += fXV; O->forcey_V += fYV;
          O->forcex_P += fXP; O->forcey_P += fYP;
          O->torque += (p[0] - Cx) * fYT - (p[1] - Cy) * fXT;
          O->torque_P += (p[0] - Cx) * fYP - (p[1] - Cy) * fXP;
          O->torque_V += (p[0] - Cx) * fYV - (p[1] - Cy) * fXV;
          const double forcePar = fXT * vel_unit[0] + fYT * vel_unit[1];
          O->thrust += .5 * (forcePar + fabs(forcePar));
          O->drag -= .5 * (forcePar - fabs(forcePar));
          const double powOut = fXT * vx + fYT * vy;
          const double powDef = fXT * vxDef + fYT * vyDef;
          O->Pout += powOut;
          O->defPower += powDef;
          O->PoutBnd += dmin(0, powOut);
          O->defPowerBnd += dmin(0, powDef);
        }
      }
      free(V); free(P);
    }
    /* Shape::computeForces */
    t->perimeter = t->forcex = t->forcey = t->forcex_P = t->forcey_P = 0;
    t->forcex_V = t->forcey_V = t->torque = t->torque_P = t->torque_V = 0;
    t->drag = t->thrust = t->Pout = t->PoutBnd = t->defPower = t->defPowerBnd = 0;
    t->circulation = 0;
    for (int b = 0; b < s->nb; ++b) {
      const oblock* O = sh->ob[b];
      if (!O) continue;
      t->circulation += O->circulation;
      t->perimeter += O->perimeter; t->torque += O->torque;
      t->forcex += O->forcex; t->forcey += O->forcey;
      t->forcex_P += O->forcex_P; t->forcey_P += O->forcey_P;
      t->forcex_V += O->forcex_V; t->forcey_V += O->forcey_V;
      t->torque_P += O->torque_P; t->torque_V += O->torque_V;
      t->drag += O->drag; t->thrust += O->thrust;
      t->Pout += O->Pout; t->defPowerBnd += O->defPowerBnd;
      t->PoutBnd += O->PoutBnd; t->defPower += O->defPower;
    }
    t->Pthrust = t->thrust * sqrt(t->u * t->u + t->v * t->v);
    t->Pdrag = t->drag * sqrt(t->u * t->u + t->v * t->v);
    const double denUnb = t->Pthrust - dmin(t->defPower, 0);
    const double demBnd = t->Pthrust - t->defPowerBnd;
    t->EffPDef = t->Pthrust / dmax(denUnb, EPS);
    t->EffPDefBnd = t->Pthrust / dmax(demBnd, EPS);
  }
}

/* Simulation::reset, Simulation.cpp:232-241 */
void orc_reset(orc_sim* s)
{
  for (int k = 0; k < s->nshapes; ++k) shape_reset(s, &s->shapes[k]);
  s->time = 0; s->step = 0; s->uinfx = 0; s->uinfy = 0;
  ic(s);
  orc_put_objects(s, 0);
  orc_apply_obj_vel(s);
  s->dt = 0;
}

/* Simulation::advance, Simulation.cpp:279-305 */
int orc_advance(orc_sim* s, double dt)
{
  s->dt = dt;
  const int err = orc_put_objects(s, dt);
  if (err) return err;
  orc_advdiff(s, dt);
  if (s->cfg.iterative_penal) orc_pressure_iterator(s, dt);
  else orc_pressure_single(s, dt);
  orc_compute_forces(s);
  s->time += s->dt;
  s->step++;
  checker(s); /* startProfiler("Dump") */
  return 0;
}

int orc_simulate(orc_sim* s, int nsteps)
{
  s->t_fft = 0;
  const double t0 = omp_get_wtime();
  for (int n = 0; n < nsteps; ++n) {
    const double dt = orc_calc_max_timestep(s);
    const int err = orc_advance(s, dt);
    if (err) { s->t_total = omp_get_wtime() - t0; return err; }
  }
  s->t_total = omp_get_wtime() - t0;
  return 0;
}
