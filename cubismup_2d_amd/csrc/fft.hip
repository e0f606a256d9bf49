// fft.hip -- FFT Poisson solvers (periodic and zero-padded free-space) on gfx950.
//
// Replaces FFTW_periodic::solve (FFTW_periodic.h:135-166), FFTW_freespace::solve
// (FFTW_freespace.cpp:14-56) and the reference's cuFFT path (CUDA_all.cu:41-128).
// Three kernels per solve, each reading and writing HBM exactly once:
//   1. rows forward: two real rows are packed as one complex row (a + i b),
//      one complex Stockham FFT in LDS, then split into both half spectra.
//      DC and Nyquist (both real) share one complex slot.  Output goes to a
//      column-group-major spectrum spec[cg][y][G] (G = 4 complex = 64 B), so
//      the two rows of a workgroup fill whole 128 B lines.
//   2. columns: one workgroup owns a column group (contiguous in memory),
//      forward FFT, spectral multiply (-1/k^2 or the Green's function
//      transform), inverse FFT, write back in place.  Free-space columns are
//      zero-pruned: only the NY data rows are read and written.
//   3. rows inverse: unpack two half-spectrum rows, inverse complex FFT, real
//      part -> row y, imaginary part -> row y+1.
// Cosine solver (FFTW_dirichlet, FFTW_dirichlet.h:23-148): a DCT-II of an
// N-point sequence x is, up to the phase exp(i pi k / 2N), the FFT of its
// 2N-point mirror image (x, reversed x), and a DCT-III of real coefficients
// multiplied by a symbol s(k) is the first N points of the inverse FFT of
// s(k) times that spectrum when s is extended evenly, s(2N-k) = s(k).  So the
// 2D cosine solve is the free-space pipeline with the mirrored grid in place
// of the zero padding (the row pass reads x reversed past NX, the column pass
// reads the spectrum rows reversed past NY) and the DCT-domain symbol.
// The FFT core is a radix-16 Stockham autosort (radix 2/4/8 for the last
// pass), 16 complex values per thread in registers, LDS padded by one slot
// every 16 entries so stride-16 stores are bank-conflict free.  Twiddles come
// from per-length tables exp(-2 pi i k/L) built in long double on the host.
#include <cmath>
#include <vector>

#include "device_util.hpp"

namespace cup2d {

constexpr int G = 4;  // complex columns per group (64 B in FP64)

struct FFTPlan {
  int Lr = 0, Lc = 0;     // row / column FFT lengths (complex)
  int NYs = 0;            // rows stored in the spectrum
  int ncg = 0;            // column groups
  int fs = 0;             // free-space
  int cs = 0;             // cosine (Neumann) solver on the mirrored 2NX x 2NY grid
  double2* spec = nullptr;
  double2* tw_r = nullptr;
  double2* tw_c = nullptr;
  double* ghat = nullptr;   // free-space: [ncg][NY+1][G]
  double* ghatN = nullptr;  // free-space Nyquist column [NY+1]
  double* cosX = nullptr;   // cosine: (float) cos(2 pi i / NX), i < NX
  double* cosY = nullptr;   // cosine: (float) cos(2 pi j / NY), j < NY
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// multiply by -i (forward) or +i (inverse)
template <bool FWD>
__device__ __forceinline__ double2 rot(double2 a) {
  return FWD ? make_double2(a.y, -a.x) : make_double2(-a.y, a.x);
}

template <bool FWD>
__device__ __forceinline__ void dft2(double2& a, double2& b) {
  const double2 t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <bool FWD>
__device__ __forceinline__ void dft4(double2& x0, double2& x1, double2& x2, double2& x3) {
  const double2 a0 = cadd(x0, x2), a1 = csub(x0, x2), a2 = cadd(x1, x3);
  const double2 a3 = rot<FWD>(csub(x1, x3));
  x0 = cadd(a0, a2);
  x1 = cadd(a1, a3);
  x2 = csub(a0, a2);
  x3 = csub(a1, a3);
}

// W_N^m = exp(-2 pi i m/N) (forward) constants
template <bool FWD>
__device__ __forceinline__ double2 w16(int m) {
  constexpr double c1 = 0.92387953251128673848, s1 = 0.38268343236508978178;
  constexpr double c2 = 0.70710678118654752440;
  double2 w;
  switch (m) {
    case 0: w = make_double2(1, 0); break;
    case 1: w = make_double2(c1, -s1); break;
    case 2: w = make_double2(c2, -c2); break;
    case 3: w = make_double2(s1, -c1); break;
    case 4: w = make_double2(0, -1); break;
    case 6: w = make_double2(-c2, -c2); break;
    case 9: w = make_double2(-c1, s1); break;
    default: w = make_double2(1, 0); break;
  }
  if (!FWD) w.y = -w.y;
  return w;
}

template <int R, bool FWD>
__device__ __forceinline__ void dft(double2* v) {
  if constexpr (R == 2) {
    dft2<FWD>(v[0], v[1]);
  } else if constexpr (R == 4) {
    dft4<FWD>(v[0], v[1], v[2], v[3]);
  } else if constexpr (R == 8) {
    double2 y[2][4];
#pragma unroll
    for (int n2 = 0; n2 < 2; ++n2) {
      y[n2][0] = v[n2]; y[n2][1] = v[n2 + 2]; y[n2][2] = v[n2 + 4]; y[n2][3] = v[n2 + 6];
      dft4<FWD>(y[n2][0], y[n2][1], y[n2][2], y[n2][3]);
    }
    // W8^k1 = W16^(2 k1)
#pragma unroll
    for (int k1 = 1; k1 < 4; ++k1) y[1][k1] = cmul(y[1][k1], w16<FWD>(2 * k1));
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      v[k1] = cadd(y[0][k1], y[1][k1]);
      v[k1 + 4] = csub(y[0][k1], y[1][k1]);
    }
  } else {  // 16
    double2 y[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
      y[n2][0] = v[n2]; y[n2][1] = v[n2 + 4]; y[n2][2] = v[n2 + 8]; y[n2][3] = v[n2 + 12];
      dft4<FWD>(y[n2][0], y[n2][1], y[n2][2], y[n2][3]);
    }
#pragma unroll
    for (int n2 = 1; n2 < 4; ++n2)
#pragma unroll
      for (int k1 = 1; k1 < 4; ++k1) y[n2][k1] = cmul(y[n2][k1], w16<FWD>(n2 * k1));
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      dft4<FWD>(y[0][k1], y[1][k1], y[2][k1], y[3][k1]);
      v[k1] = y[0][k1]; v[k1 + 4] = y[1][k1]; v[k1 + 8] = y[2][k1]; v[k1 + 12] = y[3][k1];
    }
  }
}

// Elements per thread of the FFT core (the first pass is radix E): 16 gives
// fewer LDS round trips than 8 (measured: 8 was slower on every length).
constexpr int FE = 16;

// LDS index with one pad slot every E entries (stride-E stores hit distinct banks)
template <int E>
__device__ __forceinline__ int padE(int i) { return i + i / E; }

// Passes after the first: span NS = E, E^2, ...; radix R = min(E, L/NS);
// PER = E/R butterflies per thread.  Butterfly j = t + p*NT of a pass needs
// the twiddles b^r, r = 1..R-1, with b = tw[(j % NS) * L/(NS R)]; a thread
// keeps only the bases b (one per butterfly and pass) in registers.
template <int L, int E, int NS>
struct TwCount {
  static constexpr int R = (L / NS) >= E ? E : L / NS;
  static constexpr int n = NS >= L ? 0 : (E / R) + TwCount<L, E, (NS >= L ? L : NS * R)>::n;
};
template <int L, int E>
struct TwCount<L, E, L> {
  static constexpr int n = 0;
};

// A thread's twiddle bases for every pass (forward sign; the inverse
// conjugates at use), loaded once per kernel.
template <int L, int E, int NS>
__device__ __forceinline__ void tw_load(const double2* __restrict__ tw, int t, double2* w) {
  if constexpr (NS < L) {
    constexpr int R = TwCount<L, E, NS>::R, PER = E / R, NT = L / E, TSTEP = L / (NS * R);
#pragma unroll
    for (int p = 0; p < PER; ++p) w[p] = tw[((t + p * NT) % NS) * TSTEP];
    tw_load<L, E, NS * R>(tw, t, w + PER);
  }
}

// pw[r] = b^r, r = 1..R-1, by repeated squaring and products of two lower
// powers (at most four roundings deep), conjugated for the inverse.
template <int R, bool FWD>
__device__ __forceinline__ void tw_powers(double2 b, double2* pw) {
  if (!FWD) b.y = -b.y;
  pw[1] = b;
#pragma unroll
  for (int r = 2; r < R; ++r) {
    const int hi = (r & 8) ? 8 : (r & 4) ? 4 : (r & 2) ? 2 : 1;
    pw[r] = (r == hi) ? cmul(pw[hi / 2], pw[hi / 2]) : cmul(pw[hi], pw[r - hi]);
  }
}

// One Stockham pass of radix R at span NS over an L-point sequence held in
// LDS at lds[base + padE(i)*stride]; NT = L/E threads, E/R butterflies each;
// w holds this pass's twiddle bases.
template <int L, int E, int R, int NS, bool FWD>
__device__ __forceinline__ void stockham_pass(double2* lds, int base, int stride, int t,
                                              const double2* w) {
  constexpr int NT = L / E, NBF = L / R, PER = E / R;
  double2 v[PER][R];
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int j = t + p * NT;
#pragma unroll
    for (int r = 0; r < R; ++r) v[p][r] = lds[base + padE<E>(j + r * NBF) * stride];
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int j = t + p * NT;
    const int k = j % NS;
    double2 pw[R];
    tw_powers<R, FWD>(w[p], pw);
#pragma unroll
    for (int r = 1; r < R; ++r) v[p][r] = cmul(v[p][r], pw[r]);
    dft<R, FWD>(v[p]);
    const int idxD = (j / NS) * NS * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) lds[base + padE<E>(idxD + r * NS) * stride] = v[p][r];
  }
  __syncthreads();
}

// First pass (radix E, span 1) straight from registers: x[q] holds element
// t + q*NT, which is the Stockham input of butterfly j = t.  No twiddles.
template <int L, int E, bool FWD>
__device__ __forceinline__ void pass_first(double2* x, double2* lds, int base, int stride, int t) {
  dft<E, FWD>(x);
#pragma unroll
  for (int r = 0; r < E; ++r) lds[base + padE<E>(t * E + r) * stride] = x[r];
  __syncthreads();
}

// Last pass (span NS = L/R) into registers: its outputs j + r*NS land on
// element t + q*NT with q = p + (E/R)*r, i.e. the same strided order the
// first pass reads.  Ends with a barrier so the LDS can be reused.
template <int L, int E, int R, int NS, bool FWD>
__device__ __forceinline__ void pass_last(double2* lds, int base, int stride, int t,
                                          const double2* w, double2* x) {
  constexpr int NT = L / E, NBF = L / R, PER = E / R;
  static_assert(NS * R == L, "last pass");
  double2 v[PER][R];
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int j = t + p * NT;
#pragma unroll
    for (int r = 0; r < R; ++r) v[p][r] = lds[base + padE<E>(j + r * NBF) * stride];
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    double2 pw[R];
    tw_powers<R, FWD>(w[p], pw);
#pragma unroll
    for (int r = 1; r < R; ++r) v[p][r] = cmul(v[p][r], pw[r]);
    dft<R, FWD>(v[p]);
#pragma unroll
    for (int r = 0; r < R; ++r) x[p + PER * r] = v[p][r];
  }
}

template <int L, int E, int NS, bool FWD>
__device__ __forceinline__ void fft_tail(double2* lds, int base, int stride, int t,
                                         const double2* w, double2* x) {
  constexpr int REM = L / NS;
  constexpr int R = REM >= E ? E : REM;
  if constexpr (NS * R == L) {
    pass_last<L, E, R, NS, FWD>(lds, base, stride, t, w, x);
  } else {
    stockham_pass<L, E, R, NS, FWD>(lds, base, stride, t, w);
    fft_tail<L, E, NS * R, FWD>(lds, base, stride, t, w + E / R, x);
  }
}

// L-point complex FFT (L >= 2E, power of two) by NT = L/E threads; x[q] is
// element t + q*NT on entry and the transform's element t + q*NT on exit.
// The sequence is staged in LDS at lds[base + padE(i)*stride] between
// passes; w = this thread's twiddle bases from tw_load<L, E, E>.
template <int L, int E, bool FWD>
__device__ __forceinline__ void fft_regs(double2* x, double2* lds, int base, int stride, int t,
                                         const double2* w) {
  pass_first<L, E, FWD>(x, lds, base, stride, t);
  fft_tail<L, E, E, FWD>(lds, base, stride, t, w, x);
}

template <int L, int E>
constexpr int lds_len() { return L + L / E; }
template <int L, int E>
constexpr int rows_np() { return (L / E) >= 256 ? 1 : 256 / (L / E); }

// ---------------------------------------------------------- rows forward
// in: nrows rows of ncols valid reals (row pitch in_pitch), zero beyond;
// out: spec[cg][y][G], y < spec_rows.  NP row pairs per workgroup.
//
// SRC_FADE: the input is the pressure RHS on the grid; cells of band blocks
// are faded as PoissonSolver::cub2rhs does (PoissonSolver.cpp:105-121) while
// they are loaded, and written back (the reference modifies the source in
// place).  The correction factor is st->fade_corr (k_fade_reduce).
//
// SRC_DIV: the RHS is formed here from the advDiff output (u, v), as
// PressureSingle::updatePressureRHS does (PressureSingle.cpp:112-128), with
// the deferred advDiff boundary-flux correction st->flux_corr (k_flux_reduce,
// right after advDiff); the obstacle blocks take
// the values the obstacle RHS kernels left in the RHS plane; then the fade
// as for SRC_FADE.  u of both rows of a pair goes through LDS for the x
// neighbours; v of rows y0-1 .. y0+2 stays in registers.
enum { SRC_PLAIN = 0, SRC_FADE = 1, SRC_DIV = 2 };
struct RowSource {
  Grid g;
  int npart;            // fade-sum workgroups (0: no fade)
  double* rhs;          // RHS plane: faded band cells written back; obstacle blocks read (SRC_DIV)
  const StepState* st;  // fade_corr, flux_corr, dt
  const double* u;      // SRC_DIV: advDiff output
  const double* v;
  const unsigned char* obmap;
  int mirror;           // cosine: columns past ncols read the row reversed
};

template <int L, int NP, int E, int SRC>
__global__ __launch_bounds__(NP * L / E) void k_rows_fwd(
    const double* in, int in_pitch, int ncols, int nrows, double2* __restrict__ spec,
    int spec_rows, const double2* __restrict__ tw, RowSource fd) {
  constexpr int NT = L / E, PADL = lds_len<L, E>();
  extern __shared__ double2 lds[];
  const int f = threadIdx.x / NT, t = threadIdx.x % NT;
  // SRC_DIV reads the v rows next to its pair too: consecutive pairs run on
  // one XCD (workgroup b runs on XCD b % 8) so those rows are L2 hits
  int wg = blockIdx.x;
  if (SRC == SRC_DIV && gridDim.x % 8 == 0) wg = (wg % 8) * (gridDim.x / 8) + wg / 8;
  const int y0 = 2 * (wg * NP + f);
  const bool valid = y0 < nrows;
  const int base = f * PADL;
  double2 w[TwCount<L, E, E>::n];
  tw_load<L, E, E>(tw, t, w);
  double2 x[E];
  if constexpr (SRC == SRC_DIV) {
    const Grid& g = fd.g;
    const int NX = ncols, NY = nrows;
    const int yc = valid ? y0 : 0;
    const int ym = yc > 0 ? yc - 1 : yc, yp = yc + 2 < NY ? yc + 2 : yc + 1;
    double* ud = (double*)lds + (size_t)f * 2 * PADL;  // u of rows y0, y0+1
    // every load of the six rows is issued before the first use, so a wave
    // has all of them in flight (storing u into LDS inside this loop would
    // wait for each group in turn)
    double u0[E], u1[E], vm[E], v0[E], v1[E], vp[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int i = t + q * NT;
      if (i < NX) {
        u0[q] = fd.u[(size_t)yc * NX + i];
        u1[q] = fd.u[(size_t)(yc + 1) * NX + i];
        vm[q] = fd.v[(size_t)ym * NX + i];
        v0[q] = fd.v[(size_t)yc * NX + i];
        v1[q] = fd.v[(size_t)(yc + 1) * NX + i];
        vp[q] = fd.v[(size_t)yp * NX + i];
      }
    }
    const double fcorr = fd.st->flux_corr;  // k_flux_reduce, after advDiff
    const double facDiv = 0.5 * g.h / fd.st->dt;
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int i = t + q * NT;
      if (i < NX) {
        ud[i] = u0[q];
        ud[L + i] = u1[q];
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int i = t + q * NT;
      double r0 = 0, r1 = 0;
      if (valid && i < NX) {
        const int xe = i + 1 < NX ? i + 1 : i, xw = i > 0 ? i - 1 : i;
        const double ue0 = flux_fixed(ud[xe], xe, NX, fcorr);
        const double uw0 = flux_fixed(ud[xw], xw, NX, fcorr);
        const double ue1 = flux_fixed(ud[L + xe], xe, NX, fcorr);
        const double uw1 = flux_fixed(ud[L + xw], xw, NX, fcorr);
        const double vs0 = flux_fixed(vm[q], ym, NY, fcorr);
        const double vc0 = flux_fixed(v0[q], y0, NY, fcorr);
        const double vc1 = flux_fixed(v1[q], y0 + 1, NY, fcorr);
        const double vn1 = flux_fixed(vp[q], yp, NY, fcorr);
        r0 = facDiv * ((ue0 - uw0) + (vc1 - vs0));
        r1 = facDiv * ((ue1 - uw1) + (vn1 - vc0));
        const int bx = i / BS;
        if (fd.obmap[(y0 / BS) * g.bpdx + bx]) r0 = fd.rhs[(size_t)y0 * NX + i];
        if (fd.obmap[((y0 + 1) / BS) * g.bpdx + bx]) r1 = fd.rhs[(size_t)(y0 + 1) * NX + i];
      }
      x[q] = make_double2(r0, r1);
    }
    __syncthreads();  // ud is reused below
  } else {
    const double* r0 = in + (size_t)(valid ? y0 : 0) * in_pitch;
    const double* r1 = r0 + in_pitch;
#pragma unroll
    for (int q = 0; q < E; ++q) {
      int i = t + q * NT;
      if (fd.mirror && i >= ncols) i = 2 * ncols - 1 - i;
      double a = 0, b = 0;
      if (valid && i < ncols) { a = r0[i]; b = r1[i]; }
      x[q] = make_double2(a, b);
    }
  }
  if constexpr (SRC != SRC_PLAIN) {
    if (fd.npart > 0) {
      const double corr = fd.st->fade_corr;  // k_fade_reduce
      const Grid& g = fd.g;
      const int by0 = y0 / BS, by1 = (y0 + 1) / BS;
      const bool row0 = band_row(g, by0), row1 = band_row(g, by1);
      if (valid) {
#pragma unroll
        for (int q = 0; q < E; ++q) {
          const int iq = t + q * NT;
          const int i = fd.mirror && iq >= ncols ? 2 * ncols - 1 - iq : iq;
          if (i >= ncols) continue;
          const bool own = i == iq;  // the mirrored copy is written back once
          const int bx = i / BS;
          const bool col = band_col(g, bx);
          if (!(row0 || row1 || col)) continue;
          const double px = cell_pos(bx, i % BS, g.hb, g.h);
          if (row0 || col) {
            const double py = cell_pos(by0, y0 % BS, g.hb, g.h);
            x[q].x = fade_cell(x[q].x, fade_arg(g, px, py), corr);
            if (own) fd.rhs[(size_t)y0 * in_pitch + i] = x[q].x;
          }
          if (row1 || col) {
            const double py = cell_pos(by1, (y0 + 1) % BS, g.hb, g.h);
            x[q].y = fade_cell(x[q].y, fade_arg(g, px, py), corr);
            if (own) fd.rhs[(size_t)(y0 + 1) * in_pitch + i] = x[q].y;
          }
        }
      }
    }
  }
  fft_regs<L, E, true>(x, lds, base, 1, t, w);
  // split z = a + i b: A_k = (Z_k + conj Z_{L-k})/2, B_k = (Z_k - conj Z_{L-k})/(2i);
  // the partners Z_{L-k} (upper half) go through LDS
#pragma unroll
  for (int q = E / 2; q < E; ++q) lds[base + padE<E>(t + q * NT)] = x[q];
  __syncthreads();
  if (!valid) return;
#pragma unroll
  for (int q = 0; q < E / 2; ++q) {
    const int k = t + q * NT;  // 0 .. L/2-1
    const double2 Z = x[q];
    double2 A, B;
    if (k == 0) {
      const double2 ZN = x[E / 2];  // element L/2
      A = make_double2(Z.x, ZN.x);
      B = make_double2(Z.y, ZN.y);
    } else {
      const double2 Zm = lds[base + padE<E>(L - k)];
      const double2 sm = make_double2(Z.x + Zm.x, Z.y - Zm.y);  // Z + conj(Zm)
      const double2 d = make_double2(Z.x - Zm.x, Z.y + Zm.y);   // Z - conj(Zm)
      A = make_double2(0.5 * sm.x, 0.5 * sm.y);
      B = make_double2(0.5 * d.y, -0.5 * d.x);
    }
    const int cg = k / G, c = k % G;
    double2* o = spec + ((size_t)cg * spec_rows + y0) * G + c;
    o[0] = A;
    o[G] = B;
  }
}

// ---------------------------------------------------------- rows inverse
template <int L, int NP, int E>
__global__ __launch_bounds__(NP * L / E) void k_rows_inv(const double2* __restrict__ spec,
                                                         int spec_rows, double* __restrict__ out,
                                                         int out_pitch, int ncols, int nrows,
                                                         const double2* __restrict__ tw) {
  constexpr int NT = L / E, PADL = lds_len<L, E>();
  extern __shared__ double2 lds[];
  const int f = threadIdx.x / NT, t = threadIdx.x % NT;
  const int y0 = 2 * (blockIdx.x * NP + f);
  const bool valid = y0 < nrows;
  const int base = f * PADL;
  double2 w[TwCount<L, E, E>::n];
  tw_load<L, E, E>(tw, t, w);
  double2 x[E];
  // Z_k = A_k + i B_k and Z_{L-k} = conj(A_k) + i conj(B_k); DC and Nyquist
  // of both rows are packed in slot 0 (real parts only).  All loads are
  // issued before the first LDS store, so they are in flight together.
  double2 As[E / 2], Bs[E / 2];
#pragma unroll
  for (int q = 0; q < E / 2; ++q) {
    const int k = t + q * NT;
    As[q] = make_double2(0, 0);
    Bs[q] = make_double2(0, 0);
    if (valid) {
      const double2* s = spec + ((size_t)(k / G) * spec_rows + y0) * G + (k % G);
      As[q] = s[0];
      Bs[q] = s[G];
    }
  }
#pragma unroll
  for (int q = 0; q < E / 2; ++q) {
    const int k = t + q * NT;
    const double2 A = As[q], B = Bs[q];
    if (k == 0) {
      x[0] = make_double2(A.x, B.x);
      x[E / 2] = make_double2(A.y, B.y);
    } else {
      x[q] = make_double2(A.x - B.y, A.y + B.x);
      lds[base + padE<E>(L - k)] = make_double2(A.x + B.y, -A.y + B.x);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = E / 2; q < E; ++q) {
    const int i = t + q * NT;
    if (i != L / 2) x[q] = lds[base + padE<E>(i)];
  }
  __syncthreads();
  fft_regs<L, E, false>(x, lds, base, 1, t, w);
  if (!valid) return;
  double* o0 = out + (size_t)y0 * out_pitch;
  double* o1 = o0 + out_pitch;
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int i = t + q * NT;
    if (i < ncols) {
      o0[i] = x[q].x;
      o1[i] = x[q].y;
    }
  }
}

// --------------------------------------------------------------- columns
// MODE 0: periodic solve, 1: free-space solve (zero-pruned), 2: forward only.
struct ColParams {
  int NX, NY;            // physical grid
  int rows;              // data rows stored per column (NY, or Lc for MODE 2)
  double waveFactX, waveFactY, norm;
  const double* ghat;    // MODE 1
  const double* ghatN;
  const double* cosX;    // MODE 3
  const double* cosY;
};

__device__ __forceinline__ double periodic_factor(int kx, int ky, const ColParams& P) {
  // FFTW_periodic.h:82-91: kinv*norm_factor, DC set to zero
  const int kyy = ky <= P.NY / 2 ? ky : P.NY - ky;
  if (kx == 0 && kyy == 0) return 0.0;
  const double rkx = kx * P.waveFactX, rky = kyy * P.waveFactY;
  const double kinv = -1 / (rkx * rkx + rky * rky);
  return kinv * P.norm;
}

// FFTW_dirichlet::_solve (FFTW_dirichlet.h:45-61) at DCT indices (i, j) <
// (NX, NY).  Line 51 declares the blend weight as `static constexpr tol =
// 0.1;`, which has no type; it is taken as the double 0.1 the initializer
// states.  COScoef are floats (FFTW_dirichlet.h:26-27,65-68).
__device__ __forceinline__ double cosine_factor(int i, int j, const ColParams& P) {
  if (i == 0 && j == 0) return 0.0;  // in_out[0] = 0
  const double tol = 0.1;
  const double rkx = (i + 0.5) * P.waveFactX, rky = (j + 0.5) * P.waveFactY;
  const double denomFD = 1 - P.cosX[i] / 2 - P.cosY[j] / 2;
  const double denomSP = rkx * rkx + rky * rky;
  return -P.norm / ((1 - tol) * denomFD + tol * denomSP);
}

template <int LC, int MODE>
__device__ __forceinline__ double col_symbol(int kx, int cg, int c, int ky, const ColParams& P) {
  if constexpr (MODE == 0) {
    return periodic_factor(kx, ky, P);
  } else if constexpr (MODE == 3) {
    // symbol extended evenly about NX and NY; zero at the mirror Nyquist
    if (kx >= P.NX || ky == P.NY) return 0.0;
    return cosine_factor(kx, ky < P.NY ? ky : LC - ky, P);
  } else {
    const int kyy = ky <= P.NY ? ky : LC - ky;
    return P.ghat[((size_t)cg * (P.NY + 1) + kyy) * G + c];
  }
}

// Columns per workgroup (CW) of a column group of G: CW = G/2 halves the LDS
// of a whole group so two workgroups share a CU.  The two halves of a group
// run on the same XCD (block b runs on XCD b % 8) so the second half of each
// 64 B chunk is an L2 hit.
constexpr int CW = 2;
static_assert(G % CW == 0, "CW");

__device__ __forceinline__ void col_group_of_block(int b, int nblocks, int* cg, int* half) {
  constexpr int H = G / CW;
  if (H == 1 || nblocks % (8 * H) != 0) {
    *cg = b / H;
    *half = b % H;
    return;
  }
  const int xcd = b % 8, local = b / 8;        // consecutive locals share an XCD
  const int per_xcd = nblocks / 8;             // blocks per XCD
  const int lin = xcd * per_xcd + local;       // XCD-contiguous numbering
  *cg = lin / H;
  *half = lin % H;
}

template <int LC, int MODE, int E>
__global__ __launch_bounds__(CW * LC / E) void k_cols(double2* __restrict__ spec, ColParams P,
                                                      const double2* __restrict__ tw) {
  constexpr int NT = LC / E;
  extern __shared__ double2 lds[];
  const int cl = threadIdx.x % CW, t = threadIdx.x / CW;
  int cg, half;
  col_group_of_block(blockIdx.x, gridDim.x, &cg, &half);
  const int c = half * CW + cl;  // column inside the group
  double2* col = spec + (size_t)cg * P.rows * G;
  double2 w[TwCount<LC, E, E>::n];
  tw_load<LC, E, E>(tw, t, w);
  double2 x[E];
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int i = t + q * NT;
    if constexpr (MODE == 3)  // mirrored grid: rows past NY are rows NY-1, NY-2, ...
      x[q] = col[(size_t)(i < P.rows ? i : 2 * P.rows - 1 - i) * G + c];
    else
      x[q] = (MODE != 1 || i < P.rows) ? col[(size_t)i * G + c] : make_double2(0, 0);
  }
  fft_regs<LC, E, true>(x, lds, cl, CW, t, w);
  if constexpr (MODE == 2) {
#pragma unroll
    for (int q = 0; q < E; ++q) col[(size_t)(t + q * NT) * G + c] = x[q];
    return;
  } else {
    const int kx = cg * G + c;
    if (cg == 0 && half == 0) {
      // column 0 packs the DC (kx=0) and Nyquist (kx=NX/2) columns as
      // c = dc + i nq; with real, even symbols s0, sN:
      // C'(k) = C(k) (s0+sN)/2 + conj(C(-k)) (s0-sN)/2
#pragma unroll
      for (int q = 0; q < E; ++q) lds[padE<E>(t + q * NT) * CW + cl] = x[q];
      __syncthreads();
#pragma unroll
      for (int q = 0; q < E; ++q) {
        const int i = t + q * NT;
        const double s0 = col_symbol<LC, MODE>(kx, cg, c, i, P);
        if (kx == 0) {
          double sN;
          if constexpr (MODE == 0) sN = periodic_factor(P.NX / 2, i, P);
          else if constexpr (MODE == 3) sN = 0.0;  // mirror Nyquist kx = NX
          else sN = P.ghatN[i <= P.NY ? i : LC - i];
          const double2 m = lds[padE<E>((LC - i) % LC) * CW + cl];
          const double a = 0.5 * (s0 + sN), b = 0.5 * (s0 - sN);
          x[q] = make_double2(x[q].x * a + m.x * b, x[q].y * a - m.y * b);
        } else {
          x[q] = make_double2(x[q].x * s0, x[q].y * s0);
        }
      }
      __syncthreads();
    } else {
#pragma unroll
      for (int q = 0; q < E; ++q) {
        const double sy = col_symbol<LC, MODE>(kx, cg, c, t + q * NT, P);
        x[q] = make_double2(x[q].x * sy, x[q].y * sy);
      }
    }
    fft_regs<LC, E, false>(x, lds, cl, CW, t, w);
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int i = t + q * NT;
      if ((MODE != 1 && MODE != 3) || i < P.rows) col[(size_t)i * G + c] = x[q];
    }
  }
}

// Green's function spectrum: Re(FFT G) * norm, even-folded to ky <= NY.
// spec holds the full forward transform of the (2NX)x(2NY) kernel.
__global__ void k_ghat_extract(const double2* __restrict__ spec, int Lc, int ncg, int NY,
                               double norm, double* __restrict__ ghat,
                               double* __restrict__ ghatN) {
  const int n = ncg * G * (NY + 1);
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const int c = q % G, ky = (q / G) % (NY + 1), cg = q / (G * (NY + 1));
    const double2 C = spec[((size_t)cg * Lc + ky) * G + c];
    if (cg == 0 && c == 0) {
      const double2 Cm = spec[((size_t)0 * Lc + (Lc - ky) % Lc) * G + 0];
      // DC = (C + conj Cm)/2, Nyq = (C - conj Cm)/(2i): real parts
      ghat[q] = 0.5 * (C.x + Cm.x) * norm;
      ghatN[ky] = 0.5 * (C.y + Cm.y) * norm;
    } else {
      ghat[q] = C.x * norm;
    }
  }
}

// ----------------------------------------------------------- dispatch
template <int L>
static void launch_rows_fwd(cup2d_ctx* c, const double* in, int pitch, int ncols, int nrows,
                            double2* spec, int spec_rows, const double2* tw, int src,
                            const RowSource& fd) {
  constexpr int NP = rows_np<L, FE>();
  const size_t shm = (size_t)NP * lds_len<L, FE>() * sizeof(double2);
  const dim3 grid((nrows / 2 + NP - 1) / NP), block(NP * L / FE);
  if (src == SRC_DIV)
    CUP2D_TIMED_LAUNCH(c, (k_rows_fwd<L, NP, FE, SRC_DIV>), grid, block, shm, c->stream, in, pitch,
                       ncols, nrows, spec, spec_rows, tw, fd);
  else if (src == SRC_FADE)
    CUP2D_TIMED_LAUNCH(c, (k_rows_fwd<L, NP, FE, SRC_FADE>), grid, block, shm, c->stream, in,
                       pitch, ncols, nrows, spec, spec_rows, tw, fd);
  else
    CUP2D_TIMED_LAUNCH(c, (k_rows_fwd<L, NP, FE, SRC_PLAIN>), grid, block, shm, c->stream, in,
                       pitch, ncols, nrows, spec, spec_rows, tw, fd);
}
template <int L>
static void launch_rows_inv(cup2d_ctx* c, const double2* spec, int spec_rows, double* out,
                            int pitch, int ncols, int nrows, const double2* tw) {
  constexpr int NP = rows_np<L, FE>();
  const size_t shm = (size_t)NP * lds_len<L, FE>() * sizeof(double2);
  const int pairs = nrows / 2;
  CUP2D_TIMED_LAUNCH(c, (k_rows_inv<L, NP, FE>), dim3((pairs + NP - 1) / NP), dim3(NP * L / FE),
                     shm, c->stream, spec, spec_rows, out, pitch, ncols, nrows, tw);
}
template <int LC, int MODE>
static void launch_cols(cup2d_ctx* c, double2* spec, int ncg, const ColParams& P,
                        const double2* tw) {
  const size_t shm = (size_t)lds_len<LC, FE>() * CW * sizeof(double2);
  CUP2D_TIMED_LAUNCH(c, (k_cols<LC, MODE, FE>), dim3(ncg * (G / CW)), dim3(CW * LC / FE), shm,
                     c->stream, spec, P, tw);
}

#define CUP2D_POW2_CASES(M, ...)                                                 \
  switch (L) {                                                                   \
    case 32: M<32>(__VA_ARGS__); break;                                          \
    case 64: M<64>(__VA_ARGS__); break;                                          \
    case 128: M<128>(__VA_ARGS__); break;                                        \
    case 256: M<256>(__VA_ARGS__); break;                                        \
    case 512: M<512>(__VA_ARGS__); break;                                        \
    case 1024: M<1024>(__VA_ARGS__); break;                                      \
    case 2048: M<2048>(__VA_ARGS__); break;                                      \
    case 4096: M<4096>(__VA_ARGS__); break;                                      \
    case 8192: M<8192>(__VA_ARGS__); break;                                      \
    default: return CUP2D_ERR_ARG;                                               \
  }

static int rows_fwd(cup2d_ctx* c, int L, const double* in, int pitch, int ncols, int nrows,
                    double2* spec, int spec_rows, const double2* tw, int src,
                    const RowSource& fd) {
  CUP2D_POW2_CASES(launch_rows_fwd, c, in, pitch, ncols, nrows, spec, spec_rows, tw, src, fd);
  return CUP2D_OK;
}
static int rows_inv(cup2d_ctx* c, int L, const double2* spec, int spec_rows, double* out,
                    int pitch, int ncols, int nrows, const double2* tw) {
  CUP2D_POW2_CASES(launch_rows_inv, c, spec, spec_rows, out, pitch, ncols, nrows, tw);
  return CUP2D_OK;
}
template <int MODE>
static int cols(cup2d_ctx* c, int LC, double2* spec, int ncg, const ColParams& P,
                const double2* tw) {
  switch (LC) {
    case 32: launch_cols<32, MODE>(c, spec, ncg, P, tw); break;
    case 64: launch_cols<64, MODE>(c, spec, ncg, P, tw); break;
    case 128: launch_cols<128, MODE>(c, spec, ncg, P, tw); break;
    case 256: launch_cols<256, MODE>(c, spec, ncg, P, tw); break;
    case 512: launch_cols<512, MODE>(c, spec, ncg, P, tw); break;
    case 1024: launch_cols<1024, MODE>(c, spec, ncg, P, tw); break;
    case 2048: launch_cols<2048, MODE>(c, spec, ncg, P, tw); break;
    case 4096: launch_cols<4096, MODE>(c, spec, ncg, P, tw); break;
    default: return CUP2D_ERR_ARG;
  }
  return CUP2D_OK;
}

template <int L>
static void set_rows_attr_L() {
  constexpr int NP = rows_np<L, FE>();
  const int shm = NP * lds_len<L, FE>() * (int)sizeof(double2);
  (void)hipFuncSetAttribute((const void*)k_rows_fwd<L, NP, FE, SRC_PLAIN>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, shm);
  (void)hipFuncSetAttribute((const void*)k_rows_fwd<L, NP, FE, SRC_FADE>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, shm);
  (void)hipFuncSetAttribute((const void*)k_rows_fwd<L, NP, FE, SRC_DIV>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, shm);
  (void)hipFuncSetAttribute((const void*)k_rows_inv<L, NP, FE>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, shm);
}
template <int LC>
static void set_cols_attr() {
  const int shm = lds_len<LC, FE>() * CW * (int)sizeof(double2);
  (void)hipFuncSetAttribute((const void*)k_cols<LC, 0, FE>, hipFuncAttributeMaxDynamicSharedMemorySize, shm);
  (void)hipFuncSetAttribute((const void*)k_cols<LC, 1, FE>, hipFuncAttributeMaxDynamicSharedMemorySize, shm);
  (void)hipFuncSetAttribute((const void*)k_cols<LC, 2, FE>, hipFuncAttributeMaxDynamicSharedMemorySize, shm);
  (void)hipFuncSetAttribute((const void*)k_cols<LC, 3, FE>, hipFuncAttributeMaxDynamicSharedMemorySize, shm);
}

static std::vector<double2> twiddles(int L) {
  std::vector<double2> w(L);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int k = 0; k < L; ++k) {
    const long double a = -2.0L * pi * (long double)k / (long double)L;
    w[k] = make_double2((double)cosl(a), (double)sinl(a));
  }
  return w;
}

static bool pow2(int n) { return n >= 32 && (n & (n - 1)) == 0; }

int fft_create(cup2d_ctx* c) {
  const int NX = c->g.NX, NY = c->g.NY;
  const bool fs = c->cfg.poisson == CUP2D_POISSON_FREESPACE;
  const bool cs = c->cfg.poisson == CUP2D_POISSON_COSINE;
  FFTPlan* P = new FFTPlan();
  c->fft = P;
  P->fs = fs;
  P->cs = cs;
  P->Lr = fs || cs ? 2 * NX : NX;
  P->Lc = fs || cs ? 2 * NY : NY;
  P->NYs = NY;
  P->ncg = (P->Lr / 2) / G;
  if (!pow2(NX) || !pow2(NY) || P->Lr > 8192 || P->Lc > 4096) {
    c->err = "FFT Poisson: NX and NY must be powers of two, row length <= 8192 and column "
             "length <= 4096 (periodic NX <= 8192, NY <= 4096; free-space and cosine "
             "NX <= 4096, NY <= 2048)";
    return CUP2D_ERR_ARG;
  }
  set_rows_attr_L<32>(); set_rows_attr_L<64>(); set_rows_attr_L<128>(); set_rows_attr_L<256>();
  set_rows_attr_L<512>(); set_rows_attr_L<1024>(); set_rows_attr_L<2048>();
  set_rows_attr_L<4096>(); set_rows_attr_L<8192>();
  set_cols_attr<32>(); set_cols_attr<64>(); set_cols_attr<128>(); set_cols_attr<256>();
  set_cols_attr<512>(); set_cols_attr<1024>(); set_cols_attr<2048>(); set_cols_attr<4096>();

  CUP2D_HIP_CHECK(c, hipMalloc(&P->spec, sizeof(double2) * (size_t)P->ncg * G * P->NYs));
  const std::vector<double2> twr = twiddles(P->Lr), twc = twiddles(P->Lc);
  CUP2D_HIP_CHECK(c, hipMalloc(&P->tw_r, sizeof(double2) * twr.size()));
  CUP2D_HIP_CHECK(c, hipMalloc(&P->tw_c, sizeof(double2) * twc.size()));
  CUP2D_HIP_CHECK(c, hipMemcpy(P->tw_r, twr.data(), sizeof(double2) * twr.size(), hipMemcpyHostToDevice));
  CUP2D_HIP_CHECK(c, hipMemcpy(P->tw_c, twc.data(), sizeof(double2) * twc.size(), hipMemcpyHostToDevice));
  if (cs) {  // FFTW_dirichlet.h:65-68: COScoef[i] = (float) cos(pi/M * 2 i)
    std::vector<double> cx(NX), cy(NY);
    for (int i = 0; i < NX; ++i) cx[i] = (double)(float)std::cos(M_PI / NX * 2.0 * i);
    for (int j = 0; j < NY; ++j) cy[j] = (double)(float)std::cos(M_PI / NY * 2.0 * j);
    CUP2D_HIP_CHECK(c, hipMalloc(&P->cosX, sizeof(double) * NX));
    CUP2D_HIP_CHECK(c, hipMalloc(&P->cosY, sizeof(double) * NY));
    CUP2D_HIP_CHECK(c, hipMemcpy(P->cosX, cx.data(), sizeof(double) * NX, hipMemcpyHostToDevice));
    CUP2D_HIP_CHECK(c, hipMemcpy(P->cosY, cy.data(), sizeof(double) * NY, hipMemcpyHostToDevice));
  }
  if (!fs) return CUP2D_OK;

  // Free-space Green's function on the 2NX x 2NY cyclic grid
  // (FFTW_freespace.cpp:100-137; with 2N padding the first NX x NY outputs of
  // the cyclic convolution equal the reference's (2N-1) result).
  const int MX = 2 * NX, MY = 2 * NY;
  std::vector<double> Gk((size_t)MX * MY);
  const double h = c->g.h, fac = 1 / (2 * M_PI);
  for (int j = 0; j < MY; ++j)
    for (int i = 0; i < MX; ++i) {
      const double yi = j >= NY ? MY - j : j;
      const double xi = i >= NX ? MX - i : i;
      const double r = std::sqrt(xi * xi + yi * yi);
      Gk[(size_t)j * MX + i] = fac * std::log(h * r);
    }
  Gk[0] = fac / 2 * (2 * std::log(h / std::sqrt(M_PI)) - 1);
  double* dG = nullptr;
  double2* tspec = nullptr;
  CUP2D_HIP_CHECK(c, hipMalloc(&dG, sizeof(double) * Gk.size()));
  CUP2D_HIP_CHECK(c, hipMalloc(&tspec, sizeof(double2) * (size_t)P->ncg * G * MY));
  CUP2D_HIP_CHECK(c, hipMemcpy(dG, Gk.data(), sizeof(double) * Gk.size(), hipMemcpyHostToDevice));
  CUP2D_HIP_CHECK(c, hipMalloc(&P->ghat, sizeof(double) * (size_t)P->ncg * G * (NY + 1)));
  CUP2D_HIP_CHECK(c, hipMalloc(&P->ghatN, sizeof(double) * (size_t)(NY + 1)));
  int rc = rows_fwd(c, P->Lr, dG, MX, MX, MY, tspec, MY, P->tw_r, SRC_PLAIN, RowSource{});
  if (rc) return rc;
  ColParams cp{};
  cp.NX = NX; cp.NY = NY; cp.rows = MY;
  rc = cols<2>(c, P->Lc, tspec, P->ncg, cp, P->tw_c);
  if (rc) return rc;
  const double norm = 1.0 / ((double)MY * MX);
  hipLaunchKernelGGL(k_ghat_extract, dim3(256), dim3(256), 0, c->stream, tspec, MY, P->ncg, NY,
                     norm, P->ghat, P->ghatN);
  rc = launch_check(c, "green kernel");
  if (rc) return rc;
  CUP2D_HIP_CHECK(c, hipStreamSynchronize(c->stream));
  (void)hipFree(dG);
  (void)hipFree(tspec);
  return CUP2D_OK;
}

void fft_destroy(cup2d_ctx* c) {
  FFTPlan* P = c->fft;
  if (!P) return;
  (void)hipFree(P->spec);
  (void)hipFree(P->tw_r);
  (void)hipFree(P->tw_c);
  if (P->ghat) (void)hipFree(P->ghat);
  if (P->ghatN) (void)hipFree(P->ghatN);
  if (P->cosX) (void)hipFree(P->cosX);
  if (P->cosY) (void)hipFree(P->cosY);
  delete P;
  c->fft = nullptr;
}

// src: SRC_PLAIN (rhs as given), SRC_FADE (rhs is the grid RHS: apply
// cub2rhs's boundary fade after enqueue_fade_sums and write the faded cells
// back), SRC_DIV (form the RHS from the advDiff output first, see k_rows_fwd).
int enqueue_poisson_fft(cup2d_ctx* c, double* rhs, double* sol, int src) {
  FFTPlan* P = c->fft;
  const int NX = c->g.NX, NY = c->g.NY;
  int rc;
  {
    KernelTimer kt(c, T_FFT_ROWS_FWD);
    RowSource fd{c->g, c->band_nblocks > 0 ? fade_partials(c) : 0, rhs, c->st,
                 c->u, c->v, c->obmap, P->cs};
    rc = rows_fwd(c, P->Lr, rhs, NX, NX, NY, P->spec, P->NYs, P->tw_r, src, fd);
  }
  if (rc) return rc;
  ColParams cp{};
  cp.NX = NX; cp.NY = NY; cp.rows = NY;
  {
    KernelTimer kt(c, T_FFT_COLS);
    if (P->cs) {  // FFTW_dirichlet.h:47-48
      cp.waveFactX = M_PI / NX;
      cp.waveFactY = M_PI / NY;
      cp.norm = 0.25 / ((double)NX * NY);
      cp.cosX = P->cosX;
      cp.cosY = P->cosY;
      rc = cols<3>(c, P->Lc, P->spec, P->ncg, cp, P->tw_c);
    } else if (!P->fs) {
      cp.waveFactX = 2 * M_PI / NX;
      cp.waveFactY = 2 * M_PI / NY;
      cp.norm = 1.0 / ((double)NX * NY);
      rc = cols<0>(c, P->Lc, P->spec, P->ncg, cp, P->tw_c);
    } else {
      cp.ghat = P->ghat;
      cp.ghatN = P->ghatN;
      rc = cols<1>(c, P->Lc, P->spec, P->ncg, cp, P->tw_c);
    }
  }
  if (rc) return rc;
  {
    // a ComputeForces still reading the previous pressure on the side stream
    // must finish before it is overwritten
    rc = join_forces(c);
    if (rc) return rc;
    KernelTimer kt(c, T_FFT_ROWS_INV);
    rc = rows_inv(c, P->Lr, P->spec, P->NYs, sol, NX, NX, NY, P->tw_r);
  }
  if (rc) return rc;
  return launch_check(c, "poisson_fft");
}

}  // namespace cup2d
