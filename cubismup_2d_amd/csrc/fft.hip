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
// Rows of padding after each column group of the spectrum, so that the
// groups of one row (written together by the row passes) do not sit a
// power of two apart; two rows keep every row pair on one 128 B line
// (one row, 64 B, split them and slowed the inverse row pass 27 -> 30 µs;
// 2 and 4 measured 0.7% faster per step than none at 4096x2048).
constexpr int SPEC_PAD = 2;

struct FFTPlan {
  int Lr = 0, Lc = 0;     // row / column FFT lengths (complex)
  int NYs = 0;            // rows stored in the spectrum
  int ncg = 0;            // column groups
  int fs = 0;             // free-space
  int cs = 0;             // cosine (Neumann) solver on the mirrored 2NX x 2NY grid
  double2* spec = nullptr;
  double2* tw_r = nullptr;
  double2* tw_c = nullptr;
  // FP32 precision: the FP32 solve's spectrum and twiddle tables (float2)
  void* spec32 = nullptr;
  void* tw_r32 = nullptr;
  void* tw_c32 = nullptr;
  double* ghat = nullptr;   // free-space: [ncg][Lc/2+1][G]
  double* ghatN = nullptr;  // free-space Nyquist column [Lc/2+1]
  double* cosX = nullptr;   // cosine: (float) cos(2 pi i / NX), i < NX
  double* cosY = nullptr;   // cosine: (float) cos(2 pi j / NY), j < NY
};

// The FFT core works on V = double2 (FP64) or float2 (FP32) with S the
// matching scalar; every kernel below is templated on S.
template <typename S> struct CplxOf;
template <> struct CplxOf<double> { using type = double2; };
template <> struct CplxOf<float> { using type = float2; };
template <typename S> using cplx = typename CplxOf<S>::type;
template <typename V> struct ScalarOf;
template <> struct ScalarOf<double2> { using type = double; };
template <> struct ScalarOf<float2> { using type = float; };

template <typename V>
__device__ __forceinline__ V mk(typename ScalarOf<V>::type a, typename ScalarOf<V>::type b) {
  V r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename V>
__device__ __forceinline__ V cmul(V a, V b) {
  return mk<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename V>
__device__ __forceinline__ V cadd(V a, V b) { return mk<V>(a.x + b.x, a.y + b.y); }
template <typename V>
__device__ __forceinline__ V csub(V a, V b) { return mk<V>(a.x - b.x, a.y - b.y); }
// multiply by -i (forward) or +i (inverse)
template <bool FWD, typename V>
__device__ __forceinline__ V rot(V a) {
  return FWD ? mk<V>(a.y, -a.x) : mk<V>(-a.y, a.x);
}

template <bool FWD, typename V>
__device__ __forceinline__ void dft2(V& a, V& b) {
  const V t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <bool FWD, typename V>
__device__ __forceinline__ void dft4(V& x0, V& x1, V& x2, V& x3) {
  const V a0 = cadd(x0, x2), a1 = csub(x0, x2), a2 = cadd(x1, x3);
  const V a3 = rot<FWD>(csub(x1, x3));
  x0 = cadd(a0, a2);
  x1 = cadd(a1, a3);
  x2 = csub(a0, a2);
  x3 = csub(a1, a3);
}

// W_N^m = exp(-2 pi i m/N) (forward) constants, rounded to S
template <typename V, bool FWD>
__device__ __forceinline__ V w16(int m) {
  using S = typename ScalarOf<V>::type;
  constexpr S c1 = (S)0.92387953251128673848, s1 = (S)0.38268343236508978178;
  constexpr S c2 = (S)0.70710678118654752440;
  V w;
  switch (m) {
    case 0: w = mk<V>(1, 0); break;
    case 1: w = mk<V>(c1, -s1); break;
    case 2: w = mk<V>(c2, -c2); break;
    case 3: w = mk<V>(s1, -c1); break;
    case 4: w = mk<V>(0, -1); break;
    case 6: w = mk<V>(-c2, -c2); break;
    case 9: w = mk<V>(-c1, s1); break;
    default: w = mk<V>(1, 0); break;
  }
  if (!FWD) w.y = -w.y;
  return w;
}

template <int R, bool FWD, typename V>
__device__ __forceinline__ void dft(V* v) {
  if constexpr (R == 2) {
    dft2<FWD>(v[0], v[1]);
  } else if constexpr (R == 4) {
    dft4<FWD>(v[0], v[1], v[2], v[3]);
  } else if constexpr (R == 8) {
    V y[2][4];
#pragma unroll
    for (int n2 = 0; n2 < 2; ++n2) {
      y[n2][0] = v[n2]; y[n2][1] = v[n2 + 2]; y[n2][2] = v[n2 + 4]; y[n2][3] = v[n2 + 6];
      dft4<FWD>(y[n2][0], y[n2][1], y[n2][2], y[n2][3]);
    }
    // W8^k1 = W16^(2 k1)
#pragma unroll
    for (int k1 = 1; k1 < 4; ++k1) y[1][k1] = cmul(y[1][k1], w16<V, FWD>(2 * k1));
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      v[k1] = cadd(y[0][k1], y[1][k1]);
      v[k1 + 4] = csub(y[0][k1], y[1][k1]);
    }
  } else {  // 16
    V y[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
      y[n2][0] = v[n2]; y[n2][1] = v[n2 + 4]; y[n2][2] = v[n2 + 8]; y[n2][3] = v[n2 + 12];
      dft4<FWD>(y[n2][0], y[n2][1], y[n2][2], y[n2][3]);
    }
#pragma unroll
    for (int n2 = 1; n2 < 4; ++n2)
#pragma unroll
      for (int k1 = 1; k1 < 4; ++k1) y[n2][k1] = cmul(y[n2][k1], w16<V, FWD>(n2 * k1));
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      dft4<FWD>(y[0][k1], y[1][k1], y[2][k1], y[3][k1]);
      v[k1] = y[0][k1]; v[k1 + 4] = y[1][k1]; v[k1 + 8] = y[2][k1]; v[k1 + 12] = y[3][k1];
    }
  }
}

// Elements per thread of the FFT core (the first pass is radix E), by
// transform length.  With 8 the kernels are held to 128 VGPRs (4 waves per
// SIMD, min_waves below) and two 8-wave workgroups share a CU; with 16 they
// take up to 256 VGPRs.  Measured on MI355X (µs per pass, 8 vs 16): rows of
// 4096 63.8 vs 67.5, columns of 2048 41.5 vs 42.6; rows of 8192 382 vs 350,
// inverse rows of 8192 157 vs 144, columns of 4096 205 vs 193.
template <int L>
constexpr int row_e() { return L >= 8192 ? 16 : 8; }
template <int LC>
constexpr int col_e() { return LC >= 4096 ? 16 : 8; }
template <int E>
constexpr int min_waves() { return E == 8 ? 4 : 1; }

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
template <int L, int E, int NS, typename V>
__device__ __forceinline__ void tw_load(const V* __restrict__ tw, int t, V* w) {
  if constexpr (NS < L) {
    constexpr int R = TwCount<L, E, NS>::R, PER = E / R, NT = L / E, TSTEP = L / (NS * R);
#pragma unroll
    for (int p = 0; p < PER; ++p) w[p] = tw[((t + p * NT) % NS) * TSTEP];
    tw_load<L, E, NS * R>(tw, t, w + PER);
  }
}

// pw[r] = b^r, r = 1..R-1, by repeated squaring and products of two lower
// powers (at most four roundings deep), conjugated for the inverse.
template <int R, bool FWD, typename V>
__device__ __forceinline__ void tw_powers(V b, V* pw) {
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
template <int L, int E, int R, int NS, bool FWD, typename V>
__device__ __forceinline__ void stockham_pass(V* lds, int base, int stride, int t, const V* w) {
  constexpr int NT = L / E, NBF = L / R, PER = E / R;
  V v[PER][R];
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
    V pw[R];
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
template <int L, int E, bool FWD, typename V>
__device__ __forceinline__ void pass_first(V* x, V* lds, int base, int stride, int t) {
  dft<E, FWD>(x);
#pragma unroll
  for (int r = 0; r < E; ++r) lds[base + padE<E>(t * E + r) * stride] = x[r];
  __syncthreads();
}

// Last pass (span NS = L/R) into registers: its outputs j + r*NS land on
// element t + q*NT with q = p + (E/R)*r, i.e. the same strided order the
// first pass reads.  Ends with a barrier so the LDS can be reused.
template <int L, int E, int R, int NS, bool FWD, typename V>
__device__ __forceinline__ void pass_last(V* lds, int base, int stride, int t, const V* w, V* x) {
  constexpr int NT = L / E, NBF = L / R, PER = E / R;
  static_assert(NS * R == L, "last pass");
  V v[PER][R];
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int j = t + p * NT;
#pragma unroll
    for (int r = 0; r < R; ++r) v[p][r] = lds[base + padE<E>(j + r * NBF) * stride];
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    V pw[R];
    tw_powers<R, FWD>(w[p], pw);
#pragma unroll
    for (int r = 1; r < R; ++r) v[p][r] = cmul(v[p][r], pw[r]);
    dft<R, FWD>(v[p]);
#pragma unroll
    for (int r = 0; r < R; ++r) x[p + PER * r] = v[p][r];
  }
}

template <int L, int E, int NS, bool FWD, typename V>
__device__ __forceinline__ void fft_tail(V* lds, int base, int stride, int t, const V* w, V* x) {
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
template <int L, int E, bool FWD, typename V>
__device__ __forceinline__ void fft_regs(V* x, V* lds, int base, int stride, int t, const V* w) {
  pass_first<L, E, FWD>(x, lds, base, stride, t);
  fft_tail<L, E, E, FWD>(lds, base, stride, t, w, x);
}

template <int L, int E>
constexpr int lds_len() { return L + L / E; }
template <int L, int E>
constexpr int rows_np() { return (L / E) >= 256 ? 1 : 256 / (L / E); }

// ---------------------------------------------------------- rows forward
// in: nrows rows of ncols valid reals (row pitch in_pitch), zero beyond;
// out: spec[cg][y][G], y < spec_rows.  NP row pairs per workgroup.  S is the
// transform's real type (double or float); every plane it reads is S.
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
//
// store_all (SRC_FADE, SRC_DIV): every cell's final RHS is written to the
// RHS plane, not only the faded band cells (the FP64 refinement reads it).
//
// SRC_DIV with obmap == nullptr is the early pass: every block's RHS is the
// divergence (obstacle blocks included), run before the obstacle lists
// exist.  SRC_PATCH then completes it: with st->rows_redo clear (no obstacle
// block in the fade band) only the pairs in block rows that hold an obstacle
// block transform delta = (final RHS - divergence) of the obstacle cells,
// zero elsewhere, and add it to the spectrum; with st->rows_redo set every
// pair redoes the whole SRC_DIV pass.  st->rows_early_on (set by the list
// builder's first phase from the previous step's obstacle rows) switches
// the early pass off when many block rows hold obstacles, the patch then
// costing about a whole pass: the early launch exits at once and SRC_PATCH
// does the whole pass.
enum { SRC_PLAIN = 0, SRC_FADE = 1, SRC_DIV = 2, SRC_PATCH = 3 };
struct RowSource {
  Grid g;
  int npart;            // fade-sum workgroups (0: no fade)
  void* rhs;            // RHS plane (S): faded band cells written back; obstacle blocks read (SRC_DIV)
  const StepState* st;  // fade_corr, flux_corr, dt
  const void* u;        // SRC_DIV: advDiff output (S)
  const void* v;
  const unsigned char* obmap;
  int mirror;           // cosine: columns past ncols read the row reversed
  int store_all;
  const double* fadepart;  // SRC_PATCH redo: the fade sums' partials (npart x 3)
};

// Forward transform of a packed row pair held in x (x[q] = element t + q*NT
// of row y0 + i row y0+1) and its split into the two rows' half spectra,
// stored (or with patch added) at spec[cg][y0, y0+1][G].  Every thread of the
// workgroup takes part (the barriers); invalid pairs store nothing.
template <int L, int E, typename S>
__device__ __forceinline__ void rows_fwd_transform_store(cplx<S>* x, cplx<S>* lds, int base, int t,
                                                         const cplx<S>* w, bool valid,
                                                         cplx<S>* __restrict__ spec,
                                                         int spec_rows, int y0, bool patch) {
  using V = cplx<S>;
  constexpr int NT = L / E;
  fft_regs<L, E, true>(x, lds, base, 1, t, w);
  // split z = a + i b: A_k = (Z_k + conj Z_{L-k})/2, B_k = (Z_k - conj Z_{L-k})/(2i);
  // the partners Z_{L-k} (upper half) go through LDS
#pragma unroll
  for (int q = E / 2; q < E; ++q) lds[base + padE<E>(t + q * NT)] = x[q];
  __syncthreads();
  if (!valid) return;
  const S half = (S)0.5;
#pragma unroll
  for (int q = 0; q < E / 2; ++q) {
    const int k = t + q * NT;  // 0 .. L/2-1
    const V Z = x[q];
    V A, B;
    if (k == 0) {
      const V ZN = x[E / 2];  // element L/2
      A = mk<V>(Z.x, ZN.x);
      B = mk<V>(Z.y, ZN.y);
    } else {
      const V Zm = lds[base + padE<E>(L - k)];
      const V sm = mk<V>(Z.x + Zm.x, Z.y - Zm.y);  // Z + conj(Zm)
      const V d = mk<V>(Z.x - Zm.x, Z.y + Zm.y);   // Z - conj(Zm)
      A = mk<V>(half * sm.x, half * sm.y);
      B = mk<V>(half * d.y, -half * d.x);
    }
    const int cg = k / G, c = k % G;
    V* o = spec + ((size_t)cg * spec_rows + y0) * G + c;
    if (patch) {
      const V A0 = o[0], B0 = o[G];
      o[0] = mk<V>(A0.x + A.x, A0.y + A.y);
      o[G] = mk<V>(B0.x + B.x, B0.y + B.y);
    } else {
      o[0] = A;
      o[G] = B;
    }
  }
}

// One group of NP row pairs (group wg: pairs wg*NP .. wg*NP + NP-1) of
// k_rows_fwd; tw is the length-L twiddle table.
template <int L, int NP, int E, int SRC, typename S>
__device__ __forceinline__ void rows_fwd_group(int wg, const S* in, int in_pitch, int ncols,
                                               int nrows, cplx<S>* __restrict__ spec,
                                               int spec_rows, const cplx<S>* __restrict__ tw,
                                               const RowSource& fd, cplx<S>* lds) {
  using V = cplx<S>;
  constexpr int NT = L / E, PADL = lds_len<L, E>();
  const int f = threadIdx.x / NT, t = threadIdx.x % NT;
  S* rhsp = static_cast<S*>(fd.rhs);
  constexpr bool DIV = SRC == SRC_DIV || SRC == SRC_PATCH;
  const int y0 = 2 * (wg * NP + f);
  const bool valid = y0 < nrows;
  const int base = f * PADL;
  // SRC_PATCH without a redo: delta rows, added to the spectrum
  bool patch = false;
  if constexpr (SRC == SRC_PATCH) {
    patch = fd.st->rows_redo == 0 && fd.st->rows_early_on != 0;
    if (patch) {  // groups whose block rows hold no obstacle block have nothing to add
      int any = 0;
      if (valid)
        for (int bx = t; bx < fd.g.bpdx; bx += NT) any |= fd.obmap[(y0 / BS) * fd.g.bpdx + bx];
      if (!__syncthreads_or(any)) return;
    }
  }
  // the fade factor: k_fade_reduce's, or with a redo of the early pass the
  // final one (the obstacle RHS kernel replaced the partials of the obstacle
  // blocks in the band), reduced here in one fixed order by every workgroup
  double fcorr = 0;
  if constexpr (SRC != SRC_PLAIN) fcorr = fd.st->fade_corr;
  if constexpr (SRC == SRC_PATCH) {
    if (fd.st->rows_redo != 0 && fd.npart > 0) {
      __shared__ double s_corr;
      if (threadIdx.x < 64) {
        double sums[3];
        wave_sum_table<3>(fd.fadepart, fd.npart, 3, sums);
        if (threadIdx.x == 0) s_corr = fade_corr_factor(sums[0], sums[1], sums[2]);
      }
      __syncthreads();
      fcorr = s_corr;
    }
  }
  V x[E];
  if constexpr (DIV) {
    const Grid& g = fd.g;
    const S* fu = static_cast<const S*>(fd.u);
    const S* fv = static_cast<const S*>(fd.v);
    const int NX = ncols, NY = nrows;
    const int yc = valid ? y0 : 0;
    const int ym = yc > 0 ? yc - 1 : yc, yp = yc + 2 < NY ? yc + 2 : yc + 1;
    S* ud = (S*)lds + (size_t)f * 2 * PADL;  // u of rows y0, y0+1
    // every load of the six rows is issued before the first use, so a wave
    // has all of them in flight (storing u into LDS inside this loop would
    // wait for each group in turn)
    S u0[E], u1[E], vm[E], v0[E], v1[E], vp[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int i = t + q * NT;
      if (i < NX) {
        u0[q] = fu[(size_t)yc * NX + i];
        u1[q] = fu[(size_t)(yc + 1) * NX + i];
        vm[q] = fv[(size_t)ym * NX + i];
        v0[q] = fv[(size_t)yc * NX + i];
        v1[q] = fv[(size_t)(yc + 1) * NX + i];
        vp[q] = fv[(size_t)yp * NX + i];
      }
    }
    const S fcorr = (S)fd.st->flux_corr;  // k_flux_reduce, after advDiff
    const S facDiv = (S)(0.5 * g.h / fd.st->dt);
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
      S r0 = 0, r1 = 0;
      if (valid && i < NX) {
        const int xe = i + 1 < NX ? i + 1 : i, xw = i > 0 ? i - 1 : i;
        const S ue0 = flux_fixed<S>(ud[xe], xe, NX, fcorr);
        const S uw0 = flux_fixed<S>(ud[xw], xw, NX, fcorr);
        const S ue1 = flux_fixed<S>(ud[L + xe], xe, NX, fcorr);
        const S uw1 = flux_fixed<S>(ud[L + xw], xw, NX, fcorr);
        const S vs0 = flux_fixed<S>(vm[q], ym, NY, fcorr);
        const S vc0 = flux_fixed<S>(v0[q], y0, NY, fcorr);
        const S vc1 = flux_fixed<S>(v1[q], y0 + 1, NY, fcorr);
        const S vn1 = flux_fixed<S>(vp[q], yp, NY, fcorr);
        r0 = facDiv * ((ue0 - uw0) + (vc1 - vs0));
        r1 = facDiv * ((ue1 - uw1) + (vn1 - vc0));
        const int bx = i / BS;
        if (patch) {  // a pair lies in one block row
          if (fd.obmap[(y0 / BS) * g.bpdx + bx]) {
            r0 = rhsp[(size_t)y0 * NX + i] - r0;
            r1 = rhsp[(size_t)(y0 + 1) * NX + i] - r1;
          } else {
            r0 = r1 = 0;
          }
        } else if (fd.obmap) {
          if (fd.obmap[(y0 / BS) * g.bpdx + bx]) r0 = rhsp[(size_t)y0 * NX + i];
          if (fd.obmap[((y0 + 1) / BS) * g.bpdx + bx]) r1 = rhsp[(size_t)(y0 + 1) * NX + i];
        }
      }
      x[q] = mk<V>(r0, r1);
    }
    __syncthreads();  // ud is reused below
  } else {
    const S* r0 = in + (size_t)(valid ? y0 : 0) * in_pitch;
    const S* r1 = r0 + in_pitch;
#pragma unroll
    for (int q = 0; q < E; ++q) {
      int i = t + q * NT;
      if (fd.mirror && i >= ncols) i = 2 * ncols - 1 - i;
      S a = 0, b = 0;
      if (valid && i < ncols) { a = r0[i]; b = r1[i]; }
      x[q] = mk<V>(a, b);
    }
  }
  if constexpr (SRC != SRC_PLAIN) {
    // (a patch has no band cells: an obstacle block in the band forces the redo)
    if (fd.npart > 0 && !patch) {
      const double corr = fcorr;
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
            x[q].x = (S)fade_cell((double)x[q].x, fade_arg(g, px, py), corr);
            if (own && !fd.store_all) rhsp[(size_t)y0 * in_pitch + i] = x[q].x;
          }
          if (row1 || col) {
            const double py = cell_pos(by1, (y0 + 1) % BS, g.hb, g.h);
            x[q].y = (S)fade_cell((double)x[q].y, fade_arg(g, px, py), corr);
            if (own && !fd.store_all) rhsp[(size_t)(y0 + 1) * in_pitch + i] = x[q].y;
          }
        }
      }
    }
    if (fd.store_all && valid && !patch) {
#pragma unroll
      for (int q = 0; q < E; ++q) {
        const int i = t + q * NT;
        if (i < ncols) {
          rhsp[(size_t)y0 * in_pitch + i] = x[q].x;
          rhsp[(size_t)(y0 + 1) * in_pitch + i] = x[q].y;
        }
      }
    }
  }
  // the thread's twiddle bases, loaded only now: held through the loads of
  // the source rows they would push the kernel past its register budget
  asm volatile("" ::: "memory");
  V w[TwCount<L, E, E>::n];
  tw_load<L, E, E>(tw, t, w);
  rows_fwd_transform_store<L, E, S>(x, lds, base, t, w, valid, spec, spec_rows, y0, patch);
}

template <int L, int NP, int E, int SRC, typename S>
__global__ __launch_bounds__(NP * L / E, min_waves<E>()) void k_rows_fwd(
    const S* in, int in_pitch, int ncols, int nrows, cplx<S>* __restrict__ spec,
    int spec_rows, const cplx<S>* __restrict__ tw, RowSource fd) {
  using V = cplx<S>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* lds = reinterpret_cast<V*>(smem_raw);
  // the early pass runs only when the step's lists predict a small patch
  if constexpr (SRC == SRC_DIV)
    if (!fd.obmap && fd.st->rows_early_on == 0) return;
  // SRC_DIV reads the v rows next to its pair too: consecutive pairs run on
  // one XCD (workgroup b runs on XCD b % 8) so those rows are L2 hits.  The
  // patch keeps the plain order: its few busy groups (the obstacle rows, a
  // contiguous range) are then spread over all XCDs.
  bool xcd_order = SRC == SRC_DIV;
  if constexpr (SRC == SRC_PATCH) xcd_order = fd.st->rows_redo != 0 || fd.st->rows_early_on == 0;
  int wg = blockIdx.x;
  if (xcd_order && gridDim.x % 8 == 0) wg = (wg % 8) * (gridDim.x / 8) + wg / 8;
  rows_fwd_group<L, NP, E, SRC, S>(wg, in, in_pitch, ncols, nrows, spec, spec_rows, tw, fd, lds);
}

// ---------------------------------------------------------- rows inverse
// Output of the inverse row pass (x = the transform's real output):
// OUT_SET: out (double) = x; OUT_ADD: out (double) += x (the FP64 update of a
// refinement sweep); OUT_RES: out (float) = b - x, b the float RHS plane
// (the residual of a refinement sweep, x = the FP64 operator applied to p).
// With S = double and OUT_SET the pass is the plain FP64 one.
enum { OUT_SET = 0, OUT_ADD = 1, OUT_RES = 2, OUT_RES_FWD = 3 };
// OUT_RES_FWD (FP64 transform, float residual): OUT_RES whose residual rows,
// rounded to float as OUT_RES stores them, are transformed forward in FP32
// at once into spec32 (k_rows_fwd<SRC_PLAIN, float>'s arithmetic and
// layout) instead of being stored: the FP32 solve of the residual starts at
// the column pass.  Rows without a mirror only (L == ncols).
// OUT_ADD also subtracts *mean (the solution's mean before the update, which
// the spectral operator cannot see: the solve's convention is a zero mean).
// OUT_RES also writes the workgroup's statistics of r and b to
// respart[5 wg .. 5 wg + 4]: cell count, sum r, sum (r - mean_wg r)^2, sum b,
// sum (b - mean_wg b)^2 (the residual norm of the sweep without the mean,
// which the operator does not see; k_res_reduce merges them).
template <int L, int NP, int E, typename S, int OUT>
__global__ __launch_bounds__(NP * L / E, min_waves<E>()) void k_rows_inv(const cplx<S>* __restrict__ spec,
                                                         int spec_rows, void* __restrict__ outv,
                                                         const float* __restrict__ bres,
                                                         int out_pitch, int ncols, int nrows,
                                                         const cplx<S>* __restrict__ tw,
                                                         const double* __restrict__ mean,
                                                         double* __restrict__ respart,
                                                         float2* __restrict__ spec32,
                                                         const float2* __restrict__ tw32) {
  using V = cplx<S>;
  constexpr int NT = L / E, PADL = lds_len<L, E>();
  constexpr bool RES = OUT == OUT_RES || OUT == OUT_RES_FWD;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* lds = reinterpret_cast<V*>(smem_raw);
  const int f = threadIdx.x / NT, t = threadIdx.x % NT;
  const int y0 = 2 * (blockIdx.x * NP + f);
  const bool valid = y0 < nrows;
  const int base = f * PADL;
  V w[TwCount<L, E, E>::n];
  tw_load<L, E, E>(tw, t, w);
  V x[E];
  // Z_k = A_k + i B_k and Z_{L-k} = conj(A_k) + i conj(B_k); DC and Nyquist
  // of both rows are packed in slot 0 (real parts only).  All loads are
  // issued before the first LDS store, so they are in flight together.
  V As[E / 2], Bs[E / 2];
#pragma unroll
  for (int q = 0; q < E / 2; ++q) {
    const int k = t + q * NT;
    As[q] = mk<V>(0, 0);
    Bs[q] = mk<V>(0, 0);
    if (valid) {
      const V* s = spec + ((size_t)(k / G) * spec_rows + y0) * G + (k % G);
      As[q] = s[0];
      Bs[q] = s[G];
    }
  }
#pragma unroll
  for (int q = 0; q < E / 2; ++q) {
    const int k = t + q * NT;
    const V A = As[q], B = Bs[q];
    if (k == 0) {
      x[0] = mk<V>(A.x, B.x);
      x[E / 2] = mk<V>(A.y, B.y);
    } else {
      x[q] = mk<V>(A.x - B.y, A.y + B.x);
      lds[base + padE<E>(L - k)] = mk<V>(A.x + B.y, -A.y + B.x);
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
  if constexpr (RES) {
    constexpr int NTB = NP * L / E;
    // r is held for the second moments; b is read again there (the same
    // floats, L1/L2 hits) rather than held, which would exceed the registers
    double rr[2 * E];
    double rs = 0, bs = 0, cnt = 0;
    const float* b0 = bres + (size_t)(valid ? y0 : 0) * out_pitch;
    const float* b1 = b0 + out_pitch;
    if (valid) {
      float* o0 = static_cast<float*>(outv) + (size_t)y0 * out_pitch;
      float* o1 = o0 + out_pitch;
#pragma unroll
      for (int q = 0; q < E; ++q) {
        const int i = t + q * NT;
        rr[2 * q] = rr[2 * q + 1] = 0;
        if (i < ncols) {
          const double bb0 = b0[i], bb1 = b1[i];
          const double r0 = bb0 - (double)x[q].x, r1 = bb1 - (double)x[q].y;
          if constexpr (OUT == OUT_RES) {
            o0[i] = (float)r0;
            o1[i] = (float)r1;
          }
          rr[2 * q] = r0; rr[2 * q + 1] = r1;
          rs += r0 + r1;
          bs += bb0 + bb1;
          cnt += 2;
        }
      }
    } else {
#pragma unroll
      for (int q = 0; q < 2 * E; ++q) rr[q] = 0;
    }
    // every thread reaches the reductions (no early return in this mode)
    double* red = reinterpret_cast<double*>(smem_raw);
    const double n_wg = block_sum<NTB>(cnt, red);
    rs = block_sum<NTB>(rs, red);
    bs = block_sum<NTB>(bs, red);
    const double mr = n_wg > 0 ? rs / n_wg : 0, mb = n_wg > 0 ? bs / n_wg : 0;
    double r2 = 0, b2 = 0;
    if (valid) {
#pragma unroll
      for (int q = 0; q < E; ++q) {
        const int i = t + q * NT;
        if (i < ncols) {
          const double d0 = rr[2 * q] - mr, d1 = rr[2 * q + 1] - mr;
          const double e0 = (double)b0[i] - mb, e1 = (double)b1[i] - mb;
          r2 += d0 * d0 + d1 * d1;
          b2 += e0 * e0 + e1 * e1;
        }
      }
    }
    r2 = block_sum<NTB>(r2, red);
    b2 = block_sum<NTB>(b2, red);
    if (threadIdx.x == 0) {
      double* o = respart + 5 * blockIdx.x;
      o[0] = n_wg; o[1] = rs; o[2] = r2; o[3] = bs; o[4] = b2;
    }
    if constexpr (OUT == OUT_RES_FWD) {
      __syncthreads();  // the reductions' LDS is reused by the transform
      float2* ldsf = reinterpret_cast<float2*>(smem_raw);
      float2 xf[E];
#pragma unroll
      for (int q = 0; q < E; ++q)
        xf[q] = t + q * NT < ncols ? make_float2((float)rr[2 * q], (float)rr[2 * q + 1])
                                   : make_float2(0.f, 0.f);
      asm volatile("" ::: "memory");  // the float bases are loaded only now
      float2 wf[TwCount<L, E, E>::n];
      tw_load<L, E, E>(tw32, t, wf);
      rows_fwd_transform_store<L, E, float>(xf, ldsf, base, t, wf, valid, spec32, spec_rows, y0,
                                            false);
    }
  } else {
    if (!valid) return;
    double* o0 = static_cast<double*>(outv) + (size_t)y0 * out_pitch;
    double* o1 = o0 + out_pitch;
    const double pm = OUT == OUT_ADD && mean ? *mean : 0.0;
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int i = t + q * NT;
      if (i < ncols) {
        if constexpr (OUT == OUT_ADD) {
          o0[i] += (double)x[q].x - pm;
          o1[i] += (double)x[q].y - pm;
        } else {
          o0[i] = (double)x[q].x;
          o1[i] = (double)x[q].y;
        }
      }
    }
  }
}

// relative residual norm |r - mean r| / |b - mean b| (r = b - A p) of a
// refinement sweep from the inverse row pass's per-workgroup statistics,
// merged by Chan's formula (M2 = sum M2_wg + sum n_wg (mean_wg - mean)^2),
// into *out
__global__ __launch_bounds__(256) void k_res_reduce(const double* __restrict__ respart, int n,
                                                    double* __restrict__ out) {
  __shared__ double lds[4];
  double cnt = 0, rs = 0, bs = 0;
  for (int i = threadIdx.x; i < n; i += 256) {
    cnt += respart[5 * i];
    rs += respart[5 * i + 1];
    bs += respart[5 * i + 3];
  }
  cnt = block_sum<256>(cnt, lds);
  rs = block_sum<256>(rs, lds);
  bs = block_sum<256>(bs, lds);
  const double mr = cnt > 0 ? rs / cnt : 0, mb = cnt > 0 ? bs / cnt : 0;
  double r2 = 0, b2 = 0;
  for (int i = threadIdx.x; i < n; i += 256) {
    const double nw = respart[5 * i];
    if (nw <= 0) continue;
    const double dr = respart[5 * i + 1] / nw - mr, db = respart[5 * i + 3] / nw - mb;
    r2 += respart[5 * i + 2] + nw * dr * dr;
    b2 += respart[5 * i + 4] + nw * db * db;
  }
  r2 = block_sum<256>(r2, lds);
  b2 = block_sum<256>(b2, lds);
  if (threadIdx.x == 0) *out = b2 > 0 ? sqrt(r2 / b2) : sqrt(r2);
}

// --------------------------------------------------------------- columns
// MODE 0: periodic solve, 1: free-space solve (zero-pruned), 2: forward only,
// 3: cosine solve, 4: periodic apply, 5: cosine apply.  The apply modes
// multiply by the inverse of the solve symbol (zero at the DC mode, which
// the solve discards): the FP64 spectral operator whose residual the
// refinement sweeps correct.
struct ColParams {
  int NX, NY;            // physical grid
  int rows;              // data rows stored per column (NY, or Lc for MODE 2)
  int stride;            // rows between two column groups in the spectrum (>= rows)
  double waveFactX, waveFactY, norm;
  const double* ghat;    // MODE 1: the Green's function spectrum, even-folded at GH = Lc/2
  const double* ghatN;
  int GH;
  const double* cosX;    // MODE 3, 5
  const double* cosY;
  double* dc_out;        // MODE 4, 5: receives the input's mean (its DC coefficient * norm)
};

__device__ __forceinline__ double periodic_factor(int kx, int ky, const ColParams& P) {
  // FFTW_periodic.h:82-91: kinv*norm_factor, DC set to zero
  const int kyy = ky <= P.NY / 2 ? ky : P.NY - ky;
  if (kx == 0 && kyy == 0) return 0.0;
  const double rkx = kx * P.waveFactX, rky = kyy * P.waveFactY;
  const double kinv = -1 / (rkx * rkx + rky * rky);
  return kinv * P.norm;
}
// the operator whose pseudo-inverse periodic_factor applies: -k^2 norm
__device__ __forceinline__ double periodic_apply(int kx, int ky, const ColParams& P) {
  const int kyy = ky <= P.NY / 2 ? ky : P.NY - ky;
  if (kx == 0 && kyy == 0) return 0.0;
  const double rkx = kx * P.waveFactX, rky = kyy * P.waveFactY;
  return -(rkx * rkx + rky * rky) * P.norm;
}

// FFTW_dirichlet::_solve (FFTW_dirichlet.h:45-61) at DCT indices (i, j) <
// (NX, NY).  Line 51 declares the blend weight as `static constexpr tol =
// 0.1;`, which has no type; it is taken as the double 0.1 the initializer
// states.  COScoef are floats (FFTW_dirichlet.h:26-27,65-68).
__device__ __forceinline__ double cosine_denom(int i, int j, const ColParams& P) {
  const double tol = 0.1;
  const double rkx = (i + 0.5) * P.waveFactX, rky = (j + 0.5) * P.waveFactY;
  const double denomFD = 1 - P.cosX[i] / 2 - P.cosY[j] / 2;
  const double denomSP = rkx * rkx + rky * rky;
  return (1 - tol) * denomFD + tol * denomSP;
}
__device__ __forceinline__ double cosine_factor(int i, int j, const ColParams& P) {
  if (i == 0 && j == 0) return 0.0;  // in_out[0] = 0
  return -P.norm / cosine_denom(i, j, P);
}
__device__ __forceinline__ double cosine_apply(int i, int j, const ColParams& P) {
  if (i == 0 && j == 0) return 0.0;
  return -P.norm * cosine_denom(i, j, P);
}

template <int LC, int MODE>
__device__ __forceinline__ double col_symbol(int kx, int cg, int c, int ky, const ColParams& P) {
  if constexpr (MODE == 0) {
    return periodic_factor(kx, ky, P);
  } else if constexpr (MODE == 4) {
    return periodic_apply(kx, ky, P);
  } else if constexpr (MODE == 3 || MODE == 5) {
    // symbol extended evenly about NX and NY; zero at the mirror Nyquist
    if (kx >= P.NX || ky == P.NY) return 0.0;
    const int kyy = ky < P.NY ? ky : LC - ky;
    return MODE == 3 ? cosine_factor(kx, kyy, P) : cosine_apply(kx, kyy, P);
  } else {
    const int kyy = ky <= P.GH ? ky : LC - ky;
    return P.ghat[((size_t)cg * (P.GH + 1) + kyy) * G + c];
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

template <int LC, int MODE, int E, typename S>
__global__ __launch_bounds__(CW * LC / E, min_waves<E>()) void k_cols(cplx<S>* __restrict__ spec, ColParams P,
                                                      const cplx<S>* __restrict__ tw) {
  using V = cplx<S>;
  constexpr int NT = LC / E;
  constexpr bool MIRROR = MODE == 3 || MODE == 5;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V* lds = reinterpret_cast<V*>(smem_raw);
  const int cl = threadIdx.x % CW, t = threadIdx.x / CW;
  int cg, half;
  col_group_of_block(blockIdx.x, gridDim.x, &cg, &half);
  const int c = half * CW + cl;  // column inside the group
  V* col = spec + (size_t)cg * P.stride * G;
  V w[TwCount<LC, E, E>::n];
  tw_load<LC, E, E>(tw, t, w);
  V x[E];
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int i = t + q * NT;
    if constexpr (MIRROR)  // mirrored grid: rows past NY are rows NY-1, NY-2, ...
      x[q] = col[(size_t)(i < P.rows ? i : 2 * P.rows - 1 - i) * G + c];
    else if constexpr (MODE == 1)  // zero padding: rows <= LC/2 (checked at launch)
      x[q] = q < E / 2 && i < P.rows ? col[(size_t)i * G + c] : mk<V>(0, 0);
    else
      x[q] = col[(size_t)i * G + c];
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
      // apply modes: the DC coefficient (real part of the packed DC column
      // at ky = 0) is the input's sum, times 4 on the mirrored grid; the
      // operator discards it, the refinement removes it from the solution
      if constexpr (MODE == 4 || MODE == 5)
        if (t == 0 && cl == 0 && P.dc_out) *P.dc_out = (double)x[0].x * P.norm;
      __syncthreads();
#pragma unroll
      for (int q = 0; q < E; ++q) {
        const int i = t + q * NT;
        const double s0 = col_symbol<LC, MODE>(kx, cg, c, i, P);
        if (kx == 0) {
          double sN;
          if constexpr (MODE == 0) sN = periodic_factor(P.NX / 2, i, P);
          else if constexpr (MODE == 4) sN = periodic_apply(P.NX / 2, i, P);
          else if constexpr (MIRROR) sN = 0.0;  // mirror Nyquist kx = NX
          else sN = P.ghatN[i <= P.GH ? i : LC - i];
          const V m = lds[padE<E>((LC - i) % LC) * CW + cl];
          const S a = (S)(0.5 * (s0 + sN)), b = (S)(0.5 * (s0 - sN));
          x[q] = mk<V>(x[q].x * a + m.x * b, x[q].y * a - m.y * b);
        } else {
          const S ss = (S)s0;
          x[q] = mk<V>(x[q].x * ss, x[q].y * ss);
        }
      }
      __syncthreads();
    } else {
#pragma unroll
      for (int q = 0; q < E; ++q) {
        const S sy = (S)col_symbol<LC, MODE>(kx, cg, c, t + q * NT, P);
        x[q] = mk<V>(x[q].x * sy, x[q].y * sy);
      }
    }
    // the bases again from the table (L1/L2 hits) rather than held through
    // the symbol step, where they would push the kernel past its register
    // budget
    asm volatile("" ::: "memory");
    tw_load<LC, E, E>(tw, t, w);
    fft_regs<LC, E, false>(x, lds, cl, CW, t, w);
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int i = t + q * NT;
      if (MODE == 1 && q >= E / 2) continue;
      if ((MODE != 1 && !MIRROR) || i < P.rows) col[(size_t)i * G + c] = x[q];
    }
  }
}

// Green's function spectrum: Re(FFT G) * norm, even-folded to ky <= GH =
// Lc/2.  spec holds the full forward transform of the Lr x Lc kernel.
__global__ void k_ghat_extract(const double2* __restrict__ spec, int Lc, int ncg, int GH,
                               double norm, double* __restrict__ ghat,
                               double* __restrict__ ghatN) {
  const int n = ncg * G * (GH + 1);
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const int c = q % G, ky = (q / G) % (GH + 1), cg = q / (G * (GH + 1));
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
template <int L, typename S>
static void launch_rows_fwd(cup2d_ctx* c, const S* in, int pitch, int ncols, int nrows,
                            cplx<S>* spec, int spec_rows, const cplx<S>* tw, int src,
                            const RowSource& fd, hipStream_t st) {
  constexpr int RE = row_e<L>();
  constexpr int NP = rows_np<L, RE>();
  const size_t shm = (size_t)NP * lds_len<L, RE>() * sizeof(cplx<S>);
  const dim3 grid((nrows / 2 + NP - 1) / NP), block(NP * L / RE);
  if (src == SRC_DIV)
    CUP2D_TIMED_LAUNCH(c, (k_rows_fwd<L, NP, RE, SRC_DIV, S>), grid, block, shm, st, in,
                       pitch, ncols, nrows, spec, spec_rows, tw, fd);
  else if (src == SRC_PATCH)
    CUP2D_TIMED_LAUNCH(c, (k_rows_fwd<L, NP, RE, SRC_PATCH, S>), grid, block, shm, st, in,
                       pitch, ncols, nrows, spec, spec_rows, tw, fd);
  else if (src == SRC_FADE)
    CUP2D_TIMED_LAUNCH(c, (k_rows_fwd<L, NP, RE, SRC_FADE, S>), grid, block, shm, st, in,
                       pitch, ncols, nrows, spec, spec_rows, tw, fd);
  else
    CUP2D_TIMED_LAUNCH(c, (k_rows_fwd<L, NP, RE, SRC_PLAIN, S>), grid, block, shm, st,
                       in, pitch, ncols, nrows, spec, spec_rows, tw, fd);
}
template <int L, typename S, int OUT>
static void launch_rows_inv(cup2d_ctx* c, const cplx<S>* spec, int spec_rows, void* out,
                            const float* bres, int pitch, int ncols, int nrows,
                            const cplx<S>* tw, const double* mean, float2* spec32,
                            const float2* tw32) {
  constexpr int RE = row_e<L>();
  c->rows_inv_wgs = (nrows / 2 + rows_np<L, RE>() - 1) / rows_np<L, RE>();
  constexpr int NP = rows_np<L, RE>();
  const size_t shm = (size_t)NP * lds_len<L, RE>() * sizeof(cplx<S>);
  const int pairs = nrows / 2;
  CUP2D_TIMED_LAUNCH(c, (k_rows_inv<L, NP, RE, S, OUT>), dim3((pairs + NP - 1) / NP),
                     dim3(NP * L / RE), shm, c->stream, spec, spec_rows, out, bres, pitch,
                     ncols, nrows, tw, mean, c->respart, spec32, tw32);
}
template <int LC, int MODE, typename S>
static int launch_cols(cup2d_ctx* c, cplx<S>* spec, int ncg, const ColParams& P,
                       const cplx<S>* tw) {
  constexpr int CE = col_e<LC>();
  if (MODE == 1 && 2 * P.rows > LC) return CUP2D_ERR_ARG;  // k_cols reads rows < LC/2 only
  const size_t shm = (size_t)lds_len<LC, CE>() * CW * sizeof(cplx<S>);
  CUP2D_TIMED_LAUNCH(c, (k_cols<LC, MODE, CE, S>), dim3(ncg * (G / CW)), dim3(CW * LC / CE), shm,
                     c->stream, spec, P, tw);
  return CUP2D_OK;
}

template <typename S>
static int rows_fwd(cup2d_ctx* c, int L, const S* in, int pitch, int ncols, int nrows,
                    cplx<S>* spec, int spec_rows, const cplx<S>* tw, int src,
                    const RowSource& fd, hipStream_t st = nullptr) {
  if (!st) st = c->stream;
  switch (L) {
    case 32: launch_rows_fwd<32, S>(c, in, pitch, ncols, nrows, spec, spec_rows, tw, src, fd, st); break;
    case 64: launch_rows_fwd<64, S>(c, in, pitch, ncols, nrows, spec, spec_rows, tw, src, fd, st); break;
    case 128: launch_rows_fwd<128, S>(c, in, pitch, ncols, nrows, spec, spec_rows, tw, src, fd, st); break;
    case 256: launch_rows_fwd<256, S>(c, in, pitch, ncols, nrows, spec, spec_rows, tw, src, fd, st); break;
    case 512: launch_rows_fwd<512, S>(c, in, pitch, ncols, nrows, spec, spec_rows, tw, src, fd, st); break;
    case 1024: launch_rows_fwd<1024, S>(c, in, pitch, ncols, nrows, spec, spec_rows, tw, src, fd, st); break;
    case 2048: launch_rows_fwd<2048, S>(c, in, pitch, ncols, nrows, spec, spec_rows, tw, src, fd, st); break;
    case 4096: launch_rows_fwd<4096, S>(c, in, pitch, ncols, nrows, spec, spec_rows, tw, src, fd, st); break;
    case 8192: launch_rows_fwd<8192, S>(c, in, pitch, ncols, nrows, spec, spec_rows, tw, src, fd, st); break;
    default: return CUP2D_ERR_ARG;
  }
  return CUP2D_OK;
}
template <typename S, int OUT>
static int rows_inv(cup2d_ctx* c, int L, const cplx<S>* spec, int spec_rows, void* out,
                    const float* bres, int pitch, int ncols, int nrows, const cplx<S>* tw,
                    const double* mean = nullptr, float2* spec32 = nullptr,
                    const float2* tw32 = nullptr) {
  switch (L) {
    case 32: launch_rows_inv<32, S, OUT>(c, spec, spec_rows, out, bres, pitch, ncols, nrows, tw, mean, spec32, tw32); break;
    case 64: launch_rows_inv<64, S, OUT>(c, spec, spec_rows, out, bres, pitch, ncols, nrows, tw, mean, spec32, tw32); break;
    case 128: launch_rows_inv<128, S, OUT>(c, spec, spec_rows, out, bres, pitch, ncols, nrows, tw, mean, spec32, tw32); break;
    case 256: launch_rows_inv<256, S, OUT>(c, spec, spec_rows, out, bres, pitch, ncols, nrows, tw, mean, spec32, tw32); break;
    case 512: launch_rows_inv<512, S, OUT>(c, spec, spec_rows, out, bres, pitch, ncols, nrows, tw, mean, spec32, tw32); break;
    case 1024: launch_rows_inv<1024, S, OUT>(c, spec, spec_rows, out, bres, pitch, ncols, nrows, tw, mean, spec32, tw32); break;
    case 2048: launch_rows_inv<2048, S, OUT>(c, spec, spec_rows, out, bres, pitch, ncols, nrows, tw, mean, spec32, tw32); break;
    case 4096: launch_rows_inv<4096, S, OUT>(c, spec, spec_rows, out, bres, pitch, ncols, nrows, tw, mean, spec32, tw32); break;
    case 8192: launch_rows_inv<8192, S, OUT>(c, spec, spec_rows, out, bres, pitch, ncols, nrows, tw, mean, spec32, tw32); break;
    default: return CUP2D_ERR_ARG;
  }
  return CUP2D_OK;
}
template <int MODE, typename S>
static int cols(cup2d_ctx* c, int LC, cplx<S>* spec, int ncg, const ColParams& P,
                const cplx<S>* tw) {
  switch (LC) {
    case 32: return launch_cols<32, MODE, S>(c, spec, ncg, P, tw);
    case 64: return launch_cols<64, MODE, S>(c, spec, ncg, P, tw);
    case 128: return launch_cols<128, MODE, S>(c, spec, ncg, P, tw);
    case 256: return launch_cols<256, MODE, S>(c, spec, ncg, P, tw);
    case 512: return launch_cols<512, MODE, S>(c, spec, ncg, P, tw);
    case 1024: return launch_cols<1024, MODE, S>(c, spec, ncg, P, tw);
    case 2048: return launch_cols<2048, MODE, S>(c, spec, ncg, P, tw);
    case 4096: return launch_cols<4096, MODE, S>(c, spec, ncg, P, tw);
    default: return CUP2D_ERR_ARG;
  }
  return CUP2D_OK;
}

template <int L, typename S>
static void set_rows_attr_LS() {
  constexpr int RE = row_e<L>();
  constexpr int NP = rows_np<L, RE>();
  const int shm = NP * lds_len<L, RE>() * (int)sizeof(cplx<S>);
  (void)hipFuncSetAttribute((const void*)k_rows_fwd<L, NP, RE, SRC_PLAIN, S>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, shm);
  (void)hipFuncSetAttribute((const void*)k_rows_fwd<L, NP, RE, SRC_FADE, S>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, shm);
  (void)hipFuncSetAttribute((const void*)k_rows_fwd<L, NP, RE, SRC_DIV, S>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, shm);
  (void)hipFuncSetAttribute((const void*)k_rows_fwd<L, NP, RE, SRC_PATCH, S>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, shm);
  (void)hipFuncSetAttribute((const void*)k_rows_inv<L, NP, RE, S, OUT_SET>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, shm);
  if constexpr (sizeof(S) == 4)
    (void)hipFuncSetAttribute((const void*)k_rows_inv<L, NP, RE, S, OUT_ADD>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, shm);
  else {
    (void)hipFuncSetAttribute((const void*)k_rows_inv<L, NP, RE, S, OUT_RES>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, shm);
    (void)hipFuncSetAttribute((const void*)k_rows_inv<L, NP, RE, S, OUT_RES_FWD>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, shm);
  }
}
template <int L>
static void set_rows_attr_L() {
  set_rows_attr_LS<L, double>();
  set_rows_attr_LS<L, float>();
}
template <int LC>
static void set_cols_attr() {
  constexpr int CE = col_e<LC>();
  const int shm = lds_len<LC, CE>() * CW * (int)sizeof(double2);
  const int shm32 = lds_len<LC, CE>() * CW * (int)sizeof(float2);
  const hipFuncAttribute A = hipFuncAttributeMaxDynamicSharedMemorySize;
  (void)hipFuncSetAttribute((const void*)k_cols<LC, 0, CE, double>, A, shm);
  (void)hipFuncSetAttribute((const void*)k_cols<LC, 1, CE, double>, A, shm);
  (void)hipFuncSetAttribute((const void*)k_cols<LC, 2, CE, double>, A, shm);
  (void)hipFuncSetAttribute((const void*)k_cols<LC, 3, CE, double>, A, shm);
  (void)hipFuncSetAttribute((const void*)k_cols<LC, 4, CE, double>, A, shm);
  (void)hipFuncSetAttribute((const void*)k_cols<LC, 5, CE, double>, A, shm);
  (void)hipFuncSetAttribute((const void*)k_cols<LC, 0, CE, float>, A, shm32);
  (void)hipFuncSetAttribute((const void*)k_cols<LC, 1, CE, float>, A, shm32);
  (void)hipFuncSetAttribute((const void*)k_cols<LC, 3, CE, float>, A, shm32);
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
static std::vector<float2> twiddles32(int L) {
  std::vector<float2> w(L);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int k = 0; k < L; ++k) {
    const long double a = -2.0L * pi * (long double)k / (long double)L;
    w[k] = make_float2((float)cosl(a), (float)sinl(a));
  }
  return w;
}

static bool pow2(int n) { return n >= 32 && (n & (n - 1)) == 0; }
static int next_pow2(int n) {
  int p = 1;
  while (p < n) p <<= 1;
  return p;
}

int fft_create(cup2d_ctx* c) {
  const int NX = c->g.NX, NY = c->g.NY;
  const bool fs = c->cfg.poisson == CUP2D_POISSON_FREESPACE;
  const bool cs = c->cfg.poisson == CUP2D_POISSON_COSINE;
  FFTPlan* P = new FFTPlan();
  c->fft = P;
  P->fs = fs;
  P->cs = cs;
  // Free-space: any cyclic length >= 2N - 1 gives the linear convolution on
  // the N data points (FFTW_freespace.h:25-27 uses 2N - 1), so the rows and
  // columns take the next power of two >= 2N and any grid of 32-cell blocks
  // is solved.  Periodic (N) and cosine (mirrored 2N) lengths must be powers
  // of two themselves.
  P->Lr = fs ? next_pow2(2 * NX) : cs ? 2 * NX : NX;
  P->Lc = fs ? next_pow2(2 * NY) : cs ? 2 * NY : NY;
  P->NYs = NY + SPEC_PAD;
  P->ncg = (P->Lr / 2) / G;
  if ((!fs && (!pow2(NX) || !pow2(NY))) || P->Lr > 8192 || P->Lc > 4096) {
    c->err = "FFT Poisson: periodic and cosine NX and NY must be powers of two; row length "
             "<= 8192 and column length <= 4096 (periodic NX <= 8192, NY <= 4096; cosine "
             "NX <= 4096, NY <= 2048; free-space next_pow2(2 NX) <= 8192, next_pow2(2 NY) "
             "<= 4096)";
    return CUP2D_ERR_ARG;
  }
  set_rows_attr_L<32>(); set_rows_attr_L<64>(); set_rows_attr_L<128>(); set_rows_attr_L<256>();
  set_rows_attr_L<512>(); set_rows_attr_L<1024>(); set_rows_attr_L<2048>();
  set_rows_attr_L<4096>(); set_rows_attr_L<8192>();
  set_cols_attr<32>(); set_cols_attr<64>(); set_cols_attr<128>(); set_cols_attr<256>();
  set_cols_attr<512>(); set_cols_attr<1024>(); set_cols_attr<2048>(); set_cols_attr<4096>();

  const size_t nspec = (size_t)P->ncg * G * P->NYs;
  CUP2D_HIP_CHECK(c, hipMalloc(&P->spec, sizeof(double2) * nspec));
  const std::vector<double2> twr = twiddles(P->Lr), twc = twiddles(P->Lc);
  CUP2D_HIP_CHECK(c, hipMalloc(&P->tw_r, sizeof(double2) * twr.size()));
  CUP2D_HIP_CHECK(c, hipMalloc(&P->tw_c, sizeof(double2) * twc.size()));
  CUP2D_HIP_CHECK(c, hipMemcpy(P->tw_r, twr.data(), sizeof(double2) * twr.size(), hipMemcpyHostToDevice));
  CUP2D_HIP_CHECK(c, hipMemcpy(P->tw_c, twc.data(), sizeof(double2) * twc.size(), hipMemcpyHostToDevice));
  if (c->f32) {
    // the FP32 solve's spectrum, tables rounded to float; P->spec (FP64)
    // serves the refinement's operator apply
    CUP2D_HIP_CHECK(c, hipMalloc(&P->spec32, sizeof(float2) * nspec));
    const std::vector<float2> twr32 = twiddles32(P->Lr), twc32 = twiddles32(P->Lc);
    CUP2D_HIP_CHECK(c, hipMalloc(&P->tw_r32, sizeof(float2) * twr32.size()));
    CUP2D_HIP_CHECK(c, hipMalloc(&P->tw_c32, sizeof(float2) * twc32.size()));
    CUP2D_HIP_CHECK(c, hipMemcpy(P->tw_r32, twr32.data(), sizeof(float2) * twr32.size(),
                                 hipMemcpyHostToDevice));
    CUP2D_HIP_CHECK(c, hipMemcpy(P->tw_c32, twc32.data(), sizeof(float2) * twc32.size(),
                                 hipMemcpyHostToDevice));
  }
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

  // Free-space Green's function on the Lr x Lc cyclic grid
  // (FFTW_freespace.cpp:100-137), even in both directions; with any cyclic
  // size >= 2N - 1 the first NX x NY outputs of the cyclic convolution equal
  // the reference's (2N-1) result.
  const int MX = P->Lr, MY = P->Lc, GH = MY / 2;
  std::vector<double> Gk((size_t)MX * MY);
  const double h = c->g.h, fac = 1 / (2 * M_PI);
  for (int j = 0; j < MY; ++j)
    for (int i = 0; i < MX; ++i) {
      const double yi = j > MY - j ? MY - j : j;
      const double xi = i > MX - i ? MX - i : i;
      const double r = std::sqrt(xi * xi + yi * yi);
      Gk[(size_t)j * MX + i] = fac * std::log(h * r);
    }
  Gk[0] = fac / 2 * (2 * std::log(h / std::sqrt(M_PI)) - 1);
  double* dG = nullptr;
  double2* tspec = nullptr;
  CUP2D_HIP_CHECK(c, hipMalloc(&dG, sizeof(double) * Gk.size()));
  CUP2D_HIP_CHECK(c, hipMalloc(&tspec, sizeof(double2) * (size_t)P->ncg * G * MY));
  CUP2D_HIP_CHECK(c, hipMemcpy(dG, Gk.data(), sizeof(double) * Gk.size(), hipMemcpyHostToDevice));
  CUP2D_HIP_CHECK(c, hipMalloc(&P->ghat, sizeof(double) * (size_t)P->ncg * G * (GH + 1)));
  CUP2D_HIP_CHECK(c, hipMalloc(&P->ghatN, sizeof(double) * (size_t)(GH + 1)));
  int rc = rows_fwd<double>(c, P->Lr, dG, MX, MX, MY, tspec, MY, P->tw_r, SRC_PLAIN,
                            RowSource{});
  if (rc) return rc;
  ColParams cp{};
  cp.NX = NX; cp.NY = NY; cp.rows = MY; cp.stride = MY;
  rc = cols<2, double>(c, P->Lc, tspec, P->ncg, cp, P->tw_c);
  if (rc) return rc;
  const double norm = 1.0 / ((double)MY * MX);
  hipLaunchKernelGGL(k_ghat_extract, dim3(256), dim3(256), 0, c->stream, tspec, MY, P->ncg, GH,
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
  if (P->spec32) (void)hipFree(P->spec32);
  if (P->tw_r32) (void)hipFree(P->tw_r32);
  if (P->tw_c32) (void)hipFree(P->tw_c32);
  if (P->ghat) (void)hipFree(P->ghat);
  if (P->ghatN) (void)hipFree(P->ghatN);
  if (P->cosX) (void)hipFree(P->cosX);
  if (P->cosY) (void)hipFree(P->cosY);
  delete P;
  c->fft = nullptr;
}

// The column pass's parameters and mode (solve or apply) for this plan.
static ColParams col_params(const cup2d_ctx* c, bool apply, int* mode) {
  const FFTPlan* P = c->fft;
  const int NX = c->g.NX, NY = c->g.NY;
  ColParams cp{};
  cp.NX = NX; cp.NY = NY; cp.rows = NY; cp.stride = P->NYs;
  if (P->cs) {  // FFTW_dirichlet.h:47-48
    cp.waveFactX = M_PI / NX;
    cp.waveFactY = M_PI / NY;
    cp.norm = 0.25 / ((double)NX * NY);
    cp.cosX = P->cosX;
    cp.cosY = P->cosY;
    *mode = apply ? 5 : 3;
  } else if (!P->fs) {
    cp.waveFactX = 2 * M_PI / NX;
    cp.waveFactY = 2 * M_PI / NY;
    cp.norm = 1.0 / ((double)NX * NY);
    *mode = apply ? 4 : 0;
  } else {
    cp.ghat = P->ghat;
    cp.ghatN = P->ghatN;
    cp.GH = P->Lc / 2;
    *mode = 1;
  }
  return cp;
}

template <typename S>
static int cols_solve(cup2d_ctx* c, cplx<S>* spec, const cplx<S>* tw) {
  int mode;
  const ColParams cp = col_params(c, false, &mode);
  const FFTPlan* P = c->fft;
  if (mode == 3) return cols<3, S>(c, P->Lc, spec, P->ncg, cp, tw);
  if (mode == 0) return cols<0, S>(c, P->Lc, spec, P->ncg, cp, tw);
  return cols<1, S>(c, P->Lc, spec, P->ncg, cp, tw);
}

// The early forward row pass (see SRC_PATCH): the final fade factor from
// the early fade sums (an obstacle block in the band changes it, and then
// the pass is redone), then the RHS of every block formed from the advDiff
// output and transformed, on stream s.
int enqueue_rows_early(cup2d_ctx* c, hipStream_t s) {
  c->rows_early = false;
  FFTPlan* P = c->fft;
  if (!P || c->cfg.poisson == CUP2D_POISSON_COSINE || !c->flux_deferred) return CUP2D_OK;
  const int NX = c->g.NX, NY = c->g.NY;
  const int npart = c->band_nblocks > 0 ? fade_partials(c) : 0;
  if (npart > 0) {
    if (!c->fade_early) return CUP2D_OK;
    const int rf = enqueue_fade_reduce_on(c, s);
    if (rf) return rf;
  }
  const RowSource fd{c->g, npart, c->rhs, c->st, c->u, c->v, nullptr, P->cs, c->f32 ? 1 : 0};
  int rc;
  {
    KernelTimer kt(c, T_FFT_ROWS_FWD);
    if (!c->f32)
      rc = rows_fwd<double>(c, P->Lr, c->rhs, NX, NX, NY, P->spec, P->NYs, P->tw_r, SRC_DIV, fd,
                            s);
    else
      rc = rows_fwd<float>(c, P->Lr, tp<float>(c->rhs), NX, NX, NY,
                           reinterpret_cast<float2*>(P->spec32), P->NYs,
                           reinterpret_cast<const float2*>(P->tw_r32), SRC_DIV, fd, s);
  }
  if (rc) return rc;
  c->rows_early = true;
  return launch_check(c, "early row pass");
}

// src: SRC_PLAIN (rhs as given), SRC_FADE (rhs is the grid RHS: apply
// cub2rhs's boundary fade after enqueue_fade_sums and write the faded cells
// back), SRC_DIV (form the RHS from the advDiff output first, see k_rows_fwd).
// At FP32 precision rhs is the float RHS plane, the solve runs in float and
// its result goes to the double pressure, then refine_sweeps sweeps of
//   r = b - A p  (A: the FP64 spectral operator; rows, apply column pass,
//                 inverse rows writing the float residual),
//   p += S32(r)  (the FP32 solve of the residual added in FP64),
// with b the RHS the FP32 solve transformed (stored whole by its row pass).
int enqueue_poisson_fft(cup2d_ctx* c, double* rhs, double* sol, int src) {
  FFTPlan* P = c->fft;
  const int NX = c->g.NX, NY = c->g.NY;
  int rc;
  // the forward row pass ran early on the side stream: add the obstacle
  // blocks' correction (timed with the obstacle RHS, whose work it completes)
  const bool patch = src == SRC_DIV && c->rows_early;
  c->rows_early = false;
  if (patch) src = SRC_PATCH;
  if (!c->f32) {
    {
      KernelTimer kt(c, patch ? T_RHS_OBST : T_FFT_ROWS_FWD);
      RowSource fd{c->g, c->band_nblocks > 0 ? fade_partials(c) : 0, rhs, c->st,
                   c->u, c->v, c->obmap, P->cs, 0, c->red};
      rc = rows_fwd<double>(c, P->Lr, rhs, NX, NX, NY, P->spec, P->NYs, P->tw_r, src, fd);
    }
    if (rc) return rc;
    {
      KernelTimer kt(c, T_FFT_COLS);
      rc = cols_solve<double>(c, P->spec, P->tw_c);
    }
    if (rc) return rc;
    {
      // a ComputeForces still reading the previous pressure on the side stream
      // must finish before it is overwritten
      rc = join_forces(c);
      if (rc) return rc;
      KernelTimer kt(c, T_FFT_ROWS_INV);
      rc = rows_inv<double, OUT_SET>(c, P->Lr, P->spec, P->NYs, sol, nullptr, NX, NX, NY,
                                     P->tw_r);
    }
    if (rc) return rc;
    return launch_check(c, "poisson_fft");
  }
  // FP32 solve
  float* b = tp<float>(rhs);
  float2* s32 = reinterpret_cast<float2*>(P->spec32);
  const float2* twr32 = reinterpret_cast<const float2*>(P->tw_r32);
  const float2* twc32 = reinterpret_cast<const float2*>(P->tw_c32);
  {
    KernelTimer kt(c, patch ? T_RHS_OBST : T_FFT_ROWS_FWD);
    RowSource fd{c->g, c->band_nblocks > 0 ? fade_partials(c) : 0, rhs, c->st,
                 c->u, c->v, c->obmap, P->cs, src != SRC_PLAIN ? 1 : 0, c->red};
    rc = rows_fwd<float>(c, P->Lr, b, NX, NX, NY, s32, P->NYs, twr32, src, fd);
  }
  if (rc) return rc;
  {
    KernelTimer kt(c, T_FFT_COLS);
    rc = cols_solve<float>(c, s32, twc32);
  }
  if (rc) return rc;
  rc = join_forces(c);
  if (rc) return rc;
  {
    KernelTimer kt(c, T_FFT_ROWS_INV);
    rc = rows_inv<float, OUT_SET>(c, P->Lr, s32, P->NYs, sol, nullptr, NX, NX, NY, twr32);
  }
  if (rc) return rc;
  const int sweeps = P->fs ? 0 : c->refine_sweeps;
  static_assert(sizeof(StepState::res) / sizeof(double) >= 4, "one residual per sweep");
  float* res = tp<float>(c->rres);
  for (int k = 0; k < sweeps; ++k) {
    TimerScope ts(c, T_REFINE);
    const RowSource plain{c->g, 0, nullptr, c->st, nullptr, nullptr, nullptr, P->cs, 0};
    rc = rows_fwd<double>(c, P->Lr, sol, NX, NX, NY, P->spec, P->NYs, P->tw_r, SRC_PLAIN, plain);
    if (rc) return rc;
    int mode;
    ColParams ca = col_params(c, true, &mode);
    ca.dc_out = &c->st->pmean;
    rc = mode == 5 ? cols<5, double>(c, P->Lc, P->spec, P->ncg, ca, P->tw_c)
                   : cols<4, double>(c, P->Lc, P->spec, P->ncg, ca, P->tw_c);
    if (rc) return rc;
    if (!P->cs && P->Lr == NX) {  // the residual's forward rows fused into its pass
      rc = rows_inv<double, OUT_RES_FWD>(c, P->Lr, P->spec, P->NYs, nullptr, b, NX, NX, NY,
                                         P->tw_r, nullptr, s32, twr32);
      if (rc) return rc;
      hipLaunchKernelGGL(k_res_reduce, dim3(1), dim3(256), 0, c->stream, c->respart,
                         c->rows_inv_wgs, &c->st->res[k]);
    } else {
      rc = rows_inv<double, OUT_RES>(c, P->Lr, P->spec, P->NYs, res, b, NX, NX, NY, P->tw_r);
      if (rc) return rc;
      hipLaunchKernelGGL(k_res_reduce, dim3(1), dim3(256), 0, c->stream, c->respart,
                         c->rows_inv_wgs, &c->st->res[k]);
      rc = rows_fwd<float>(c, P->Lr, res, NX, NX, NY, s32, P->NYs, twr32, SRC_PLAIN, plain);
      if (rc) return rc;
    }
    rc = cols_solve<float>(c, s32, twc32);
    if (rc) return rc;
    rc = rows_inv<float, OUT_ADD>(c, P->Lr, s32, P->NYs, sol, nullptr, NX, NX, NY, twr32,
                                  &c->st->pmean);
    if (rc) return rc;
  }
  return launch_check(c, "poisson_fft");
}

}  // namespace cup2d
