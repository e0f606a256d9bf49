/*
 * cup2d_fft.c -- TEST INFRASTRUCTURE ONLY: the oracle's own FFT.
 *
 * FFTW3 is not available, so the oracle restates the transforms the
 * reference calls (fftw_plan_dft_r2c_2d / c2r_2d on an in-place buffer with
 * row stride 2*(MX/2+1), FFTW_periodic.h:118-119, FFTW_freespace.cpp:83-95):
 * unnormalized, forward sign -1, backward sign +1.
 *   - power-of-two lengths: iterative radix-2 Cooley-Tukey (decimation in
 *     time, bit-reversed input order), twiddles exp(-2 pi i k/n) from libm;
 *   - other lengths: Bluestein's chirp-z algorithm (L. Bluestein 1970) on a
 *     power-of-two convolution, chirp angles reduced with k^2 mod 2n;
 *   - real transforms of even length use the standard half-length complex
 *     packing (Numerical Recipes 12.3 "realft" split).
 * Checked in tests against numpy.fft and closed-form Fourier modes.
 */
#include "cup2d_fft.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

static void radix2_init(fft1d* p, int n)
{
  p->n = n;
  p->tw = (double*)malloc(sizeof(double) * (size_t)(n > 1 ? n : 2));
  for (int k = 0; k < n / 2; ++k) {
    const double a = -2.0 * M_PI * (double)k / (double)n;
    p->tw[2 * k] = cos(a);
    p->tw[2 * k + 1] = sin(a);
  }
  p->rev = (int*)malloc(sizeof(int) * (size_t)n);
  int lg = 0;
  while ((1 << lg) < n) ++lg;
  for (int i = 0; i < n; ++i) {
    int r = 0;
    for (int b = 0; b < lg; ++b) r |= ((i >> b) & 1) << (lg - 1 - b);
    p->rev[i] = r;
  }
}

/* in-place radix-2, x interleaved complex */
static void radix2_exec(const fft1d* p, double* x, int sign)
{
  const int n = p->n;
  for (int i = 0; i < n; ++i) {
    const int r = p->rev[i];
    if (r > i) {
      double t0 = x[2 * i], t1 = x[2 * i + 1];
      x[2 * i] = x[2 * r]; x[2 * i + 1] = x[2 * r + 1];
      x[2 * r] = t0; x[2 * r + 1] = t1;
    }
  }
  for (int m = 1; m < n; m <<= 1) {
    const int step = n / (2 * m);
    for (int k = 0; k < n; k += 2 * m) {
      for (int j = 0; j < m; ++j) {
        const double wr = p->tw[2 * j * step];
        const double wi = -sign * p->tw[2 * j * step + 1]; /* table holds exp(-i..) */
        double* a = x + 2 * (k + j);
        double* b = x + 2 * (k + j + m);
        const double tr = wr * b[0] - wi * b[1];
        const double ti = wr * b[1] + wi * b[0];
        b[0] = a[0] - tr; b[1] = a[1] - ti;
        a[0] = a[0] + tr; a[1] = a[1] + ti;
      }
    }
  }
}

void fft1d_init(fft1d* p, int n)
{
  memset(p, 0, sizeof(*p));
  p->n = n;
  if (is_pow2(n)) { radix2_init(p, n); p->pow2 = 1; return; }
  p->pow2 = 0;
  int m = 1;
  while (m < 2 * n - 1) m <<= 1;
  p->sub = (fft1d*)malloc(sizeof(fft1d));
  memset(p->sub, 0, sizeof(fft1d));
  radix2_init(p->sub, m);
  p->sub->pow2 = 1;
  p->m = m;
  p->chirp = (double*)malloc(sizeof(double) * 2 * (size_t)n);
  for (int k = 0; k < n; ++k) {
    const long long k2 = ((long long)k * k) % (2LL * n);
    const double a = -M_PI * (double)k2 / (double)n;
    p->chirp[2 * k] = cos(a);
    p->chirp[2 * k + 1] = sin(a);
  }
  /* bhat_fwd = FFT(conj chirp), bhat_bwd = FFT(chirp) */
  p->bhat[0] = (double*)calloc(2 * (size_t)m, sizeof(double));
  p->bhat[1] = (double*)calloc(2 * (size_t)m, sizeof(double));
  for (int s = 0; s < 2; ++s) {
    const double sg = s == 0 ? -1.0 : 1.0; /* conj for forward */
    double* b = p->bhat[s];
    for (int k = 0; k < n; ++k) {
      b[2 * k] = p->chirp[2 * k];
      b[2 * k + 1] = sg * p->chirp[2 * k + 1];
      if (k > 0) {
        b[2 * (m - k)] = b[2 * k];
        b[2 * (m - k) + 1] = b[2 * k + 1];
      }
    }
    radix2_exec(p->sub, b, -1);
  }
}

void fft1d_free(fft1d* p)
{
  free(p->tw); free(p->rev); free(p->chirp);
  free(p->bhat[0]); free(p->bhat[1]);
  if (p->sub) { fft1d_free(p->sub); free(p->sub); }
  memset(p, 0, sizeof(*p));
}

size_t fft1d_work_size(const fft1d* p) { return p->pow2 ? 0 : 2 * (size_t)p->m; }

void fft1d_exec(const fft1d* p, double* x, int sign, double* work)
{
  if (p->pow2) { radix2_exec(p, x, sign); return; }
  const int n = p->n, m = p->m;
  /* forward: chirp c_k = exp(-i pi k^2/n); a_k = x_k c_k; X = c . (a * conj c) */
  const double cs = sign < 0 ? 1.0 : -1.0; /* backward uses conj chirp */
  memset(work, 0, sizeof(double) * 2 * (size_t)m);
  for (int k = 0; k < n; ++k) {
    const double cr = p->chirp[2 * k], ci = cs * p->chirp[2 * k + 1];
    work[2 * k] = x[2 * k] * cr - x[2 * k + 1] * ci;
    work[2 * k + 1] = x[2 * k] * ci + x[2 * k + 1] * cr;
  }
  radix2_exec(p->sub, work, -1);
  const double* b = p->bhat[sign < 0 ? 0 : 1];
  for (int k = 0; k < m; ++k) {
    const double ar = work[2 * k], ai = work[2 * k + 1];
    work[2 * k] = ar * b[2 * k] - ai * b[2 * k + 1];
    work[2 * k + 1] = ar * b[2 * k + 1] + ai * b[2 * k];
  }
  radix2_exec(p->sub, work, +1);
  const double inv = 1.0 / (double)m;
  for (int k = 0; k < n; ++k) {
    const double cr = p->chirp[2 * k], ci = cs * p->chirp[2 * k + 1];
    const double wr = work[2 * k] * inv, wi = work[2 * k + 1] * inv;
    x[2 * k] = wr * cr - wi * ci;
    x[2 * k + 1] = wr * ci + wi * cr;
  }
}

/* ---------------- 2D real transforms on the FFTW in-place layout -------- */

void fft2d_init(fft2d* p, int MX, int MY, size_t stride)
{
  p->MX = MX; p->MY = MY; p->MXh = MX / 2 + 1; p->stride = stride;
  p->even = (MX % 2) == 0;
  fft1d_init(&p->row, p->even ? MX / 2 : MX);
  fft1d_init(&p->col, MY);
  /* split twiddles for the even real transform: exp(-2 pi i k / MX) */
  p->rtw = (double*)malloc(sizeof(double) * 2 * (size_t)(MX / 2 + 1));
  for (int k = 0; k <= MX / 2; ++k) {
    const double a = -2.0 * M_PI * (double)k / (double)MX;
    p->rtw[2 * k] = cos(a);
    p->rtw[2 * k + 1] = sin(a);
  }
}

void fft2d_free(fft2d* p)
{
  fft1d_free(&p->row); fft1d_free(&p->col); free(p->rtw);
}

/* real row (length MX) -> MX/2+1 complex, in place in a row of 2*MXh reals */
static void r2c_row(const fft2d* p, double* row, double* tmp, double* work)
{
  const int MX = p->MX;
  if (!p->even) {
    for (int i = 0; i < MX; ++i) { tmp[2 * i] = row[i]; tmp[2 * i + 1] = 0; }
    fft1d_exec(&p->row, tmp, -1, work);
    memcpy(row, tmp, sizeof(double) * 2 * (size_t)p->MXh);
    return;
  }
  const int H = MX / 2;
  for (int i = 0; i < H; ++i) { tmp[2 * i] = row[2 * i]; tmp[2 * i + 1] = row[2 * i + 1]; }
  fft1d_exec(&p->row, tmp, -1, work);
  /* X_k = (Z_k + conj Z_{H-k})/2 - i/2 w^k (Z_k - conj Z_{H-k}) */
  for (int k = 0; k <= H; ++k) {
    const int k1 = k % H, k2 = (H - k) % H;
    const double zr = tmp[2 * k1], zi = tmp[2 * k1 + 1];
    const double cr = tmp[2 * k2], ci = -tmp[2 * k2 + 1];
    const double er = 0.5 * (zr + cr), ei = 0.5 * (zi + ci);
    const double or_ = 0.5 * (zr - cr), oi = 0.5 * (zi - ci);
    /* -i * (or + i oi) = oi - i or */
    const double tr = oi, ti = -or_;
    const double wr = p->rtw[2 * k], wi = p->rtw[2 * k + 1];
    row[2 * k] = er + (wr * tr - wi * ti);
    row[2 * k + 1] = ei + (wr * ti + wi * tr);
  }
}

/* MX/2+1 complex (Hermitian half) -> real row of length MX, in place */
static void c2r_row(const fft2d* p, double* row, double* tmp, double* work)
{
  const int MX = p->MX;
  if (!p->even) {
    const int MXh = p->MXh;
    for (int k = 0; k < MXh; ++k) {
      tmp[2 * k] = row[2 * k];
      tmp[2 * k + 1] = k == 0 ? 0.0 : row[2 * k + 1];
    }
    for (int k = MXh; k < MX; ++k) {
      tmp[2 * k] = row[2 * (MX - k)];
      tmp[2 * k + 1] = -row[2 * (MX - k) + 1];
    }
    fft1d_exec(&p->row, tmp, +1, work);
    for (int i = 0; i < MX; ++i) row[i] = tmp[2 * i];
    return;
  }
  const int H = MX / 2;
  /* Z_k = E_k + i O_k with E_k = X_k + conj X_{H-k}, O_k = w^-k (X_k - conj X_{H-k}) */
  for (int k = 0; k < H; ++k) {
    const double xr = row[2 * k], xi = k == 0 ? 0.0 : row[2 * k + 1];
    const double yr = row[2 * (H - k)], yi = (H - k) == H ? 0.0 : -row[2 * (H - k) + 1];
    /* yr + i yi = conj X_{H-k}, with Im of DC/Nyquist dropped */
    const double er = xr + yr, ei = xi + yi;
    const double dr = xr - yr, di = xi - yi;
    const double wr = p->rtw[2 * k], wi = -p->rtw[2 * k + 1];
    const double or_ = wr * dr - wi * di, oi = wr * di + wi * dr;
    tmp[2 * k] = er - oi;
    tmp[2 * k + 1] = ei + or_;
  }
  fft1d_exec(&p->row, tmp, +1, work);
  for (int i = 0; i < H; ++i) { row[2 * i] = tmp[2 * i]; row[2 * i + 1] = tmp[2 * i + 1]; }
}

#define COLB 8

void fft2d_exec(const fft2d* p, double* buf, int sign, int nrows_in)
{
  const int MX = p->MX, MY = p->MY, MXh = p->MXh;
  const size_t stride = p->stride;
  if (nrows_in < 0 || nrows_in > MY) nrows_in = MY;
  if (sign < 0) {
    #pragma omp parallel
    {
      double* tmp = (double*)malloc(sizeof(double) * 2 * (size_t)MX + 16);
      double* work = (double*)malloc(sizeof(double) * (fft1d_work_size(&p->row) + 2));
      #pragma omp for schedule(static)
      for (int j = 0; j < nrows_in; ++j) r2c_row(p, buf + (size_t)j * stride, tmp, work);
      free(tmp); free(work);
    }
  }
  /* columns: complex length MY, MXh columns at complex stride MXh */
  {
    const int ncb = (MXh + COLB - 1) / COLB;
    #pragma omp parallel
    {
      double* col = (double*)malloc(sizeof(double) * 2 * (size_t)MY * COLB);
      double* work = (double*)malloc(sizeof(double) * (fft1d_work_size(&p->col) + 2));
      #pragma omp for schedule(static)
      for (int cb = 0; cb < ncb; ++cb) {
        const int c0 = cb * COLB, nc = (c0 + COLB <= MXh) ? COLB : MXh - c0;
        for (int j = 0; j < MY; ++j) {
          const double* src = buf + (size_t)j * stride + 2 * (size_t)c0;
          for (int c = 0; c < nc; ++c) {
            col[2 * ((size_t)c * MY + j)] = src[2 * c];
            col[2 * ((size_t)c * MY + j) + 1] = src[2 * c + 1];
          }
        }
        for (int c = 0; c < nc; ++c) fft1d_exec(&p->col, col + 2 * (size_t)c * MY, sign, work);
        for (int j = 0; j < MY; ++j) {
          double* dst = buf + (size_t)j * stride + 2 * (size_t)c0;
          for (int c = 0; c < nc; ++c) {
            dst[2 * c] = col[2 * ((size_t)c * MY + j)];
            dst[2 * c + 1] = col[2 * ((size_t)c * MY + j) + 1];
          }
        }
      }
      free(col); free(work);
    }
  }
  if (sign > 0) {
    #pragma omp parallel
    {
      double* tmp = (double*)malloc(sizeof(double) * 2 * (size_t)MX + 16);
      double* work = (double*)malloc(sizeof(double) * (fft1d_work_size(&p->row) + 2));
      #pragma omp for schedule(static)
      for (int j = 0; j < MY; ++j) c2r_row(p, buf + (size_t)j * stride, tmp, work);
      free(tmp); free(work);
    }
  }
}
