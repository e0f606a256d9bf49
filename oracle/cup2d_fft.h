/* cup2d_fft.h -- TEST INFRASTRUCTURE ONLY: the oracle's own FFT (FFTW stand-in). */
#ifndef CUP2D_ORACLE_FFT_H
#define CUP2D_ORACLE_FFT_H
#include <stddef.h>

typedef struct fft1d {
  int n, pow2, m;
  double* tw;           /* radix-2 twiddles, n/2 complex */
  int* rev;             /* bit reversal permutation */
  double* chirp;        /* Bluestein chirp, n complex */
  double* bhat[2];      /* FFT of the convolution kernel (fwd, bwd) */
  struct fft1d* sub;    /* power-of-two sub-plan of length m */
} fft1d;

typedef struct {
  int MX, MY, MXh, even;
  size_t stride;        /* row stride in reals (2*MXh for FFTW in-place) */
  fft1d row, col;
  double* rtw;          /* real-split twiddles */
} fft2d;

void fft1d_init(fft1d* p, int n);
void fft1d_free(fft1d* p);
size_t fft1d_work_size(const fft1d* p);
void fft1d_exec(const fft1d* p, double* x, int sign, double* work);

void fft2d_init(fft2d* p, int MX, int MY, size_t stride);
void fft2d_free(fft2d* p);
/* sign -1: r2c (rows then columns); sign +1: c2r (columns then rows).
 * nrows_in: forward only; rows >= nrows_in are known zero (skip row pass). */
void fft2d_exec(const fft2d* p, double* buf, int sign, int nrows_in);
#endif
