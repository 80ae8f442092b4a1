/*
 * cpu_util.c -- host utilities of the xgpu.h API (plain C, build-time sized).
 *
 * Each function keeps the observable behaviour of its reference counterpart
 * (reference src/cpu_util.c): same random stream for a given srand() seed,
 * same output layouts, same report text and tolerance.  The only intended
 * difference: xgpuReorderMatrix leaves the entries past triLength zero (the
 * reference leaves them unspecified).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "xgpu.h"
#include "xgpu_sizing.h"

/* Legacy test data (reference src/cpu_util.c:13-55): Box-Muller normal
 * deviates with sigma 2.5 from two rand() draws per sample, rounded,
 * clipped to +-7 and scaled by 16, real part from the cosine branch. */
void xgpuRandomComplex(ComplexInput *random_num, long long unsigned int length) {
  const double sigma = 2.5;
  for (long long unsigned int n = 0; n < length; ++n) {
    double u1 = rand() / (double)RAND_MAX;
    double u2 = rand() / (double)RAND_MAX;
    if (u1 == 0.0) u1 = 0.5 / RAND_MAX;
    if (u2 == 0.0) u2 = 0.5 / RAND_MAX;
    const double radius = sigma * sqrt(-2.0 * log(u1));
    const double phase = 2 * M_PI * u2;
    double re = round(radius * cos(phase));
    double im = round(radius * sin(phase));
    re = re > 7.0 ? 7.0 : (re < -7.0 ? -7.0 : re);
    im = im > 7.0 ? 7.0 : (im < -7.0 ? -7.0 : im);
    random_num[n].real = (ReImInput)(((int)re) * 16);
    random_num[n].imag = (ReImInput)(((int)im) * 16);
  }
}

/* Register-tile or real/imag split-plane output -> interleaved triangular
 * order, in place (reference src/cpu_util.c:57-132).  Visibility (f, i>=j,
 * p1, p2) lives at triangular index ((f*NBASELINE + i(i+1)/2 + j)*2+p1)*2+p2. */
void xgpuReorderMatrix(Complex *matrix) {
#if MATRIX_ORDER == TRIANGULAR_ORDER
  (void)matrix;
#else
  const size_t mat_len = (size_t)XG_MAT_LENGTH;
  Complex *tri = (Complex *)calloc(mat_len, sizeof(Complex));
  if (!tri) return;
#ifdef DP4A
  const int *plane = (const int *)matrix;
#else
  const float *plane = (const float *)matrix;
#endif
  for (size_t f = 0; f < NFREQUENCY; ++f)
    for (size_t i = 0; i < NSTATION; ++i)
      for (size_t j = 0; j <= i; ++j) {
        const size_t k = f * NBASELINE + i * (i + 1) / 2 + j;
#if MATRIX_ORDER == REGISTER_TILE_TRIANGULAR_ORDER
        const size_t RT = (NSTATION / 2 + 1) * (NSTATION / 4);
        const size_t R = i / 2, C = j / 2, q = 2 * (j & 1) + (i & 1);
        const size_t src = (f * 4 + q) * RT + R * (R + 1) / 2 + C;
#else
        const size_t src = k;
#endif
        for (size_t p = 0; p < 4; ++p) {
          tri[k * 4 + p].real = plane[src * 4 + p];
          tri[k * 4 + p].imag = plane[src * 4 + p + mat_len];
        }
      }
  memcpy(matrix, tri, mat_len * sizeof(Complex));
  free(tri);
#endif
}

#define XG_TOL 1e-5 /* FIXED_POINT inputs (reference src/cpu_util.c:142-146) */

/* |z| as the reference computes it: the squares and their sum in the
 * element type for float builds (so reported errors match digit for digit),
 * in double for int builds (int squares would overflow). */
#ifndef DP4A
static double cabs_c(Complex z) { return sqrt(z.real * z.real + z.imag * z.imag); }
#else
static double cabs_c(Complex z) { return sqrt((double)z.real * z.real + (double)z.imag * z.imag); }
#endif

/* Per-visibility relative error report (reference src/cpu_util.c:147-232). */
void xgpuCheckResult(Complex *gpu, Complex *cpu, int verbose, ComplexInput *array_h) {
  printf("Checking result (tolerance == %g)...\n", XG_TOL);
  fflush(stdout);
  int nbad = 0;
  double worst = 0.0;
  for (int i = 0; i < NSTATION; ++i)
    for (int j = 0; j <= i; ++j)
      for (int p1 = 0; p1 < NPOL; ++p1)
        for (int p2 = 0; p2 < NPOL; ++p2)
          for (int f = 0; f < NFREQUENCY; ++f) {
            const int k = f * NBASELINE + i * (i + 1) / 2 + j;
            const int idx = (k * NPOL + p1) * NPOL + p2;
#ifndef DP4A
            gpu[idx].real = round(gpu[idx].real);
            gpu[idx].imag = round(gpu[idx].imag);
#endif
            double err;
            if (cabs_c(cpu[idx]) == 0) {
              err = cabs_c(gpu[idx]);
            } else {
              Complex d;
              d.real = cpu[idx].real - gpu[idx].real;
              d.imag = cpu[idx].imag - gpu[idx].imag;
              err = cabs_c(d) / cabs_c(cpu[idx]);
            }
            if (err > worst) worst = err;
            if (err > XG_TOL) {
              if (verbose > 0) {
                if (nbad == 0)
                  printf("freq  i   j    k px py  linear   cpu.real      gpu.real     xcpu.imag      gpu.imag abs(cpu) abs(gpu)\n");
#ifndef DP4A
                printf("%3d %3d %3d %4d %2d %2d %5d %12g  %12g  %12g  %12g (%g %g)\n", f, i, j, k, p1,
                       p2, idx, cpu[idx].real, gpu[idx].real, cpu[idx].imag, gpu[idx].imag,
                       cabs_c(cpu[idx]), cabs_c(gpu[idx]));
#else
                printf("%3d %3d %3d %4d %2d %2d %5d %12d  %12d  %12d  %12d (%g %g)\n", f, i, j, k, p1,
                       p2, idx, cpu[idx].real, gpu[idx].real, cpu[idx].imag, gpu[idx].imag,
                       cabs_c(cpu[idx]), cabs_c(gpu[idx]));
#endif
                if (verbose > 1 && array_h) {
                  double sre = 0, sim = 0;
                  for (int t = 0; t < NTIME; ++t) {
                    const ComplexInput a = array_h[((t * NFREQUENCY + f) * NSTATION + i) * 2 + p1];
                    const ComplexInput b = array_h[((t * NFREQUENCY + f) * NSTATION + j) * 2 + p2];
                    const double pre = a.real * b.real + a.imag * b.imag;
                    const double pim = a.imag * b.real - a.real * b.imag;
                    sre += pre;
                    sim += pim;
                    printf(" %4d (%4g,%4g) (%4g,%4g) -> (%6g, %6g)\n", t, (float)a.real,
                           (float)a.imag, (float)b.real, (float)b.imag, (float)pre, (float)pim);
                  }
#ifndef DP4A
                  printf("                                 (%6g, %6g)\n", sre, sim);
#else
                  printf("                                 (%6d, %6d)\n", (int)sre, (int)sim);
#endif
                }
              }
              nbad++;
            }
          }
  if (nbad)
    printf("Outer product summation failed with %d deviations (max error %g)\n\n", nbad, worst);
  else
    printf("Outer product summation successful (max error %g)\n\n", worst);
}

/* dp4a input layout: components split and time corner-turned by 4
 * (reference src/cpu_util.c:235-255). */
void xgpuSwizzleInput(ComplexInput *out, const ComplexInput *in) {
  printf("Swizzling input\n");
  signed char *o = (signed char *)out;
  const signed char *x = (const signed char *)in;
  const size_t row = (size_t)NFREQUENCY * NSTATION * NPOL * 2; /* bytes per sample */
  for (size_t t = 0; t < NTIME; ++t)
    for (size_t e = 0; e < row; ++e) o[((t / 4) * row + e) * 4 + (t % 4)] = x[t * row + e];
}

/* Packed Hermitian -> full matrix [f][i][j][p1][p2] (reference
 * src/cpu_util.c:258-279). */
void xgpuExtractMatrix(Complex *matrix, Complex *packed) {
  for (size_t f = 0; f < NFREQUENCY; ++f)
    for (size_t i = 0; i < NSTATION; ++i)
      for (size_t j = 0; j <= i; ++j) {
        const size_t k = f * NBASELINE + i * (i + 1) / 2 + j;
        for (size_t p1 = 0; p1 < NPOL; ++p1)
          for (size_t p2 = 0; p2 < NPOL; ++p2) {
            const Complex v = packed[(k * NPOL + p1) * NPOL + p2];
            Complex *lo = &matrix[(((f * NSTATION + i) * NSTATION + j) * NPOL + p1) * NPOL + p2];
            Complex *up = &matrix[(((f * NSTATION + j) * NSTATION + i) * NPOL + p2) * NPOL + p1];
            lo->real = v.real;
            lo->imag = v.imag;
            up->real = v.real;
            up->imag = -v.imag;
          }
      }
}
