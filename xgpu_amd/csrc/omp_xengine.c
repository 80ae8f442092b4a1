/*
 * omp_xengine.c -- CPU X-engine entry point of the xgpu.h API.
 *
 * Same contract as the reference CPU X-engine (reference
 * src/omp_xengine.c:36-93): input [t][f][s][p] ComplexInput, output
 * triangular order, OpenMP over (channel, baseline).  The float build
 * accumulates in float like the reference; the DP4A build (int Complex)
 * accumulates exactly in 64-bit integers and stores the result modulo 2^32,
 * where the reference would round the running sum through float once it
 * passes 2^24.
 *
 * The channel's samples are first gathered station-major so the inner loop
 * walks contiguous memory; the arithmetic and its order per visibility
 * (t ascending) are unchanged.  With COMPLEX_BLOCK_SIZE 32 the gather also
 * decodes the blocked input (reference src/omp_xengine.c:64-78): in each
 * block of 32 entries, byte q holds the real part of plain entry q and byte
 * 32 + q its imaginary part.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "xgpu.h"
#include "xgpu_sizing.h"

void xgpuOmpXengine(Complex *matrix_h, ComplexInput *array_h) {
  const size_t nin = (size_t)NSTATION * NPOL;
  for (int f = 0; f < NFREQUENCY; ++f) {
    /* [input][t] copy of channel f */
    ComplexInput *ch = (ComplexInput *)malloc(nin * NTIME * sizeof(ComplexInput));
    if (!ch) return;
#pragma omp parallel for schedule(static)
    for (size_t m = 0; m < nin; ++m)
      for (size_t t = 0; t < NTIME; ++t) {
        const size_t i = (t * NFREQUENCY + f) * nin + m;
#if COMPLEX_BLOCK_SIZE == 32
        const signed char *blk = (const signed char *)array_h + 64 * (i / 32);
        ch[m * NTIME + t].real = blk[i % 32];
        ch[m * NTIME + t].imag = blk[32 + i % 32];
#else
        ch[m * NTIME + t] = array_h[i];
#endif
      }

#pragma omp parallel for schedule(dynamic, 16)
    for (long b = 0; b < (long)NBASELINE; ++b) {
      int i = 0;
      while ((long)(i + 1) * (i + 2) / 2 <= b) ++i;
      const int j = (int)(b - (long)i * (i + 1) / 2);
      Complex *out = &matrix_h[((size_t)f * NBASELINE + (size_t)b) * 4];
      for (int p1 = 0; p1 < NPOL; ++p1)
        for (int p2 = 0; p2 < NPOL; ++p2) {
          const ComplexInput *x = &ch[((size_t)i * NPOL + p1) * NTIME];
          const ComplexInput *y = &ch[((size_t)j * NPOL + p2) * NTIME];
#ifdef DP4A
          int64_t re = 0, im = 0;
          for (int t = 0; t < NTIME; ++t) {
            re += (int)x[t].real * y[t].real + (int)x[t].imag * y[t].imag;
            im += (int)x[t].imag * y[t].real - (int)x[t].real * y[t].imag;
          }
          out[p1 * 2 + p2].real = (int)(uint32_t)(uint64_t)re;
          out[p1 * 2 + p2].imag = (int)(uint32_t)(uint64_t)im;
#else
          float re = 0.f, im = 0.f;
          for (int t = 0; t < NTIME; ++t) {
            re += (float)x[t].real * (float)y[t].real + (float)x[t].imag * (float)y[t].imag;
            im += (float)x[t].imag * (float)y[t].real - (float)x[t].real * (float)y[t].imag;
          }
          out[p1 * 2 + p2].real = re;
          out[p1 * 2 + p2].imag = im;
#endif
        }
    }
    free(ch);
  }
}
