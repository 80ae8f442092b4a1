/*
 * xgpu_sizing.h -- build-time sizing of the xgpu.h drop-in library.
 *
 * Same make-variable interface and checks as the reference's sizing header
 * (reference src/xgpu_info.h:6-81): NPOL, NSTATION, NFREQUENCY, NTIME,
 * NTIME_PIPE with defaults 2/256/10/1024/128, NTIME a multiple of NTIME_PIPE,
 * and the matrix order chosen by MATRIX_ORDER_TRIANGULAR / _REAL_IMAG /
 * (default) _REGISTER_TILE.  DP4A selects the int8 compute path.
 */
#ifndef XGPU_SIZING_H
#define XGPU_SIZING_H

#include "xgpu.h"

#ifndef NPOL
#define NPOL 2
#endif
#if NPOL != 2
#error NPOL must be 2
#endif

#ifndef NSTATION
#define NSTATION 256
#endif
#ifndef NFREQUENCY
#define NFREQUENCY 10
#endif
#ifndef NTIME
#define NTIME 1024
#endif
#ifndef NTIME_PIPE
#define NTIME_PIPE 128
#endif

#if (NSTATION % 4) != 0
#error NSTATION must be a multiple of 4
#endif
#if (NTIME_PIPE % 4) != 0
#error NTIME_PIPE must be a multiple of 4
#endif
#if (NTIME % NTIME_PIPE) != 0
#error NTIME must be a multiple of NTIME_PIPE
#endif
#define PIPE_LENGTH (NTIME / NTIME_PIPE)

#define NBASELINE ((NSTATION + 1) * (NSTATION / 2))

/* Input layout (reference src/omp_xengine.c:30-34, src/kernel.cuh:212-218):
 * 1 = plain [t][f][s][p] ComplexInput, 32 = blocks of 32 entries holding the
 * real parts of 16 stations' (X, Y) followed by their imaginary parts. */
#ifndef COMPLEX_BLOCK_SIZE
#define COMPLEX_BLOCK_SIZE 1
#endif
#if COMPLEX_BLOCK_SIZE != 1 && COMPLEX_BLOCK_SIZE != 32
#error COMPLEX_BLOCK_SIZE must be 1 or 32
#endif
#if COMPLEX_BLOCK_SIZE == 32 && (NSTATION % 16) != 0
#error COMPLEX_BLOCK_SIZE 32 needs NSTATION to be a multiple of 16
#endif

#if defined MATRIX_ORDER_TRIANGULAR
#define MATRIX_ORDER TRIANGULAR_ORDER
#elif defined MATRIX_ORDER_REAL_IMAG
#define MATRIX_ORDER REAL_IMAG_TRIANGULAR_ORDER
#else
#define MATRIX_ORDER REGISTER_TILE_TRIANGULAR_ORDER
#endif

#define XG_VEC_LENGTH ((unsigned long long)NFREQUENCY * NTIME * NSTATION * NPOL)
#define XG_VEC_LENGTH_PIPE ((unsigned long long)NFREQUENCY * NTIME_PIPE * NSTATION * NPOL)
#define XG_TRI_LENGTH ((unsigned long long)NFREQUENCY * NBASELINE * NPOL * NPOL)
#if MATRIX_ORDER == REGISTER_TILE_TRIANGULAR_ORDER
#define XG_MAT_LENGTH \
  ((unsigned long long)NFREQUENCY * ((NSTATION / 2 + 1) * (NSTATION / 4)) * NPOL * NPOL * 4)
#else
#define XG_MAT_LENGTH XG_TRI_LENGTH
#endif

#endif /* XGPU_SIZING_H */
