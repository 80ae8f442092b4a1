/*
 * xgpu_fft.h -- batched single-precision complex FFTs on MI355X (gfx950):
 * the F-engine step that produces an X-engine's channelized input.
 *
 * Library: libxgpufft.so (xgpu_amd/lib).  Replaces the reference's FFT
 * prototype (src/fft/FFT{8..8192}.cu, src/fft/IFFT{8..8192}.cu; entry
 * points declared at src/fft/main.cpp:37-55).  Forward: X_k = sum_j x_j
 * exp(-2 pi i jk/N); inverse: the same with +i, unnormalized (a round trip
 * returns N x).  Data: interleaved (re, im) float pairs, in place.
 */
#ifndef XGPU_FFT_H
#define XGPU_FFT_H

#ifdef __cplusplus
extern "C" {
#endif

/* layouts */
#define XGPU_FFT_NATURAL 0   /* batch transforms back to back, natural order in and out */
#define XGPU_FFT_PROTOTYPE 1 /* the reference entry points' layouts (below) */

/* return codes */
#define XGPU_FFT_OK 0
#define XGPU_FFT_BAD_SIZE 1     /* n not a power of two in 8..8192 (prototype: 8, 16, 64, 256..8192) */
#define XGPU_FFT_BAD_BATCH 2    /* prototype layout: batch not a multiple of its group */
#define XGPU_FFT_BAD_ARGUMENT 3 /* NULL / unaligned work, negative batch, unknown layout */
#define XGPU_FFT_ERROR 4        /* HIP error (message on stderr) */

/* In-place batched FFT of `batch` transforms of n points at the device
 * pointer `work` (8-byte aligned), queued on `stream` (a hipStream_t; NULL =
 * default stream).  inverse = 0 forward, 1 inverse.  Each element is read
 * from and written to HBM once.  No counterpart in the reference: the
 * stream-aware, error-returning form of the entry points below. */
int xgpuAmdFft(void *work, int n, long long batch, int inverse, int layout, void *stream);

/* Measurement: average milliseconds of `reps` xgpuAmdFft calls on the
 * default stream (HIP events), after one warm-up call; work is transformed
 * reps + 1 times.  No counterpart in the reference (its harness times with
 * CUDA events, src/fft/main.cpp:14-33). */
int xgpuAmdFftTime(void *work, int n, long long batch, int inverse, int layout, int reps,
                   float *ms);

/* The reference entry points (src/fft/main.cpp:37-55; kernels
 * src/fft/FFT<N>.cu, src/fft/IFFT<N>.cu): in place, default stream, the
 * prototype layouts the reference harness transposes to and from
 * (src/fft/main.cpp:249-264):
 *   N = 8, 16: groups of 64 transforms; element y of transform z at group
 *              offset z + 64 y (batch a multiple of 64);
 *   N = 64:    groups of 8 transforms; element 8a + x of transform t at group
 *              offset 64 a + 8 t + x (batch a multiple of 8);
 *   N = 256 .. 1024: natural;
 *   N >= 2048: forward input natural, forward output with bin y (N/512) + z
 *              at offset 512 z + y; the inverse reads that order and writes
 *              natural order.
 * A rejected call prints the reason on stderr and leaves `work` untouched. */
void FFT8(void *work, int batch);
void FFT16(void *work, int batch);
void FFT64(void *work, int batch);
void FFT256(void *work, int batch);
void FFT512(void *work, int batch);
void FFT1024(void *work, int batch);
void FFT2048(void *work, int batch);
void FFT4096(void *work, int batch);
void FFT8192(void *work, int batch);
void IFFT8(void *work, int batch);
void IFFT16(void *work, int batch);
void IFFT64(void *work, int batch);
void IFFT256(void *work, int batch);
void IFFT512(void *work, int batch);
void IFFT1024(void *work, int batch);
void IFFT2048(void *work, int batch);
void IFFT4096(void *work, int batch);
void IFFT8192(void *work, int batch);

#ifdef __cplusplus
}
#endif

#endif
