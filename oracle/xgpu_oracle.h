/*
 * xgpu_oracle.h -- CPU ORACLE for the X-engine.  TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline may load
 * this library, and only as the checker.  The product library (xgpu_amd/)
 * never links or calls it.
 *
 * Restates, with run-time sizes, the reference's CPU algorithms
 * (reference src/omp_xengine.c, src/cpu_util.c) plus the exact variants the
 * GPU paths are checked against.  How each function is pinned:
 *   oracle_xengine_f32        golden vectors from the compiled reference
 *                             xgpuOmpXengine (tests/golden), brute force
 *   oracle_xengine_dp4a_f32acc golden vectors of the reference built -DDP4A,
 *                             including its float rounding past 2^24
 *   oracle_xengine_i64        numpy int64 brute force, closed forms (constant
 *                             input, single station, -128), agreement with
 *                             the reference below 2^24
 *   oracle_xengine_i64_pairs  equals oracle_xengine_i64 on every pair
 *   oracle_xengine_i64_channels  equals oracle_xengine_i64 on the listed channels
 *   oracle_xengine_f64_f16    numpy complex128 brute force, closed forms
 *   oracle_reorder            golden vectors of the reference xgpuReorderMatrix
 *                             (register tile and real/imag builds)
 *   oracle_regtile_from_tri   inverse of oracle_reorder (round trip) and the
 *                             layout formula of the reference kernel writer
 *   oracle_swizzle            golden vectors of the reference xgpuSwizzleInput
 *   oracle_unblock32          golden vectors of the reference xgpuOmpXengine
 *                             built -DCOMPLEX_BLOCK_SIZE=32 (float and DP4A),
 *                             through oracle_xengine_f32 / _i64 of its output
 *   oracle_extract            golden vectors of the reference xgpuExtractMatrix
 *   oracle_random_legacy      golden vectors of the reference xgpuRandomComplex
 *   oracle_fill_uniform       numpy restatement of the same integer formula
 *   oracle_check              golden text/values of the reference
 *                             xgpuCheckResult on planted errors
 *   oracle_fft_f64            numpy.fft (a library FFT) on random data,
 *                             closed forms (impulse, constant, single tone),
 *                             Parseval, forward/inverse round trip
 */
#ifndef XGPU_ORACLE_H
#define XGPU_ORACLE_H

#include <stddef.h>
#include <stdint.h>

/* Legacy float X-engine: out[2*(((f*NB + k)*2+p1)*2+p2) + {0,1}] (triangular,
 * interleaved), float products and float accumulation in t order
 * (reference src/omp_xengine.c:24-28, 36-93).  nthreads <= 0: all cores. */
void oracle_xengine_f32(float *out, const int8_t *in, int ns, int nf, int nt, int nthreads);

/* The reference compiled with -DDP4A: int accumulators updated through float
 * arithmetic (C compound assignment), reproducing its rounding past 2^24. */
void oracle_xengine_dp4a_f32acc(int32_t *out, const int8_t *in, int ns, int nf, int nt,
                                int nthreads);

/* Exact X-engine in int64, same layout as oracle_xengine_f32. */
void oracle_xengine_i64(int64_t *out, const int8_t *in, int ns, int nf, int nt, int nthreads);

/* Exact int64 visibilities for selected (f, i, j) triples (i >= j):
 * out[(n*4 + p1*2+p2)*2 + {re,im}]. */
void oracle_xengine_i64_pairs(int64_t *out, const int8_t *in, int ns, int nf, int nt,
                              const int32_t *fij, int npairs, int nthreads);

/* Exact int64 visibilities of whole channels chans[0 .. nchan-1]: out holds
 * one channel triangle after another, each laid out as in oracle_xengine_i64. */
void oracle_xengine_i64_channels(int64_t *out, const int8_t *in, int ns, int nf, int nt,
                                 const int32_t *chans, int nchan, int nthreads);

/* FP64 X-engine of FP16 complex input (uint16 halves re,im), same layout. */
void oracle_xengine_f64_f16(double *out, const uint16_t *in, int ns, int nf, int nt,
                            int nthreads);

/* Split-plane REGISTER_TILE (order 3000) or REAL_IMAG (order 2000) buffer of
 * 4- or 8-byte scalars -> interleaved triangular order (reference
 * src/cpu_util.c:57-132).  plane = matLength of the source. */
void oracle_reorder(int order, int elem_bytes, void *tri, const void *src, int ns, int nf,
                    long long plane);

/* Inverse: triangular -> split-plane REGISTER_TILE, quadrant "12" of diagonal
 * tiles zero (reference src/kernel.cuh:49-71). */
void oracle_regtile_from_tri(int elem_bytes, void *dst, const void *tri, int ns, int nf,
                             long long plane);

/* dp4a input corner turn (reference src/cpu_util.c:235-255). */
void oracle_swizzle(int8_t *out, const int8_t *in, int ns, int nf, int nt);

/* COMPLEX_BLOCK_SIZE = 32 input -> plain [t][f][s][p][re,im] input.  The
 * flat ComplexInput array is cut into blocks of 32 entries (64 bytes); entry
 * q of a block in the plain layout (station q/2, polarization q%2 of the
 * block's 16 stations) has its real part at byte q of the block and its
 * imaginary part at byte 32 + q (reference src/omp_xengine.c:64-78).
 * nbytes is a multiple of 64. */
void oracle_unblock32(int8_t *out, const int8_t *in, size_t nbytes);

/* Packed triangular -> full Hermitian matrix, float complex
 * (reference src/cpu_util.c:258-279). */
void oracle_extract(float *full, const float *tri, int ns, int nf);

/* Legacy test data after srand(seed) (reference src/cpu_util.c:13-55). */
void oracle_random_legacy(int8_t *out, size_t ncomplex, unsigned seed);

/* Counter-based uniform integers, identical to the device generator:
 * byte i = lo + ((hi32(mix64(mix64(seed) + i)) * (hi-lo+1)) >> 32). */
void oracle_fill_uniform(int8_t *out, size_t n, uint64_t seed, int lo, int hi);

/* Relative-error check of the reference xgpuCheckResult
 * (src/cpu_util.c:147-232) on triangular float data: returns the number of
 * visibilities above tol and stores the maximum error.  round_gpu mirrors the
 * in-place rounding the reference applies to non-DP4A fixed-point results. */
int oracle_check(const float *gpu, const float *cpu, int ns, int nf, double tol, int round_gpu,
                 double *max_err);

/* Double-precision reference FFT of the reference FFT harness: radix-2
 * Stockham from float input, result in double (reference
 * src/fft/main.cpp:80-117; the harness rounds it to float).  Forward uses
 * exp(-2 pi i jk/n), inverse exp(+2 pi i jk/n), unnormalized.  in/out:
 * batch transforms of n interleaved (re, im) values, natural order; n a
 * power of two. */
void oracle_fft_f64(double *out, const float *in, int n, int batch, int inverse);

#endif /* XGPU_ORACLE_H */
