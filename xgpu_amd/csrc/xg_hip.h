/*
 * xg_hip.h -- thin C-ABI layer between the C host runtime and HIP.
 *
 * The host runtime (engine.c, xgpu_api.c) is plain C compiled by gcc and never
 * includes a HIP header; every device interaction goes through the functions
 * below, implemented in xg_hip.hip (HIP runtime calls) and xg_kernels.hip
 * (kernel launches).  Streams, events and graphs are opaque pointers.
 *
 * All functions return 0 on success and a nonzero HIP error code otherwise;
 * xgh_last_error() gives the message of the most recent failure.
 */
#ifndef XG_HIP_H
#define XG_HIP_H

#include <stddef.h>
#include <stdint.h>
#include "xg_launch.h"

#ifdef __cplusplus
extern "C" {
#endif

const char *xgh_last_error(void);

/* devices */
int xgh_device_count(int *count);
int xgh_set_device(int device);
int xgh_get_device(int *device);
int xgh_device_name(int device, char *buf, size_t len);
int xgh_device_cu_count(int device, int *cus);

/* memory */
int xgh_malloc(void **ptr, size_t bytes);
int xgh_free(void *ptr);
int xgh_host_alloc(void **ptr, size_t bytes);   /* pinned */
int xgh_host_free(void *ptr);
int xgh_memcpy_h2d_async(void *dst, const void *src, size_t bytes, void *stream);
int xgh_memcpy_d2h_async(void *dst, const void *src, size_t bytes, void *stream);
int xgh_memcpy_d2d_async(void *dst, const void *src, size_t bytes, void *stream);
int xgh_memcpy_d2h(void *dst, const void *src, size_t bytes);
int xgh_memcpy_h2d(void *dst, const void *src, size_t bytes);

/* streams and events (default priority only) */
int xgh_stream_create(void **stream);
int xgh_stream_destroy(void *stream);
int xgh_stream_sync(void *stream);
int xgh_stream_wait_event(void *stream, void *event);
int xgh_event_create(void **event, int timing);
int xgh_event_destroy(void *event);
int xgh_event_record(void *event, void *stream);
int xgh_event_sync(void *event);
int xgh_event_query(void *event);              /* 0 done, 1 pending, <0 error */
int xgh_event_elapsed_ms(float *ms, void *start, void *stop);
int xgh_device_sync(void);

/* graphs */
int xgh_graph_begin(void *stream);
int xgh_graph_end(void *stream, void **exec);
int xgh_graph_launch(void *exec, void *stream);
int xgh_graph_destroy(void *exec);

/* the tile the kernels take as the k-th of XCD xcd (CPU only, for tests);
 * 0 past the end of that XCD's share */
int xgh_schedule(const XgLaunch *launch, int xcd, int k, int *job, int *f, int *rb, int *cb);

/* kernels (xg_kernels.hip) */
int xgh_launch_xengine(int family, const XgLaunch *launch, void *stream);
int xgh_launch_fill_zero(void *ptr, size_t bytes, void *stream);
/* plain 16-byte streaming copy (16-byte aligned pointers, bytes % 4 == 0) */
int xgh_launch_copy(void *dst, const void *src, size_t bytes, void *stream);
/* dst[i] += src[i]; kind XG_OUT_I64 (int32 -> int64) or XG_OUT_F64 (f32 -> f64) */
/* dst (int64 / f64) += src (int32 / f32) when add, else dst = src */
int xgh_launch_spill(int kind, int add, void *dst, const void *src, size_t n, void *stream);
/* dst[i] = spill[i] + cur[i] (int64 or f64), or the low 32 bits / f32 rounding
 * when narrow != 0 (dst is then int32 / f32) */
int xgh_launch_combine(int kind, int narrow, void *dst, const void *spill,
                       const void *cur, size_t n, void *stream);
/* bytes of uniform integers in [lo,hi] from the counter-based generator */
int xgh_launch_fill_uniform_i8(void *ptr, size_t n, uint64_t seed, int lo, int hi,
                               void *stream);
/* device reorder of a split-plane REGISTER_TILE or REAL_IMAG buffer into
 * interleaved TRIANGULAR order (elem: 4 = int32/f32, 8 = int64/f64) */
/* Triangular dump of a spilled integration: dst = spill + cur, reordered from
 * split planes (int64 + int32, or double + float, by kind), written wide or,
 * with narrow, as int32 (modulo 2^32) / float. */
int xgh_launch_reorder_combine(int src_order, int kind, int narrow, void *dst, const void *spill,
                               const void *cur, int nstation, int nfreq, long long plane,
                               void *stream);
int xgh_launch_reorder(int src_order, int elem_bytes, void *dst, const void *src,
                       int nstation, int nfreq, long long plane, void *stream);
/* MFMA-only probe: grid workgroups of 256 threads, each wave `iters` x 8
 * independent int8 32x32x32 MFMAs; out holds grid*256 ints */
int xgh_launch_mfma_peak(int grid, int iters, int *out, void *stream);
int xgh_launch_mfma_peak_data(int f16, int grid, int iters, const void *ops, int *out,
                              void *stream);
/* device corner turn between the plain input layout and the depth-4
 * swizzled layout of the dp4a input format, [t/4][f][s][p][c][t%4]:
 * plain -> swizzled, or swizzled -> plain when unswizzle != 0 (ntime % 4 ==
 * 0; swizzled side 16-byte aligned, plain side 4-byte aligned) */
int xgh_launch_swizzle(int unswizzle, void *dst, const void *src, int nstation, int nfreq,
                       int ntime, void *stream);
/* packed triangular (interleaved complex) -> full matrix [f][i][j][p1][p2],
 * element kind XG_OUT_* */
int xgh_launch_extract(int kind, void *dst, const void *packed, int nstation, int nfreq,
                       void *stream);
/* COMPLEX_BLOCK_SIZE=32 input -> plain [t][f][s][p] input, in 64-byte blocks
 * (bytes % 64 == 0, 16-byte aligned pointers; dst may equal src) */
int xgh_launch_unblock32(void *dst, const void *src, size_t bytes, void *stream);

/* Static device scratch of XG_STATIC_SCRATCH_BYTES (a file-scope __device__
 * array, one per device, for callers that must not allocate): its device
 * address on `device`, looked up once per device and cached. */
#define XG_STATIC_SCRATCH_BYTES (64u << 20)
int xgh_static_scratch(int device, void **ptr);

#ifdef __cplusplus
}
#endif

#endif /* XG_HIP_H */
