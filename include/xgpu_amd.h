/*
 * xgpu_amd.h -- runtime-sized, device-pointer extension of the xgpu.h API.
 *
 * xgpu.h fixes every size when the library is built and moves data only from
 * host buffers.  This header adds (no xGPU counterpart unless cited):
 *   - engines sized at run time (any even station count, any channel count,
 *     any block length), one per subarray, independent of each other;
 *   - correlation of input already resident in device memory (the hot path);
 *   - the three compute paths: exact int8 MFMA, FP16 MFMA, plain FP32;
 *   - long integrations past the int32 bound, spilled to int64 / FP64;
 *   - ragged batches of several engines in one persistent launch;
 *   - a continuous-batching correlation server (xgpu_serve.h).
 *
 * All functions return XGPU_OK (0) or an xgpu.h error code; the message of
 * the last failure on the calling thread is xgpuAmdLastError().
 *
 * Tracing: with the environment variable XGPU_TRACE set ("1" or "stderr":
 * stderr; any other value except "0": a file appended to), the library logs
 * one line per X-engine launch (issued, captured into a hipGraph or replayed),
 * spill and dump: microseconds since the first line, kernel, workgroups,
 * stream and per job the channels, stations, samples, store / accumulate and
 * tiles.  Read once per process.
 */
#ifndef XGPU_AMD_H
#define XGPU_AMD_H

#include <stddef.h>
#include "xgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Additional error codes (xgpu.h uses 0-5). */
#define XGPU_INVALID_CONFIG  (6)   /* unsupported sizes or types */
#define XGPU_NO_DEVICE       (7)   /* no HIP device visible */

/* Dump flags for xgpuAmdDump. */
#define XGPU_DUMP_DEVICE     (1)   /* dst is device memory (default: host) */
#define XGPU_DUMP_WIDE       (2)   /* int64 (int8 path) / double (float paths) */
#define XGPU_DUMP_TRIANGULAR (4)   /* interleaved TRIANGULAR_ORDER instead of the
                                      engine's matrix_order (device reorder) */

/* Copy directions for xgpuAmdMemcpy. */
#define XGPU_H2D 1
#define XGPU_D2H 2
#define XGPU_D2D 3

typedef struct XGPUAmdConfigStruct {
  unsigned int npol;          /* must be 2 */
  unsigned int nstation;      /* even, >= 2; REGISTER_TILE order needs a multiple of 4 */
  unsigned int nfrequency;    /* >= 1 */
  unsigned int ntime;         /* samples per xgpuAmdCorrelateHost call */
  unsigned int ntimepipe;     /* samples per host->device transfer (divides ntime) */
  unsigned int compute_type;  /* XGPU_INT8, XGPU_FLOAT32 or XGPU_FLOAT16 */
  unsigned int matrix_order;  /* TRIANGULAR_ORDER, REAL_IMAG_TRIANGULAR_ORDER or
                                 REGISTER_TILE_TRIANGULAR_ORDER */
  int device;
  unsigned int complex_block_size;  /* input layout: 0 or 1 = plain [t][f][s][p]; 32 = the
                                       COMPLEX_BLOCK_SIZE=32 layout (per 32-entry block, the
                                       real parts of 16 stations' (X, Y), then their
                                       imaginary parts); int8 input, nstation % 16 == 0 */
  unsigned int time_swizzle;        /* 0 = samples in order; 4 = the depth-4 time corner turn
                                       of xgpuSwizzleInput ([t/4][f][s][p][c][t%4], the input
                                       of the reference's DP4A build, src/cpu_util.c:235-255),
                                       applied to the bytes of the layout above; int8 input,
                                       blocks of a multiple of 4 samples, 16-byte aligned */
} XGPUAmdConfig;

typedef struct XGPUAmdEngineStruct XGPUAmdEngine;

/* Sizes a config implies, in XGPUInfo terms (like xgpuInfo, reference
 * src/cuda_xengine.cu:113-142).  Input type: XGPU_INT8 for the int8 and FP32
 * paths (2-byte complex), XGPU_FLOAT16 for the FP16 path (4-byte complex).
 * Output element: int32 (int8 path) or float (FP32/FP16 paths), matLength
 * complex values in split planes (REGISTER_TILE, REAL_IMAG) or interleaved
 * (TRIANGULAR). */
int xgpuAmdInfo(const XGPUAmdConfig *cfg, XGPUInfo *info);

/* Create / destroy an engine: device integration buffer, streams, and the
 * staging buffers of the host-buffer path. */
int xgpuAmdCreate(XGPUAmdEngine **engine, const XGPUAmdConfig *cfg);
void xgpuAmdDestroy(XGPUAmdEngine *engine);

/* Start a new integration (reference xgpuClearDeviceIntegrationBuffer,
 * src/cuda_xengine.cu:334-348).  O(1): the next correlation stores instead of
 * accumulating. */
int xgpuAmdClear(XGPUAmdEngine *engine);

/* Correlate ntime samples of device-resident input laid out
 * [t][f][s][p][re,im] (reference order, src/omp_xengine.c:57-60) into the
 * integration, asynchronously on the engine's stream.  input_d must be
 * 8-byte aligned.  Blocks longer than the int32 bound (int8 path: 65535
 * samples) are split and spilled to int64 transparently. */
int xgpuAmdCorrelate(XGPUAmdEngine *engine, const void *input_d, unsigned int ntime);

/* xgpuAmdCorrelate without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  The conversion of a
 * block-32 or time-swizzled input layout goes through a static scratch of 64
 * MiB per device (XG_STATIC_SCRATCH_BYTES), in chunks of whole 4-sample groups
 * that fit it; plain input needs none.  Refused with XGPU_OUT_OF_MEMORY when
 * the call would cross the int32 bound before the engine's spill buffer
 * exists (an xgpuAmdCorrelate past the bound allocates it), and with
 * XGPU_INVALID_CONFIG when 4 samples of input exceed the scratch.  Every
 * engine on a device shares that device's scratch, which keeps the last
 * call's contents: calls that overlap on one device (different streams, or
 * threads) race.  Same results as xgpuAmdCorrelate. */
int xgpuAmdCorrelateNoAlloc(XGPUAmdEngine *engine, const void *input_d, unsigned int ntime);

/* Host-buffer path with the semantics of xgpuCudaXengine (reference
 * src/cuda_xengine.cu:543-677): copies cfg.ntime samples from input_h in
 * ntimepipe chunks, overlapping transfer and compute, then applies syncOp;
 * SYNCOP_DUMP writes matLength values to output_h. */
int xgpuAmdCorrelateHost(XGPUAmdEngine *engine, const void *input_h, int syncOp,
                         void *output_h);

/* Copy the current integration to dst (flags XGPU_DUMP_*); waits for the
 * engine's queued work.  Without XGPU_DUMP_WIDE the values are int32 (wrapped
 * exactly as int32 accumulation would wrap) or float. */
int xgpuAmdDump(XGPUAmdEngine *engine, void *dst, int flags);

/* Wait for all work queued on the engine. */
int xgpuAmdSynchronize(XGPUAmdEngine *engine);

/* The engine's HIP stream (hipStream_t). */
void *xgpuAmdStream(XGPUAmdEngine *engine);

/* Samples integrated since the last clear. */
unsigned long long xgpuAmdSamples(const XGPUAmdEngine *engine);

/* Message of the last failure on this thread ("" if none). */
const char *xgpuAmdLastError(void);

/* ---- ragged batching ---------------------------------------------------- */
typedef struct XGPUAmdBatchItemStruct {
  XGPUAmdEngine *engine;    /* all items must use one compute type */
  const void *input_d;      /* device input block */
  unsigned int ntime;       /* samples in the block (<= int32 bound) */
} XGPUAmdBatchItem;

/* Correlate up to 16 blocks of different engines (station counts, channel
 * counts, block lengths) in one persistent launch on `stream` (NULL: the first
 * item's engine stream).  Each engine's integration advances as if its block
 * had been correlated alone. */
int xgpuAmdCorrelateBatch(const XGPUAmdBatchItem *items, int nitems, void *stream);

/* ---- device memory and data helpers (no HIP headers needed) ------------- */
int xgpuAmdMalloc(void **ptr, size_t bytes);
int xgpuAmdFree(void *ptr);
int xgpuAmdMemcpy(void *dst, const void *src, size_t bytes, int kind);
int xgpuAmdDeviceSynchronize(void);
/* Fill nbytes of device memory with uniform integers in [lo, hi] from the
 * counter-based generator: byte i = lo + ((hi32(mix64(mix64(seed) + i)) *
 * (hi-lo+1)) >> 32), mix64 = SplitMix64 finaliser. */
int xgpuAmdFillUniformInt8(void *ptr_d, size_t nbytes, unsigned long long seed, int lo, int hi);
/* Measured MFMA-only int8 throughput of `device` in complex MACs per second
 * (4 int8 MACs each): the on-board ceiling the X-engine's rate is compared
 * with.  Best of `reps` launches after a warm-up. */
int xgpuAmdMfmaPeak(int device, int reps, double *cmac_per_s);
/* The same MFMA-only stream with operands taken from `operands_d`, 65536
 * bytes in device memory on `device` (fill them with the input distribution
 * of interest): int8 values for compute_type XGPU_INT8 (int8 MFMA), fp16
 * values for XGPU_FLOAT16 (fp16 MFMA with fp32 accumulation).  The operands change every instruction,
 * and the chip's power management sets the clock by what the datapath
 * toggles.  So this is the MFMA-only ceiling for input of that
 * distribution: the X-engine's rate on such input is compared with it.  One
 * workgroup per CU and one wave per SIMD, as the X-engine.  Best of `reps`
 * launches after a warm-up.  No counterpart in the reference. */
#define XGPU_PEAK_OPERAND_BYTES 65536
int xgpuAmdMfmaPeakData(int device, const void *operands_d, unsigned int compute_type, int reps,
                        double *cmac_per_s);

/* ---- device converters ------------------------------------------------- */
/* Device xgpuSwizzleInput (reference src/cpu_util.c:235-255): ntime samples
 * of int8 input, plain [t][f][s][p][re,im] at src_d -> the depth-4 swizzled
 * layout [t/4][f][s][p][c][t%4] at dst_d; with unswizzle != 0 the inverse.
 * ntime % 4 == 0; the swizzled side 16-byte aligned, the plain side 4-byte
 * aligned; dst_d and src_d do not overlap.  Asynchronous on `stream` (NULL:
 * the null stream). */
int xgpuAmdSwizzleInput(void *dst_d, const void *src_d, unsigned int nstation,
                        unsigned int nfrequency, unsigned int ntime, int unswizzle, void *stream);

/* Element kinds of the device converters */
#define XGPU_ELEM_INT32   0
#define XGPU_ELEM_INT64   1
#define XGPU_ELEM_FLOAT32 2
#define XGPU_ELEM_FLOAT64 3

/* Device xgpuExtractMatrix (reference src/cpu_util.c:258-279): the packed
 * triangular matrix (TRIANGULAR_ORDER, interleaved (re, im), nfrequency *
 * nstation(nstation+1)/2 * 4 complex values, e.g. an XGPU_DUMP_TRIANGULAR
 * dump) -> the full Hermitian matrix [f][i][j][p1][p2] of nfrequency *
 * nstation^2 * 4 complex values, the upper half as conjugates.  Elements of
 * kind XGPU_ELEM_*.  Asynchronous on `stream` (NULL: the null stream). */
int xgpuAmdExtractMatrix(void *matrix_d, const void *packed_d, unsigned int nstation,
                         unsigned int nfrequency, int elem_kind, void *stream);

/* ---- measurement -------------------------------------------------------- */
/* Run xgpuAmdCorrelate(engine, input_d, ntime) warmup + reps times on the
 * engine stream after an xgpuAmdClear each (one kernel per call), timing the
 * reps with HIP events on that stream; *ms = mean device time per call. */
int xgpuAmdTimeCorrelate(XGPUAmdEngine *engine, const void *input_d, unsigned int ntime,
                         int warmup, int reps, float *ms);

/* Kernel tuning of an engine (0 fields = default). */
typedef struct XGPUAmdTuningStruct {
  int wg_per_cu;   /* persistent workgroups per CU: 1 or 2 */
  int prefetch;    /* MFMA kernels: 1 = touch the input 2 stages ahead into L2, 2 = off */
  int graph;       /* host-buffer path: 1 = capture each call's transfers and kernels in
                      a hipGraph once and replay it, 2 = issue them one by one */
  int kernel;      /* int8 and fp16: 1 = complex-plane kernel (any input), 2 = real-Gram
                      kernel where it applies (16-byte aligned input and, for int8,
                      nstation % 4 == 0; default); int8: 3 = K-interleaved kernel
                      (pair tiles of 128 x 256 inputs) where the real-Gram one applies */
  int stage;       /* real-Gram kernel: samples per LDS stage: int8 64 or 128 with
                      register staging, 32 or 64 with DMA staging; fp16 32 or 64 with
                      register staging, 16 or 32 with DMA staging (0 = default: int8
                      DMA 32, fp16 DMA 16) */
  int staging;     /* real-Gram kernel: 1 = global -> VGPR -> LDS, 2 = global -> LDS DMA
                      (0 = default: DMA for int8 and fp16) */
  int ntimepipe;   /* host-buffer path: samples per host -> device transfer and kernel
                      launch (the time-pipe length; divides the config's ntime, a multiple
                      of 4 for time-swizzled input; 0 = the config's ntimepipe) */
  int mfma;        /* int8 real-Gram DMA kernel: MFMA tile shape, 1 = 32x32x32 (default),
                      2 = 16x16x64 (with staging 2, stage 64) (0 = default) */
} XGPUAmdTuning;
int xgpuAmdSetTuning(XGPUAmdEngine *engine, const XGPUAmdTuning *tuning);
int xgpuAmdGetTuning(const XGPUAmdEngine *engine, XGPUAmdTuning *tuning);

/* Autotuner of the device path: time every candidate tuning on ntime samples
 * of input_d with HIP events (warmup + reps calls each), keep the fastest on
 * the engine and return it with its time.  Candidates: the complex-plane
 * kernel x workgroups per CU x L2 prefetch; the real-Gram kernel x staging
 * (registers or global -> LDS DMA) x LDS stage depth, and for the int8 DMA
 * kernel x MFMA tile shape (32x32x32, 16x16x64); the K-interleaved int8
 * kernel.  Candidates run on a scratch engine of the same configuration, so
 * the caller's integration is left as it was.  When cache_path is not NULL
 * the result is appended to that text tuning cache, keyed by device name,
 * station count, channel count, block length and compute type, and an engine
 * created later with the same key picks it up (xgpuAmdLoadTuning). */
int xgpuAmdAutotune(XGPUAmdEngine *engine, const void *input_d, unsigned int ntime, int warmup,
                    int reps, const char *cache_path, XGPUAmdTuning *best, float *best_ms);

/* Autotuner of the host-buffer path (xgpuAmdCorrelateHost): the time-pipe
 * length (samples per transfer and launch, the divisors of the config's
 * ntime from ntime down to ntime/64, at least 32 samples) x hipGraph replay
 * on or off, each timed on the wall clock over warmup + reps calls of
 * input_h with SYNCOP_SYNC_COMPUTE, on a scratch engine of the same
 * configuration and kernel tuning.  Keeps the fastest (ntimepipe, graph) on
 * the engine and returns it with its time per call; cache as above (key
 * marked as the host path). */
int xgpuAmdAutotuneHost(XGPUAmdEngine *engine, const void *input_h, int warmup, int reps,
                        const char *cache_path, XGPUAmdTuning *best, float *best_ms);
/* Apply the tuning cached for this engine's key and ntime; returns
 * XGPU_INVALID_CONFIG (and leaves the engine unchanged) when there is none. */
int xgpuAmdLoadTuning(XGPUAmdEngine *engine, const char *cache_path, unsigned int ntime);

#ifdef __cplusplus
}
#endif

#endif /* XGPU_AMD_H */
