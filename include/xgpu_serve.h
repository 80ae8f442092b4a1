/*
 * xgpu_serve.h -- continuous-batching correlation server (no xGPU counterpart).
 *
 * Many subarray streams, each with its own device-resident integration
 * (station count, channel count and block length differ per stream), submit
 * correlation blocks as they arrive.  A scheduler thread admits a block when
 * the HBM staging pool can hold its input (admission by footprint), copies it
 * to the device, and packs the oldest admitted block of each stream into one
 * persistent ragged launch (xgpuAmdCorrelateBatch, up to 16 streams).  A
 * stream's consecutive pending blocks of one integration are fused into one
 * job of that launch (host blocks are staged back to back; device blocks are
 * fused when they already lie back to back in memory), so one kernel pass
 * integrates them and updates the stream's integration buffer once.  While a
 * batch computes, the next batch's inputs are copied and finished streams'
 * dumps are copied back, so the work of concurrent requests overlaps on the
 * device.  After blocks_per_dump blocks a stream's integration is dumped
 * (exact: int64 for the int8 path, double for FP16; triangular order) into a
 * caller ring and restarted.  Block-to-dump latency = completion of the dump
 * on the device minus the arrival (submit) time of the integration's last
 * block; p50/p99 are reported with the sustained CMAC/s.
 *
 * Thread safety: all functions may be called from any thread.  A submitted
 * block (host or device) must stay valid and unchanged until
 * xgpuServerBlocksDone() reports it done (or xgpuServerDrain returns).
 */
#ifndef XGPU_SERVE_H
#define XGPU_SERVE_H

#include <stddef.h>
#include "xgpu_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct XGPUServerStruct XGPUServer;

typedef struct XGPUServerConfigStruct {
  int device;
  size_t staging_bytes;  /* HBM budget for admitted input blocks (0: 4 GiB) */
  int max_batch;         /* jobs (streams) per launch, 1..16 (0: 16) */
  int max_inflight;      /* launches queued on the device at once (0: 2) */
} XGPUServerConfig;

typedef struct XGPUStreamConfigStruct {
  unsigned int nstation;        /* even, >= 2 */
  unsigned int nfrequency;
  unsigned int compute_type;    /* XGPU_INT8 or XGPU_FLOAT16 */
  unsigned int blocks_per_dump; /* K >= 1 */
  unsigned int dump_device;     /* 0: the dump ring is host memory (dumps are copied back);
                                   1: device memory (dumps stay in HBM for a device-side
                                   consumer: no transfer over the host link) */
} XGPUStreamConfig;

typedef struct XGPUServerStatsStruct {
  unsigned long long blocks_submitted;
  unsigned long long blocks_done;
  unsigned long long dumps_done;
  unsigned long long launches;
  unsigned long long cmac_done;     /* complex MACs of completed blocks */
  double busy_s;                    /* first submit -> last completion */
  double tcmac_per_s;               /* cmac_done / busy_s / 1e12 */
  double p50_ms, p99_ms, max_ms;    /* block-to-dump latency over dumps */
  double mean_batch;                /* jobs (streams) per launch */
  double mean_fused;                /* blocks per job (consecutive blocks fused) */
  size_t staging_peak_bytes;        /* high-water mark of admitted inputs */
} XGPUServerStats;

int xgpuServerCreate(XGPUServer **server, const XGPUServerConfig *cfg);
/* Drains, then releases everything (stream engines included). */
int xgpuServerDestroy(XGPUServer *server);

/* Open a stream with a device-resident integration.  dump_ring is host memory
 * (or, with cfg->dump_device, device memory) of dump_slots *
 * xgpuServerDumpBytes() bytes; dump k lands in slot k % dump_slots as
 * interleaved (re, im) int64 (int8) or double (FP16) pairs in
 * TRIANGULAR_ORDER.  xgpuServerDumpsDone(stream) > k means dump k is complete
 * in its slot. */
int xgpuServerOpenStream(XGPUServer *server, const XGPUStreamConfig *cfg, void *dump_ring,
                         unsigned int dump_slots, int *stream_id);
size_t xgpuServerDumpBytes(XGPUServer *server, int stream_id);

/* Enqueue one block of ntime samples ([t][f][s][p] order, host memory) for a
 * stream.  Blocks of one stream are integrated in submission order. */
int xgpuServerSubmit(XGPUServer *server, int stream_id, const void *block_h, unsigned int ntime);
/* The same for a block already in device memory (8-byte aligned; 16-byte
 * aligned blocks run on the real-Gram kernels): no staging copy, and no share
 * of the staging pool.  Consecutive blocks of a stream that follow each other
 * in memory (block k+1 starts where block k ends, e.g. a ring of blocks in
 * HBM) are fused into one job. */
int xgpuServerSubmitDevice(XGPUServer *server, int stream_id, const void *block_d,
                           unsigned int ntime);

/* Wait until every submitted block and every due dump has completed. */
int xgpuServerDrain(XGPUServer *server);

unsigned long long xgpuServerBlocksDone(XGPUServer *server, int stream_id);
unsigned long long xgpuServerDumpsDone(XGPUServer *server, int stream_id);
int xgpuServerStats(XGPUServer *server, XGPUServerStats *stats);
int xgpuServerResetStats(XGPUServer *server);

/* Pinned host memory for blocks and dump rings (fast asynchronous copies). */
int xgpuAmdHostAlloc(void **ptr, size_t bytes);
int xgpuAmdHostFree(void *ptr);

#ifdef __cplusplus
}
#endif

#endif /* XGPU_SERVE_H */
