/*
 * xg_launch.h -- launch descriptors shared by the C host runtime and the HIP
 * kernels.  Plain C, no HIP types, so the host side compiles with gcc.
 *
 * One launch correlates up to XG_MAX_JOBS "jobs".  A job is one block of
 * ntime samples of one subarray: input layout [t][f][s][p][re,im]
 * (reference src/omp_xengine.c:57-60), output into that subarray's
 * integration buffer.  Jobs in one launch may have different station counts,
 * channel counts and block lengths (ragged batching).
 *
 * Work unit: a tile = (job, channel f, row block rb, column block cb) with
 * rb >= cb, where a block is XG_TILE consecutive inputs (input = 2*station +
 * pol); the K-interleaved kernel takes pair tiles instead (XgJob.pair_tiles).
 * Tiles are assigned to the 8 XCDs by channel (all tiles of a channel on one
 * XCD, so the channel's input is fetched from HBM into one L2), and
 * each XCD's tiles are dealt round-robin to the persistent workgroups that the
 * hardware places on that XCD (workgroup w runs on XCD w % 8).
 */
#ifndef XG_LAUNCH_H
#define XG_LAUNCH_H

#include <stdint.h>

#define XG_MAX_JOBS 16
#define XG_NXCD 8
#define XG_TILE 128         /* inputs per tile side */
#define XG_WG_THREADS 256   /* 4 wave64 per workgroup */

/* output orders (values are the xgpu.h codes) */
#define XG_ORDER_TRI      1000
#define XG_ORDER_REALIMAG 2000
#define XG_ORDER_REGTILE  3000

/* output element kinds */
#define XG_OUT_I32 0
#define XG_OUT_I64 1
#define XG_OUT_F32 2
#define XG_OUT_F64 3

/* kernel families */
#define XG_KERNEL_I8_MFMA   0
#define XG_KERNEL_F16_MFMA  1
#define XG_KERNEL_F32_PLAIN 2
#define XG_KERNEL_I8_RG     3  /* int8 real-Gram MFMA (nstation % 4 == 0) */
#define XG_KERNEL_F16_RG    4  /* fp16 real-Gram MFMA (16-byte aligned input) */
#define XG_KERNEL_I8_KG     5  /* int8 MFMA with (re, im) interleaved along K, pair tiles
                                  (nstation % 4 == 0, 16-byte aligned input) */

typedef struct XgJob {
  const void *input;   /* sample (t=0,f=0,s=0,p=0) of this block */
  void *out;           /* integration buffer (real plane for split orders) */
  long long tstride;   /* bytes between consecutive time samples */
  long long plane;     /* elements from real plane to imag plane (split orders) */
  int nstation;        /* even, >= 2 */
  int nfreq;
  int ntime;           /* samples in this block */
  int nblk;            /* ceil(2*nstation / XG_TILE) */
  int tiles_per_chan;  /* nblk*(nblk+1)/2 (square tiles) */
  int order;           /* XG_ORDER_* */
  int accumulate;      /* 0: store, 1: add into out */
  int out_kind;        /* XG_OUT_* */
  int chan_base;       /* global channel number of f=0 (XCD assignment) */
  int pair_tiles;      /* 0: square tiles (rb, cb), rb >= cb;  1: pair tiles (rb, cp) of
                          XG_TILE row inputs x 2 XG_TILE column inputs, 2 cp <= rb,
                          tiles_per_chan = xg_pair_count(nblk) */
} XgJob;

/* Pair tiles per channel of the K-interleaved kernel: (rb, cp) with 2 cp < nblk
 * and 2 cp <= rb, nblk - 2 cp of them in column cp (the order xg_pair_order in
 * xg_kernels.hip walks). */
static inline int xg_pair_count(int nblk) {
  int n = 0;
  for (int cp = 0; 2 * cp < nblk; ++cp) n += nblk - 2 * cp;
  return n;
}

typedef struct XgLaunch {
  XgJob jobs[XG_MAX_JOBS];
  int njobs;
  short wg_per_xcd;                             /* grid = 8 * wg_per_xcd */
  char prefetch;                                /* MFMA kernels: L2 prefetch 2 stages ahead */
  char stage;                                   /* real-Gram kernel: 0 / 1 = register
                                                   staging, 64 / 128 samples per stage;
                                                   3 / 4 = global->LDS DMA staging, 64 /
                                                   32 samples per stage; 5 = DMA staging,
                                                   64 samples per stage, 16x16x64 MFMA
                                                   tiles; fp16: 0 / 1 =
                                                   32 / 64 samples per stage, register
                                                   staging; 2 = DMA staging, 32 */
  int xcd_start[XG_NXCD][XG_MAX_JOBS + 1];      /* per-XCD prefix of tile counts */
} XgLaunch;

/* The host (gcc) and device (hipcc) compilers must agree on these layouts:
 * the launch descriptor is passed by value as the kernel argument. */
#ifdef __cplusplus
static_assert(sizeof(XgJob) == 72, "XgJob layout");
static_assert(sizeof(XgLaunch) == 72 * XG_MAX_JOBS + 8 + 4 * XG_NXCD * (XG_MAX_JOBS + 1),
              "XgLaunch layout");
#else
_Static_assert(sizeof(XgJob) == 72, "XgJob layout");
_Static_assert(sizeof(XgLaunch) == 72 * XG_MAX_JOBS + 8 + 4 * XG_NXCD * (XG_MAX_JOBS + 1),
               "XgLaunch layout");
#endif

#endif /* XG_LAUNCH_H */
