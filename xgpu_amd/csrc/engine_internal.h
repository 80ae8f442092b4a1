/*
 * engine_internal.h -- private engine state shared by the C host runtime
 * files (engine.c, serve.c).  Not installed.
 */
#ifndef XG_ENGINE_INTERNAL_H
#define XG_ENGINE_INTERNAL_H

#include <stddef.h>
#include "xg_launch.h"
#include "xgpu_amd.h"

struct XGPUAmdEngineStruct {
  XGPUAmdConfig cfg;
  XGPUInfo info;
  int family;        /* XG_KERNEL_* */
  int out_kind;      /* XG_OUT_I32 or XG_OUT_F32 */
  int wide_kind;     /* XG_OUT_I64 or XG_OUT_F64 */
  unsigned bound;    /* samples an integration buffer takes before a spill */
  size_t cplx_bytes; /* 2 (int8) or 4 (fp16) */
  long long tstride; /* input bytes per time sample */
  size_t integ_bytes;
  int cus;
  int wg_per_cu;
  int prefetch;      /* MFMA kernels: L2 prefetch 2 stages ahead */
  int kernel_pref;   /* int8: 1 = complex-plane kernel, 2 = real-Gram when eligible */
  int stage;         /* real-Gram kernel: samples per LDS stage */
  int staging;       /* real-Gram kernel: 1 registers, 2 DMA */
  int mfma;          /* int8 real-Gram DMA kernel: 1 = 32x32x32, 2 = 16x16x64 MFMA tiles */
  unsigned ntp;      /* host-buffer path: samples per transfer and launch */

  void *integ;       /* int32 / f32 integration, 2*matLength scalars */
  void *wide;        /* int64 / f64 spill, allocated on first spill */
  void *scratch;     /* dump staging */
  void *scratch2;

  int cleared;       /* integ is logically zero: next kernel stores */
  int spilled;       /* wide holds part of the integration */
  unsigned samples_in_integ;
  unsigned long long samples;

  void *stream;      /* compute */
  void *copy_stream; /* host -> device transfers */
  void *ev_kernel[2];
  int ev_kernel_used[2];
  void **ev_chunk;
  int npipe_events;
  void *ev_t0, *ev_t1, *ev_aux, *ev_join;
  void *dev_in[2];
  int parity;

  /* COMPLEX_BLOCK_SIZE=32 and / or depth-4 time-swizzled input: converted to
   * the plain layout before the kernels.  Block-32 only: in place in dev_in on
   * the host path, else into `unblock`; swizzled: always into `unblock`
   * (unswizzled there, then unblocked in place) */
  int block32;
  int swizzled;
  void *unblock;
  size_t unblock_bytes;
  /* the device's static scratch (xgh_static_scratch): conversion buffer of
   * xgpuAmdCorrelateNoAlloc, shared by every engine on the device */
  void *static_scratch;

  /* device-path and batch launch steps captured in hipGraphs (graph knob),
   * keyed by everything the step launches; the batch cache lives in the
   * batch's first engine */
  struct XgStepGraph *steps;
  int nsteps;
  int host_graph_off;  /* a host-path capture failed (e.g. pageable source) */

  /* host-buffer path graphs, keyed by (input pointer, cleared state, buffer) */
  int use_graph;
  struct {
    const void *input_h;
    int cleared, buf;
    void *exec;
  } graphs[8];
  int ngraphs;
};

/* default of the L2 prefetch knob of the MFMA kernels */
#define XG_DEFAULT_PREFETCH 0
/* default staging and samples per LDS stage of the int8 real-Gram kernel */
#define XG_DEFAULT_RG_STAGING 2
#define XG_DEFAULT_RG_STAGE 32
/* samples per LDS stage of the fp16 real-Gram kernel with register staging */
#define XG_DEFAULT_RGH_STAGE 32
/* fp16 real-Gram kernel: global -> LDS DMA staging by default (bit-identical to
 * register staging, 0-8 % faster at NS=256 x 2 pol, NF=128, NT=2048) */
#define XG_DEFAULT_RGH_STAGING 2
/* samples per LDS stage of the fp16 DMA kernel: 16 (8 stage buffers; 2-3 %
 * faster than 32 with 4 at NS=256 x 2 pol, NF=128, NT=2048, bit-identical) */
#define XG_DEFAULT_RGDH_STAGE 16

int xg_fail(int code, const char *fmt, ...);
void xg_plan_launch(XgLaunch *L, int cus, int wg_per_cu);
int xg_spill(XGPUAmdEngine *e);
int xg_launch_single(XGPUAmdEngine *e, const void *input_d, unsigned ntime);
/* kernel family that correlates this engine's input at input_d */
int xg_launch_family(const XGPUAmdEngine *e, const void *input_d);
/* queue a dump (flags XGPU_DUMP_*) on the engine stream without waiting */
int xg_dump_async(XGPUAmdEngine *e, void *dst, int flags);
size_t xg_dump_bytes(const XGPUAmdEngine *e, int flags);

#endif /* XG_ENGINE_INTERNAL_H */
