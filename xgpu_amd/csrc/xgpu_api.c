/*
 * xgpu_api.c -- the xgpu.h drop-in entry points (plain C).
 *
 * Sizes are fixed at build time (xgpu_sizing.h).  Each entry point keeps the
 * contract of the reference function it replaces (reference
 * src/cuda_xengine.cu:149-677) and delegates the device work to a runtime-sized
 * engine (engine.c).  Deliberate differences, all invisible to a correct
 * caller:
 *   - caller-supplied host buffers are used in place and never page-locked
 *     (the reference registers them with cudaHostRegister); buffers the
 *     library allocates are pinned, as in the reference;
 *   - no device -> XGPU_CUDA_ERROR instead of exit(-1);
 *   - the device list is printed only when XGPU_VERBOSE is set;
 *   - a DP4A build (int Complex) accumulates exactly, with an int64 spill past
 *     65535 samples; a dump reports the integration modulo 2^32, which is what
 *     int32 accumulation holds.  Like the reference's DP4A build it takes the
 *     input in the layout of xgpuSwizzleInput; the corner turn is undone on
 *     the device before the kernels.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "engine_internal.h"
#include "xg_hip.h"
#include "xgpu.h"
#include "xgpu_amd.h"
#include "xgpu_sizing.h"

#ifndef XGPU_VERSION
#define XGPU_VERSION "xgpu-mi355x-1.0"
#endif

static const XGPUInfo g_info = {
    .npol = NPOL,
    .nstation = NSTATION,
    .nbaseline = NBASELINE,
    .nfrequency = NFREQUENCY,
    .ntime = NTIME,
    .ntimepipe = NTIME_PIPE,
    .input_type = XGPU_INT8,
#ifdef DP4A
    .compute_type = XGPU_INT8,
#else
    .compute_type = XGPU_FLOAT32,
#endif
    .vecLength = XG_VEC_LENGTH,
    .vecLengthPipe = XG_VEC_LENGTH_PIPE,
    .matLength = XG_MAT_LENGTH,
    .triLength = XG_TRI_LENGTH,
    .matrix_order = MATRIX_ORDER,
    .shared_atomic_size = 4,
    .complex_block_size = COMPLEX_BLOCK_SIZE,
};

typedef struct {
  XGPUAmdEngine *engine;
  int device;
  int register_array, register_matrix; /* adopt caller buffers at init */
  int array_set, matrix_set;
  ComplexInput *own_array;             /* pinned, allocated by the library */
  Complex *own_matrix;
  int adopted_array, adopted_matrix;   /* caller buffers in use */
} Internal;

const char *xgpuVersionString() { return XGPU_VERSION; }

void xgpuInfo(XGPUInfo *pcxs) {
  if (pcxs) *pcxs = g_info;
}

static int report(int rc, const char *where) {
  if (rc) fprintf(stderr, "(xgpu) %s: %s\n", where, xgpuAmdLastError());
  return rc;
}

int xgpuSetHostInputBuffer(XGPUContext *context) {
  Internal *in = (Internal *)context->internal;
  if (!in) return XGPU_NOT_INITIALIZED;
  in->array_set = 1;
  xgh_set_device(in->device);
  if (in->own_array) {
    xgh_host_free(in->own_array);
    in->own_array = NULL;
  }
  in->adopted_array = 0;
  if (context->array_h) {
    in->adopted_array = in->register_array;
  } else {
    context->array_len = g_info.vecLength;
    void *p = NULL;
    if (xgh_host_alloc(&p, context->array_len * sizeof(ComplexInput))) {
      fprintf(stderr, "(xgpu) host input allocation: %s\n", xgh_last_error());
      return XGPU_CUDA_ERROR;
    }
    context->array_h = (ComplexInput *)p;
    in->own_array = context->array_h;
  }
  context->input_offset = 0;
  return XGPU_OK;
}

int xgpuSetHostOutputBuffer(XGPUContext *context) {
  Internal *in = (Internal *)context->internal;
  if (!in) return XGPU_NOT_INITIALIZED;
  in->matrix_set = 1;
  xgh_set_device(in->device);
  if (in->own_matrix) {
    xgh_host_free(in->own_matrix);
    in->own_matrix = NULL;
  }
  in->adopted_matrix = 0;
  if (context->matrix_h) {
    in->adopted_matrix = in->register_matrix;
  } else {
    context->matrix_len = g_info.matLength;
    void *p = NULL;
    if (xgh_host_alloc(&p, context->matrix_len * sizeof(Complex))) {
      fprintf(stderr, "(xgpu) host output allocation: %s\n", xgh_last_error());
      return XGPU_CUDA_ERROR;
    }
    context->matrix_h = (Complex *)p;
    in->own_matrix = context->matrix_h;
  }
  context->output_offset = 0;
  return XGPU_OK;
}

int xgpuInit(XGPUContext *context, int device_flags) {
  Internal *in = (Internal *)calloc(1, sizeof(Internal));
  if (!in) return XGPU_OUT_OF_MEMORY;
  context->internal = in;
  in->device = device_flags & XGPU_DEVICE_MASK;
  in->register_array = !(device_flags & XGPU_DONT_REGISTER_ARRAY);
  in->register_matrix = !(device_flags & XGPU_DONT_REGISTER_MATRIX);

  int ndev = 0;
  if (xgh_device_count(&ndev) != 0 || ndev == 0) {
    fprintf(stderr, "(xgpu) no HIP devices found\n");
    free(in);
    context->internal = NULL;
    return XGPU_CUDA_ERROR;
  }
  if (getenv("XGPU_VERBOSE")) {
    char name[256];
    for (int i = 0; i < ndev; ++i)
      if (xgh_device_name(i, name, sizeof name) == 0) printf("Found device %d: %s\n", i, name);
    if (xgh_device_name(in->device, name, sizeof name) == 0)
      printf("Using device %d: %s\n", in->device, name);
  }
  XGPUAmdConfig cfg;
  memset(&cfg, 0, sizeof cfg);
  cfg.npol = NPOL;
  cfg.nstation = NSTATION;
  cfg.nfrequency = NFREQUENCY;
  cfg.ntime = NTIME;
  cfg.ntimepipe = NTIME_PIPE;
#ifdef DP4A
  cfg.compute_type = XGPU_INT8;
#else
  cfg.compute_type = XGPU_FLOAT32;
#endif
  cfg.matrix_order = MATRIX_ORDER;
  cfg.device = in->device;
  cfg.complex_block_size = COMPLEX_BLOCK_SIZE;
#ifdef DP4A
  /* the reference's DP4A build reads input that the caller corner-turned
   * with xgpuSwizzleInput (src/cuda_correlator.c:172-173); the engine undoes
   * the corner turn on the device */
  cfg.time_swizzle = 4;
#endif
  int rc = xgpuAmdCreate(&in->engine, &cfg);
  if (rc) {
    report(rc, "xgpuInit");
    free(in);
    context->internal = NULL;
    return rc == XGPU_OUT_OF_MEMORY ? XGPU_OUT_OF_MEMORY : XGPU_CUDA_ERROR;
  }
  if ((in->register_array && (rc = xgpuSetHostInputBuffer(context)) != XGPU_OK) ||
      (in->register_matrix && (rc = xgpuSetHostOutputBuffer(context)) != XGPU_OK) ||
      (rc = xgpuClearDeviceIntegrationBuffer(context)) != XGPU_OK) {
    /* no half-initialised context: a failed init leaves nothing to free */
    xgpuFree(context);
    return rc;
  }
  return XGPU_OK;
}

int xgpuClearDeviceIntegrationBuffer(XGPUContext *context) {
  Internal *in = (Internal *)context->internal;
  if (!in) return XGPU_NOT_INITIALIZED;
  /* the queued work of earlier calls still reads its own state; the next
   * kernel stores instead of accumulating */
  return xgpuAmdClear(in->engine);
}

void xgpuFree(XGPUContext *context) {
  Internal *in = (Internal *)context->internal;
  if (!in) return;
  xgh_set_device(in->device);
  xgpuAmdDestroy(in->engine);
  if (in->own_array) {
    xgh_host_free(in->own_array);
    context->array_h = NULL;
  }
  if (in->adopted_array) context->array_h = NULL;
  if (in->own_matrix) {
    xgh_host_free(in->own_matrix);
    context->matrix_h = NULL;
  }
  if (in->adopted_matrix) context->matrix_h = NULL;
  free(in);
  context->internal = NULL;
}

int xgpuCudaXengine(XGPUContext *context, int syncOp) {
  Internal *in = (Internal *)context->internal;
  if (!in) return XGPU_NOT_INITIALIZED;
  if (!in->array_set || !in->matrix_set) return XGPU_HOST_BUFFER_NOT_SET;
  int rc = xgpuAmdCorrelateHost(in->engine, context->array_h + context->input_offset, syncOp,
                                context->matrix_h + context->output_offset);
  if (rc) {
    report(rc, "xgpuCudaXengine");
    return rc == XGPU_HOST_BUFFER_NOT_SET ? rc : XGPU_CUDA_ERROR;
  }
  return XGPU_OK;
}
