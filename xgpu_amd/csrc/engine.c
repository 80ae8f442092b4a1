/*
 * engine.c -- runtime-sized X-engine host runtime (plain C).
 *
 * Owns one subarray's device integration buffer and drives the gfx950 kernels
 * through the thin C-ABI layer (xg_hip.h).  It is the MI355X counterpart of the
 * internal context and pipeline of the reference host code
 * (reference src/cuda_xengine.cu:30-65 context, :180-331 init, :334-348 clear,
 * :543-677 the correlation pipeline and dump), with these differences:
 *   - sizes are run-time values; several engines coexist (no globals);
 *   - one kernel launch integrates a whole device-resident block; the host-
 *     buffer path launches one kernel per transferred chunk, like the
 *     reference, overlapping H2D copies with compute on two streams;
 *   - clear is O(1): the next kernel stores instead of accumulating;
 *   - int32 (or f32) integrations are spilled into an int64 (or f64) buffer
 *     before they could overflow, so integrations of any length are exact.
 */
#include <pthread.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "engine_internal.h"
#include "xg_hip.h"
#include "xg_launch.h"
#include "xgpu_amd.h"

static _Thread_local char g_msg[512];

int xg_fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_msg, sizeof g_msg, fmt, ap);
  va_end(ap);
  return code;
}

static int hipfail(const char *what) {
  return xg_fail(XGPU_CUDA_ERROR, "%s: %s", what, xgh_last_error());
}

#define HIPCALL(expr, what)            \
  do {                                 \
    if ((expr) != 0) return hipfail(what); \
  } while (0)

const char *xgpuAmdLastError(void) { return g_msg; }

/* ---- XGPU_TRACE ----------------------------------------------------------
 * The environment variable XGPU_TRACE, read once per process, switches on a
 * log of the engines' device work: "1" (or "stderr") writes to stderr, any
 * other value except "0" is a file that lines are appended to.  One line per
 * X-engine launch (issued, captured into a hipGraph, or replayed from one),
 * spill and dump, with microseconds since the first line, the kernel, its
 * grid and stream and, per job, the channels, stations, samples, whether it
 * stores or accumulates and its tiles.  It records what was issued and when;
 * kernel durations come from the HIP events of xgpuAmdTimeCorrelate or a
 * profiler.  (The reference's counterpart is its compile-time CUBE macros and
 * TIME_CUDA_CALLS / RUNTIME_STATS switches, reference src/cube/cube.h:301-420,
 * src/cuda_xengine.cu:103-109.) */
static pthread_once_t g_trace_once = PTHREAD_ONCE_INIT;
static FILE *g_trace;
static double g_trace_t0;
static _Thread_local const char *g_trace_how = "launch"; /* "capture" while capturing */

static double mono_us(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return 1e6 * (double)ts.tv_sec + 1e-3 * (double)ts.tv_nsec;
}

static void trace_open(void) {
  const char *v = getenv("XGPU_TRACE");
  if (!v || !*v || strcmp(v, "0") == 0) return;
  g_trace = (strcmp(v, "1") == 0 || strcmp(v, "stderr") == 0) ? stderr : fopen(v, "a");
  g_trace_t0 = mono_us();
}

static FILE *trace_out(void) {
  pthread_once(&g_trace_once, trace_open);
  return g_trace;
}

/* The kernel an X-engine launch runs (the dispatch of xgh_launch_xengine). */
static const char *kernel_label(int family, const XgLaunch *L) {
  switch (family) {
    case XG_KERNEL_I8_MFMA: return "xg_xengine_mfma<int8>";
    case XG_KERNEL_F16_MFMA: return "xg_xengine_mfma<fp16>";
    case XG_KERNEL_F32_PLAIN: return "xg_xengine_f32";
    case XG_KERNEL_I8_KG: return "xg_xengine_kg";
    case XG_KERNEL_I8_RG:
      return L->stage == 3 ? "xg_xengine_rgd<64,4>" : L->stage == 4 ? "xg_xengine_rgd<32,8>"
           : L->stage == 5 ? "xg_xengine_rgd16<4>" : L->stage == 1 ? "xg_xengine_rg<128>"
           : "xg_xengine_rg<64>";
    case XG_KERNEL_F16_RG:
      return L->stage == 3 ? "xg_xengine_rgdh<16,8>" : L->stage == 2 ? "xg_xengine_rgdh<32,4>"
           : L->stage == 1 ? "xg_xengine_rgh<64>" : "xg_xengine_rgh<32>";
    default: return "unknown";
  }
}

static void trace_launch(const char *how, int family, const XgLaunch *L, const void *stream) {
  FILE *f = trace_out();
  if (!f) return;
  char line[160 + 96 * XG_MAX_JOBS];
  int n = snprintf(line, sizeof line, "[xgpu-trace] %12.1f us %s kernel=%s workgroups=%d threads=%d stream=%p jobs=%d",
                   mono_us() - g_trace_t0, how, kernel_label(family, L), 8 * L->wg_per_xcd,
                   XG_WG_THREADS, stream, L->njobs);
  for (int i = 0; i < L->njobs && n < (int)sizeof line - 96; ++i) {
    const XgJob *J = &L->jobs[i];
    n += snprintf(line + n, sizeof line - n, " | nf=%d ns=%d nt=%d %s tiles=%d", J->nfreq,
                  J->nstation, J->ntime, J->accumulate ? "accumulate" : "store",
                  J->tiles_per_chan * J->nfreq);
  }
  fprintf(f, "%s\n", line);
  fflush(f);
}

static void trace_event(const XGPUAmdEngine *e, const char *what, const char *fmt, ...) {
  FILE *f = trace_out();
  if (!f) return;
  char detail[160];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(detail, sizeof detail, fmt, ap);
  va_end(ap);
  fprintf(f, "[xgpu-trace] %12.1f us %s engine=%p stream=%p %s\n", mono_us() - g_trace_t0, what,
          (const void *)e, e->stream, detail);
  fflush(f);
}

/* Largest number of samples an int32 integration can hold for any int8
 * input: |ar*br + ai*bi| <= 2*128*128 per sample, and both accumulators of
 * the kernel (re and the reordered-plane im) stay within 2^31-1 for
 * 65535*32768. */
#define XG_I32_SAMPLE_BOUND 65535u
/* Float paths: spill the f32 integration into f64 every 2^16 samples, keeping
 * f32 partial sums short. */
#define XG_F32_SPILL_SAMPLES 65536u

int xgpuAmdInfo(const XGPUAmdConfig *c, XGPUInfo *info) {
  if (!c || !info) return xg_fail(XGPU_INVALID_CONFIG, "NULL argument");
  if (c->npol != 2) return xg_fail(XGPU_INVALID_CONFIG, "npol must be 2 (got %u)", c->npol);
  if (c->nstation < 2 || (c->nstation & 1))
    return xg_fail(XGPU_INVALID_CONFIG, "nstation must be even and >= 2 (got %u)", c->nstation);
  if (c->nstation > 2 * 255 * XG_TILE / 2)
    return xg_fail(XGPU_INVALID_CONFIG, "nstation too large (%u)", c->nstation);
  if (c->nfrequency < 1) return xg_fail(XGPU_INVALID_CONFIG, "nfrequency must be >= 1");
  if (c->compute_type != XGPU_INT8 && c->compute_type != XGPU_FLOAT32 &&
      c->compute_type != XGPU_FLOAT16)
    return xg_fail(XGPU_INVALID_CONFIG, "unknown compute_type %u", c->compute_type);
  if (c->matrix_order != TRIANGULAR_ORDER && c->matrix_order != REAL_IMAG_TRIANGULAR_ORDER &&
      c->matrix_order != REGISTER_TILE_TRIANGULAR_ORDER)
    return xg_fail(XGPU_INVALID_CONFIG, "unknown matrix_order %u", c->matrix_order);
  if (c->matrix_order == REGISTER_TILE_TRIANGULAR_ORDER && (c->nstation % 4))
    return xg_fail(XGPU_INVALID_CONFIG,
                   "REGISTER_TILE order needs nstation %% 4 == 0 (got %u)", c->nstation);
  if (c->complex_block_size > 1 && c->complex_block_size != 32)
    return xg_fail(XGPU_INVALID_CONFIG, "complex_block_size must be 0, 1 or 32 (got %u)",
                   c->complex_block_size);
  if (c->complex_block_size == 32 && (c->compute_type == XGPU_FLOAT16 || c->nstation % 16))
    return xg_fail(XGPU_INVALID_CONFIG,
                   "complex_block_size 32 needs int8 input and nstation %% 16 == 0");
  if (c->time_swizzle != 0 && c->time_swizzle != 4)
    return xg_fail(XGPU_INVALID_CONFIG, "time_swizzle must be 0 or 4 (got %u)", c->time_swizzle);
  if (c->time_swizzle == 4 && (c->compute_type == XGPU_FLOAT16 || (c->ntimepipe % 4) ||
                               (c->ntime % 4)))
    return xg_fail(XGPU_INVALID_CONFIG,
                   "time_swizzle 4 needs int8 input and ntime, ntimepipe multiples of 4");
  if (c->ntimepipe && c->ntime % c->ntimepipe)
    return xg_fail(XGPU_INVALID_CONFIG, "ntime (%u) must be a multiple of ntimepipe (%u)",
                   c->ntime, c->ntimepipe);
  const unsigned long long ns = c->nstation, nf = c->nfrequency;
  memset(info, 0, sizeof *info);
  info->npol = 2;
  info->nstation = c->nstation;
  info->nbaseline = (unsigned)((ns + 1) * (ns / 2));
  info->nfrequency = c->nfrequency;
  info->ntime = c->ntime;
  info->ntimepipe = c->ntimepipe ? c->ntimepipe : c->ntime;
  info->input_type = (c->compute_type == XGPU_FLOAT16) ? XGPU_FLOAT16 : XGPU_INT8;
  info->compute_type = c->compute_type;
  info->vecLength = nf * c->ntime * ns * 2;
  info->vecLengthPipe = nf * info->ntimepipe * ns * 2;
  info->triLength = nf * ((ns + 1) * (ns / 2)) * 4;
  if (c->matrix_order == REGISTER_TILE_TRIANGULAR_ORDER)
    info->matLength = nf * ((ns / 2 + 1) * (ns / 4)) * 4 * 4;
  else
    info->matLength = info->triLength;
  info->matrix_order = c->matrix_order;
  info->shared_atomic_size = 4;
  info->complex_block_size = c->complex_block_size == 32 ? 32 : 1;
  return XGPU_OK;
}

static void job_fill(XgJob *J, const XGPUAmdEngine *e, const void *input, unsigned ntime,
                     int accumulate, int chan_base) {
  memset(J, 0, sizeof *J);
  J->input = input;
  J->out = e->integ;
  J->tstride = e->tstride;
  J->plane = (long long)e->info.matLength;
  J->nstation = (int)e->cfg.nstation;
  J->nfreq = (int)e->cfg.nfrequency;
  J->ntime = (int)ntime;
  J->nblk = (int)((2 * e->cfg.nstation + XG_TILE - 1) / XG_TILE);
  J->tiles_per_chan = J->nblk * (J->nblk + 1) / 2;
  J->order = (int)e->cfg.matrix_order;
  J->accumulate = accumulate;
  J->out_kind = e->out_kind;
  J->chan_base = chan_base;
}

/* Tile kind of a job for the kernel family that runs it. */
static void job_tiles(XgJob *J, int family) {
  J->pair_tiles = (family == XG_KERNEL_I8_KG);
  J->tiles_per_chan = J->pair_tiles ? xg_pair_count(J->nblk) : J->nblk * (J->nblk + 1) / 2;
}

void xg_plan_launch(XgLaunch *L, int cus, int wg_per_cu) {
  int maxx = 0;
  for (int x = 0; x < XG_NXCD; ++x) {
    L->xcd_start[x][0] = 0;
    for (int j = 0; j < L->njobs; ++j) {
      const XgJob *J = &L->jobs[j];
      const int f0 = (((x - J->chan_base) % XG_NXCD) + XG_NXCD) % XG_NXCD;
      const int nch = (f0 < J->nfreq) ? (J->nfreq - 1 - f0) / XG_NXCD + 1 : 0;
      L->xcd_start[x][j + 1] = L->xcd_start[x][j] + nch * J->tiles_per_chan;
    }
    for (int j = L->njobs + 1; j <= XG_MAX_JOBS; ++j) L->xcd_start[x][j] = L->xcd_start[x][L->njobs];
    if (L->xcd_start[x][L->njobs] > maxx) maxx = L->xcd_start[x][L->njobs];
  }
  int slots = (cus / XG_NXCD) * wg_per_cu;
  if (slots < 1) slots = 1;
  L->wg_per_xcd = (short)(maxx < slots ? maxx : slots);
}

int xgpuAmdCreate(XGPUAmdEngine **out, const XGPUAmdConfig *cfg) {
  if (!out) return xg_fail(XGPU_INVALID_CONFIG, "NULL engine pointer");
  *out = NULL;
  XGPUInfo info;
  int rc = xgpuAmdInfo(cfg, &info);
  if (rc) return rc;
  int ndev = 0;
  if (xgh_device_count(&ndev) != 0 || ndev == 0)
    return xg_fail(XGPU_NO_DEVICE, "no HIP device found (%s)", xgh_last_error());
  if (cfg->device < 0 || cfg->device >= ndev)
    return xg_fail(XGPU_INVALID_CONFIG, "device %d out of range (%d devices)", cfg->device, ndev);
  XGPUAmdEngine *e = (XGPUAmdEngine *)calloc(1, sizeof *e);
  if (!e) return xg_fail(XGPU_OUT_OF_MEMORY, "out of host memory");
  e->cfg = *cfg;
  e->info = info;
  e->family = (cfg->compute_type == XGPU_INT8)      ? XG_KERNEL_I8_MFMA
              : (cfg->compute_type == XGPU_FLOAT16) ? XG_KERNEL_F16_MFMA
                                                    : XG_KERNEL_F32_PLAIN;
  e->out_kind = (e->family == XG_KERNEL_I8_MFMA) ? XG_OUT_I32 : XG_OUT_F32;
  e->wide_kind = (e->family == XG_KERNEL_I8_MFMA) ? XG_OUT_I64 : XG_OUT_F64;
  e->bound = (e->family == XG_KERNEL_I8_MFMA) ? XG_I32_SAMPLE_BOUND : XG_F32_SPILL_SAMPLES;
  e->cplx_bytes = (e->family == XG_KERNEL_F16_MFMA) ? 4 : 2;
  e->block32 = (cfg->complex_block_size == 32);
  e->swizzled = (cfg->time_swizzle == 4);
  e->tstride = (long long)cfg->nfrequency * cfg->nstation * 2 * (long long)e->cplx_bytes;
  e->integ_bytes = (size_t)info.matLength * 2 * 4;
  e->cleared = 1;
  e->ntp = info.ntimepipe;
  e->mfma = 1;
  e->wg_per_cu = 2;
  e->prefetch = XG_DEFAULT_PREFETCH;
  e->kernel_pref = 2;
  if (e->family == XG_KERNEL_F16_MFMA) {
    e->staging = XG_DEFAULT_RGH_STAGING;
    e->stage = e->staging == 2 ? XG_DEFAULT_RGDH_STAGE : XG_DEFAULT_RGH_STAGE;
  } else {
    e->stage = XG_DEFAULT_RG_STAGE;
    e->staging = XG_DEFAULT_RG_STAGING;
  }
  if ((rc = xgh_set_device(cfg->device)) != 0) goto hip_error;
  if ((rc = xgh_device_cu_count(cfg->device, &e->cus)) != 0) goto hip_error;
  if ((rc = xgh_static_scratch(cfg->device, &e->static_scratch)) != 0) goto hip_error;
  if ((rc = xgh_malloc(&e->integ, e->integ_bytes)) != 0) goto hip_error;
  if ((rc = xgh_stream_create(&e->stream)) != 0) goto hip_error;
  if ((rc = xgh_stream_create(&e->copy_stream)) != 0) goto hip_error;
  for (int i = 0; i < 2; ++i)
    if ((rc = xgh_event_create(&e->ev_kernel[i], 0)) != 0) goto hip_error;
  if ((rc = xgh_event_create(&e->ev_t0, 1)) != 0) goto hip_error;
  if ((rc = xgh_event_create(&e->ev_t1, 1)) != 0) goto hip_error;
  if ((rc = xgh_event_create(&e->ev_aux, 0)) != 0) goto hip_error;
  if ((rc = xgh_event_create(&e->ev_join, 0)) != 0) goto hip_error;
  *out = e;
  return XGPU_OK;
hip_error:
  rc = hipfail("xgpuAmdCreate");
  xgpuAmdDestroy(e);
  return rc;
}

static void drop_steps(XGPUAmdEngine *e);

void xgpuAmdDestroy(XGPUAmdEngine *e) {
  if (!e) return;
  xgh_set_device(e->cfg.device);
  if (e->stream) xgh_stream_sync(e->stream);
  if (e->copy_stream) xgh_stream_sync(e->copy_stream);
  for (int i = 0; i < 2; ++i) {
    xgh_free(e->dev_in[i]);
    xgh_event_destroy(e->ev_kernel[i]);
  }
  for (int p = 0; p < e->npipe_events; ++p) xgh_event_destroy(e->ev_chunk[p]);
  free(e->ev_chunk);
  xgh_event_destroy(e->ev_t0);
  xgh_event_destroy(e->ev_t1);
  xgh_event_destroy(e->ev_aux);
  xgh_event_destroy(e->ev_join);
  for (int g = 0; g < e->ngraphs; ++g) xgh_graph_destroy(e->graphs[g].exec);
  drop_steps(e);
  free(e->steps);
  xgh_free(e->integ);
  xgh_free(e->wide);
  xgh_free(e->scratch);
  xgh_free(e->scratch2);
  xgh_free(e->unblock);
  xgh_stream_destroy(e->stream);
  xgh_stream_destroy(e->copy_stream);
  free(e);
}

int xgpuAmdClear(XGPUAmdEngine *e) {
  if (!e) return xg_fail(XGPU_NOT_INITIALIZED, "NULL engine");
  e->cleared = 1;
  e->spilled = 0;
  e->samples = 0;
  e->samples_in_integ = 0;
  return XGPU_OK;
}

unsigned long long xgpuAmdSamples(const XGPUAmdEngine *e) { return e ? e->samples : 0; }
void *xgpuAmdStream(XGPUAmdEngine *e) { return e ? e->stream : NULL; }

static void drop_graphs(XGPUAmdEngine *e);
static void drop_steps(XGPUAmdEngine *e);

int xgpuAmdSetTuning(XGPUAmdEngine *e, const XGPUAmdTuning *t) {
  if (!e || !t) return xg_fail(XGPU_NOT_INITIALIZED, "NULL argument");
  if (t->wg_per_cu < 0 || t->wg_per_cu > 2)
    return xg_fail(XGPU_INVALID_CONFIG, "wg_per_cu must be 0..2");
  if (t->prefetch < 0 || t->prefetch > 2) return xg_fail(XGPU_INVALID_CONFIG, "prefetch must be 0..2");
  if (t->graph < 0 || t->graph > 2) return xg_fail(XGPU_INVALID_CONFIG, "graph must be 0..2");
  if (t->kernel < 0 || t->kernel > 3) return xg_fail(XGPU_INVALID_CONFIG, "kernel must be 0..3");
  if (t->kernel == 3 && e->family != XG_KERNEL_I8_MFMA)
    return xg_fail(XGPU_INVALID_CONFIG, "kernel 3 (K-interleaved) is int8 only");
  if (t->staging < 0 || t->staging > 2) return xg_fail(XGPU_INVALID_CONFIG, "staging must be 0..2");
  int staging, stage;
  if (e->family == XG_KERNEL_F16_MFMA) {
    /* fp16 real-Gram kernel: register staging (32 or 64 samples per stage)
     * or global -> LDS DMA (16 or 32) */
    staging = t->staging ? t->staging : XG_DEFAULT_RGH_STAGING;
    if (staging == 1 && t->stage != 0 && t->stage != 32 && t->stage != 64)
      return xg_fail(XGPU_INVALID_CONFIG, "fp16 stage must be 0, 32 or 64 with register staging");
    if (staging == 2 && t->stage != 0 && t->stage != 16 && t->stage != 32)
      return xg_fail(XGPU_INVALID_CONFIG, "fp16 stage must be 0, 16 or 32 with DMA staging");
    stage = t->stage ? t->stage : (staging == 2 ? XG_DEFAULT_RGDH_STAGE : XG_DEFAULT_RGH_STAGE);
  } else {
    staging = t->staging ? t->staging : XG_DEFAULT_RG_STAGING;
    if (t->stage != 0 && !(staging == 1 && (t->stage == 64 || t->stage == 128)) &&
        !(staging == 2 && (t->stage == 32 || t->stage == 64)))
      return xg_fail(XGPU_INVALID_CONFIG,
                     "stage must be 0, or 64 / 128 with register staging, 32 / 64 with DMA staging");
    stage = t->stage ? t->stage : (staging == XG_DEFAULT_RG_STAGING ? XG_DEFAULT_RG_STAGE : 64);
  }
  if (t->ntimepipe < 0 || (t->ntimepipe > 0 && (e->cfg.ntime % (unsigned)t->ntimepipe ||
                                                 (e->swizzled && t->ntimepipe % 4))))
    return xg_fail(XGPU_INVALID_CONFIG, "ntimepipe %d must divide ntime %u%s", t->ntimepipe,
                   e->cfg.ntime, e->swizzled ? " and be a multiple of 4" : "");
  if (t->mfma < 0 || t->mfma > 2) return xg_fail(XGPU_INVALID_CONFIG, "mfma must be 0..2");
  if (t->mfma == 2 && !(e->family == XG_KERNEL_I8_MFMA && staging == 2 && stage == 64))
    return xg_fail(XGPU_INVALID_CONFIG,
                   "mfma 2 (16x16x64) is the int8 real-Gram DMA kernel's (staging 2, stage 64)");
  /* graphs captured with the old kernel choice would keep replaying it */
  drop_graphs(e);
  e->ntp = t->ntimepipe ? (unsigned)t->ntimepipe : e->info.ntimepipe;
  e->mfma = t->mfma ? t->mfma : 1;
  e->use_graph = (t->graph == 1);
  e->wg_per_cu = t->wg_per_cu ? t->wg_per_cu : 2;
  e->prefetch = t->prefetch == 0 ? XG_DEFAULT_PREFETCH : (t->prefetch == 1);
  e->kernel_pref = t->kernel ? t->kernel : 2;
  e->staging = staging;
  e->stage = stage;
  return XGPU_OK;
}

int xgpuAmdGetTuning(const XGPUAmdEngine *e, XGPUAmdTuning *t) {
  if (!e || !t) return xg_fail(XGPU_NOT_INITIALIZED, "NULL argument");
  memset(t, 0, sizeof *t);
  t->wg_per_cu = e->wg_per_cu;
  t->prefetch = e->prefetch ? 1 : 2;
  t->graph = e->use_graph ? 1 : 2;
  t->kernel = e->kernel_pref;
  t->stage = e->stage;
  t->staging = e->staging;
  t->ntimepipe = (int)e->ntp;
  t->mfma = e->mfma;
  return XGPU_OK;
}

/* wide += integ; integ = 0 (the int32 / f32 part restarts from zero).  The
 * first spill of an integration stores the widened values instead of adding
 * them to a cleared buffer. */
int xg_spill(XGPUAmdEngine *e) {
  const size_t n = e->info.matLength * 2;
  trace_event(e, "spill", "samples=%u %s", e->samples_in_integ,
              e->spilled ? "add" : "first");
  if (!e->wide) HIPCALL(xgh_malloc(&e->wide, n * 8), "spill buffer");
  if (e->cleared) { /* nothing integrated since the clear */
    if (!e->spilled) HIPCALL(xgh_launch_fill_zero(e->wide, n * 8, e->stream), "spill clear");
  } else {
    HIPCALL(xgh_launch_spill(e->wide_kind, e->spilled, e->wide, e->integ, n, e->stream), "spill");
  }
  e->spilled = 1;
  e->cleared = 1; /* integ now logically zero: next kernel stores */
  e->samples_in_integ = 0;
  return XGPU_OK;
}

/* The real-Gram kernels stage 16-byte chunks (8 int8 or 4 fp16 inputs) with
 * 16-byte buffer loads: they need a 16-byte aligned block and, for int8,
 * nstation % 4 == 0 (fp16 rows are 16-byte multiples for every even
 * nstation), and keep every in-stage byte offset (< 128 * tstride) below
 * 2^31. */
int xg_launch_family(const XGPUAmdEngine *e, const void *input_d) {
  if (e->family == XG_KERNEL_F32_PLAIN || e->kernel_pref == 1) return e->family;
  /* these kernels store 32-bit output only (64-bit dumps come from the spill kernels) */
  if (e->out_kind != XG_OUT_I32 && e->out_kind != XG_OUT_F32) return e->family;
  if ((uintptr_t)input_d & 15) return e->family;
  if (e->family == XG_KERNEL_I8_MFMA && e->cfg.nstation % 4) return e->family;
  if (128LL * e->tstride >= (1LL << 31) - 4096) return e->family;
  if (e->family == XG_KERNEL_I8_MFMA) return e->kernel_pref == 3 ? XG_KERNEL_I8_KG : XG_KERNEL_I8_RG;
  return XG_KERNEL_F16_RG;
}

/* XgLaunch.stage code of the real-Gram kernel variant */
static char rg_stage_code(const XGPUAmdEngine *e) {
  if (e->family == XG_KERNEL_F16_MFMA)
    return (char)(e->staging == 2 ? (e->stage == 16 ? 3 : 2) : e->stage == 64 ? 1 : 0);
  if (e->staging == 2) return (char)(e->stage == 32 ? 4 : e->mfma == 2 ? 5 : 3);
  return (char)(e->stage == 128 ? 1 : 0);
}

/* one workgroup per CU: the real-Gram and K-interleaved kernels hold 512
 * registers per lane */
static int wg_per_cu_for(const XGPUAmdEngine *e, int family) {
  return (family == XG_KERNEL_I8_RG || family == XG_KERNEL_F16_RG || family == XG_KERNEL_I8_KG)
             ? 1
             : e->wg_per_cu;
}

/* The launch that correlates ntime samples of plain input at input_d into
 * the engine's integration (store or accumulate by its cleared state). */
static int single_launch(XGPUAmdEngine *e, const void *input_d, unsigned ntime, XgLaunch *L) {
  memset(L, 0, sizeof *L);
  L->njobs = 1;
  const int family = xg_launch_family(e, input_d);
  job_fill(&L->jobs[0], e, input_d, ntime, !e->cleared, 0);
  job_tiles(&L->jobs[0], family);
  xg_plan_launch(L, e->cus, wg_per_cu_for(e, family));
  L->prefetch = (char)e->prefetch;
  L->stage = rg_stage_code(e);
  return family;
}

static void advance(XGPUAmdEngine *e, unsigned ntime) {
  e->cleared = 0;
  e->samples_in_integ += ntime;
  e->samples += ntime;
}

int xg_launch_single(XGPUAmdEngine *e, const void *input_d, unsigned ntime) {
  XgLaunch L;
  const int family = single_launch(e, input_d, ntime, &L);
  trace_launch(g_trace_how, family, &L, e->stream);
  HIPCALL(xgh_launch_xengine(family, &L, e->stream), "xengine kernel");
  advance(e, ntime);
  return XGPU_OK;
}

/* ---- launch steps captured in hipGraphs --------------------------------- */
/* A step: the input-layout conversions of the jobs that need one (raw input
 * -> the job's plain input), then one X-engine launch.  With the graph knob
 * on, a step is captured once and replayed; the key is the whole launch
 * descriptor (buffers, shapes, store or accumulate, kernel and tuning), the
 * raw inputs and the stream, so a replay always repeats exactly the launch
 * an eager call would issue.  Spills are never part of a step. */
#define XG_STEP_GRAPHS 8
struct XgStepGraph {
  XgLaunch L;
  const void *raw[XG_MAX_JOBS];
  int family;
  void *stream;
  void *exec;
};

static void drop_steps(XGPUAmdEngine *e) {
  for (int g = 0; g < e->nsteps; ++g) xgh_graph_destroy(e->steps[g].exec);
  e->nsteps = 0;
}

/* Forget the captured graphs (they hold buffer pointers and the kernel
 * tuning of the moment they were captured). */
static void drop_graphs(XGPUAmdEngine *e) {
  for (int g = 0; g < e->ngraphs; ++g) xgh_graph_destroy(e->graphs[g].exec);
  e->ngraphs = 0;
  drop_steps(e);
}

static int convert_input(XGPUAmdEngine *e, void *dst, const void *src, unsigned ntime,
                         void *stream);

/* Issue a step eagerly: conversions (engines[i] converts when raw[i] differs
 * from the job's input), then the launch. */
static int issue_step(XGPUAmdEngine *const *engines, const void *const *raw, const XgLaunch *L,
                      int family, void *s, const char *how) {
  for (int i = 0; i < L->njobs; ++i)
    if (raw[i] != L->jobs[i].input && L->jobs[i].ntime > 0) {
      int rc = convert_input(engines[i], (void *)L->jobs[i].input, raw[i], (unsigned)L->jobs[i].ntime, s);
      if (rc) return rc;
    }
  trace_launch(how, family, L, s);
  HIPCALL(xgh_launch_xengine(family, L, s), "xengine kernel");
  return XGPU_OK;
}

/* Run a step on stream s: replay owner's captured graph of it, capturing it
 * first when the owner's graph knob is on; otherwise issue it eagerly. */
static int run_step(XGPUAmdEngine *owner, XGPUAmdEngine *const *engines, const void *const *raw,
                    const XgLaunch *L, int family, void *s) {
  if (!owner->use_graph) return issue_step(engines, raw, L, family, s, "launch");
  if (!owner->steps) {
    owner->steps = (struct XgStepGraph *)calloc(XG_STEP_GRAPHS, sizeof *owner->steps);
    if (!owner->steps) return xg_fail(XGPU_OUT_OF_MEMORY, "out of host memory");
  }
  int g = -1;
  for (int i = 0; i < owner->nsteps && g < 0; ++i) {
    const struct XgStepGraph *t = &owner->steps[i];
    if (t->family == family && t->stream == s && memcmp(&t->L, L, sizeof *L) == 0 &&
        memcmp(t->raw, raw, sizeof(void *) * (size_t)L->njobs) == 0)
      g = i;
  }
  if (g < 0) {
    if (owner->nsteps == XG_STEP_GRAPHS) { /* drop the oldest */
      xgh_graph_destroy(owner->steps[0].exec);
      memmove(&owner->steps[0], &owner->steps[1], sizeof *owner->steps * (XG_STEP_GRAPHS - 1));
      owner->nsteps--;
    }
    HIPCALL(xgh_graph_begin(s), "begin capture");
    const int rc = issue_step(engines, raw, L, family, s, "capture");
    void *exec = NULL;
    const int rc_end = xgh_graph_end(s, &exec);
    if (rc || rc_end) {
      xgh_graph_destroy(exec);
      return rc ? rc : xg_fail(XGPU_CUDA_ERROR, "graph capture: %s", xgh_last_error());
    }
    g = owner->nsteps++;
    struct XgStepGraph *t = &owner->steps[g];
    memcpy(&t->L, L, sizeof *L);
    memset(t->raw, 0, sizeof t->raw);
    memcpy(t->raw, raw, sizeof(void *) * (size_t)L->njobs);
    t->family = family;
    t->stream = s;
    t->exec = exec;
  }
  trace_launch("replay", family, L, s);
  HIPCALL(xgh_graph_launch(owner->steps[g].exec, s), "graph launch");
  return XGPU_OK;
}

/* The engine's conversion buffer, at least `bytes` long. */
static int ensure_unblock(XGPUAmdEngine *e, size_t bytes) {
  if (bytes <= e->unblock_bytes) return XGPU_OK;
  if (e->unblock) { /* queued kernels, and captured graphs, may still read the old buffer */
    HIPCALL(xgh_device_sync(), "sync");
    drop_graphs(e);
    xgh_free(e->unblock);
    e->unblock = NULL;
    e->unblock_bytes = 0;
  }
  HIPCALL(xgh_malloc(&e->unblock, bytes), "input conversion buffer");
  e->unblock_bytes = bytes;
  return XGPU_OK;
}

/* Queue on `stream` the conversion of ntime samples of this engine's input
 * layout at src (device memory) into the plain layout at dst: unswizzle the
 * depth-4 time corner turn, then undo the block-32 layout (in place). */
static int convert_input(XGPUAmdEngine *e, void *dst, const void *src, unsigned ntime,
                         void *stream) {
  const size_t bytes = (size_t)ntime * (size_t)e->tstride;
  if (e->swizzled) {
    if (ntime % 4) return xg_fail(XGPU_INVALID_CONFIG, "swizzled input: ntime (%u) %% 4 != 0", ntime);
    if (((uintptr_t)src) & 15) return xg_fail(XGPU_INVALID_CONFIG, "swizzled input must be 16-byte aligned");
    HIPCALL(xgh_launch_swizzle(1, dst, src, (int)e->cfg.nstation, (int)e->cfg.nfrequency, (int)ntime,
                               stream),
            "input unswizzle");
    src = dst;
  }
  if (e->block32) {
    if (((uintptr_t)src) & 15) return xg_fail(XGPU_INVALID_CONFIG, "block-32 input must be 16-byte aligned");
    HIPCALL(xgh_launch_unblock32(dst, src, bytes, stream), "block-32 conversion");
  }
  return XGPU_OK;
}

/* Convert ntime samples of device input in the engine's layout into the
 * engine's plain-layout buffer on `stream`; *plain points at the result. */
static int unblock_input(XGPUAmdEngine *e, const void *input_d, unsigned ntime, void *stream,
                         const void **plain) {
  int rc = ensure_unblock(e, (size_t)ntime * (size_t)e->tstride);
  if (rc) return rc;
  if ((rc = convert_input(e, e->unblock, input_d, ntime, stream))) return rc;
  *plain = e->unblock;
  return XGPU_OK;
}

/* Correlate plain-layout device input (splitting at the integration bound).
 * A call that fits a fresh integration but not the room left spills first
 * and runs whole: splitting it would cost one more launch, whose sweep over
 * every tile and read-modify-write of the whole output take as long for one
 * sample as for hundreds. */
static int correlate_plain(XGPUAmdEngine *e, const void *input_d, unsigned int ntime) {
  const char *p = (const char *)input_d;
  while (ntime > 0) {
    unsigned room = e->bound - e->samples_in_integ;
    if (room == 0 || (room < ntime && ntime <= e->bound && e->samples_in_integ > 0)) {
      int rc = xg_spill(e);
      if (rc) return rc;
      continue;
    }
    unsigned n = ntime < room ? ntime : room;
    int rc = xg_launch_single(e, p, n);
    if (rc) return rc;
    p += (size_t)n * (size_t)e->tstride;
    ntime -= n;
  }
  return XGPU_OK;
}

int xgpuAmdCorrelate(XGPUAmdEngine *e, const void *input_d, unsigned int ntime) {
  if (!e) return xg_fail(XGPU_NOT_INITIALIZED, "NULL engine");
  if (!input_d && ntime) return xg_fail(XGPU_INVALID_CONFIG, "NULL input");
  if (((uintptr_t)input_d) & 7) return xg_fail(XGPU_INVALID_CONFIG, "input must be 8-byte aligned");
  HIPCALL(xgh_set_device(e->cfg.device), "set device");
  int rc;
  if (e->use_graph && ntime > 0 && e->samples_in_integ + ntime <= e->bound) {
    /* one step (conversion + launch), captured once and replayed */
    const void *plain = input_d;
    if (e->block32 || e->swizzled) {
      if ((rc = ensure_unblock(e, (size_t)ntime * (size_t)e->tstride))) return rc;
      plain = e->unblock;
    }
    XgLaunch L;
    const int family = single_launch(e, plain, ntime, &L);
    if ((rc = run_step(e, &e, &input_d, &L, family, e->stream))) return rc;
    advance(e, ntime);
    return XGPU_OK;
  }
  if ((e->block32 || e->swizzled) && ntime) {
    const void *plain;
    if ((rc = unblock_input(e, input_d, ntime, e->stream, &plain))) return rc;
    input_d = plain;
  }
  return correlate_plain(e, input_d, ntime);
}

/* xgpuAmdCorrelate without allocating: the input-layout conversion goes
 * through the device's static scratch, in chunks of whole 4-sample groups
 * that fit it, and the call is refused when it would need the spill buffer
 * before it exists.  The scratch is shared by every engine on the device, so
 * calls on different streams must not overlap. */
int xgpuAmdCorrelateNoAlloc(XGPUAmdEngine *e, const void *input_d, unsigned int ntime) {
  if (!e) return xg_fail(XGPU_NOT_INITIALIZED, "NULL engine");
  if (!input_d && ntime) return xg_fail(XGPU_INVALID_CONFIG, "NULL input");
  if (((uintptr_t)input_d) & 7) return xg_fail(XGPU_INVALID_CONFIG, "input must be 8-byte aligned");
  if (!e->wide && (unsigned long long)e->samples_in_integ + ntime > e->bound)
    return xg_fail(XGPU_OUT_OF_MEMORY,
                   "this call passes the int32 bound (%u samples) and needs the spill buffer, "
                   "which xgpuAmdCorrelateNoAlloc does not allocate", e->bound);
  HIPCALL(xgh_set_device(e->cfg.device), "set device");
  if (!(e->block32 || e->swizzled) || ntime == 0) return correlate_plain(e, input_d, ntime);
  const unsigned unit = e->swizzled ? 4u : 1u;
  const size_t unit_bytes = (size_t)unit * (size_t)e->tstride;
  if (unit_bytes > XG_STATIC_SCRATCH_BYTES)
    return xg_fail(XGPU_INVALID_CONFIG,
                   "%u samples of %lld bytes exceed the static scratch (%u bytes)", unit,
                   e->tstride, XG_STATIC_SCRATCH_BYTES);
  const unsigned chunk = (unsigned)(XG_STATIC_SCRATCH_BYTES / unit_bytes) * unit;
  const char *src = (const char *)input_d;
  for (unsigned done = 0; done < ntime;) {
    const unsigned n = ntime - done < chunk ? ntime - done : chunk;
    int rc = convert_input(e, e->static_scratch, src + (size_t)done * (size_t)e->tstride, n, e->stream);
    if (rc) return rc;
    if ((rc = correlate_plain(e, e->static_scratch, n))) return rc;
    done += n;
  }
  return XGPU_OK;
}

static int ensure_scratch(XGPUAmdEngine *e, void **buf, size_t bytes) {
  if (*buf) return XGPU_OK;
  HIPCALL(xgh_malloc(buf, bytes), "dump scratch");
  return XGPU_OK;
}

/* Stage the integration for a dump (flags XGPU_DUMP_*): combine spilled
 * parts, widen, and reorder to triangular order as the flags ask, all queued
 * on the engine stream.  When `direct` is non-NULL (a device destination),
 * the last pass writes there and *out is set to NULL, so no extra copy
 * follows; otherwise *out names the staged device buffer and *bytes its size. */
static int dump_stage(XGPUAmdEngine *e, int flags, void *direct, const void **out, size_t *bytes) {
  const int wide = (flags & XGPU_DUMP_WIDE) != 0;
  const int tri = (flags & XGPU_DUMP_TRIANGULAR) && e->cfg.matrix_order != TRIANGULAR_ORDER;
  const size_t n = e->info.matLength * 2; /* scalars in the integration */
  const size_t esz = wide ? 8 : 4;
  const void *val = NULL;
  int rc;
  if (tri && e->spilled && !e->cleared) { /* spill + current part, reordered in one pass */
    void *target = direct;
    if (!target) {
      if ((rc = ensure_scratch(e, &e->scratch2, n * 8))) return rc;
      target = e->scratch2;
    }
    HIPCALL(xgh_launch_reorder_combine((int)e->cfg.matrix_order, e->wide_kind, !wide, target,
                                       e->wide, e->integ, (int)e->cfg.nstation,
                                       (int)e->cfg.nfrequency, (long long)e->info.matLength,
                                       e->stream),
            "dump reorder combine");
    *out = direct ? NULL : target;
    *bytes = e->info.triLength * 2 * esz;
    return XGPU_OK;
  }
  /* Without a reorder, a device destination takes the combined, widened or
   * zeroed values straight from the kernel that makes them (no copy after). */
  void *stage = (!tri && direct) ? direct : NULL;
  if (e->cleared && !e->spilled) {
    if (!stage) {
      if ((rc = ensure_scratch(e, &e->scratch, n * 8))) return rc;
      stage = e->scratch;
    }
    HIPCALL(xgh_launch_fill_zero(stage, n * esz, e->stream), "dump zero");
    val = stage;
  } else if (!wide && !e->spilled) {
    val = e->integ;
    stage = NULL;
  } else {
    if (!stage) {
      if ((rc = ensure_scratch(e, &e->scratch, n * 8))) return rc;
      stage = e->scratch;
    }
    if (e->spilled && !e->cleared) {
      HIPCALL(xgh_launch_combine(e->wide_kind, !wide, stage, e->wide, e->integ, n, e->stream),
              "dump combine");
    } else if (e->spilled) { /* only the spilled part is live */
      if (wide) {
        HIPCALL(xgh_memcpy_d2d_async(stage, e->wide, n * 8, e->stream), "dump copy");
      } else {
        if ((rc = ensure_scratch(e, &e->scratch2, n * 8))) return rc;
        HIPCALL(xgh_launch_fill_zero(e->scratch2, n * 4, e->stream), "dump zero");
        HIPCALL(xgh_launch_combine(e->wide_kind, 1, stage, e->wide, e->scratch2, n, e->stream),
                "dump combine");
      }
    } else { /* wide requested, nothing spilled: widen integ */
      HIPCALL(xgh_launch_spill(e->wide_kind, 0, stage, e->integ, n, e->stream), "dump widen");
    }
    val = stage;
  }
  if (stage && stage == direct) { /* written in place */
    *out = NULL;
    *bytes = n * esz;
    return XGPU_OK;
  }
  if (tri) {
    void *target = direct;
    if (!target) {
      if ((rc = ensure_scratch(e, &e->scratch2, n * 8))) return rc;
      target = e->scratch2;
    }
    HIPCALL(xgh_launch_reorder((int)e->cfg.matrix_order, (int)esz, target, val,
                               (int)e->cfg.nstation, (int)e->cfg.nfrequency,
                               (long long)e->info.matLength, e->stream),
            "dump reorder");
    val = direct ? NULL : target;
  }
  *out = val;
  *bytes = (tri ? e->info.triLength * 2 : n) * esz;
  return XGPU_OK;
}

int xg_dump_async(XGPUAmdEngine *e, void *dst, int flags) {
  const void *val = NULL;
  size_t bytes = 0;
  trace_event(e, "dump", "dst=%p%s%s%s samples=%llu", dst, (flags & XGPU_DUMP_DEVICE) ? " device" : " host",
              (flags & XGPU_DUMP_WIDE) ? " wide" : "", (flags & XGPU_DUMP_TRIANGULAR) ? " triangular" : "",
              (unsigned long long)e->samples);
  /* A dump to device memory is written straight into dst by the kernel that
   * makes it (reorder, combine, widen or zero); only a plain copy of the
   * integration goes through a memcpy. */
  int rc = dump_stage(e, flags, (flags & XGPU_DUMP_DEVICE) ? dst : NULL, &val, &bytes);
  if (rc) return rc;
  if (!val) return XGPU_OK;
  if (flags & XGPU_DUMP_DEVICE)
    HIPCALL(xgh_memcpy_d2d_async(dst, val, bytes, e->stream), "dump D2D");
  else
    HIPCALL(xgh_memcpy_d2h_async(dst, val, bytes, e->stream), "dump D2H");
  return XGPU_OK;
}

size_t xg_dump_bytes(const XGPUAmdEngine *e, int flags) {
  const int tri = (flags & XGPU_DUMP_TRIANGULAR) && e->cfg.matrix_order != TRIANGULAR_ORDER;
  return (tri ? e->info.triLength * 2 : e->info.matLength * 2) * ((flags & XGPU_DUMP_WIDE) ? 8 : 4);
}

int xgpuAmdDump(XGPUAmdEngine *e, void *dst, int flags) {
  if (!e) return xg_fail(XGPU_NOT_INITIALIZED, "NULL engine");
  if (!dst) return xg_fail(XGPU_INVALID_CONFIG, "NULL dump destination");
  HIPCALL(xgh_set_device(e->cfg.device), "set device");
  int rc = xg_dump_async(e, dst, flags);
  if (rc) return rc;
  HIPCALL(xgh_stream_sync(e->stream), "dump sync");
  return XGPU_OK;
}

int xgpuAmdSynchronize(XGPUAmdEngine *e) {
  if (!e) return xg_fail(XGPU_NOT_INITIALIZED, "NULL engine");
  HIPCALL(xgh_set_device(e->cfg.device), "set device");
  HIPCALL(xgh_stream_sync(e->copy_stream), "sync copy stream");
  HIPCALL(xgh_stream_sync(e->stream), "sync compute stream");
  return XGPU_OK;
}

/* Host-buffer path (reference src/cuda_xengine.cu:543-677).  Two device
 * input buffers alternate between calls; within a call, chunk p's transfer on
 * the copy stream is followed by its kernel on the compute stream, so the
 * transfer of chunk p+1 overlaps the kernel of chunk p. */
/* Issue one host-path call: per chunk, H2D on the copy stream, then the
 * kernel of that chunk on the compute stream. */
static int host_issue(XGPUAmdEngine *e, const void *input_h, char *dbase, int npipe, unsigned ntp,
                      size_t chunk) {
  const char *hbase = (const char *)input_h;
  for (int p = 0; p < npipe; ++p) {
    HIPCALL(xgh_memcpy_h2d_async(dbase + p * chunk, hbase + p * chunk, chunk, e->copy_stream), "H2D");
    HIPCALL(xgh_event_record(e->ev_chunk[p], e->copy_stream), "record");
    HIPCALL(xgh_stream_wait_event(e->stream, e->ev_chunk[p]), "wait");
    const char *in = dbase + p * chunk;
    int rc;
    if (e->swizzled) { /* into the conversion buffer, sized for the whole call */
      if ((rc = convert_input(e, (char *)e->unblock + p * chunk, in, ntp, e->stream))) return rc;
      in = (const char *)e->unblock + p * chunk;
    } else if (e->block32) { /* the library's own buffer: convert in place */
      if ((rc = convert_input(e, (void *)in, in, ntp, e->stream))) return rc;
    }
    if ((rc = correlate_plain(e, in, ntp))) return rc;
  }
  return XGPU_OK;
}

/* Host-buffer path (reference src/cuda_xengine.cu:543-677).  Two device
 * input buffers alternate between calls; within a call, chunk p's transfer on
 * the copy stream is followed by its kernel on the compute stream, so the
 * transfer of chunk p+1 overlaps the kernel of chunk p.  With the graph knob
 * on, the call is captured once per (input pointer, clear state, buffer) into
 * a two-stream hipGraph and replayed on the compute stream; calls that would
 * cross a spill are issued eagerly. */
int xgpuAmdCorrelateHost(XGPUAmdEngine *e, const void *input_h, int syncOp, void *output_h) {
  if (!e) return xg_fail(XGPU_NOT_INITIALIZED, "NULL engine");
  if (!input_h) return xg_fail(XGPU_HOST_BUFFER_NOT_SET, "NULL host input");
  if (syncOp == SYNCOP_DUMP && !output_h) return xg_fail(XGPU_HOST_BUFFER_NOT_SET, "NULL host output");
  if (e->cfg.ntime == 0) return XGPU_OK;
  HIPCALL(xgh_set_device(e->cfg.device), "set device");
  const unsigned ntp = e->ntp;
  const int npipe = (int)(e->cfg.ntime / ntp);
  const size_t chunk = (size_t)ntp * (size_t)e->tstride;
  const int buf = e->parity;
  e->parity ^= 1;
  if (!e->dev_in[buf]) HIPCALL(xgh_malloc(&e->dev_in[buf], chunk * npipe), "input buffer");
  if (e->npipe_events < npipe) {
    void **ev = (void **)realloc(e->ev_chunk, sizeof(void *) * npipe);
    if (!ev) return xg_fail(XGPU_OUT_OF_MEMORY, "out of host memory");
    e->ev_chunk = ev;
    for (int p = e->npipe_events; p < npipe; ++p) {
      HIPCALL(xgh_event_create(&e->ev_chunk[p], 0), "event");
      e->npipe_events = p + 1;
    }
  }
  char *dbase = (char *)e->dev_in[buf];
  int rc;
  if (e->swizzled && (rc = ensure_unblock(e, chunk * npipe))) return rc;
  const int graphable = e->use_graph && !e->host_graph_off &&
                        e->samples_in_integ + e->cfg.ntime <= e->bound &&
                        (e->cfg.ntime <= e->bound);
  if (graphable) {
    int g = -1;
    for (int i = 0; i < e->ngraphs; ++i)
      if (e->graphs[i].input_h == input_h && e->graphs[i].cleared == e->cleared &&
          e->graphs[i].buf == buf)
        g = i;
    /* the graph orders after everything queued on the compute stream, and the
     * copy stream's earlier transfers */
    HIPCALL(xgh_event_record(e->ev_aux, e->copy_stream), "record");
    HIPCALL(xgh_stream_wait_event(e->stream, e->ev_aux), "wait");
    if (g < 0) {
      if (e->ngraphs == (int)(sizeof e->graphs / sizeof e->graphs[0])) {
        xgh_graph_destroy(e->graphs[0].exec);
        memmove(&e->graphs[0], &e->graphs[1], sizeof e->graphs[0] * (e->ngraphs - 1));
        e->ngraphs--;
      }
      const int was_cleared = e->cleared;
      const unsigned s_in = e->samples_in_integ;
      const unsigned long long s_all = e->samples;
      HIPCALL(xgh_graph_begin(e->stream), "begin capture");
      g_trace_how = "capture";
      rc = xgh_event_record(e->ev_join, e->stream) || xgh_stream_wait_event(e->copy_stream, e->ev_join)
               ? XGPU_CUDA_ERROR
               : host_issue(e, input_h, dbase, npipe, ntp, chunk);
      g_trace_how = "launch";
      if (rc == XGPU_OK && (xgh_event_record(e->ev_join, e->copy_stream) ||
                            xgh_stream_wait_event(e->stream, e->ev_join)))
        rc = XGPU_CUDA_ERROR;
      void *exec = NULL;
      const int rc_end = xgh_graph_end(e->stream, &exec);
      /* capture only recorded the work: restore the state it advanced */
      e->cleared = was_cleared;
      e->samples_in_integ = s_in;
      e->samples = s_all;
      if (rc == XGPU_OK && rc_end == 0) {
        g = e->ngraphs++;
        e->graphs[g].input_h = input_h;
        e->graphs[g].cleared = was_cleared;
        e->graphs[g].buf = buf;
        e->graphs[g].exec = exec;
      } else {
        /* not capturable (e.g. a pageable source): issue host-path calls eagerly */
        xgh_graph_destroy(exec);
        e->host_graph_off = 1;
      }
    }
    if (g >= 0) {
      trace_event(e, "replay", "host-path graph: input=%p, %d chunks of %u samples", input_h, npipe, ntp);
      HIPCALL(xgh_graph_launch(e->graphs[g].exec, e->stream), "graph launch");
      e->cleared = 0;
      e->samples_in_integ += e->cfg.ntime;
      e->samples += e->cfg.ntime;
    } else {
      /* eager issue: this buffer is free once the kernels of the call that
       * last used it have finished */
      if (e->ev_kernel_used[buf])
        HIPCALL(xgh_stream_wait_event(e->copy_stream, e->ev_kernel[buf]), "wait");
      if ((rc = host_issue(e, input_h, dbase, npipe, ntp, chunk))) return rc;
    }
  } else {
    /* the buffer is free once the kernels of the call that used it finished */
    if (e->ev_kernel_used[buf])
      HIPCALL(xgh_stream_wait_event(e->copy_stream, e->ev_kernel[buf]), "wait");
    if ((rc = host_issue(e, input_h, dbase, npipe, ntp, chunk))) return rc;
  }
  HIPCALL(xgh_event_record(e->ev_kernel[buf], e->stream), "record");
  e->ev_kernel_used[buf] = 1;
  switch (syncOp) {
    case SYNCOP_DUMP:
      return xgpuAmdDump(e, output_h, 0);
    case SYNCOP_SYNC_COMPUTE:
      HIPCALL(xgh_stream_sync(e->stream), "sync compute");
      break;
    case SYNCOP_SYNC_TRANSFER: /* graph replays carry their transfers on the compute stream */
      HIPCALL(xgh_stream_sync(graphable ? e->stream : e->copy_stream), "sync transfer");
      break;
    default:
      break;
  }
  return XGPU_OK;
}

int xgpuAmdCorrelateBatch(const XGPUAmdBatchItem *items, int nitems, void *stream) {
  if (!items || nitems < 1) return xg_fail(XGPU_INVALID_CONFIG, "empty batch");
  if (nitems > XG_MAX_JOBS) return xg_fail(XGPU_INVALID_CONFIG, "batch larger than %d", XG_MAX_JOBS);
  XGPUAmdEngine *e0 = items[0].engine;
  if (!e0) return xg_fail(XGPU_NOT_INITIALIZED, "NULL engine in batch");
  for (int i = 0; i < nitems; ++i) {
    XGPUAmdEngine *e = items[i].engine;
    if (!e) return xg_fail(XGPU_NOT_INITIALIZED, "NULL engine in batch");
    if (e->family != e0->family || e->cfg.device != e0->cfg.device)
      return xg_fail(XGPU_INVALID_CONFIG, "batch items must share compute type and device");
    if (items[i].ntime > e->bound)
      return xg_fail(XGPU_INVALID_CONFIG, "batch block of %u samples exceeds the bound %u",
                     items[i].ntime, e->bound);
    if (((uintptr_t)items[i].input_d) & 7)
      return xg_fail(XGPU_INVALID_CONFIG, "input must be 8-byte aligned");
    for (int k = 0; k < i; ++k)
      if (items[k].engine == e) return xg_fail(XGPU_INVALID_CONFIG, "engine twice in one batch");
  }
  HIPCALL(xgh_set_device(e0->cfg.device), "set device");
  void *s = stream ? stream : e0->stream;
  XgLaunch L;
  memset(&L, 0, sizeof L);
  XGPUAmdEngine *engines[XG_MAX_JOBS];
  const void *raw[XG_MAX_JOBS];
  int chan = 0;
  for (int i = 0; i < nitems; ++i) {
    XGPUAmdEngine *e = items[i].engine;
    if (e->samples_in_integ + items[i].ntime > e->bound) {
      int rc = xg_spill(e);
      if (rc) return rc;
    }
    if (e->stream != s) { /* order after the engine's own queued work */
      HIPCALL(xgh_event_record(e->ev_aux, e->stream), "record");
      HIPCALL(xgh_stream_wait_event(s, e->ev_aux), "wait");
    }
    const void *in = items[i].input_d;
    engines[i] = e;
    raw[i] = in;
    if ((e->block32 || e->swizzled) && items[i].ntime) { /* converted by the step */
      int rc = ensure_unblock(e, (size_t)items[i].ntime * (size_t)e->tstride);
      if (rc) return rc;
      in = e->unblock;
    }
    job_fill(&L.jobs[i], e, in, items[i].ntime, !e->cleared, chan);
    chan += (int)e->cfg.nfrequency;
  }
  L.njobs = nitems;
  /* the real-Gram kernel only when it takes every item */
  int family = xg_launch_family(e0, L.jobs[0].input);
  for (int i = 1; i < nitems; ++i)
    if (xg_launch_family(items[i].engine, L.jobs[i].input) != family) family = e0->family;
  for (int i = 0; i < nitems; ++i) job_tiles(&L.jobs[i], family);
  xg_plan_launch(&L, e0->cus, wg_per_cu_for(e0, family));
  L.prefetch = (char)e0->prefetch;
  L.stage = rg_stage_code(e0);
  int rc = run_step(e0, engines, raw, &L, family, s);
  if (rc) return rc;
  HIPCALL(xgh_event_record(e0->ev_aux, s), "record");
  for (int i = 0; i < nitems; ++i) {
    XGPUAmdEngine *e = items[i].engine;
    if (e->stream != s) HIPCALL(xgh_stream_wait_event(e->stream, e0->ev_aux), "wait");
    e->cleared = 0;
    e->samples_in_integ += items[i].ntime;
    e->samples += items[i].ntime;
  }
  return XGPU_OK;
}

int xgpuAmdTimeCorrelate(XGPUAmdEngine *e, const void *input_d, unsigned int ntime, int warmup,
                         int reps, float *ms) {
  if (!e || !ms || reps < 1) return xg_fail(XGPU_INVALID_CONFIG, "bad arguments");
  int rc;
  for (int i = 0; i < warmup; ++i) {
    xgpuAmdClear(e);
    if ((rc = xgpuAmdCorrelate(e, input_d, ntime))) return rc;
  }
  HIPCALL(xgh_event_record(e->ev_t0, e->stream), "record");
  for (int i = 0; i < reps; ++i) {
    xgpuAmdClear(e);
    if ((rc = xgpuAmdCorrelate(e, input_d, ntime))) return rc;
  }
  HIPCALL(xgh_event_record(e->ev_t1, e->stream), "record");
  HIPCALL(xgh_event_sync(e->ev_t1), "event sync");
  float total = 0.f;
  HIPCALL(xgh_event_elapsed_ms(&total, e->ev_t0, e->ev_t1), "elapsed");
  *ms = total / (float)reps;
  return XGPU_OK;
}

/* MFMA-only throughput of the device (int8 32x32x32, 8 chains per wave, 8
 * workgroups of 4 waves per CU), in complex MACs per second (4 int8 MACs
 * each), best of `reps` timed launches after one warm-up. */
int xgpuAmdMfmaPeakData(int device, const void *operands_d, unsigned compute_type, int reps,
                        double *cmac_per_s) {
  if (!cmac_per_s || !operands_d || reps < 1) return xg_fail(XGPU_INVALID_CONFIG, "bad arguments");
  if (compute_type != XGPU_INT8 && compute_type != XGPU_FLOAT16)
    return xg_fail(XGPU_INVALID_CONFIG, "compute_type must be XGPU_INT8 or XGPU_FLOAT16");
  const int f16 = compute_type == XGPU_FLOAT16;
  HIPCALL(xgh_set_device(device), "set device");
  int cus = 0;
  HIPCALL(xgh_device_cu_count(device, &cus), "CU count");
  const int grid = cus, iters = 8192;
  void *out = NULL, *stream = NULL, *e0 = NULL, *e1 = NULL;
  int rc = XGPU_OK;
  if (xgh_malloc(&out, (size_t)grid * 256 * 4) || xgh_stream_create(&stream) ||
      xgh_event_create(&e0, 1) || xgh_event_create(&e1, 1)) {
    rc = hipfail("mfma peak setup");
    goto done;
  }
  float best = 0.f;
  for (int r = 0; r <= reps; ++r) {
    float ms = 0.f;
    if (xgh_event_record(e0, stream) ||
        xgh_launch_mfma_peak_data(f16, grid, iters, operands_d, (int *)out, stream) ||
        xgh_event_record(e1, stream) || xgh_event_sync(e1) || xgh_event_elapsed_ms(&ms, e0, e1)) {
      rc = hipfail("mfma peak");
      goto done;
    }
    if (r > 0 && (best == 0.f || ms < best)) best = ms;
  }
  /* MACs: grid * 4 waves * iters * 8 MFMAs * (32768 int8 / 16384 fp16) */
  *cmac_per_s = (double)grid * 4 * iters * 8 * (f16 ? 16384.0 : 32768.0) / 4.0 / (best * 1e-3);
done:
  xgh_event_destroy(e0);
  xgh_event_destroy(e1);
  xgh_stream_destroy(stream);
  xgh_free(out);
  return rc;
}

int xgpuAmdMfmaPeak(int device, int reps, double *cmac_per_s) {
  if (!cmac_per_s || reps < 1) return xg_fail(XGPU_INVALID_CONFIG, "bad arguments");
  HIPCALL(xgh_set_device(device), "set device");
  int cus = 0;
  HIPCALL(xgh_device_cu_count(device, &cus), "CU count");
  const int grid = 8 * cus, iters = 2048;
  void *out = NULL, *stream = NULL, *e0 = NULL, *e1 = NULL;
  int rc = XGPU_OK;
  if (xgh_malloc(&out, (size_t)grid * 256 * 4) || xgh_stream_create(&stream) ||
      xgh_event_create(&e0, 1) || xgh_event_create(&e1, 1)) {
    rc = hipfail("mfma peak setup");
    goto done;
  }
  float best = 0.f;
  for (int r = 0; r <= reps; ++r) {
    float ms = 0.f;
    if (xgh_event_record(e0, stream) || xgh_launch_mfma_peak(grid, iters, (int *)out, stream) ||
        xgh_event_record(e1, stream) || xgh_event_sync(e1) || xgh_event_elapsed_ms(&ms, e0, e1)) {
      rc = hipfail("mfma peak");
      goto done;
    }
    if (r > 0 && (best == 0.f || ms < best)) best = ms;
  }
  /* MACs: grid * 4 waves * iters * 8 MFMAs * 32768 */
  *cmac_per_s = (double)grid * 4 * iters * 8 * 32768.0 / 4.0 / (best * 1e-3);
done:
  xgh_event_destroy(e0);
  xgh_event_destroy(e1);
  xgh_stream_destroy(stream);
  xgh_free(out);
  return rc;
}

/* ---- device memory helpers ---------------------------------------------- */
int xgpuAmdMalloc(void **ptr, size_t bytes) {
  if (!ptr) return xg_fail(XGPU_INVALID_CONFIG, "NULL pointer");
  HIPCALL(xgh_malloc(ptr, bytes), "xgpuAmdMalloc");
  return XGPU_OK;
}
int xgpuAmdFree(void *ptr) {
  HIPCALL(xgh_free(ptr), "xgpuAmdFree");
  return XGPU_OK;
}
int xgpuAmdMemcpy(void *dst, const void *src, size_t bytes, int kind) {
  switch (kind) {
    case XGPU_H2D: HIPCALL(xgh_memcpy_h2d(dst, src, bytes), "xgpuAmdMemcpy"); break;
    case XGPU_D2H: HIPCALL(xgh_memcpy_d2h(dst, src, bytes), "xgpuAmdMemcpy"); break;
    case XGPU_D2D:
      if (((((uintptr_t)dst | (uintptr_t)src) & 15) == 0) && (bytes & 3) == 0)
        HIPCALL(xgh_launch_copy(dst, src, bytes, NULL), "xgpuAmdMemcpy");
      else
        HIPCALL(xgh_memcpy_d2d_async(dst, src, bytes, NULL), "xgpuAmdMemcpy");
      HIPCALL(xgh_device_sync(), "xgpuAmdMemcpy");
      break;
    default: return xg_fail(XGPU_INVALID_CONFIG, "bad copy kind %d", kind);
  }
  return XGPU_OK;
}
int xgpuAmdDeviceSynchronize(void) {
  HIPCALL(xgh_device_sync(), "xgpuAmdDeviceSynchronize");
  return XGPU_OK;
}
int xgpuAmdSwizzleInput(void *dst_d, const void *src_d, unsigned nstation, unsigned nfrequency,
                        unsigned ntime, int unswizzle, void *stream) {
  if (!dst_d || !src_d) return xg_fail(XGPU_INVALID_CONFIG, "NULL pointer");
  if (ntime % 4) return xg_fail(XGPU_INVALID_CONFIG, "ntime (%u) must be a multiple of 4", ntime);
  const void *sw = unswizzle ? src_d : dst_d, *pl = unswizzle ? dst_d : src_d;
  if (((uintptr_t)sw & 15) || ((uintptr_t)pl & 3))
    return xg_fail(XGPU_INVALID_CONFIG, "swizzled buffer must be 16-byte and plain buffer 4-byte aligned");
  const size_t bytes = (size_t)ntime * nfrequency * nstation * 4;
  if ((const char *)dst_d < (const char *)src_d + bytes && (const char *)src_d < (const char *)dst_d + bytes)
    return xg_fail(XGPU_INVALID_CONFIG, "source and destination overlap");
  HIPCALL(xgh_launch_swizzle(unswizzle != 0, dst_d, src_d, (int)nstation, (int)nfrequency, (int)ntime,
                             stream),
          "xgpuAmdSwizzleInput");
  return XGPU_OK;
}

int xgpuAmdExtractMatrix(void *matrix_d, const void *packed_d, unsigned nstation, unsigned nfrequency,
                         int elem_kind, void *stream) {
  if (!matrix_d || !packed_d) return xg_fail(XGPU_INVALID_CONFIG, "NULL pointer");
  if (elem_kind < XGPU_ELEM_INT32 || elem_kind > XGPU_ELEM_FLOAT64)
    return xg_fail(XGPU_INVALID_CONFIG, "unknown element kind %d", elem_kind);
  const size_t esz = (elem_kind == XGPU_ELEM_INT64 || elem_kind == XGPU_ELEM_FLOAT64) ? 8 : 4;
  if ((((uintptr_t)matrix_d) | ((uintptr_t)packed_d)) & (esz - 1))
    return xg_fail(XGPU_INVALID_CONFIG, "buffers must be aligned to the element size");
  /* XGPU_ELEM_* equal the XG_OUT_* kinds of the kernels */
  HIPCALL(xgh_launch_extract(elem_kind, matrix_d, packed_d, (int)nstation, (int)nfrequency, stream),
          "xgpuAmdExtractMatrix");
  return XGPU_OK;
}

int xgpuAmdFillUniformInt8(void *ptr_d, size_t nbytes, unsigned long long seed, int lo, int hi) {
  if (hi < lo || lo < -128 || hi > 127) return xg_fail(XGPU_INVALID_CONFIG, "bad range");
  HIPCALL(xgh_launch_fill_uniform_i8(ptr_d, nbytes, seed, lo, hi, NULL), "fill");
  HIPCALL(xgh_device_sync(), "fill sync");
  return XGPU_OK;
}
