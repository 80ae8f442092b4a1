/*
 * xg_hip_fake.c -- a CPU fake of the thin HIP C-ABI layer (xgpu_amd/csrc/
 * xg_hip.h), for testing the host runtime (engine.c, serve.c, autotune.c,
 * xgpu_api.c) without a GPU, under AddressSanitizer and UBSan, with fault
 * injection.  TEST INFRASTRUCTURE ONLY: never linked into the product.
 *
 * "Device" memory is host memory; every operation runs synchronously when it
 * is issued (so every stream and event is always complete), except on a
 * stream that is capturing a graph: its operations, and those of streams
 * that joined the capture by waiting on one of its events, are recorded and
 * run again at each graph launch.  The kernels are CPU restatements of what
 * the gfx950 kernels compute: the X-engine writes every visibility of each
 * job through the same index formulas (xg_kernels.hip vis_index, reference
 * src/kernel.cuh:49-117), in int32 (wrapping), float or the wide kinds, and
 * the layout kernels move the same elements.
 *
 * Fault injection: xgh_fake_fail(category, after) makes the call of that
 * category that follows `after` successful ones fail (once); categories:
 * "malloc" (device and pinned host allocations), "copy", "launch", "event",
 * "stream", "graph", "sync".
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "xg_hip.h"

#define FAKE_ERR 2 /* the code a failing call returns */

static _Thread_local char g_err[256];
static long long g_live;  /* live device + pinned allocations (atomic updates) */
static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;  /* the fault table */

static struct {
  const char *name;
  int armed, after;
  long long calls;
} g_fault[] = {{"malloc"}, {"copy"}, {"launch"}, {"event"}, {"stream"}, {"graph"}, {"sync"}};
#define NFAULT ((int)(sizeof g_fault / sizeof g_fault[0]))

void xgh_fake_fail(const char *category, int after) {
  pthread_mutex_lock(&g_mu);
  for (int i = 0; i < NFAULT; ++i)
    if (!strcmp(g_fault[i].name, category)) {
      g_fault[i].armed = 1;
      g_fault[i].after = after;
    }
  pthread_mutex_unlock(&g_mu);
}

void xgh_fake_reset(void) {
  pthread_mutex_lock(&g_mu);
  for (int i = 0; i < NFAULT; ++i) {
    g_fault[i].armed = 0;
    g_fault[i].calls = 0;
  }
  pthread_mutex_unlock(&g_mu);
}

long long xgh_fake_live_allocations(void) { return __atomic_load_n(&g_live, __ATOMIC_SEQ_CST); }

long long xgh_fake_calls(const char *category) {
  long long n = -1;
  pthread_mutex_lock(&g_mu);
  for (int i = 0; i < NFAULT; ++i)
    if (!strcmp(g_fault[i].name, category)) n = g_fault[i].calls;
  pthread_mutex_unlock(&g_mu);
  return n;
}

/* 0, or FAKE_ERR when an armed fault of this category fires now */
static int fault(const char *category, const char *what) {
  int rc = 0;
  pthread_mutex_lock(&g_mu);
  for (int i = 0; i < NFAULT; ++i) {
    if (strcmp(g_fault[i].name, category)) continue;
    g_fault[i].calls++;
    if (g_fault[i].armed && g_fault[i].after-- == 0) {
      g_fault[i].armed = 0;
      snprintf(g_err, sizeof g_err, "%s: injected fault (%s)", what, category);
      rc = FAKE_ERR;
    }
  }
  pthread_mutex_unlock(&g_mu);
  return rc;
}

const char *xgh_last_error(void) { return g_err; }

/* ---- devices -------------------------------------------------------------- */
int xgh_device_count(int *count) {
  *count = 1;
  return 0;
}
int xgh_set_device(int device) { return device == 0 ? 0 : FAKE_ERR; }
int xgh_get_device(int *device) {
  *device = 0;
  return 0;
}
int xgh_device_name(int device, char *buf, size_t len) {
  (void)device;
  snprintf(buf, len, "cpu fake of gfx950");
  return 0;
}
int xgh_device_cu_count(int device, int *cus) {
  (void)device;
  *cus = 256;
  return 0;
}

/* ---- memory --------------------------------------------------------------- */
int xgh_malloc(void **ptr, size_t bytes) {
  *ptr = NULL;
  if (fault("malloc", "xgh_malloc")) return FAKE_ERR;
  *ptr = malloc(bytes ? bytes : 1);
  if (!*ptr) return FAKE_ERR;
  __atomic_add_fetch(&g_live, 1, __ATOMIC_SEQ_CST);
  return 0;
}
int xgh_free(void *ptr) {
  if (ptr) {
    free(ptr);
    __atomic_sub_fetch(&g_live, 1, __ATOMIC_SEQ_CST);
  }
  return 0;
}
int xgh_host_alloc(void **ptr, size_t bytes) { return xgh_malloc(ptr, bytes); }
int xgh_host_free(void *ptr) { return xgh_free(ptr); }

/* ---- operations, streams, graphs ----------------------------------------- */
enum { OP_COPY, OP_XENGINE, OP_FILL0, OP_SPILL, OP_COMBINE, OP_REORDER, OP_REORDER_COMBINE,
       OP_SWIZZLE, OP_EXTRACT, OP_UNBLOCK32, OP_UNIFORM, OP_NOP };

typedef struct {
  int kind;
  int a, b, c;           /* small arguments (family, kind, narrow, order, ...) */
  void *dst;
  const void *src, *src2;
  size_t n;
  int ns, nf, nt, lo, hi;
  long long plane;
  uint64_t seed;
  XgLaunch *L;           /* OP_XENGINE: a copy of the descriptor */
} Op;

typedef struct Graph {
  Op *ops;
  int nops, cap;
} Graph;

typedef struct Stream {
  Graph *capture;        /* non-NULL while recording */
  int origin;            /* the stream that began the capture */
} Stream;

typedef struct Event {
  Graph *capture;        /* recorded inside a capture */
} Event;

static void run_op(const Op *op);

static int issue(void *stream, Op op) {
  Stream *s = (Stream *)stream;
  if (s && s->capture) {
    Graph *g = s->capture;
    if (g->nops == g->cap) {
      const int cap = g->cap ? 2 * g->cap : 16;
      Op *p = (Op *)realloc(g->ops, sizeof(Op) * (size_t)cap);
      if (!p) return FAKE_ERR;
      g->ops = p;
      g->cap = cap;
    }
    if (op.L) { /* the graph owns its copy of the descriptor */
      XgLaunch *c = (XgLaunch *)malloc(sizeof *c);
      if (!c) return FAKE_ERR;
      memcpy(c, op.L, sizeof *c);
      op.L = c;
    }
    g->ops[g->nops++] = op;
    return 0;
  }
  run_op(&op);
  return 0;
}

int xgh_stream_create(void **stream) {
  *stream = NULL;
  if (fault("stream", "xgh_stream_create")) return FAKE_ERR;
  *stream = calloc(1, sizeof(Stream));
  return *stream ? 0 : FAKE_ERR;
}
int xgh_stream_destroy(void *stream) {
  free(stream);
  return 0;
}
int xgh_stream_sync(void *stream) {
  (void)stream;
  return fault("sync", "xgh_stream_sync");
}
/* Streams that joined a capture by waiting on one of its events: released
 * when the capture ends. */
#define MAX_JOINED 64
static Stream *g_joined[MAX_JOINED];

int xgh_stream_wait_event(void *stream, void *event) {
  Stream *s = (Stream *)stream;
  Event *e = (Event *)event;
  if (s && e && e->capture && !s->capture) { /* joins the capture */
    s->capture = e->capture;
    for (int i = 0; i < MAX_JOINED; ++i)
      if (!g_joined[i]) {
        g_joined[i] = s;
        break;
      }
  }
  return 0;
}
int xgh_event_create(void **event, int timing) {
  (void)timing;
  *event = NULL;
  if (fault("event", "xgh_event_create")) return FAKE_ERR;
  *event = calloc(1, sizeof(Event));
  return *event ? 0 : FAKE_ERR;
}
int xgh_event_destroy(void *event) {
  free(event);
  return 0;
}
int xgh_event_record(void *event, void *stream) {
  if (fault("event", "xgh_event_record")) return FAKE_ERR;
  Stream *s = (Stream *)stream;
  ((Event *)event)->capture = (s && s->capture) ? s->capture : NULL;
  return 0;
}
int xgh_event_sync(void *event) {
  (void)event;
  return fault("sync", "xgh_event_sync");
}
int xgh_event_query(void *event) {
  (void)event;
  return 0;
}
int xgh_event_elapsed_ms(float *ms, void *start, void *stop) {
  (void)start;
  (void)stop;
  *ms = 0.001f;
  return 0;
}
int xgh_device_sync(void) { return fault("sync", "xgh_device_sync"); }

int xgh_graph_begin(void *stream) {
  if (fault("graph", "xgh_graph_begin")) return FAKE_ERR;
  Stream *s = (Stream *)stream;
  if (!s || s->capture) return FAKE_ERR;
  s->capture = (Graph *)calloc(1, sizeof(Graph));
  s->origin = 1;
  return s->capture ? 0 : FAKE_ERR;
}

int xgh_graph_end(void *stream, void **exec) {
  Stream *s = (Stream *)stream;
  *exec = NULL;
  if (!s || !s->capture || !s->origin) return FAKE_ERR;
  Graph *g = s->capture;
  s->capture = NULL;
  s->origin = 0;
  for (int i = 0; i < MAX_JOINED; ++i)
    if (g_joined[i] && g_joined[i]->capture == g) {
      g_joined[i]->capture = NULL;
      g_joined[i] = NULL;
    }
  *exec = g;
  return fault("graph", "xgh_graph_end");
}

int xgh_graph_launch(void *exec, void *stream) {
  if (fault("graph", "xgh_graph_launch")) return FAKE_ERR;
  (void)stream;
  const Graph *g = (const Graph *)exec;
  for (int i = 0; i < g->nops; ++i) run_op(&g->ops[i]);
  return 0;
}

int xgh_graph_destroy(void *exec) {
  Graph *g = (Graph *)exec;
  if (!g) return 0;
  for (int i = 0; i < g->nops; ++i) free(g->ops[i].L);
  free(g->ops);
  free(g);
  return 0;
}

int xgh_memcpy_h2d_async(void *dst, const void *src, size_t bytes, void *stream) {
  if (fault("copy", "xgh_memcpy_h2d_async")) return FAKE_ERR;
  return issue(stream, (Op){.kind = OP_COPY, .dst = dst, .src = src, .n = bytes});
}
int xgh_memcpy_d2h_async(void *dst, const void *src, size_t bytes, void *stream) {
  if (fault("copy", "xgh_memcpy_d2h_async")) return FAKE_ERR;
  return issue(stream, (Op){.kind = OP_COPY, .dst = dst, .src = src, .n = bytes});
}
int xgh_memcpy_d2d_async(void *dst, const void *src, size_t bytes, void *stream) {
  if (fault("copy", "xgh_memcpy_d2d_async")) return FAKE_ERR;
  return issue(stream, (Op){.kind = OP_COPY, .dst = dst, .src = src, .n = bytes});
}
int xgh_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  return xgh_memcpy_d2h_async(dst, src, bytes, NULL);
}
int xgh_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  return xgh_memcpy_h2d_async(dst, src, bytes, NULL);
}

/* ---- kernels -------------------------------------------------------------- */
int xgh_schedule(const XgLaunch *launch, int xcd, int k, int *job, int *f, int *rb, int *cb) {
  (void)launch, (void)xcd, (void)k, (void)job, (void)f, (void)rb, (void)cb;
  return 0;
}

static int launch(void *stream, Op op) {
  if (fault("launch", "kernel launch")) return FAKE_ERR;
  return issue(stream, op);
}

int xgh_launch_xengine(int family, const XgLaunch *L, void *stream) {
  return launch(stream, (Op){.kind = OP_XENGINE, .a = family, .L = (XgLaunch *)L});
}
int xgh_launch_fill_zero(void *ptr, size_t bytes, void *stream) {
  return launch(stream, (Op){.kind = OP_FILL0, .dst = ptr, .n = bytes});
}
int xgh_launch_copy(void *dst, const void *src, size_t bytes, void *stream) {
  return launch(stream, (Op){.kind = OP_COPY, .dst = dst, .src = src, .n = bytes});
}
int xgh_launch_spill(int kind, int add, void *dst, const void *src, size_t n, void *stream) {
  return launch(stream, (Op){.kind = OP_SPILL, .a = kind, .b = add, .dst = dst, .src = src, .n = n});
}
int xgh_launch_combine(int kind, int narrow, void *dst, const void *spill, const void *cur,
                       size_t n, void *stream) {
  return launch(stream, (Op){.kind = OP_COMBINE, .a = kind, .b = narrow, .dst = dst, .src = spill,
                             .src2 = cur, .n = n});
}
int xgh_launch_fill_uniform_i8(void *ptr, size_t n, uint64_t seed, int lo, int hi, void *stream) {
  if (hi < lo) return FAKE_ERR;
  return launch(stream, (Op){.kind = OP_UNIFORM, .dst = ptr, .n = n, .seed = seed, .lo = lo, .hi = hi});
}
int xgh_launch_reorder_combine(int src_order, int kind, int narrow, void *dst, const void *spill,
                               const void *cur, int nstation, int nfreq, long long plane,
                               void *stream) {
  return launch(stream, (Op){.kind = OP_REORDER_COMBINE, .a = src_order, .b = kind, .c = narrow,
                             .dst = dst, .src = spill, .src2 = cur, .ns = nstation, .nf = nfreq,
                             .plane = plane});
}
int xgh_launch_reorder(int src_order, int elem_bytes, void *dst, const void *src, int nstation,
                       int nfreq, long long plane, void *stream) {
  return launch(stream, (Op){.kind = OP_REORDER, .a = src_order, .b = elem_bytes, .dst = dst,
                             .src = src, .ns = nstation, .nf = nfreq, .plane = plane});
}
int xgh_launch_mfma_peak(int grid, int iters, int *out, void *stream) {
  (void)iters;
  if (fault("launch", "mfma peak")) return FAKE_ERR;
  (void)stream;
  memset(out, 0, (size_t)grid * 256 * sizeof(int));
  return 0;
}
int xgh_launch_mfma_peak_data(int f16, int grid, int iters, const void *ops, int *out,
                              void *stream) {
  (void)f16, (void)ops;
  return xgh_launch_mfma_peak(grid, iters, out, stream);
}
int xgh_launch_swizzle(int unswizzle, void *dst, const void *src, int nstation, int nfreq,
                       int ntime, void *stream) {
  if (ntime % 4) return FAKE_ERR;
  return launch(stream, (Op){.kind = OP_SWIZZLE, .a = unswizzle, .dst = dst, .src = src,
                             .ns = nstation, .nf = nfreq, .nt = ntime});
}
int xgh_launch_extract(int kind, void *dst, const void *packed, int nstation, int nfreq,
                       void *stream) {
  return launch(stream, (Op){.kind = OP_EXTRACT, .a = kind, .dst = dst, .src = packed,
                             .ns = nstation, .nf = nfreq});
}
int xgh_launch_unblock32(void *dst, const void *src, size_t bytes, void *stream) {
  if (bytes % 64) return FAKE_ERR;
  return launch(stream, (Op){.kind = OP_UNBLOCK32, .dst = dst, .src = src, .n = bytes});
}

#define FAKE_SCRATCH_BYTES ((size_t)XG_STATIC_SCRATCH_BYTES)
static void *g_scratch;
int xgh_static_scratch(int device, void **ptr) {
  if (device != 0) return FAKE_ERR;
  if (!g_scratch) g_scratch = malloc(FAKE_SCRATCH_BYTES); /* held until the process ends */
  *ptr = g_scratch;
  return g_scratch ? 0 : FAKE_ERR;
}

/* ---- CPU restatements of the kernels --------------------------------------- */
static double half_to_double(uint16_t h) {
  const int sign = (h >> 15) & 1, ex = (h >> 10) & 0x1f, man = h & 0x3ff;
  double v = ex == 0 ? ldexp((double)man, -24) : ex == 31 ? (man ? NAN : INFINITY)
                                                           : ldexp((double)(man | 0x400), ex - 25);
  return sign ? -v : v;
}

/* (re, im) sample of input m (2 station + pol) at time t of channel f */
static void sample(const XgJob *J, int fp16, int t, int f, int m, double *re, double *im) {
  const long long at = (long long)t * J->tstride + ((long long)f * 2 * J->nstation + m) * (fp16 ? 4 : 2);
  if (fp16) {
    const uint16_t *p = (const uint16_t *)((const char *)J->input + at);
    *re = half_to_double(p[0]);
    *im = half_to_double(p[1]);
  } else {
    const int8_t *p = (const int8_t *)J->input + at;
    *re = p[0];
    *im = p[1];
  }
}

/* Output index of the real part and the distance to the imaginary part;
 * 0 when (m, n) is not stored (the kernels' vis_index). */
static int vis_index(const XgJob *J, int f, int m, int n, long long *ire, long long *dim,
                     int *zero) {
  const int i = m >> 1, p1 = m & 1, j = n >> 1, p2 = n & 1;
  *zero = 0;
  if (J->order == XG_ORDER_REGTILE) {
    const long long P = J->nstation >> 1, RT = P * (P + 1) / 2;
    const int R = i >> 1, rx = i & 1, C = j >> 1, ry = j & 1;
    if (i < j) {
      if (R != C || J->accumulate) return 0;
      *zero = 1;
    }
    const long long l = ((long long)f * 4 + (2 * ry + rx)) * RT + (long long)R * (R + 1) / 2 + C;
    *ire = (l * 2 + p1) * 2 + p2;
    *dim = J->plane;
    return 1;
  }
  if (i < j) return 0;
  const long long NB = (long long)J->nstation * (J->nstation + 1) / 2;
  const long long k = (long long)f * NB + (long long)i * (i + 1) / 2 + j;
  if (J->order == XG_ORDER_REALIMAG) {
    *ire = k * 4 + p1 * 2 + p2;
    *dim = J->plane;
  } else {
    *ire = ((k * 2 + p1) * 2 + p2) * 2;
    *dim = 1;
  }
  return 1;
}

static void put(const XgJob *J, long long idx, double v, int64_t iv) {
  switch (J->out_kind) {
    case XG_OUT_I32: {
      uint32_t *p = (uint32_t *)J->out + idx;
      *p = J->accumulate ? *p + (uint32_t)iv : (uint32_t)iv;
      break;
    }
    case XG_OUT_I64: {
      int64_t *p = (int64_t *)J->out + idx;
      *p = J->accumulate ? *p + iv : iv;
      break;
    }
    case XG_OUT_F32: {
      float *p = (float *)J->out + idx;
      *p = J->accumulate ? *p + (float)v : (float)v;
      break;
    }
    default: {
      double *p = (double *)J->out + idx;
      *p = J->accumulate ? *p + v : v;
    }
  }
}

static void run_xengine(int family, const XgLaunch *L) {
  const int fp16 = family == XG_KERNEL_F16_MFMA || family == XG_KERNEL_F16_RG;
  for (int q = 0; q < L->njobs; ++q) {
    const XgJob *J = &L->jobs[q];
    const int nin = 2 * J->nstation;
    for (int f = 0; f < J->nfreq; ++f)
      for (int m = 0; m < nin; ++m)
        for (int n = 0; n < nin; ++n) {
          long long ire, dim;
          int zero;
          if (!vis_index(J, f, m, n, &ire, &dim, &zero)) continue;
          double re = 0, im = 0;
          int64_t ire64 = 0, iim64 = 0;
          for (int t = 0; t < J->ntime && !zero; ++t) {
            double ar, ai, br, bi;
            sample(J, fp16, t, f, m, &ar, &ai);
            sample(J, fp16, t, f, n, &br, &bi);
            re += ar * br + ai * bi;
            im += ai * br - ar * bi;
            ire64 += (int64_t)(ar * br + ai * bi);
            iim64 += (int64_t)(ai * br - ar * bi);
          }
          put(J, ire, re, ire64);
          put(J, ire + dim, im, iim64);
        }
  }
}

static long long reorder_src(int order, long long x, long long NB, long long RT) {
  const int pp = (int)(x & 3);
  const long long kk = x >> 2;
  if (order != XG_ORDER_REGTILE) return kk * 4 + pp;
  const int f = (int)(kk / NB);
  const long long b = kk - (long long)f * NB;
  int i = 0;
  while ((long long)(i + 1) * (i + 2) / 2 <= b) ++i;
  const int j = (int)(b - (long long)i * (i + 1) / 2);
  const int R = i >> 1, rx = i & 1, C = j >> 1, ry = j & 1;
  return (((long long)f * 4 + (2 * ry + rx)) * RT + (long long)R * (R + 1) / 2 + C) * 4 + pp;
}

static void run_op(const Op *op) {
  switch (op->kind) {
    case OP_COPY:
      memmove(op->dst, op->src, op->n);
      break;
    case OP_XENGINE:
      run_xengine(op->a, op->L);
      break;
    case OP_FILL0:
      memset(op->dst, 0, op->n);
      break;
    case OP_SPILL:
      for (size_t i = 0; i < op->n; ++i) {
        if (op->a == XG_OUT_I64) {
          int64_t *d = (int64_t *)op->dst;
          d[i] = (op->b ? d[i] : 0) + ((const int32_t *)op->src)[i];
        } else {
          double *d = (double *)op->dst;
          d[i] = (op->b ? d[i] : 0.0) + ((const float *)op->src)[i];
        }
      }
      break;
    case OP_COMBINE:
      for (size_t i = 0; i < op->n; ++i) {
        if (op->a == XG_OUT_I64) {
          const uint64_t v = (uint64_t)((const int64_t *)op->src)[i] + (uint64_t)(int64_t)((const int32_t *)op->src2)[i];
          if (op->b) ((uint32_t *)op->dst)[i] = (uint32_t)v;
          else ((int64_t *)op->dst)[i] = (int64_t)v;
        } else {
          const double v = ((const double *)op->src)[i] + ((const float *)op->src2)[i];
          if (op->b) ((float *)op->dst)[i] = (float)v;
          else ((double *)op->dst)[i] = v;
        }
      }
      break;
    case OP_REORDER:
    case OP_REORDER_COMBINE: {
      const long long NB = (long long)op->ns * (op->ns + 1) / 2, P = op->ns >> 1;
      const long long RT = P * (P + 1) / 2, total = (long long)op->nf * NB * 4;
      for (long long x = 0; x < total; ++x) {
        const long long s = reorder_src(op->a, x, NB, RT);
        if (op->kind == OP_REORDER) {
          const int eb = op->b;
          char *d = (char *)op->dst;
          const char *c = (const char *)op->src;
          memcpy(d + (2 * x) * eb, c + s * eb, (size_t)eb);
          memcpy(d + (2 * x + 1) * eb, c + (s + op->plane) * eb, (size_t)eb);
        } else if (op->b == XG_OUT_I64) {
          const int64_t *w = (const int64_t *)op->src;
          const int32_t *c = (const int32_t *)op->src2;
          const uint64_t re = (uint64_t)w[s] + (uint64_t)(int64_t)c[s];
          const uint64_t im = (uint64_t)w[s + op->plane] + (uint64_t)(int64_t)c[s + op->plane];
          if (op->c) {
            ((uint32_t *)op->dst)[2 * x] = (uint32_t)re;
            ((uint32_t *)op->dst)[2 * x + 1] = (uint32_t)im;
          } else {
            ((int64_t *)op->dst)[2 * x] = (int64_t)re;
            ((int64_t *)op->dst)[2 * x + 1] = (int64_t)im;
          }
        } else {
          const double *w = (const double *)op->src;
          const float *c = (const float *)op->src2;
          const double re = w[s] + c[s], im = w[s + op->plane] + c[s + op->plane];
          if (op->c) {
            ((float *)op->dst)[2 * x] = (float)re;
            ((float *)op->dst)[2 * x + 1] = (float)im;
          } else {
            ((double *)op->dst)[2 * x] = re;
            ((double *)op->dst)[2 * x + 1] = im;
          }
        }
      }
      break;
    }
    case OP_SWIZZLE: {
      const size_t row = (size_t)op->nf * op->ns * 4;
      unsigned char *d = (unsigned char *)op->dst;
      const unsigned char *c = (const unsigned char *)op->src;
      for (size_t t = 0; t < (size_t)op->nt; ++t)
        for (size_t e = 0; e < row; ++e) {
          const size_t sw = ((t / 4) * row + e) * 4 + t % 4, pl = t * row + e;
          if (op->a) d[pl] = c[sw]; else d[sw] = c[pl];
        }
      break;
    }
    case OP_EXTRACT: {
      const int ns = op->ns;
      const size_t esz = (op->a == XG_OUT_I64 || op->a == XG_OUT_F64) ? 8 : 4;
      const long long nb = (long long)ns * (ns + 1) / 2;
      for (int f = 0; f < op->nf; ++f)
        for (int i = 0; i < ns; ++i)
          for (int j = 0; j < ns; ++j)
            for (int p1 = 0; p1 < 2; ++p1)
              for (int p2 = 0; p2 < 2; ++p2) {
                const int direct = i > j || (i == j && p1 > p2);
                const int a = i >= j ? i : j, b = i >= j ? j : i;
                const long long cell = (long long)f * nb + (long long)a * (a + 1) / 2 + b;
                const long long s = (cell * 4 + (direct ? p1 * 2 + p2 : p2 * 2 + p1)) * 2;
                const long long d = ((((long long)f * ns + i) * ns + j) * 4 + p1 * 2 + p2) * 2;
                for (int c = 0; c < 2; ++c) {
                  const char *from = (const char *)op->src + (s + c) * (long long)esz;
                  char *to = (char *)op->dst + (d + c) * (long long)esz;
                  if (c == 0 || direct) {
                    memcpy(to, from, esz);
                  } else if (op->a == XG_OUT_I32) {
                    int32_t v;
                    memcpy(&v, from, 4);
                    v = (int32_t)(0u - (uint32_t)v);
                    memcpy(to, &v, 4);
                  } else if (op->a == XG_OUT_I64) {
                    int64_t v;
                    memcpy(&v, from, 8);
                    v = (int64_t)(0ull - (uint64_t)v);
                    memcpy(to, &v, 8);
                  } else if (op->a == XG_OUT_F32) {
                    float v;
                    memcpy(&v, from, 4);
                    v = -v;
                    memcpy(to, &v, 4);
                  } else {
                    double v;
                    memcpy(&v, from, 8);
                    v = -v;
                    memcpy(to, &v, 8);
                  }
                }
              }
      break;
    }
    case OP_UNBLOCK32: {
      const size_t nblk = op->n / 64;
      for (size_t k = 0; k < nblk; ++k) {
        unsigned char tmp[64];
        memcpy(tmp, (const unsigned char *)op->src + 64 * k, 64);
        for (int e = 0; e < 32; ++e) {
          ((unsigned char *)op->dst)[64 * k + 2 * e] = tmp[e];
          ((unsigned char *)op->dst)[64 * k + 2 * e + 1] = tmp[32 + e];
        }
      }
      break;
    }
    case OP_UNIFORM: {
      uint64_t x = op->seed + 0x9E3779B97F4A7C15ull;
      x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
      x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
      const uint64_t key = x ^ (x >> 31), span = (uint64_t)(op->hi - op->lo + 1);
      for (size_t i = 0; i < op->n; ++i) {
        uint64_t r = key + i + 0x9E3779B97F4A7C15ull;
        r = (r ^ (r >> 30)) * 0xBF58476D1CE4E5B9ull;
        r = (r ^ (r >> 27)) * 0x94D049BB133111EBull;
        r ^= r >> 31;
        ((unsigned char *)op->dst)[i] = (unsigned char)((op->lo + (int)(((r >> 32) * span) >> 32)) & 0xff);
      }
      break;
    }
    default:
      break;
  }
}
