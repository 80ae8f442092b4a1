/*
 * test_fault_injection.c -- the host runtime (engine.c, serve.c,
 * autotune.c, xgpu_api.c) against the CPU fake of the HIP layer
 * (xg_hip_fake.c), built with AddressSanitizer and UBSan.  TEST
 * INFRASTRUCTURE ONLY.
 *
 * Scenarios: the xgpu.h pipeline (xgpuInit -> xgpuCudaXengine(SYNCOP_DUMP) ->
 * xgpuReorderMatrix) against xgpuOmpXengine; every allocation, copy, launch,
 * event and stream of xgpuInit / xgpuAmdCreate / xgpuCudaXengine / dumps /
 * spills / the server failing in turn, with the error code and message the
 * reference's checkCudaError path reports as XGPU_CUDA_ERROR (reference
 * src/cuda_xengine.cu:94-101) but without exit(), and with nothing leaked;
 * graph-captured steps, the batch, the allocation-free entry point and the
 * autotuner on the fake.  Prints "ok <scenario>" per scenario; exits nonzero
 * on the first failure (and LeakSanitizer fails the run on any leak).
 *
 * Built with the xgpu.h sizing NSTATION=16, NFREQUENCY=2, NTIME=64,
 * NTIME_PIPE=16 (float Complex).
 */
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "xgpu.h"
#include "xgpu_amd.h"
#include "xgpu_serve.h"

void xgh_fake_fail(const char *category, int after);
void xgh_fake_reset(void);
long long xgh_fake_live_allocations(void);
long long xgh_fake_calls(const char *category);

#define CHECK(cond, ...)                                                   \
  do {                                                                     \
    if (!(cond)) {                                                         \
      fprintf(stderr, "FAIL %s:%d: %s: ", __FILE__, __LINE__, #cond);      \
      fprintf(stderr, __VA_ARGS__);                                        \
      fprintf(stderr, "\n");                                               \
      exit(1);                                                             \
    }                                                                      \
  } while (0)

static const char *categories[] = {"malloc", "copy", "launch", "event", "stream", "sync", "graph"};
#define NCAT ((int)(sizeof categories / sizeof categories[0]))

/* deterministic int8 test data in [-20, 20] */
static void fill(int8_t *x, size_t n, unsigned seed) {
  uint64_t s = 0x9E3779B97F4A7C15ull * (seed + 1);
  for (size_t i = 0; i < n; ++i) {
    s ^= s << 13, s ^= s >> 7, s ^= s << 17;
    x[i] = (int8_t)((int)(s % 41) - 20);
  }
}

/* exact triangular visibilities (int64 interleaved) of int8 input */
static void exact(int64_t *out, const int8_t *x, int ns, int nf, int nt) {
  const long long nb = (long long)ns * (ns + 1) / 2;
  for (int f = 0; f < nf; ++f)
    for (int i = 0; i < ns; ++i)
      for (int j = 0; j <= i; ++j)
        for (int p1 = 0; p1 < 2; ++p1)
          for (int p2 = 0; p2 < 2; ++p2) {
            int64_t re = 0, im = 0;
            for (int t = 0; t < nt; ++t) {
              const int8_t *a = x + ((((size_t)t * nf + f) * ns + i) * 2 + p1) * 2;
              const int8_t *b = x + ((((size_t)t * nf + f) * ns + j) * 2 + p2) * 2;
              re += a[0] * b[0] + a[1] * b[1];
              im += a[1] * b[0] - a[0] * b[1];
            }
            const long long k = ((f * nb + (long long)i * (i + 1) / 2 + j) * 4 + p1 * 2 + p2) * 2;
            out[k] = re;
            out[k + 1] = im;
          }
}

/* ---- xgpu.h pipeline ------------------------------------------------------ */
static void xgpu_h_pipeline(void) {
  XGPUInfo info;
  xgpuInfo(&info);
  XGPUContext ctx;
  memset(&ctx, 0, sizeof ctx);
  CHECK(xgpuInit(&ctx, 0) == XGPU_OK, "init");
  fill((int8_t *)ctx.array_h, info.vecLength * 2, 1);
  Complex *cpu = malloc(info.triLength * sizeof(Complex));
  xgpuOmpXengine(cpu, ctx.array_h);
  for (int call = 0; call < 2; ++call) { /* a fresh integration each time */
    CHECK(xgpuClearDeviceIntegrationBuffer(&ctx) == XGPU_OK, "clear");
    CHECK(xgpuCudaXengine(&ctx, SYNCOP_DUMP) == XGPU_OK, "xengine");
    xgpuReorderMatrix(ctx.matrix_h);
    CHECK(memcmp(ctx.matrix_h, cpu, info.triLength * sizeof(Complex)) == 0, "result call %d", call);
  }
  free(cpu);
  xgpuFree(&ctx);
  CHECK(xgh_fake_live_allocations() == 0, "leaked %lld", xgh_fake_live_allocations());
  printf("ok xgpu.h pipeline matches xgpuOmpXengine\n");
}

/* xgpuInit with the n-th call of each category failing: an error code, no
 * context, nothing leaked; then a clean init still works. */
static void xgpu_init_faults(void) {
  int failures = 0;
  for (int c = 0; c < NCAT; ++c)
    for (int n = 0; n < 12; ++n) {
      xgh_fake_reset();
      xgh_fake_fail(categories[c], n);
      XGPUContext ctx;
      memset(&ctx, 0, sizeof ctx);
      const int rc = xgpuInit(&ctx, 0);
      if (rc != XGPU_OK) {
        ++failures;
        CHECK(rc == XGPU_CUDA_ERROR || rc == XGPU_OUT_OF_MEMORY, "code %d", rc);
        CHECK(ctx.internal == NULL, "context left set after a failed init");
      }
      xgpuFree(&ctx);
      CHECK(xgh_fake_live_allocations() == 0, "%s #%d leaked %lld", categories[c], n,
            xgh_fake_live_allocations());
    }
  xgh_fake_reset();
  CHECK(failures > 0, "no injected fault reached xgpuInit");
  printf("ok xgpuInit fails cleanly on %d injected faults\n", failures);
}

/* xgpuCudaXengine with a failing copy / launch / event: XGPU_CUDA_ERROR and
 * a message; the context stays valid, a later call succeeds. */
static void xgpu_xengine_faults(void) {
  XGPUInfo info;
  xgpuInfo(&info);
  int failures = 0;
  for (int c = 0; c < NCAT; ++c)
    for (int n = 0; n < 8; ++n) {
      xgh_fake_reset();
      XGPUContext ctx;
      memset(&ctx, 0, sizeof ctx);
      CHECK(xgpuInit(&ctx, 0) == XGPU_OK, "init");
      fill((int8_t *)ctx.array_h, info.vecLength * 2, 2);
      xgh_fake_fail(categories[c], n);
      const int rc = xgpuCudaXengine(&ctx, SYNCOP_DUMP);
      if (rc != XGPU_OK) {
        ++failures;
        CHECK(rc == XGPU_CUDA_ERROR, "code %d (%s #%d)", rc, categories[c], n);
        CHECK(strstr(xgpuAmdLastError(), "injected") != NULL, "message '%s'", xgpuAmdLastError());
      }
      xgh_fake_reset();
      CHECK(xgpuClearDeviceIntegrationBuffer(&ctx) == XGPU_OK, "clear");
      CHECK(xgpuCudaXengine(&ctx, SYNCOP_DUMP) == XGPU_OK, "retry after %s #%d", categories[c], n);
      xgpuFree(&ctx);
      CHECK(xgh_fake_live_allocations() == 0, "leaked %lld", xgh_fake_live_allocations());
    }
  CHECK(failures > 0, "no injected fault reached xgpuCudaXengine");
  printf("ok xgpuCudaXengine returns XGPU_CUDA_ERROR on %d injected faults and recovers\n",
         failures);
}

/* ---- runtime-sized engine -------------------------------------------------- */
static XGPUAmdConfig config(unsigned ns, unsigned nf, unsigned order) {
  XGPUAmdConfig c;
  memset(&c, 0, sizeof c);
  c.npol = 2;
  c.nstation = ns;
  c.nfrequency = nf;
  c.compute_type = XGPU_INT8;
  c.matrix_order = order;
  return c;
}

static void engine_create_faults(void) {
  int failures = 0;
  for (int c = 0; c < NCAT; ++c)
    for (int n = 0; n < 12; ++n) {
      xgh_fake_reset();
      xgh_fake_fail(categories[c], n);
      XGPUAmdConfig cfg = config(8, 2, REGISTER_TILE_TRIANGULAR_ORDER);
      XGPUAmdEngine *e = NULL;
      const int rc = xgpuAmdCreate(&e, &cfg);
      if (rc) {
        ++failures;
        CHECK(e == NULL && rc == XGPU_CUDA_ERROR, "rc %d", rc);
      }
      xgpuAmdDestroy(e);
      CHECK(xgh_fake_live_allocations() == 0, "leaked %lld", xgh_fake_live_allocations());
    }
  xgh_fake_reset();
  printf("ok xgpuAmdCreate fails cleanly on %d injected faults\n", failures);
}

/* Long integrations spill to int64; a failing spill allocation or dump copy
 * is reported, and the exact dump afterwards is right. */
static void engine_spill_and_dump(void) {
  const int ns = 6, nf = 2, nt = 40000;
  const size_t nb = (size_t)nt * nf * ns * 4;
  int8_t *x = malloc(nb);
  fill(x, nb, 3);
  int64_t *ref = malloc(sizeof(int64_t) * nf * ns * (ns + 1) / 2 * 8);
  exact(ref, x, ns, nf, nt);
  XGPUAmdConfig cfg = config(ns, nf, TRIANGULAR_ORDER);
  XGPUAmdEngine *e = NULL;
  CHECK(xgpuAmdCreate(&e, &cfg) == XGPU_OK, "create");
  void *d = NULL;
  CHECK(xgpuAmdMalloc(&d, nb) == XGPU_OK, "malloc");
  CHECK(xgpuAmdMemcpy(d, x, nb, XGPU_H2D) == XGPU_OK, "upload");
  CHECK(xgpuAmdCorrelate(e, d, nt) == XGPU_OK, "first");
  xgh_fake_fail("malloc", 0); /* the spill buffer of the second call */
  CHECK(xgpuAmdCorrelate(e, d, nt) == XGPU_CUDA_ERROR, "spill allocation failure");
  xgh_fake_reset();
  xgpuAmdClear(e);
  for (int k = 0; k < 3; ++k) CHECK(xgpuAmdCorrelate(e, d, nt) == XGPU_OK, "call %d", k);
  const size_t n = (size_t)nf * ns * (ns + 1) / 2 * 8;
  int64_t *got = malloc(n * sizeof(int64_t));
  xgh_fake_fail("copy", 0);
  CHECK(xgpuAmdDump(e, got, XGPU_DUMP_WIDE) == XGPU_CUDA_ERROR, "dump copy failure");
  xgh_fake_reset();
  CHECK(xgpuAmdDump(e, got, XGPU_DUMP_WIDE) == XGPU_OK, "dump");
  for (size_t i = 0; i < n; ++i) CHECK(got[i] == 3 * ref[i], "value %zu", i);
  free(got);
  free(ref);
  free(x);
  xgpuAmdFree(d);
  xgpuAmdDestroy(e);
  CHECK(xgh_fake_live_allocations() == 0, "leaked %lld", xgh_fake_live_allocations());
  printf("ok spill to int64 and dumps report injected faults, then stay exact\n");
}

/* graph-captured device steps, a batch and the allocation-free entry point */
static void engine_graphs_batch_noalloc(void) {
  const int ns = 8, nf = 3, nt = 128;
  const size_t nb = (size_t)nt * nf * ns * 4;
  int8_t *x = malloc(nb);
  fill(x, nb, 4);
  const size_t n = (size_t)nf * ns * (ns + 1) / 2 * 8;
  int64_t *ref = malloc(n * sizeof(int64_t)), *got = malloc(n * sizeof(int64_t));
  exact(ref, x, ns, nf, nt);
  void *d = NULL;
  CHECK(xgpuAmdMalloc(&d, nb) == XGPU_OK, "malloc");
  CHECK(xgpuAmdMemcpy(d, x, nb, XGPU_H2D) == XGPU_OK, "upload");
  XGPUAmdConfig cfg = config(ns, nf, REGISTER_TILE_TRIANGULAR_ORDER);
  XGPUAmdEngine *a = NULL, *b = NULL;
  CHECK(xgpuAmdCreate(&a, &cfg) == XGPU_OK && xgpuAmdCreate(&b, &cfg) == XGPU_OK, "create");
  XGPUAmdTuning t;
  xgpuAmdGetTuning(a, &t);
  t.graph = 1;
  CHECK(xgpuAmdSetTuning(a, &t) == XGPU_OK, "graph knob");
  for (int rep = 0; rep < 3; ++rep) {
    xgpuAmdClear(a);
    CHECK(xgpuAmdCorrelate(a, d, nt) == XGPU_OK, "graph step");
    CHECK(xgpuAmdCorrelate(a, d, nt) == XGPU_OK, "graph step (accumulate)");
    CHECK(xgpuAmdDump(a, got, XGPU_DUMP_WIDE | XGPU_DUMP_TRIANGULAR) == XGPU_OK, "dump");
    for (size_t i = 0; i < n; ++i) CHECK(got[i] == 2 * ref[i], "graph rep %d value %zu", rep, i);
  }
  XGPUAmdBatchItem items[2] = {{a, d, (unsigned)nt}, {b, d, (unsigned)nt}};
  xgpuAmdClear(a);
  for (int rep = 0; rep < 2; ++rep) CHECK(xgpuAmdCorrelateBatch(items, 2, NULL) == XGPU_OK, "batch");
  CHECK(xgpuAmdDump(b, got, XGPU_DUMP_WIDE | XGPU_DUMP_TRIANGULAR) == XGPU_OK, "dump b");
  for (size_t i = 0; i < n; ++i) CHECK(got[i] == 2 * ref[i], "batch value %zu", i);
  /* a failing launch of an eager batch (owner b: no graph knob), and a
   * failing replay of a's captured batch */
  XGPUAmdBatchItem eager[2] = {{b, d, (unsigned)nt}, {a, d, (unsigned)nt}};
  xgh_fake_fail("launch", 0);
  CHECK(xgpuAmdCorrelateBatch(eager, 2, NULL) == XGPU_CUDA_ERROR, "batch launch failure");
  xgh_fake_fail("graph", 0);
  CHECK(xgpuAmdCorrelateBatch(items, 2, NULL) == XGPU_CUDA_ERROR, "batch replay failure");
  xgh_fake_reset();
  xgpuAmdClear(b);
  CHECK(xgpuAmdCorrelateNoAlloc(b, d, nt) == XGPU_OK, "no-alloc");
  CHECK(xgpuAmdDump(b, got, XGPU_DUMP_WIDE | XGPU_DUMP_TRIANGULAR) == XGPU_OK, "dump");
  for (size_t i = 0; i < n; ++i) CHECK(got[i] == ref[i], "no-alloc value %zu", i);
  const long long mallocs = xgh_fake_calls("malloc");
  CHECK(xgpuAmdCorrelateNoAlloc(b, d, nt) == XGPU_OK, "no-alloc again");
  CHECK(xgh_fake_calls("malloc") == mallocs, "xgpuAmdCorrelateNoAlloc allocated");
  /* the autotuner keeps the caller's integration */
  XGPUAmdTuning best;
  float ms = 0.f;
  CHECK(xgpuAmdAutotune(b, d, nt, 0, 1, NULL, &best, &ms) == XGPU_OK, "autotune");
  CHECK(xgpuAmdDump(b, got, XGPU_DUMP_WIDE | XGPU_DUMP_TRIANGULAR) == XGPU_OK, "dump");
  for (size_t i = 0; i < n; ++i) CHECK(got[i] == 2 * ref[i], "after autotune value %zu", i);
  xgpuAmdDestroy(a);
  xgpuAmdDestroy(b);
  xgpuAmdFree(d);
  free(x);
  free(ref);
  free(got);
  CHECK(xgh_fake_live_allocations() == 0, "leaked %lld", xgh_fake_live_allocations());
  printf("ok graph steps, batch, allocation-free entry point and autotuner on the fake\n");
}

/* ---- server --------------------------------------------------------------- */
static void server_runs_and_reports_faults(void) {
  const int ns = 8, nf = 2, nt = 64, K = 3;
  const size_t nb = (size_t)nt * nf * ns * 4;
  int8_t *x = malloc(nb);
  fill(x, nb, 5);
  const size_t n = (size_t)nf * ns * (ns + 1) / 2 * 8;
  int64_t *ref = malloc(n * sizeof(int64_t));
  exact(ref, x, ns, nf, nt);
  /* mode 0: host blocks, host ring; 1: the same with a failing launch;
   * 2: device blocks, device ring */
  for (int mode = 0; mode < 3; ++mode) {
    const int inject = mode == 1, dev = mode == 2;
    XGPUServerConfig sc = {0, (size_t)1 << 20, 0, 1};
    XGPUServer *s = NULL;
    CHECK(xgpuServerCreate(&s, &sc) == XGPU_OK, "server");
    XGPUStreamConfig st = {ns, nf, XGPU_INT8, K, (unsigned)dev};
    int64_t *ring = malloc(n * sizeof(int64_t) * 2);
    void *ring_d = NULL, *x_d = NULL;
    if (dev) {
      CHECK(xgpuAmdMalloc(&ring_d, n * sizeof(int64_t) * 2) == XGPU_OK, "device ring");
      CHECK(xgpuAmdMalloc(&x_d, nb) == XGPU_OK && xgpuAmdMemcpy(x_d, x, nb, XGPU_H2D) == XGPU_OK,
            "device block");
    }
    int id = -1;
    CHECK(xgpuServerOpenStream(s, &st, dev ? ring_d : (void *)ring, 2, &id) == XGPU_OK, "open");
    if (inject) xgh_fake_fail("launch", 1);
    int rc = XGPU_OK;
    for (int k = 0; k < 2 * K && rc == XGPU_OK; ++k)
      rc = dev ? xgpuServerSubmitDevice(s, id, x_d, nt) : xgpuServerSubmit(s, id, x, nt);
    const int drc = xgpuServerDrain(s);
    if (inject) {
      CHECK(drc == XGPU_CUDA_ERROR || rc != XGPU_OK, "the failure is reported");
    } else {
      CHECK(drc == XGPU_OK, "drain");
      CHECK(xgpuServerDumpsDone(s, id) == 2, "dumps");
      if (dev)
        CHECK(xgpuAmdMemcpy(ring, ring_d, n * sizeof(int64_t) * 2, XGPU_D2H) == XGPU_OK, "ring");
      for (int dmp = 0; dmp < 2; ++dmp)
        for (size_t i = 0; i < n; ++i)
          CHECK(ring[dmp * n + i] == K * ref[i], "mode %d dump %d %zu", mode, dmp, i);
    }
    xgh_fake_reset();
    xgpuServerDestroy(s);
    free(ring);
    if (dev) {
      xgpuAmdFree(ring_d);
      xgpuAmdFree(x_d);
    }
    CHECK(xgh_fake_live_allocations() == 0, "leaked %lld", xgh_fake_live_allocations());
  }
  free(ref);
  free(x);
  printf("ok server: exact dumps (host and device rings), and an injected launch failure is "
         "reported by drain\n");
}

/* ---- randomized serving ----------------------------------------------------
 * Streams of random shapes, host or device blocks (device blocks back to back
 * in memory, so they fuse), host or device dump rings of 2 slots, random
 * server limits, submitted by two threads in a random interleaving (each
 * thread owns half the streams, so each stream's blocks stay in order).  Every
 * stream's last two dumps are checked against the exact sums of their
 * integrations' blocks. */
#define RS_STREAMS 6
#define RS_DUMPS 3
typedef struct {
  int ns, nf, nt, K, dev, dump_dev, id;
  size_t nb, nout;
  int8_t *x;          /* RS_DUMPS * K blocks, back to back */
  void *x_d, *ring_d;
  int64_t *ring;
} RsStream;

typedef struct {
  XGPUServer *s;
  RsStream *st;
  int first, count;
  uint64_t seed;
  int rc;
} RsSubmitter;

static uint64_t rs_next(uint64_t *s) {
  *s ^= *s << 13, *s ^= *s >> 7, *s ^= *s << 17;
  return *s;
}

static void *rs_submit(void *arg) {
  RsSubmitter *w = (RsSubmitter *)arg;
  int next[RS_STREAMS] = {0};
  int left = 0;
  for (int i = 0; i < w->count; ++i) left += RS_DUMPS * w->st[w->first + i].K;
  uint64_t r = w->seed;
  while (left > 0 && w->rc == XGPU_OK) {
    const int i = w->first + (int)(rs_next(&r) % (uint64_t)w->count);
    RsStream *q = &w->st[i];
    if (next[i] == RS_DUMPS * q->K) continue;
    const int b = next[i]++;
    w->rc = q->dev ? xgpuServerSubmitDevice(w->s, q->id, (char *)q->x_d + (size_t)b * q->nb, q->nt)
                   : xgpuServerSubmit(w->s, q->id, q->x + (size_t)b * q->nb, q->nt);
    --left;
  }
  return NULL;
}

static void server_randomized(uint64_t seed) {
  uint64_t r = 0x9E3779B97F4A7C15ull ^ (seed * 0xD1B54A32D192ED03ull);
  static const int NS[] = {4, 8, 12, 16}, NT[] = {16, 32, 48};
  RsStream st[RS_STREAMS];
  XGPUServerConfig sc = {0, (rs_next(&r) & 1) ? ((size_t)64 << 10) : ((size_t)1 << 20),
                         1 + (int)(rs_next(&r) % 16), 1 + (int)(rs_next(&r) % 3)};
  XGPUServer *s = NULL;
  CHECK(xgpuServerCreate(&s, &sc) == XGPU_OK, "server");
  for (int i = 0; i < RS_STREAMS; ++i) {
    RsStream *q = &st[i];
    q->ns = NS[rs_next(&r) % 4];
    q->nf = 1 + (int)(rs_next(&r) % 3);
    q->nt = NT[rs_next(&r) % 3];
    q->K = 1 + (int)(rs_next(&r) % 6);
    q->dev = (int)(rs_next(&r) & 1);
    q->dump_dev = (int)(rs_next(&r) & 1);
    q->nb = (size_t)q->nt * q->nf * q->ns * 4;
    q->nout = (size_t)q->nf * q->ns * (q->ns + 1) / 2 * 8;
    q->x = malloc(RS_DUMPS * q->K * q->nb);
    fill(q->x, RS_DUMPS * q->K * q->nb, (unsigned)(1000 * seed + i));
    q->x_d = q->ring_d = NULL;
    q->ring = calloc(2 * q->nout, sizeof(int64_t));
    if (q->dev) {
      CHECK(xgpuAmdMalloc(&q->x_d, RS_DUMPS * q->K * q->nb) == XGPU_OK, "blocks");
      CHECK(xgpuAmdMemcpy(q->x_d, q->x, RS_DUMPS * q->K * q->nb, XGPU_H2D) == XGPU_OK, "upload");
    }
    if (q->dump_dev) CHECK(xgpuAmdMalloc(&q->ring_d, 2 * q->nout * sizeof(int64_t)) == XGPU_OK, "ring");
    XGPUStreamConfig cfg = {(unsigned)q->ns, (unsigned)q->nf, XGPU_INT8, (unsigned)q->K,
                            (unsigned)q->dump_dev};
    CHECK(xgpuServerOpenStream(s, &cfg, q->dump_dev ? q->ring_d : (void *)q->ring, 2, &q->id) == XGPU_OK,
          "open: %s", xgpuAmdLastError());
  }
  RsSubmitter w[2] = {{s, st, 0, RS_STREAMS / 2, seed * 2 + 1, XGPU_OK},
                      {s, st, RS_STREAMS / 2, RS_STREAMS - RS_STREAMS / 2, seed * 2 + 2, XGPU_OK}};
  pthread_t th[2];
  for (int t = 0; t < 2; ++t) CHECK(pthread_create(&th[t], NULL, rs_submit, &w[t]) == 0, "thread");
  for (int t = 0; t < 2; ++t) pthread_join(th[t], NULL);
  CHECK(w[0].rc == XGPU_OK && w[1].rc == XGPU_OK, "submit: %s", xgpuAmdLastError());
  CHECK(xgpuServerDrain(s) == XGPU_OK, "drain: %s", xgpuAmdLastError());
  XGPUServerStats stats;
  CHECK(xgpuServerStats(s, &stats) == XGPU_OK, "stats");
  unsigned long long total = 0;
  for (int i = 0; i < RS_STREAMS; ++i) total += RS_DUMPS * st[i].K;
  CHECK(stats.blocks_done == total, "blocks done %llu of %llu", stats.blocks_done, total);
  for (int i = 0; i < RS_STREAMS; ++i) {
    RsStream *q = &st[i];
    CHECK(xgpuServerDumpsDone(s, q->id) == RS_DUMPS, "stream %d dumps", i);
    if (q->dump_dev)
      CHECK(xgpuAmdMemcpy(q->ring, q->ring_d, 2 * q->nout * sizeof(int64_t), XGPU_D2H) == XGPU_OK, "ring");
    int64_t *ref = calloc(q->nout, sizeof(int64_t)), *one = malloc(q->nout * sizeof(int64_t));
    for (int d = RS_DUMPS - 2; d < RS_DUMPS; ++d) {  /* dump d sits in slot d % 2 */
      memset(ref, 0, q->nout * sizeof(int64_t));
      for (int b = 0; b < q->K; ++b) {
        exact(one, q->x + (size_t)(d * q->K + b) * q->nb, q->ns, q->nf, q->nt);
        for (size_t k = 0; k < q->nout; ++k) ref[k] += one[k];
      }
      const int64_t *got = q->ring + (size_t)(d % 2) * q->nout;
      for (size_t k = 0; k < q->nout; ++k)
        CHECK(got[k] == ref[k], "seed %llu stream %d (ns %d nf %d nt %d K %d dev %d ring %d) dump %d "
              "value %zu", (unsigned long long)seed, i, q->ns, q->nf, q->nt, q->K, q->dev, q->dump_dev,
              d, k);
    }
    free(ref);
    free(one);
  }
  xgpuServerDestroy(s);
  for (int i = 0; i < RS_STREAMS; ++i) {
    free(st[i].x);
    free(st[i].ring);
    if (st[i].x_d) xgpuAmdFree(st[i].x_d);
    if (st[i].ring_d) xgpuAmdFree(st[i].ring_d);
  }
  CHECK(xgh_fake_live_allocations() == 0, "leaked %lld", xgh_fake_live_allocations());
}

static void server_randomized_runs(int nseeds) {
  for (int seed = 1; seed <= nseeds; ++seed) server_randomized((uint64_t)seed);
  printf("ok randomized serving: %d seeds of streams, limits and submission orders, every dump exact\n",
         nseeds);
}

/* ---- XGPU_TRACE (run as "test_fault_injection trace <file>": the switch is
 * read once per process) ---------------------------------------------------- */
static int count_lines(const char *path, const char *needle) {
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  char line[4096];
  int n = 0;
  while (fgets(line, sizeof line, f))
    if (strstr(line, needle)) ++n;
  fclose(f);
  return n;
}

static void trace_log(const char *path) {
  const int ns = 8, nf = 3, nt = 64;
  const size_t nb = (size_t)nt * nf * ns * 4;
  int8_t *x = malloc(nb);
  fill(x, nb, 5);
  const size_t n = (size_t)nf * ns * (ns + 1) / 2 * 8;
  int64_t *got = malloc(n * sizeof(int64_t));
  void *d = NULL;
  CHECK(xgpuAmdMalloc(&d, nb) == XGPU_OK && xgpuAmdMemcpy(d, x, nb, XGPU_H2D) == XGPU_OK, "upload");
  XGPUAmdConfig cfg = config(ns, nf, REGISTER_TILE_TRIANGULAR_ORDER);
  XGPUAmdEngine *e = NULL;
  CHECK(xgpuAmdCreate(&e, &cfg) == XGPU_OK, "create");
  CHECK(xgpuAmdCorrelate(e, d, nt) == XGPU_OK, "eager");             /* 1 launch */
  CHECK(xgpuAmdDump(e, got, XGPU_DUMP_WIDE | XGPU_DUMP_TRIANGULAR) == XGPU_OK, "dump");
  XGPUAmdTuning t;
  xgpuAmdGetTuning(e, &t);
  t.graph = 1;
  CHECK(xgpuAmdSetTuning(e, &t) == XGPU_OK, "graph knob");
  CHECK(xgpuAmdClear(e) == XGPU_OK, "clear");
  CHECK(xgpuAmdCorrelate(e, d, nt) == XGPU_OK, "capture + replay"); /* store step */
  CHECK(xgpuAmdClear(e) == XGPU_OK, "clear");
  CHECK(xgpuAmdCorrelate(e, d, nt) == XGPU_OK, "replay");           /* same step again */
  xgpuAmdDestroy(e);
  xgpuAmdFree(d);
  free(x);
  free(got);
  CHECK(count_lines(path, "[xgpu-trace]") >= 5, "trace lines");
  CHECK(count_lines(path, " launch kernel=") == 1, "eager launch lines: %d", count_lines(path, " launch kernel="));
  CHECK(count_lines(path, " capture kernel=") == 1, "capture lines");
  CHECK(count_lines(path, " replay kernel=") == 2, "replay lines");
  CHECK(count_lines(path, "nf=3 ns=8 nt=64 store") == 4, "job fields");
  CHECK(count_lines(path, " dump engine=") == 1, "dump lines");
  printf("ok XGPU_TRACE logs launches, captures, replays and dumps\n");
}

int main(int argc, char **argv) {
  if (argc == 3 && strcmp(argv[1], "trace") == 0) {
    remove(argv[2]);
    setenv("XGPU_TRACE", argv[2], 1);
    trace_log(argv[2]);
    return 0;
  }
  if (argc == 3 && strcmp(argv[1], "serve") == 0) {  /* the randomized serving runs only */
    server_randomized_runs(atoi(argv[2]));
    return 0;
  }
  xgpu_h_pipeline();
  xgpu_init_faults();
  xgpu_xengine_faults();
  engine_create_faults();
  engine_spill_and_dump();
  engine_graphs_batch_noalloc();
  server_runs_and_reports_faults();
  server_randomized_runs(8);
  printf("all fault-injection scenarios passed\n");
  return 0;
}
