/*
 * autotune.c -- kernel-parameter autotuner and tuning cache (no xGPU
 * counterpart).
 *
 * Device path (xgpuAmdAutotune): the engine's kernel knobs, for int8 the
 * kernel variant (complex-plane kernel x persistent workgroups per CU x L2
 * prefetch; the real-Gram kernel x staging path (registers or global->LDS
 * DMA) x LDS stage depth, and the DMA kernel's MFMA tile shape; the
 * K-interleaved kernel), for fp16 the same with the fp16 real-Gram kernel's
 * stage depths, for the FP32 plain path workgroups per CU.  Each is timed with
 * HIP events over `reps` single-kernel integrations of the caller's device
 * input.  Host path (xgpuAmdAutotuneHost): the time-pipe length (samples per
 * transfer and launch) x hipGraph replay, timed on the wall clock per
 * xgpuAmdCorrelateHost call.  Candidates run on a scratch engine of the same
 * configuration, so the caller's integration is left as it was; the fastest
 * is kept on the caller's engine.  The cache is a text file, one line per
 * result:
 *   <device>|<ns>|<nf>|<ntime>|<compute>[|host]|<wg_per_cu>|<prefetch>|<kernel>|<stage>|
 *   <staging>|<ntimepipe>|<mfma>|<graph>|<ms>
 * Later lines win, so re-tuning simply appends.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "engine_internal.h"
#include "xg_hip.h"
#include "xgpu_amd.h"

static void key_of(const XGPUAmdEngine *e, unsigned ntime, int host, char *buf, size_t len) {
  char name[128] = "unknown";
  xgh_device_name(e->cfg.device, name, sizeof name);
  for (char *p = name; *p; ++p)
    if (*p == '|' || *p == '\n') *p = '_';
  snprintf(buf, len, "%s|%u|%u|%u|%u%s", name, e->cfg.nstation, e->cfg.nfrequency, ntime,
           e->cfg.compute_type, host ? "|host" : "");
}

static int append_cache(const XGPUAmdEngine *e, const char *path, unsigned ntime, int host,
                        const XGPUAmdTuning *t, float ms) {
  FILE *f = fopen(path, "a");
  if (!f) return xg_fail(XGPU_INVALID_CONFIG, "cannot write tuning cache %s", path);
  char key[256];
  key_of(e, ntime, host, key, sizeof key);
  fprintf(f, "%s|%d|%d|%d|%d|%d|%d|%d|%d|%.4f\n", key, t->wg_per_cu, t->prefetch, t->kernel,
          t->stage, t->staging, t->ntimepipe, t->mfma, t->graph, ms);
  fclose(f);
  return XGPU_OK;
}

/* an engine of the same configuration, for timing candidates */
static int scratch_engine(const XGPUAmdEngine *e, XGPUAmdEngine **out) {
  XGPUAmdConfig c = e->cfg;
  return xgpuAmdCreate(out, &c);
}

/* The last cached line of `key`, parsed into *t; 1 when found. */
static int find_cached(FILE *f, const char *key, XGPUAmdTuning *t) {
  const size_t klen = strlen(key);
  char line[512];
  int found = 0;
  while (fgets(line, sizeof line, f)) {
    if (strncmp(line, key, klen) != 0 || line[klen] != '|') continue;
    int v[8];
    float ms = 0.f;
    if (sscanf(line + klen + 1, "%d|%d|%d|%d|%d|%d|%d|%d|%f", &v[0], &v[1], &v[2], &v[3], &v[4],
               &v[5], &v[6], &v[7], &ms) == 9) {
      t->wg_per_cu = v[0];
      t->prefetch = v[1];
      t->kernel = v[2];
      t->stage = v[3];
      t->staging = v[4];
      t->ntimepipe = v[5];
      t->mfma = v[6];
      t->graph = v[7];
      found = 1;
    }
  }
  return found;
}

int xgpuAmdLoadTuning(XGPUAmdEngine *e, const char *path, unsigned ntime) {
  if (!e || !path) return xg_fail(XGPU_INVALID_CONFIG, "NULL argument");
  FILE *f = fopen(path, "r");
  if (!f) return xg_fail(XGPU_INVALID_CONFIG, "no tuning cache at %s", path);
  char key[256];
  key_of(e, ntime, 0, key, sizeof key);
  XGPUAmdTuning dev, host, cur;
  memset(&dev, 0, sizeof dev);
  memset(&host, 0, sizeof host);
  const int have_dev = find_cached(f, key, &dev);
  rewind(f);
  char hkey[256];
  key_of(e, e->cfg.ntime, 1, hkey, sizeof hkey);
  const int have_host = e->cfg.ntime > 0 && find_cached(f, hkey, &host);
  fclose(f);
  if (!have_dev) return xg_fail(XGPU_INVALID_CONFIG, "no cached tuning for %s", key);
  xgpuAmdGetTuning(e, &cur);
  /* the host-path knobs from the host-path line when there is one */
  dev.ntimepipe = have_host ? host.ntimepipe : cur.ntimepipe;
  dev.graph = have_host ? host.graph : cur.graph;
  return xgpuAmdSetTuning(e, &dev);
}

int xgpuAmdAutotune(XGPUAmdEngine *e, const void *input_d, unsigned ntime, int warmup, int reps,
                    const char *cache_path, XGPUAmdTuning *best, float *best_ms) {
  if (!e || !input_d || ntime == 0 || reps < 1)
    return xg_fail(XGPU_INVALID_CONFIG, "bad autotune arguments");
  XGPUAmdTuning keep;
  xgpuAmdGetTuning(e, &keep);
  XGPUAmdTuning win = keep;
  float win_ms = 0.f;
  int have = 0;
  XGPUAmdTuning cand[16];
  int ncand = 0;
  keep.mfma = 0;
  const int npf = (e->family == XG_KERNEL_F32_PLAIN) ? 1 : 2;
  for (int wg = 1; wg <= 2; ++wg)
    for (int pf = 1; pf <= npf; ++pf) {
      XGPUAmdTuning t = keep;
      t.wg_per_cu = wg;
      t.prefetch = (e->family == XG_KERNEL_F32_PLAIN) ? 2 : pf;
      t.kernel = 1;
      t.stage = 0;
      t.staging = 0;
      cand[ncand++] = t;
    }
  XGPUAmdEngine *x = NULL;
  int rc = scratch_engine(e, &x);
  if (rc) return rc;
  if (e->family != XG_KERNEL_F32_PLAIN) {
    XGPUAmdTuning probe = keep;
    probe.kernel = 2;
    probe.staging = 0;
    probe.stage = 0;
    if ((rc = xgpuAmdSetTuning(x, &probe))) {
      xgpuAmdDestroy(x);
      return rc;
    }
    /* staging, stage */
    static const int v_i8[4][2] = {{1, 64}, {1, 128}, {2, 32}, {2, 64}};
    static const int v_f16[4][2] = {{1, 32}, {1, 64}, {2, 16}, {2, 32}};
    const int fam = xg_launch_family(x, input_d);
    const int nv = fam == XG_KERNEL_I8_RG ? 4 : fam == XG_KERNEL_F16_RG ? 4 : 0;
    for (int v = 0; v < nv; ++v) {
      XGPUAmdTuning t = keep;
      t.kernel = 2;
      t.staging = fam == XG_KERNEL_I8_RG ? v_i8[v][0] : v_f16[v][0];
      t.stage = fam == XG_KERNEL_I8_RG ? v_i8[v][1] : v_f16[v][1];
      cand[ncand++] = t;
    }
    if (fam == XG_KERNEL_I8_RG) { /* 16x16x64 MFMA tiles in the DMA kernel */
      XGPUAmdTuning t = keep;
      t.kernel = 2;
      t.staging = 2;
      t.stage = 64;
      t.mfma = 2;
      cand[ncand++] = t;
    }
    if (fam == XG_KERNEL_I8_RG) { /* the K-interleaved kernel takes the same inputs */
      XGPUAmdTuning t = keep;
      t.kernel = 3;
      t.staging = 0;
      t.stage = 0;
      cand[ncand++] = t;
    }
  }
  for (int c = 0; c < ncand && !rc; ++c) {
    const XGPUAmdTuning t = cand[c];
    float ms = 0.f;
    if ((rc = xgpuAmdSetTuning(x, &t))) break;
    if ((rc = xgpuAmdTimeCorrelate(x, input_d, ntime, warmup, reps, &ms))) break;
    if (!have || ms < win_ms) {
      have = 1;
      win_ms = ms;
      win = t;
    }
  }
  xgpuAmdDestroy(x);
  if (rc) return rc;
  if ((rc = xgpuAmdSetTuning(e, &win))) return rc;
  xgpuAmdGetTuning(e, &win);
  if (best) *best = win;
  if (best_ms) *best_ms = win_ms;
  if (cache_path && (rc = append_cache(e, cache_path, ntime, 0, &win, win_ms))) return rc;
  return XGPU_OK;
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int xgpuAmdAutotuneHost(XGPUAmdEngine *e, const void *input_h, int warmup, int reps,
                        const char *cache_path, XGPUAmdTuning *best, float *best_ms) {
  if (!e || !input_h || reps < 1 || e->cfg.ntime == 0)
    return xg_fail(XGPU_INVALID_CONFIG, "bad host autotune arguments (the engine needs an ntime)");
  XGPUAmdTuning keep;
  xgpuAmdGetTuning(e, &keep);
  XGPUAmdEngine *x = NULL;
  int rc = scratch_engine(e, &x);
  if (rc) return rc;
  XGPUAmdTuning win = keep;
  double win_s = 0.0;
  int have = 0;
  const unsigned nt = e->cfg.ntime;
  for (unsigned d = 1; d <= 64 && !rc; d *= 2) {
    if (nt % d) break;
    const unsigned ntp = nt / d;
    if (ntp < 32 && d > 1) break;
    if (e->swizzled && ntp % 4) continue;
    for (int graph = 1; graph <= 2 && !rc; ++graph) {
      XGPUAmdTuning t = keep;
      t.ntimepipe = (int)ntp;
      t.graph = graph;
      if ((rc = xgpuAmdSetTuning(x, &t))) break;
      for (int i = 0; i < warmup && !rc; ++i)
        rc = xgpuAmdCorrelateHost(x, input_h, SYNCOP_SYNC_COMPUTE, NULL);
      const double t0 = now_s();
      for (int i = 0; i < reps && !rc; ++i)
        rc = xgpuAmdCorrelateHost(x, input_h, SYNCOP_SYNC_COMPUTE, NULL);
      const double sec = (now_s() - t0) / reps;
      if (!rc && (!have || sec < win_s)) {
        have = 1;
        win_s = sec;
        win = t;
      }
      xgpuAmdClear(x);
    }
  }
  xgpuAmdDestroy(x);
  if (rc) return rc;
  if ((rc = xgpuAmdSetTuning(e, &win))) return rc;
  xgpuAmdGetTuning(e, &win);
  if (best) *best = win;
  if (best_ms) *best_ms = (float)(win_s * 1e3);
  if (cache_path && (rc = append_cache(e, cache_path, nt, 1, &win, (float)(win_s * 1e3)))) return rc;
  return XGPU_OK;
}
