/*
 * autotune.c -- kernel-parameter autotuner and tuning cache (no xGPU
 * counterpart).
 *
 * Candidates are the engine's kernel knobs: for int8 the kernel variant
 * (complex-plane kernel x persistent workgroups per CU x L2 prefetch, and
 * the real-Gram kernel x staging path (registers or global->LDS DMA) x LDS
 * stage depth where it applies), for fp16 the same with the fp16 real-Gram
 * kernel's stage depths, for the FP32 plain path workgroups per CU.  Each is timed with HIP events on the
 * engine stream over `reps` single-kernel integrations of the caller's
 * device input; the fastest is kept.  The cache is a text file, one line per
 * result:
 *   <device>|<ns>|<nf>|<ntime>|<compute>|<wg_per_cu>|<prefetch>|<kernel>|<stage>|<staging>|<ms>
 * Later lines win, so re-tuning simply appends.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "engine_internal.h"
#include "xg_hip.h"
#include "xgpu_amd.h"

static void key_of(const XGPUAmdEngine *e, unsigned ntime, char *buf, size_t len) {
  char name[128] = "unknown";
  xgh_device_name(e->cfg.device, name, sizeof name);
  for (char *p = name; *p; ++p)
    if (*p == '|' || *p == '\n') *p = '_';
  snprintf(buf, len, "%s|%u|%u|%u|%u", name, e->cfg.nstation, e->cfg.nfrequency, ntime,
           e->cfg.compute_type);
}

int xgpuAmdLoadTuning(XGPUAmdEngine *e, const char *path, unsigned ntime) {
  if (!e || !path) return xg_fail(XGPU_INVALID_CONFIG, "NULL argument");
  FILE *f = fopen(path, "r");
  if (!f) return xg_fail(XGPU_INVALID_CONFIG, "no tuning cache at %s", path);
  char key[256], line[512];
  key_of(e, ntime, key, sizeof key);
  const size_t klen = strlen(key);
  int found = 0;
  XGPUAmdTuning t;
  memset(&t, 0, sizeof t);
  while (fgets(line, sizeof line, f)) {
    if (strncmp(line, key, klen) == 0 && line[klen] == '|') {
      int wg = 0, pf = 0, kern = 0, stg = 0, stgg = 0;
      float ms = 0.f;
      if (sscanf(line + klen + 1, "%d|%d|%d|%d|%d|%f", &wg, &pf, &kern, &stg, &stgg, &ms) == 6) {
        t.wg_per_cu = wg;
        t.prefetch = pf;
        t.kernel = kern;
        t.stage = stg;
        t.staging = stgg;
        found = 1;
      }
    }
  }
  fclose(f);
  if (!found) return xg_fail(XGPU_INVALID_CONFIG, "no cached tuning for %s", key);
  XGPUAmdTuning cur;
  xgpuAmdGetTuning(e, &cur);
  t.graph = cur.graph;
  return xgpuAmdSetTuning(e, &t);
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
  if (e->family != XG_KERNEL_F32_PLAIN) {
    XGPUAmdTuning probe = keep;
    probe.kernel = 2;
    probe.staging = 0;
    probe.stage = 0;
    xgpuAmdSetTuning(e, &probe);
    /* staging, stage */
    static const int v_i8[4][2] = {{1, 64}, {1, 128}, {2, 32}, {2, 64}};
    static const int v_f16[3][2] = {{1, 32}, {1, 64}, {2, 32}};
    const int fam = xg_launch_family(e, input_d);
    const int nv = fam == XG_KERNEL_I8_RG ? 4 : fam == XG_KERNEL_F16_RG ? 3 : 0;
    for (int v = 0; v < nv; ++v) {
      XGPUAmdTuning t = keep;
      t.kernel = 2;
      t.staging = fam == XG_KERNEL_I8_RG ? v_i8[v][0] : v_f16[v][0];
      t.stage = fam == XG_KERNEL_I8_RG ? v_i8[v][1] : v_f16[v][1];
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
  for (int c = 0; c < ncand; ++c) {
    const XGPUAmdTuning t = cand[c];
    int rc = xgpuAmdSetTuning(e, &t);
    if (rc) return rc;
    float ms = 0.f;
    rc = xgpuAmdTimeCorrelate(e, input_d, ntime, warmup, reps, &ms);
    if (rc) {
      xgpuAmdSetTuning(e, &keep);
      return rc;
    }
    if (!have || ms < win_ms) {
      have = 1;
      win_ms = ms;
      win = t;
    }
  }
  int rc = xgpuAmdSetTuning(e, &win);
  if (rc) return rc;
  xgpuAmdClear(e);
  if (best) *best = win;
  if (best_ms) *best_ms = win_ms;
  if (cache_path) {
    FILE *f = fopen(cache_path, "a");
    if (!f) return xg_fail(XGPU_INVALID_CONFIG, "cannot write tuning cache %s", cache_path);
    char key[256];
    key_of(e, ntime, key, sizeof key);
    fprintf(f, "%s|%d|%d|%d|%d|%d|%.4f\n", key, win.wg_per_cu, win.prefetch, win.kernel,
            win.stage, win.staging, win_ms);
    fclose(f);
  }
  return XGPU_OK;
}
