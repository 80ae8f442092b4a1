/*
 * cuda_correlator.c -- command-line driver of the xgpu.h library: runs
 * integrations of synthetic data through xgpuCudaXengine, times them, and
 * checks a single dumped integration against the CPU X-engine.
 *
 * Interface kept from the reference driver (reference
 * src/cuda_correlator.c:38-283), because run scripts pass these options and
 * read these lines: the option letters -c -C -d -f -o -r -s -v -h, and the
 * report lines "Correlating ...", "Sending fixed point data to GPU.",
 * "Calling CPU X-Engine", "Calling GPU X-Engine", "Elapsed time ...",
 * "Theoretical BW_max ..." and xgpuCheckResult's verdict.  DP4A builds take
 * the corner-turned input of xgpuSwizzleInput, as the reference's do.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "xgpu.h"

typedef struct {
  int device;
  int calls;       /* xgpuCudaXengine calls per integration (-c) */
  int integrations; /* (-C) */
  int sync_op;     /* sync op of every call but an integration's last (-o) */
  int final_op;    /* sync op of an integration's last call (-f) */
  int caller_buffers; /* -r: host buffers allocated here, not by the library */
  unsigned seed;
  int verbose;
} Options;

static void usage(FILE *to, const char *prog) {
  fprintf(to,
          "usage: %s [-c calls] [-C integrations] [-d device] [-f syncop] [-o syncop]\n"
          "          [-r] [-s seed] [-v level] [-h]\n"
          "\n"
          "Correlates synthetic data (xgpuRandomComplex) with the X-engine sized when the\n"
          "library was built, reports the call rate, and -- when exactly one call dumps\n"
          "one integration -- compares the result with xgpuOmpXengine.\n"
          "\n"
          "  -c calls         xgpuCudaXengine calls that make up one integration (1)\n"
          "  -C integrations  integrations to run (1)\n"
          "  -d device        device number (0)\n"
          "  -f syncop        sync op of the last call of each integration (1)\n"
          "  -o syncop        sync op of the other calls (2)\n"
          "                   syncop: 0 none, 1 wait and dump, 2 wait for the input\n"
          "                   transfer, 3 wait for the kernels\n"
          "  -r               use host buffers allocated by this program (default: let\n"
          "                   the library allocate them)\n"
          "  -s seed          srand seed of the synthetic data (1)\n"
          "  -v level         detail of the comparison report, 0..3 (0)\n"
          "  -h               this text\n",
          prog);
}

/* parses argv into *o; returns 0, or the exit status the program should end with */
static int parse_options(int argc, char **argv, Options *o) {
  o->device = 0;
  o->calls = o->integrations = 1;
  o->sync_op = SYNCOP_SYNC_TRANSFER;
  o->final_op = SYNCOP_DUMP;
  o->caller_buffers = 0;
  o->seed = 1;
  o->verbose = 0;
  for (int c; (c = getopt(argc, argv, "C:c:d:f:ho:rs:v:")) != -1;) {
    const long v = optarg ? strtol(optarg, NULL, 0) : 0;
    if (c == 'c' || c == 'C') {
      if (v < 1) {
        fprintf(stderr, "%s must be positive\n", c == 'c' ? "count" : "outer count");
        return 1;
      }
      *(c == 'c' ? &o->calls : &o->integrations) = (int)v;
    } else if (c == 'd') {
      o->device = (int)v;
    } else if (c == 'f') {
      o->final_op = (int)v;
    } else if (c == 'o') {
      o->sync_op = (int)v;
    } else if (c == 'r') {
      o->caller_buffers = 1;
    } else if (c == 's') {
      o->seed = (unsigned)v;
    } else if (c == 'v') {
      o->verbose = (int)v;
    } else {
      usage(c == 'h' ? stdout : stderr, argv[0]);
      exit(c == 'h' ? EXIT_SUCCESS : EXIT_FAILURE);
    }
  }
  return 0;
}

static double now_ms(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return 1e3 * (double)ts.tv_sec + 1e-6 * (double)ts.tv_nsec;
}

/* The two rate lines for `calls` calls that took `ms`: the call time, the
 * widest band the engine keeps up with (NTIME x NFREQUENCY channel samples
 * per call, in MHz) and the input rate in Gbit/s. */
static void report_rate(const XGPUInfo *info, double ms, long calls) {
  const double per_call = ms / (double)calls;
  const double band_mhz = (double)info->ntime * info->nfrequency / per_call / 1e3;
  const double gbit = 8.0 * (double)info->vecLength * sizeof(ComplexInput) * (double)calls / 1e9;
  printf("Elapsed time %.6f ms total, %.6f ms/call average\n", ms, per_call);
  printf("Theoretical BW_max %.3f MHz, throughput %.3f Gbps\n", band_mhz, gbit / (ms * 1e-3));
}

/* One integration: o->calls calls, the last with o->final_op.  Returns the
 * first xgpu error, or 0. */
static int integrate(XGPUContext *ctx, const Options *o) {
  for (int i = 0; i < o->calls; ++i) {
    const int rc = xgpuCudaXengine(ctx, i + 1 == o->calls ? o->final_op : o->sync_op);
    if (rc) {
      fprintf(stderr, "xgpuCudaXengine returned error code %d\n", rc);
      return rc;
    }
  }
  return 0;
}

int main(int argc, char **argv) {
  Options o;
  int rc = parse_options(argc, argv, &o);
  if (rc) return rc;
  srand(o.seed);

  XGPUInfo info;
  xgpuInfo(&info);
  printf("Correlating %u stations with %u channels and integration length %u\n", info.nstation,
         info.nfrequency, info.ntime);
  printf("Sending fixed point data to GPU.\n");

  XGPUContext ctx;
  memset(&ctx, 0, sizeof ctx);
  if (o.caller_buffers) {
    ctx.array_len = info.vecLength;
    ctx.matrix_len = info.matLength;
    ctx.array_h = (ComplexInput *)malloc(ctx.array_len * sizeof(ComplexInput));
    ctx.matrix_h = (Complex *)malloc(ctx.matrix_len * sizeof(Complex));
  }
  ComplexInput *caller_array = ctx.array_h;
  Complex *caller_matrix = ctx.matrix_h;
  ComplexInput *data = NULL; /* the samples in plain order */
  Complex *cpu = NULL;
  const int check = o.final_op == SYNCOP_DUMP && o.calls == 1 && o.integrations == 1;

  if ((rc = xgpuInit(&ctx, o.device)) != 0) {
    fprintf(stderr, "xgpuInit returned error code %d\n", rc);
    goto done;
  }
#ifdef DP4A
  /* the library reads the corner-turned layout; keep the plain samples for
   * the CPU engine and the check */
  data = (ComplexInput *)malloc(ctx.array_len * sizeof(ComplexInput));
  if (!data) {
    fprintf(stderr, "out of host memory\n");
    rc = XGPU_OUT_OF_MEMORY;
    goto done;
  }
  xgpuRandomComplex(data, info.vecLength);
  xgpuSwizzleInput(ctx.array_h, data);
#else
  data = ctx.array_h;
  xgpuRandomComplex(data, info.vecLength);
#endif

  /* the CPU engine writes TRIANGULAR_ORDER */
  cpu = (Complex *)malloc(info.triLength * sizeof(Complex));
  if (!cpu) {
    fprintf(stderr, "error allocating output buffer for xgpuOmpXengine\n");
    rc = XGPU_OUT_OF_MEMORY;
    goto done;
  }
  if (check) {
    printf("Calling CPU X-Engine\n");
    xgpuOmpXengine(cpu, data);
  }

  printf("Calling GPU X-Engine\n");
  const double t_all = now_ms();
  for (int j = 0; j < o.integrations; ++j) {
    const double t0 = now_ms();
    if ((rc = integrate(&ctx, &o)) != 0) goto done;
    report_rate(&info, now_ms() - t0, o.calls);
  }
  if (o.integrations > 1) report_rate(&info, now_ms() - t_all, (long)o.calls * o.integrations);

  if (check) {
    xgpuReorderMatrix(ctx.matrix_h);
    xgpuCheckResult(ctx.matrix_h, cpu, o.verbose, data);
  }

done:
  free(cpu);
#ifdef DP4A
  free(data);
#endif
  xgpuFree(&ctx);
  free(caller_array);
  free(caller_matrix);
  return rc;
}
