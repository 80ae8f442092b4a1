/*
 * cuda_correlator.c -- command-line correctness and timing driver for the
 * xgpu.h library, with the options and report lines of the reference driver
 * (reference src/cuda_correlator.c:38-283):
 *   -c INTEG_CALLS  calls to xgpuCudaXengine per integration [1]
 *   -C INTEG_COUNT  integrations [1]
 *   -d DEVNUM       device [0]
 *   -f FINAL_SYNCOP sync op of the last call [1]
 *   -o SYNCOP       sync op of the other calls [2]
 *   -r              caller-allocated host buffers
 *   -s SEED         srand seed of the test data [1]
 *   -v LEVEL        verbosity of the check
 * With a single dumping call the result is reordered and checked against
 * xgpuOmpXengine.  Input is always in [t][f][s][p] order: the int8 path
 * reads it directly, so DP4A builds do not swizzle it.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>
#include <unistd.h>

#include "xgpu.h"

#define ELAPSED_MS(a, b) \
  ((((int64_t)(b).tv_sec - (a).tv_sec) * 1000000000 + ((b).tv_nsec - (a).tv_nsec)) / 1e6)

int main(int argc, char **argv) {
  int opt, device = 0, outer_count = 1, count = 1, verbose = 0, host_alloc = 0;
  int sync_op = SYNCOP_SYNC_TRANSFER, final_sync_op = SYNCOP_DUMP;
  unsigned int seed = 1;
  while ((opt = getopt(argc, argv, "C:c:d:f:ho:rs:v:")) != -1) {
    switch (opt) {
      case 'c':
        count = (int)strtoul(optarg, NULL, 0);
        if (count < 1) { fprintf(stderr, "count must be positive\n"); return 1; }
        break;
      case 'C':
        outer_count = (int)strtoul(optarg, NULL, 0);
        if (outer_count < 1) { fprintf(stderr, "outer count must be positive\n"); return 1; }
        break;
      case 'd': device = (int)strtoul(optarg, NULL, 0); break;
      case 'f': final_sync_op = (int)strtoul(optarg, NULL, 0); break;
      case 'o': sync_op = (int)strtoul(optarg, NULL, 0); break;
      case 'r': host_alloc = 1; break;
      case 's': seed = (unsigned)strtoul(optarg, NULL, 0); break;
      case 'v': verbose = (int)strtoul(optarg, NULL, 0); break;
      default:
        fprintf(stderr,
                "Usage: %s [options]\n"
                "Options:\n"
                "  -c INTEG_CALLS    Calls to xgpuCudaXengine per integration [1]\n"
                "  -C INTEG_COUNT    Number of integrations [1]\n"
                "  -d DEVNUM         GPU device to use [0]\n"
                "  -f FINAL_SYNCOP   Sync operation for final call [1]\n"
                "  -o SYNCOP         Sync operation for all but final call [1]\n"
                "                    Sync operation values are:\n"
                "                         0 (no sync)\n"
                "                         1 (sync and dump)\n"
                "                         2 (sync host to device transfer)\n"
                "                         3 (sync kernel computations)\n"
                "  -r                Register host allocated memory [false]\n"
                "                    (otherwise use library allocated memory)\n"
                "  -s SEED           Random number seed [1]\n"
                "  -v {0|1|2|3}      Verbosity level (debug only) [0]\n"
                "  -h                Show this message\n",
                argv[0]);
        exit(EXIT_FAILURE);
    }
  }
  srand(seed);
  XGPUInfo info;
  xgpuInfo(&info);
  printf("Correlating %u stations with %u channels and integration length %u\n", info.nstation,
         info.nfrequency, info.ntime);
  printf("Sending fixed point data to GPU.\n");

  XGPUContext context;
  if (host_alloc) {
    context.array_len = info.vecLength;
    context.matrix_len = info.matLength;
    context.array_h = malloc(context.array_len * sizeof(ComplexInput));
    context.matrix_h = malloc(context.matrix_len * sizeof(Complex));
  } else {
    context.array_h = NULL;
    context.matrix_h = NULL;
  }
  Complex *omp_matrix = NULL;
  int err = xgpuInit(&context, device);
  if (err) {
    fprintf(stderr, "xgpuInit returned error code %d\n", err);
    goto cleanup;
  }
  ComplexInput *array_h = context.array_h;
  Complex *gpu_matrix = context.matrix_h;
  xgpuRandomComplex(array_h, info.vecLength);

  omp_matrix = (Complex *)malloc(info.triLength * sizeof(Complex));
  if (!omp_matrix) {
    fprintf(stderr, "error allocating output buffer for xgpuOmpXengine\n");
    goto cleanup;
  }
  if (final_sync_op == SYNCOP_DUMP && count * outer_count == 1) {
    printf("Calling CPU X-Engine\n");
    xgpuOmpXengine(omp_matrix, array_h);
  }

  printf("Calling GPU X-Engine\n");
  struct timespec outer_start, start, stop;
  clock_gettime(CLOCK_MONOTONIC, &outer_start);
  for (int j = 0; j < outer_count; ++j) {
    clock_gettime(CLOCK_MONOTONIC, &start);
    for (int i = 0; i < count; ++i) {
      err = xgpuCudaXengine(&context, i == count - 1 ? final_sync_op : sync_op);
      if (err) {
        fprintf(stderr, "xgpuCudaXengine returned error code %d\n", err);
        goto cleanup;
      }
    }
    clock_gettime(CLOCK_MONOTONIC, &stop);
    const double total = ELAPSED_MS(start, stop), per_call = total / count;
    const double max_bw = info.ntime * info.nfrequency / per_call / 1000;
    const double gbps =
        ((float)(8 * context.array_len * sizeof(ComplexInput) * count)) / total / 1e6;
    printf("Elapsed time %.6f ms total, %.6f ms/call average\n", total, per_call);
    printf("Theoretical BW_max %.3f MHz, throughput %.3f Gbps\n", max_bw, gbps);
  }
  if (outer_count > 1) {
    clock_gettime(CLOCK_MONOTONIC, &stop);
    const double total = ELAPSED_MS(outer_start, stop);
    const double per_call = total / (count * outer_count);
    const double max_bw = info.ntime * info.nfrequency / per_call / 1000;
    const double gbps =
        ((float)(8 * context.array_len * sizeof(ComplexInput) * count * outer_count)) / total / 1e6;
    printf("Elapsed time %.6f ms total, %.6f ms/call average\n", total, per_call);
    printf("Theoretical BW_max %.3f MHz, throughput %.3f Gbps\n", max_bw, gbps);
  }
  if (final_sync_op == SYNCOP_DUMP && count * outer_count == 1) {
    xgpuReorderMatrix(gpu_matrix);
    xgpuCheckResult(gpu_matrix, omp_matrix, verbose, array_h);
  }

cleanup:
  free(omp_matrix);
  xgpuFree(&context);
  if (host_alloc) {
    free(context.array_h);
    free(context.matrix_h);
  }
  return err;
}
