/*
 * xgpuinfo.c -- print the build-time sizing of the linked xgpu library.
 *
 * The line texts are an interface: autoconf's AX_CHECK_XGPUINFO extracts
 * "Number of stations: ", "Number of frequencies: " and
 * "time samples per GPU integration: " with sed (reference src/xgpu.m4:57-67),
 * so every line matches the reference tool (reference src/xgpuinfo.c:5-67).
 */
#include <stdio.h>

#include "xgpu.h"

int main(void) {
  XGPUInfo info;
  xgpuInfo(&info);
  printf("xGPU library version: %s\n", xgpuVersionString());
  printf("Number of polarizations: %u\n", info.npol);
  printf("Number of stations: %u\n", info.nstation);
  printf("Number of baselines: %u\n", info.nbaseline);
  printf("Number of frequencies: %u\n", info.nfrequency);
  printf("Number of time samples per GPU integration: %u\n", info.ntime);
  printf("Number of time samples per transfer to GPU: %u\n", info.ntimepipe);
  printf("Type of ComplexInput components: ");
  switch (info.input_type) {
    case XGPU_INT8: printf("8 bit integers\n"); break;
    case XGPU_FLOAT32: printf("32 bit floats\n"); break;
    case XGPU_INT32: printf("32 bit integers\n"); break;
    case XGPU_FLOAT16: printf("16 bit floats\n"); break;
    default: printf("<unknown type code: %d>\n", info.input_type);
  }
  printf("Type of computation: ");
  switch (info.compute_type) {
    case XGPU_INT8: printf("8 bit multiply, 32-bit accumulate\n"); break;
    case XGPU_FLOAT32: printf("FP32 multiply, FP32 accumulate\n"); break;
    case XGPU_FLOAT16: printf("FP16 multiply, FP32 accumulate\n"); break;
    default: printf("<unknown type code: %d>\n", info.compute_type);
  }
  printf("Number of ComplexInput elements in GPU input vector: %llu\n", info.vecLength);
  printf("Number of ComplexInput elements per transfer to GPU: %llu\n", info.vecLengthPipe);
  printf("Number of Complex elements in GPU output vector: %llu\n", info.matLength);
  printf("Number of Complex elements in reordered output vector: %llu\n", info.triLength);
  printf("Output matrix order: ");
  switch (info.matrix_order) {
    case TRIANGULAR_ORDER: printf("triangular\n"); break;
    case REAL_IMAG_TRIANGULAR_ORDER: printf("real imaginary triangular\n"); break;
    case REGISTER_TILE_TRIANGULAR_ORDER: printf("register tile triangular\n"); break;
    default: printf("<unknown order code: %d>\n", info.matrix_order);
  }
  printf("Shared atomic transfer size: %lu\n", (unsigned long)info.shared_atomic_size);
  printf("Complex block size: %lu\n", (unsigned long)info.complex_block_size);
  return 0;
}
