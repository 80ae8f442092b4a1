/*
 * xgpu.h -- drop-in public C API of the MI355X X-engine.
 *
 * Every type, constant and prototype in this header has the same name, value,
 * field order and meaning as the xGPU library's public header
 * (reference src/xgpu.h:1-241), so that code written against xGPU compiles and
 * links against this library unchanged.  The library behind it runs on HIP
 * (gfx950) instead of CUDA; the entry point keeps its historical name
 * xgpuCudaXengine.
 *
 * Sizing is fixed when the library is built (make variables NSTATION,
 * NFREQUENCY, NTIME, NTIME_PIPE, DP4A), exactly as in xGPU; xgpuInfo() reports
 * it.  The runtime-sized, device-pointer API lives in xgpu_amd.h.
 *
 * Additions (additive only, no existing value changes): XGPU_FLOAT16 as an
 * input/compute type code.
 */
#include <stddef.h>

#ifndef XGPU_H
#define XGPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* Inputs are always 8-bit two's-complement integers (src/xgpu.h:13). */
#define FIXED_POINT

#ifndef FIXED_POINT
typedef float ReImInput;
#else
typedef signed char ReImInput;
#endif

/* One complex input sample: real byte then imaginary byte (src/xgpu.h:22-25). */
typedef struct ComplexInputStruct {
  ReImInput real;
  ReImInput imag;
} ComplexInput;

/* One visibility.  Float by default; int when the consumer defines DP4A,
 * which must match a library built with DP4A=yes (src/xgpu.h:27-37). */
#ifndef DP4A
typedef struct ComplexStruct {
  float real;
  float imag;
} Complex;
#else
typedef struct ComplexStruct {
  int real;
  int imag;
} Complex;
#endif

/* Input / compute type codes (src/xgpu.h:40-42). */
#define XGPU_INT8    (0)
#define XGPU_FLOAT32 (1)
#define XGPU_INT32   (2)
/* Additive: FP16 complex input / FP16 MFMA compute (no xGPU counterpart). */
#define XGPU_FLOAT16 (3)

/* Output matrix orders (src/xgpu.h:45-47). */
#define TRIANGULAR_ORDER 1000
#define REAL_IMAG_TRIANGULAR_ORDER 2000
#define REGISTER_TILE_TRIANGULAR_ORDER 3000

/* xgpuInit flags: low 16 bits select the device (src/xgpu.h:50-54). */
#define XGPU_DEVICE_MASK          ((1<<16)-1)
#define XGPU_DONT_REGISTER_ARRAY  (1<<16)
#define XGPU_DONT_REGISTER_MATRIX (1<<17)
#define XGPU_DONT_REGISTER        (XGPU_DONT_REGISTER_ARRAY | \
                                   XGPU_DONT_REGISTER_MATRIX)

/* Build-time sizing reported by xgpuInfo() (src/xgpu.h:61-92).  Input values
 * of the fields are ignored. */
typedef struct XGPUInfoStruct {
  unsigned int npol;          /* polarizations per station */
  unsigned int nstation;      /* stations */
  unsigned int nbaseline;     /* (nstation+1)*(nstation/2) */
  unsigned int nfrequency;    /* frequency channels */
  unsigned int ntime;         /* time samples per xgpuCudaXengine call */
  unsigned int ntimepipe;     /* time samples per host-to-device transfer */
  unsigned int input_type;    /* XGPU_INT8 (or XGPU_FLOAT16, additive) */
  unsigned int compute_type;  /* XGPU_INT8 or XGPU_FLOAT32 (or XGPU_FLOAT16) */
  long long unsigned int vecLength;     /* ComplexInput elements per call */
  long long unsigned int vecLengthPipe; /* ComplexInput elements per transfer */
  long long unsigned int matLength;     /* Complex elements in the output */
  long long unsigned int triLength;     /* Complex elements in triangular order */
  unsigned int matrix_order;  /* one of the *_ORDER codes */
  size_t shared_atomic_size;  /* kept for ABI compatibility */
  size_t complex_block_size;  /* kept for ABI compatibility */
} XGPUInfo;

/* Caller-allocated context (src/xgpu.h:94-113).  Field order is ABI. */
typedef struct XGPUContextStruct {
  ComplexInput *array_h;   /* host input buffer */
  Complex *matrix_h;       /* host output buffer */
  size_t array_len;        /* in ComplexInput elements */
  size_t matrix_len;       /* in Complex elements */
  /* Element offsets into the host buffers; reset to 0 by xgpuSetHost*Buffer. */
  size_t input_offset;
  size_t output_offset;
  void *internal;          /* library-private state */
} XGPUContext;

/* Return codes (src/xgpu.h:116-121).  XGPU_INSUFFICIENT_TEXTURE_MEMORY is
 * never returned: this library uses no textures. */
#define XGPU_OK                          (0)
#define XGPU_OUT_OF_MEMORY               (1)
#define XGPU_CUDA_ERROR                  (2)
#define XGPU_INSUFFICIENT_TEXTURE_MEMORY (3)
#define XGPU_NOT_INITIALIZED             (4)
#define XGPU_HOST_BUFFER_NOT_SET         (5)

/* syncOp values for xgpuCudaXengine (src/xgpu.h:124-127). */
#define SYNCOP_NONE           0
#define SYNCOP_DUMP           1
#define SYNCOP_SYNC_TRANSFER  2
#define SYNCOP_SYNC_COMPUTE   3

/* Static version string; do not modify or free (src/xgpu.h:134). */
const char * xgpuVersionString();

/* Fill *pcxs with the build-time sizing (src/xgpu.h:140). */
void xgpuInfo(XGPUInfo *pcxs);

/* Allocate device state and set the host buffers (src/xgpu.h:142-168).
 * A NULL array_h / matrix_h makes the library allocate pinned host memory and
 * set the *_len field.  A caller buffer is used in place; unless the matching
 * XGPU_DONT_REGISTER_* flag is given it is adopted now, otherwise the caller
 * must call xgpuSetHost*Buffer before the first xgpuCudaXengine. */
int xgpuInit(XGPUContext *context, int device_flags);

/* Zero the device integration buffer: start a new integration
 * (src/xgpu.h:170-174). */
int xgpuClearDeviceIntegrationBuffer(XGPUContext *context);

/* Replace the host input buffer with context->array_h (NULL: allocate)
 * (src/xgpu.h:176-188). */
int xgpuSetHostInputBuffer(XGPUContext *context);

/* Replace the host output buffer with context->matrix_h (NULL: allocate)
 * (src/xgpu.h:190-202). */
int xgpuSetHostOutputBuffer(XGPUContext *context);

/* Release everything the library allocated (src/xgpu.h:204). */
void xgpuFree(XGPUContext *context);

/* Correlate the ntime samples at array_h + input_offset into the device
 * integration buffer, then: SYNCOP_NONE return; SYNCOP_DUMP wait and copy the
 * integration buffer to matrix_h + output_offset; SYNCOP_SYNC_TRANSFER wait
 * for the input transfers; SYNCOP_SYNC_COMPUTE wait for the computation
 * (src/xgpu.h:206-219). */
int xgpuCudaXengine(XGPUContext *context, int syncOp);

/* CPU utilities (src/xgpu.h:223-231). */
void xgpuRandomComplex(ComplexInput* random_num, long long unsigned int length);
void xgpuReorderMatrix(Complex *matrix);
void xgpuCheckResult(Complex *gpu, Complex *cpu, int verbose, ComplexInput *array_h);
void xgpuSwizzleInput(ComplexInput *out, const ComplexInput *in);
void xgpuExtractMatrix(Complex *matrix, Complex *packed);

/* CPU X-engine, triangular order output (src/xgpu.h:235). */
void xgpuOmpXengine(Complex *matrix_h, ComplexInput *array_h);

#ifdef __cplusplus
}
#endif

#endif /* XGPU_H */
