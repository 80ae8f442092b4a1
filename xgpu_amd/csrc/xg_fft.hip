// xg_fft.hip -- batched single-precision complex FFTs for gfx950 (the F-engine
// step before the X-engine).
//
// Replaces the FFT prototype of the reference (src/fft/FFT{8..8192}.cu,
// src/fft/IFFT{8..8192}.cu, harness src/fft/main.cpp): in-place batched
// complex-to-complex transforms of N = 8 .. 8192 points, forward
// (exp(-2 pi i jk/N)) and unnormalized inverse.
//
// Design: one pass over HBM.  A workgroup (256 threads; 512 for N = 8192)
// loads a chunk of max(N, 2048) elements (2048/N whole transforms, or one
// transform) into LDS (16, 32 or 64 KiB), runs every radix pass there and
// stores the chunk once,
// so each element is read and written exactly once whatever N is (the
// reference needs two kernels, hence two HBM round trips, for N >= 2048).
// The passes are self-sorting Stockham radix-8/4/2 steps,
//   Y[j R L + s L + k] = sum_r w_R^{rs} w_{RL}^{rk} X[j L + k + r N/R],
// with L the product of the radices already done; the result is in natural
// order.  Twiddles come from a table of exp(-2 pi i m / 8192) computed in
// double precision on the device and rounded once to float.  The inverse is
// conj(FFT(conj(x))).
//
// Two layouts:
//   natural   transforms contiguous, natural order in and out (xgpuAmdFft);
//   prototype the reference entry points' layouts (FFT<N>/IFFT<N>), read off
//             the harness's transposes (src/fft/main.cpp:249-264):
//             N = 8, 16: groups of 64 transforms, element y of transform z
//               at group offset z + 64 y;
//             N = 64: groups of 8 transforms, element 8a + x of transform t
//               at group offset 64 a + 8 t + x;
//             N = 256 .. 1024: natural;
//             N >= 2048: the forward output (and so the inverse input) has
//               bin y (N/512) + z at offset 512 z + y; the forward input and
//               the inverse output are natural.
#include <hip/hip_runtime.h>
#include <mutex>
#include <stdint.h>
#include <stdio.h>

#include "xgpu_fft.h"

#define XG_FFT_TWN 8192  // twiddle table size: exp(-2 pi i m / 8192)

__device__ float2 xg_fft_tw[XG_FFT_TWN];

__global__ void xg_fft_twiddle_init_k() {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= XG_FFT_TWN) return;
  double s, c;
  sincospi(2.0 * m / XG_FFT_TWN, &s, &c);
  xg_fft_tw[m] = make_float2((float)c, (float)-s);
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by -i
__device__ __forceinline__ float2 cmi(float2 a) { return make_float2(a.y, -a.x); }

// In-register forward DFTs of size 2, 4 and 8: b_s = sum_r a_r w_R^{rs}.
__device__ __forceinline__ void dft2(float2 *a) {
  const float2 t = a[0];
  a[0] = cadd(t, a[1]);
  a[1] = csub(t, a[1]);
}
__device__ __forceinline__ void dft4(float2 *a) {
  const float2 t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
  const float2 t2 = cadd(a[1], a[3]), t3 = cmi(csub(a[1], a[3]));
  a[0] = cadd(t0, t2);
  a[2] = csub(t0, t2);
  a[1] = cadd(t1, t3);
  a[3] = csub(t1, t3);
}
__device__ __forceinline__ void dft8(float2 *a) {
  float2 e[4] = {a[0], a[2], a[4], a[6]}, o[4] = {a[1], a[3], a[5], a[7]};
  dft4(e);
  dft4(o);
  const float c = 0.70710678118654752f;  // w_8 = (c, -c)
  const float2 w1o = make_float2(c * (o[1].x + o[1].y), c * (o[1].y - o[1].x));
  const float2 w2o = cmi(o[2]);
  const float2 w3o = make_float2(c * (o[3].y - o[3].x), -c * (o[3].x + o[3].y));
  a[0] = cadd(e[0], o[0]);
  a[4] = csub(e[0], o[0]);
  a[1] = cadd(e[1], w1o);
  a[5] = csub(e[1], w1o);
  a[2] = cadd(e[2], w2o);
  a[6] = csub(e[2], w2o);
  a[3] = cadd(e[3], w3o);
  a[7] = csub(e[3], w3o);
}
template <int R>
__device__ __forceinline__ void dft(float2 *a) {
  if constexpr (R == 2) dft2(a);
  else if constexpr (R == 4) dft4(a);
  else dft8(a);
}

// Radix of pass p for size N (0 past the last pass).  The product of a
// plan's radices is N, and every pass has at least one butterfly per thread
// of the workgroup.
__host__ __device__ constexpr int fft_radix(int n, int p) {
  switch (n) {
    case 8: return p == 0 ? 8 : 0;
    case 16: return p == 0 ? 8 : p == 1 ? 2 : 0;
    case 32: return p == 0 ? 8 : p == 1 ? 4 : 0;
    case 64: return p < 2 ? 8 : 0;
    case 128: return p == 0 ? 8 : p < 3 ? 4 : 0;
    case 256: return p < 2 ? 8 : p == 2 ? 4 : 0;
    case 512: return p < 3 ? 8 : 0;
    case 1024: return p < 2 ? 8 : p < 4 ? 4 : 0;
    case 2048: return p < 3 ? 8 : p == 3 ? 4 : 0;
    case 4096: return p < 4 ? 8 : 0;
    case 8192: return p < 4 ? 8 : p == 4 ? 2 : 0;
    default: return 0;
  }
}

template <int N>
struct FftCfg {
  static constexpr int kT = N >= 2048 ? 1 : 2048 / N;     // transforms per workgroup
  static constexpr int kC = kT * N;                        // elements per workgroup
  static constexpr int kThreads = N == 8192 ? 512 : 256;  // more waves where LDS limits
  static constexpr int kE = kC / kThreads;                 // elements per thread
};

// Chunk offset, in layout LAYOUT, of the workgroup's natural-order element q
// (transform q / N, element or bin q % N) on the input side (IN = true) or
// the output side of a transform.
template <int N, bool INV, int LAYOUT, bool IN>
__device__ __forceinline__ int fft_gpos(int q) {
  if (LAYOUT == XGPU_FFT_NATURAL) return q;
  if (N == 8 || N == 16) {  // groups of 64 transforms: element y of z at z + 64 y
    const int S = 64 * N, grp = q / S, rem = q % S;
    return grp * S + rem / N + 64 * (rem % N);
  }
  if (N == 64) {  // groups of 8 transforms: element 8a + x of t at 64 a + 8 t + x
    const int grp = q / 512, rem = q % 512, e = rem % 64;
    return grp * 512 + 64 * (e / 8) + 8 * (rem / 64) + e % 8;
  }
  if (N >= 2048 && (IN == INV)) {  // spectrum side: bin y (N/512) + z at 512 z + y
    return 512 * (q % (N / 512)) + q / (N / 512);
  }
  return q;
}

// LDS slot of natural-order element q.  The prototype layout of N = 8 and 16
// interleaves 64 transforms (element y of transform z at z + 64 y), so the
// staging loops, coalesced on the HBM side, touch LDS with a stride of N
// elements between lanes: N = 16 hits 2 of the 64 banks.  An XOR of the
// element index with bits of the transform index spreads them: a block of N
// float2 spans 2N of the 64 banks, so 32 / N consecutive transforms already
// start on different banks, and the XOR takes the transform index bits above
// those.  The slot stays in its N-element block, so every index stays below kC.
// The radix passes of these small sizes gain from it in natural order too.
template <int N, int LAYOUT>
__device__ __forceinline__ int fft_lds(int q) {
  if constexpr (N == 8 || N == 16) return q ^ (((q / N) >> (N == 8 ? 2 : 1)) & (N - 1));
  // N = 8192 (512 threads, radix 8 passes): an XOR within 16-element blocks
  // measured faster; for the other sizes it measured slower
  if constexpr (N == 8192) return q ^ ((q >> 4) & 15);
  return q;
}

// Inverse of fft_gpos: natural-order element held at chunk offset i.
template <int N, bool INV, int LAYOUT, bool IN>
__device__ __forceinline__ int fft_natural(int i) {
  if (LAYOUT == XGPU_FFT_NATURAL) return i;
  if (N == 8 || N == 16) {
    const int S = 64 * N, grp = i / S, w = i % S;
    return (grp * 64 + w % 64) * N + w / 64;
  }
  if (N == 64) {
    const int grp = i / 512, w = i % 512;
    return (grp * 8 + (w / 8) % 8) * 64 + 8 * (w / 64) + w % 8;
  }
  if (N >= 2048 && (IN == INV)) return (i % 512) * (N / 512) + i / 512;
  return i;
}

// One Stockham pass of radix R at span L over the workgroup's transforms:
// every thread reads its butterflies' inputs, the workgroup syncs, then every
// thread writes its outputs (in place in LDS).  Where a thread's inputs are at
// least 8 elements apart (N/R >= 8), the first pass reads straight from HBM
// (FIRST), and where its outputs are (L >= 8), the last pass writes straight
// to HBM (LAST): consecutive butterflies then touch runs of consecutive
// elements.  Other sizes stage through LDS with coalesced loops.
template <int N, int R, int L, bool FIRST, bool LAST, bool INV, int LAYOUT>
__device__ __forceinline__ void fft_pass(float2 *X, float2 *work, long long base, long long nelem,
                                         int tid) {
  typedef FftCfg<N> C;
  constexpr int kB = C::kC / R / C::kThreads;  // butterflies per thread
  static_assert(kB >= 1 && kB * C::kThreads * R == C::kC, "pass shape");
  float2 v[kB][R];
  int dst[kB];
#pragma unroll
  for (int p = 0; p < kB; ++p) {
    const int b = tid + C::kThreads * p;
    const int t = b / (N / R), rem = b % (N / R);
    const int k = rem % L, j = rem / L;
    const int q0 = t * N + j * L + k;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int q = q0 + r * (N / R);
      if constexpr (FIRST) {
        const long long g = base + fft_gpos<N, INV, LAYOUT, true>(q);
        float2 x = g < nelem ? work[g] : make_float2(0.f, 0.f);
        if (INV) x.y = -x.y;  // inverse = conj(FFT(conj(x)))
        v[p][r] = x;
      } else {
        v[p][r] = X[fft_lds<N, LAYOUT>(q)];
      }
    }
    if (L > 1) {
      // w^{rk}, w = exp(-2 pi i / RL): w^k, w^2k, w^4k from the table (m < 8192 / R,
      // so 4m < 8192), the other powers as products of those (at most two roundings)
      const int m = k * (XG_FFT_TWN / (R * L));
      float2 w[R];
      w[1] = xg_fft_tw[m];
      if constexpr (R >= 4) {
        w[2] = xg_fft_tw[2 * m];
        w[3] = cmul(w[1], w[2]);
      }
      if constexpr (R == 8) {
        w[4] = xg_fft_tw[4 * m];
        w[5] = cmul(w[1], w[4]);
        w[6] = cmul(w[2], w[4]);
        w[7] = cmul(w[3], w[4]);
      }
#pragma unroll
      for (int r = 1; r < R; ++r) v[p][r] = cmul(v[p][r], w[r]);
    }
    dst[p] = t * N + j * R * L + k;
  }
  if constexpr (!FIRST && !LAST) __syncthreads();  // every read of X before any write
#pragma unroll
  for (int p = 0; p < kB; ++p) {
    dft<R>(v[p]);
#pragma unroll
    for (int s = 0; s < R; ++s) {
      const int q = dst[p] + s * L;
      if constexpr (LAST) {
        const long long g = base + fft_gpos<N, INV, LAYOUT, false>(q);
        if (g < nelem) {
          float2 y = v[p][s];
          if (INV) y.y = -y.y;
          work[g] = y;
        }
      } else {
        X[fft_lds<N, LAYOUT>(q)] = v[p][s];
      }
    }
  }
  if constexpr (!LAST) __syncthreads();  // this pass's writes before the next pass's reads
}

// passes of the plan from pass P (span L) on; FUSE_IN / FUSE_OUT: the first
// pass may read HBM and the last may write it directly
template <int N, int P, int L, bool INV, int LAYOUT, bool FUSE_IN, bool FUSE_OUT>
__device__ __forceinline__ void fft_passes(float2 *X, float2 *work, long long base, long long nelem,
                                           int tid) {
  constexpr int R = fft_radix(N, P);
  if constexpr (R > 0) {
    constexpr bool FIRST = FUSE_IN && P == 0 && N / R >= 8;
    constexpr bool LAST = FUSE_OUT && fft_radix(N, P + 1) == 0 && L >= 8;
    fft_pass<N, R, L, FIRST, LAST, INV, LAYOUT>(X, work, base, nelem, tid);
    fft_passes<N, P + 1, L * R, INV, LAYOUT, FUSE_IN, FUSE_OUT>(X, work, base, nelem, tid);
  }
}

template <int N, bool INV, int LAYOUT>
__global__ __launch_bounds__(FftCfg<N>::kThreads) void xg_fft_k(float2 *work, long long nelem) {
  typedef FftCfg<N> C;
  constexpr int R0 = fft_radix(N, 0);
  constexpr int kPasses = fft_radix(N, 1) == 0 ? 1 : fft_radix(N, 2) == 0 ? 2
                        : fft_radix(N, 3) == 0 ? 3 : fft_radix(N, 4) == 0 ? 4 : 5;
  constexpr int RL = fft_radix(N, kPasses - 1);          // radix of the last pass
  // The spectrum side of the prototype layout at N >= 2048 is transposed
  // (bin y N/512 + z at 512 z + y): it goes through LDS too, so HBM sees
  // whole consecutive runs instead of four lines 4 KiB apart per wave access.
  // N = 8 and 16 stage both sides (their last pass would write 64-byte runs).
  constexpr bool kProtoSpec = LAYOUT == XGPU_FFT_PROTOTYPE && N >= 2048;
  constexpr bool kSmall = N <= 16;
  constexpr bool kStageIn = N / R0 < 8 || (kProtoSpec && INV) || kSmall;    // first pass reads LDS
  constexpr bool kStageOut = N / RL < 8 || (kProtoSpec && !INV) || kSmall;  // last pass writes LDS
  __shared__ float2 X[C::kC];
  const int tid = threadIdx.x;
  const long long base = (long long)blockIdx.x * C::kC;
  if constexpr (kStageIn) {  // coalesced HBM -> LDS in natural order
#pragma unroll
    for (int e = 0; e < C::kE; ++e) {
      const int i = tid + C::kThreads * e;
      float2 v = base + i < nelem ? work[base + i] : make_float2(0.f, 0.f);
      if (INV) v.y = -v.y;
      X[fft_lds<N, LAYOUT>(fft_natural<N, INV, LAYOUT, true>(i))] = v;
    }
    __syncthreads();
  }
  fft_passes<N, 0, 1, INV, LAYOUT, !kStageIn, !kStageOut>(X, work, base, nelem, tid);
  if constexpr (kStageOut) {  // coalesced LDS -> HBM
#pragma unroll
    for (int e = 0; e < C::kE; ++e) {
      const int i = tid + C::kThreads * e;
      if (base + i < nelem) {
        float2 v = X[fft_lds<N, LAYOUT>(fft_natural<N, INV, LAYOUT, false>(i))];
        if (INV) v.y = -v.y;
        work[base + i] = v;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static int fft_fail(hipError_t e, const char *what) {
  fprintf(stderr, "xgpu fft: %s: %s\n", what, hipGetErrorString(e));
  return XGPU_FFT_ERROR;
}

// The twiddle table is filled once per device, on the caller's stream.
static int twiddles_ready(hipStream_t s) {
  static std::mutex mu;
  static bool done[256];
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fft_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= 256) return fft_fail(hipErrorInvalidDevice, "device index");
  std::lock_guard<std::mutex> lock(mu);
  if (done[dev]) return XGPU_FFT_OK;
  hipLaunchKernelGGL(xg_fft_twiddle_init_k, dim3(XG_FFT_TWN / 256), dim3(256), 0, s);
  if ((e = hipGetLastError()) != hipSuccess) return fft_fail(e, "twiddle table");
  done[dev] = true;
  return XGPU_FFT_OK;
}

template <int N, bool INV, int LAYOUT>
static int fft_launch(void *work, long long batch, hipStream_t s) {
  const long long nelem = batch * N;
  const long long grid = (nelem + FftCfg<N>::kC - 1) / FftCfg<N>::kC;
  if (grid == 0) return XGPU_FFT_OK;
  if (grid > 0x7fffffffLL) return fft_fail(hipErrorInvalidValue, "batch too large");
  hipLaunchKernelGGL((xg_fft_k<N, INV, LAYOUT>), dim3((unsigned)grid), dim3(FftCfg<N>::kThreads),
                     0, s, (float2 *)work, nelem);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? XGPU_FFT_OK : fft_fail(e, "fft launch");
}

template <bool INV, int LAYOUT>
static int fft_dispatch(void *work, int n, long long batch, hipStream_t s) {
  switch (n) {
    case 8: return fft_launch<8, INV, LAYOUT>(work, batch, s);
    case 16: return fft_launch<16, INV, LAYOUT>(work, batch, s);
    case 32: return fft_launch<32, INV, LAYOUT>(work, batch, s);
    case 64: return fft_launch<64, INV, LAYOUT>(work, batch, s);
    case 128: return fft_launch<128, INV, LAYOUT>(work, batch, s);
    case 256: return fft_launch<256, INV, LAYOUT>(work, batch, s);
    case 512: return fft_launch<512, INV, LAYOUT>(work, batch, s);
    case 1024: return fft_launch<1024, INV, LAYOUT>(work, batch, s);
    case 2048: return fft_launch<2048, INV, LAYOUT>(work, batch, s);
    case 4096: return fft_launch<4096, INV, LAYOUT>(work, batch, s);
    case 8192: return fft_launch<8192, INV, LAYOUT>(work, batch, s);
    default: return XGPU_FFT_BAD_SIZE;
  }
}

// transforms per group of the prototype layouts
static long long proto_group(int n) { return (n == 8 || n == 16) ? 64 : n == 64 ? 8 : 1; }

extern "C" int xgpuAmdFft(void *work, int n, long long batch, int inverse, int layout,
                          void *stream) {
  if (fft_radix(n, 0) == 0) return XGPU_FFT_BAD_SIZE;
  if (layout != XGPU_FFT_NATURAL && layout != XGPU_FFT_PROTOTYPE) return XGPU_FFT_BAD_ARGUMENT;
  if (layout == XGPU_FFT_PROTOTYPE && (n == 32 || n == 128)) return XGPU_FFT_BAD_SIZE;
  if (batch < 0 || (batch > 0 && !work)) return XGPU_FFT_BAD_ARGUMENT;
  if (((uintptr_t)work) & 7) return XGPU_FFT_BAD_ARGUMENT;
  if (layout == XGPU_FFT_PROTOTYPE && batch % proto_group(n)) return XGPU_FFT_BAD_BATCH;
  if (batch == 0) return XGPU_FFT_OK;
  hipStream_t s = (hipStream_t)stream;
  const int rc = twiddles_ready(s);
  if (rc) return rc;
  if (layout == XGPU_FFT_NATURAL)
    return inverse ? fft_dispatch<true, XGPU_FFT_NATURAL>(work, n, batch, s)
                   : fft_dispatch<false, XGPU_FFT_NATURAL>(work, n, batch, s);
  return inverse ? fft_dispatch<true, XGPU_FFT_PROTOTYPE>(work, n, batch, s)
                 : fft_dispatch<false, XGPU_FFT_PROTOTYPE>(work, n, batch, s);
}

extern "C" int xgpuAmdFftTime(void *work, int n, long long batch, int inverse, int layout,
                              int reps, float *ms) {
  if (!ms || reps < 1) return XGPU_FFT_BAD_ARGUMENT;
  int rc = xgpuAmdFft(work, n, batch, inverse, layout, 0);  // warm-up, and argument checks
  if (rc) return rc;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return fft_fail(hipErrorOutOfMemory, "event");
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fft_fail(hipErrorOutOfMemory, "event");
  }
  hipError_t e = hipEventRecord(e0, 0);
  for (int i = 0; i < reps && e == hipSuccess && rc == XGPU_FFT_OK; ++i)
    rc = xgpuAmdFft(work, n, batch, inverse, layout, 0);
  if (e == hipSuccess) e = hipEventRecord(e1, 0);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float total = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&total, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  if (e != hipSuccess) return fft_fail(e, "timing");
  *ms = total / (float)reps;
  return XGPU_FFT_OK;
}

// Reference entry points (src/fft/main.cpp:37-55): in place on the default
// stream, prototype layouts; a rejected call prints the reason and does
// nothing.
#define XG_FFT_ENTRY(NAME, N, INV)                                                  \
  extern "C" void NAME(void *work, int batch) {                                     \
    const int rc = xgpuAmdFft(work, N, batch, INV, XGPU_FFT_PROTOTYPE, 0);            \
    if (rc) fprintf(stderr, #NAME ": rejected (%d), batch %d\n", rc, batch);        \
  }
XG_FFT_ENTRY(FFT8, 8, 0)
XG_FFT_ENTRY(FFT16, 16, 0)
XG_FFT_ENTRY(FFT64, 64, 0)
XG_FFT_ENTRY(FFT256, 256, 0)
XG_FFT_ENTRY(FFT512, 512, 0)
XG_FFT_ENTRY(FFT1024, 1024, 0)
XG_FFT_ENTRY(FFT2048, 2048, 0)
XG_FFT_ENTRY(FFT4096, 4096, 0)
XG_FFT_ENTRY(FFT8192, 8192, 0)
XG_FFT_ENTRY(IFFT8, 8, 1)
XG_FFT_ENTRY(IFFT16, 16, 1)
XG_FFT_ENTRY(IFFT64, 64, 1)
XG_FFT_ENTRY(IFFT256, 256, 1)
XG_FFT_ENTRY(IFFT512, 512, 1)
XG_FFT_ENTRY(IFFT1024, 1024, 1)
XG_FFT_ENTRY(IFFT2048, 2048, 1)
XG_FFT_ENTRY(IFFT4096, 4096, 1)
XG_FFT_ENTRY(IFFT8192, 8192, 1)
