// xg_kernels.hip -- gfx950 X-engine kernels.
//
// Replaces the CUDA kernel shared2x2 and its LDS loaders
// (reference src/kernel.cuh:238-591, src/shared_transfer_4.cuh,
// src/shared_transfer_4_dp4a.cuh, src/shared_transfer_8.cuh).  The
// arithmetic is the same per-channel Hermitian rank-K update
//   V[f][i][j][p1][p2] = sum_t x[t][f][i][p1] * conj(x[t][f][j][p2]),  i >= j
// (reference src/omp_xengine.c:24-28,55-84), re = ar*br + ai*bi,
// im = ai*br - ar*bi, but the implementation is MI355X-native:
//
//  * One persistent launch covers a list of jobs (subarray blocks) and, for
//    each, every channel's lower-triangle of 128x128-input tiles.  Workgroups
//    stay resident and walk their XCD's share of the tile list.
//  * Accumulators stay in registers for the whole block of ntime samples
//    (the reference re-reads and re-writes global memory every NTIME_PIPE).
//  * Input is time-major ([t][f][s][p][c]); the corner turn to the
//    time-contiguous MFMA operand layout is done in registers with v_perm
//    while staging global -> LDS, never as a separate HBM pass.
//  * LDS rows are 128 bytes (one stage of one input), with a 16-byte-chunk
//    XOR swizzle that makes both the staging writes and the operand reads
//    bank-conflict free.
//
// Integer path (int8 MFMA, int32 accumulation, exact):
//   Re:  acc_re += A  . B   with A = (ar,ai) pairs of the row input,
//                                 B = (br,bi) pairs of the column input
//   Im:  acc_im += A' . B   with A' = (ai, ~ar)   (planes reordered)
//   ~ar = -ar - 1 for every int8 including -128, so
//   A'.B = sum(ai*br - ar*bi) - sum(bi), and the exact conjugate term is
//   acc_im + sum_t bi(column).  No int8 value is ever negated, so -128 is
//   exact.  The column sums come from v_dot4 on the B operand.
//   |per-sample term| <= 2*128*128, so int32 cannot overflow for blocks of
//   up to 65535 samples; the host splits longer blocks.
//
// FP16 path: A' = (ai, -ar) by a sign flip (exact in floating point),
// FP16 MFMA with FP32 accumulation.
//
// FP32 plain path: SIMT FP32 FMA, exact int8 -> float conversion, 8x8
// complex register tile per thread.
#include <hip/hip_runtime.h>
#include <pthread.h>
#include <stdint.h>

#include "xg_hip.h"
#include "xg_launch.h"

int xgh_report(hipError_t e, const char *what);

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

#define XG_ROWB 128                         // LDS bytes per input per stage
#define XG_SIDE_BYTES (XG_TILE * XG_ROWB)   // 16 KiB: one side of a tile
#define XG_STAGE_BYTES (2 * XG_SIDE_BYTES)  // rows + columns
#define XG_LDS_BYTES (2 * XG_STAGE_BYTES)   // double buffered: 64 KiB

// ---------------------------------------------------------------------------
// Tile schedule
// ---------------------------------------------------------------------------
struct TileRef {
  int job, f, rb, cb;
};

// Order of the tiles of one channel: the lower triangle of nblk x nblk blocks
// is cut into super-blocks of XG_SB x XG_SB blocks, taken column by column
// (lower triangle only), and the blocks inside a super-block column by column.
// Tiles that run at the same time share input chunks (a tile (rb, cb) reads
// chunks rb and cb), and the last tiles of the order are packed into few
// chunks: at nblk = 8 the last 4 tiles touch 3 chunks and the last 8 touch 4.
#define XG_SB 8
__host__ __device__ inline void xg_tile_order(int nblk, int u, int *rb, int *cb) {
  const int nsb = (nblk + XG_SB - 1) / XG_SB;
  for (int C = 0; C < nsb; ++C) {
    const int c0 = C * XG_SB, c1 = min(c0 + XG_SB, nblk);
    for (int R = C; R < nsb; ++R) {
      const int r0 = R * XG_SB, r1 = min(r0 + XG_SB, nblk);
      const int m = r1 - r0;
      const int count = (R == C) ? m * (m + 1) / 2 : m * (c1 - c0);
      if (u >= count) {
        u -= count;
        continue;
      }
      for (int c = c0; c < c1; ++c) {
        const int top = (R == C) ? c : r0;  // first row of column c in this super-block
        const int n = r1 - top;
        if (u < n) {
          *rb = top + u;
          *cb = c;
          return;
        }
        u -= n;
      }
    }
  }
  *rb = *cb = -1;  // u past the triangle (not reached for u < nblk (nblk+1) / 2)
}

// Pair tiles (rb, cp): XG_TILE row inputs [XG_TILE rb, +XG_TILE) by 2 XG_TILE
// column inputs [2 XG_TILE cp, +2 XG_TILE), for 2 cp <= rb: every input pair
// (m, n) with n <= m + 1 lies in exactly one of them (rb = m / 128,
// cp = n / 256).  Order: column by column, rows from the diagonal down.
__host__ __device__ inline void xg_pair_order(int nblk, int u, int *rb, int *cp) {
  for (int c = 0; 2 * c < nblk; ++c) {
    const int n = nblk - 2 * c;
    if (u < n) {
      *rb = 2 * c + u;
      *cp = c;
      return;
    }
    u -= n;
  }
  *rb = *cp = -1;  // u past the last pair tile (not reached for u < xg_pair_count)
}

// k-th tile of XCD `xcd`'s share (see xg_launch.h).  Returns false past the end.
// Within a job the XCD's channels are visited in two passes: first, for each
// channel in turn, the largest multiple W of its tiles that fills whole rounds
// of the XCD's W persistent workgroups (so a round works on one channel and
// its tiles share their input in L2), then the remaining tiles of all the
// channels together.
__host__ __device__ inline bool xg_tile(const XgLaunch &L, int xcd, int k, TileRef &t) {
  const int *start = L.xcd_start[xcd];
  if (k >= start[L.njobs]) return false;
  int j = 0;
  while (k >= start[j + 1]) ++j;
  const XgJob &J = L.jobs[j];
  const int local = k - start[j];
  const int T = J.tiles_per_chan;
  const int nch = (start[j + 1] - start[j]) / T;
  const int W = L.wg_per_xcd > 0 ? L.wg_per_xcd : 1;
  const int full = (T / W) * W;  // tiles per channel in the first pass
  int ci, u;
  if (local < nch * full) {
    ci = local / full;
    u = local - ci * full;
  } else {
    const int rest = T - full, l2 = local - nch * full;
    ci = l2 / rest;
    u = full + (l2 - ci * rest);
  }
  const int f0 = (((xcd - J.chan_base) % XG_NXCD) + XG_NXCD) % XG_NXCD;
  t.job = j;
  t.f = f0 + XG_NXCD * ci;
  if (J.pair_tiles)
    xg_pair_order(J.nblk, u, &t.rb, &t.cb);
  else
    xg_tile_order(J.nblk, u, &t.rb, &t.cb);
  return true;
}

// Host view of the schedule, for tests: the tile the kernels take as the k-th
// of XCD `xcd`.  Returns 0 past the end of the XCD's share.
extern "C" int xgh_schedule(const XgLaunch *L, int xcd, int k, int *job, int *f, int *rb,
                            int *cb) {
  TileRef t;
  if (!xg_tile(*L, xcd, k, t)) return 0;
  *job = t.job;
  *f = t.f;
  *rb = t.rb;
  *cb = t.cb;
  return 1;
}

// ---------------------------------------------------------------------------
// LDS addressing: row = input within the tile side, chunk = 16-byte group of
// samples.  slot = chunk ^ ((row + row/8) & 7) gives distinct 16-byte bank
// groups both for 8 consecutive rows at one chunk (operand reads) and for the
// rows 4g+q / 2g+q written by 8 consecutive staging lanes.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lds_off(int side, int row, int chunk) {
  const int slot = chunk ^ ((row + (row >> 3)) & 7);
  return side * XG_SIDE_BYTES + row * XG_ROWB + (slot << 4);
}

// ---------------------------------------------------------------------------
// Visibility store: input indices m (row) and n (column) within channel f.
// Handles the triangle cut, the three output orders and store/accumulate.
// Orders follow reference src/kernel.cuh:49-117 and src/cpu_util.c:57-132.
// ---------------------------------------------------------------------------
template <typename V>
__device__ __forceinline__ void put(void *base, long long idx, V v, int acc) {
  V *p = (V *)base + idx;
  if (acc) *p = *p + v; else *p = v;
}

__device__ __forceinline__ void put_i32(void *base, long long idx, int v, int acc) {
  unsigned *p = (unsigned *)base + idx;
  if (acc) *p = *p + (unsigned)v; else *p = (unsigned)v;
}

// Output index of the real part and the distance to the imaginary part.
__device__ __forceinline__ bool vis_index(const XgJob &J, int f, int m, int n,
                                          long long &ire, long long &dim, bool &zero_only) {
  const int ninput = 2 * J.nstation;
  zero_only = false;
  if (m >= ninput || n >= ninput) return false;
  const int i = m >> 1, p1 = m & 1, j = n >> 1, p2 = n & 1;
  if (J.order == XG_ORDER_REGTILE) {
    const long long P = J.nstation >> 1;
    const long long RT = P * (P + 1) / 2;
    const int R = i >> 1, rx = i & 1, C = j >> 1, ry = j & 1;
    if (i < j) {
      // Quadrant "12" of a diagonal register tile is never written by the
      // reference (kernel.cuh:65-71) and stays zero after a clear.
      if (R != C || J.accumulate) return false;
      zero_only = true;
    }
    const long long l = ((long long)f * 4 + (2 * ry + rx)) * RT + (long long)R * (R + 1) / 2 + C;
    ire = (l * 2 + p1) * 2 + p2;
    dim = J.plane;
    return true;
  }
  if (i < j) return false;
  const long long NB = (long long)J.nstation * (J.nstation + 1) / 2;
  const long long k = (long long)f * NB + (long long)i * (i + 1) / 2 + j;
  if (J.order == XG_ORDER_REALIMAG) {
    ire = k * 4 + p1 * 2 + p2;
    dim = J.plane;
  } else {  // TRIANGULAR: interleaved complex
    ire = ((k * 2 + p1) * 2 + p2) * 2;
    dim = 1;
  }
  return true;
}

__device__ __forceinline__ void store_int(const XgJob &J, int f, int m, int n, int re, int im) {
  long long ire, dim;
  bool zero;
  if (!vis_index(J, f, m, n, ire, dim, zero)) return;
  if (zero) { re = 0; im = 0; }
  if (J.out_kind == XG_OUT_I64) {
    put<long long>(J.out, ire, (long long)re, J.accumulate);
    put<long long>(J.out, ire + dim, (long long)im, J.accumulate);
  } else {
    put_i32(J.out, ire, re, J.accumulate);
    put_i32(J.out, ire + dim, im, J.accumulate);
  }
}

__device__ __forceinline__ void store_flt(const XgJob &J, int f, int m, int n, float re, float im) {
  long long ire, dim;
  bool zero;
  if (!vis_index(J, f, m, n, ire, dim, zero)) return;
  if (zero) { re = 0.f; im = 0.f; }
  if (J.out_kind == XG_OUT_F64) {
    put<double>(J.out, ire, (double)re, J.accumulate);
    put<double>(J.out, ire + dim, (double)im, J.accumulate);
  } else {
    put<float>(J.out, ire, re, J.accumulate);
    put<float>(J.out, ire + dim, im, J.accumulate);
  }
}

// ---------------------------------------------------------------------------
// MFMA kernels (int8 and fp16 share the skeleton)
// ---------------------------------------------------------------------------
template <bool F16>
struct MfmaTraits;

// int8: a complex sample is 2 bytes; a 16-byte chunk holds 8 samples; one
// stage is 64 samples.  Staging task: 4 inputs x 8 samples (8-byte loads).
template <>
struct MfmaTraits<false> {
  static constexpr int kCplxBytes = 2;
  static constexpr int kStageSamples = 64;
  typedef v16i Acc;
};

// fp16: a complex sample is 4 bytes; a chunk holds 4 samples; one stage is 32
// samples.  Staging task: 2 inputs x 8 samples (8-byte loads).
template <>
struct MfmaTraits<true> {
  static constexpr int kCplxBytes = 4;
  static constexpr int kStageSamples = 32;
  typedef v16f Acc;
};

template <bool F16>
__device__ __forceinline__ void stage_load(uint2 (&d)[8], const unsigned char *src, bool valid,
                                           int t0, int ntime, long long tstride) {
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int t = t0 + k;
    if (valid && t < ntime)
      d[k] = *(const uint2 *)(src + (long long)t * tstride);
    else
      d[k] = make_uint2(0u, 0u);
  }
}

// Corner turn of one staging task into LDS.
//   int8: d[k] = samples t0+k of inputs 4g..4g+3 (2 bytes each).  Input q's 8
//         samples become one 16-byte chunk (chunk index h) of row 4g+q.
//   fp16: d[k] = samples t0+k of inputs 2g, 2g+1 (4 bytes each).  Input q's
//         8 samples become chunks 2h and 2h+1 of row 2g+q.
template <bool F16>
__device__ __forceinline__ void stage_store(unsigned char *lds, int side, int g, int h,
                                            const uint2 (&d)[8]) {
  if (!F16) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const unsigned sel = (q & 1) ? 0x07060302u : 0x05040100u;
      v4i w;
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const unsigned lo = (q >> 1) ? d[2 * s].y : d[2 * s].x;
        const unsigned hi = (q >> 1) ? d[2 * s + 1].y : d[2 * s + 1].x;
        w[s] = (int)__builtin_amdgcn_perm(hi, lo, sel);
      }
      *(v4i *)(lds + lds_off(side, 4 * g + q, h)) = w;
    }
  } else {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        v4i w;
#pragma unroll
        for (int s = 0; s < 4; ++s) w[s] = (int)(q ? d[4 * c + s].y : d[4 * c + s].x);
        *(v4i *)(lds + lds_off(side, 2 * g + q, 2 * h + c)) = w;
      }
    }
  }
}

// PF: also touch the lines of stage s+2 with one dword load per 128-byte line
// per thread (consumed a stage later into a sink), so the real loads of stage
// s+1 find their lines in L2.  Returns the sink.
template <bool F16, bool PF>
__device__ __forceinline__ unsigned mfma_tile(const XgJob &J, const TileRef &T, unsigned char *lds) {
  typedef MfmaTraits<F16> Tr;
  typedef typename Tr::Acc Acc;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int ninput = 2 * J.nstation;
  const bool diag = (T.rb == T.cb);

  // staging assignment
  const int g = F16 ? (tid & 63) : (tid & 31);
  const int h = F16 ? (tid >> 6) : (tid >> 5);
  const int per_task = F16 ? 2 : 4;  // inputs per staging task
  const unsigned char *in_f =
      (const unsigned char *)J.input + (long long)T.f * J.nstation * 2 * Tr::kCplxBytes;
  const int inp0 = T.rb * XG_TILE + per_task * g;
  const int inp1 = T.cb * XG_TILE + per_task * g;
  const bool v0 = inp0 < ninput;
  const bool v1 = (!diag) && inp1 < ninput;
  const unsigned char *src0 = in_f + (long long)inp0 * Tr::kCplxBytes;
  const unsigned char *src1 = in_f + (long long)inp1 * Tr::kCplxBytes;
  const int nstage = (J.ntime + Tr::kStageSamples - 1) / Tr::kStageSamples;

  // prefetch assignment: thread -> (side, sample, 128-byte segment) of a stage
  const int pf_side = tid >> 7, pf_u = tid & 127;
  const int pf_sample = pf_u / Tr::kCplxBytes, pf_seg = pf_u % Tr::kCplxBytes;
  const int pf_inp = (pf_side ? T.cb : T.rb) * XG_TILE + pf_seg * (128 / Tr::kCplxBytes);
  const bool pf_ok = PF && pf_inp < ninput && !(pf_side && diag);
  const unsigned char *pf_src = in_f + (long long)pf_inp * Tr::kCplxBytes;
  unsigned pf_prev = 0, pf_sink = 0;
  auto prefetch = [&](int st) {
    const int t = st * Tr::kStageSamples + pf_sample;
    pf_sink ^= pf_prev;
    pf_prev = (pf_ok && t < J.ntime) ? *(const unsigned *)(pf_src + (long long)t * J.tstride) : 0u;
  };

  const int wr = wave >> 1, wc = wave & 1;
  const bool active = !(diag && wr < wc);
  const int half = lane >> 5, r32 = lane & 31;
  const int colside = diag ? 0 : 1;

  Acc accR[2][2], accI[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        accR[a][b][e] = 0;
        accI[a][b][e] = 0;
      }
  int csum[2] = {0, 0};

  uint2 d0[8], d1[8];
  stage_load<F16>(d0, src0, v0, 8 * h, J.ntime, J.tstride);
  stage_load<F16>(d1, src1, v1, 8 * h, J.ntime, J.tstride);
  if (PF) prefetch(1);
  stage_store<F16>(lds, 0, g, h, d0);
  if (!diag) stage_store<F16>(lds, 1, g, h, d1);
  __syncthreads();

  for (int s = 0; s < nstage; ++s) {
    const int buf = s & 1;
    const bool more = (s + 1) < nstage;
    if (more) {
      const int t0 = (s + 1) * Tr::kStageSamples + 8 * h;
      stage_load<F16>(d0, src0, v0, t0, J.ntime, J.tstride);
      stage_load<F16>(d1, src1, v1, t0, J.ntime, J.tstride);
    }
    if (PF && s + 2 < nstage) prefetch(s + 2);
    if (active) {
      const unsigned char *base = lds + buf * XG_STAGE_BYTES;
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        const int chunk = 2 * ks + half;
        v4i A[2], Ap[2], B[2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
          A[a] = *(const v4i *)(base + lds_off(0, wr * 64 + a * 32 + r32, chunk));
#pragma unroll
        for (int b = 0; b < 2; ++b)
          B[b] = *(const v4i *)(base + lds_off(colside, wc * 64 + b * 32 + r32, chunk));
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const unsigned x = (unsigned)A[a][e];
            if (!F16)  // (ar,ai) -> (ai,~ar) per complex byte pair
              Ap[a][e] = (int)(__builtin_amdgcn_perm(x, x, 0x02030001u) ^ 0xFF00FF00u);
            else       // (re,im) -> (im,-re) per complex half pair
              Ap[a][e] = (int)(((x >> 16) | (x << 16)) ^ 0x80000000u);
          }
        if (!F16) {
#pragma unroll
          for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int e = 0; e < 4; ++e)
              csum[b] = __builtin_amdgcn_sdot4(B[b][e], 0x01000100, csum[b], false);
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            if (!F16) {
              accR[a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[a], B[b], accR[a][b], 0, 0, 0);
              accI[a][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(Ap[a], B[b], accI[a][b], 0, 0, 0);
            } else {
              accR[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(
                  __builtin_bit_cast(v8h, A[a]), __builtin_bit_cast(v8h, B[b]), accR[a][b], 0, 0, 0);
              accI[a][b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(
                  __builtin_bit_cast(v8h, Ap[a]), __builtin_bit_cast(v8h, B[b]), accI[a][b], 0, 0, 0);
            }
          }
      }
    }
    if (more) {
      unsigned char *nb = lds + (buf ^ 1) * XG_STAGE_BYTES;
      stage_store<F16>(nb, 0, g, h, d0);
      if (!diag) stage_store<F16>(nb, 1, g, h, d1);
    }
    __syncthreads();
  }

  pf_sink ^= pf_prev;
  if (!active) return pf_sink;
  if (!F16) {
    // lanes l and l^32 hold the two K-halves of column l&31
    csum[0] += __shfl_xor(csum[0], 32);
    csum[1] += __shfl_xor(csum[1], 32);
  }
  // 32x32 accumulator layout: element e of lane l is row 8*(e/4)+4*(l/32)+e%4,
  // column l%32.
  const int m0 = T.rb * XG_TILE + wr * 64;
  const int n0 = T.cb * XG_TILE + wc * 64;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int m = m0 + a * 32 + 8 * (e >> 2) + 4 * half + (e & 3);
        const int n = n0 + b * 32 + r32;
        if (!F16) {
          const int im = (int)((unsigned)accI[a][b][e] + (unsigned)csum[b]);
          store_int(J, T.f, m, n, accR[a][b][e], im);
        } else {
          store_flt(J, T.f, m, n, accR[a][b][e], accI[a][b][e]);
        }
      }
  return pf_sink;
}

template <bool F16, bool PF>
__global__ __attribute__((amdgpu_flat_work_group_size(256, 256), amdgpu_waves_per_eu(2, 2)))
void xg_xengine_mfma(XgLaunch L) {
  __shared__ __attribute__((aligned(16))) unsigned char lds[XG_LDS_BYTES];
  const int xcd = blockIdx.x % XG_NXCD;
  const int wl = blockIdx.x / XG_NXCD;
  TileRef T;
  unsigned sink = 0;
  for (int k = wl; xg_tile(L, xcd, k, T); k += L.wg_per_xcd) {
    sink ^= mfma_tile<F16, PF>(L.jobs[T.job], T, lds);
    __syncthreads();
  }
  // njobs is never negative: keeps the prefetch loads alive without a store
  if (PF && L.njobs < 0) *(unsigned *)L.jobs[0].out = sink;
}

// ---------------------------------------------------------------------------
// Real-Gram int8 MFMA kernel (the default int8 path for nstation % 4 == 0).
//
// Each complex input is two real rows (re, im); one int8 MFMA stream
// accumulates the real Gram matrix G[r][c] = sum_t y_r(t) y_c(t) of the
// tile's 256 row-reals x 256 column-reals, and the epilogue combines
//   Re V(m,n) = G[m_re][n_re] + G[m_im][n_im]
//   Im V(m,n) = G[m_im][n_re] - G[m_re][n_im]
// in int32: the real and imaginary planes are reordered into two
// accumulators and subtracted, never negated, so -128 is exact.
//
// Data path: raw input bytes go global -> VGPR -> LDS unchanged (16-byte
// buffer loads; the hardware range check returns zeros past the block and for
// inputs past the station count), one 256-byte LDS row per sample and tile
// side.  MFMA operands come out of LDS through the gfx950 transpose load
// ds_read_b64_tr_b8: a 16-lane group supplies 16 eight-byte rows (sample k,
// half h) and lane j receives byte j of the chunk for 8 consecutive samples --
// one real row, time-contiguous.  4 waves of 128x128 reals (4x4 blocks of
// v_mfma_i32_32x32x32_i8, 256 accumulator registers), one workgroup per CU.
//
// Software pipeline (one wave per SIMD, so nothing else hides latency):
//  * two operand register sets: the 16 transpose reads of kstep k+1 are
//    interleaved with the 16 MFMAs of kstep k;
//  * one barrier per stage, placed before the last kstep, whose MFMAs then
//    cover the reads of the next stage's first kstep, the LDS writes of stage
//    s+2 and the global loads of stage s+4 (two whole stages before their data
//    is written to LDS; two staging register sets).
// ---------------------------------------------------------------------------
typedef int v2i __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) v2i lds_v2i;

#define RG_ROW 256  // LDS bytes per sample per tile side (128 inputs x 2 bytes)

// E: element format (RgInt8 or RgHalf, below); STG samples per stage.
template <class E, int STG>
struct RgCfg {
  static constexpr int kSide = STG * E::kRow;            // bytes per side per stage
  static constexpr int kBuf = 2 * kSide;                 // rows + columns
  static constexpr int kLds = 2 * kBuf;                  // double buffered
  static constexpr int kPass = XG_WG_THREADS / E::kChunks;  // sample rows per staging pass
  static constexpr int kLoads = STG / kPass;             // 16-byte loads per thread per side
  static constexpr int kSteps = STG / E::kK;             // MFMA ksteps per stage
};

// 16-byte slot of chunk c in a sample row with (sample & 7) == s7.  The 16
// chunks of a row fill the 16 slots (staging writes are conflict free), and a
// transpose read of chunks (c, c+1), c even, over 8 consecutive samples lands
// on 8 slots of each half of the row: all 64 banks.
__device__ __forceinline__ int rg_slot(int chunk, int s7) {
  return chunk ^ s7 ^ ((chunk & 1) << 3);
}

// Per-lane LDS offset of a 32-real-row block operand (chunks cb, cb+1 of the
// 256-byte sample rows) for K samples 0..31 of a kstep.  In a 16-lane group,
// supplier r gives the 8-byte half (r & 1) of sample (r >> 1) + 8 k; since
// every sample offset added later is a multiple of 8, (sample & 7) is the
// lane constant r >> 1.
__device__ __forceinline__ int rg_lane_off(int chunk_base, int lane) {
  const int g = (lane >> 4) & 1, kg = lane >> 5, r = lane & 15;
  const int chunk = chunk_base + g;
  return (kg * 16 + (r >> 1)) * RG_ROW + (rg_slot(chunk, r >> 1) << 4) + ((r & 1) << 3);
}

// Two transpose loads: samples kbase + 16 kg + {0..7} and {8..15}.
__device__ __forceinline__ v4i rg_operand(const unsigned char *p) {
  const v2i lo = __builtin_amdgcn_ds_read_tr8_b64_v2i32((lds_v2i *)p);
  const v2i hi = __builtin_amdgcn_ds_read_tr8_b64_v2i32((lds_v2i *)(p + 8 * RG_ROW));
  v4i out;
  out[0] = lo.x;
  out[1] = lo.y;
  out[2] = hi.x;
  out[3] = hi.y;
  return out;
}

// FP16 format: a complex sample is 4 bytes (re, im halves), so a sample row
// of one tile side is 512 bytes and real row R of the tile is half R of the
// row, as in the int8 format.  32 chunks of 16 bytes (4 inputs) per row.
#define RGH_ROW 512

typedef short v4s __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) v4s lds_v4s;

// 16-byte slot of chunk c in a sample row with (sample & 3) == s3.  A staging
// pass writes 16 consecutive chunks of a row to 16 distinct slots; an operand
// read touches chunks 4m .. 4m+3 of 4 consecutive samples, which land on the 16
// slots of distinct bank positions (slot & 15): all 64 banks.
__device__ __forceinline__ int rgh_slot(int chunk, int s3) { return chunk ^ (s3 << 2); }

// Per-lane LDS offset of the operand of the 32-real-row block `blk` (chunks
// 4 blk .. 4 blk + 3) for samples 0..15 of a kstep.  ds_read_b64_tr_b16 in a
// 16-lane group: output lane j receives, as its half k, half (j & 3) of the
// 8 bytes that lane 4k + (j >> 2) supplies.  Supplier r = 4k + q (k = sample,
// q = 4-row group) gives real rows 16g + 4q .. +3, so output lane j holds
// real row 16g + j for 4 consecutive samples; lanes 32..63 take samples 8..15.
__device__ __forceinline__ int rgh_lane_off(int blk, int lane) {
  const int g = (lane >> 4) & 1, kg = lane >> 5, k = (lane >> 2) & 3, q = lane & 3;
  const int chunk = 4 * blk + 2 * g + (q >> 1);
  return (8 * kg + k) * RGH_ROW + (rgh_slot(chunk, k) << 4) + ((q & 1) << 3);
}

// Two transpose loads: samples kbase + 8 kg + {0..3} and {4..7}, i.e. the
// 8 halves of one real row that v_mfma_f32_32x32x16_f16 takes from a lane.
__device__ __forceinline__ v4i rgh_operand(const unsigned char *p) {
  const v4s lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4s *)p);
  const v4s hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4s *)(p + 4 * RGH_ROW));
  const v2i l2 = __builtin_bit_cast(v2i, lo), h2 = __builtin_bit_cast(v2i, hi);
  v4i out;
  out[0] = l2.x;
  out[1] = l2.y;
  out[2] = h2.x;
  out[3] = h2.y;
  return out;
}

// Element formats of the real-Gram kernels: LDS row of one sample on one tile
// side, its chunk swizzle, the operand of a 32-real-row block and the MFMA.
// Both keep 256 real rows per tile and a 16-byte operand per lane.
struct RgInt8 {
  typedef v16i Acc;
  static constexpr int kRow = RG_ROW;
  static constexpr int kChunks = 16;   // 16-byte chunks per row
  static constexpr int kInputs = 8;    // inputs per chunk
  static constexpr int kK = 32;        // samples per MFMA
  // the DMA kernel's diagonal waves issue only the 10 MFMA blocks on or below
  // the diagonal (1.7-2 % faster at NS=512, NF=64, NT=8192)
  static constexpr bool kDiagLower = true;
  __device__ static int slot(int chunk, int srow) { return rg_slot(chunk, srow & 7); }
  // the chunk that belongs in `slot` of sample row `srow` (inverse of slot)
  __device__ static int dma_chunk(int slot, int srow) {
    const int x = slot ^ (srow & 7);
    return x ^ ((x & 1) << 3);
  }
  __device__ static int lane_off(int blk, int lane) { return rg_lane_off(2 * blk, lane); }
  __device__ static v4i operand(const unsigned char *p) { return rg_operand(p); }
  __device__ static v16i mfma(v4i a, v4i b, v16i c) {
    return __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c, 0, 0, 0);
  }
};

struct RgHalf {
  typedef v16f Acc;
  static constexpr int kRow = RGH_ROW;
  static constexpr int kChunks = 32;
  static constexpr int kInputs = 4;
  static constexpr int kK = 16;
  // the DMA kernel's diagonal waves issue only the 10 MFMA blocks on or below
  // the diagonal (2 % faster at NS=256, NF=128, NT=2048)
  static constexpr bool kDiagLower = true;
  __device__ static int slot(int chunk, int srow) { return rgh_slot(chunk, srow & 3); }
  __device__ static int dma_chunk(int slot, int srow) { return slot ^ ((srow & 3) << 2); }
  __device__ static int lane_off(int blk, int lane) { return rgh_lane_off(blk, lane); }
  __device__ static v4i operand(const unsigned char *p) { return rgh_operand(p); }
  __device__ static v16f mfma(v4i a, v4i b, v16f c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(v8h, a),
                                                   __builtin_bit_cast(v8h, b), c, 0, 0, 0);
  }
};

// Workgroup barrier for LDS traffic only: waits for this wave's LDS operations
// (not its global loads, which stay in flight across it).  The memory clobber
// keeps the compiler from moving LDS accesses across it.
__device__ __forceinline__ void rg_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

template <typename T, T V>
struct integral_c {
  static constexpr T value = V;
};

// Epilogue of the real-Gram kernels.  Element e of lane l in block (a,b) is
// G[R][Cc] with R = 128wr + 32a + 8(e/4) + 4(l/32) + e%4 and
// Cc = 128wc + 32b + l%32.  Even lanes (column n_re) finish input row m = R/2
// for e%4 in {0,1}, odd lanes (column n_im) the row m+1 (e%4 in {2,3}); each
// takes the two values it lacks from its partner lane.  Integer sums wrap
// modulo 2^32 (the host bounds the block so that they do not overflow).
template <class Acc>
struct RgElem;
template <>
struct RgElem<v16i> {
  typedef int T;       // accumulator element
  typedef unsigned S;  // 32-bit output element
  // Non-temporal output stores (the int8 kernels: plain stores measured
  // 0.15 % slower at NS=512 and 0.45 % at NS=2048, where the input is
  // re-read from beyond L2 and the output would displace it).
  static constexpr bool kStreamOut = true;
};
template <>
struct RgElem<v16f> {
  typedef float T;
  typedef float S;
  // Plain output stores (the fp16 kernels: 0.6-1.2 % faster than
  // non-temporal ones at NS=256, NF=128, NT=2048).
  static constexpr bool kStreamOut = false;
};

__device__ __forceinline__ int rg_add(int a, int b) { return (int)((unsigned)a + (unsigned)b); }
__device__ __forceinline__ int rg_sub(int a, int b) { return (int)((unsigned)a - (unsigned)b); }
__device__ __forceinline__ float rg_add(float a, float b) { return a + b; }
__device__ __forceinline__ float rg_sub(float a, float b) { return a - b; }

// Exchange of a value between lanes l and l ^ 1 with DPP quad_perm(1, 0, 3, 2):
// no LDS-queue round trip, and the compiler folds it into the add or subtract
// that uses it.
struct RgDppSwap {
  __device__ __forceinline__ int operator()(int x) const {
    return __builtin_amdgcn_update_dpp(0, x, 0xB1, 0xF, 0xF, false);
  }
  __device__ __forceinline__ float operator()(float x) const {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0xB1, 0xF, 0xF, false));
  }
};

// The same exchange through ds_bpermute, with an address the caller computed
// from its own lane index.
struct RgBperm {
  int addr;  // 4 * (lane ^ 1)
  __device__ __forceinline__ int operator()(int x) const {
    return __builtin_amdgcn_ds_bpermute(addr, x);
  }
  __device__ __forceinline__ float operator()(float x) const {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(x)));
  }
};

template <class Acc, class SW, class T = typename RgElem<Acc>::T>
__device__ __forceinline__ void rg_combine(const Acc &blk, int q, bool odd, const SW &swap, T &re,
                                           T &im) {
  const T x0 = blk[4 * q + 0], x1 = blk[4 * q + 1];
  const T x2 = blk[4 * q + 2], x3 = blk[4 * q + 3];
  // each lane sends what its partner lacks: odd lanes their e%4 = 0,1
  // (column n_im of row m; e%4 = 0 negated), even lanes their e%4 = 2,3
  // (column n_re of m+1).  The negation on the sending side makes every
  // received value an addend or a minuend: with the DPP swap, a received
  // subtrahend compiled to v_subrev_u32_dpp, whose int32 results were wrong
  // on gfx950 (the float form was not affected).
  const T y0 = swap(odd ? rg_sub(T(0), x0) : x2);
  const T y1 = swap(odd ? x1 : x3);
  if (!odd) {  // x0 = G[mr][nr], x1 = G[mi][nr], y0 = -G[mr][ni], y1 = G[mi][ni]
    re = rg_add(x0, y1);
    im = rg_add(x1, y0);
  } else {     // x2 = G[m'r][ni], x3 = G[m'i][ni], y0 = G[m'r][nr], y1 = G[m'i][nr]
    re = rg_add(y0, x3);
    im = rg_sub(y1, x2);
  }
}


// 32-bit (int32 or float) stores of the three output orders, one whole
// cell per lane (the engine gives these kernels 32-bit output only; 64-bit
// dumps come from the spill / combine kernels).  A cell is a station pair
// (i, j) with its four polarization products (element 2 p1 + p2): 16 bytes
// per plane in register-tile and real/imag order, 32 bytes of interleaved
// (re, im) in triangular order.  The four lanes of a quad (p1 = l & 1,
// p2 = (l >> 1) & 1) hold one product each of the cells of column blocks
// b = 0..3 (station j = j0 + 8 b); a wave-private LDS transpose (ds_write_b32,
// ds_read_b128 per plane) gives lane l the whole cell of block b = l & 3;
// triangular order interleaves the two planes in registers.  One code path
// can serve the three orders (ORDER = 0).  Cells of consecutive j are
// consecutive in every order, so each store instruction writes runs of 16
// (register tile) or 32 cells instead of 4-byte pieces.  Index formulas as in vis_index (reference
// src/kernel.cuh:49-117).
#define RG_CELL_GROUP 144                     // LDS bytes per quad group: 2 x 64 + 16 padding
#define RG_CELL_SCRATCH (16 * RG_CELL_GROUP)  // per wave and step (16 quad groups)
#define RG_EPI_LDS (4 * 2 * RG_CELL_SCRATCH)  // 4 waves x 2 halves: the epilogue's LDS

// ORDER: an XG_ORDER_* constant, or 0 to read the order from the job.
template <int ORDER, class Acc, class SW>
__device__ __forceinline__ void rg_store_cells(const XgJob &J, const TileRef &T,
                                               Acc (&acc)[4][4], int wr, int wc, int lane,
                                               const SW &swap, unsigned char *lds) {
  typedef typename RgElem<Acc>::T E;
  typedef typename RgElem<Acc>::S S;
  typedef S S4 __attribute__((ext_vector_type(4)));
  const int order = ORDER ? ORDER : J.order;
  const bool tri = order == XG_ORDER_TRI, regt = order == XG_ORDER_REGTILE;
  const bool odd = lane & 1;
  const int hi = lane >> 5, mj = (lane & 31) >> 2;   // station offsets: i = i0 + hi, j = j0 + mj
  const int e = 2 * (lane & 1) + ((lane >> 1) & 1);  // this lane's product: 2 p1 + p2
  const int bq = lane & 3;                           // the cell (column block) this lane stores
  // quad group (hi, mj); the 144-byte group stride spreads the writes over the banks
  unsigned char *scr0 =
      lds + (wr * 2 + wc) * 2 * RG_CELL_SCRATCH + (hi * 8 + mj) * RG_CELL_GROUP;
  const int ns = J.nstation;
  const int accum = J.accumulate;
  const long long plane = J.plane;
  const int j = T.cb * 64 + 32 * wc + mj + 8 * bq;
  const bool diag = T.rb == T.cb;
  const bool full = !diag && (T.rb * 64 + 64 <= ns);  // every (i, j) valid, i > j
  S *base;
  if (regt) {  // cell (R, C) of quadrant (i & 1, j & 1) of channel f
    const long long RT = (long long)(ns >> 1) * ((ns >> 1) + 1) / 2;
    base = (S *)J.out + 4 * (((long long)T.f * 4 + 2 * (j & 1) + hi) * RT + (j >> 1));
  } else {     // cell i(i+1)/2 + j of channel f
    const long long NB = (long long)ns * (ns + 1) / 2;
    base = (S *)J.out + (tri ? 8 : 4) * ((long long)T.f * NB + j);
  }
  // Cells of step n = 4 a + q (row block a, row pair q): combined, sent
  // through scratch half n & 1 and read back whole.
  auto gather = [&](int n, S4 &cre, S4 &cim) {
    const int a = n >> 2, q = n & 3;
    unsigned char *scr = scr0 + (n & 1) * RG_CELL_SCRATCH;
    E re[4], im[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) rg_combine(acc[a][b], q, odd, swap, re[b], im[b]);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      *(S *)(scr + 16 * b + 4 * e) = (S)re[b];
      *(S *)(scr + 64 + 16 * b + 4 * e) = (S)im[b];
    }
    cre = *(const S4 *)(scr + 16 * bq);
    cim = *(const S4 *)(scr + 64 + 16 * bq);
  };
  auto put = [&](int n, const S4 &cre, const S4 &cim) {
    const int a = n >> 2, q = n & 3;
    S4 v0 = cre, v1 = cim;
    if (tri) {  // triangular order interleaves (re, im) by product
      v0 = S4{cre[0], cim[0], cre[1], cim[1]};
      v1 = S4{cre[2], cim[2], cre[3], cim[3]};
    }
    const int i = T.rb * 64 + 32 * wr + hi + 8 * a + 2 * q;
    if (!full) {
      if (i >= ns || j >= ns) return;
      if (i < j) {
        // quadrant "12" of a diagonal register-tile cell is never written by
        // the reference (kernel.cuh:65-71): zero after a clear
        if (!regt || (i >> 1) != (j >> 1) || accum) return;
        v0 = (S4)(S)0;
        v1 = (S4)(S)0;
      }
    }
    S *p, *p1;
    if (regt) {
      const long long R = i >> 1;
      p = base + 4 * (R * (R + 1) / 2);
      p1 = p + plane;
    } else {
      const long long rowi = (long long)i * (i + 1) / 2;
      p = base + (tri ? 8 : 4) * rowi;
      p1 = tri ? p + 4 : p + plane;
    }
    if (accum) {
      v0 += *(const S4 *)p;
      v1 += *(const S4 *)p1;
    }
    if constexpr (RgElem<Acc>::kStreamOut) {
      // streaming stores: the tile's output is not read again by this
      // launch and should not evict the input the next round reuses
      __builtin_nontemporal_store(v0, (S4 *)p);
      __builtin_nontemporal_store(v1, (S4 *)p1);
    } else {
      *(S4 *)p = v0;
      *(S4 *)p1 = v1;
    }
  };
  // Software-pipelined by one step: step n + 1's lane exchange and LDS
  // round trip are issued before step n's bounds checks and stores, whose
  // branches would otherwise keep the two steps apart.
  S4 c0, c1;
  gather(0, c0, c1);
#pragma unroll
  for (int n = 0; n < 15; ++n) {
    S4 n0, n1;
    gather(n + 1, n0, n1);
    put(n, c0, c1);
    c0 = n0;
    c1 = n1;
  }
  put(15, c0, c1);
}

// One copy of the stores per output order, int32 and float.  With the
// epilogue pipelined (each step's accumulator reads feed the next step's
// exchange), the per-order copies no longer make the compiler hoist those
// reads and spill them; they leave the int8 DMA kernels with no spill at all
// and measure faster than one code path that reads the order at run time.
template <class SW>
__device__ __forceinline__ void rg_epilogue(const XgJob &J, const TileRef &T, v16i (&acc)[4][4],
                                            int wr, int wc, int lane, const SW &swap,
                                            unsigned char *lds) {
  if (J.order == XG_ORDER_REGTILE)
    rg_store_cells<XG_ORDER_REGTILE>(J, T, acc, wr, wc, lane, swap, lds);
  else if (J.order == XG_ORDER_TRI)
    rg_store_cells<XG_ORDER_TRI>(J, T, acc, wr, wc, lane, swap, lds);
  else
    rg_store_cells<XG_ORDER_REALIMAG>(J, T, acc, wr, wc, lane, swap, lds);
}
template <class SW>
__device__ __forceinline__ void rg_epilogue(const XgJob &J, const TileRef &T, v16f (&acc)[4][4],
                                            int wr, int wc, int lane, const SW &swap,
                                            unsigned char *lds) {
  if (J.order == XG_ORDER_REGTILE)
    rg_store_cells<XG_ORDER_REGTILE>(J, T, acc, wr, wc, lane, swap, lds);
  else if (J.order == XG_ORDER_TRI)
    rg_store_cells<XG_ORDER_TRI>(J, T, acc, wr, wc, lane, swap, lds);
  else
    rg_store_cells<XG_ORDER_REALIMAG>(J, T, acc, wr, wc, lane, swap, lds);
}

// Output stage of a tile.  The lane index reaches the epilogue through an
// opaque copy, so the values derived from it are computed here instead of
// being hoisted out of the tile loop and spilled across it: a spill reload
// waits for every store and load issued before it, which would serialize the
// epilogue's stores.  On a diagonal tile the wave above the diagonal
// (wr < wc) duplicates its mirror and stores nothing.
// The 32-bit stores use the first RG_EPI_LDS bytes of LDS: every wave must
// be done with the stage buffers (the caller's barrier).
template <class Acc>
__device__ __forceinline__ void rg_finish(const XgJob &J, const TileRef &T, Acc (&acc)[4][4],
                                          unsigned char *lds) {
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  if (!(T.rb == T.cb && wr < wc))
    rg_epilogue(J, T, acc, wr, wc, lane, RgDppSwap{}, lds);
}

// Column block of the i-th of a kstep's 16 MFMAs (row block i / 4): the
// columns run backwards on odd rows, so consecutive MFMAs share an operand
// (one operand's bits change per MFMA instead of two every fourth).
__device__ __forceinline__ constexpr int rg_col(int i) { return ((i >> 2) & 1) ? 3 - (i & 3) : (i & 3); }

// Row and column block of the i-th of the 10 MFMA blocks on or below the
// diagonal (row block >= column block), in the same snake order: (0,0),
// (1,1), (1,0), (2,0), (2,1), (2,2), (3,3), (3,2), (3,1), (3,0).
__device__ __forceinline__ constexpr int rg_lrow(int i) { return i < 1 ? 0 : i < 3 ? 1 : i < 6 ? 2 : 3; }
__device__ __forceinline__ constexpr int rg_lcol(int i) {
  return i == 0 ? 0 : i == 1 ? 1 : i == 2 ? 0 : i < 6 ? i - 3 : 9 - i;
}

// scheduling-group masks (llvm.amdgcn.sched.group.barrier)
#define SG_MFMA 0x008
#define SG_VMEM_READ 0x020
#define SG_DS_READ 0x100
#define SG_DS_WRITE 0x200

// Global-load state of one tile for the register-staged real-Gram kernel.
// Staging: thread -> (chunk, first sample row); load i covers sample
// srow + kPass i of a stage.  Each load has its own buffer resource, based at
// its first sample row and sized to end at the block's last valid sample, and
// a zero scalar offset: the range check (per-lane offset vs size) then zeroes
// every sample past the block and every invalid chunk, and all offsets stay
// below 16 * tstride.  The fields are copies of the job's (kept in registers:
// the barriers' memory clobber would otherwise make the compiler reload them
// from the kernel arguments every stage).
template <class E, int STG>
struct RgLoader {
  typedef RgCfg<E, STG> C;
  const unsigned char *row_base, *col_base;
  long long load_step;   // bytes between loads i and i+1
  long long stage_step;  // bytes between stages
  int ntime, ts, row_bytes, voff0, voff1;

  __device__ __forceinline__ RgLoader(const XgJob &J, const TileRef &T) {
    const int tid = threadIdx.x;
    const int ninput = 2 * J.nstation;
    const bool diag = (T.rb == T.cb);
    ntime = J.ntime;
    ts = (int)J.tstride;  // < 2^24: the host keeps 128 * tstride < 2^31
    const int chunk = tid % E::kChunks, srow = tid / E::kChunks;
    // a sample row of one tile side holds 128 inputs: E::kRow / 128 bytes per input
    const unsigned char *in_f =
        (const unsigned char *)J.input + (long long)T.f * ninput * (E::kRow / XG_TILE);
    row_base = in_f + T.rb * E::kRow;
    col_base = in_f + T.cb * E::kRow;
    load_step = (long long)C::kPass * ts;
    stage_step = (long long)STG * ts;
    // bytes of a sample row that a valid chunk can touch (<= tstride)
    row_bytes = ts < E::kRow ? ts : E::kRow;
    // chunk covers inputs blk*128 + kInputs*chunk and the next kInputs - 1
    // (ninput is a multiple of kInputs here); an invalid chunk gets an offset
    // past any resource's size and reads zeros
    const int voff = srow * ts + 16 * chunk;
    voff0 = (T.rb * XG_TILE + E::kInputs * chunk < ninput) ? voff : 0x7fffff00;
    // diagonal tiles read their columns from the row side; the column loads
    // then fall outside the buffer and return zeros (no branch in the stage loop)
    voff1 = (!diag && T.cb * XG_TILE + E::kInputs * chunk < ninput) ? voff : 0x7fffff00;
  }

  // stage st of both tile sides into (r0, r1); past the block: zeros
  __device__ __forceinline__ void gload(v4i *r0, v4i *r1, int st) const {
    const long long sbase = st * stage_step;
#pragma unroll
    for (int i = 0; i < C::kLoads; ++i) {
      const int t0 = st * STG + C::kPass * i;       // first sample of this load
      // valid samples of the kPass, clamped at 0 before the product: past the
      // block ntime - t0 is far below 0 and (ntime - t0) * ts would wrap
      const int rows = max(min(ntime - t0, C::kPass), 0);
      // (rows - 1) full sample rows + the last one's reachable bytes; none: 0
      const int nrec = rows > 0 ? rows * ts - ts + row_bytes : 0;
      const long long boff = sbase + i * load_step;
      const __amdgpu_buffer_rsrc_t rs0 = __builtin_amdgcn_make_buffer_rsrc(
          (void *)(row_base + boff), (short)0, nrec, 0x00020000);
      const __amdgpu_buffer_rsrc_t rs1 = __builtin_amdgcn_make_buffer_rsrc(
          (void *)(col_base + boff), (short)0, nrec, 0x00020000);
      r0[i] = __builtin_amdgcn_raw_buffer_load_b128(rs0, voff0, 0, 0);
      r1[i] = __builtin_amdgcn_raw_buffer_load_b128(rs1, voff1, 0, 0);
    }
  }
};

// MFMA pipeline of one tile.  On entry the staging registers hold stages 0
// (set a) and 1 (set b), or have their loads in flight; on exit acc holds the
// tile's real Gram block and the staging registers hold loads past the block.
template <class E, int STG>
__device__ __forceinline__ void rg_tile(const XgJob &J, const TileRef &T, const RgLoader<E, STG> &G,
                                        unsigned char *lds, v4i *ra0, v4i *ra1, v4i *rb0,
                                        v4i *rb1, typename E::Acc (&acc)[4][4]) {
  typedef RgCfg<E, STG> C;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const bool diag = (T.rb == T.cb);
  const int nstage = (J.ntime + STG - 1) / STG;
  const int chunk = tid % E::kChunks, srow = tid / E::kChunks;
  const int wpos = srow * E::kRow + (E::slot(chunk, srow) << 4);  // LDS write offset
  auto lwrite = [&](const v4i *r0, const v4i *r1, int buf) {
    unsigned char *b = lds + buf * C::kBuf;
#pragma unroll
    for (int i = 0; i < C::kLoads; ++i) {
      *(v4i *)(b + wpos + C::kPass * i * E::kRow) = r0[i];
      *(v4i *)(b + C::kSide + wpos + C::kPass * i * E::kRow) = r1[i];
    }
  };

  const int wr = wave >> 1, wc = wave & 1;
  int aoff[4], boff[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) aoff[a] = E::lane_off(4 * wr + a, lane);
#pragma unroll
  for (int b = 0; b < 4; ++b) boff[b] = E::lane_off(4 * wc + b, lane) + (diag ? 0 : C::kSide);
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[a][b][e] = 0;

  v4i A0[4], B0[4], A1[4], B1[4];
  auto lread = [&](const unsigned char *base, v4i *A, v4i *B) {
#pragma unroll
    for (int a = 0; a < 4; ++a) A[a] = E::operand(base + aoff[a]);
#pragma unroll
    for (int b = 0; b < 4; ++b) B[b] = E::operand(base + boff[b]);
  };
  // One kstep: the 16 MFMAs on (A, B), with the 16 transpose reads of the next
  // kstep into (NA, NB) over the first 8 MFMAs and, when FILL, the staging
  // writes of one stage into LDS and the global loads of a later one over the
  // last 8.
  auto kstep = [&](auto fill_c, const v4i *A, const v4i *B, v4i *NA, v4i *NB,
                   const unsigned char *nbase, auto fill) {
    constexpr bool FILL = decltype(fill_c)::value;
    constexpr int kFill = 2 * C::kLoads / 8;  // staging writes (and loads) per MFMA
#pragma unroll
    for (int i = 0; i < 16; ++i)
      acc[i >> 2][rg_col(i)] = E::mfma(A[i >> 2], B[rg_col(i)], acc[i >> 2][rg_col(i)]);
    lread(nbase, NA, NB);
    fill();
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      __builtin_amdgcn_sched_group_barrier(SG_MFMA, 1, 0);
      __builtin_amdgcn_sched_group_barrier(SG_DS_READ, 2, 0);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      __builtin_amdgcn_sched_group_barrier(SG_MFMA, 1, 0);
      if constexpr (FILL) {
        __builtin_amdgcn_sched_group_barrier(SG_DS_WRITE, kFill, 0);
        __builtin_amdgcn_sched_group_barrier(SG_VMEM_READ, kFill, 0);
      }
    }
  };
  typedef integral_c<bool, true> yes;
  typedef integral_c<bool, false> no;
  auto none = [] {};

  // Pipeline.  Stage s+1 is in LDS buffer (s+1)&1 while stage s is read from
  // buffer s&1; after the barrier that ends the reads of buffer s&1, the last
  // kstep of stage s writes stage s+2 into it from the staging registers and
  // issues the loads of stage s+4 into the same registers, so a global load
  // has two whole stages to land.  Even stages use register set a, odd ones
  // set b (the stage loop is unrolled by two; a block with an odd number of
  // stages runs one more stage of zeros).  Prologue: stages 0 and 1 from the
  // staging registers into LDS, stages 2 and 3 in flight, kstep 0 operands in
  // set 0.
  //
  // The stage body has no branches (a branch around the MFMAs makes the
  // compiler copy the accumulators): past the last stage the global loads have
  // an empty range and read zeros without touching memory, and the staging
  // writes and operand reads go to the idle LDS buffer.
  lwrite(ra0, ra1, 0);
  G.gload(ra0, ra1, 2);
  __builtin_amdgcn_sched_barrier(0);  // set a's loads first: they are waited for first
  lwrite(rb0, rb1, 1);
  G.gload(rb0, rb1, 3);
  rg_barrier();
  if (diag && wr < wc) {
    // On a diagonal tile this wave's block is the mirror of wave (1, 0)'s and
    // is never stored: it does its share of the staging (the same writes,
    // loads and stage barriers, in the same order) and no operand reads or
    // MFMAs, which the chip's power budget gives to the other waves.
    for (int s = 0; s < nstage; s += 2) {
      rg_barrier();
      lwrite(ra0, ra1, 0);
      G.gload(ra0, ra1, s + 4);
      rg_barrier();
      lwrite(rb0, rb1, 1);
      G.gload(rb0, rb1, s + 5);
    }
    return;  // nothing to store (rg_finish skips this wave on a diagonal tile)
  }
  lread(lds, A0, B0);
  auto stage = [&](int s, const unsigned char *cur, const unsigned char *nxt, v4i *r0, v4i *r1,
                   int buf) {
    // ksteps 0 .. kSteps-2 of this stage
#pragma unroll
    for (int ks = 0; ks < C::kSteps - 1; ++ks) {
      if (ks & 1)
        kstep(no(), A1, B1, A0, B0, cur + E::kK * (ks + 1) * E::kRow, none);
      else
        kstep(no(), A0, B0, A1, B1, cur + E::kK * (ks + 1) * E::kRow, none);
    }
    rg_barrier();  // every wave is done reading this stage's buffer
    // last kstep: reads kstep 0 of stage s+1 into set 0, writes stage s+2 into
    // this stage's buffer and issues the global loads of stage s+4
    auto refill = [&] {
      lwrite(r0, r1, buf);
      G.gload(r0, r1, s + 4);
    };
    if constexpr ((C::kSteps - 1) & 1)
      kstep(yes(), A1, B1, A0, B0, nxt, refill);
    else
      kstep(yes(), A0, B0, A1, B1, nxt, refill);
  };
  for (int s = 0; s < nstage; s += 2) {
    stage(s, lds, lds + C::kBuf, ra0, ra1, 0);
    stage(s + 1, lds + C::kBuf, lds, rb0, rb1, 1);
  }
}

// Persistent loop of the register-staged real-Gram kernels.  The first two
// stages of the next tile are loaded into the staging registers while the
// current tile's output is stored, so a tile starts with its data on chip.
template <class E, int STG>
__device__ __forceinline__ void rg_run(const XgLaunch &L, unsigned char *lds) {
  typedef RgCfg<E, STG> C;
  static_assert(C::kLds >= RG_EPI_LDS, "the epilogue's scratch fits in the stage buffers");
  const int xcd = blockIdx.x % XG_NXCD;
  const int wl = blockIdx.x / XG_NXCD;
  TileRef T;
  if (!xg_tile(L, xcd, wl, T)) return;
  v4i ra0[C::kLoads], ra1[C::kLoads], rb0[C::kLoads], rb1[C::kLoads];
  RgLoader<E, STG> G(L.jobs[T.job], T);
  G.gload(ra0, ra1, 0);
  G.gload(rb0, rb1, 1);
  for (int k = wl;; k += L.wg_per_xcd) {
    typename E::Acc acc[4][4];
    rg_tile<E, STG>(L.jobs[T.job], T, G, lds, ra0, ra1, rb0, rb1, acc);
    TileRef N;
    const bool more = xg_tile(L, xcd, k + L.wg_per_xcd, N);
    if (more) {
      G = RgLoader<E, STG>(L.jobs[N.job], N);
      G.gload(ra0, ra1, 0);
      G.gload(rb0, rb1, 1);
    }
    rg_barrier();  // the last stage's writes into the idle buffer are done
    rg_finish(L.jobs[T.job], T, acc, lds);
    rg_barrier();  // every wave is done with this tile's LDS buffers
    if (!more) break;
    T = N;
  }
}

// int8 input (nstation % 4 == 0)
template <int STG>
__global__ __attribute__((amdgpu_flat_work_group_size(256, 256), amdgpu_waves_per_eu(1, 1)))
void xg_xengine_rg(XgLaunch L) {
  __shared__ __attribute__((aligned(16))) unsigned char lds[RgCfg<RgInt8, STG>::kLds];
  rg_run<RgInt8, STG>(L, lds);
}

// fp16 input: v_mfma_f32_32x32x16_f16, float accumulation
template <int STG>
__global__ __attribute__((amdgpu_flat_work_group_size(256, 256), amdgpu_waves_per_eu(1, 1)))
void xg_xengine_rgh(XgLaunch L) {
  __shared__ __attribute__((aligned(16))) unsigned char lds[RgCfg<RgHalf, STG>::kLds];
  rg_run<RgHalf, STG>(L, lds);
}


// ---------------------------------------------------------------------------
// Real-Gram kernel with global -> LDS DMA staging.  Same arithmetic, LDS
// layout and MFMA pipeline as rg_tile, but the input goes straight from memory
// into LDS (buffer_load_dwordx4 ... lds: lane L writes LDS bytes
// [M0 + 16 L, +16), an out-of-range lane writes zeros), so there are no
// staging registers and no LDS write instructions, and NBUF stage buffers
// (128 KiB) let each stage's loads be issued NBUF-1 stages before its data is
// read.  Each lane loads the chunk that belongs in its LDS slot (E::dma_chunk,
// the inverse of E::slot): for int8, slot = rg_slot(chunk, sample & 7), i.e.
// chunk = g(slot ^ (sample & 7)) with g(x) = x ^ ((x & 1) << 3); for fp16,
// chunk = slot ^ 4 (sample & 3).  Both element formats (int8 xg_xengine_rgd,
// fp16 xg_xengine_rgdh) share the tile function.
//
// The DMAs are issued from inline assembly: the compiler would otherwise
// wait for every outstanding LDS DMA before each LDS read.  Each wave waits
// for its own DMAs of the next stage (vmcnt) in the stage barrier; nothing
// else in the stage loop uses the vector-memory counter.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void rg_dma16(v4i rsrc, int voff, unsigned m0) {
  asm volatile("buffer_load_dwordx4 %0, %1, 0 offen lds" ::"v"(voff), "s"(rsrc), "{m0}"(m0));
}

// Stage barrier of the DMA kernel: this wave's DMAs except the newest N have
// landed in LDS, its LDS reads are done, and every wave got here.  No
// instruction is scheduled across it: MFMAs (which touch no memory) would
// otherwise move over it, and a stage would reach its barrier with the
// operand reads of the next stage just issued instead of 8 MFMAs earlier
// (1-2 % slower, int8 and fp16).
template <int N>
__device__ __forceinline__ void rgd_barrier() {
  __builtin_amdgcn_sched_barrier(0);
  asm volatile("s_waitcnt vmcnt(%0)\n\ts_waitcnt lgkmcnt(0)\n\ts_barrier" ::"n"(N) : "memory");
  __builtin_amdgcn_sched_barrier(0);
}

// E: element format (RgInt8 or RgHalf).  A 1 KiB wave DMA covers kSpd
// consecutive sample rows (4 int8 rows of 256 bytes, 2 fp16 rows of 512).
template <class E, int STG, int NBUF>
struct RgdCfg {
  static constexpr int kSide = STG * E::kRow;
  static constexpr int kBuf = 2 * kSide;
  static constexpr int kLds = NBUF * kBuf;
  static constexpr int kSpd = 1024 / E::kRow;     // sample rows per DMA
  static constexpr int kDma = STG / (4 * kSpd);   // DMAs per wave per side per stage
  static constexpr int kSteps = STG / E::kK;
};

template <class E, int STG, int NBUF>
__device__ __forceinline__ void rgd_tile(const XgJob &J, const TileRef &T, unsigned char *lds) {
  typedef RgdCfg<E, STG, NBUF> C;
  static_assert((NBUF & (NBUF - 1)) == 0 && NBUF >= 2, "NBUF: power of two");
  static_assert(C::kLds >= RG_EPI_LDS, "the epilogue's scratch fits in the stage buffers");
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int ninput = 2 * J.nstation;
  const bool diag = (T.rb == T.cb);
  const int nstage = (J.ntime + STG - 1) / STG;
  const int ntime = J.ntime;
  const int ts = (int)J.tstride;  // < 2^24: the host keeps 128 * tstride < 2^31
  const unsigned char *in_f =
      (const unsigned char *)J.input + (long long)T.f * J.nstation * (2 * E::kRow / XG_TILE);
  const unsigned long long row_base = (unsigned long long)(in_f + T.rb * E::kRow);
  const unsigned long long col_base = (unsigned long long)(in_f + T.cb * E::kRow);
  const int row_bytes = ts < E::kRow ? ts : E::kRow;
  const unsigned lds_base = (unsigned)(unsigned long long)lds;

  // DMA lane mapping: wave DMA j of a stage is DMA g = 4 j + wave, sample rows
  // g kSpd .. +kSpd-1; lane L fills slot L % kChunks of row L / kChunks.  The
  // row's swizzle phase (sample & 7, or & 3 for fp16) is the same for every
  // DMA of this wave, so each lane loads one fixed chunk.
  const int srow = lane / E::kChunks, slot = lane % E::kChunks;
  const int chunk = E::dma_chunk(slot, wave * C::kSpd + srow);
  const int dvoff = srow * ts + 16 * chunk;
  const int dvoff0 = (T.rb * XG_TILE + E::kInputs * chunk < ninput) ? dvoff : 0x7fffff00;
  const int dvoff1 = (!diag && T.cb * XG_TILE + E::kInputs * chunk < ninput) ? dvoff : 0x7fffff00;
  // the DMAs of stage st, in order: side 0 then side 1, j = 0 .. kDma-1
  auto dma = [&](int st, int side, int j) {
    const int g = 4 * j + wave;
    const int s0 = st * STG + g * C::kSpd;  // first sample of this DMA
    // valid sample rows, clamped at 0 before the product: past the block
    // ntime - s0 is far below 0 and (ntime - s0) * ts would wrap to a large
    // positive range
    const int rows = max(min(ntime - s0, C::kSpd), 0);
    const int nrec = rows > 0 ? rows * ts - ts + row_bytes : 0;
    const unsigned long long a = (side ? col_base : row_base) + (unsigned long long)(long long)s0 * ts;
    v4i rs;
    rs.x = (int)(unsigned)a;
    rs.y = (int)((a >> 32) & 0xffff);
    rs.z = nrec;
    rs.w = 0x00020000;
    const unsigned m0 = lds_base + (st & (NBUF - 1)) * C::kBuf + side * C::kSide + g * 1024;
    rg_dma16(rs, side ? dvoff1 : dvoff0, m0);
  };
  auto dma_stage = [&](int st) {
#pragma unroll
    for (int side = 0; side < 2; ++side)
#pragma unroll
      for (int j = 0; j < C::kDma; ++j) dma(st, side, j);
  };

  const int wr = wave >> 1, wc = wave & 1;
  int aoff[4], boff[4];  // operand read offsets, set after the prologue barrier
  typename E::Acc acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[a][b][e] = 0;

  v4i A0[4], B0[4], A1[4], B1[4];
  auto lread = [&](const unsigned char *base, v4i *A, v4i *B) {
#pragma unroll
    for (int a = 0; a < 4; ++a) A[a] = E::operand(base + aoff[a]);
#pragma unroll
    for (int b = 0; b < 4; ++b) B[b] = E::operand(base + boff[b]);
  };
  // One kstep: 16 MFMAs on (A, B), the 16 transpose reads of the next kstep
  // over the first 8, and (when FILL) the DMAs of stage `st` over the last 8.
  // LOWER (a wave on the diagonal of a diagonal tile, E::kDiagLower): only
  // the 10 MFMA blocks on or below the diagonal, with the reads over the
  // first 5 and the DMAs over the last 5.  The 6 blocks above it hold
  // products of station pairs i < j, which no output order stores.
  auto kstep = [&](auto fill_c, auto lower_c, const v4i *A, const v4i *B, v4i *NA, v4i *NB,
                   const unsigned char *nbase, int st) {
    constexpr bool FILL = decltype(fill_c)::value;
    constexpr int kTot = 2 * C::kDma;  // DMAs of one stage
    if constexpr (!decltype(lower_c)::value) {
#pragma unroll
      for (int i = 0; i < 8; ++i) acc[i >> 2][rg_col(i)] = E::mfma(A[i >> 2], B[rg_col(i)], acc[i >> 2][rg_col(i)]);
      lread(nbase, NA, NB);
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        __builtin_amdgcn_sched_group_barrier(SG_MFMA, 1, 0);
        __builtin_amdgcn_sched_group_barrier(SG_DS_READ, 2, 0);
      }
#pragma unroll
      for (int i = 8; i < 16; ++i) {
        acc[i >> 2][rg_col(i)] = E::mfma(A[i >> 2], B[rg_col(i)], acc[i >> 2][rg_col(i)]);
        if constexpr (FILL) {
          // DMAs [(i-8) kTot / 8, (i-7) kTot / 8): every DMA exactly once
#pragma unroll
          for (int d = (i - 8) * kTot / 8; d < (i - 7) * kTot / 8; ++d) dma(st, d / C::kDma, d % C::kDma);
        }
      }
    } else {
#pragma unroll
      for (int i = 0; i < 5; ++i)
        acc[rg_lrow(i)][rg_lcol(i)] = E::mfma(A[rg_lrow(i)], B[rg_lcol(i)], acc[rg_lrow(i)][rg_lcol(i)]);
      lread(nbase, NA, NB);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        __builtin_amdgcn_sched_group_barrier(SG_MFMA, 1, 0);
        __builtin_amdgcn_sched_group_barrier(SG_DS_READ, 3, 0);
      }
      __builtin_amdgcn_sched_group_barrier(SG_MFMA, 1, 0);
      __builtin_amdgcn_sched_group_barrier(SG_DS_READ, 4, 0);
#pragma unroll
      for (int i = 5; i < 10; ++i) {
        acc[rg_lrow(i)][rg_lcol(i)] = E::mfma(A[rg_lrow(i)], B[rg_lcol(i)], acc[rg_lrow(i)][rg_lcol(i)]);
        if constexpr (FILL) {
          // DMAs [(i-5) kTot / 5, (i-4) kTot / 5): every DMA exactly once
#pragma unroll
          for (int d = (i - 5) * kTot / 5; d < (i - 4) * kTot / 5; ++d) dma(st, d / C::kDma, d % C::kDma);
        }
      }
    }
  };
  typedef integral_c<bool, true> yes;
  typedef integral_c<bool, false> no;

  // prologue: DMAs of stages 0 .. NBUF-1, wait for stage 0
#pragma unroll
  for (int st = 0; st < NBUF; ++st) dma_stage(st);
  rgd_barrier<(NBUF - 1) * 2 * C::kDma>();
  if (diag && wr < wc) {
    // On a diagonal tile this wave's block is the mirror of wave (1, 0)'s and
    // is never stored: it keeps its share of the DMAs and the stage barriers
    // (same order, same counts) and issues no MFMAs, which the chip's power
    // budget gives to the other waves.
    const int nrun = (C::kSteps & 1) ? (nstage + 1) & ~1 : nstage;
    for (int s = 0; s < nrun; ++s) {
      rgd_barrier<(NBUF - 2) * 2 * C::kDma>();
      dma_stage(s + NBUF);
    }
    rgd_barrier<0>();
    return;  // nothing to store (rg_finish skips this wave on a diagonal tile)
  }
  {
    // The operand offsets are computed here, from an opaque copy of the lane,
    // not hoisted out of the tile loop: kept live across it they were spilled,
    // and the reload's wait (vmcnt(0)) then waited for every prologue DMA
    // instead of stage 0's.
    int lane_o = threadIdx.x & 63;
    asm volatile("" : "+v"(lane_o));
#pragma unroll
    for (int a = 0; a < 4; ++a) aoff[a] = E::lane_off(4 * wr + a, lane_o);
#pragma unroll
    for (int b = 0; b < 4; ++b) boff[b] = E::lane_off(4 * wc + b, lane_o) + (diag ? 0 : C::kSide);
  }
  lread(lds, A0, B0);
  // One stage whose kstep-0 operands are in operand set P.  The last kstep
  // reads the next stage's kstep 0 into set (P + kSteps) & 1.
  auto run_stage = [&](int s, auto p_c, auto lower_c) {
    constexpr int P = decltype(p_c)::value;
    const unsigned char *cur = lds + (s & (NBUF - 1)) * C::kBuf;
    const unsigned char *nxt = lds + ((s + 1) & (NBUF - 1)) * C::kBuf;
#pragma unroll
    for (int ks = 0; ks < C::kSteps - 1; ++ks) {
      if ((P + ks) & 1)
        kstep(no(), lower_c, A1, B1, A0, B0, cur + E::kK * (ks + 1) * E::kRow, 0);
      else
        kstep(no(), lower_c, A0, B0, A1, B1, cur + E::kK * (ks + 1) * E::kRow, 0);
    }
    // stage s+1 has landed (this wave's DMAs; the barrier covers the others'),
    // and every wave is done reading stage s's buffer, which is refilled with
    // stage s+NBUF during the last kstep
    rgd_barrier<(NBUF - 2) * 2 * C::kDma>();
    if constexpr ((P + C::kSteps - 1) & 1)
      kstep(yes(), lower_c, A1, B1, A0, B0, nxt, s + NBUF);
    else
      kstep(yes(), lower_c, A0, B0, A1, B1, nxt, s + NBUF);
  };
  auto stage_loop = [&](auto lower_c) {
    if constexpr ((C::kSteps & 1) == 0) {
      for (int s = 0; s < nstage; ++s) run_stage(s, integral_c<int, 0>(), lower_c);
    } else {
      // the operand sets alternate between stages: two stages per iteration (a
      // block with an odd number of stages runs one more, zero-filled stage)
      for (int s = 0; s < nstage; s += 2) {
        run_stage(s, integral_c<int, 0>(), lower_c);
        run_stage(s + 1, integral_c<int, 1>(), lower_c);
      }
    }
  };
  if constexpr (E::kDiagLower) {
    if (diag && wr == wc)
      stage_loop(yes());
    else
      stage_loop(no());
  } else {
    stage_loop(no());
  }
  // DMAs of stages past the block (zero-filled) may still be in flight: drain
  // them before the next tile reuses the buffers
  rgd_barrier<0>();
  rg_finish(J, T, acc, lds);
}

// ---------------------------------------------------------------------------
// int8 real-Gram DMA kernel with 16x16x64 MFMA tiles (tuning mfma = 2).  The
// same LDS layout, DMA staging and stage barriers as rgd_tile at 64-sample
// stages (one kstep per stage), but each wave's 128 x 128 reals are 8 x 8
// blocks of v_mfma_i32_16x16x64_i8 (4 accumulators per lane and block, 256
// in all).  Operand of a 16-real block (one 16-byte chunk of the sample row):
// lane l takes real row l % 16 for samples 16 (l / 16) .. +15, two transpose
// reads of 8 samples each (the 16-lane-group rule of rg_lane_off with the
// whole wave on one chunk).  Accumulator element e of lane l is real row
// 4 (l / 16) + e, real column l % 16 of the block.  Twice the operand bytes
// per MAC of the 32x32x32 tiles: an autotuner alternative.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int rg16_lane_off(int chunk, int lane) {
  const int q = lane >> 4, r = lane & 15;
  return (q * 16 + (r >> 1)) * RG_ROW + (rg_slot(chunk, r >> 1) << 4) + ((r & 1) << 3);
}

// 32-bit stores of a 16x16-tile wave: for row block bu and a group of four
// column blocks (bv = 4 bvg + 0..3), a lane quad (l & ~3) holds the four
// polarization products (2 p1 + p2 = 2 (l & 1) + ((l >> 1) & 1)) of cell
// (i, j) of each block, i = 64 rb + 32 wr + 4 bu + (l >> 4), j = 64 cb +
// 32 wc + 4 bv + ((l & 15) >> 2).  A wave-private LDS transpose gives lane l
// the whole cell of block bv = 4 bvg + (l & 3); its 16-lane group then
// writes 16 cells of consecutive j.  Orders, bounds and the triangle cut as
// rg_store_cells (reference src/kernel.cuh:49-117).
__device__ __forceinline__ void rg16_store_cells(const XgJob &J, const TileRef &T, v4i (&acc)[8][8],
                                                 int wr, int wc, int lane, const RgBperm &swap,
                                                 unsigned char *lds) {
  typedef unsigned S4 __attribute__((ext_vector_type(4)));
  const int order = J.order;
  const bool tri = order == XG_ORDER_TRI, regt = order == XG_ORDER_REGTILE;
  const bool odd = lane & 1;
  const int e = 2 * (lane & 1) + ((lane >> 1) & 1);
  const int bq = lane & 3;
  unsigned char *scr = lds + (wr * 2 + wc) * RG_CELL_SCRATCH + (lane >> 2) * RG_CELL_GROUP;
  const int ns = J.nstation;
  const int accum = J.accumulate;
  const long long plane = J.plane;
  const bool diag = T.rb == T.cb;
  const bool full = !diag && (T.rb * 64 + 64 <= ns);
  const long long RT = (long long)(ns >> 1) * ((ns >> 1) + 1) / 2;
  const long long NB = (long long)ns * (ns + 1) / 2;
#pragma unroll
  for (int bu = 0; bu < 8; ++bu)
#pragma unroll
    for (int bvg = 0; bvg < 2; ++bvg) {
      __builtin_amdgcn_sched_barrier(0);
      int re[4], im[4];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        v16i blk;
        blk[0] = acc[bu][4 * bvg + b][0];
        blk[1] = acc[bu][4 * bvg + b][1];
        blk[2] = acc[bu][4 * bvg + b][2];
        blk[3] = acc[bu][4 * bvg + b][3];
        rg_combine(blk, 0, odd, swap, re[b], im[b]);
      }
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        *(unsigned *)(scr + 16 * b + 4 * e) = (unsigned)re[b];
        *(unsigned *)(scr + 64 + 16 * b + 4 * e) = (unsigned)im[b];
      }
      const S4 cre = *(const S4 *)(scr + 16 * bq);
      const S4 cim = *(const S4 *)(scr + 64 + 16 * bq);
      S4 v0 = cre, v1 = cim;
      if (tri) {
        v0 = S4{cre[0], cim[0], cre[1], cim[1]};
        v1 = S4{cre[2], cim[2], cre[3], cim[3]};
      }
      const int i = T.rb * 64 + 32 * wr + 4 * bu + (lane >> 4);
      const int j = T.cb * 64 + 32 * wc + 16 * bvg + 4 * bq + ((lane & 15) >> 2);
      if (!full) {
        if (i >= ns || j >= ns) continue;
        if (i < j) {
          if (!regt || (i >> 1) != (j >> 1) || accum) continue;
          v0 = (S4)0u;
          v1 = (S4)0u;
        }
      }
      unsigned *p, *p1;
      if (regt) {
        const long long R = i >> 1;
        p = (unsigned *)J.out +
            4 * (((long long)T.f * 4 + 2 * (j & 1) + (i & 1)) * RT + R * (R + 1) / 2 + (j >> 1));
        p1 = p + plane;
      } else {
        const long long k = (long long)T.f * NB + (long long)i * (i + 1) / 2 + j;
        p = (unsigned *)J.out + (tri ? 8 : 4) * k;
        p1 = tri ? p + 4 : p + plane;
      }
      if (accum) {
        v0 += *(const S4 *)p;
        v1 += *(const S4 *)p1;
      }
      __builtin_nontemporal_store(v0, (S4 *)p);
      __builtin_nontemporal_store(v1, (S4 *)p1);
    }
}

template <int NBUF>
__device__ __forceinline__ void rgd16_tile(const XgJob &J, const TileRef &T, unsigned char *lds) {
  typedef RgInt8 E;
  constexpr int STG = 64;
  typedef RgdCfg<E, STG, NBUF> C;
  static_assert(C::kDma == 4, "4 DMAs per wave per side per stage");
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int ninput = 2 * J.nstation;
  const bool diag = (T.rb == T.cb);
  const int nstage = (J.ntime + STG - 1) / STG;
  const int ntime = J.ntime;
  const int ts = (int)J.tstride;  // < 2^24: the host keeps 128 * tstride < 2^31
  const unsigned char *in_f =
      (const unsigned char *)J.input + (long long)T.f * J.nstation * (2 * E::kRow / XG_TILE);
  const unsigned long long row_base = (unsigned long long)(in_f + T.rb * E::kRow);
  const unsigned long long col_base = (unsigned long long)(in_f + T.cb * E::kRow);
  const int row_bytes = ts < E::kRow ? ts : E::kRow;
  const unsigned lds_base = (unsigned)(unsigned long long)lds;

  // DMA lane mapping and resources exactly as rgd_tile
  const int srow = lane / E::kChunks, slot = lane % E::kChunks;
  const int chunk = E::dma_chunk(slot, wave * C::kSpd + srow);
  const int dvoff = srow * ts + 16 * chunk;
  const int dvoff0 = (T.rb * XG_TILE + E::kInputs * chunk < ninput) ? dvoff : 0x7fffff00;
  const int dvoff1 = (!diag && T.cb * XG_TILE + E::kInputs * chunk < ninput) ? dvoff : 0x7fffff00;
  auto dma = [&](int st, int side, int j) {
    const int g = 4 * j + wave;
    const int s0 = st * STG + g * C::kSpd;
    const int rows = max(min(ntime - s0, C::kSpd), 0);  // clamped before the product
    const int nrec = rows > 0 ? rows * ts - ts + row_bytes : 0;
    const unsigned long long a = (side ? col_base : row_base) + (unsigned long long)(long long)s0 * ts;
    v4i rs;
    rs.x = (int)(unsigned)a;
    rs.y = (int)((a >> 32) & 0xffff);
    rs.z = nrec;
    rs.w = 0x00020000;
    const unsigned m0 = lds_base + (st & (NBUF - 1)) * C::kBuf + side * C::kSide + g * 1024;
    rg_dma16(rs, side ? dvoff1 : dvoff0, m0);
  };
  auto dma_stage = [&](int st) {
#pragma unroll
    for (int side = 0; side < 2; ++side)
#pragma unroll
      for (int j = 0; j < C::kDma; ++j) dma(st, side, j);
  };

  const int wr = wave >> 1, wc = wave & 1;
  // operand offsets: 16-real block b of the wave is chunk 8 w + b of its side
  const int aoff = rg16_lane_off(8 * wr, lane) - (rg_slot(8 * wr, (lane & 15) >> 1) << 4);
  const int boff = rg16_lane_off(8 * wc, lane) - (rg_slot(8 * wc, (lane & 15) >> 1) << 4) +
                   (diag ? 0 : C::kSide);
  const int s7 = (lane & 15) >> 1;  // the sample's swizzle phase (sample offsets are multiples of 8)
  v4i acc[8][8];
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int b = 0; b < 8; ++b) acc[a][b] = (v4i)0;

  v4i A0[8], B0[8], A1[8], B1[8];
  auto operand = [&](const unsigned char *base, int off, int blk) {
    return rg_operand(base + off + (rg_slot(blk, s7) << 4));
  };
  // One kstep (64 samples): MFMA i on (A, B) in snake order, the 16 operand
  // reads of the next stage over the first 32, the stage's 8 DMAs over the
  // last 32.
  auto kstep = [&](const v4i *A, const v4i *B, v4i *NA, v4i *NB, const unsigned char *nbase,
                   int st) {
#pragma unroll
    for (int i = 0; i < 64; ++i) {
      const int a = i >> 3, b = ((i >> 3) & 1) ? 7 - (i & 7) : (i & 7);
      acc[a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[a], B[b], acc[a][b], 0, 0, 0);
      if (i < 32 && (i & 1) == 0) {
        const int k = i >> 1;  // operand k: A0..A7 then B0..B7
        if (k < 8)
          NA[k] = operand(nbase, aoff, 8 * wr + k);
        else
          NB[k - 8] = operand(nbase, boff, 8 * wc + k - 8);
      }
      if (i >= 32 && (i & 3) == 0) {
        const int d = (i - 32) >> 2;  // DMA d of the stage: side d / 4, j = d % 4
        dma(st, d >> 2, d & 3);
      }
    }
  };

  // prologue: DMAs of stages 0 .. NBUF-1, wait for stage 0
#pragma unroll
  for (int st = 0; st < NBUF; ++st) dma_stage(st);
  rgd_barrier<(NBUF - 1) * 2 * C::kDma>();
  const int nrun = (nstage + 1) & ~1;  // operand sets alternate: stages in pairs
  if (diag && wr < wc) {
    // the mirror wave of a diagonal tile: its DMAs and barriers, no MFMAs
    for (int s = 0; s < nrun; ++s) {
      rgd_barrier<(NBUF - 2) * 2 * C::kDma>();
      dma_stage(s + NBUF);
    }
    rgd_barrier<0>();
    return;
  }
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    A0[k] = operand(lds, aoff, 8 * wr + k);
    B0[k] = operand(lds, boff, 8 * wc + k);
  }
  for (int s = 0; s < nrun; s += 2) {
    // stage s + 1 has landed (this wave's DMAs; the barrier covers the
    // others'), and every wave is done reading stage s's buffer, which is
    // refilled with stage s + NBUF during the kstep
    rgd_barrier<(NBUF - 2) * 2 * C::kDma>();
    kstep(A0, B0, A1, B1, lds + ((s + 1) & (NBUF - 1)) * C::kBuf, s + NBUF);
    rgd_barrier<(NBUF - 2) * 2 * C::kDma>();
    kstep(A1, B1, A0, B0, lds + ((s + 2) & (NBUF - 1)) * C::kBuf, s + 1 + NBUF);
  }
  rgd_barrier<0>();
  int tid2 = threadIdx.x;
  asm volatile("" : "+v"(tid2));  // lane-derived values computed here, not hoisted
  const int lane2 = tid2 & 63;
  if (!(diag && wr < wc)) rg16_store_cells(J, T, acc, wr, wc, lane2, RgBperm{4 * (lane2 ^ 1)}, lds);
}

template <int NBUF>
__global__ __attribute__((amdgpu_flat_work_group_size(256, 256), amdgpu_waves_per_eu(1, 1)))
void xg_xengine_rgd16(XgLaunch L) {
  __shared__ __attribute__((aligned(16))) unsigned char lds[RgdCfg<RgInt8, 64, NBUF>::kLds];
  const int xcd = blockIdx.x % XG_NXCD;
  const int wl = blockIdx.x / XG_NXCD;
  TileRef T;
  for (int k = wl; xg_tile(L, xcd, k, T); k += L.wg_per_xcd) {
    rgd16_tile<NBUF>(L.jobs[T.job], T, lds);
    rg_barrier();  // every wave is done with this tile's LDS buffers
  }
}

template <int STG, int NBUF>
__global__ __attribute__((amdgpu_flat_work_group_size(256, 256), amdgpu_waves_per_eu(1, 1)))
void xg_xengine_rgd(XgLaunch L) {
  __shared__ __attribute__((aligned(16))) unsigned char lds[RgdCfg<RgInt8, STG, NBUF>::kLds];
  const int xcd = blockIdx.x % XG_NXCD;
  const int wl = blockIdx.x / XG_NXCD;
  TileRef T;
  for (int k = wl; xg_tile(L, xcd, k, T); k += L.wg_per_xcd) {
    rgd_tile<RgInt8, STG, NBUF>(L.jobs[T.job], T, lds);
    rg_barrier();  // every wave is done with this tile's LDS buffers
  }
}

// fp16 input, global -> LDS DMA staging (16-byte aligned input)
template <int STG, int NBUF>
__global__ __attribute__((amdgpu_flat_work_group_size(256, 256), amdgpu_waves_per_eu(1, 1)))
void xg_xengine_rgdh(XgLaunch L) {
  __shared__ __attribute__((aligned(16))) unsigned char lds[RgdCfg<RgHalf, STG, NBUF>::kLds];
  const int xcd = blockIdx.x % XG_NXCD;
  const int wl = blockIdx.x / XG_NXCD;
  TileRef T;
  for (int k = wl; xg_tile(L, xcd, k, T); k += L.wg_per_xcd) {
    rgd_tile<RgHalf, STG, NBUF>(L.jobs[T.job], T, lds);
    rg_barrier();  // every wave is done with this tile's LDS buffers
  }
}

// ---------------------------------------------------------------------------
// K-interleaved int8 kernel (family XG_KERNEL_I8_KG).  The MFMA's K runs over
// (sample, re/im) pairs: a complex input is one 16-bit element (re byte, im
// byte), as the input stores it, so a tile needs one accumulator per output
// value instead of the real-Gram's four per complex product:
//   Re(m, n) = sum_t re_m re_n + im_m im_n = A . B
//   Im(m, n) = sum_t im_m re_n - re_m im_n = A'' . B + C_n
// with A = (re_m, im_m) and B = (re_n, im_n) per sample, A'' = (im_m, ~re_m)
// (~x = -x - 1 is exact for every int8, where -x is not for -128) and the
// column sum C_n = sum_t im_n, which the lanes accumulate with dot4 on the B
// operands.  With one accumulator per output value, a workgroup holds a pair
// tile of 128 row inputs x 256 column inputs (2 x 4 Re and 2 x 4 Im 32 x 32
// blocks per wave): 768 input bytes per sample for 131,072 MACs, against 512
// bytes for 65,536 in the real-Gram kernel.
//
// LDS: stages of 32 samples, 4 buffers (3 stages of DMA in flight).  A sample
// row is 256 bytes (row side) or 512 bytes (column side) of 16-byte chunks
// (8 inputs); chunk c of a row of sample s sits in slot c ^ ((s & 3) << 2).
// Operands: ds_read_b64_tr_b16 pairs as in the fp16 kernel (each lane gets 8
// consecutive samples of one input).  Waves: 2 x 2, wave (wr, wc) owns rows
// 64 wr .. +63 and columns 128 wc .. +127 of the tile.
// ---------------------------------------------------------------------------
#define KG_AROW 256                     // LDS bytes per sample, row side (128 inputs)
#define KG_BROW 512                     // column side (256 inputs)
#define KG_STG 32                       // samples per stage
#define KG_NBUF 4                       // stage buffers
#define KG_SIDEA (KG_STG * KG_AROW)     // 8 KiB
#define KG_BUF (KG_SIDEA + KG_STG * KG_BROW)  // 24 KiB per stage
#define KG_DMA 6                        // DMAs per wave per stage: 2 row-side, 4 column-side

__device__ __forceinline__ int kg_slot(int chunk, int s3) { return chunk ^ (s3 << 2); }

// Per-lane LDS offset of the operand of 32-input block `blk` for samples
// 0..15 of a kstep, in rows of `row` bytes (the tr16 lane rule of
// rgh_lane_off, with inputs in place of real rows).
__device__ __forceinline__ int kg_lane_off(int blk, int lane, int row) {
  const int g = (lane >> 4) & 1, kg = lane >> 5, k = (lane >> 2) & 3, q = lane & 3;
  const int chunk = 4 * blk + 2 * g + (q >> 1);
  return (8 * kg + k) * row + (kg_slot(chunk, k) << 4) + ((q & 1) << 3);
}

template <int ROW>
__device__ __forceinline__ v4i kg_operand(const unsigned char *p) {
  const v4s lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4s *)p);
  const v4s hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4s *)(p + 4 * ROW));
  const v2i l2 = __builtin_bit_cast(v2i, lo), h2 = __builtin_bit_cast(v2i, hi);
  v4i out;
  out[0] = l2.x;
  out[1] = l2.y;
  out[2] = h2.x;
  out[3] = h2.y;
  return out;
}

// (re, im) -> (im, ~re) in every 16-bit element
__device__ __forceinline__ v4i kg_rotate(v4i a) {
  v4i o;
#pragma unroll
  for (int i = 0; i < 4; ++i)
    o[i] = (int)(__builtin_amdgcn_perm((unsigned)a[i], (unsigned)a[i], 0x02030001u) ^ 0xFF00FF00u);
  return o;
}

// DPP exchange with the neighbouring lane (l ^ 1)
__device__ __forceinline__ int kg_swap1(int x) {
  return __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, true);  // quad_perm [1, 0, 3, 2]
}

// 32-bit stores of one pair tile, one whole cell per lane.  For block (a, b)
// and element group q, lane l holds column input n = 256 cp + 128 wc + 32 b +
// (l & 31) (station j, polarization p2 = l & 1) against 4 row inputs
// (stations i0, i0 + 1, both p1), i0 = 64 rb + 32 wr + 16 a + 4 q +
// 2 (l >> 5).  With its neighbour l ^ 1 (the other p2) it completes cell
// (i0, j) (even lane) and (i0 + 1, j) (odd lane): 4 products in order
// 2 p1 + p2.  One code path serves the three orders (only the address and,
// for triangular order, the re/im interleave differ), so the compiler has
// no per-order copies of the accumulator reads to hoist.  Index formulas as
// in vis_index (reference src/kernel.cuh:49-117).
__device__ __forceinline__ void kg_store_cells(const XgJob &J, const TileRef &T, v16i (&are)[2][4],
                                               v16i (&aim)[2][4], const int (&corr)[4], int wr,
                                               int wc, int lane) {
  typedef unsigned S4 __attribute__((ext_vector_type(4)));
  const bool odd = lane & 1;
  const int ns = J.nstation;
  const bool accum = J.accumulate;
  const bool regt = J.order == XG_ORDER_REGTILE, tri = J.order == XG_ORDER_TRI;
  const long long plane = J.plane;
  const int jl = (lane & 31) >> 1;
  const int i_lane = T.rb * 64 + 32 * wr + 2 * (lane >> 5) + (odd ? 1 : 0);
  const bool full = (T.cb * 128 + 127 < T.rb * 64) && (T.rb * 64 + 64 <= ns);
  const long long RT = (long long)(ns >> 1) * ((ns >> 1) + 1) / 2;
  const long long NB = (long long)ns * (ns + 1) / 2;
  unsigned *const out = (unsigned *)J.out;
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        __builtin_amdgcn_sched_barrier(0);
        int re[4], im[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          re[e] = are[a][b][4 * q + e];
          im[e] = rg_add(aim[a][b][4 * q + e], corr[b]);
        }
        // odd lanes send their e = 0, 1 (row i0), even lanes e = 2, 3 (row i0 + 1)
        const int r0 = kg_swap1(odd ? re[0] : re[2]), r1 = kg_swap1(odd ? re[1] : re[3]);
        const int m0 = kg_swap1(odd ? im[0] : im[2]), m1 = kg_swap1(odd ? im[1] : im[3]);
        S4 cr, ci;
        if (!odd) {
          cr = S4{(unsigned)re[0], (unsigned)r0, (unsigned)re[1], (unsigned)r1};
          ci = S4{(unsigned)im[0], (unsigned)m0, (unsigned)im[1], (unsigned)m1};
        } else {
          cr = S4{(unsigned)r0, (unsigned)re[2], (unsigned)r1, (unsigned)re[3]};
          ci = S4{(unsigned)m0, (unsigned)im[2], (unsigned)m1, (unsigned)im[3]};
        }
        const int i = i_lane + 16 * a + 4 * q;
        const int j = T.cb * 128 + 64 * wc + 16 * b + jl;
        if (!full) {
          if (i >= ns || j >= ns) continue;
          if (i < j) {
            // quadrant "12" of a diagonal register-tile cell is never written by
            // the reference (kernel.cuh:65-71): zero after a clear
            if (!regt || (i >> 1) != (j >> 1) || accum) continue;
            cr = (S4)0u;
            ci = (S4)0u;
          }
        }
        long long cell;  // first element of the cell (real plane / interleaved)
        if (regt)
          cell = 4 * ((((long long)T.f * 4 + 2 * (j & 1) + (i & 1)) * RT) +
                      (long long)(i >> 1) * ((i >> 1) + 1) / 2 + (j >> 1));
        else
          cell = (tri ? 8 : 4) * ((long long)T.f * NB + (long long)i * (i + 1) / 2 + j);
        unsigned *const p = out + cell;
        unsigned *const p1 = tri ? p + 4 : p + plane;
        S4 v0 = cr, v1 = ci;
        if (tri) {
          v0 = S4{cr[0], ci[0], cr[1], ci[1]};
          v1 = S4{cr[2], ci[2], cr[3], ci[3]};
        }
        if (accum) {
          v0 += *(const S4 *)p;
          v1 += *(const S4 *)p1;
        }
        // streaming stores: the tile's output is not read again by this
        // launch and should not evict the input the next round reuses
        __builtin_nontemporal_store(v0, (S4 *)p);
        __builtin_nontemporal_store(v1, (S4 *)p1);
      }
}

// Output stage of a pair tile (int32 output, the only kind the engine gives
// this kernel); the lane index reaches it through an opaque copy (see
// rg_finish).  The column correction needs both sample halves of the
// column's sums (lanes l and l ^ 32).
__device__ __forceinline__ void kg_finish(const XgJob &J, const TileRef &T, v16i (&are)[2][4],
                                          v16i (&aim)[2][4], const int (&csum)[4]) {
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  int corr[4];
#pragma unroll
  for (int b = 0; b < 4; ++b)
    corr[b] = rg_add(csum[b], __builtin_amdgcn_ds_bpermute(4 * (lane ^ 32), csum[b]));
  kg_store_cells(J, T, are, aim, corr, wr, wc, lane);
}

__device__ __forceinline__ void kg_tile(const XgJob &J, const TileRef &T, unsigned char *lds) {
  // the lane constants below are recomputed per tile (opaque copy of the
  // thread index) instead of being hoisted out of the tile loop and spilled
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int ninput = 2 * J.nstation;
  const int nstage = (J.ntime + KG_STG - 1) / KG_STG;
  const int ntime = J.ntime;
  const int ts = (int)J.tstride;  // < 2^24: the host keeps 128 * tstride < 2^31
  const unsigned char *in_f = (const unsigned char *)J.input + (long long)T.f * J.nstation * 4;
  const unsigned long long abase = (unsigned long long)(in_f + T.rb * KG_AROW);
  const unsigned long long bbase = (unsigned long long)(in_f + T.cb * KG_BROW);
  const int arow_bytes = ts < KG_AROW ? ts : KG_AROW;
  const int brow_bytes = ts < KG_BROW ? ts : KG_BROW;
  const unsigned lds_base = (unsigned)(unsigned long long)lds;

  // DMA lane mapping.  Row side: a DMA covers 4 samples of 256 bytes, lane L
  // fills slot L & 15 of sample row L >> 4 (sample & 3 == L >> 4).  Column
  // side: 2 samples of 512 bytes, lane L fills slot L & 31 of sample row
  // L >> 5; sample & 3 is (L >> 5) or 2 + (L >> 5) for even / odd DMAs.
  // Chunks past the last input read past the range: zeros.
  int voa, vob[2];
  {
    const int r = lane >> 4, c = kg_slot(lane & 15, r);
    voa = (T.rb * XG_TILE + 8 * c < ninput) ? r * ts + 16 * c : 0x7fffff00;
#pragma unroll
    for (int par = 0; par < 2; ++par) {
      const int rr = lane >> 5, cc = kg_slot(lane & 31, (2 * par + rr) & 3);
      vob[par] = (T.cb * 2 * XG_TILE + 8 * cc < ninput) ? rr * ts + 16 * cc : 0x7fffff00;
    }
  }
  // DMA d of stage st for this wave: d = 0, 1: row-side DMA 2 wave + d;
  // d = 2..5: column-side DMA 4 wave + d - 2.
  auto dma = [&](int st, int d) {
    v4i rs;
    int voff;
    unsigned m0 = lds_base + (st & (KG_NBUF - 1)) * KG_BUF;
    unsigned long long a;
    if (d < 2) {
      const int g = 2 * wave + d;
      const int s0 = st * KG_STG + 4 * g;
      const int rows = max(min(ntime - s0, 4), 0);  // clamped before the product
      rs.z = rows > 0 ? rows * ts - ts + arow_bytes : 0;
      a = abase + (unsigned long long)(long long)s0 * ts;
      voff = voa;
      m0 += 4 * g * KG_AROW;
    } else {
      const int j = 4 * wave + d - 2;
      const int s0 = st * KG_STG + 2 * j;
      const int rows = max(min(ntime - s0, 2), 0);  // clamped before the product
      rs.z = rows > 0 ? rows * ts - ts + brow_bytes : 0;
      a = bbase + (unsigned long long)(long long)s0 * ts;
      voff = vob[(d - 2) & 1];
      m0 += KG_SIDEA + 2 * j * KG_BROW;
    }
    rs.x = (int)(unsigned)a;
    rs.y = (int)((a >> 32) & 0xffff);
    rs.w = 0x00020000;
    rg_dma16(rs, voff, m0);
  };

  const int wr = wave >> 1, wc = wave & 1;
  int aoff[2], boff[4];
#pragma unroll
  for (int a = 0; a < 2; ++a) aoff[a] = kg_lane_off(2 * wr + a, lane, KG_AROW);
#pragma unroll
  for (int b = 0; b < 4; ++b) boff[b] = KG_SIDEA + kg_lane_off(4 * wc + b, lane, KG_BROW);
  v16i are[2][4], aim[2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        are[a][b][e] = 0;
        aim[a][b][e] = 0;
      }
  int csum[4] = {0, 0, 0, 0};

  v4i A0[2], B0[4], A1[2], B1[4];
  // operands of kstep kk (0: samples 0..15, 1: 16..31) of the stage at base
  auto lread = [&](const unsigned char *base, int kk, v4i *A, v4i *B) {
#pragma unroll
    for (int a = 0; a < 2; ++a) A[a] = kg_operand<KG_AROW>(base + aoff[a] + kk * 16 * KG_AROW);
#pragma unroll
    for (int b = 0; b < 4; ++b) B[b] = kg_operand<KG_BROW>(base + boff[b] + kk * 16 * KG_BROW);
  };
  // One kstep: 8 Re MFMAs, the next kstep's operand reads, 8 Im MFMAs over
  // which (FILL) the stage's DMAs are issued.
  auto kstep = [&](auto fill_c, const v4i *A, const v4i *B, v4i *NA, v4i *NB,
                   const unsigned char *nbase, int nkk, int st) {
    constexpr bool FILL = decltype(fill_c)::value;
    v4i Ar[2];
    Ar[0] = kg_rotate(A[0]);
    Ar[1] = kg_rotate(A[1]);
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i) csum[b] = __builtin_amdgcn_sdot4(B[b][i], 0x01000100, csum[b], false);
#pragma unroll
    for (int i = 0; i < 8; ++i)
      are[i >> 2][i & 3] =
          __builtin_amdgcn_mfma_i32_32x32x32_i8(A[i >> 2], B[i & 3], are[i >> 2][i & 3], 0, 0, 0);
    lread(nbase, nkk, NA, NB);
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      aim[i >> 2][i & 3] =
          __builtin_amdgcn_mfma_i32_32x32x32_i8(Ar[i >> 2], B[i & 3], aim[i >> 2][i & 3], 0, 0, 0);
      if constexpr (FILL) {
        if (i < KG_DMA) dma(st, i);
      }
    }
  };
  typedef integral_c<bool, true> yes;
  typedef integral_c<bool, false> no;

  // prologue: DMAs of stages 0 .. NBUF-1, wait for stage 0
#pragma unroll
  for (int st = 0; st < KG_NBUF; ++st)
#pragma unroll
    for (int d = 0; d < KG_DMA; ++d) dma(st, d);
  rgd_barrier<(KG_NBUF - 1) * KG_DMA>();
  lread(lds, 0, A0, B0);
  for (int s = 0; s < nstage; ++s) {
    const unsigned char *cur = lds + (s & (KG_NBUF - 1)) * KG_BUF;
    const unsigned char *nxt = lds + ((s + 1) & (KG_NBUF - 1)) * KG_BUF;
    kstep(no(), A0, B0, A1, B1, cur, 1, 0);
    // stage s+1 has landed (this wave's DMAs; the barrier covers the others'),
    // and every wave is done reading stage s's buffer, which is refilled with
    // stage s+NBUF during the last kstep
    rgd_barrier<(KG_NBUF - 2) * KG_DMA>();
    kstep(yes(), A1, B1, A0, B0, nxt, 0, s + KG_NBUF);
  }
  // DMAs of stages past the block (zero-filled) may still be in flight
  rgd_barrier<0>();
  kg_finish(J, T, are, aim, csum);
}

__global__ __attribute__((amdgpu_flat_work_group_size(256, 256), amdgpu_waves_per_eu(1, 1)))
void xg_xengine_kg(XgLaunch L) {
  __shared__ __attribute__((aligned(16))) unsigned char lds[KG_NBUF * KG_BUF];
  const int xcd = blockIdx.x % XG_NXCD;
  const int wl = blockIdx.x / XG_NXCD;
  TileRef T;
  for (int k = wl; xg_tile(L, xcd, k, T); k += L.wg_per_xcd) {
    kg_tile(L.jobs[T.job], T, lds);
    rg_barrier();  // every wave is done with this tile's LDS buffers
  }
}

// ---------------------------------------------------------------------------
// FP32 plain kernel: exact int8 -> float, SIMT FMA.  Thread (ty,tx) of a
// 16x16 layout owns rows ty+16a and columns tx+16b (a,b < 8) of the
// 128x128-input tile.  Stage: 16 samples, float2 per input per sample.
// ---------------------------------------------------------------------------
#define XG_F32_TK 16
#define XG_F32_SIDE (XG_F32_TK * XG_TILE)  // float2 elements per side per stage

__device__ __forceinline__ void f32_tile(const XgJob &J, const TileRef &T, float2 *lds) {
  const int tid = threadIdx.x;
  const int ty = tid >> 4, tx = tid & 15;
  const int ninput = 2 * J.nstation;
  const bool diag = (T.rb == T.cb);
  const int g = tid & 31, h = tid >> 5;  // staging: inputs 4g..4g+3, samples 2h,2h+1
  const unsigned char *in_f = (const unsigned char *)J.input + (long long)T.f * J.nstation * 4;
  const int inp0 = T.rb * XG_TILE + 4 * g, inp1 = T.cb * XG_TILE + 4 * g;
  const bool v0 = inp0 < ninput, v1 = (!diag) && inp1 < ninput;
  const unsigned char *src0 = in_f + (long long)inp0 * 2;
  const unsigned char *src1 = in_f + (long long)inp1 * 2;
  const int nstage = (J.ntime + XG_F32_TK - 1) / XG_F32_TK;

  float accR[8][8], accI[8][8];
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      accR[a][b] = 0.f;
      accI[a][b] = 0.f;
    }

  uint2 d0[2], d1[2];
  auto load = [&](int s) {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int t = s * XG_F32_TK + 2 * h + k;
      const bool ok = t < J.ntime;
      d0[k] = (v0 && ok) ? *(const uint2 *)(src0 + (long long)t * J.tstride) : make_uint2(0u, 0u);
      d1[k] = (v1 && ok) ? *(const uint2 *)(src1 + (long long)t * J.tstride) : make_uint2(0u, 0u);
    }
  };
  auto cvt_store = [&](float2 *side, const uint2 (&d)[2]) {
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      float4 lo, hi;
      const unsigned x = d[k].x, y = d[k].y;
      lo.x = (float)(signed char)(x);
      lo.y = (float)(signed char)(x >> 8);
      lo.z = (float)(signed char)(x >> 16);
      lo.w = (float)(signed char)(x >> 24);
      hi.x = (float)(signed char)(y);
      hi.y = (float)(signed char)(y >> 8);
      hi.z = (float)(signed char)(y >> 16);
      hi.w = (float)(signed char)(y >> 24);
      float4 *row = (float4 *)(side + (2 * h + k) * XG_TILE + 4 * g);
      row[0] = lo;
      row[1] = hi;
    }
  };

  load(0);
  cvt_store(lds, d0);
  if (!diag) cvt_store(lds + XG_F32_SIDE, d1);
  __syncthreads();
  for (int s = 0; s < nstage; ++s) {
    const int buf = s & 1;
    const bool more = (s + 1) < nstage;
    if (more) load(s + 1);
    const float2 *R = lds + buf * 2 * XG_F32_SIDE;
    const float2 *C = diag ? R : R + XG_F32_SIDE;
#pragma unroll 2
    for (int t = 0; t < XG_F32_TK; ++t) {
      float2 r[8], c[8];
#pragma unroll
      for (int a = 0; a < 8; ++a) r[a] = R[t * XG_TILE + ty + 16 * a];
#pragma unroll
      for (int b = 0; b < 8; ++b) c[b] = C[t * XG_TILE + tx + 16 * b];
#pragma unroll
      for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 8; ++b) {
          accR[a][b] = fmaf(r[a].x, c[b].x, accR[a][b]);
          accR[a][b] = fmaf(r[a].y, c[b].y, accR[a][b]);
          accI[a][b] = fmaf(r[a].y, c[b].x, accI[a][b]);
          accI[a][b] = fmaf(-r[a].x, c[b].y, accI[a][b]);
        }
    }
    if (more) {
      float2 *nb = lds + (buf ^ 1) * 2 * XG_F32_SIDE;
      cvt_store(nb, d0);
      if (!diag) cvt_store(nb + XG_F32_SIDE, d1);
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int b = 0; b < 8; ++b)
      store_flt(J, T.f, T.rb * XG_TILE + ty + 16 * a, T.cb * XG_TILE + tx + 16 * b, accR[a][b],
                accI[a][b]);
}

__global__ __attribute__((amdgpu_flat_work_group_size(256, 256), amdgpu_waves_per_eu(1, 2)))
void xg_xengine_f32(XgLaunch L) {
  __shared__ __attribute__((aligned(16))) float2 lds[4 * XG_F32_SIDE];
  const int xcd = blockIdx.x % XG_NXCD;
  const int wl = blockIdx.x / XG_NXCD;
  TileRef T;
  for (int k = wl; xg_tile(L, xcd, k, T); k += L.wg_per_xcd) {
    f32_tile(L.jobs[T.job], T, lds);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Support kernels
// ---------------------------------------------------------------------------
__global__ void xg_fill_zero_k(uint4 *p, size_t n16, unsigned *tail, size_t ntail) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
    p[i] = make_uint4(0u, 0u, 0u, 0u);
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ntail) tail[t] = 0u;
}

// Plain streaming copy (16-byte loads and stores, grid-stride).
__global__ void xg_copy_k(uint4 *dst, const uint4 *src, size_t n16, unsigned *dtail,
                          const unsigned *stail, size_t ntail) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
    dst[i] = src[i];
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ntail) dtail[t] = stail[t];
}

// Four consecutive 4- or 8-byte elements through 16-byte accesses (p 16-byte aligned).
template <typename T>
__device__ __forceinline__ void xg_ld4(const T *p, T (&o)[4]) {
  if constexpr (sizeof(T) == 4) {
    typedef T T4 __attribute__((ext_vector_type(4)));
    const T4 v = *(const T4 *)p;
    o[0] = v[0]; o[1] = v[1]; o[2] = v[2]; o[3] = v[3];
  } else {
    typedef T T2 __attribute__((ext_vector_type(2)));
    const T2 a = ((const T2 *)p)[0], b = ((const T2 *)p)[1];
    o[0] = a[0]; o[1] = a[1]; o[2] = b[0]; o[3] = b[1];
  }
}
template <typename T>
__device__ __forceinline__ void xg_st4(T *p, const T (&v)[4]) {
  if constexpr (sizeof(T) == 4) {
    typedef T T4 __attribute__((ext_vector_type(4)));
    *(T4 *)p = T4{v[0], v[1], v[2], v[3]};
  } else {
    typedef T T2 __attribute__((ext_vector_type(2)));
    ((T2 *)p)[0] = T2{v[0], v[1]};
    ((T2 *)p)[1] = T2{v[2], v[3]};
  }
}

// dst += src (ADD) or dst = src, widening each element.  V = 4: each thread
// takes 4 consecutive elements with 16-byte accesses (nv = n / 4, every
// pointer 16-byte aligned); V = 1: one element per thread (nv = n).
template <typename D, typename S, bool ADD, int V>
__global__ void xg_spill_k(D *dst, const S *src, size_t nv) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += stride) {
    if constexpr (V == 1) {
      dst[i] = ADD ? dst[i] + (D)src[i] : (D)src[i];
    } else {
      S s[4];
      D d[4];
      xg_ld4(src + 4 * i, s);
      if (ADD) xg_ld4(dst + 4 * i, d);
#pragma unroll
      for (int k = 0; k < 4; ++k) d[k] = ADD ? d[k] + (D)s[k] : (D)s[k];
      xg_st4(dst + 4 * i, d);
    }
  }
}

// dst = spill + cur, in the wide type W, stored as O (W, or its low 32 bits /
// float rounding for a narrow dump).  V as for xg_spill_k.
template <typename W, typename S, typename O, int V>
__global__ void xg_combine_k(O *dst, const W *spill, const S *cur, size_t nv) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += stride) {
    if constexpr (V == 1) {
      dst[i] = (O)(spill[i] + (W)cur[i]);
    } else {
      W w[4];
      S s[4];
      O o[4];
      xg_ld4(spill + 4 * i, w);
      xg_ld4(cur + 4 * i, s);
#pragma unroll
      for (int k = 0; k < 4; ++k) o[k] = (O)(w[k] + (W)s[k]);
      xg_st4(dst + 4 * i, o);
    }
  }
}

// Counter-based generator: byte i = lo + ((hi32(splitmix64(key + i)) * span) >> 32)
__device__ __forceinline__ uint64_t xg_mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__global__ void xg_fill_uniform_k(uint32_t *p, size_t nwords, uint64_t key, int lo, uint64_t span,
                                  unsigned char *tailp, size_t ntail, size_t tail0) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x; w < nwords; w += stride) {
    uint32_t v = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const uint64_t r = xg_mix64(key + 4 * w + b);
      const int x = lo + (int)(((r >> 32) * span) >> 32);
      v |= ((uint32_t)(x & 0xff)) << (8 * b);
    }
    p[w] = v;
  }
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ntail) {
    const uint64_t r = xg_mix64(key + tail0 + t);
    tailp[t] = (unsigned char)((lo + (int)(((r >> 32) * span) >> 32)) & 0xff);
  }
}

// Reorder split-plane REGISTER_TILE / REAL_IMAG into interleaved TRIANGULAR
// (reference src/cpu_util.c:57-132, run on the device).  One thread per
// triangular visibility (f, i>=j, p1, p2).  PAIR (64-bit elements, dst
// 16-byte aligned): the visibility's (re, im) go out as one 16-byte store, so
// a wave's store covers 1 KiB without gaps.
typedef unsigned long long xg_u64x2 __attribute__((ext_vector_type(2)));

// Source index (real plane) of triangular visibility x = (f, baseline, p1 p2):
// NB baselines per channel, RT station-pair cells per register-tile quadrant.
__device__ __forceinline__ long long xg_reorder_src(int src_order, long long x, long long NB,
                                                    long long RT) {
  const int pp = (int)(x & 3);
  const long long kk = x >> 2;
  if (src_order != XG_ORDER_REGTILE) return kk * 4 + pp;
  const int f = (int)(kk / NB);
  const long long b = kk - (long long)f * NB;
  int i = (int)((sqrt(8.0 * (double)b + 1.0) - 1.0) * 0.5);
  while ((long long)i * (i + 1) / 2 > b) --i;
  while ((long long)(i + 1) * (i + 2) / 2 <= b) ++i;
  const int j = (int)(b - (long long)i * (i + 1) / 2);
  const int R = i >> 1, rx = i & 1, C = j >> 1, ry = j & 1;
  const long long l = ((long long)f * 4 + (2 * ry + rx)) * RT + (long long)R * (R + 1) / 2 + C;
  return l * 4 + pp;
}

template <typename E, bool PAIR = false>
__global__ void xg_reorder_k(int src_order, E *dst, const E *src, int nstation, int nfreq,
                             long long plane) {
  const long long NB = (long long)nstation * (nstation + 1) / 2;
  const long long total = (long long)nfreq * NB * 4;
  const long long P = nstation >> 1;
  const long long RT = P * (P + 1) / 2;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x; x < total; x += stride) {
    const long long s = xg_reorder_src(src_order, x, NB, RT);
    if constexpr (PAIR) {
      __builtin_nontemporal_store(xg_u64x2{src[s], src[s + plane]}, (xg_u64x2 *)dst + x);
    } else {
      dst[2 * x] = src[s];
      dst[2 * x + 1] = src[s + plane];
    }
  }
}

// Dump of a spilled integration in one pass: dst (triangular, O) =
// spill (split planes, W) + cur (split planes, S), one thread per visibility;
// O is W (wide dump) or S (narrow: the sum modulo 2^32 for int32).  PAIR: dst
// aligned to 2 * sizeof(O), (re, im) as one store.
template <typename W, typename S, typename O, bool PAIR>
__global__ void xg_reorder_combine_k(int src_order, O *dst, const W *spill, const S *cur,
                                     int nstation, int nfreq, long long plane) {
  const long long NB = (long long)nstation * (nstation + 1) / 2;
  const long long total = (long long)nfreq * NB * 4;
  const long long P = nstation >> 1;
  const long long RT = P * (P + 1) / 2;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long x = (long long)blockIdx.x * blockDim.x + threadIdx.x; x < total; x += stride) {
    const long long s = xg_reorder_src(src_order, x, NB, RT);
    const O re = (O)(spill[s] + (W)cur[s]);
    const O im = (O)(spill[s + plane] + (W)cur[s + plane]);
    if constexpr (PAIR) {
      typedef O O2 __attribute__((ext_vector_type(2)));
      __builtin_nontemporal_store(O2{re, im}, (O2 *)dst + x);
    } else {
      dst[2 * x] = re;
      dst[2 * x + 1] = im;
    }
  }
}

// The same reorder for 32-bit elements, one half cell per thread.  A cell's
// 4 products are 16 consecutive bytes in each source plane (register tile and
// real/imag order alike) and 32 interleaved bytes in triangular order; thread
// y moves half h = y & 1 of cell y >> 1 (products 2h, 2h+1) with two 8-byte
// loads and one 16-byte store, so a wave's store covers 1 KiB of the
// destination without gaps.  src, dst 16-byte aligned and plane % 4 == 0
// (checked by the launcher).
typedef unsigned xg_u2 __attribute__((ext_vector_type(2)));
typedef unsigned xg_u4 __attribute__((ext_vector_type(4)));
__global__ void xg_reorder_cells_k(int src_order, xg_u4 *dst, const unsigned *src, int nstation,
                                   int nfreq, long long plane) {
  const long long NB = (long long)nstation * (nstation + 1) / 2;
  const long long total = 2 * (long long)nfreq * NB;
  const long long P = nstation >> 1;
  const long long RT = P * (P + 1) / 2;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long y = (long long)blockIdx.x * blockDim.x + threadIdx.x; y < total; y += stride) {
    const long long x = y >> 1;
    const int f = (int)(x / NB);
    const long long b = x - (long long)f * NB;
    int i = (int)((sqrt(8.0 * (double)b + 1.0) - 1.0) * 0.5);
    while ((long long)i * (i + 1) / 2 > b) --i;
    while ((long long)(i + 1) * (i + 2) / 2 <= b) ++i;
    const int j = (int)(b - (long long)i * (i + 1) / 2);
    long long s;
    if (src_order == XG_ORDER_REGTILE) {
      const int R = i >> 1, rx = i & 1, C = j >> 1, ry = j & 1;
      s = 4 * (((long long)f * 4 + (2 * ry + rx)) * RT + (long long)R * (R + 1) / 2 + C);
    } else {
      s = 4 * x;
    }
    s += 2 * (y & 1);
    const xg_u2 re = *(const xg_u2 *)(src + s);
    const xg_u2 im = *(const xg_u2 *)(src + s + plane);
    __builtin_nontemporal_store(xg_u4{re[0], im[0], re[1], im[1]}, dst + y);
  }
}

// dp4a input corner turn of depth 4 (reference src/cpu_util.c:235-255) on
// the device, both directions.  With row = bytes per time sample (NF NS 4)
// and e a byte within a sample row, swizzled byte ((t/4) row + e) 4 + t%4 is
// plain byte t row + e.  Thread w handles 4 consecutive bytes e = 4 e4 .. +3
// of the 4 samples of group t4 (w = t4 row/4 + e4): plain side four 32-bit
// words (one per sample row), swizzled side one 16-byte word (16 w), related
// by a 4 x 4 byte transpose.  UNSWIZZLE: swizzled -> plain.
__device__ __forceinline__ void xg_transpose4x4(const uint32_t (&in)[4], uint32_t (&out)[4]) {
#pragma unroll
  for (int e = 0; e < 4; ++e)
    out[e] = ((in[0] >> (8 * e)) & 0xffu) | (((in[1] >> (8 * e)) & 0xffu) << 8) |
             (((in[2] >> (8 * e)) & 0xffu) << 16) | (((in[3] >> (8 * e)) & 0xffu) << 24);
}

template <bool UNSWIZZLE>
__global__ void xg_swizzle_k(unsigned char *dst, const unsigned char *src, long long row4,
                             long long total) {
  const long long row = 4 * row4;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long w = (long long)blockIdx.x * blockDim.x + threadIdx.x; w < total; w += stride) {
    const long long t4 = w / row4, e4 = w - t4 * row4;
    unsigned char *plain = UNSWIZZLE ? dst : (unsigned char *)src;
    const long long base = 4 * t4 * row + 4 * e4;  // plain byte of (sample 4 t4, e = 4 e4)
    uint32_t in[4], out[4];
    if (UNSWIZZLE) {
      const uint4 v = ((const uint4 *)src)[w];
      in[0] = v.x;
      in[1] = v.y;
      in[2] = v.z;
      in[3] = v.w;
    } else {
#pragma unroll
      for (int b = 0; b < 4; ++b) in[b] = *(const uint32_t *)(plain + base + b * row);
    }
    xg_transpose4x4(in, out);
    if (UNSWIZZLE) {
#pragma unroll
      for (int b = 0; b < 4; ++b) *(uint32_t *)(plain + base + b * row) = out[b];
    } else {
      ((uint4 *)dst)[w] = make_uint4(out[0], out[1], out[2], out[3]);
    }
  }
}

// Packed Hermitian -> full matrix (reference xgpuExtractMatrix,
// src/cpu_util.c:258-279): out[f][i][j][p1][p2] = packed[f][k(i,j)][p1][p2]
// for i > j and conj(packed[f][k(j,i)][p2][p1]) for i < j, k(i,j) = i(i+1)/2
// + j, interleaved (re, im).  A diagonal cell (i = j) is written twice per
// polarization pair by the reference loop, and its last writes leave
// [a][b] = packed[a][b] for a > b and conj(packed[b][a]) for a <= b; the
// same values are produced here (for Hermitian input both readings agree).
// Negation as the reference does it on its element type (integer negation
// wraps for the most negative value; floating point flips the sign).
//
// One workgroup per tile of 16 x 16 stations on or below the diagonal (row
// tile a >= column tile b) of one channel: thread (il, jl) reads packed cell
// (i, j) = (16 a + il, 16 b + jl), i >= j, once (the 16 cells of a row are
// contiguous), writes out[f][i][j] (the rule above for a diagonal cell), and
// keeps the cell in LDS; after the barrier thread (il, jl) writes the mirror
// cell out[f][16 b + il][16 a + jl] (row < column) from the packed cell
// thread (jl, il) read.  Both writes store 16 contiguous cells per row.  Rows
// of the LDS tile are padded by one element, so the transposed reads of a
// wave fall on different banks.
#define XG_XT 16
template <typename T>
__global__ __launch_bounds__(256) void xg_extract_tile_k(T *out, const T *packed, int ns, int nf) {
  __shared__ T cell[XG_XT][XG_XT * 8 + 1];
  const int nt = (ns + XG_XT - 1) / XG_XT;
  const long long npair = (long long)nt * (nt + 1) / 2;
  const long long nb = (long long)ns * (ns + 1) / 2;
  const int il = threadIdx.x >> 4, jl = threadIdx.x & 15;
  for (long long w = blockIdx.x; w < npair * nf; w += gridDim.x) {
    const int f = (int)(w / npair);
    const long long p = w - (long long)f * npair;
    // pair p -> (a, b), b <= a: p = a (a + 1) / 2 + b
    int a = (int)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
    while ((long long)a * (a + 1) / 2 > p) --a;
    while ((long long)(a + 1) * (a + 2) / 2 <= p) ++a;
    const int b = (int)(p - (long long)a * (a + 1) / 2);
    const int i = a * XG_XT + il, j = b * XG_XT + jl;
    if (i < ns && j <= i) {
      const T *src = packed + ((long long)f * nb + (long long)i * (i + 1) / 2 + j) * 8;
      T v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = src[q];
      T o[8];
#pragma unroll
      for (int p1 = 0; p1 < 2; ++p1)
#pragma unroll
        for (int p2 = 0; p2 < 2; ++p2) {
          const bool direct = (i > j) || (p1 > p2);
          const int d = (p1 * 2 + p2) * 2, s2 = direct ? d : (p2 * 2 + p1) * 2;
          o[d] = v[s2];
          o[d + 1] = direct ? v[s2 + 1] : (T)(T(0) - v[s2 + 1]);
        }
      T *dst = out + (((long long)f * ns + i) * ns + j) * 8;
#pragma unroll
      for (int q = 0; q < 8; ++q) dst[q] = o[q];
#pragma unroll
      for (int q = 0; q < 8; ++q) cell[il][jl * 8 + q] = v[q];
    }
    __syncthreads();
    // mirror cell (row jm < column im) from packed cell (im, jm), read by thread (jl, il)
    const int jm = b * XG_XT + il, im = a * XG_XT + jl;
    if (im < ns && jm < im) {
      T v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = cell[jl][il * 8 + q];
      T o[8];
#pragma unroll
      for (int p1 = 0; p1 < 2; ++p1)
#pragma unroll
        for (int p2 = 0; p2 < 2; ++p2) {
          const int d = (p1 * 2 + p2) * 2, s2 = (p2 * 2 + p1) * 2;
          o[d] = v[s2];
          o[d + 1] = (T)(T(0) - v[s2 + 1]);
        }
      T *dst = out + (((long long)f * ns + jm) * ns + im) * 8;
#pragma unroll
      for (int q = 0; q < 8; ++q) dst[q] = o[q];
    }
    __syncthreads();  // the tile's LDS is reused by the next pair
  }
}

// COMPLEX_BLOCK_SIZE = 32 input -> plain [t][f][s][p][re,im] (reference
// src/omp_xengine.c:64-78 reads the same layout).  Each 64-byte block holds
// the real parts of 32 inputs (16 stations x 2 polarizations) in bytes 0..31
// and their imaginary parts in bytes 32..63; the plain layout interleaves
// them.  One thread per block reads all of it before writing, so dst may
// equal src.
__global__ void xg_unblock32_k(uint4 *dst, const uint4 *src, long long nblk) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x; b < nblk; b += stride) {
    const uint4 r0 = src[4 * b], r1 = src[4 * b + 1], i0 = src[4 * b + 2], i1 = src[4 * b + 3];
    const unsigned re[8] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
    const unsigned im[8] = {i0.x, i0.y, i0.z, i0.w, i1.x, i1.y, i1.z, i1.w};
    uint4 o[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      // bytes (re.b0, im.b0, re.b1, im.b1) and (re.b2, im.b2, re.b3, im.b3)
      o[j].x = __builtin_amdgcn_perm(im[2 * j], re[2 * j], 0x05010400u);
      o[j].y = __builtin_amdgcn_perm(im[2 * j], re[2 * j], 0x07030602u);
      o[j].z = __builtin_amdgcn_perm(im[2 * j + 1], re[2 * j + 1], 0x05010400u);
      o[j].w = __builtin_amdgcn_perm(im[2 * j + 1], re[2 * j + 1], 0x07030602u);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) dst[4 * b + j] = o[j];
  }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static inline unsigned grid_for(size_t n, unsigned block) {
  size_t g = (n + block - 1) / block;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return (unsigned)g;
}

extern "C" int xgh_launch_xengine(int family, const XgLaunch *L, void *stream) {
  const unsigned grid = (unsigned)(XG_NXCD * L->wg_per_xcd);
  if (grid == 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  switch (family) {
    case XG_KERNEL_I8_MFMA:
      if (L->prefetch)
        hipLaunchKernelGGL((xg_xengine_mfma<false, true>), dim3(grid), dim3(XG_WG_THREADS), 0, s, *L);
      else
        hipLaunchKernelGGL((xg_xengine_mfma<false, false>), dim3(grid), dim3(XG_WG_THREADS), 0, s, *L);
      break;
    case XG_KERNEL_F16_MFMA:
      if (L->prefetch)
        hipLaunchKernelGGL((xg_xengine_mfma<true, true>), dim3(grid), dim3(XG_WG_THREADS), 0, s, *L);
      else
        hipLaunchKernelGGL((xg_xengine_mfma<true, false>), dim3(grid), dim3(XG_WG_THREADS), 0, s, *L);
      break;
    case XG_KERNEL_F32_PLAIN:
      hipLaunchKernelGGL(xg_xengine_f32, dim3(grid), dim3(XG_WG_THREADS), 0, s, *L);
      break;
    case XG_KERNEL_I8_RG:
      if (L->stage == 3)
        hipLaunchKernelGGL((xg_xengine_rgd<64, 4>), dim3(grid), dim3(XG_WG_THREADS), 0, s, *L);
      else if (L->stage == 4)
        hipLaunchKernelGGL((xg_xengine_rgd<32, 8>), dim3(grid), dim3(XG_WG_THREADS), 0, s, *L);
      else if (L->stage == 5)
        hipLaunchKernelGGL((xg_xengine_rgd16<4>), dim3(grid), dim3(XG_WG_THREADS), 0, s, *L);
      else if (L->stage == 1)
        hipLaunchKernelGGL(xg_xengine_rg<128>, dim3(grid), dim3(XG_WG_THREADS), 0, s, *L);
      else
        hipLaunchKernelGGL(xg_xengine_rg<64>, dim3(grid), dim3(XG_WG_THREADS), 0, s, *L);
      break;
    case XG_KERNEL_I8_KG:
      hipLaunchKernelGGL(xg_xengine_kg, dim3(grid), dim3(XG_WG_THREADS), 0, s, *L);
      break;
    case XG_KERNEL_F16_RG:
      if (L->stage == 3)
        hipLaunchKernelGGL((xg_xengine_rgdh<16, 8>), dim3(grid), dim3(XG_WG_THREADS), 0, s, *L);
      else if (L->stage == 2)
        hipLaunchKernelGGL((xg_xengine_rgdh<32, 4>), dim3(grid), dim3(XG_WG_THREADS), 0, s, *L);
      else if (L->stage == 1)
        hipLaunchKernelGGL(xg_xengine_rgh<64>, dim3(grid), dim3(XG_WG_THREADS), 0, s, *L);
      else
        hipLaunchKernelGGL(xg_xengine_rgh<32>, dim3(grid), dim3(XG_WG_THREADS), 0, s, *L);
      break;
    default:
      return xgh_report(hipErrorInvalidValue, "xgh_launch_xengine(family)");
  }
  return xgh_report(hipGetLastError(), "xg_xengine launch");
}

extern "C" int xgh_launch_fill_zero(void *ptr, size_t bytes, void *stream) {
  if (!bytes) return 0;
  const size_t n16 = bytes / 16;
  const size_t ntail = (bytes - n16 * 16) / 4;
  unsigned *tail = (unsigned *)((char *)ptr + n16 * 16);
  hipLaunchKernelGGL(xg_fill_zero_k, dim3(grid_for(n16 > ntail ? n16 : ntail, 256)), dim3(256), 0,
                     (hipStream_t)stream, (uint4 *)ptr, n16, tail, ntail);
  return xgh_report(hipGetLastError(), "xg_fill_zero launch");
}

extern "C" int xgh_launch_copy(void *dst, const void *src, size_t bytes, void *stream) {
  if (!bytes) return 0;
  if ((((uintptr_t)dst | (uintptr_t)src) & 15) || (bytes & 3))
    return xgh_report(hipErrorInvalidValue, "xgh_launch_copy(alignment)");
  const size_t n16 = bytes / 16, ntail = (bytes - n16 * 16) / 4;
  hipLaunchKernelGGL(xg_copy_k, dim3(grid_for(n16 > ntail ? n16 : ntail, 256)), dim3(256), 0,
                     (hipStream_t)stream, (uint4 *)dst, (const uint4 *)src, n16,
                     (unsigned *)((char *)dst + n16 * 16),
                     (const unsigned *)((const char *)src + n16 * 16), ntail);
  return xgh_report(hipGetLastError(), "xg_copy launch");
}

template <typename D, typename S, bool ADD>
static void launch_spill(D *dst, const S *src, size_t n, hipStream_t s) {
  if ((((uintptr_t)dst | (uintptr_t)src) & 15) == 0 && (n & 3) == 0)
    hipLaunchKernelGGL((xg_spill_k<D, S, ADD, 4>), dim3(grid_for(n / 4, 256)), dim3(256), 0, s,
                       dst, src, n / 4);
  else
    hipLaunchKernelGGL((xg_spill_k<D, S, ADD, 1>), dim3(grid_for(n, 256)), dim3(256), 0, s, dst,
                       src, n);
}

template <typename W, typename S, typename O>
static void launch_combine(O *dst, const W *spill, const S *cur, size_t n, hipStream_t s) {
  if ((((uintptr_t)dst | (uintptr_t)spill | (uintptr_t)cur) & 15) == 0 && (n & 3) == 0)
    hipLaunchKernelGGL((xg_combine_k<W, S, O, 4>), dim3(grid_for(n / 4, 256)), dim3(256), 0, s,
                       dst, spill, cur, n / 4);
  else
    hipLaunchKernelGGL((xg_combine_k<W, S, O, 1>), dim3(grid_for(n, 256)), dim3(256), 0, s, dst,
                       spill, cur, n);
}

extern "C" int xgh_launch_spill(int kind, int add, void *dst, const void *src, size_t n,
                                void *stream) {
  if (!n) return 0;
  hipStream_t s = (hipStream_t)stream;
  if (kind == XG_OUT_I64 && add)
    launch_spill<long long, int, true>((long long *)dst, (const int *)src, n, s);
  else if (kind == XG_OUT_I64)
    launch_spill<long long, int, false>((long long *)dst, (const int *)src, n, s);
  else if (kind == XG_OUT_F64 && add)
    launch_spill<double, float, true>((double *)dst, (const float *)src, n, s);
  else if (kind == XG_OUT_F64)
    launch_spill<double, float, false>((double *)dst, (const float *)src, n, s);
  else
    return xgh_report(hipErrorInvalidValue, "xgh_launch_spill(kind)");
  return xgh_report(hipGetLastError(), "xg_spill launch");
}

extern "C" int xgh_launch_combine(int kind, int narrow, void *dst, const void *spill,
                                  const void *cur, size_t n, void *stream) {
  if (!n) return 0;
  hipStream_t s = (hipStream_t)stream;
  if (kind == XG_OUT_I64) {
    if (narrow)  // low 32 bits: what wrapping int32 accumulation would hold
      launch_combine((unsigned *)dst, (const unsigned long long *)spill, (const int *)cur, n, s);
    else
      launch_combine((long long *)dst, (const long long *)spill, (const int *)cur, n, s);
  } else if (kind == XG_OUT_F64) {
    if (narrow)
      launch_combine((float *)dst, (const double *)spill, (const float *)cur, n, s);
    else
      launch_combine((double *)dst, (const double *)spill, (const float *)cur, n, s);
  } else {
    return xgh_report(hipErrorInvalidValue, "xgh_launch_combine(kind)");
  }
  return xgh_report(hipGetLastError(), "xg_combine launch");
}

extern "C" int xgh_launch_fill_uniform_i8(void *ptr, size_t n, uint64_t seed, int lo, int hi,
                                          void *stream) {
  if (!n) return 0;
  if (hi < lo || lo < -128 || hi > 127)
    return xgh_report(hipErrorInvalidValue, "xgh_launch_fill_uniform_i8(range)");
  const size_t nwords = n / 4, ntail = n - nwords * 4;
  // key = mix64(seed); the same definition is used by the host generator
  uint64_t x = seed + 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  const uint64_t key = x ^ (x >> 31);
  hipLaunchKernelGGL(xg_fill_uniform_k, dim3(grid_for(nwords > ntail ? nwords : ntail, 256)),
                     dim3(256), 0, (hipStream_t)stream, (uint32_t *)ptr, nwords, key, lo,
                     (uint64_t)(hi - lo + 1), (unsigned char *)ptr + nwords * 4, ntail, nwords * 4);
  return xgh_report(hipGetLastError(), "xg_fill_uniform launch");
}

extern "C" int xgh_launch_reorder(int src_order, int elem_bytes, void *dst, const void *src,
                                  int nstation, int nfreq, long long plane, void *stream) {
  const long long total = (long long)nfreq * nstation * (nstation + 1) / 2 * 4;
  if (total <= 0) return 0;
  if (src_order != XG_ORDER_REGTILE && src_order != XG_ORDER_REALIMAG)
    return xgh_report(hipErrorInvalidValue, "xgh_launch_reorder(order)");
  const dim3 g(grid_for((size_t)total, 256)), b(256);
  const bool cells = elem_bytes == 4 && !((uintptr_t)dst & 15) && !((uintptr_t)src & 15) &&
                     (plane & 3) == 0;
  if (cells)
    hipLaunchKernelGGL(xg_reorder_cells_k, dim3(grid_for((size_t)(total / 2), 256)), b, 0,
                       (hipStream_t)stream, src_order, (xg_u4 *)dst, (const unsigned *)src,
                       nstation, nfreq, plane);
  else if (elem_bytes == 4)
    hipLaunchKernelGGL((xg_reorder_k<unsigned>), g, b, 0, (hipStream_t)stream, src_order,
                       (unsigned *)dst, (const unsigned *)src, nstation, nfreq, plane);
  else if (elem_bytes == 8 && !((uintptr_t)dst & 15))
    hipLaunchKernelGGL((xg_reorder_k<unsigned long long, true>), g, b, 0, (hipStream_t)stream,
                       src_order, (unsigned long long *)dst, (const unsigned long long *)src,
                       nstation, nfreq, plane);
  else if (elem_bytes == 8)
    hipLaunchKernelGGL((xg_reorder_k<unsigned long long>), g, b, 0, (hipStream_t)stream, src_order,
                       (unsigned long long *)dst, (const unsigned long long *)src, nstation, nfreq,
                       plane);
  else
    return xgh_report(hipErrorInvalidValue, "xgh_launch_reorder(elem)");
  return xgh_report(hipGetLastError(), "xg_reorder launch");
}

template <typename W, typename S, typename O>
static void reorder_combine_as(int src_order, void *dst, const void *spill, const void *cur,
                               int nstation, int nfreq, long long plane, dim3 g, hipStream_t st) {
  if (!((uintptr_t)dst & (2 * sizeof(O) - 1)))
    hipLaunchKernelGGL((xg_reorder_combine_k<W, S, O, true>), g, dim3(256), 0, st, src_order,
                       (O *)dst, (const W *)spill, (const S *)cur, nstation, nfreq, plane);
  else
    hipLaunchKernelGGL((xg_reorder_combine_k<W, S, O, false>), g, dim3(256), 0, st, src_order,
                       (O *)dst, (const W *)spill, (const S *)cur, nstation, nfreq, plane);
}

extern "C" int xgh_launch_reorder_combine(int src_order, int kind, int narrow, void *dst,
                                          const void *spill, const void *cur, int nstation,
                                          int nfreq, long long plane, void *stream) {
  const long long total = (long long)nfreq * nstation * (nstation + 1) / 2 * 4;
  if (total <= 0) return 0;
  if (src_order != XG_ORDER_REGTILE && src_order != XG_ORDER_REALIMAG)
    return xgh_report(hipErrorInvalidValue, "xgh_launch_reorder_combine(order)");
  const dim3 g(grid_for((size_t)total, 256));
  hipStream_t st = (hipStream_t)stream;
  if (kind == XG_OUT_I64 && narrow)
    reorder_combine_as<long long, int, int>(src_order, dst, spill, cur, nstation, nfreq, plane, g, st);
  else if (kind == XG_OUT_I64)
    reorder_combine_as<long long, int, long long>(src_order, dst, spill, cur, nstation, nfreq,
                                                  plane, g, st);
  else if (kind == XG_OUT_F64 && narrow)
    reorder_combine_as<double, float, float>(src_order, dst, spill, cur, nstation, nfreq, plane,
                                             g, st);
  else if (kind == XG_OUT_F64)
    reorder_combine_as<double, float, double>(src_order, dst, spill, cur, nstation, nfreq, plane,
                                              g, st);
  else
    return xgh_report(hipErrorInvalidValue, "xgh_launch_reorder_combine(kind)");
  return xgh_report(hipGetLastError(), "xg_reorder_combine launch");
}

extern "C" int xgh_launch_swizzle(int unswizzle, void *dst, const void *src, int nstation,
                                  int nfreq, int ntime, void *stream) {
  if (ntime % 4) return xgh_report(hipErrorInvalidValue, "xgh_launch_swizzle(ntime % 4)");
  const void *sw = unswizzle ? src : dst, *pl = unswizzle ? dst : src;
  if (((uintptr_t)sw & 15) || ((uintptr_t)pl & 3))
    return xgh_report(hipErrorInvalidValue, "xgh_launch_swizzle(alignment)");
  const long long row4 = (long long)nfreq * nstation;  // 4-byte words per sample row
  const long long total = (long long)(ntime / 4) * row4;
  if (total == 0) return 0;
  if (unswizzle)
    hipLaunchKernelGGL(xg_swizzle_k<true>, dim3(grid_for((size_t)total, 256)), dim3(256), 0,
                       (hipStream_t)stream, (unsigned char *)dst, (const unsigned char *)src, row4,
                       total);
  else
    hipLaunchKernelGGL(xg_swizzle_k<false>, dim3(grid_for((size_t)total, 256)), dim3(256), 0,
                       (hipStream_t)stream, (unsigned char *)dst, (const unsigned char *)src, row4,
                       total);
  return xgh_report(hipGetLastError(), "xg_swizzle launch");
}

extern "C" int xgh_launch_extract(int kind, void *dst, const void *packed, int nstation, int nfreq,
                                  void *stream) {
  const long long nt = (nstation + XG_XT - 1) / XG_XT;
  const long long nwork = nt * (nt + 1) / 2 * nfreq;  // tiles on or below the diagonal
  if (nwork == 0) return 0;
  const dim3 g((unsigned)(nwork < 65536 ? nwork : 65536)), b(256);
  hipStream_t s = (hipStream_t)stream;
  switch (kind) {
    case XG_OUT_I32:
      hipLaunchKernelGGL(xg_extract_tile_k<int>, g, b, 0, s, (int *)dst, (const int *)packed,
                         nstation, nfreq);
      break;
    case XG_OUT_I64:
      hipLaunchKernelGGL(xg_extract_tile_k<long long>, g, b, 0, s, (long long *)dst,
                         (const long long *)packed, nstation, nfreq);
      break;
    case XG_OUT_F32:
      hipLaunchKernelGGL(xg_extract_tile_k<float>, g, b, 0, s, (float *)dst, (const float *)packed,
                         nstation, nfreq);
      break;
    case XG_OUT_F64:
      hipLaunchKernelGGL(xg_extract_tile_k<double>, g, b, 0, s, (double *)dst,
                         (const double *)packed, nstation, nfreq);
      break;
    default:
      return xgh_report(hipErrorInvalidValue, "xgh_launch_extract(kind)");
  }
  return xgh_report(hipGetLastError(), "xg_extract launch");
}

extern "C" int xgh_launch_unblock32(void *dst, const void *src, size_t bytes, void *stream) {
  if ((bytes % 64) || (((uintptr_t)dst | (uintptr_t)src) & 15))
    return xgh_report(hipErrorInvalidValue, "xgh_launch_unblock32(alignment)");
  const long long nblk = (long long)(bytes / 64);
  if (nblk == 0) return 0;
  hipLaunchKernelGGL(xg_unblock32_k, dim3(grid_for((size_t)nblk, 256)), dim3(256), 0,
                     (hipStream_t)stream, (uint4 *)dst, (const uint4 *)src, nblk);
  return xgh_report(hipGetLastError(), "xg_unblock32 launch");
}

// ---------------------------------------------------------------------------
// MFMA-only throughput probe (measurement): 8 independent int8 32x32x32
// accumulator chains per wave, no memory traffic in the loop.  It gives the
// on-board ceiling against which the X-engine's MFMA rate is compared.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void xg_mfma_peak_k(int iters, int *out) {
  const v4i a = {(int)threadIdx.x, 1, 2, 3}, b = {3, (int)blockIdx.x, 1, 0};
  v16i c[8];
#pragma unroll
  for (int j = 0; j < 8; ++j)
#pragma unroll
    for (int e = 0; e < 16; ++e) c[j][e] = 0;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int j = 0; j < 8; ++j) c[j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, c[j], 0, 0, 0);
  }
  int s = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j)
#pragma unroll
    for (int e = 0; e < 16; ++e) s += c[j][e];
  out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}

// The same MFMA stream with operands of a chosen bit activity: each lane
// holds 8 A and 8 B operands from `ops` (16 x 256 x 16 bytes), and MFMA j of
// step r uses A[(r + j) & 7] and B[j], so the operands change every
// instruction.  One workgroup per CU, one wave per SIMD, as the X-engine.
// The chip's power management sets the clock by what the datapath toggles,
// so this is the MFMA-only ceiling for input of that distribution.
// F16: v_mfma_f32_32x32x16_f16 on the same 16 bytes per lane (8 halves).
template <bool F16>
__global__ __attribute__((amdgpu_flat_work_group_size(256, 256), amdgpu_waves_per_eu(1, 1)))
void xg_mfma_peak_data_k(const v4i *ops, int iters, int *out) {
  const int t = threadIdx.x;
  v4i A[8], B[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    A[j] = ops[j * 256 + t];
    B[j] = ops[(8 + j) * 256 + t];
  }
  if constexpr (F16) {
    v16f c[8];
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) c[j][e] = 0.f;
    for (int i = 0; i < iters; i += 8) {
#pragma unroll
      for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int j = 0; j < 8; ++j)
          c[j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(v8h, A[(r + j) & 7]),
                                                        __builtin_bit_cast(v8h, B[j]), c[j], 0, 0, 0);
    }
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) s += c[j][e];
    out[(size_t)blockIdx.x * 256 + t] = __float_as_int(s);
  } else {
    v16i c[8];
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) c[j][e] = 0;
    for (int i = 0; i < iters; i += 8) {
#pragma unroll
      for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int j = 0; j < 8; ++j)
          c[j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[(r + j) & 7], B[j], c[j], 0, 0, 0);
    }
    int s = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) s += c[j][e];
    out[(size_t)blockIdx.x * 256 + t] = s;
  }
}

// iters: a multiple of 8; ops: XG_PEAK_OPERAND_BYTES of device memory
extern "C" int xgh_launch_mfma_peak_data(int f16, int grid, int iters, const void *ops, int *out,
                                         void *stream) {
  if (grid <= 0 || iters <= 0) return 0;
  if (f16)
    hipLaunchKernelGGL(xg_mfma_peak_data_k<true>, dim3(grid), dim3(256), 0, (hipStream_t)stream,
                       (const v4i *)ops, iters, out);
  else
    hipLaunchKernelGGL(xg_mfma_peak_data_k<false>, dim3(grid), dim3(256), 0, (hipStream_t)stream,
                       (const v4i *)ops, iters, out);
  return xgh_report(hipGetLastError(), "xg_mfma_peak_data_k launch");
}

extern "C" int xgh_launch_mfma_peak(int grid, int iters, int *out, void *stream) {
  if (grid <= 0 || iters <= 0) return 0;
  hipLaunchKernelGGL(xg_mfma_peak_k, dim3(grid), dim3(256), 0, (hipStream_t)stream, iters, out);
  return xgh_report(hipGetLastError(), "xg_mfma_peak_k launch");
}

// ---------------------------------------------------------------------------
// Static device scratch for the entry point that allocates nothing
// (xgpuAmdCorrelateNoAlloc): the input-layout conversion buffer lives in this
// file-scope array, one copy per device, held until the process ends.  Host
// code never uses the array's name as a device address: its device address is
// looked up once per device with hipGetSymbolAddress (with that device
// current) and kept here.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[XG_STATIC_SCRATCH_BYTES];

#define XG_MAX_DEVICES 64
static void *g_scratch_ptr[XG_MAX_DEVICES];
static pthread_mutex_t g_scratch_mu = PTHREAD_MUTEX_INITIALIZER;

extern "C" int xgh_static_scratch(int device, void **ptr) {
  if (device < 0 || device >= XG_MAX_DEVICES)
    return xgh_report(hipErrorInvalidDevice, "xgh_static_scratch(device)");
  pthread_mutex_lock(&g_scratch_mu);
  int rc = 0;
  if (!g_scratch_ptr[device]) {
    int cur = 0;
    rc = xgh_report(hipGetDevice(&cur), "hipGetDevice");
    if (!rc) rc = xgh_report(hipSetDevice(device), "hipSetDevice");
    if (!rc) {
      void *p = nullptr;
      rc = xgh_report(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
      if (!rc) g_scratch_ptr[device] = p;
      const int rc2 = xgh_report(hipSetDevice(cur), "hipSetDevice");
      if (!rc) rc = rc2;
    }
  }
  *ptr = g_scratch_ptr[device];
  pthread_mutex_unlock(&g_scratch_mu);
  return rc;
}
