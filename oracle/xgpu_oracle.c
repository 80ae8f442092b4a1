/*
 * xgpu_oracle.c -- CPU ORACLE for the X-engine.  TEST INFRASTRUCTURE ONLY
 * (see xgpu_oracle.h for who may use it and how each function is pinned).
 *
 * Build: oracle/Makefile, gcc -O2 -fopenmp -ffp-contract=off (no FMA
 * contraction, so float arithmetic rounds exactly where the reference's C
 * rounds).
 */
#include "xgpu_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static void set_threads(int nthreads) {
#ifdef _OPENMP
  omp_set_num_threads(nthreads > 0 ? nthreads : omp_get_num_procs());
#else
  (void)nthreads;
#endif
}

/* baseline k -> (station1 >= station2); reference src/omp_xengine.c:48-49
 * computes the same pair with a square root */
static void tri_pair(long long k, int *s1, int *s2) {
  int i = (int)(-0.5 + sqrt(0.25 + 2.0 * (double)k));
  while ((long long)i * (i + 1) / 2 > k) --i;
  while ((long long)(i + 1) * (i + 2) / 2 <= k) ++i;
  *s1 = i;
  *s2 = (int)(k - (long long)i * (i + 1) / 2);
}

static inline size_t in_index(int t, int f, int s, int p, int ns, int nf) {
  /* reference src/omp_xengine.c:57-60, in bytes of the re/im pair */
  return ((((size_t)t * nf + f) * ns + s) * 2 + p) * 2;
}

void oracle_xengine_f32(float *out, const int8_t *in, int ns, int nf, int nt, int nthreads) {
  const long long nb = (long long)(ns + 1) * (ns / 2);
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic)
  for (long long x = 0; x < (long long)nf * nb; ++x) {
    const int f = (int)(x / nb);
    const long long k = x - (long long)f * nb;
    int s1, s2;
    tri_pair(k, &s1, &s2);
    float acc[4][2] = {{0}};
    for (int t = 0; t < nt; ++t) {
      for (int p1 = 0; p1 < 2; ++p1)
        for (int p2 = 0; p2 < 2; ++p2) {
          const int8_t *a = in + in_index(t, f, s1, p1, ns, nf);
          const int8_t *b = in + in_index(t, f, s2, p2, ns, nf);
          /* cxmac, reference src/omp_xengine.c:24-28 */
          acc[p1 * 2 + p2][0] += (float)a[0] * (float)b[0] + (float)a[1] * (float)b[1];
          acc[p1 * 2 + p2][1] += (float)a[1] * (float)b[0] - (float)a[0] * (float)b[1];
        }
    }
    for (int p = 0; p < 4; ++p) {
      out[(x * 4 + p) * 2] = acc[p][0];
      out[(x * 4 + p) * 2 + 1] = acc[p][1];
    }
  }
}

void oracle_xengine_dp4a_f32acc(int32_t *out, const int8_t *in, int ns, int nf, int nt,
                                int nthreads) {
  const long long nb = (long long)(ns + 1) * (ns / 2);
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic)
  for (long long x = 0; x < (long long)nf * nb; ++x) {
    const int f = (int)(x / nb);
    const long long k = x - (long long)f * nb;
    int s1, s2;
    tri_pair(k, &s1, &s2);
    int acc[4][2] = {{0}};
    for (int t = 0; t < nt; ++t) {
      for (int p1 = 0; p1 < 2; ++p1)
        for (int p2 = 0; p2 < 2; ++p2) {
          const int8_t *a = in + in_index(t, f, s1, p1, ns, nf);
          const int8_t *b = in + in_index(t, f, s2, p2, ns, nf);
          /* int lvalue += float expression: the sum is formed in float and
           * converted back to int */
          int *r = &acc[p1 * 2 + p2][0];
          r[0] = (int)((float)r[0] + ((float)a[0] * (float)b[0] + (float)a[1] * (float)b[1]));
          r[1] = (int)((float)r[1] + ((float)a[1] * (float)b[0] - (float)a[0] * (float)b[1]));
        }
    }
    for (int p = 0; p < 4; ++p) {
      out[(x * 4 + p) * 2] = acc[p][0];
      out[(x * 4 + p) * 2 + 1] = acc[p][1];
    }
  }
}

/* One channel gathered input-major in int16 planes: row m = 2s+p holds the
 * nt real parts, then the nt imaginary parts. */
static int16_t *gather_channel(const int8_t *in, int f, int ns, int nf, int nt) {
  const size_t nin = (size_t)ns * 2;
  int16_t *ch = (int16_t *)malloc(nin * nt * 2 * sizeof(int16_t));
  if (!ch) return NULL;
  for (int t = 0; t < nt; ++t) {
    const int8_t *row = in + in_index(t, f, 0, 0, ns, nf);
    for (size_t m = 0; m < nin; ++m) {
      ch[m * 2 * nt + t] = row[m * 2];
      ch[m * 2 * nt + nt + t] = row[m * 2 + 1];
    }
  }
  return ch;
}

/* Exact sums of one pair.  Each sample adds at most 2 * 128 * 128 = 32768 in
 * magnitude, so int32 partial sums over 16384 samples cannot overflow; they
 * are added into int64. */
#define EXACT_CHUNK 16384
static void exact_pair(int64_t *o, const int16_t *ch, int nt, int s1, int s2) {
  for (int p1 = 0; p1 < 2; ++p1)
    for (int p2 = 0; p2 < 2; ++p2) {
      const int16_t *ar = ch + (size_t)(2 * s1 + p1) * 2 * nt, *ai = ar + nt;
      const int16_t *br = ch + (size_t)(2 * s2 + p2) * 2 * nt, *bi = br + nt;
      int64_t re = 0, im = 0;
      for (int t0 = 0; t0 < nt; t0 += EXACT_CHUNK) {
        const int t1 = t0 + EXACT_CHUNK < nt ? t0 + EXACT_CHUNK : nt;
        int32_t r = 0, i = 0;
        for (int t = t0; t < t1; ++t) {
          r += (int32_t)ar[t] * br[t] + (int32_t)ai[t] * bi[t];
          i += (int32_t)ai[t] * br[t] - (int32_t)ar[t] * bi[t];
        }
        re += r;
        im += i;
      }
      o[(p1 * 2 + p2) * 2] = re;
      o[(p1 * 2 + p2) * 2 + 1] = im;
    }
}

void oracle_xengine_i64(int64_t *out, const int8_t *in, int ns, int nf, int nt, int nthreads) {
  const long long nb = (long long)(ns + 1) * (ns / 2);
  set_threads(nthreads);
  for (int f = 0; f < nf; ++f) {
    int16_t *ch = gather_channel(in, f, ns, nf, nt);
    if (!ch) return;
#pragma omp parallel for schedule(dynamic, 8)
    for (long long k = 0; k < nb; ++k) {
      int s1, s2;
      tri_pair(k, &s1, &s2);
      exact_pair(out + ((size_t)f * nb + k) * 8, ch, nt, s1, s2);
    }
    free(ch);
  }
}

/* Every visibility of the listed channels: out[c] is channel chans[c]'s
 * triangle (as oracle_xengine_i64 lays out one channel). */
void oracle_xengine_i64_channels(int64_t *out, const int8_t *in, int ns, int nf, int nt,
                                 const int32_t *chans, int nchan, int nthreads) {
  const long long nb = (long long)(ns + 1) * (ns / 2);
  set_threads(nthreads);
  for (int c = 0; c < nchan; ++c) {
    int16_t *ch = gather_channel(in, chans[c], ns, nf, nt);
    if (!ch) return;
#pragma omp parallel for schedule(dynamic, 8)
    for (long long k = 0; k < nb; ++k) {
      int s1, s2;
      tri_pair(k, &s1, &s2);
      exact_pair(out + ((size_t)c * nb + k) * 8, ch, nt, s1, s2);
    }
    free(ch);
  }
}

void oracle_xengine_i64_pairs(int64_t *out, const int8_t *in, int ns, int nf, int nt,
                              const int32_t *fij, int npairs, int nthreads) {
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
  for (int n = 0; n < npairs; ++n) {
    const int f = fij[3 * n], s1 = fij[3 * n + 1], s2 = fij[3 * n + 2];
    for (int p1 = 0; p1 < 2; ++p1)
      for (int p2 = 0; p2 < 2; ++p2) {
        int64_t re = 0, im = 0;
        for (int t = 0; t < nt; ++t) {
          const int8_t *a = in + in_index(t, f, s1, p1, ns, nf);
          const int8_t *b = in + in_index(t, f, s2, p2, ns, nf);
          re += (int32_t)a[0] * b[0] + (int32_t)a[1] * b[1];
          im += (int32_t)a[1] * b[0] - (int32_t)a[0] * b[1];
        }
        out[(n * 4 + p1 * 2 + p2) * 2] = re;
        out[(n * 4 + p1 * 2 + p2) * 2 + 1] = im;
      }
  }
}

/* IEEE binary16 -> double, by the definition of the format */
static double half_to_double(uint16_t h) {
  const int sign = (h >> 15) & 1, ex = (h >> 10) & 0x1f, man = h & 0x3ff;
  double v;
  if (ex == 0)
    v = ldexp((double)man, -24);
  else if (ex == 31)
    v = man ? NAN : INFINITY;
  else
    v = ldexp((double)(man | 0x400), ex - 25);
  return sign ? -v : v;
}

void oracle_xengine_f64_f16(double *out, const uint16_t *in, int ns, int nf, int nt,
                            int nthreads) {
  const long long nb = (long long)(ns + 1) * (ns / 2);
  const size_t nin = (size_t)ns * 2;
  set_threads(nthreads);
  for (int f = 0; f < nf; ++f) {
    double *ch = (double *)malloc(nin * nt * 2 * sizeof(double));
    if (!ch) return;
    for (int t = 0; t < nt; ++t)
      for (size_t m = 0; m < nin; ++m) {
        const uint16_t *v = in + ((((size_t)t * nf + f) * nin) + m) * 2;
        ch[(m * nt + t) * 2] = half_to_double(v[0]);
        ch[(m * nt + t) * 2 + 1] = half_to_double(v[1]);
      }
#pragma omp parallel for schedule(dynamic, 8)
    for (long long k = 0; k < nb; ++k) {
      int s1, s2;
      tri_pair(k, &s1, &s2);
      for (int p1 = 0; p1 < 2; ++p1)
        for (int p2 = 0; p2 < 2; ++p2) {
          const double *a = ch + (size_t)(2 * s1 + p1) * nt * 2;
          const double *b = ch + (size_t)(2 * s2 + p2) * nt * 2;
          double re = 0, im = 0;
          for (int t = 0; t < nt; ++t) {
            re += a[2 * t] * b[2 * t] + a[2 * t + 1] * b[2 * t + 1];
            im += a[2 * t + 1] * b[2 * t] - a[2 * t] * b[2 * t + 1];
          }
          out[(((size_t)f * nb + k) * 4 + p1 * 2 + p2) * 2] = re;
          out[(((size_t)f * nb + k) * 4 + p1 * 2 + p2) * 2 + 1] = im;
        }
    }
    free(ch);
  }
}

/* register-tile location of triangular visibility (f, i>=j) */
static long long regtile_l(int f, int i, int j, int ns) {
  const long long P = ns / 2, RT = P * (P + 1) / 2;
  const int R = i / 2, C = j / 2, q = 2 * (j & 1) + (i & 1);
  return ((long long)f * 4 + q) * RT + (long long)R * (R + 1) / 2 + C;
}

void oracle_reorder(int order, int elem_bytes, void *tri, const void *src, int ns, int nf,
                    long long plane) {
  const long long nb = (long long)ns * (ns + 1) / 2;
  for (int f = 0; f < nf; ++f)
    for (int i = 0; i < ns; ++i)
      for (int j = 0; j <= i; ++j) {
        const long long k = (long long)f * nb + (long long)i * (i + 1) / 2 + j;
        const long long l = (order == 3000) ? regtile_l(f, i, j, ns) : k;
        for (int p = 0; p < 4; ++p) {
          if (elem_bytes == 4) {
            ((uint32_t *)tri)[(k * 4 + p) * 2] = ((const uint32_t *)src)[l * 4 + p];
            ((uint32_t *)tri)[(k * 4 + p) * 2 + 1] = ((const uint32_t *)src)[l * 4 + p + plane];
          } else {
            ((uint64_t *)tri)[(k * 4 + p) * 2] = ((const uint64_t *)src)[l * 4 + p];
            ((uint64_t *)tri)[(k * 4 + p) * 2 + 1] = ((const uint64_t *)src)[l * 4 + p + plane];
          }
        }
      }
}

void oracle_regtile_from_tri(int elem_bytes, void *dst, const void *tri, int ns, int nf,
                             long long plane) {
  memset(dst, 0, (size_t)plane * 2 * elem_bytes);
  const long long nb = (long long)ns * (ns + 1) / 2;
  for (int f = 0; f < nf; ++f)
    for (int i = 0; i < ns; ++i)
      for (int j = 0; j <= i; ++j) {
        const long long k = (long long)f * nb + (long long)i * (i + 1) / 2 + j;
        const long long l = regtile_l(f, i, j, ns);
        for (int p = 0; p < 4; ++p) {
          if (elem_bytes == 4) {
            ((uint32_t *)dst)[l * 4 + p] = ((const uint32_t *)tri)[(k * 4 + p) * 2];
            ((uint32_t *)dst)[l * 4 + p + plane] = ((const uint32_t *)tri)[(k * 4 + p) * 2 + 1];
          } else {
            ((uint64_t *)dst)[l * 4 + p] = ((const uint64_t *)tri)[(k * 4 + p) * 2];
            ((uint64_t *)dst)[l * 4 + p + plane] = ((const uint64_t *)tri)[(k * 4 + p) * 2 + 1];
          }
        }
      }
}

void oracle_swizzle(int8_t *out, const int8_t *in, int ns, int nf, int nt) {
  for (int t = 0; t < nt; ++t)
    for (int f = 0; f < nf; ++f)
      for (int s = 0; s < ns; ++s)
        for (int p = 0; p < 2; ++p)
          for (int c = 0; c < 2; ++c)
            out[((((size_t)(t / 4) * nf + f) * ns + s) * 2 + p) * 2 * 4 + (size_t)c * 4 + t % 4] =
                in[((((size_t)t * nf + f) * ns + s) * 2 + p) * 2 + c];
}

void oracle_unblock32(int8_t *out, const int8_t *in, size_t nbytes) {
  for (size_t b = 0; b < nbytes; b += 64)
    for (int q = 0; q < 32; ++q) {
      out[b + 2 * q] = in[b + q];
      out[b + 2 * q + 1] = in[b + 32 + q];
    }
}

void oracle_extract(float *full, const float *tri, int ns, int nf) {
  const long long nb = (long long)ns * (ns + 1) / 2;
  for (int f = 0; f < nf; ++f)
    for (int i = 0; i < ns; ++i)
      for (int j = 0; j <= i; ++j) {
        const long long k = (long long)f * nb + (long long)i * (i + 1) / 2 + j;
        for (int p1 = 0; p1 < 2; ++p1)
          for (int p2 = 0; p2 < 2; ++p2) {
            const float re = tri[(k * 4 + p1 * 2 + p2) * 2];
            const float im = tri[(k * 4 + p1 * 2 + p2) * 2 + 1];
            const size_t lo = ((((size_t)f * ns + i) * ns + j) * 2 + p1) * 2 + p2;
            const size_t up = ((((size_t)f * ns + j) * ns + i) * 2 + p2) * 2 + p1;
            full[lo * 2] = re;
            full[lo * 2 + 1] = im;
            full[up * 2] = re;
            full[up * 2 + 1] = -im;
          }
      }
}

void oracle_random_legacy(int8_t *out, size_t ncomplex, unsigned seed) {
  srand(seed);
  for (size_t n = 0; n < ncomplex; ++n) {
    double u1 = rand() / (double)RAND_MAX;
    double u2 = rand() / (double)RAND_MAX;
    if (u1 == 0.0) u1 = 0.5 / RAND_MAX;
    if (u2 == 0.0) u2 = 0.5 / RAND_MAX;
    const double r = 2.5 * sqrt(-2.0 * log(u1));
    const double th = 2 * M_PI * u2;
    double a = round(r * cos(th)), b = round(r * sin(th));
    if (a > 7.0) a = 7.0;
    if (a < -7.0) a = -7.0;
    if (b > 7.0) b = 7.0;
    if (b < -7.0) b = -7.0;
    out[2 * n] = (int8_t)(((int)a) << 4);
    out[2 * n + 1] = (int8_t)(((int)b) << 4);
  }
}

static uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

void oracle_fill_uniform(int8_t *out, size_t n, uint64_t seed, int lo, int hi) {
  const uint64_t key = mix64(seed), span = (uint64_t)(hi - lo + 1);
  for (size_t i = 0; i < n; ++i)
    out[i] = (int8_t)(lo + (int)(((mix64(key + i) >> 32) * span) >> 32));
}

/* zabs of the reference: the squares and their sum are float arithmetic */
static double zabs_f(float re, float im) { return sqrt(re * re + im * im); }

int oracle_check(const float *gpu, const float *cpu, int ns, int nf, double tol, int round_gpu,
                 double *max_err) {
  const long long nb = (long long)(ns + 1) * (ns / 2);
  int bad = 0;
  double worst = 0.0;
  for (int i = 0; i < ns; ++i)
    for (int j = 0; j <= i; ++j)
      for (int p = 0; p < 4; ++p)
        for (int f = 0; f < nf; ++f) {
          const long long idx = ((long long)f * nb + (long long)i * (i + 1) / 2 + j) * 4 + p;
          float gr = gpu[2 * idx], gi = gpu[2 * idx + 1];
          const float cr = cpu[2 * idx], ci = cpu[2 * idx + 1];
          if (round_gpu) {
            gr = (float)round(gr);
            gi = (float)round(gi);
          }
          double err;
          if (zabs_f(cr, ci) == 0)
            err = zabs_f(gr, gi);
          else
            err = zabs_f(cr - gr, ci - gi) / zabs_f(cr, ci);
          if (err > worst) worst = err;
          if (err > tol) bad++;
        }
  if (max_err) *max_err = worst;
  return bad;
}

void oracle_fft_f64(double *out, const float *in, int n, int batch, int inverse) {
  double *X = (double *)malloc(sizeof(double) * 2 * (size_t)n);
  double *Y = (double *)malloc(sizeof(double) * 2 * (size_t)n);
  if (!X || !Y) {
    free(X);
    free(Y);
    return;
  }
  const double sign = inverse ? 1.0 : -1.0;
  for (int b = 0; b < batch; ++b) {
    const float *src = in + (size_t)b * 2 * n;
    for (int i = 0; i < 2 * n; ++i) X[i] = src[i];
    /* pass with span L: Y[2jL + k] = X[jL + k] + w X[jL + n/2 + k],
     * Y[2jL + L + k] = X[jL + k] - w X[jL + n/2 + k], w = exp(sign i pi k / L) */
    for (int L = 1; L < n; L *= 2) {
      for (int k = 0; k < L; ++k) {
        const double phi = sign * M_PI * k / L;
        const double wr = cos(phi), wi = sin(phi);
        for (int j = 0; j < n / (2 * L); ++j) {
          const double *a = X + 2 * (j * L + k), *c = X + 2 * (j * L + n / 2 + k);
          const double tr = wr * c[0] - wi * c[1], ti = wr * c[1] + wi * c[0];
          double *y0 = Y + 2 * (2 * j * L + k), *y1 = Y + 2 * (2 * j * L + L + k);
          y0[0] = a[0] + tr;
          y0[1] = a[1] + ti;
          y1[0] = a[0] - tr;
          y1[1] = a[1] - ti;
        }
      }
      double *T = X;
      X = Y;
      Y = T;
    }
    memcpy(out + (size_t)b * 2 * n, X, sizeof(double) * 2 * (size_t)n);
  }
  free(X);
  free(Y);
}
