/*
 * serve.c -- continuous-batching correlation server (include/xgpu_serve.h).
 *
 * Two threads per server:
 *   scheduler   admits pending blocks into the HBM staging pool (first fit),
 *               copies them to the device, launches one ragged persistent
 *               batch (xgpuAmdCorrelateBatch), and queues due dumps;
 *   completion  waits on the batch / dump events in issue order, frees
 *               staging, and records block-to-dump latency from device event
 *               timestamps.
 * A stream's consecutive pending blocks of one integration are fused into one
 * job: host blocks are staged back to back, device blocks when they already
 * lie back to back in memory.  One kernel pass then integrates them all and
 * read-modify-writes the stream's integration buffer once instead of once per
 * block (for short blocks of a large subarray that pass is most of the cost).
 * Engines are only touched by the scheduler thread.  Locking: s->mu guards
 * the queues, the pool, the streams' counters and the statistics; the event
 * free list has a lock of its own (ev_mu), because the scheduler takes events
 * after releasing s->mu while the completion thread returns them.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "engine_internal.h"
#include "xg_hip.h"
#include "xgpu_amd.h"
#include "xgpu_serve.h"

#define XS_MAX_STREAMS 64
#define XS_MAX_RANGES 1024
#define XS_ALIGN 256
#define XS_MAX_FUSE 64        /* blocks fused into one job */

typedef struct Block {
  int stream;
  const void *src;         /* host block, or device block when on_device */
  int on_device;
  unsigned ntime;
  double t_arrival;
  size_t bytes;
  struct Block *next;      /* global arrival order */
} Block;

/* One job of a batch: nblk consecutive blocks of one stream, fused */
typedef struct {
  Block *blk[XS_MAX_FUSE];
  int nblk;
  unsigned ntime;
  size_t off, staged;      /* staging range (staged = 0: device blocks, none) */
} Job;

typedef struct {
  XGPUAmdEngine *eng;
  XGPUStreamConfig cfg;
  char *ring;
  unsigned slots;
  size_t dump_bytes;
  size_t bytes_per_sample;
  unsigned long long cmac_per_sample;
  unsigned long long submitted, done, dumps_issued, dumps_done;
  unsigned in_integ;       /* blocks launched into the current integration */
  double last_arrival;
} Stream;

typedef struct Record {
  int kind;                /* 0 batch, 1 dump */
  void *event;
  int n;
  int stream[XG_MAX_JOBS];
  int nblk[XG_MAX_JOBS];   /* blocks fused into job i */
  size_t off[XG_MAX_JOBS], bytes[XG_MAX_JOBS];  /* staging to release (bytes 0: none) */
  unsigned long long cmac;
  double t_arrival;
  int sid;
  struct Record *next;
} Record;

typedef struct {
  size_t off, len;
} Range;

struct XGPUServerStruct {
  XGPUServerConfig cfg;
  pthread_mutex_t mu;
  pthread_cond_t cv;
  pthread_t sched, comp;
  int threads_started;
  int stop;
  int error;
  char errmsg[256];

  Stream streams[XS_MAX_STREAMS];
  int nstreams;

  Block *pend_head, *pend_tail;
  unsigned long long pending;   /* submitted, not yet completed */

  char *pool;
  Range ranges[XS_MAX_RANGES];  /* free ranges, sorted by offset */
  int nranges;
  size_t used, peak;

  void *copy_stream, *compute_stream;
  void *ev_ref;
  double t_ref;
  pthread_mutex_t ev_mu;        /* guards ev_free / nev_free only */
  void *ev_free[512];
  int nev_free;

  Record *rec_head, *rec_tail;
  int inflight;

  double *lat;
  size_t nlat, caplat;
  unsigned long long blocks_submitted, blocks_done, dumps_done, launches, launched_blocks, cmac_done;
  unsigned long long launched_jobs;
  double t_first, t_last;
};

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ---- staging pool (first fit over sorted free ranges) ------------------- */
static int pool_alloc(XGPUServer *s, size_t bytes, size_t *off) {
  bytes = (bytes + XS_ALIGN - 1) / XS_ALIGN * XS_ALIGN;
  for (int i = 0; i < s->nranges; ++i) {
    if (s->ranges[i].len >= bytes) {
      *off = s->ranges[i].off;
      s->ranges[i].off += bytes;
      s->ranges[i].len -= bytes;
      if (s->ranges[i].len == 0) {
        memmove(&s->ranges[i], &s->ranges[i + 1], sizeof(Range) * (s->nranges - i - 1));
        s->nranges--;
      }
      s->used += bytes;
      if (s->used > s->peak) s->peak = s->used;
      return 1;
    }
  }
  return 0;
}

static void pool_free(XGPUServer *s, size_t off, size_t bytes) {
  bytes = (bytes + XS_ALIGN - 1) / XS_ALIGN * XS_ALIGN;
  s->used -= bytes;
  int i = 0;
  while (i < s->nranges && s->ranges[i].off < off) ++i;
  /* merge with the previous and / or next range when adjacent */
  int merged_prev = (i > 0 && s->ranges[i - 1].off + s->ranges[i - 1].len == off);
  int merged_next = (i < s->nranges && off + bytes == s->ranges[i].off);
  if (merged_prev && merged_next) {
    s->ranges[i - 1].len += bytes + s->ranges[i].len;
    memmove(&s->ranges[i], &s->ranges[i + 1], sizeof(Range) * (s->nranges - i - 1));
    s->nranges--;
  } else if (merged_prev) {
    s->ranges[i - 1].len += bytes;
  } else if (merged_next) {
    s->ranges[i].off = off;
    s->ranges[i].len += bytes;
  } else {
    /* XS_MAX_RANGES bounds fragmentation: at most one range per live block + 1 */
    memmove(&s->ranges[i + 1], &s->ranges[i], sizeof(Range) * (s->nranges - i));
    s->ranges[i].off = off;
    s->ranges[i].len = bytes;
    s->nranges++;
  }
}

/* Event free list, under ev_mu: the scheduler takes events outside s->mu
 * while the completion thread returns them. */
static void *get_event(XGPUServer *s) {
  void *e = NULL;
  pthread_mutex_lock(&s->ev_mu);
  if (s->nev_free > 0) e = s->ev_free[--s->nev_free];
  pthread_mutex_unlock(&s->ev_mu);
  if (e) return e;
  if (xgh_event_create(&e, 1)) return NULL;
  return e;
}

static void put_event(XGPUServer *s, void *e) {
  pthread_mutex_lock(&s->ev_mu);
  const int keep = s->nev_free < (int)(sizeof s->ev_free / sizeof s->ev_free[0]);
  if (keep) s->ev_free[s->nev_free++] = e;
  pthread_mutex_unlock(&s->ev_mu);
  if (!keep) xgh_event_destroy(e);
}

static void push_record(XGPUServer *s, Record *r) {
  r->next = NULL;
  if (s->rec_tail)
    s->rec_tail->next = r;
  else
    s->rec_head = r;
  s->rec_tail = r;
}

static void set_error(XGPUServer *s, const char *what) {
  if (!s->error) {
    s->error = 1;
    snprintf(s->errmsg, sizeof s->errmsg, "%s: %s", what, xgpuAmdLastError());
    fprintf(stderr, "(xgpu serve) %s\n", s->errmsg);
  }
}

/* ---- scheduler ----------------------------------------------------------- */
/* Unlink b, whose predecessor in the pending list is prev (NULL: the head). */
static void unlink_pending(XGPUServer *s, Block *prev, Block *b) {
  if (prev) prev->next = b->next; else s->pend_head = b->next;
  if (s->pend_tail == b) s->pend_tail = prev;
}

/* Form the job of the stream whose oldest pending block is `first` (with
 * predecessor `prev`): `first` and the stream's next pending blocks, as long as
 * they stay in the current integration (the dump after blocks_per_dump blocks
 * closes it), fit the samples one launch may integrate (the engine's bound),
 * have the same source kind and, for device blocks, follow each other in
 * memory.  A host job takes one staging range for all its blocks; when that
 * does not fit, fewer blocks are taken.  On success the blocks leave the
 * pending list.  Returns 0 when not even `first` is admitted.  Under s->mu. */
static int take_job(XGPUServer *s, Block *prev, Block *first, Job *job) {
  Stream *st = &s->streams[first->stream];
  Block *cand[XS_MAX_FUSE], *pred[XS_MAX_FUSE];
  unsigned room = st->cfg.blocks_per_dump - st->in_integ;
  if (room > XS_MAX_FUSE) room = XS_MAX_FUSE;
  int n = 0;
  unsigned ntime = 0;
  for (Block *p = prev, *b = first; b && (unsigned)n < room; p = b, b = b->next) {
    if (b->stream != first->stream) continue;
    if (n > 0) {
      const Block *last = cand[n - 1];
      if (b->on_device != first->on_device || ntime + b->ntime > st->eng->bound) break;
      if (b->on_device && (const char *)b->src != (const char *)last->src + last->bytes) break;
    }
    cand[n] = b;
    pred[n] = p;
    ntime += b->ntime;
    ++n;
  }
  job->staged = 0;
  job->off = 0;
  if (!first->on_device) {
    for (;;) {
      size_t bytes = 0;
      for (int k = 0; k < n; ++k) bytes += cand[k]->bytes;
      if (pool_alloc(s, bytes, &job->off)) {
        job->staged = bytes;
        break;
      }
      if (n == 1) return 0;
      n /= 2;
    }
  }
  job->nblk = n;
  job->ntime = 0;
  for (int k = n - 1; k >= 0; --k) { /* later blocks first: earlier ones may be their predecessors */
    unlink_pending(s, pred[k], cand[k]);
    job->blk[k] = cand[k];
    job->ntime += cand[k]->ntime;
  }
  return 1;
}

static void *scheduler_main(void *arg) {
  XGPUServer *s = (XGPUServer *)arg;
  xgh_set_device(s->cfg.device);
  Job jobs[XG_MAX_JOBS];
  pthread_mutex_lock(&s->mu);
  for (;;) {
    while (!s->stop && !s->error && (!s->pend_head || s->inflight >= s->cfg.max_inflight))
      pthread_cond_wait(&s->cv, &s->mu);
    if (s->error || (s->stop && !s->pend_head)) break;
    /* Form a batch: per stream, in order of its oldest pending block, one job
     * of one compute type, as long as the staging pool admits it. */
    int n = 0;
    int seen[XS_MAX_STREAMS] = {0};
    unsigned type = 0;
    Block *prev = NULL, *b = s->pend_head;
    while (b && n < s->cfg.max_batch) {
      Stream *st = &s->streams[b->stream];
      if (!seen[b->stream] && (n == 0 || st->cfg.compute_type == type)) {
        seen[b->stream] = 1;
        if (take_job(s, prev, b, &jobs[n])) {
          if (n == 0) type = st->cfg.compute_type;
          ++n;
          /* b and its fused blocks left the list; prev (never one of them)
           * now precedes the next block to look at */
          b = prev ? prev->next : s->pend_head;
          continue;
        }
      }
      seen[b->stream] = 1;
      prev = b;
      b = b->next;
    }
    if (n == 0) { /* nothing fits: wait for a completion to free staging */
      if (s->inflight == 0 && s->rec_head == NULL) {
        snprintf(s->errmsg, sizeof s->errmsg, "a pending block does not fit the staging pool");
        s->error = 1;
        break;
      }
      pthread_cond_wait(&s->cv, &s->mu);
      continue;
    }
    pthread_mutex_unlock(&s->mu);

    /* device work, outside the lock */
    Record *rb = (Record *)calloc(1, sizeof(Record));
    XGPUAmdBatchItem items[XG_MAX_JOBS];
    Record *dumps[XG_MAX_JOBS];
    int ndump = 0, ncopies = 0;
    int ok = rb != NULL;
    for (int i = 0; ok && i < n; ++i) {
      Job *jb = &jobs[i];
      items[i].engine = s->streams[jb->blk[0]->stream].eng;
      items[i].ntime = jb->ntime;
      if (jb->staged) {
        size_t at = jb->off;
        for (int k = 0; ok && k < jb->nblk; ++k) {
          if (xgh_memcpy_h2d_async(s->pool + at, jb->blk[k]->src, jb->blk[k]->bytes, s->copy_stream)) {
            xg_fail(XGPU_CUDA_ERROR, "staging copy: %s", xgh_last_error());
            ok = 0;
          }
          at += jb->blk[k]->bytes;
          ++ncopies;
        }
        items[i].input_d = s->pool + jb->off;
      } else {
        items[i].input_d = jb->blk[0]->src;
      }
    }
    void *ev_copy = NULL;
    if (ok && ncopies) {
      ev_copy = get_event(s);
      if (!ev_copy || xgh_event_record(ev_copy, s->copy_stream) ||
          xgh_stream_wait_event(s->compute_stream, ev_copy))
        ok = 0;
    }
    if (ok && xgpuAmdCorrelateBatch(items, n, s->compute_stream)) ok = 0;
    if (ok) {
      rb->kind = 0;
      rb->n = n;
      rb->event = get_event(s);
      if (!rb->event || xgh_event_record(rb->event, s->compute_stream)) ok = 0;
    }
    unsigned long long nblocks = 0;
    for (int i = 0; ok && i < n; ++i) {
      Job *jb = &jobs[i];
      const int sid = jb->blk[0]->stream;
      Stream *st = &s->streams[sid];
      rb->stream[i] = sid;
      rb->nblk[i] = jb->nblk;
      rb->off[i] = jb->off;
      rb->bytes[i] = jb->staged;
      rb->cmac += st->cmac_per_sample * jb->ntime;
      nblocks += (unsigned long long)jb->nblk;
      st->in_integ += (unsigned)jb->nblk;
      st->last_arrival = jb->blk[jb->nblk - 1]->t_arrival;
      if (st->in_integ == st->cfg.blocks_per_dump) {
        Record *rd = (Record *)calloc(1, sizeof(Record));
        if (!rd) { ok = 0; break; }
        rd->kind = 1;
        rd->sid = sid;
        rd->t_arrival = st->last_arrival;
        char *slot = st->ring + (st->dumps_issued % st->slots) * st->dump_bytes;
        if (xg_dump_async(st->eng, slot, XGPU_DUMP_WIDE | XGPU_DUMP_TRIANGULAR |
                                             (st->cfg.dump_device ? XGPU_DUMP_DEVICE : 0))) {
          free(rd);
          ok = 0;
          break;
        }
        rd->event = get_event(s);
        if (!rd->event || xgh_event_record(rd->event, xgpuAmdStream(st->eng))) { free(rd); ok = 0; break; }
        xgpuAmdClear(st->eng);
        st->in_integ = 0;
        st->dumps_issued++;
        dumps[ndump++] = rd;
      }
    }
    pthread_mutex_lock(&s->mu);
    if (!ok) {
      set_error(s, "batch launch");
      if (ev_copy) put_event(s, ev_copy);
      if (rb && rb->event) put_event(s, rb->event);
      for (int i = 0; i < ndump; ++i) {
        put_event(s, dumps[i]->event);
        free(dumps[i]);
      }
      free(rb);
      for (int i = 0; i < n; ++i)
        for (int k = 0; k < jobs[i].nblk; ++k) free(jobs[i].blk[k]);
      break;
    }
    push_record(s, rb);
    for (int i = 0; i < ndump; ++i) push_record(s, dumps[i]);
    /* the copy event is complete once the batch event is; recycle it then */
    if (ev_copy) {
      Record *rc = (Record *)calloc(1, sizeof(Record));
      if (rc) {
        rc->kind = 2;
        rc->event = ev_copy;
        push_record(s, rc);
      } else {
        put_event(s, ev_copy); /* re-recording a pending event is allowed */
      }
    }
    s->inflight++;
    s->launches++;
    s->launched_jobs += (unsigned long long)n;
    s->launched_blocks += nblocks;
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < jobs[i].nblk; ++k) free(jobs[i].blk[k]);
    pthread_cond_broadcast(&s->cv);
  }
  pthread_cond_broadcast(&s->cv);
  pthread_mutex_unlock(&s->mu);
  return NULL;
}

/* ---- completion ---------------------------------------------------------- */
static void *completion_main(void *arg) {
  XGPUServer *s = (XGPUServer *)arg;
  xgh_set_device(s->cfg.device);
  pthread_mutex_lock(&s->mu);
  for (;;) {
    while (!s->rec_head && !(s->stop && s->pending == 0) && !s->error)
      pthread_cond_wait(&s->cv, &s->mu);
    if (!s->rec_head) break;
    Record *r = s->rec_head;
    pthread_mutex_unlock(&s->mu);
    int rc = xgh_event_sync(r->event);
    float ms = 0.f;
    if (!rc && r->kind == 1) rc = xgh_event_elapsed_ms(&ms, s->ev_ref, r->event);
    const double t = now_s();
    pthread_mutex_lock(&s->mu);
    s->rec_head = r->next;
    if (!s->rec_head) s->rec_tail = NULL;
    if (rc) {
      set_error(s, "completion");
    } else if (r->kind == 0) {
      unsigned long long nb = 0;
      for (int i = 0; i < r->n; ++i) {
        if (r->bytes[i]) pool_free(s, r->off[i], r->bytes[i]);
        s->streams[r->stream[i]].done += (unsigned long long)r->nblk[i];
        nb += (unsigned long long)r->nblk[i];
      }
      s->blocks_done += nb;
      s->pending -= nb;
      s->cmac_done += r->cmac;
      s->inflight--;
      s->t_last = t;
    } else if (r->kind == 1) {
      const double t_dev = s->t_ref + 1e-3 * (double)ms;
      if (s->nlat == s->caplat) {
        size_t cap = s->caplat ? 2 * s->caplat : 1024;
        double *p = (double *)realloc(s->lat, cap * sizeof(double));
        if (p) {
          s->lat = p;
          s->caplat = cap;
        }
      }
      if (s->nlat < s->caplat) s->lat[s->nlat++] = t_dev - r->t_arrival;
      s->streams[r->sid].dumps_done++;
      s->dumps_done++;
      if (t_dev > s->t_last) s->t_last = t_dev;
    }
    put_event(s, r->event);
    free(r);
    pthread_cond_broadcast(&s->cv);
  }
  pthread_mutex_unlock(&s->mu);
  return NULL;
}

/* ---- API ------------------------------------------------------------------ */
int xgpuAmdHostAlloc(void **ptr, size_t bytes) {
  if (!ptr) return xg_fail(XGPU_INVALID_CONFIG, "NULL pointer");
  if (xgh_host_alloc(ptr, bytes)) return xg_fail(XGPU_OUT_OF_MEMORY, "%s", xgh_last_error());
  return XGPU_OK;
}

int xgpuAmdHostFree(void *ptr) {
  if (xgh_host_free(ptr)) return xg_fail(XGPU_CUDA_ERROR, "%s", xgh_last_error());
  return XGPU_OK;
}

int xgpuServerCreate(XGPUServer **out, const XGPUServerConfig *cfg) {
  if (!out || !cfg) return xg_fail(XGPU_INVALID_CONFIG, "NULL argument");
  *out = NULL;
  XGPUServer *s = (XGPUServer *)calloc(1, sizeof *s);
  if (!s) return xg_fail(XGPU_OUT_OF_MEMORY, "out of host memory");
  s->cfg = *cfg;
  if (s->cfg.staging_bytes == 0) s->cfg.staging_bytes = (size_t)4 << 30;
  if (s->cfg.max_batch <= 0 || s->cfg.max_batch > XG_MAX_JOBS) s->cfg.max_batch = XG_MAX_JOBS;
  if (s->cfg.max_inflight <= 0) s->cfg.max_inflight = 2;
  s->cfg.staging_bytes = (s->cfg.staging_bytes + XS_ALIGN - 1) / XS_ALIGN * XS_ALIGN;
  pthread_mutex_init(&s->mu, NULL);
  pthread_mutex_init(&s->ev_mu, NULL);
  pthread_cond_init(&s->cv, NULL);
  int ndev = 0;
  if (xgh_device_count(&ndev) || ndev == 0) {
    pthread_cond_destroy(&s->cv);
    pthread_mutex_destroy(&s->ev_mu);
    pthread_mutex_destroy(&s->mu);
    free(s);
    return xg_fail(XGPU_NO_DEVICE, "no HIP device found");
  }
  if (xgh_set_device(cfg->device) || xgh_malloc((void **)&s->pool, s->cfg.staging_bytes) ||
      xgh_stream_create(&s->copy_stream) || xgh_stream_create(&s->compute_stream) ||
      xgh_event_create(&s->ev_ref, 1) || xgh_event_record(s->ev_ref, s->compute_stream) ||
      xgh_event_sync(s->ev_ref)) {
    int rc = xg_fail(XGPU_CUDA_ERROR, "server setup: %s", xgh_last_error());
    xgpuServerDestroy(s);
    return rc;
  }
  s->t_ref = now_s();
  s->ranges[0].off = 0;
  s->ranges[0].len = s->cfg.staging_bytes;
  s->nranges = 1;
  if (pthread_create(&s->sched, NULL, scheduler_main, s) ||
      pthread_create(&s->comp, NULL, completion_main, s)) {
    xgpuServerDestroy(s);
    return xg_fail(XGPU_OUT_OF_MEMORY, "cannot start server threads");
  }
  s->threads_started = 1;
  *out = s;
  return XGPU_OK;
}

int xgpuServerDrain(XGPUServer *s) {
  if (!s) return xg_fail(XGPU_NOT_INITIALIZED, "NULL server");
  pthread_mutex_lock(&s->mu);
  while (!s->error && (s->pending > 0 || s->rec_head)) pthread_cond_wait(&s->cv, &s->mu);
  const int err = s->error;
  pthread_mutex_unlock(&s->mu);
  if (err) return xg_fail(XGPU_CUDA_ERROR, "server: %s", s->errmsg);
  return XGPU_OK;
}

int xgpuServerDestroy(XGPUServer *s) {
  if (!s) return XGPU_OK;
  if (s->threads_started) {
    xgpuServerDrain(s);
    pthread_mutex_lock(&s->mu);
    s->stop = 1;
    pthread_cond_broadcast(&s->cv);
    pthread_mutex_unlock(&s->mu);
    pthread_join(s->sched, NULL);
    pthread_join(s->comp, NULL);
  }
  xgh_set_device(s->cfg.device);
  if (s->compute_stream) xgh_stream_sync(s->compute_stream);
  if (s->copy_stream) xgh_stream_sync(s->copy_stream);
  for (int i = 0; i < s->nstreams; ++i) xgpuAmdDestroy(s->streams[i].eng);
  while (s->pend_head) {
    Block *b = s->pend_head;
    s->pend_head = b->next;
    free(b);
  }
  while (s->rec_head) {
    Record *r = s->rec_head;
    s->rec_head = r->next;
    xgh_event_destroy(r->event);
    free(r);
  }
  for (int i = 0; i < s->nev_free; ++i) xgh_event_destroy(s->ev_free[i]);
  xgh_event_destroy(s->ev_ref);
  xgh_stream_destroy(s->copy_stream);
  xgh_stream_destroy(s->compute_stream);
  xgh_free(s->pool);
  free(s->lat);
  pthread_cond_destroy(&s->cv);
  pthread_mutex_destroy(&s->ev_mu);
  pthread_mutex_destroy(&s->mu);
  free(s);
  return XGPU_OK;
}

int xgpuServerOpenStream(XGPUServer *s, const XGPUStreamConfig *cfg, void *dump_ring,
                         unsigned dump_slots, int *stream_id) {
  if (!s || !cfg || !stream_id) return xg_fail(XGPU_INVALID_CONFIG, "NULL argument");
  if (!dump_ring || dump_slots == 0) return xg_fail(XGPU_INVALID_CONFIG, "no dump ring");
  if (cfg->blocks_per_dump == 0) return xg_fail(XGPU_INVALID_CONFIG, "blocks_per_dump must be >= 1");
  if (cfg->dump_device > 1) return xg_fail(XGPU_INVALID_CONFIG, "dump_device must be 0 or 1");
  if (cfg->compute_type != XGPU_INT8 && cfg->compute_type != XGPU_FLOAT16)
    return xg_fail(XGPU_INVALID_CONFIG, "server streams are XGPU_INT8 or XGPU_FLOAT16");
  XGPUAmdConfig ec;
  memset(&ec, 0, sizeof ec);
  ec.npol = 2;
  ec.nstation = cfg->nstation;
  ec.nfrequency = cfg->nfrequency;
  ec.compute_type = cfg->compute_type;
  ec.matrix_order = TRIANGULAR_ORDER;
  ec.device = s->cfg.device;
  XGPUAmdEngine *e = NULL;
  int rc = xgpuAmdCreate(&e, &ec);
  if (rc) return rc;
  pthread_mutex_lock(&s->mu);
  if (s->nstreams == XS_MAX_STREAMS) {
    pthread_mutex_unlock(&s->mu);
    xgpuAmdDestroy(e);
    return xg_fail(XGPU_INVALID_CONFIG, "too many streams (max %d)", XS_MAX_STREAMS);
  }
  const int id = s->nstreams;
  Stream *st = &s->streams[id];
  memset(st, 0, sizeof *st);
  st->eng = e;
  st->cfg = *cfg;
  st->ring = (char *)dump_ring;
  st->slots = dump_slots;
  st->dump_bytes = xg_dump_bytes(e, XGPU_DUMP_WIDE | XGPU_DUMP_TRIANGULAR);
  st->bytes_per_sample = (size_t)e->tstride;
  st->cmac_per_sample = (unsigned long long)cfg->nfrequency * cfg->nstation * (cfg->nstation + 1) / 2 * 4;
  s->nstreams++;
  *stream_id = id;
  pthread_mutex_unlock(&s->mu);
  return XGPU_OK;
}

size_t xgpuServerDumpBytes(XGPUServer *s, int id) {
  if (!s || id < 0 || id >= s->nstreams) return 0;
  return s->streams[id].dump_bytes;
}

static int submit(XGPUServer *s, int id, const void *src, int on_device, unsigned ntime) {
  if (!s) return xg_fail(XGPU_NOT_INITIALIZED, "NULL server");
  if (id < 0 || id >= s->nstreams) return xg_fail(XGPU_INVALID_CONFIG, "bad stream id %d", id);
  if (!src || ntime == 0) return xg_fail(XGPU_INVALID_CONFIG, "empty block");
  Stream *st = &s->streams[id];
  if (ntime > st->eng->bound)
    return xg_fail(XGPU_INVALID_CONFIG, "block of %u samples exceeds %u", ntime, st->eng->bound);
  const size_t bytes = st->bytes_per_sample * ntime;
  if (on_device) {
    if ((uintptr_t)src & 7) return xg_fail(XGPU_INVALID_CONFIG, "device block must be 8-byte aligned");
  } else if (bytes > s->cfg.staging_bytes) {
    return xg_fail(XGPU_OUT_OF_MEMORY, "block of %zu bytes exceeds the staging pool", bytes);
  }
  Block *b = (Block *)calloc(1, sizeof *b);
  if (!b) return xg_fail(XGPU_OUT_OF_MEMORY, "out of host memory");
  b->stream = id;
  b->src = src;
  b->on_device = on_device;
  b->ntime = ntime;
  b->bytes = bytes;
  pthread_mutex_lock(&s->mu);
  if (s->error) {
    pthread_mutex_unlock(&s->mu);
    free(b);
    return xg_fail(XGPU_CUDA_ERROR, "server: %s", s->errmsg);
  }
  b->t_arrival = now_s();
  if (s->blocks_submitted == 0 || s->t_first == 0.0) s->t_first = b->t_arrival;
  if (s->pend_tail)
    s->pend_tail->next = b;
  else
    s->pend_head = b;
  s->pend_tail = b;
  s->pending++;
  s->blocks_submitted++;
  st->submitted++;
  pthread_cond_broadcast(&s->cv);
  pthread_mutex_unlock(&s->mu);
  return XGPU_OK;
}

int xgpuServerSubmit(XGPUServer *s, int id, const void *block_h, unsigned ntime) {
  return submit(s, id, block_h, 0, ntime);
}

int xgpuServerSubmitDevice(XGPUServer *s, int id, const void *block_d, unsigned ntime) {
  return submit(s, id, block_d, 1, ntime);
}

unsigned long long xgpuServerBlocksDone(XGPUServer *s, int id) {
  if (!s || id < 0 || id >= s->nstreams) return 0;
  pthread_mutex_lock(&s->mu);
  unsigned long long v = s->streams[id].done;
  pthread_mutex_unlock(&s->mu);
  return v;
}

unsigned long long xgpuServerDumpsDone(XGPUServer *s, int id) {
  if (!s || id < 0 || id >= s->nstreams) return 0;
  pthread_mutex_lock(&s->mu);
  unsigned long long v = s->streams[id].dumps_done;
  pthread_mutex_unlock(&s->mu);
  return v;
}

static int cmp_double(const void *a, const void *b) {
  const double x = *(const double *)a, y = *(const double *)b;
  return (x > y) - (x < y);
}

static double pct(const double *v, size_t n, double p) {
  if (n == 0) return 0.0;
  double k = p * (double)(n - 1);
  size_t i = (size_t)k;
  if (i + 1 >= n) return v[n - 1];
  return v[i] + (k - (double)i) * (v[i + 1] - v[i]);
}

int xgpuServerStats(XGPUServer *s, XGPUServerStats *st) {
  if (!s || !st) return xg_fail(XGPU_INVALID_CONFIG, "NULL argument");
  memset(st, 0, sizeof *st);
  pthread_mutex_lock(&s->mu);
  st->blocks_submitted = s->blocks_submitted;
  st->blocks_done = s->blocks_done;
  st->dumps_done = s->dumps_done;
  st->launches = s->launches;
  st->cmac_done = s->cmac_done;
  st->busy_s = (s->t_last > s->t_first) ? s->t_last - s->t_first : 0.0;
  st->tcmac_per_s = st->busy_s > 0 ? (double)s->cmac_done / st->busy_s / 1e12 : 0.0;
  st->mean_batch = s->launches ? (double)s->launched_jobs / (double)s->launches : 0.0;
  st->mean_fused = s->launched_jobs ? (double)s->launched_blocks / (double)s->launched_jobs : 0.0;
  st->staging_peak_bytes = s->peak;
  double *v = NULL;
  size_t n = s->nlat;
  if (n) {
    v = (double *)malloc(n * sizeof(double));
    if (v) memcpy(v, s->lat, n * sizeof(double));
  }
  pthread_mutex_unlock(&s->mu);
  if (v) {
    qsort(v, n, sizeof(double), cmp_double);
    st->p50_ms = 1e3 * pct(v, n, 0.50);
    st->p99_ms = 1e3 * pct(v, n, 0.99);
    st->max_ms = 1e3 * v[n - 1];
    free(v);
  }
  return XGPU_OK;
}

int xgpuServerResetStats(XGPUServer *s) {
  if (!s) return xg_fail(XGPU_NOT_INITIALIZED, "NULL server");
  pthread_mutex_lock(&s->mu);
  s->nlat = 0;
  s->blocks_submitted = s->blocks_done = s->dumps_done = 0;
  s->launches = s->launched_blocks = s->launched_jobs = s->cmac_done = 0;
  s->t_first = s->t_last = 0.0;
  s->peak = s->used;
  pthread_mutex_unlock(&s->mu);
  return XGPU_OK;
}
