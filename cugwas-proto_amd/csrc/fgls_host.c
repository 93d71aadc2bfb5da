/*
 * fgls_host.c -- C host side of the Cholesky FGLS solver for one MI355X.
 *
 * Implements fgls_chol_gpu() (include/fgls_chol_gpu.h) with the reference's
 * arguments, files and output layout (src/fgls_chol_gpu.c:97-795), as a
 * single-GPU streaming pipeline:
 *
 *   reader thread:  pread(XR.in block c)  -> pinned host slot c%2
 *                   (pieces of up to 32 MB over io_threads OpenMP threads)
 *   copy stream:    hipMemcpy2DAsync slot -> device X_R buffer c%2
 *   compute stream: conversions / tiles, then fgls_hip_solve_block* (product
 *                   through L^-1 + per-SNP stage) per trait -> device B/V c%2
 *                   hipMemcpyAsync B/V -> pinned host B/V c%2
 *   main thread:    pwrite B/V of block c-1 at the reference's offsets
 *
 * Ordering is by hipEvents; every buffer is double buffered (NSLOT), so disk
 * reads, PCIe uploads, device compute and output writes of neighbouring
 * blocks overlap.  Per trait: by default Phi is uploaded once and each
 * trait's M is built, factorised and inverted on the device
 * (fgls_hip_set_trait_phi); potrf=host forms and factorises M on the host
 * (LAPACK dpotrf, as fgls_chol_gpu.c:406-430) and uploads L.  Several traits
 * share one read of X_R (traits_per_pass; the CPU twin re-reads it per
 * trait, fgls_chol.c:223 and :336-337).
 *
 * fgls_eigen() (include/fgls_eigen.h, the reference's src/fgls_eigen.c:37-236)
 * runs the same pipeline in the eigenbasis of Phi: one host dsyevr of Phi
 * (fgls_eigen.c:780-806), Z and its eigenvalues uploaded once, and one dense
 * product Z' x_R per block serving every trait of a pass
 * (fgls_hip_eig_solve_block).
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <errno.h>
#include <fcntl.h>
#include <math.h>
#include <omp.h>
#include <pthread.h>
#include <sched.h>
#include <semaphore.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>
#include <unistd.h>

#include <hip/hip_runtime_api.h>

#include "fgls_chol_gpu.h"
#include "fgls_eigen.h"
#include "fgls_hip.h"
#include "fgls_internal.h"

/* LP64 LAPACK of the OpenBLAS build the library is linked with. */
extern void scipy_dpotrf_(const char *uplo, const int *n, double *a, const int *lda, int *info);
extern void scipy_openblas_set_num_threads(int n);
extern void scipy_dsyevr_(const char *jobz, const char *range, const char *uplo, const int *n,
                          double *a, const int *lda, const double *vl, const double *vu,
                          const int *il, const int *iu, const double *abstol, int *m, double *w,
                          double *z, const int *ldz, int *isuppz, double *work, const int *lwork,
                          int *iwork, const int *liwork, int *info);

typedef struct {
  int64_t block;     /* SNPs per device launch */
  int device;
  int serial;
  int stats;
  int mixed;         /* precision=mixed: int8 genotypes x sliced L^-1 */
  int io_threads;    /* parallel preads of X_R per block */
  int host_potrf;    /* potrf=host: M and its factor on the host (OpenBLAS) */
  int traits_per_pass;  /* traits sharing one read of X_R (0: as many as fit) */
  int xr_i8;         /* xr_format: 0 f64, 1 i8 (n bytes per column), 2 u2 (ceil(n/4) bytes per
                        column), 3 i16 (dosages x * dosage_scale, 2 n bytes per column) */
  double dosage_scale;  /* precision=mixed for dosages: x * scale = 128 hi + lo (0: genotypes) */
  int eig_host;      /* eig=host: eigendecomposition of Phi by host LAPACK dsyevr (default: device) */
  int no_roctx;      /* roctx=0: no roctx ranges around the pipeline stages */
  int slices_six;    /* slices=6: six digits of L^-1 in the mixed mode (default auto: five where the
                        error bound allows, fgls_hip_ctx_set_slice_form) */
} options_t;

/* roctx ranges around the pipeline's stages, the counterpart of the
 * reference's VampirTrace regions (fgls_chol_gpu.c:57-59, :69-71): a
 * profiler (rocprofv3 --marker-trace) shows the trait setups, the block loop
 * and, per block, the read, the enqueue and the write-back.  The roctx
 * library is resolved at the first use; without it the ranges are no-ops.
 * Ranges nest per thread. */
typedef int (*roctx_push_t)(const char *);
typedef int (*roctx_pop_t)(void);
static roctx_push_t g_roctx_push;
static roctx_pop_t g_roctx_pop;
static pthread_once_t g_roctx_once = PTHREAD_ONCE_INIT;

static void roctx_load(void) {
  static const char *const libs[] = {"librocprofiler-sdk-roctx.so.1", "libroctx64.so.4"};
  for (size_t i = 0; i < sizeof(libs) / sizeof(libs[0]); ++i) {
    void *h = dlopen(libs[i], RTLD_NOW | RTLD_LOCAL);
    if (!h) continue;
    roctx_push_t push = (roctx_push_t)dlsym(h, "roctxRangePushA");
    roctx_pop_t pop = (roctx_pop_t)dlsym(h, "roctxRangePop");
    if (push && pop) {
      g_roctx_push = push;
      g_roctx_pop = pop;
      return;
    }
    dlclose(h);
  }
}

static void range_push(int on, const char *name) {
  if (!on) return;
  pthread_once(&g_roctx_once, roctx_load);
  if (g_roctx_push) g_roctx_push(name);
}

static void range_pop(int on) {
  if (on && g_roctx_pop) g_roctx_pop();
}

/* Timings of the calling thread's last completed fgls_*_gpu_ex call. */
static __thread fgls_run_stats g_last_stats;

int fgls_last_run_stats(fgls_run_stats *out) {
  if (!out) return fgls_set_error(FGLS_ERR_ARG, "NULL output pointer");
  *out = g_last_stats;
  return FGLS_OK;
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static int host_cores(void) {
  cpu_set_t set;
  if (sched_getaffinity(0, sizeof(set), &set) == 0) return CPU_COUNT(&set);
  long n = sysconf(_SC_NPROCESSORS_ONLN);
  return n > 0 ? (int)n : 1;
}

static void parse_options(const char *s, options_t *o) {
  if (!s) return;
  char buf[512];
  snprintf(buf, sizeof(buf), "%s", s);
  char *save = NULL;
  for (char *tok = strtok_r(buf, ",", &save); tok; tok = strtok_r(NULL, ",", &save)) {
    char *eq = strchr(tok, '=');
    if (!eq) continue;
    *eq = 0;
    if (!strcmp(tok, "xr_format")) {
      o->xr_i8 = !strcmp(eq + 1, "i8") ? 1 : !strcmp(eq + 1, "u2") ? 2 : !strcmp(eq + 1, "i16") ? 3
                                                                   : (!strcmp(eq + 1, "f64") ? 0 : -1);
      continue;
    }
    if (!strcmp(tok, "potrf")) {
      o->host_potrf = !strcmp(eq + 1, "host") ? 1 : (!strcmp(eq + 1, "gpu") ? 0 : -1);
      continue;
    }
    if (!strcmp(tok, "eig")) {
      o->eig_host = !strcmp(eq + 1, "host") ? 1 : (!strcmp(eq + 1, "gpu") ? 0 : -1);
      continue;
    }
    if (!strcmp(tok, "slices")) {
      o->slices_six = !strcmp(eq + 1, "6") ? 1 : (!strcmp(eq + 1, "auto") ? 0 : -1);
      continue;
    }
    if (!strcmp(tok, "precision")) {
      o->mixed = !strcmp(eq + 1, "mixed") ? 1 : (!strcmp(eq + 1, "fp64") ? 0 : -1);
      continue;
    }
    if (!strcmp(tok, "dosage_scale")) {
      o->dosage_scale = atof(eq + 1);
      if (!(o->dosage_scale > 0.0)) o->dosage_scale = -1.0;
      continue;
    }
    long long v = atoll(eq + 1);
    if (!strcmp(tok, "block")) o->block = v;
    else if (!strcmp(tok, "device")) o->device = (int)v;
    else if (!strcmp(tok, "serial")) o->serial = (int)v;
    else if (!strcmp(tok, "stats")) o->stats = (int)v;
    else if (!strcmp(tok, "io_threads")) o->io_threads = (int)v;
    else if (!strcmp(tok, "traits_per_pass")) o->traits_per_pass = (int)v;
    else if (!strcmp(tok, "roctx")) o->no_roctx = (v == 0);
  }
}

/* Read exactly `count` doubles at element offset `start` (io.c:7-19, but a
 * short read is an error here instead of a printed warning). */
/* pread of [off, off + want) into dst; returns the bytes read (short at EOF) */
static size_t pread_full(int fd, char *dst, size_t want, off_t off) {
  size_t got = 0;
  while (got < want) {
    ssize_t r = pread(fd, dst + got, want - got, off + (off_t)got);
    if (r < 0 && errno == EINTR) continue;
    if (r <= 0) break;
    got += (size_t)r;
  }
  return got;
}

/* count doubles from element `start` of a raw FP64 file (io.c:7-19).  Reads
 * of 64 MB and more are split into 64 MB pieces read by `threads` threads
 * (the first touch of a fresh buffer and the page-cache copy run in
 * parallel: Phi.in is n^2 doubles, 7.2 GB at n = 30,000). */
static int read_exact_mt(const char *path, double *buf, size_t count, size_t start, int threads) {
  int fd = open(path, O_RDONLY);
  if (fd < 0) return fgls_set_error(FGLS_ERR_IO, "cannot open %s: %s", path, strerror(errno));
  const size_t want = count * sizeof(double);
  const off_t off = (off_t)(start * sizeof(double));
  const size_t piece = (size_t)64 << 20;
  const long npieces = (long)((want + piece - 1) / piece);
  size_t got = 0;
  if (npieces <= 1 || threads <= 1) {
    got = pread_full(fd, (char *)buf, want, off);
  } else {
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1) reduction(+ : got)
    for (long i = 0; i < npieces; ++i) {
      const size_t o = (size_t)i * piece, w = (want - o < piece) ? want - o : piece;
      got += pread_full(fd, (char *)buf + o, w, off + (off_t)o);
    }
  }
  close(fd);
  if (got != want)
    return fgls_set_error(FGLS_ERR_IO, "%s: read %zu elements - %zu expected", path,
                          got / sizeof(double), count);
  return FGLS_OK;
}

static int read_exact(const char *path, double *buf, size_t count, size_t start) {
  return read_exact_mt(path, buf, count, start, 1);
}

/* The file holds at least count doubles (the error read_exact would give). */
static int check_size(const char *path, size_t count) {
  int fd = open(path, O_RDONLY);
  if (fd < 0) return fgls_set_error(FGLS_ERR_IO, "cannot open %s: %s", path, strerror(errno));
  struct stat sb;
  const int ok = fstat(fd, &sb) == 0;
  close(fd);
  if (!ok) return fgls_set_error(FGLS_ERR_IO, "cannot stat %s: %s", path, strerror(errno));
  if ((size_t)sb.st_size < count * sizeof(double))
    return fgls_set_error(FGLS_ERR_IO, "%s: read %zu elements - %zu expected", path,
                          (size_t)sb.st_size / sizeof(double), count);
  return FGLS_OK;
}

/* count doubles of a raw FP64 file into device memory at dst, through two
 * 256 MB pinned staging buffers: `threads` threads read one buffer while the
 * other is copied (Phi.in goes to the device this way when no host copy is
 * needed; it is never held whole in host memory). */
static int upload_file(const char *path, size_t count, double *dst, int threads) {
  const size_t want = count * sizeof(double), chunk = (size_t)256 << 20, piece = (size_t)16 << 20;
  const size_t bsz = want < chunk ? want : chunk;
  char *stage[2] = {NULL, NULL};
  hipEvent_t ev[2] = {NULL, NULL};
  hipStream_t st = NULL;
  int rc = FGLS_OK;
  int fd = open(path, O_RDONLY);
  if (fd < 0) return fgls_set_error(FGLS_ERR_IO, "cannot open %s: %s", path, strerror(errno));
  if (hipStreamCreate(&st) != hipSuccess || hipEventCreateWithFlags(&ev[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&ev[1], hipEventDisableTiming) != hipSuccess) {
    rc = fgls_set_error(FGLS_ERR_HIP, "stream or event creation failed");
    goto out;
  }
  for (int s = 0; s < 2; ++s)
    if (hipHostMalloc((void **)&stage[s], bsz, hipHostMallocDefault) != hipSuccess) {
      rc = fgls_set_error(FGLS_ERR_NOMEM, "pinned allocation of %zu bytes failed", bsz);
      goto out;
    }
  for (size_t o = 0, c = 0; o < want; o += chunk, ++c) {
    const int s = (int)(c & 1);
    const size_t w = (want - o < chunk) ? want - o : chunk;
    if (c >= 2 && hipEventSynchronize(ev[s]) != hipSuccess) {
      rc = fgls_set_error(FGLS_ERR_HIP, "upload of %s failed", path);
      break;
    }
    const long np = (long)((w + piece - 1) / piece);
    size_t got = 0;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1) reduction(+ : got)
    for (long i = 0; i < np; ++i) {
      const size_t po = (size_t)i * piece, pw = (w - po < piece) ? w - po : piece;
      got += pread_full(fd, stage[s] + po, pw, (off_t)(o + po));
    }
    if (got != w) {
      rc = fgls_set_error(FGLS_ERR_IO, "%s: read %zu elements - %zu expected", path,
                          (o + got) / sizeof(double), count);
      break;
    }
    if (hipMemcpyAsync((char *)dst + o, stage[s], w, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipEventRecord(ev[s], st) != hipSuccess) {
      rc = fgls_set_error(FGLS_ERR_HIP, "upload of %s failed", path);
      break;
    }
  }
out:
  if (st) hipStreamSynchronize(st);
  for (int s = 0; s < 2; ++s) {
    if (stage[s]) hipHostFree(stage[s]);
    if (ev[s]) hipEventDestroy(ev[s]);
  }
  if (st) hipStreamDestroy(st);
  close(fd);
  return rc;
}

static int write_at(int fd, const char *name, const void *buf, size_t bytes, off_t off) {
  size_t done = 0;
  while (done < bytes) {
    ssize_t w = pwrite(fd, (const char *)buf + done, bytes - done, off + (off_t)done);
    if (w < 0 && errno == EINTR) continue;
    if (w <= 0) return fgls_set_error(FGLS_ERR_IO, "write to %s failed: %s", name, strerror(errno));
    done += (size_t)w;
  }
  return FGLS_OK;
}

#define HCHK(call)                                                                       \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess) {                                                              \
      rc = fgls_set_error(FGLS_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));  \
      goto done;                                                                         \
    }                                                                                    \
  } while (0)

#define RCHK(expr)          \
  do {                      \
    rc = (expr);            \
    if (rc != FGLS_OK) goto done; \
  } while (0)

/* ---------------------------------------------------------------- reader */
/* X_R blocks in flight: pinned read slots, device X_R buffers and B/V
 * buffers, used in turn.  Three measured 5-10 % slower than two on the
 * reference's FP64 file (the reader running further ahead competes with
 * the uploads for host memory bandwidth) and the same on compact files. */
#define NSLOT 2
typedef struct {
  int fd;
  const char *path;
  int n, wXR, t;
  size_t col_bytes;         /* bytes per X_R column: 8 n (FP64), n (int8) or ceil(n/4) (2-bit) */
  int64_t m, block, nchunks;
  void *slot[NSLOT];            /* pinned */
  sem_t free_[NSLOT], full[NSLOT];
  int status[NSLOT];            /* FGLS_OK or FGLS_ERR_IO per filled slot */
  char err[NSLOT][256];
  volatile int stop;
  int io_threads;           /* parallel preads per slot */
  int trace;                /* roctx ranges */
  double busy_s;            /* time spent in pread */
} reader_t;

static void *reader_main(void *arg) {
  reader_t *R = (reader_t *)arg;
  const size_t col_bytes = R->col_bytes;
  for (int64_t it = 0; it < (int64_t)R->t * R->nchunks; ++it) {
    const int s = (int)(it % NSLOT);
    sem_wait(&R->free_[s]);
    if (R->stop) break;
    const int64_t c = it % R->nchunks;
    const int64_t i0 = c * R->block;
    const int64_t cnt = (R->m - i0) < R->block ? (R->m - i0) : R->block;
    const size_t bytes = (size_t)cnt * R->wXR * col_bytes;
    const off_t off = (off_t)i0 * R->wXR * (off_t)col_bytes;
    double t0 = now_s();
    range_push(R->trace, "fgls: read X_R block");
    /* preads of at most 32 MB handed out to io_threads threads in turn
     * (page-cache reads are memory-copy bound per thread; small pieces keep
     * every thread busy to the end of the block): about two pieces per
     * thread, whole MBs, at least 1 MB, so that the blocks of the compact
     * files (an int8 or 2-bit block is 8 or 32 times smaller than an FP64
     * one) still spread over all threads */
    size_t got = 0;
    const int T = R->io_threads > 0 ? R->io_threads : 1;
    const size_t mb = (size_t)1 << 20;
    size_t piece = (bytes / (2 * (size_t)T) + mb - 1) / mb * mb;
    if (piece < mb) piece = mb;
    if (piece > 32 * mb) piece = 32 * mb;
    const long np = (long)((bytes + piece - 1) / piece);
#pragma omp parallel for num_threads(T) schedule(dynamic, 1) reduction(+ : got)
    for (long i = 0; i < np; ++i) {
      const size_t o = (size_t)i * piece, w = (bytes - o < piece) ? bytes - o : piece;
      got += pread_full(R->fd, (char *)R->slot[s] + o, w, off + (off_t)o);
    }
    range_pop(R->trace);
    R->busy_s += now_s() - t0;
    if (got != bytes) {
      R->status[s] = FGLS_ERR_IO;
      snprintf(R->err[s], sizeof(R->err[s]),
               "%s: data not read completely (SNPs %lld..%lld: %zu of %zu bytes)", R->path,
               (long long)i0, (long long)(i0 + cnt - 1), got, bytes);
    } else {
      R->status[s] = FGLS_OK;
    }
    sem_post(&R->full[s]);
  }
  return NULL;
}

/* ---------------------------------------------------------------- solver */
int fgls_factor_M(int n, const double *Phi, int64_t ldphi, double h2, double sigma2, double *L,
                  int64_t ldl, int threads) {
  if (n < 1 || !Phi || !L || ldphi < n || ldl < n)
    return fgls_set_error(FGLS_ERR_ARG, "fgls_factor_M: bad arguments");
  if (threads < 1) threads = host_cores();
  /* M := sigma*h*Phi; M_ii += sigma*(1-h)   (fgls_chol_gpu.c:406-414) */
  const double alpha = h2 * sigma2, beta = (1.0 - h2) * sigma2;
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int64_t c = 0; c < n; ++c) {
    const double *src = Phi + (size_t)c * ldphi;
    double *dst = L + (size_t)c * ldl;
    for (int64_t r = 0; r < n; ++r) dst[r] = alpha * src[r];
    dst[c] += beta;
  }
  scipy_openblas_set_num_threads(threads);
  int info = 0, ld = (int)ldl;
  scipy_dpotrf_("L", &n, L, &ld, &info);                 /* fgls_chol_gpu.c:416-430 */
  if (info != 0) return fgls_set_error(FGLS_ERR_NOTSPD, "dpotrf(M) failed (info: %d)", info);
  return FGLS_OK;
}

int fgls_eigen_decompose(int n, const double *Phi, int64_t ldphi, double *Z, int64_t ldz, double *W,
                         int threads) {
  if (n < 1 || !Phi || !Z || !W || ldphi < n || ldz < n)
    return fgls_set_error(FGLS_ERR_ARG, "fgls_eigen_decompose: bad arguments");
  if (threads < 1) threads = host_cores();
  /* dsyevr destroys its input: work on a copy of the lower triangle */
  double *A = (double *)malloc((size_t)n * n * sizeof(double));
  int *isuppz = (int *)malloc(2 * (size_t)n * sizeof(int));
  double *work = NULL;
  int *iwork = NULL;
  int rc = FGLS_OK;
  if (!A || !isuppz) {
    rc = fgls_set_error(FGLS_ERR_NOMEM, "host allocation for dsyevr failed");
    goto out;
  }
  for (int64_t c = 0; c < n; ++c) memcpy(A + (size_t)c * n, Phi + (size_t)c * ldphi, (size_t)n * sizeof(double));
  scipy_openblas_set_num_threads(threads);
  {
    /* eigenDec, fgls_eigen.c:780-806: dsyevr("V", "A", "L"); workspace by query */
    const double dd = -1.0;
    const int id = 0, ld = (int)ldz;
    int nfound = 0, info = 0, lwork = -1, liwork = -1, iw = 0;
    double wq = 0.0;
    scipy_dsyevr_("V", "A", "L", &n, A, &n, &dd, &dd, &id, &id, &dd, &nfound, W, Z, &ld, isuppz, &wq,
                  &lwork, &iw, &liwork, &info);
    lwork = (int)wq + 1;
    liwork = iw > 0 ? iw : 10 * n;
    work = (double *)malloc((size_t)lwork * sizeof(double));
    iwork = (int *)malloc((size_t)liwork * sizeof(int));
    if (!work || !iwork) {
      rc = fgls_set_error(FGLS_ERR_NOMEM, "host allocation for dsyevr failed");
      goto out;
    }
    scipy_dsyevr_("V", "A", "L", &n, A, &n, &dd, &dd, &id, &id, &dd, &nfound, W, Z, &ld, isuppz, work,
                  &lwork, iwork, &liwork, &info);
    if (info != 0 || nfound != n)
      rc = fgls_set_error(FGLS_ERR_NOTSPD, "dsyevr failed (info: %d)", info);
  }
out:
  free(A);
  free(isuppz);
  free(work);
  free(iwork);
  return rc;
}

/* Device bytes of one eigen-context trait beyond the shared buffers. */
static size_t eig_trait_bytes(int n, int p, int64_t block, int mixed) {
  const size_t ldn = ((size_t)n + 127) / 128 * 128;
  const size_t rt = mixed ? ldn / 32 : ldn / 128;
  size_t b = rt * (size_t)(p + 1) * (size_t)block * sizeof(double);   /* partial sums */
  b += ldn * (size_t)(2 * p + 3) * sizeof(double);                   /* weights, Y, Z'Y */
  b += (size_t)block * ((size_t)p + (size_t)p * p) * sizeof(double) * 2;   /* B, V slots */
  return b;
}

/* Device bytes of one trait's context beyond the shared buffers. */
static size_t ctx_bytes(int n, int p, int wXR, int64_t block, int mixed) {
  const size_t ldn = ((size_t)n + 127) / 128 * 128;
  const size_t rt = mixed ? ldn / 32 : ldn / 128;
  size_t b = 2 * ldn * ldn * sizeof(double);                    /* L, L^-1 */
  b += rt * (size_t)(p + 1) * (size_t)block * sizeof(double);    /* partial sums */
  b += (size_t)block * ((size_t)p + (size_t)p * p) * sizeof(double) * 2;   /* B, V slots */
  if (mixed) b += ldn * ldn * 6 + ldn * (size_t)block * wXR;    /* int8 slices, tiles */
  if (wXR > 1) b += ldn * (size_t)block * wXR * sizeof(double);  /* L^-1 X_R of a block */
  return b + (64u << 20);
}

/* The streaming solver; eig = 0: Cholesky variant, 1: eigenbasis variant. */
static int fgls_run(int eig, int n, int p, int m, int t, int wXL, int wXR, int x_b, int y_b,
                    int num_threads, const char *Phi_path, const char *h2_path,
                    const char *sigma2_path, const char *XL_path, const char *XR_path,
                    const char *Y_path, const char *B_path, const char *V_path,
                    const char *options) {
  int rc = FGLS_OK;
  options_t opt = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0.0, 0, 0};
  parse_options(options ? options : getenv("FGLS_OPTIONS"), &opt);
  if (opt.eig_host < 0) return fgls_set_error(FGLS_ERR_ARG, "eig must be gpu or host");
  if (opt.mixed < 0) return fgls_set_error(FGLS_ERR_ARG, "precision must be fp64 or mixed");
  if (opt.slices_six < 0) return fgls_set_error(FGLS_ERR_ARG, "slices must be auto or 6");
  if (opt.host_potrf < 0) return fgls_set_error(FGLS_ERR_ARG, "potrf must be gpu or host");
  if (opt.xr_i8 < 0) return fgls_set_error(FGLS_ERR_ARG, "xr_format must be f64, i8, u2 or i16");
  if (opt.dosage_scale < 0.0) return fgls_set_error(FGLS_ERR_ARG, "dosage_scale must be positive");
  const int xr16 = opt.xr_i8 == 3;   /* 16-bit dosage file */
  if (xr16 && !(opt.dosage_scale > 0.0))
    return fgls_set_error(FGLS_ERR_ARG, "xr_format=i16 needs dosage_scale");
  if (opt.dosage_scale > 0.0 && !xr16 && (!opt.mixed || opt.xr_i8))
    return fgls_set_error(FGLS_ERR_ARG, "dosage_scale needs precision=mixed and an FP64 X_R file, or "
                                        "xr_format=i16");
  const int dos = opt.dosage_scale > 0.0 && opt.mixed;
  if (opt.mixed && wXR != 1 && (dos || eig))
    return fgls_set_error(FGLS_ERR_ARG, "precision=mixed with wXR > 1 takes genotypes on the Cholesky path "
                                        "(dosages and the eigen variant need wXR = 1; got %d)", wXR);

  if (n < 1 || m < 0 || t < 1 || wXL < 0 || wXR < 1 || p != wXL + wXR || x_b < 1)
    return fgls_set_error(FGLS_ERR_ARG,
                          "invalid dimensions n=%d p=%d m=%d t=%d wXL=%d wXR=%d x_b=%d", n, p, m,
                          t, wXL, wXR, x_b);
  if (!Phi_path || !h2_path || !sigma2_path || !XR_path || !Y_path || !B_path || !V_path ||
      (wXL > 0 && !XL_path))
    return fgls_set_error(FGLS_ERR_ARG, "NULL path");
  if (y_b != 1) fprintf(stderr, "[Warning] y_b not used (set to 1)\n");

  /* SNPs per device launch: a multiple of x_b near 16384 (bounded by m);
   * near 65536 for a 2-bit file in the mixed mode, whose blocks are small
   * enough to read and upload at that size (measured 11 % faster through
   * files than 16384; the FP64 and int8 files run faster at 16384) */
  int64_t block = opt.block > 0 ? opt.block : 0;
  if (block <= 0) {
    const int64_t target = (opt.xr_i8 == 2 && opt.mixed == 1) ? 65536 : 16384;
    block = x_b;
    while (block < target && block + x_b <= (int64_t)m) block += x_b;
  }
  if (m > 0 && block > m) block = m;
  if (block < 1) block = 1;
  const int64_t nchunks = m > 0 ? ((int64_t)m + block - 1) / block : 0;
  const int ldx = (n + 3) & ~3;  /* device X_R leading dimension (32-byte rows) */
  const int64_t pp = (int64_t)p * p;
  const int threads = num_threads > 0 ? num_threads : host_cores();

  double *Phi = NULL, *M = NULL, *h = NULL, *sig = NULL, *XL = NULL, *y = NULL;
  /* per slot s and trait k of a pass: index s * K + k, views into one
   * buffer per slot (dBs[s] + k * block * p, ...) */
  double **hB = NULL, **hV = NULL, **dB = NULL, **dV = NULL;
  double *hBs[NSLOT] = {NULL}, *hVs[NSLOT] = {NULL}, *dBs[NSLOT] = {NULL}, *dVs[NSLOT] = {NULL};
  fgls_hip_eig *ectx = NULL;       /* eigen variant: one context for the K traits of a pass */
  double *Zh = NULL, *Wh = NULL, *dW = NULL, *yK = NULL;
  double t_eig = 0.0;
  unsigned long long *hErr = NULL;  /* pinned copies of the failure records */
  unsigned int *hBad = NULL;        /* pinned: genotypes of the slot that are not int8-exact */
  int8_t *dX8[NSLOT] = {NULL};
  uint8_t *dP2[NSLOT] = {NULL};   /* xr_format=u2: packed columns of the slot */
  int8_t *dX8T[NSLOT] = {NULL};   /* mixed: the slot's genotypes tiled once for every trait */
  uint8_t *dX2T[NSLOT] = {NULL};  /* mixed with a 2-bit file: the slot's 2-bit tiles */
  const int u2t = opt.mixed && opt.xr_i8 == 2;   /* 2-bit file straight to 2-bit tiles */
  const size_t col_bytes = xr16 ? (size_t)n * 2 : opt.xr_i8 == 2 ? (size_t)(n + 3) / 4
                                               : opt.xr_i8 == 1 ? (size_t)n : (size_t)n * sizeof(double);
  unsigned int *dBad = NULL;
  const int ld8 = (n + 63) & ~63;
  double *dL = NULL, *dXL = NULL, *dy = NULL, *dX[NSLOT] = {NULL};
  hipStream_t s_copy = NULL, s_comp = NULL;
  hipEvent_t ev_h2d[NSLOT] = {NULL}, ev_out[NSLOT] = {NULL}, ev_free[NSLOT] = {NULL};
  fgls_hip_ctx **ctx = NULL;
  int K = 1;                        /* traits per pass over X_R */
  reader_t R;
  memset(&R, 0, sizeof(R));
  R.fd = -1;
  int fdB = -1, fdV = -1;
  pthread_t rthread;
  int reader_started = 0;
  double t_potrf = 0.0, t_trait = 0.0, t_loop = 0.0, h2d_ms = 0.0;
  const double t_call0 = now_s();
  /* roctx ranges of this thread; an error exit closes the open ones */
  const int trace = !opt.no_roctx;
  int rdepth = 0;
#define RANGE_PUSH(name)       \
  do {                         \
    range_push(trace, (name)); \
    ++rdepth;                  \
  } while (0)
#define RANGE_POP()        \
  do {                     \
    if (rdepth > 0) {      \
      range_pop(trace);    \
      --rdepth;            \
    }                      \
  } while (0)
  RANGE_PUSH("fgls: inputs and device setup");
  hipEvent_t ev_c0[NSLOT] = {NULL}, ev_c1[NSLOT] = {NULL};   /* timed uploads */

  const size_t nn = (size_t)n * n;
  /* Phi is held on the host only for the host factor (M per trait) or the
   * host eigendecomposition (Z); otherwise it is streamed to the device */
  const int host_phi = opt.host_potrf || (eig && opt.eig_host);
  if (host_phi) {
    Phi = (double *)malloc(nn * sizeof(double));
    M = (double *)malloc(nn * sizeof(double));
  }
  h = (double *)calloc((size_t)t, sizeof(double));
  sig = (double *)calloc((size_t)t, sizeof(double));
  XL = (double *)calloc((size_t)n * (wXL > 0 ? wXL : 1), sizeof(double));
  y = (double *)calloc((size_t)n, sizeof(double));
  if ((host_phi && (!Phi || !M)) || !h || !sig || !XL || !y) {
    rc = fgls_set_error(FGLS_ERR_NOMEM, "host allocation of %zu bytes failed", 2 * nn * sizeof(double));
    goto done;
  }
  if (host_phi) RCHK(read_exact_mt(Phi_path, Phi, nn, 0, threads));
  else RCHK(check_size(Phi_path, nn));
  RCHK(read_exact(h2_path, h, (size_t)t, 0));
  RCHK(read_exact(sigma2_path, sig, (size_t)t, 0));
  if (wXL > 0) RCHK(read_exact(XL_path, XL, (size_t)wXL * n, 0));

  R.fd = open(XR_path, O_RDONLY);
  if (R.fd < 0) {
    rc = fgls_set_error(FGLS_ERR_IO, "cannot open %s: %s", XR_path, strerror(errno));
    goto done;
  }
  fdB = open(B_path, O_WRONLY | O_CREAT | O_TRUNC, 0644);
  fdV = open(V_path, O_WRONLY | O_CREAT | O_TRUNC, 0644);
  if (fdB < 0 || fdV < 0) {
    rc = fgls_set_error(FGLS_ERR_IO, "cannot create %s / %s: %s", B_path, V_path, strerror(errno));
    goto done;
  }

  /* device state: K trait contexts share each X_R block (one read per pass);
   * K is all t traits when they fit in 60 % of the free device memory */
  HCHK(hipSetDevice(opt.device));
  {
    size_t free_b = 0, total_b = 0;
    HCHK(hipMemGetInfo(&free_b, &total_b));
    const size_t ldn = ((size_t)n + 127) / 128 * 128;
    const size_t per = eig ? eig_trait_bytes(n, p, block, opt.mixed) : ctx_bytes(n, p, wXR, block, opt.mixed);
    size_t shared = (size_t)block * wXR * ldx * sizeof(double) * NSLOT + nn * sizeof(double);
    if (eig) shared += ldn * ldn * (sizeof(double) + (opt.mixed ? 6 : 0)) + (64u << 20);
    if (eig && wXR > 1) shared += ldn * (size_t)block * wXR * sizeof(double);   /* Z' X_R of a block */
    long long fit = free_b > shared ? (long long)((free_b - shared) * 0.6 / per) : 1;
    if (fit < 1) fit = 1;
    K = opt.traits_per_pass > 0 ? opt.traits_per_pass : (int)(fit < t ? fit : t);
    if (K > t) K = t;
  }
  ctx = (fgls_hip_ctx **)calloc((size_t)K, sizeof(*ctx));
  hB = (double **)calloc((size_t)NSLOT * K, sizeof(double *));
  hV = (double **)calloc((size_t)NSLOT * K, sizeof(double *));
  dB = (double **)calloc((size_t)NSLOT * K, sizeof(double *));
  dV = (double **)calloc((size_t)NSLOT * K, sizeof(double *));
  if (!ctx || !hB || !hV || !dB || !dV) {
    rc = fgls_set_error(FGLS_ERR_NOMEM, "host allocation failed");
    goto done;
  }
  if (eig) {
    RCHK(fgls_hip_eig_create(&ectx, n, wXL, K, (int)(block > 0 ? block : 1), opt.device));
    if (opt.mixed) RCHK(fgls_hip_eig_set_mode(ectx, FGLS_MODE_MIXED_I8));
  } else {
    for (int k = 0; k < K; ++k) {
      RCHK(fgls_hip_ctx_create(&ctx[k], n, wXL, wXR, (int)(block > 0 ? block : 1), opt.device));
      if (opt.mixed) {
        if (opt.slices_six) RCHK(fgls_hip_ctx_set_slice_form(ctx[k], FGLS_SLICES_SIX));
        RCHK(fgls_hip_ctx_set_mode(ctx[k], FGLS_MODE_MIXED_I8));
      }
    }
  }
  HCHK(hipStreamCreateWithFlags(&s_copy, hipStreamNonBlocking));
  HCHK(hipStreamCreateWithFlags(&s_comp, hipStreamNonBlocking));
  for (int s = 0; s < NSLOT; ++s) {
    HCHK(hipEventCreateWithFlags(&ev_h2d[s], hipEventDisableTiming));
    HCHK(hipEventCreate(&ev_c0[s]));
    HCHK(hipEventCreate(&ev_c1[s]));
    HCHK(hipEventCreateWithFlags(&ev_out[s], hipEventDisableTiming));
    HCHK(hipEventCreateWithFlags(&ev_free[s], hipEventDisableTiming));
  }
  /* device potrf: Phi is uploaded once and every trait's M is built and
   * factored on the device; host potrf: each trait's factor is uploaded */
  HCHK(hipMalloc((void **)&dL, nn * sizeof(double)));
  if (!opt.host_potrf && !eig) RCHK(upload_file(Phi_path, nn, dL, threads));
  HCHK(hipMalloc((void **)&dXL, (size_t)n * (wXL > 0 ? wXL : 1) * sizeof(double)));
  HCHK(hipMalloc((void **)&dy, (size_t)n * K * sizeof(double)));
  if (wXL > 0) HCHK(hipMemcpy(dXL, XL, (size_t)n * wXL * sizeof(double), hipMemcpyHostToDevice));
  if (eig) {
    /* Phi = Z diag(W) Z' once (eigenDec, fgls_eigen.c:780-806): by default on
     * the device (rocSOLVER dsyevd), with eig=host by LAPACK dsyevr on the
     * host as the reference; Z and W stay on the device for every trait */
    double te = now_s();
    Zh = M;
    Wh = (double *)malloc((size_t)n * sizeof(double));
    yK = (double *)malloc((size_t)n * K * sizeof(double));
    if (!Wh || !yK) {
      rc = fgls_set_error(FGLS_ERR_NOMEM, "host allocation failed");
      goto done;
    }
    HCHK(hipMalloc((void **)&dW, (size_t)n * sizeof(double)));
    if (opt.eig_host) {
      RCHK(fgls_eigen_decompose(n, Phi, n, Zh, n, Wh, threads));
      HCHK(hipMemcpy(dL, Zh, nn * sizeof(double), hipMemcpyHostToDevice));
      HCHK(hipMemcpy(dW, Wh, (size_t)n * sizeof(double), hipMemcpyHostToDevice));
    } else {
      double *dPhi = NULL;
      HCHK(hipMalloc((void **)&dPhi, nn * sizeof(double)));
      rc = upload_file(Phi_path, nn, dPhi, threads);
      if (rc == FGLS_OK) rc = fgls_hip_eigen_decompose(n, dPhi, n, dL, n, dW, opt.device, s_comp);
      hipFree(dPhi);
      if (rc != FGLS_OK) goto done;
    }
    t_eig = now_s() - te;
    RCHK(fgls_hip_eig_set_basis(ectx, dL, n, dW, dXL, n, s_comp));
  }
  HCHK(hipHostMalloc((void **)&hErr, (size_t)NSLOT * K * sizeof(unsigned long long), hipHostMallocDefault));
  for (int i = 0; i < NSLOT * K; ++i) hErr[i] = ~0ull;
  HCHK(hipHostMalloc((void **)&hBad, NSLOT * sizeof(unsigned int), hipHostMallocDefault));
  for (int s = 0; s < NSLOT; ++s) hBad[s] = 0;
  if ((opt.mixed || opt.xr_i8) && nchunks > 0) {
    HCHK(hipMalloc((void **)&dBad, NSLOT * sizeof(unsigned int)));
    for (int s = 0; s < NSLOT; ++s) {
      /* rows n..ld8-1 stay zero: uploads and conversions write rows < n;
       * dosages take two digit planes */
      const size_t b8 = (size_t)block * wXR * ld8 * (dos ? 2 : 1);
      HCHK(hipMalloc((void **)&dX8[s], b8));
      HCHK(hipMemset(dX8[s], 0, b8));
      if (opt.xr_i8 >= 2) HCHK(hipMalloc((void **)&dP2[s], (size_t)block * wXR * col_bytes));
      if (u2t) HCHK(hipMalloc((void **)&dX2T[s], fgls_hip_x2t_bytes(n, (int)block * wXR)));
      else if (opt.mixed) HCHK(hipMalloc((void **)&dX8T[s], fgls_hip_x8t_bytes(n, (int)block * wXR, dos ? 2 : 1)));
    }
  }
  if (nchunks > 0) {
    const size_t xbytes = (size_t)block * wXR * ldx * sizeof(double);
    for (int s = 0; s < NSLOT; ++s) {
      HCHK(hipMalloc((void **)&dX[s], xbytes));
      HCHK(hipMemset(dX[s], 0, xbytes));
      HCHK(hipHostMalloc((void **)&R.slot[s], (size_t)block * wXR * col_bytes, hipHostMallocDefault));
      HCHK(hipMalloc((void **)&dBs[s], (size_t)K * block * p * sizeof(double)));
      HCHK(hipMalloc((void **)&dVs[s], (size_t)K * block * pp * sizeof(double)));
      HCHK(hipHostMalloc((void **)&hBs[s], (size_t)K * block * p * sizeof(double), hipHostMallocDefault));
      HCHK(hipHostMalloc((void **)&hVs[s], (size_t)K * block * pp * sizeof(double), hipHostMallocDefault));
      for (int k = 0; k < K; ++k) {
        const int q = s * K + k;
        dB[q] = dBs[s] + (size_t)k * block * p;
        dV[q] = dVs[s] + (size_t)k * block * pp;
        hB[q] = hBs[s] + (size_t)k * block * p;
        hV[q] = hVs[s] + (size_t)k * block * pp;
      }
    }
    R.path = XR_path;
    R.n = n;
    R.col_bytes = col_bytes;
    R.wXR = wXR;
    R.t = (t + K - 1) / K;   /* passes over X_R */
    R.m = m;
    R.block = block;
    R.nchunks = nchunks;
    R.io_threads = opt.io_threads > 0 ? opt.io_threads : (host_cores() < 16 ? host_cores() : 16);
    R.trace = trace;
    for (int s = 0; s < NSLOT; ++s) {
      sem_init(&R.free_[s], 0, 1);
      sem_init(&R.full[s], 0, 0);
    }
    if (pthread_create(&rthread, NULL, reader_main, &R) != 0) {
      rc = fgls_set_error(FGLS_ERR_IO, "cannot start the reader thread");
      goto done;
    }
    reader_started = 1;
  }
  scipy_openblas_set_num_threads(threads);
  RANGE_POP();

  int64_t it = 0;  /* global chunk counter across passes (reader order) */
  for (int j0 = 0; j0 < t; j0 += K) {
    const int kn = (t - j0) < K ? (t - j0) : K;   /* traits of this pass */
    RANGE_PUSH("fgls: trait setup");
    if (eig) {
      /* Z' Y, weights, B_t, V_tl and res_sigma of the pass's traits */
      double t0 = now_s();
      RCHK(read_exact(Y_path, yK, (size_t)n * kn, (size_t)j0 * n));
      HCHK(hipMemcpyAsync(dy, yK, (size_t)n * kn * sizeof(double), hipMemcpyHostToDevice, s_comp));
      RCHK(fgls_hip_eig_set_traits(ectx, kn, h + j0, sig + j0, dy, n, s_comp));
      t_trait += now_s() - t0;
    }
    for (int k = 0; k < kn && !eig; ++k) {
      const int j = j0 + k;
      double t0 = now_s();
      if (opt.host_potrf) {
        if ((rc = fgls_factor_M(n, Phi, n, h[j], sig[j], M, n, threads)) != FGLS_OK) goto done;
        t_potrf += now_s() - t0;
        t0 = now_s();
        RCHK(read_exact(Y_path, y, (size_t)n, (size_t)j * n));
        HCHK(hipMemcpyAsync(dL, M, nn * sizeof(double), hipMemcpyHostToDevice, s_comp));
        HCHK(hipMemcpyAsync(dy, y, (size_t)n * sizeof(double), hipMemcpyHostToDevice, s_comp));
        RCHK(fgls_hip_set_trait(ctx[k], dL, n, dXL, n, dy, s_comp));
      } else {
        RCHK(read_exact(Y_path, y, (size_t)n, (size_t)j * n));
        HCHK(hipMemcpyAsync(dy, y, (size_t)n * sizeof(double), hipMemcpyHostToDevice, s_comp));
        RCHK(fgls_hip_set_trait_phi(ctx[k], dL, n, h[j], sig[j], dXL, n, dy, s_comp));
      }
      t_trait += now_s() - t0;
    }
    RANGE_POP();

    double tl0 = now_s();
    RANGE_PUSH("fgls: block loop");
    for (int64_t c = 0; c <= nchunks; ++c, ++it) {
      const int s = (int)(it % NSLOT);
      if (c < nchunks) {
        RANGE_PUSH("fgls: upload and enqueue block");
        const int64_t i0 = c * block;
        const int cnt = (int)((m - i0) < block ? (m - i0) : block);
        sem_wait(&R.full[s]);
        if (R.status[s] != FGLS_OK) {
          rc = fgls_set_error(FGLS_ERR_IO, "%s", R.err[s]);
          goto done;
        }
        /* device X_R slot s is free once the solves of chunk c - NSLOT have run */
        HCHK(hipStreamWaitEvent(s_copy, ev_free[s], 0));
        HCHK(hipEventRecord(ev_c0[s], s_copy));
        if (opt.xr_i8 >= 2)
          HCHK(hipMemcpyAsync(dP2[s], R.slot[s], (size_t)cnt * wXR * col_bytes, hipMemcpyHostToDevice, s_copy));
        else if (opt.xr_i8)
          HCHK(hipMemcpy2DAsync(dX8[s], (size_t)ld8, R.slot[s], (size_t)n, (size_t)n,
                                (size_t)cnt * wXR, hipMemcpyHostToDevice, s_copy));
        else
          HCHK(hipMemcpy2DAsync(dX[s], (size_t)ldx * sizeof(double), R.slot[s],
                                (size_t)n * sizeof(double), (size_t)n * sizeof(double),
                                (size_t)cnt * wXR, hipMemcpyHostToDevice, s_copy));
        HCHK(hipEventRecord(ev_c1[s], s_copy));
        HCHK(hipEventRecord(ev_h2d[s], s_copy));
        HCHK(hipStreamWaitEvent(s_comp, ev_h2d[s], 0));
        if (u2t) {
          /* 2-bit codes -> the product's 2-bit tiles; code 3 is counted and
           * fails the run at write-back */
          HCHK(hipMemsetAsync(dBad + s, 0, sizeof(unsigned int), s_comp));
          RCHK(fgls_hip_tile_u2_packed(dP2[s], (int64_t)col_bytes, n, cnt * wXR, dX2T[s], dBad + s, s_comp));
          HCHK(hipMemcpyAsync(&hBad[s], dBad + s, sizeof(unsigned int), hipMemcpyDeviceToHost, s_comp));
        } else if (opt.xr_i8 == 2) {
          /* 2-bit codes -> int8 on the device; code 3 is counted and fails
           * the run at write-back */
          HCHK(hipMemsetAsync(dBad + s, 0, sizeof(unsigned int), s_comp));
          RCHK(fgls_hip_u2_to_i8(dP2[s], (int64_t)col_bytes, n, cnt * wXR, dX8[s], ld8, dBad + s, s_comp));
          HCHK(hipMemcpyAsync(&hBad[s], dBad + s, sizeof(unsigned int), hipMemcpyDeviceToHost, s_comp));
        } else if (xr16 && opt.mixed) {
          /* 16-bit dosages -> the two int8 digit planes; out-of-range values fail the run */
          HCHK(hipMemsetAsync(dBad + s, 0, sizeof(unsigned int), s_comp));
          RCHK(fgls_hip_i16_to_dosage_i8((const int16_t *)dP2[s], n, n, cnt, dX8[s], ld8, dBad + s, s_comp));
          HCHK(hipMemcpyAsync(&hBad[s], dBad + s, sizeof(unsigned int), hipMemcpyDeviceToHost, s_comp));
        } else if (xr16) {
          RCHK(fgls_hip_i16_to_f64((const int16_t *)dP2[s], n, n, cnt * wXR, opt.dosage_scale, dX[s], ldx, s_comp));
        }
        if (opt.mixed && !opt.xr_i8) {
          /* exact FP64 -> int8 on the device; a value that is not an integer
           * in [-127, 127] is counted and fails the run at write-back */
          HCHK(hipMemsetAsync(dBad + s, 0, sizeof(unsigned int), s_comp));
          if (dos)
            RCHK(fgls_hip_to_dosage_i8(dX[s], ldx, n, cnt, opt.dosage_scale, dX8[s], ld8, dBad + s, s_comp));
          else
            RCHK(fgls_hip_to_i8(dX[s], ldx, n, cnt * wXR, dX8[s], ld8, dBad + s, s_comp));
          HCHK(hipMemcpyAsync(&hBad[s], dBad + s, sizeof(unsigned int), hipMemcpyDeviceToHost, s_comp));
        } else if (!opt.mixed && (opt.xr_i8 == 1 || opt.xr_i8 == 2)) {
          RCHK(fgls_hip_i8_to_f64(dX8[s], ld8, n, cnt * wXR, dX[s], ldx, s_comp));
        } else if (opt.mixed && opt.xr_i8 == 1) {
          /* an int8 file goes to the product as is: -128 is out of its range */
          HCHK(hipMemsetAsync(dBad + s, 0, sizeof(unsigned int), s_comp));
          RCHK(fgls_hip_check_i8(dX8[s], ld8, n, cnt * wXR, dBad + s, s_comp));
          HCHK(hipMemcpyAsync(&hBad[s], dBad + s, sizeof(unsigned int), hipMemcpyDeviceToHost, s_comp));
        }
        const int planes = dos ? 2 : 1;
        if (opt.mixed && !u2t) RCHK(fgls_hip_tile_i8(dX8[s], ld8, n, cnt * wXR, planes, dX8T[s], s_comp));
        if (eig) {
          const int64_t ldb = block * p, ldv = block * pp;
          if (u2t)
            RCHK(fgls_hip_eig_solve_block_u2t(ectx, dX2T[s], cnt, dBs[s], ldb, dVs[s], ldv, i0, s_comp));
          else if (opt.mixed)
            RCHK(fgls_hip_eig_solve_block_i8t(ectx, dX8T[s], cnt, planes, dos ? opt.dosage_scale : 1.0, dBs[s],
                                              ldb, dVs[s], ldv, i0, s_comp));
          else if (wXR > 1)
            RCHK(fgls_hip_eig_solve_block_wide(ectx, dX[s], ldx, cnt, wXR, dBs[s], ldb, dVs[s], ldv, i0,
                                               s_comp));
          else
            RCHK(fgls_hip_eig_solve_block(ectx, dX[s], ldx, cnt, dBs[s], ldb, dVs[s], ldv, i0, s_comp));
          HCHK(hipMemcpyAsync(&hErr[s * K], fgls_eig_err_ptr(ectx), (size_t)kn * sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, s_comp));
        }
        for (int k = 0; k < kn; ++k) {
          const int q = s * K + k;
          const int64_t base = (int64_t)(j0 + k) * m + i0;
          if (eig) {
            HCHK(hipMemcpyAsync(hB[q], dB[q], (size_t)cnt * p * sizeof(double), hipMemcpyDeviceToHost, s_comp));
            HCHK(hipMemcpyAsync(hV[q], dV[q], (size_t)cnt * pp * sizeof(double), hipMemcpyDeviceToHost, s_comp));
            continue;
          }
          if (u2t)
            RCHK(fgls_hip_solve_block_u2t(ctx[k], dX2T[s], cnt, dB[q], dV[q], base, s_comp));
          else if (opt.mixed)
            RCHK(fgls_hip_solve_block_i8t(ctx[k], dX8T[s], cnt, planes, dos ? opt.dosage_scale : 1.0, dB[q],
                                          dV[q], base, s_comp));
          else
            RCHK(fgls_hip_solve_block(ctx[k], dX[s], ldx, cnt, dB[q], dV[q], base, s_comp));
          HCHK(hipMemcpyAsync(hB[q], dB[q], (size_t)cnt * p * sizeof(double), hipMemcpyDeviceToHost, s_comp));
          HCHK(hipMemcpyAsync(hV[q], dV[q], (size_t)cnt * pp * sizeof(double), hipMemcpyDeviceToHost, s_comp));
          HCHK(hipMemcpyAsync(&hErr[q], fgls_ctx_err_ptr(ctx[k]), sizeof(unsigned long long),
                              hipMemcpyDeviceToHost, s_comp));
        }
        HCHK(hipEventRecord(ev_free[s], s_comp));
        HCHK(hipEventRecord(ev_out[s], s_comp));
        if (opt.serial) HCHK(hipDeviceSynchronize());
        /* the pinned read slot can be refilled once its upload has landed */
        HCHK(hipEventSynchronize(ev_h2d[s]));
        {
          float ms = 0.0f;
          HCHK(hipEventElapsedTime(&ms, ev_c0[s], ev_c1[s]));
          h2d_ms += ms;
        }
        sem_post(&R.free_[s]);
        RANGE_POP();
      }
      if (c > 0) {  /* write back chunk c-1 */
        RANGE_PUSH("fgls: write back block");
        const int ps = (int)((it - 1) % NSLOT);
        const int64_t i0 = (c - 1) * block;
        const int cnt = (int)((m - i0) < block ? (m - i0) : block);
        HCHK(hipEventSynchronize(ev_out[ps]));
        if (hBad[ps] != 0) {
          if (opt.xr_i8 == 2)
            rc = fgls_set_error(FGLS_ERR_ARG, "%s: %u 2-bit codes of SNPs %lld.. are 3 (not a genotype)",
                                XR_path, hBad[ps], (long long)i0);
          else if (xr16)
            rc = fgls_set_error(FGLS_ERR_ARG, "%s: %u 16-bit dosage values of SNPs %lld.. are outside "
                                "[-16256, 16383]", XR_path, hBad[ps], (long long)i0);
          else if (dos)
            rc = fgls_set_error(FGLS_ERR_ARG,
                                "precision=mixed: %u dosage values of SNPs %lld.. are not multiples of "
                                "1/%g in range (use precision=fp64)", hBad[ps], (long long)i0,
                                opt.dosage_scale);
          else
            rc = fgls_set_error(FGLS_ERR_ARG,
                                "precision=mixed: %u genotype values of SNPs %lld.. are not integers in "
                                "[-127, 127] (use precision=fp64)", hBad[ps], (long long)i0);
          goto done;
        }
        for (int k = 0; k < kn; ++k) {
          const int q = ps * K + k;
          const int j = j0 + k;
          if (hErr[q] != ~0ull) {  /* first failing SNP so far (fgls_chol_gpu.c:664-670) */
            const long long bad = (long long)(hErr[q] >> 8) - (eig ? 0 : (long long)j * m);
            rc = fgls_set_error(FGLS_ERR_NOTSPD, "dpotrf(V) failed (info: %d) - i: %lld",
                                (int)(hErr[q] & 0xff), bad);
            goto done;
          }
          RCHK(write_at(fdB, B_path, hB[q], (size_t)cnt * p * sizeof(double),
                        (off_t)(((int64_t)j * m + i0) * p) * (off_t)sizeof(double)));
          RCHK(write_at(fdV, V_path, hV[q], (size_t)cnt * pp * sizeof(double),
                        (off_t)(((int64_t)j * m + i0) * pp) * (off_t)sizeof(double)));
        }
        RANGE_POP();
      }
    }
    --it;  /* the c == nchunks pass consumed no slot */
    t_loop += now_s() - tl0;
    RANGE_POP();
  }
  HCHK(hipStreamSynchronize(s_comp));
  {
    fgls_run_stats st;
    memset(&st, 0, sizeof(st));
    st.total_s = now_s() - t_call0;
    st.factor_s = eig ? t_eig : t_potrf;
    st.trait_s = t_trait;
    st.loop_s = t_loop;
    st.read_busy_s = R.busy_s;
    st.h2d_busy_s = h2d_ms * 1e-3;
    st.xr_bytes = (double)m * R.t * wXR * (double)col_bytes;
    st.h2d_bytes = st.xr_bytes;
    st.trait_snps = (int64_t)m * t;
    st.block = block;
    st.passes = R.t;
    st.traits_per_pass = K;
    st.io_threads = R.io_threads;
    g_last_stats = st;
  }
  if (opt.stats) {
    const double snps = (double)m * t;
    fprintf(stderr,
            "[fgls] %s n=%d p=%d m=%d t=%d block=%lld traits/pass=%d: %s %.3fs, trait setup %.3fs, "
            "block loop %.3fs (%.0f SNPs/s, %.2f GB/s X_R), reader busy %.3fs, threads %d\n",
            eig ? "eigen" : "chol", n, p, m, t, (long long)block, K,
            eig ? (opt.eig_host ? "host dsyevr" : "device dsyevd") : "host potrf",
            eig ? t_eig : t_potrf, t_trait, t_loop, t_loop > 0 ? snps / t_loop : 0.0,
            t_loop > 0 ? (double)m * R.t * wXR * R.col_bytes / t_loop / 1e9 : 0.0, R.busy_s, threads);
  }

done:
  while (rdepth > 0) RANGE_POP();
#undef RANGE_PUSH
#undef RANGE_POP
  if (reader_started) {
    R.stop = 1;
    for (int s = 0; s < NSLOT; ++s) sem_post(&R.free_[s]);
    pthread_join(rthread, NULL);
  }
  if (nchunks > 0 && R.slot[0] != NULL) {
    for (int s = 0; s < NSLOT; ++s) {
      sem_destroy(&R.free_[s]);
      sem_destroy(&R.full[s]);
    }
  }
  if (s_comp) hipStreamSynchronize(s_comp);
  if (s_copy) hipStreamSynchronize(s_copy);
  for (int s = 0; s < NSLOT; ++s) {
    if (ev_h2d[s]) hipEventDestroy(ev_h2d[s]);
    if (ev_c0[s]) hipEventDestroy(ev_c0[s]);
    if (ev_c1[s]) hipEventDestroy(ev_c1[s]);
    if (ev_out[s]) hipEventDestroy(ev_out[s]);
    if (ev_free[s]) hipEventDestroy(ev_free[s]);
    if (dX[s]) hipFree(dX[s]);
    if (R.slot[s]) hipHostFree(R.slot[s]);
    if (dX8[s]) hipFree(dX8[s]);
    if (dP2[s]) hipFree(dP2[s]);
    if (dX8T[s]) hipFree(dX8T[s]);
    if (dX2T[s]) hipFree(dX2T[s]);
  }
  for (int s = 0; s < NSLOT; ++s) {
    if (dBs[s]) hipFree(dBs[s]);
    if (dVs[s]) hipFree(dVs[s]);
    if (hBs[s]) hipHostFree(hBs[s]);
    if (hVs[s]) hipHostFree(hVs[s]);
  }
  free(dB);
  free(dV);
  free(hB);
  free(hV);
  if (hErr) hipHostFree(hErr);
  if (hBad) hipHostFree(hBad);
  if (dBad) hipFree(dBad);
  if (s_copy) hipStreamDestroy(s_copy);
  if (s_comp) hipStreamDestroy(s_comp);
  if (dL) hipFree(dL);
  if (dXL) hipFree(dXL);
  if (dy) hipFree(dy);
  if (ctx) {
    for (int k = 0; k < K; ++k)
      if (ctx[k]) fgls_hip_ctx_destroy(ctx[k]);
    free(ctx);
  }
  if (ectx) fgls_hip_eig_destroy(ectx);
  if (dW) hipFree(dW);
  free(Wh);
  free(yK);
  if (R.fd >= 0) close(R.fd);
  if (fdB >= 0) close(fdB);
  if (fdV >= 0) close(fdV);
  free(Phi);
  if (M != Phi) free(M);
  free(h);
  free(sig);
  free(XL);
  free(y);
  return rc;
}

int fgls_chol_gpu_ex(int n, int p, int m, int t, int wXL, int wXR, int x_b, int y_b,
                     int num_threads, const char *Phi_path, const char *h2_path,
                     const char *sigma2_path, const char *XL_path, const char *XR_path,
                     const char *Y_path, const char *B_path, const char *V_path,
                     const char *options) {
  return fgls_run(0, n, p, m, t, wXL, wXR, x_b, y_b, num_threads, Phi_path, h2_path, sigma2_path,
                  XL_path, XR_path, Y_path, B_path, V_path, options);
}

int fgls_eigen_gpu_ex(int n, int p, int m, int t, int wXL, int wXR, int x_b, int y_b,
                      int num_threads, const char *Phi_path, const char *h2_path,
                      const char *sigma2_path, const char *XL_path, const char *XR_path,
                      const char *Y_path, const char *B_path, const char *V_path,
                      const char *options) {
  return fgls_run(1, n, p, m, t, wXL, wXR, x_b, y_b, num_threads, Phi_path, h2_path, sigma2_path,
                  XL_path, XR_path, Y_path, B_path, V_path, options);
}

int fgls_eigen(int n, int p, int m, int t, int wXL, int wXR, int x_b, int y_b, int num_threads,
               char *Phi_path, char *h2_path, char *sigma2_path, char *XL_path, char *XR_path,
               char *Y_path, char *B_path, char *V_path) {
  int rc = fgls_eigen_gpu_ex(n, p, m, t, wXL, wXR, x_b, y_b, num_threads, Phi_path, h2_path,
                             sigma2_path, XL_path, XR_path, Y_path, B_path, V_path, NULL);
  if (rc != FGLS_OK) {
    fprintf(stderr, "[Error]: %s\n", fgls_hip_last_error());   /* common.c:11-17 */
    exit(EXIT_FAILURE);
  }
  return 0;
}

int fgls_chol_gpu(int n, int p, int m, int t, int wXL, int wXR, int x_b, int y_b, int num_threads,
                  char *Phi_path, char *h2_path, char *sigma2_path, char *XL_path, char *XR_path,
                  char *Y_path, char *B_path, char *V_path) {
  int rc = fgls_chol_gpu_ex(n, p, m, t, wXL, wXR, x_b, y_b, num_threads, Phi_path, h2_path,
                            sigma2_path, XL_path, XR_path, Y_path, B_path, V_path, NULL);
  if (rc != FGLS_OK) {
    fprintf(stderr, "[Error]: %s\n", fgls_hip_last_error());   /* common.c:11-17 */
    exit(EXIT_FAILURE);
  }
  return 0;
}
