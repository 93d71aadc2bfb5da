/*
 * driver.c -- the reference's command-line contract (src/driver.c:15-157)
 * over the MI355X solver.
 *
 *   driver.x <variant> <data_set_directory> <#repetitions> [--check]
 *   stdin:   n p m t wXL x_b y_b num_threads          (driver.c:35-49)
 *
 * Reads <dir>/{Phi,H,Sig,XL,XR,Y}.in, writes <dir>/{B,V}.out
 * (driver.c:56-66) and prints the mean wall milliseconds per repetition as
 * the last stdout line (driver.c:98-100, get_diff_ms rounding).
 *
 * Variants: 'g' (the GPU solver, driver.c:78-86) and 'c' run fgls_chol_gpu;
 * 'e' runs fgls_eigen (driver.c:70-77), the eigenbasis variant on the GPU.
 * --check reproduces the -DDEBUG comparison (driver.c:102-154): the max
 * element-wise relative error of B.out against B_exp.out and of V.out
 * against V_exp.out when those files exist (V over its lower triangle).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "fgls_chol_gpu.h"
#include "fgls_eigen.h"

#define PATH_LEN 4096

static long diff_ms(struct timeval *a, struct timeval *b) {
  long s = b->tv_sec - a->tv_sec;
  long us = b->tv_usec - a->tv_usec;
  return (long)((s * 1000 + us / 1000.0) + 0.5);
}

static double *read_all(const char *path, size_t count) {
  FILE *f = fopen(path, "rb");
  if (!f) return NULL;
  double *buf = (double *)malloc(count * sizeof(double));
  size_t got = buf ? fread(buf, sizeof(double), count, f) : 0;
  fclose(f);
  if (got != count) {
    free(buf);
    return NULL;
  }
  return buf;
}

/* driver.c:159-177 over the selected entries (every entry, or the lower
 * triangle of each p x p block when p > 0). */
static double compare(const double *a, const double *b, size_t size, int p) {
  double maxerr = 0.0;
  for (size_t i = 0; i < size; ++i) {
    if (p > 0) {
      size_t e = i % ((size_t)p * p);
      size_t c = e / p, r = e % p;
      if (r < c) continue;
    }
    double rel = fabs((a[i] - b[i]) / b[i]);
    if (rel > maxerr) maxerr = rel;
  }
  return maxerr;
}

int main(int argc, char *argv[]) {
  if (argc != 4 && !(argc == 5 && strcmp(argv[4], "--check") == 0)) {
    fprintf(stderr, "Usage: %s <variant - [g|c|e]> <data_set_directory> <#repetitions> [--check]\n", argv[0]);
    fprintf(stderr, "\tnote: data_set_directory must contain files:\n");
    fprintf(stderr, "- Phi.in\n- H.in\n- Sig.in\n- XL.in\n- XR.in\n- Y.in\n");
    return EXIT_FAILURE;
  }
  const char var = argv[1][0];
  const char *dir = argv[2];
  const int nrep = atoi(argv[3]);
  int n, p, m, t, wXL, x_b, y_b, nths;
  if (scanf("%d %d %d %d %d %d %d %d", &n, &p, &m, &t, &wXL, &x_b, &y_b, &nths) != 8) {
    fprintf(stderr, "[Error]: expected 8 integers on stdin: n p m t wXL x_b y_b threads\n");
    return EXIT_FAILURE;
  }
  const int wXR = p - wXL;
  char Phi[PATH_LEN], H[PATH_LEN], Sig[PATH_LEN], XL[PATH_LEN], XR[PATH_LEN], Y[PATH_LEN],
      B[PATH_LEN], V[PATH_LEN];
  snprintf(XL, PATH_LEN, "%s/XL.in", dir);
  snprintf(XR, PATH_LEN, "%s/XR.in", dir);
  snprintf(Y, PATH_LEN, "%s/Y.in", dir);
  snprintf(Phi, PATH_LEN, "%s/Phi.in", dir);
  snprintf(H, PATH_LEN, "%s/H.in", dir);
  snprintf(Sig, PATH_LEN, "%s/Sig.in", dir);
  snprintf(B, PATH_LEN, "%s/B.out", dir);
  snprintf(V, PATH_LEN, "%s/V.out", dir);

  struct timeval start, end;
  gettimeofday(&start, NULL);
  for (int i = 0; i < nrep; ++i) {
    if (var == 'e')
      fgls_eigen(n, p, m, t, wXL, wXR, x_b, y_b, nths, Phi, H, Sig, XL, XR, Y, B, V);
    else
      fgls_chol_gpu(n, p, m, t, wXL, wXR, x_b, y_b, nths, Phi, H, Sig, XL, XR, Y, B, V);
  }
  gettimeofday(&end, NULL);
  long total = nrep > 0 ? diff_ms(&start, &end) / nrep : 0;

  if (argc == 5) {
    char e1[PATH_LEN];
    size_t nb = (size_t)m * t * p, nv = (size_t)m * t * p * p;
    snprintf(e1, PATH_LEN, "%s/B_exp.out", dir);
    double *mine = read_all(B, nb), *exp = read_all(e1, nb);
    if (mine && exp) printf("Checking B\nMax elemental diff: %.16e\n", compare(mine, exp, nb, 0));
    else printf("Checking B\nERROR opening %s and/or %s\n", B, e1);
    free(mine);
    free(exp);
    snprintf(e1, PATH_LEN, "%s/V_exp.out", dir);
    mine = read_all(V, nv);
    exp = read_all(e1, nv);
    if (mine && exp) printf("Checking V\nMax elemental diff: %.16e\n", compare(mine, exp, nv, p));
    else printf("Checking V\nERROR opening %s and/or %s\n", V, e1);
    free(mine);
    free(exp);
  }
  printf("%ld\n", total);
  return 0;
}
