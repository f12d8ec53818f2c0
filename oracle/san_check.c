/*
 * san_check.c -- drives every routine of fdtd_oracle.c over small grids with
 * exactly-sized heap arrays, for an AddressSanitizer / UBSan build
 * (`make -C oracle san`).  TEST INFRASTRUCTURE ONLY.
 *
 * The cases reach the places where the restatement clamps or skips indices:
 * sources and receivers on the first and last cells of every axis, grids with
 * nz > nx (the A2 read of b_x_half at j >= nx, seismicfdtd.f95:255), runs
 * split into two segments that share one psi workspace, the split C-PML
 * variant, attenuation and all six stress-rate series.  Any out-of-bounds
 * access or undefined behaviour aborts the program with a sanitizer report;
 * a clean run prints one line per case and exits 0.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fdtd_oracle.h"

static double* arr(size_t n, double v) {
  double* a = (double*)malloc(n * sizeof(double));
  if (!a) {
    fprintf(stderr, "out of memory\n");
    exit(2);
  }
  for (size_t i = 0; i < n; ++i) a[i] = v * (1.0 + 0.01 * (double)(i % 7));
  return a;
}

static int ncomp(int kind) { return (kind == 0 || kind == 2) ? 2 : 3; }
static int nfields(int kind) {
  static const int nf[4] = {5, 9, 3, 6};
  return nf[kind];
}

/* one case: kind on an nx * ny * nz grid, source at (si, sj, sk) */
static int run_case(int kind, int nx, int ny, int nz, int split, int corner) {
  const int nt = 12;
  const int three = (kind == 1 || kind == 3);
  if (!three) ny = 1;
  const size_t n2 = (size_t)nx * nz, n3 = n2 * (size_t)ny;
  const size_t nf = three ? n3 : n2;
  or_run_t r;
  memset(&r, 0, sizeof r);
  r.nx = nx;
  r.ny = ny;
  r.nz = nz;
  r.dx = r.dy = r.dz = 1.0;
  r.nt = nt;
  r.threads = 1;
  r.split = split;
  double* own[64];
  int nown = 0;
#define KEEP(p) (own[nown++] = (p))
  if (kind <= 1) {
    r.dt = 1.0e-4;
    for (int c = 0; c < 21; ++c) r.coef[c] = KEEP(arr(n2, (c == 0 || c == 11 || c == 20 || c == 6) ? 9.0e9 : 1.0e8));
    r.coef[21] = KEEP(arr(n2, 2000.0));
    for (int a = 0; a < 3; ++a) r.gam[a] = KEEP(arr(n2, 1.0e-3));
    for (int m = 0; m < 6; ++m) r.msrc[m] = KEEP(arr((size_t)nt, 1.0e3));
  } else {
    r.dt = 1.0e-10;
    int nc = kind == 2 ? 6 : 12;
    for (int c = 0; c < nc; ++c) r.coef[c] = KEEP(arr(n2, c < nc / 2 ? 4.0 : 1.0e-3));
  }
  const int nlen[3] = {nx, ny, nz};
  for (int a = 0; a < 3; ++a) {
    if (a == 1 && !three) continue;
    static const double pv[6] = {1.0, -0.2, 0.8, 1.0, -0.2, 0.8};
    for (int q = 0; q < 6; ++q) r.prof[a][q] = KEEP(arr((size_t)nlen[a], pv[q]));
  }
  /* corner 0: first cell of every axis, 1: last cell, 2: centre */
  r.isrc = corner == 1 ? nx - 1 : corner == 2 ? nx / 2 : 0;
  r.jsrc = corner == 1 ? ny - 1 : corner == 2 ? ny / 2 : 0;
  r.ksrc = corner == 1 ? nz - 1 : corner == 2 ? nz / 2 : 0;
  for (int c = 0; c < 3; ++c) r.src[c] = KEEP(arr((size_t)nt, 1.0));
  /* receivers on every corner of the grid and one beside the centre */
  int rec[9 * 3];
  r.nrec = 9;
  for (int q = 0; q < 8; ++q) {
    rec[3 * q + 0] = (q & 1) ? nx - 1 : 0;
    rec[3 * q + 1] = (q & 2) ? ny - 1 : 0;
    rec[3 * q + 2] = (q & 4) ? nz - 1 : 0;
  }
  rec[24] = nx / 2 > 0 ? nx / 2 - 1 : 0;
  rec[25] = ny / 2;
  rec[26] = nz / 2;
  r.rec = rec;
  r.trace = KEEP(arr((size_t)nt * ncomp(kind) * r.nrec, 0.0));
  for (int f = 0; f < nfields(kind); ++f) r.f[f] = KEEP(arr(nf, 0.0));
  /* two segments over one psi workspace, then a whole run with psi = NULL */
  r.psi = KEEP(arr((size_t)or_psi_len(kind, nx, ny, nz), 0.0));
  r.it0 = 0;
  r.nsteps = nt / 2;
  r.blowup_step = -1;
  int rc = or_run(kind, &r);
  if (!rc) {
    r.it0 = nt / 2;
    r.nsteps = nt - nt / 2;
    rc = or_run(kind, &r);
  }
  if (!rc) {
    r.psi = NULL;
    r.it0 = 0;
    r.nsteps = nt;
    rc = or_run(kind, &r);
  }
  /* finite, non-blown-up fields: the small cases are all stable */
  if (!rc && r.blowup_step >= 0) rc = 100;
  double sum = 0.0;
  for (size_t i = 0; i < (size_t)nt * ncomp(kind) * r.nrec; ++i) sum += r.trace[i];
  printf("kind %d grid %dx%dx%d split %d source %s: rc %d blowup %d trace sum %.6e\n", kind, nx, ny,
         nz, split, corner == 1 ? "last" : corner == 2 ? "centre" : "first", rc, r.blowup_step, sum);
  for (int i = 0; i < nown; ++i) free(own[i]);
#undef KEEP
  return rc;
}

int main(void) {
  static const int grids[3][3] = {{12, 5, 9}, {7, 4, 15}, {3, 3, 3}};
  int bad = 0;
  for (int kind = 0; kind < 4; ++kind)
    for (int g = 0; g < 3; ++g)
      for (int split = 0; split <= (kind == 0); ++split)
        for (int corner = 0; corner < 3; ++corner)
          bad |= run_case(kind, grids[g][0], grids[g][1], grids[g][2], split, corner);
  printf(bad ? "FAILED\n" : "ok\n");
  return bad;
}
