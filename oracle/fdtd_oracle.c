/*
 * fdtd_oracle.c -- CPU restatement of SeidarT's legacy FDTD time loops.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP kernels are
 * compared against, and the CPU baseline that bench.py times beside them.
 * The product (seidart_amd) never links or calls it.
 *
 * The loops below follow the legacy Fortran one for one; each block cites the
 * reference lines it restates.  Floating-point expressions keep the Fortran
 * evaluation order (left to right, divide by 4 / 6 / 8 after the sum) so the
 * restatement agrees with the amdflang build of the legacy source to a few
 * ulp per step.  Build with -ffp-contract=off.
 */
#include "fdtd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { C11 = 0, C12, C13, C14, C15, C16, C22, C23, C24, C25, C26, C33, C34, C35,
       C36, C44, C45, C46, C55, C56, C66, RHO };
enum { PK = 0, PA, PB, PKH, PAH, PBH };

#define STAB_THRESHOLD 1.0e25
/* electromagfdtd.f95:77-84 */
#define OR_PI 3.141592653589793238462643
#define OR_EPS0 8.85418782e-12

long long or_psi_len(int kind, int nx, int ny, int nz) {
  long long n2 = (long long)nx * nz, n3 = n2 * ny;
  switch (kind) {
    case 0: return 12 * n2; /* 8 legacy arrays + 4 of the split variant */
    case 1: return 18 * n3;
    case 2: return 4 * n2;
    case 3: return 12 * n3;
  }
  return 0;
}

static int ncomp_of(int kind) { return (kind == 0 || kind == 2) ? 2 : 3; }

static void set_threads(const or_run_t* r) {
#ifdef _OPENMP
  if (r->threads > 0) omp_set_num_threads(r->threads);
#else
  (void)r;
#endif
}

/* ------------------------------------------------------------------------ */
/* seismic2: legacy/fortran/seismicfdtd.f95:12-385                            */
/* ------------------------------------------------------------------------ */
static void elastic2(or_run_t* r) {
  const int nx = r->nx, nz = r->nz;
  const double dx = r->dx, dz = r->dz, DT = r->dt;
  const double *c11 = r->coef[C11], *c13 = r->coef[C13], *c15 = r->coef[C15];
  const double *c33 = r->coef[C33], *c35 = r->coef[C35], *c55 = r->coef[C55];
  const double* rho = r->coef[RHO];
  const double *gx = r->gam[0], *gz = r->gam[2];
  const double *kx = r->prof[0][PK], *ax = r->prof[0][PA], *bx = r->prof[0][PB];
  const double *kxh = r->prof[0][PKH], *axh = r->prof[0][PAH], *bxh = r->prof[0][PBH];
  const double *kz = r->prof[2][PK], *az = r->prof[2][PA], *bz = r->prof[2][PB];
  const double *kzh = r->prof[2][PKH], *azh = r->prof[2][PAH], *bzh = r->prof[2][PBH];
  double *vx = r->f[0], *vz = r->f[1], *sxx = r->f[2], *szz = r->f[3], *sxz = r->f[4];
  const long long n2 = (long long)nx * nz;
  double* m_dvx_dx = r->psi + 0 * n2;
  double* m_dvx_dz = r->psi + 1 * n2;
  double* m_dvz_dx = r->psi + 2 * n2;
  double* m_dvz_dz = r->psi + 3 * n2;
  double* m_dsxx_dx = r->psi + 4 * n2;
  double* m_dszz_dz = r->psi + 5 * n2;
  double* m_dsxz_dx = r->psi + 6 * n2;
  double* m_dsxz_dz = r->psi + 7 * n2;
  /* shear-loop memory: the normal loop's arrays (A1), or their own (split) */
  const int split = r->split;
  double* s_dvx_dx = split ? r->psi + 8 * n2 : m_dvx_dx;
  double* s_dvx_dz = split ? r->psi + 9 * n2 : m_dvx_dz;
  double* s_dvz_dx = split ? r->psi + 10 * n2 : m_dvz_dx;
  double* s_dvz_dz = split ? r->psi + 11 * n2 : m_dvz_dz;
#define I2(i, j) ((long long)(i) + (long long)nx * (j))

  for (int s = 0; s < r->nsteps; ++s) {
    const int it = r->it0 + s;
    /* normal stresses, seismicfdtd.f95:247-281 */
#pragma omp parallel for schedule(static)
    for (int j = 1; j < nz; ++j) {
      const int jb = j < nx ? j : nx - 1; /* A2: b_x_half(j) */
      for (int i = 0; i < nx - 1; ++i) {
        const long long c = I2(i, j);
        double dvx_dx = (vx[I2(i + 1, j)] - vx[c]) / dx;
        double dvz_dz = (vz[c] - vz[I2(i, j - 1)]) / dz;
        double dvz_dx = (vz[I2(i + 1, j)] - vz[c]) / dx;
        double dvx_dz = (vx[c] - vx[I2(i, j - 1)]) / dz;
        m_dvx_dx[c] = bxh[split ? i : jb] * m_dvx_dx[c] + axh[i] * dvx_dx;
        m_dvz_dz[c] = bz[j] * m_dvz_dz[c] + az[j] * dvz_dz;
        m_dvx_dz[c] = bzh[j] * m_dvx_dz[c] + azh[j] * dvx_dz;
        m_dvz_dx[c] = bx[i] * m_dvz_dx[c] + ax[i] * dvz_dx;
        dvx_dx = dvx_dx / kxh[i] + m_dvx_dx[c];
        dvz_dz = dvz_dz / kz[j] + m_dvz_dz[c];
        dvz_dx = dvz_dx / kx[i] + m_dvz_dx[c];
        dvx_dz = dvx_dz / kzh[j] + m_dvx_dz[c];
        const long long e = I2(i + 1, j), w = I2(i, j - 1);
        const double a11 = (c11[e] + 2 * c11[c] + c11[w]) / 4;
        const double a13 = (c13[e] + 2 * c13[c] + c13[w]) / 4;
        const double a15 = (c15[e] + 2 * c15[c] + c15[w]) / 4;
        const double a33 = (c33[e] + 2 * c33[c] + c33[w]) / 4;
        const double a35 = (c35[e] + 2 * c35[c] + c35[w]) / 4;
        sxx[c] = sxx[c] + (a11 * dvx_dx + a13 * dvz_dz + a15 * (dvz_dx + dvx_dz)) * DT;
        szz[c] = szz[c] + (a13 * dvx_dx + a33 * dvz_dz + a35 * (dvz_dx + dvx_dz)) * DT;
      }
    }
    /* shear stress, seismicfdtd.f95:283-312 */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nz - 1; ++j) {
      for (int i = 1; i < nx; ++i) {
        const long long c = I2(i, j);
        double dvx_dx = (vx[c] - vx[I2(i - 1, j)]) / dx;
        double dvz_dz = (vz[I2(i, j + 1)] - vz[c]) / dz;
        double dvz_dx = (vz[c] - vz[I2(i - 1, j)]) / dx;
        double dvx_dz = (vx[I2(i, j + 1)] - vx[c]) / dz;
        s_dvx_dx[c] = bxh[i] * s_dvx_dx[c] + axh[i] * dvx_dx;
        s_dvz_dz[c] = bz[j] * s_dvz_dz[c] + az[j] * dvz_dz;
        s_dvx_dz[c] = bzh[j] * s_dvx_dz[c] + azh[j] * dvx_dz;
        s_dvz_dx[c] = bx[i] * s_dvz_dx[c] + ax[i] * dvz_dx;
        dvx_dx = dvx_dx / kxh[i] + s_dvx_dx[c];
        dvz_dz = dvz_dz / kz[j] + s_dvz_dz[c];
        dvz_dx = dvz_dx / kx[i] + s_dvz_dx[c];
        dvx_dz = dvx_dz / kzh[j] + s_dvx_dz[c];
        const long long n = I2(i, j + 1), w = I2(i - 1, j);
        const double a15 = (c15[n] + 2 * c15[c] + c15[w]) / 4;
        const double a35 = (c35[n] + 2 * c35[c] + c35[w]) / 4;
        const double a55 = (c55[n] + 2 * c55[c] + c55[w]) / 4;
        sxz[c] = sxz[c] + (a15 * dvx_dx + a35 * dvz_dz + a55 * (dvz_dx + dvx_dz)) * DT;
      }
    }
    /* stress-rate (moment-tensor) source: no counterpart in the legacy code;
       rate * dt (xx, zz, xz) added at the source cell after the stress update */
    {
      const long long c = I2(r->isrc, r->ksrc);
      if (r->msrc[0]) sxx[c] = sxx[c] + r->msrc[0][it] * DT;
      if (r->msrc[2]) szz[c] = szz[c] + r->msrc[2][it] * DT;
      if (r->msrc[4]) sxz[c] = sxz[c] + r->msrc[4][it] * DT;
    }
    /* vx, seismicfdtd.f95:317-337 (A6: divides by rho(i,j)) */
#pragma omp parallel for schedule(static)
    for (int j = 1; j < nz; ++j) {
      for (int i = 1; i < nx; ++i) {
        const long long c = I2(i, j);
        double dsxx_dx = (sxx[c] - sxx[I2(i - 1, j)]) / dx;
        double dsxz_dz = (sxz[c] - sxz[I2(i, j - 1)]) / dz;
        m_dsxx_dx[c] = bx[i] * m_dsxx_dx[c] + ax[i] * dsxx_dx;
        m_dsxz_dz[c] = bz[j] * m_dsxz_dz[c] + az[j] * dsxz_dz;
        dsxx_dx = dsxx_dx / kx[i] + m_dsxx_dx[c];
        dsxz_dz = dsxz_dz / kz[j] + m_dsxz_dz[c];
        const double g = gx ? gx[c] : 0.0;
        vx[c] = vx[c] * (1 - g) + (dsxx_dx + dsxz_dz) * DT / rho[c];
      }
    }
    /* vz, seismicfdtd.f95:339-357 */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nz - 1; ++j) {
      for (int i = 0; i < nx - 1; ++i) {
        const long long c = I2(i, j);
        const double deltarho = (2 * rho[c] + rho[I2(i + 1, j)] + rho[I2(i, j + 1)]) / 4;
        double dsxz_dx = (sxz[I2(i + 1, j)] - sxz[c]) / dx;
        double dszz_dz = (szz[I2(i, j + 1)] - szz[c]) / dz;
        m_dsxz_dx[c] = bxh[i] * m_dsxz_dx[c] + axh[i] * dsxz_dx;
        m_dszz_dz[c] = bzh[j] * m_dszz_dz[c] + azh[j] * dszz_dz;
        dsxz_dx = dsxz_dx / kxh[i] + m_dsxz_dx[c];
        dszz_dz = dszz_dz / kzh[j] + m_dszz_dz[c];
        const double g = gz ? gz[c] : 0.0;
        vz[c] = vz[c] * (1 - g) + (dsxz_dx + dszz_dz) * DT / deltarho;
      }
    }
    /* source, seismicfdtd.f95:359-361 */
    {
      const long long c = I2(r->isrc, r->ksrc);
      if (r->src[0]) vx[c] = vx[c] + r->src[0][it] * DT / rho[c];
      if (r->src[2]) vz[c] = vz[c] + r->src[2][it] * DT / rho[c];
    }
    /* Dirichlet edges, seismicfdtd.f95:365-375 */
    for (int j = 0; j < nz; ++j) {
      vx[I2(0, j)] = 0; vx[I2(nx - 1, j)] = 0;
      vz[I2(0, j)] = 0; vz[I2(nx - 1, j)] = 0;
    }
    for (int i = 0; i < nx; ++i) {
      vx[I2(i, 0)] = 0; vx[I2(i, nz - 1)] = 0;
      vz[I2(i, 0)] = 0; vz[I2(i, nz - 1)] = 0;
    }
    /* blow-up check, seismicfdtd.f95:378-379 */
    if (r->blowup_step < 0) {
      double vmax = 0.0;
      int bad = 0;
#pragma omp parallel for reduction(max : vmax) reduction(| : bad)
      for (long long c = 0; c < n2; ++c) {
        const double v = sqrt(vx[c] * vx[c] + vz[c] * vz[c]);
        if (v > vmax) vmax = v;
        if (!isfinite(v)) bad = 1;
      }
      if (vmax > STAB_THRESHOLD || bad) r->blowup_step = it;
    }
    if (r->trace) {
      double* t = r->trace + (long long)it * 2 * r->nrec;
      for (int q = 0; q < r->nrec; ++q) {
        const long long c = I2(r->rec[3 * q], r->rec[3 * q + 2]);
        t[q] = vx[c];
        t[r->nrec + q] = vz[c];
      }
    }
  }
#undef I2
}

/* ------------------------------------------------------------------------ */
/* seismic25: legacy/fortran/seismicfdtd.f95:388-949                          */
/* ------------------------------------------------------------------------ */
static void elastic25(or_run_t* r) {
  const int nx = r->nx, ny = r->ny, nz = r->nz;
  const double dx = r->dx, dy = r->dy, dz = r->dz, dt = r->dt;
  const double* const* C = r->coef;
  const double* rho = r->coef[RHO];
  const double *gx = r->gam[0], *gy = r->gam[1], *gz = r->gam[2];
  const double *kx = r->prof[0][PK], *ax = r->prof[0][PA], *bx = r->prof[0][PB];
  const double *kxh = r->prof[0][PKH], *axh = r->prof[0][PAH], *bxh = r->prof[0][PBH];
  const double *ky = r->prof[1][PK], *ay = r->prof[1][PA], *by = r->prof[1][PB];
  const double *kyh = r->prof[1][PKH], *ayh = r->prof[1][PAH], *byh = r->prof[1][PBH];
  const double *kz = r->prof[2][PK], *az = r->prof[2][PA], *bz = r->prof[2][PB];
  const double *kzh = r->prof[2][PKH], *azh = r->prof[2][PAH], *bzh = r->prof[2][PBH];
  double *vx = r->f[0], *vy = r->f[1], *vz = r->f[2];
  double *sxx = r->f[3], *syy = r->f[4], *szz = r->f[5];
  double *sxy = r->f[6], *sxz = r->f[7], *syz = r->f[8];
  const long long n3 = (long long)nx * ny * nz;
  double* P = r->psi;
  double *m_dvx_dx = P + 0 * n3, *m_dvx_dy = P + 1 * n3, *m_dvx_dz = P + 2 * n3;
  double *m_dvy_dx = P + 3 * n3, *m_dvy_dy = P + 4 * n3, *m_dvy_dz = P + 5 * n3;
  double *m_dvz_dx = P + 6 * n3, *m_dvz_dy = P + 7 * n3, *m_dvz_dz = P + 8 * n3;
  double *m_dsxx_dx = P + 9 * n3, *m_dsyy_dy = P + 10 * n3, *m_dszz_dz = P + 11 * n3;
  double *m_dsxy_dx = P + 12 * n3, *m_dsxy_dy = P + 13 * n3;
  double *m_dsxz_dx = P + 14 * n3, *m_dsxz_dz = P + 15 * n3;
  double *m_dsyz_dy = P + 16 * n3, *m_dsyz_dz = P + 17 * n3;
#define I3(i, j, k) ((long long)(i) + (long long)nx * ((long long)(j) + (long long)ny * (k)))
#define I2(i, k) ((long long)(i) + (long long)nx * (k))
/* sum of four 2-D coefficient entries, in the legacy order */
#define SUM4(cc, a, b, c2, d) ((((cc)[a] + (cc)[b]) + (cc)[c2]) + (cc)[d])

  for (int s = 0; s < r->nsteps; ++s) {
    const int it = r->it0 + s;
    /* sigma xx, yy, zz: seismicfdtd.f95:617-674 */
#pragma omp parallel for schedule(static)
    for (int k = 1; k < nz; ++k)
      for (int j = 1; j < ny; ++j)
        for (int i = 0; i < nx - 1; ++i) {
          const long long c = I3(i, j, k);
          const long long xp = I3(i + 1, j, k), ym = I3(i, j - 1, k), zm = I3(i, j, k - 1);
          double dvx_dx = (vx[xp] - vx[c]) / dx;
          double dvy_dx = (vy[xp] - vy[c]) / dx;
          double dvz_dx = (vz[xp] - vz[c]) / dx;
          double dvy_dy = (vy[c] - vy[ym]) / dy;
          double dvx_dy = (vx[c] - vx[ym]) / dy;
          double dvz_dy = (vz[c] - vz[ym]) / dy;
          double dvz_dz = (vz[c] - vz[zm]) / dz;
          double dvx_dz = (vx[c] - vx[zm]) / dz;
          double dvy_dz = (vy[c] - vy[zm]) / dz;
          m_dvx_dx[c] = bxh[i] * m_dvx_dx[c] + axh[i] * dvx_dx;
          m_dvy_dx[c] = bxh[i] * m_dvy_dx[c] + axh[i] * dvy_dx;
          m_dvz_dx[c] = bxh[i] * m_dvz_dx[c] + axh[i] * dvz_dx;
          m_dvy_dy[c] = by[j] * m_dvy_dy[c] + ay[j] * dvy_dy;
          m_dvx_dy[c] = by[j] * m_dvx_dy[c] + ay[j] * dvx_dy;
          m_dvz_dy[c] = by[j] * m_dvz_dy[c] + ay[j] * dvz_dy;
          m_dvz_dz[c] = bz[k] * m_dvz_dz[c] + az[k] * dvz_dz;
          m_dvx_dz[c] = bz[k] * m_dvx_dz[c] + az[k] * dvx_dz;
          m_dvy_dz[c] = bz[k] * m_dvy_dz[c] + az[k] * dvy_dz;
          dvx_dx = dvx_dx / kxh[i] + m_dvx_dx[c];
          dvy_dx = dvy_dx / kxh[i] + m_dvy_dx[c];
          dvz_dx = dvz_dx / kxh[i] + m_dvz_dx[c];
          dvy_dy = dvy_dy / ky[j] + m_dvy_dy[c];
          dvx_dy = dvx_dy / ky[j] + m_dvx_dy[c];
          dvz_dy = dvz_dy / ky[j] + m_dvz_dy[c];
          dvz_dz = dvz_dz / kz[k] + m_dvz_dz[c];
          dvx_dz = dvx_dz / kz[k] + m_dvx_dz[c];
          dvy_dz = dvy_dz / kz[k] + m_dvy_dz[c];
          const long long p0 = I2(i + 1, k), p1 = I2(i, k), p2 = I2(i, k - 1), p3 = I2(i + 1, k - 1);
#define A4(n) (SUM4(C[n], p0, p1, p2, p3) / 4)
#define A8(n) (SUM4(C[n], p0, p1, p2, p3) / 8)
          sxx[c] = sxx[c] + (A4(C11) * dvx_dx + A4(C12) * dvy_dy + A4(C13) * dvz_dz +
                             A8(C14) * (dvy_dz + dvz_dy) + A8(C15) * (dvx_dz + dvz_dx) +
                             A8(C16) * (dvx_dy + dvz_dy)) * dt; /* A3 */
          syy[c] = syy[c] + (A4(C12) * dvx_dx + A4(C22) * dvy_dy + A4(C23) * dvz_dz +
                             A8(C24) * (dvy_dz + dvz_dy) + A8(C25) * (dvx_dz + dvz_dx) +
                             A8(C26) * (dvy_dx + dvx_dy)) * dt;
          szz[c] = szz[c] + (A4(C13) * dvx_dx + A4(C23) * dvy_dy + A4(C33) * dvz_dz +
                             A8(C34) * (dvy_dz + dvz_dy) + A8(C35) * (dvx_dz + dvz_dx) +
                             A8(C36) * (dvy_dx + dvx_dy)) * dt;
#undef A4
#undef A8
        }
    /* sigma xy: seismicfdtd.f95:677-725 (no CPML correction of dvz_dx, dvx_dz) */
#pragma omp parallel for schedule(static)
    for (int k = 1; k < nz; ++k)
      for (int j = 0; j < ny - 1; ++j)
        for (int i = 1; i < nx; ++i) {
          const long long c = I3(i, j, k);
          const long long xm = I3(i - 1, j, k), yp = I3(i, j + 1, k), zm = I3(i, j, k - 1);
          double dvx_dx = (vx[c] - vx[xm]) / dx;
          double dvy_dx = (vy[c] - vy[xm]) / dx;
          double dvz_dx = (vz[c] - vz[xm]) / dx;
          double dvy_dy = (vy[yp] - vy[c]) / dy;
          double dvx_dy = (vx[yp] - vx[c]) / dy;
          double dvz_dy = (vz[yp] - vz[c]) / dy;
          double dvz_dz = (vz[c] - vz[zm]) / dz;
          double dvx_dz = (vx[c] - vx[zm]) / dz;
          double dvy_dz = (vy[c] - vy[zm]) / dz;
          m_dvx_dx[c] = bx[i] * m_dvx_dx[c] + ax[i] * dvx_dx;
          m_dvy_dx[c] = bx[i] * m_dvy_dx[c] + ax[i] * dvy_dx;
          m_dvz_dx[c] = bx[i] * m_dvz_dx[c] + ax[i] * dvz_dx;
          m_dvy_dy[c] = byh[j] * m_dvy_dy[c] + ayh[j] * dvy_dy;
          m_dvx_dy[c] = byh[j] * m_dvx_dy[c] + ayh[j] * dvx_dy;
          m_dvz_dy[c] = byh[j] * m_dvz_dy[c] + ayh[j] * dvz_dy;
          m_dvz_dz[c] = bzh[k] * m_dvz_dz[c] + azh[k] * dvz_dz;
          m_dvx_dz[c] = bzh[k] * m_dvx_dz[c] + azh[k] * dvx_dz;
          m_dvy_dz[c] = bzh[k] * m_dvy_dz[c] + azh[k] * dvy_dz;
          dvx_dx = dvx_dx / kx[i] + m_dvx_dx[c];
          dvy_dx = dvy_dx / kx[i] + m_dvy_dx[c];
          dvy_dy = dvy_dy / kyh[j] + m_dvy_dy[c];
          dvx_dy = dvx_dy / kyh[j] + m_dvx_dy[c];
          dvz_dy = dvz_dy / kyh[j] + m_dvz_dy[c];
          dvz_dz = dvz_dz / kzh[k] + m_dvz_dz[c];
          dvy_dz = dvy_dz / kzh[k] + m_dvy_dz[c];
          const long long p0 = I2(i, k), p1 = I2(i - 1, k), p2 = I2(i, k - 1), p3 = I2(i - 1, k - 1);
#define A4(n) (SUM4(C[n], p0, p1, p2, p3) / 4)
#define A8(n) (SUM4(C[n], p0, p1, p2, p3) / 8)
          sxy[c] = sxy[c] + (A4(C16) * dvx_dx + A4(C26) * dvy_dy + A4(C36) * dvz_dz +
                             A8(C46) * (dvz_dy + dvy_dz) + A8(C56) * (dvz_dx + dvx_dz) +
                             A8(C66) * (dvy_dx + dvx_dy)) * dt;
#undef A4
#undef A8
        }
    /* sigma xz and sigma yz: seismicfdtd.f95:728-815 */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz - 1; ++k) {
      for (int j = 1; j < ny; ++j)
        for (int i = 1; i < nx; ++i) {
          const long long c = I3(i, j, k);
          const long long xm = I3(i - 1, j, k), ym = I3(i, j - 1, k), zp = I3(i, j, k + 1);
          double dvx_dx = (vx[c] - vx[xm]) / dx;
          double dvy_dx = (vy[c] - vy[xm]) / dx;
          double dvz_dx = (vz[c] - vz[xm]) / dx;
          double dvy_dy = (vy[c] - vy[ym]) / dy;
          double dvz_dy = (vz[c] - vz[ym]) / dy;
          double dvx_dy = (vx[c] - vx[ym]) / dy;
          double dvz_dz = (vz[zp] - vz[c]) / dz;
          double dvx_dz = (vx[zp] - vx[c]) / dz;
          double dvy_dz = (vy[zp] - vy[c]) / dz;
          m_dvx_dx[c] = bx[i] * m_dvx_dx[c] + ax[i] * dvx_dx;
          m_dvy_dx[c] = bx[i] * m_dvy_dx[c] + ax[i] * dvy_dx;
          m_dvz_dx[c] = bx[i] * m_dvz_dx[c] + ax[i] * dvz_dx;
          m_dvy_dy[c] = by[j] * m_dvy_dy[c] + ay[j] * dvy_dy;
          m_dvx_dy[c] = by[j] * m_dvx_dy[c] + ay[j] * dvx_dy;
          m_dvz_dy[c] = by[j] * m_dvz_dy[c] + ay[j] * dvz_dy;
          m_dvz_dz[c] = bzh[k] * m_dvz_dz[c] + azh[k] * dvz_dz;
          m_dvx_dz[c] = bzh[k] * m_dvx_dz[c] + azh[k] * dvx_dz;
          m_dvy_dz[c] = bzh[k] * m_dvy_dz[c] + azh[k] * dvy_dz;
          dvx_dx = dvx_dx / kx[i] + m_dvx_dx[c];
          dvy_dx = dvy_dx / kx[i] + m_dvy_dx[c];
          dvz_dx = dvz_dx / kx[i] + m_dvz_dx[c];
          dvy_dy = dvy_dy / ky[j] + m_dvy_dy[c];
          dvx_dy = dvx_dy / ky[j] + m_dvx_dy[c];
          dvz_dy = dvz_dy / ky[j] + m_dvz_dy[c];
          dvz_dz = dvz_dz / kzh[k] + m_dvz_dz[c];
          dvx_dz = dvx_dz / kzh[k] + m_dvx_dz[c];
          dvy_dz = dvy_dz / kzh[k] + m_dvy_dz[c];
          const long long p0 = I2(i, k + 1), p1 = I2(i, k), p2 = I2(i - 1, k), p3 = I2(i - 1, k + 1);
#define A4(n) (SUM4(C[n], p0, p1, p2, p3) / 4)
#define A8(n) (SUM4(C[n], p0, p1, p2, p3) / 8)
          sxz[c] = sxz[c] + (A4(C15) * dvx_dx + A4(C25) * dvy_dy + A4(C35) * dvz_dz +
                             A8(C45) * (dvx_dz + dvz_dx) + A8(C55) * (dvx_dz + dvz_dx) +
                             A8(C56) * (dvx_dy + dvy_dx)) * dt; /* A3 */
#undef A4
#undef A8
        }
      for (int j = 0; j < ny - 1; ++j)
        for (int i = 0; i < nx - 1; ++i) {
          const long long c = I3(i, j, k);
          const long long xp = I3(i + 1, j, k), yp = I3(i, j + 1, k), zp = I3(i, j, k + 1);
          double dvx_dx = (vx[xp] - vx[c]) / dx;
          double dvy_dx = (vy[xp] - vy[c]) / dx;
          double dvz_dx = (vz[xp] - vz[c]) / dx;
          double dvy_dy = (vy[yp] - vy[c]) / dy;
          double dvx_dy = (vx[yp] - vx[c]) / dy;
          double dvz_dy = (vz[yp] - vz[c]) / dy;
          double dvz_dz = (vz[zp] - vz[c]) / dz;
          double dvx_dz = (vx[zp] - vx[c]) / dz;
          double dvy_dz = (vy[zp] - vy[c]) / dz;
          m_dvx_dx[c] = bxh[i] * m_dvx_dx[c] + axh[i] * dvx_dx;
          m_dvy_dx[c] = bxh[i] * m_dvy_dx[c] + axh[i] * dvy_dx;
          m_dvz_dx[c] = bxh[i] * m_dvz_dx[c] + axh[i] * dvz_dx;
          m_dvy_dy[c] = byh[j] * m_dvy_dy[c] + ayh[j] * dvy_dy;
          m_dvx_dy[c] = byh[j] * m_dvx_dy[c] + ayh[j] * dvx_dy;
          m_dvz_dy[c] = byh[j] * m_dvz_dy[c] + ayh[j] * dvz_dy;
          m_dvz_dz[c] = bzh[k] * m_dvz_dz[c] + azh[k] * dvz_dz;
          m_dvx_dz[c] = bzh[k] * m_dvx_dz[c] + azh[k] * dvx_dz;
          m_dvy_dz[c] = bzh[k] * m_dvy_dz[c] + azh[k] * dvy_dz;
          dvx_dx = dvx_dx / kxh[i] + m_dvx_dx[c];
          dvy_dx = dvy_dx / kxh[i] + m_dvy_dx[c];
          dvz_dx = dvz_dx / kxh[i] + m_dvz_dx[c];
          dvy_dy = dvy_dy / kyh[j] + m_dvy_dy[c];
          dvx_dy = dvx_dy / kyh[j] + m_dvx_dy[c];
          dvz_dy = dvz_dy / kyh[j] + m_dvz_dy[c];
          dvz_dz = dvz_dz / kzh[k] + m_dvz_dz[c];
          dvx_dz = dvx_dz / kzh[k] + m_dvx_dz[c];
          dvy_dz = dvy_dz / kzh[k] + m_dvy_dz[c];
          const long long p0 = I2(i + 1, k), p1 = I2(i, k), p2 = I2(i, k + 1), p3 = I2(i + 1, k + 1);
#define A4(n) (SUM4(C[n], p0, p1, p2, p3) / 4)
#define A8(n) (SUM4(C[n], p0, p1, p2, p3) / 8)
          syz[c] = syz[c] + (A4(C14) * dvx_dx + A4(C24) * dvy_dy + A4(C34) * dvz_dz +
                             A8(C44) * (dvy_dz + dvz_dy) + A8(C45) * (dvx_dz + dvz_dx) +
                             A8(C46) * (dvy_dx + dvx_dy)) * dt;
#undef A4
#undef A8
        }
    }
    /* stress-rate (moment-tensor) source: no counterpart in the legacy code;
       rate * dt (xx, yy, zz, xy, xz, yz) added at the source cell */
    {
      const long long c = I3(r->isrc, r->jsrc, r->ksrc);
      double* sig[6] = {sxx, syy, szz, sxy, sxz, syz};
      for (int q = 0; q < 6; ++q)
        if (r->msrc[q]) sig[q][c] = sig[q][c] + r->msrc[q][it] * dt;
    }
    /* vx, vy: seismicfdtd.f95:821-871 */
#pragma omp parallel for schedule(static)
    for (int k = 1; k < nz; ++k) {
      for (int j = 1; j < ny; ++j)
        for (int i = 1; i < nx; ++i) {
          const long long c = I3(i, j, k);
          const double deltarho = (4 * rho[I2(i, k)] + rho[I2(i - 1, k)] + rho[I2(i, k - 1)]) / 6;
          double dsxx_dx = (sxx[c] - sxx[I3(i - 1, j, k)]) / dx;
          double dsxy_dy = (sxy[c] - sxy[I3(i, j - 1, k)]) / dy;
          double dsxz_dz = (sxz[c] - sxz[I3(i, j, k - 1)]) / dz;
          m_dsxx_dx[c] = bx[i] * m_dsxx_dx[c] + ax[i] * dsxx_dx;
          m_dsxy_dy[c] = by[j] * m_dsxy_dy[c] + ay[j] * dsxy_dy;
          m_dsxz_dz[c] = bz[k] * m_dsxz_dz[c] + az[k] * dsxz_dz;
          dsxx_dx = dsxx_dx / kx[i] + m_dsxx_dx[c];
          dsxy_dy = dsxy_dy / ky[j] + m_dsxy_dy[c];
          dsxz_dz = dsxz_dz / kz[k] + m_dsxz_dz[c];
          const double g = gx ? gx[I2(i, k)] : 0.0;
          vx[c] = vx[c] * (1 - g) + (dsxx_dx + dsxy_dy + dsxz_dz) * dt / deltarho;
        }
      for (int j = 0; j < ny - 1; ++j)
        for (int i = 0; i < nx - 1; ++i) {
          const long long c = I3(i, j, k);
          const double deltarho = (4 * rho[I2(i, k)] + rho[I2(i + 1, k)] + rho[I2(i, k - 1)]) / 6;
          double dsxy_dx = (sxy[I3(i + 1, j, k)] - sxy[c]) / dx;
          double dsyy_dy = (syy[I3(i, j + 1, k)] - syy[c]) / dy;
          double dsyz_dz = (syz[c] - syz[I3(i, j, k - 1)]) / dz;
          m_dsxy_dx[c] = bxh[i] * m_dsxy_dx[c] + axh[i] * dsxy_dx;
          m_dsyy_dy[c] = byh[j] * m_dsyy_dy[c] + ayh[j] * dsyy_dy;
          m_dsyz_dz[c] = bz[k] * m_dsyz_dz[c] + az[k] * dsyz_dz;
          dsxy_dx = dsxy_dx / kxh[i] + m_dsxy_dx[c];
          dsyy_dy = dsyy_dy / kyh[j] + m_dsyy_dy[c];
          dsyz_dz = dsyz_dz / kz[k] + m_dsyz_dz[c];
          const double g = gy ? gy[I2(i, k)] : 0.0;
          vy[c] = vy[c] * (1 - g) + (dsxy_dx + dsyy_dy + dsyz_dz) * dt / deltarho;
        }
    }
    /* vz: seismicfdtd.f95:873-896 */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz - 1; ++k)
      for (int j = 1; j < ny; ++j)
        for (int i = 0; i < nx - 1; ++i) {
          const long long c = I3(i, j, k);
          const double deltarho = (rho[I2(i + 1, k)] + rho[I2(i, k + 1)] + 4 * rho[I2(i, k)]) / 6;
          double dsxz_dx = (sxz[I3(i + 1, j, k)] - sxz[c]) / dx;
          double dsyz_dy = (syz[c] - syz[I3(i, j - 1, k)]) / dy;
          double dszz_dz = (szz[I3(i, j, k + 1)] - szz[c]) / dz;
          m_dsxz_dx[c] = bxh[i] * m_dsxz_dx[c] + axh[i] * dsxz_dx;
          m_dsyz_dy[c] = by[j] * m_dsyz_dy[c] + ay[j] * dsyz_dy;
          m_dszz_dz[c] = bzh[k] * m_dszz_dz[c] + azh[k] * dszz_dz;
          dsxz_dx = dsxz_dx / kxh[i] + m_dsxz_dx[c];
          dsyz_dy = dsyz_dy / ky[j] + m_dsyz_dy[c];
          dszz_dz = dszz_dz / kzh[k] + m_dszz_dz[c];
          const double g = gz ? gz[I2(i, k)] : 0.0;
          vz[c] = vz[c] * (1 - g) + (dsxz_dx + dsyz_dy + dszz_dz) * dt / deltarho;
        }
    /* source: seismicfdtd.f95:898-900 */
    {
      const long long c = I3(r->isrc, r->jsrc, r->ksrc);
      const double rs = rho[I2(r->isrc, r->ksrc)];
      if (r->src[0]) vx[c] = vx[c] + r->src[0][it] * dt / rs;
      if (r->src[1]) vy[c] = vy[c] + r->src[1][it] * dt / rs;
      if (r->src[2]) vz[c] = vz[c] + r->src[2][it] * dt / rs;
    }
    /* Dirichlet faces: seismicfdtd.f95:903-925 */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          if (i == 0 || j == 0 || k == 0 || i == nx - 1 || j == ny - 1 || k == nz - 1) {
            const long long c = I3(i, j, k);
            vx[c] = 0; vy[c] = 0; vz[c] = 0;
          }
        }
    if (r->blowup_step < 0) {
      double vmax = 0.0;
      int bad = 0;
#pragma omp parallel for reduction(max : vmax) reduction(| : bad)
      for (long long c = 0; c < n3; ++c) {
        const double v = sqrt(vx[c] * vx[c] + vy[c] * vy[c] + vz[c] * vz[c]);
        if (v > vmax) vmax = v;
        if (!isfinite(v)) bad = 1;
      }
      if (vmax > STAB_THRESHOLD || bad) r->blowup_step = it;
    }
    if (r->trace) {
      double* t = r->trace + (long long)it * 3 * r->nrec;
      for (int q = 0; q < r->nrec; ++q) {
        const long long c = I3(r->rec[3 * q], r->rec[3 * q + 1], r->rec[3 * q + 2]);
        t[q] = vx[c];
        t[r->nrec + q] = vy[c];
        t[2 * r->nrec + q] = vz[c];
      }
    }
  }
#undef SUM4
#undef I3
#undef I2
}

/* EM update coefficients, electromagfdtd.f95:160-182 (2-D) and :750-788 (2.5-D) */
static void em_cacb(double sig, double epsr_sum, double dt, double* ca, double* cb) {
  const double eps = epsr_sum * OR_EPS0;
  *ca = (1.0 - sig * dt / (2.0 * eps)) / (1.0 + sig * dt / (2.0 * eps));
  *cb = (dt / eps) / (1.0 + sig * dt / (2.0 * eps));
}

/* ------------------------------------------------------------------------ */
/* electromag2: legacy/fortran/electromagfdtd.f95:10-326                      */
/* ------------------------------------------------------------------------ */
static void em2(or_run_t* r) {
  const int nx = r->nx, nz = r->nz;
  const double dx = r->dx, dz = r->dz, DT = r->dt;
  const double *e11 = r->coef[0], *e13 = r->coef[1], *e33 = r->coef[2];
  const double *s11 = r->coef[3], *s13 = r->coef[4], *s33 = r->coef[5];
  const double *kx = r->prof[0][PK], *ax = r->prof[0][PA], *bx = r->prof[0][PB];
  const double *kxh = r->prof[0][PKH], *axh = r->prof[0][PAH], *bxh = r->prof[0][PBH];
  const double *kz = r->prof[2][PK], *az = r->prof[2][PA], *bz = r->prof[2][PB];
  const double *kzh = r->prof[2][PKH], *azh = r->prof[2][PAH], *bzh = r->prof[2][PBH];
  double *Ex = r->f[0], *Ez = r->f[1], *Hy = r->f[2];
  const long long n2 = (long long)nx * nz;
  double *m_dEx_dz = r->psi, *m_dEz_dx = r->psi + n2;
  double *m_dHy_dx = r->psi + 2 * n2, *m_dHy_dz = r->psi + 3 * n2;
  double* caEx = (double*)malloc(sizeof(double) * 4 * n2);
  double *cbEx = caEx + n2, *caEz = caEx + 2 * n2, *cbEz = caEx + 3 * n2;
  for (long long c = 0; c < n2; ++c) {
    em_cacb(s11[c] + s13[c], e11[c] + e13[c], DT, &caEx[c], &cbEx[c]);
    em_cacb(s13[c] + s33[c], e13[c] + e33[c], DT, &caEz[c], &cbEz[c]);
  }
  const double mu0 = 4.0 * OR_PI * 1.0e-7;
  const double daHy = 1.0, dbHy = DT / mu0;
#define I2(i, j) ((long long)(i) + (long long)nx * (j))

  for (int s = 0; s < r->nsteps; ++s) {
    const int it = r->it0 + s;
    /* Hy, electromagfdtd.f95:249-266 */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nz - 1; ++j)
      for (int i = 0; i < nx - 1; ++i) {
        const long long c = I2(i, j);
        double dEx_dz = (Ex[I2(i, j + 1)] - Ex[c]) / dz;
        m_dEx_dz[c] = bz[j] * m_dEx_dz[c] + az[j] * dEx_dz;
        dEx_dz = dEx_dz / kz[j] + m_dEx_dz[c];
        double dEz_dx = (Ez[I2(i + 1, j)] - Ez[c]) / dx;
        m_dEz_dx[c] = bx[i] * m_dEz_dx[c] + ax[i] * dEz_dx;
        dEz_dx = dEz_dx / kx[i] + m_dEz_dx[c];
        Hy[c] = daHy * Hy[c] + dbHy * (dEz_dx + dEx_dz);
      }
    /* Ex, electromagfdtd.f95:273-283 */
#pragma omp parallel for schedule(static)
    for (int j = 1; j < nz - 1; ++j)
      for (int i = 0; i < nx - 1; ++i) {
        const long long c = I2(i, j), cm = I2(i, j - 1);
        double dHy_dz = (Hy[c] - Hy[cm]) / dz;
        m_dHy_dz[c] = bzh[j] * m_dHy_dz[c] + azh[j] * dHy_dz;
        dHy_dz = dHy_dz / kzh[j] + m_dHy_dz[c];
        Ex[c] = ((caEx[c] + caEx[cm]) / 2) * Ex[c] + ((cbEx[c] + cbEx[cm]) / 2) * dHy_dz;
      }
    /* Ez, electromagfdtd.f95:285-295 */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < nz - 1; ++j)
      for (int i = 1; i < nx - 1; ++i) {
        const long long c = I2(i, j), cm = I2(i - 1, j);
        double dHy_dx = (Hy[c] - Hy[cm]) / dx;
        m_dHy_dx[c] = bxh[i] * m_dHy_dx[c] + axh[i] * dHy_dx;
        dHy_dx = dHy_dx / kxh[i] + m_dHy_dx[c];
        Ez[c] = ((caEz[c] + caEz[cm]) / 2) * Ez[c] + ((cbEz[c] + cbEz[cm]) / 2) * dHy_dx;
      }
    /* source, electromagfdtd.f95:300-301 (relative permittivity divisor) */
    {
      const long long c = I2(r->isrc, r->ksrc);
      if (r->src[0]) Ex[c] = Ex[c] + r->src[0][it] * DT / e11[c];
      if (r->src[2]) Ez[c] = Ez[c] + r->src[2][it] * DT / e33[c];
    }
    /* Dirichlet, electromagfdtd.f95:304-317 */
    for (int j = 0; j < nz; ++j) {
      Ex[I2(0, j)] = 0; Ex[I2(nx - 2, j)] = 0;
      Ez[I2(0, j)] = 0; Ez[I2(nx - 1, j)] = 0;
      Hy[I2(0, j)] = 0; Hy[I2(nx - 2, j)] = 0;
    }
    for (int i = 0; i < nx; ++i) {
      Ex[I2(i, 0)] = 0; Ex[I2(i, nz - 1)] = 0;
      Ez[I2(i, 0)] = 0; Ez[I2(i, nz - 2)] = 0;
      Hy[I2(i, 0)] = 0; Hy[I2(i, nz - 2)] = 0;
    }
    if (r->blowup_step < 0) {
      double vmax = 0.0;
      int bad = 0;
#pragma omp parallel for reduction(max : vmax) reduction(| : bad)
      for (long long c = 0; c < n2; ++c) {
        const double v = sqrt(Ex[c] * Ex[c] + Ez[c] * Ez[c]);
        if (v > vmax) vmax = v;
        if (!isfinite(v)) bad = 1;
      }
      if (vmax > STAB_THRESHOLD || bad) r->blowup_step = it;
    }
    if (r->trace) {
      double* t = r->trace + (long long)it * 2 * r->nrec;
      for (int q = 0; q < r->nrec; ++q) {
        const long long c = I2(r->rec[3 * q], r->rec[3 * q + 2]);
        t[q] = Ex[c];
        t[r->nrec + q] = Ez[c];
      }
    }
  }
  free(caEx);
#undef I2
}

/* ------------------------------------------------------------------------ */
/* electromag25: legacy/fortran/electromagfdtd.f95:619-1092                   */
/* ------------------------------------------------------------------------ */
static void em25(or_run_t* r) {
  const int nx = r->nx, ny = r->ny, nz = r->nz;
  const double dx = r->dx, dy = r->dy, dz = r->dz, DT = r->dt;
  const double* const* E = r->coef; /* e11 e12 e13 e22 e23 e33 s11 s12 s13 s22 s23 s33 */
  const double *kx = r->prof[0][PK], *ax = r->prof[0][PA], *bx = r->prof[0][PB];
  const double *kxh = r->prof[0][PKH], *axh = r->prof[0][PAH], *bxh = r->prof[0][PBH];
  const double *ky = r->prof[1][PK], *ay = r->prof[1][PA], *by = r->prof[1][PB];
  const double *kyh = r->prof[1][PKH], *ayh = r->prof[1][PAH], *byh = r->prof[1][PBH];
  const double *kz = r->prof[2][PK], *az = r->prof[2][PA], *bz = r->prof[2][PB];
  const double *kzh = r->prof[2][PKH], *azh = r->prof[2][PAH], *bzh = r->prof[2][PBH];
  double *Ex = r->f[0], *Ey = r->f[1], *Ez = r->f[2];
  double *Hx = r->f[3], *Hy = r->f[4], *Hz = r->f[5];
  const long long n2 = (long long)nx * nz, n3 = n2 * ny;
  double* P = r->psi;
  double *m_dEy_dx = P, *m_dEx_dy = P + n3, *m_dEz_dx = P + 2 * n3, *m_dEx_dz = P + 3 * n3;
  double *m_dEy_dz = P + 4 * n3, *m_dEz_dy = P + 5 * n3;
  double *m_dHz_dx = P + 6 * n3, *m_dHx_dz = P + 7 * n3, *m_dHy_dx = P + 8 * n3;
  double *m_dHx_dy = P + 9 * n3, *m_dHy_dz = P + 10 * n3, *m_dHz_dy = P + 11 * n3;
  double* ca = (double*)malloc(sizeof(double) * 6 * n2);
  double *caEx = ca, *cbEx = ca + n2, *caEy = ca + 2 * n2, *cbEy = ca + 3 * n2;
  double *caEz = ca + 4 * n2, *cbEz = ca + 5 * n2;
  for (long long c = 0; c < n2; ++c) {
    em_cacb(E[6][c] + E[7][c] + E[8][c], E[0][c] + E[1][c] + E[2][c], DT, &caEx[c], &cbEx[c]);
    em_cacb(E[7][c] + E[9][c] + E[10][c], E[1][c] + E[3][c] + E[4][c], DT, &caEy[c], &cbEy[c]);
    em_cacb(E[8][c] + E[10][c] + E[11][c], E[2][c] + E[4][c] + E[5][c], DT, &caEz[c], &cbEz[c]);
  }
  const double mu0 = 4.0 * OR_PI * 1.0e-7;
  const double daH = 1.0, dbH = DT / mu0;
#define I3(i, j, k) ((long long)(i) + (long long)nx * ((long long)(j) + (long long)ny * (k)))
#define I2(i, k) ((long long)(i) + (long long)nx * (k))

  for (int s = 0; s < r->nsteps; ++s) {
    const int it = r->it0 + s;
    /* Hx, electromagfdtd.f95:901-920 */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz - 1; ++k)
      for (int i = 0; i < nx - 1; ++i)
        for (int j = 0; j < ny - 1; ++j) {
          const long long c = I3(i, j, k);
          double dEz_dy = (Ez[c] - Ez[I3(i, j + 1, k)]) / dy;
          m_dEz_dy[c] = byh[j] * m_dEz_dy[c] + ayh[j] * dEz_dy;
          dEz_dy = dEz_dy / kyh[j] + m_dEz_dy[c];
          double dEy_dz = (Ey[I3(i, j, k + 1)] - Ey[c]) / dz;
          m_dEy_dz[c] = bzh[k] * m_dEy_dz[c] + azh[k] * dEy_dz;
          dEy_dz = dEy_dz / kzh[k] + m_dEy_dz[c];
          Hx[c] = daH * Hx[c] + dbH * (dEy_dz + dEz_dy);
        }
    /* Hy, electromagfdtd.f95:922-942 */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz - 1; ++k)
      for (int i = 0; i < nx - 1; ++i)
        for (int j = 0; j < ny - 1; ++j) {
          const long long c = I3(i, j, k);
          double dEx_dz = (Ex[c] - Ex[I3(i, j, k + 1)]) / dz;
          m_dEx_dz[c] = bz[k] * m_dEx_dz[c] + az[k] * dEx_dz;
          dEx_dz = dEx_dz / kz[k] + m_dEx_dz[c];
          double dEz_dx = (Ez[I3(i + 1, j, k)] - Ez[c]) / dx;
          m_dEz_dx[c] = bx[i] * m_dEz_dx[c] + ax[i] * dEz_dx;
          dEz_dx = dEz_dx / kx[i] + m_dEz_dx[c];
          Hy[c] = daH * Hy[c] + dbH * (dEz_dx + dEx_dz);
        }
    /* Hz, electromagfdtd.f95:944-962 */
#pragma omp parallel for schedule(static)
    for (int k = 1; k < nz - 1; ++k)
      for (int i = 0; i < nx - 1; ++i)
        for (int j = 0; j < ny - 1; ++j) {
          const long long c = I3(i, j, k);
          double dEx_dy = (Ex[I3(i, j + 1, k)] - Ex[c]) / dy;
          m_dEx_dy[c] = by[j] * m_dEx_dy[c] + ay[j] * dEx_dy;
          dEx_dy = dEx_dy / ky[j] + m_dEx_dy[c];
          double dEy_dx = (Ey[c] - Ey[I3(i + 1, j, k)]) / dx;
          m_dEy_dx[c] = bx[i] * m_dEy_dx[c] + ax[i] * dEy_dx;
          dEy_dx = dEy_dx / kx[i] + m_dEy_dx[c];
          Hz[c] = daH * Hz[c] + dbH * (dEy_dx + dEx_dy);
        }
    /* Ex and Ey, electromagfdtd.f95:968-1008 */
#pragma omp parallel for schedule(static)
    for (int k = 1; k < nz - 1; ++k) {
      for (int i = 0; i < nx - 1; ++i)
        for (int j = 1; j < ny - 1; ++j) {
          const long long c = I3(i, j, k);
          double dHz_dy = (Hz[c] - Hz[I3(i, j - 1, k)]) / dy;
          m_dHz_dy[c] = byh[j] * m_dHz_dy[c] + ayh[j] * dHz_dy;
          dHz_dy = dHz_dy / kyh[j] + m_dHz_dy[c];
          double dHy_dz = (Hy[I3(i, j, k - 1)] - Hy[c]) / dz;
          m_dHy_dz[c] = bz[k] * m_dHy_dz[c] + az[k] * dHy_dz;
          dHy_dz = dHy_dz / kz[k] + m_dHy_dz[c];
          Ex[c] = caEx[I2(i, k)] * Ex[c] + cbEx[I2(i, k)] * (dHz_dy + dHy_dz);
        }
      for (int i = 1; i < nx - 1; ++i)
        for (int j = 0; j < ny - 1; ++j) {
          const long long c = I3(i, j, k);
          double dHz_dx = (Hz[I3(i - 1, j, k)] - Hz[c]) / dx;
          m_dHz_dx[c] = bxh[i] * m_dHz_dx[c] + axh[i] * dHz_dx;
          dHz_dx = dHz_dx / kxh[i] + m_dHz_dx[c];
          double dHx_dz = (Hx[c] - Hx[I3(i, j, k - 1)]) / dz;
          m_dHx_dz[c] = bzh[k] * m_dHx_dz[c] + azh[k] * dHx_dz;
          dHx_dz = dHx_dz / kzh[k] + m_dHx_dz[c];
          const long long p = I2(i, k), pw = I2(i - 1, k), pd = I2(i, k - 1);
          Ey[c] = ((4 * caEy[p] + caEy[pw] + caEy[pd]) / 6) * Ey[c] +
                  ((4 * cbEy[p] + cbEy[pw] + cbEy[pd]) / 6) * (dHz_dx + dHx_dz);
        }
    }
    /* Ez, electromagfdtd.f95:1011-1029 */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz - 1; ++k)
      for (int i = 1; i < nx - 1; ++i)
        for (int j = 1; j < ny - 1; ++j) {
          const long long c = I3(i, j, k);
          double dHx_dy = (Hx[I3(i, j - 1, k)] - Hx[c]) / dy;
          m_dHx_dy[c] = byh[j] * m_dHx_dy[c] + ayh[j] * dHx_dy;
          dHx_dy = dHx_dy / kyh[j] + m_dHx_dy[c];
          double dHy_dx = (Hy[c] - Hy[I3(i - 1, j, k)]) / dx;
          m_dHy_dx[c] = bxh[i] * m_dHy_dx[c] + axh[i] * dHy_dx;
          dHy_dx = dHy_dx / kxh[i] + m_dHy_dx[c];
          const long long p = I2(i, k), pw = I2(i - 1, k), pu = I2(i, k + 1);
          Ez[c] = ((4 * caEz[p] + caEz[pw] + caEz[pu]) / 6) * Ez[c] +
                  ((4 * cbEz[p] + cbEz[pw] + cbEz[pu]) / 6) * (dHx_dy + dHy_dx);
        }
    /* source, electromagfdtd.f95:1033-1035 */
    {
      const long long c = I3(r->isrc, r->jsrc, r->ksrc);
      const long long p = I2(r->isrc, r->ksrc);
      if (r->src[0]) Ex[c] = Ex[c] + r->src[0][it] * DT / E[0][p];
      if (r->src[1]) Ey[c] = Ey[c] + r->src[1][it] * DT / E[3][p];
      if (r->src[2]) Ez[c] = Ez[c] + r->src[2][it] * DT / E[5][p];
    }
    /* Dirichlet faces, electromagfdtd.f95:1038-1078 */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          const long long c = I3(i, j, k);
          const int lo = (i == 0 || j == 0 || k == 0);
          if (lo || i == nx - 2 || j == ny - 1 || k == nz - 1) Ex[c] = 0;
          if (lo || i == nx - 1 || j == ny - 2 || k == nz - 1) Ey[c] = 0;
          if (lo || i == nx - 1 || j == ny - 1 || k == nz - 2) Ez[c] = 0;
          if (lo || i == nx - 1 || j == ny - 2 || k == nz - 2) Hx[c] = 0;
          if (lo || i == nx - 2 || j == ny - 1 || k == nz - 2) Hy[c] = 0;
          if (lo || i == nx - 2 || j == ny - 2 || k == nz - 1) Hz[c] = 0;
        }
    if (r->blowup_step < 0) {
      double vmax = 0.0;
      int bad = 0;
#pragma omp parallel for reduction(max : vmax) reduction(| : bad)
      for (long long c = 0; c < n3; ++c) {
        const double v = sqrt(Ex[c] * Ex[c] + Ey[c] * Ey[c] + Ez[c] * Ez[c]);
        if (v > vmax) vmax = v;
        if (!isfinite(v)) bad = 1;
      }
      if (vmax > STAB_THRESHOLD || bad) r->blowup_step = it;
    }
    if (r->trace) {
      double* t = r->trace + (long long)it * 3 * r->nrec;
      for (int q = 0; q < r->nrec; ++q) {
        const long long c = I3(r->rec[3 * q], r->rec[3 * q + 1], r->rec[3 * q + 2]);
        t[q] = Ex[c];
        t[r->nrec + q] = Ey[c];
        t[2 * r->nrec + q] = Ez[c];
      }
    }
  }
  free(ca);
#undef I3
#undef I2
}

int or_run(int kind, or_run_t* r) {
  if (!r || kind < 0 || kind > 3) return 1;
  if (r->nx < 3 || r->nz < 3 || ((kind == 1 || kind == 3) && r->ny < 3)) return 2;
  set_threads(r);
  int own = 0;
  if (!r->psi) {
    r->psi = (double*)calloc((size_t)or_psi_len(kind, r->nx, r->ny, r->nz), sizeof(double));
    if (!r->psi) return 3;
    own = 1;
  }
  (void)ncomp_of;
  switch (kind) {
    case 0: elastic2(r); break;
    case 1: elastic25(r); break;
    case 2: em2(r); break;
    case 3: em25(r); break;
  }
  if (own) {
    free(r->psi);
    r->psi = NULL;
  }
  return 0;
}
