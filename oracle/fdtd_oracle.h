/*
 * fdtd_oracle.h -- CPU restatement (FP64, OpenMP) of SeidarT's legacy FDTD
 * time loops.  TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the
 * CPU baseline timed beside it.  Nothing in the shipped product links this.
 *
 * Every routine restates one reference subroutine loop for loop:
 *   or_elastic2   legacy/fortran/seismicfdtd.f95:12-385     (seismic2)
 *   or_elastic25  legacy/fortran/seismicfdtd.f95:388-949    (seismic25)
 *   or_em2        legacy/fortran/electromagfdtd.f95:10-326  (electromag2)
 *   or_em25       legacy/fortran/electromagfdtd.f95:619-1092 (electromag25)
 *
 * Conventions: grids are the PADDED grid (nx = domain.nx + 2*cpml ...), arrays
 * are x-fastest (Fortran column-major), 0-based indices (Fortran i -> i-1).
 * 2-D coefficient arrays are (nx, nz); 2.5-D coefficients are y-invariant
 * (nx, nz); 3-D fields are (nx, ny, nz).  CPML coefficients are the 1-D
 * profiles the legacy solver streams in (readwrite_routines.f95:26-37).
 *
 * Kept exactly as the legacy code does them:
 *   A1  one psi array per derivative, updated by every stress loop that
 *       computes that derivative (seismicfdtd.f95:255-262 and :291-298);
 *   A2  memory_dvx_dx in the normal-stress loop decays with b_x_half(j)
 *       (seismicfdtd.f95:255); j >= nx clamps to nx-1 (the legacy reads
 *       past the end there; every config has nz <= nx);
 *   A3  the c16 / c45 coupling terms of seismic25 (seismicfdtd.f95:657,766);
 *   A6  vx divides by rho(i,j), vz by the 3-point average (:334,:354);
 *   the sigmaxy loop of seismic25 applies no CPML correction to
 *       dvz_dx, dvx_dz (seismicfdtd.f95:707-713).
 * Deliberate departures (the legacy behaviour is undefined there):
 *   A4  gamma_z is its own input (the legacy reads gammas_z into gamma_x and
 *       leaves gamma_z uninitialised, seismicfdtd.f95:158-159);
 *   2.5-D elastic multiplies each velocity by (1 - gamma) exactly as the 2-D
 *       code does (seismicfdtd.f95:334); gamma = 0 reproduces seismic25.
 *   dt is an input (the legacy recomputes it, :166-167, :511-512).
 * Optional variant (split = 1, elastic2 only; not in the legacy): the shear
 *   loop keeps its own four velocity-derivative memory arrays (undoing A1)
 *   and the normal loop's memory_dvx_dx decays with b_x_half(i) (undoing
 *   A2).  With both quirks the x and top C-PML can grow without bound in
 *   long runs (a VTI ice model with a snow surface blows up after ~4400
 *   steps); the variant is the stable scheme the gradient workloads use.
 */
#ifndef SEIDART_FDTD_ORACLE_H
#define SEIDART_FDTD_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int nx, ny, nz;
  double dx, dy, dz, dt;
  int nt;
  int it0;                 /* first step index (0-based) for segmented runs */
  int nsteps;              /* steps to run in this call */
  const double* coef[22];  /* elastic: c11..c66 (21, Voigt order of the JSON) + rho */
  const double* gam[3];    /* elastic: gamma_x, gamma_y, gamma_z (NULL = 0)   */
  const double* prof[3][6];/* [axis x,y,z][kappa, a, b, kappa_h, a_h, b_h]    */
  int isrc, jsrc, ksrc;    /* padded 0-based source cell                      */
  const double* src[3];    /* nt samples per component (x, y, z); NULL = 0    */
  int nrec;
  const int* rec;          /* nrec * 3 padded 0-based (i, j, k)               */
  double* trace;           /* nt * ncomp * nrec, written for it in segment    */
  double* f[9];            /* state fields (in/out), see per-routine lists    */
  double* psi;             /* workspace (or NULL -> allocated per call)       */
  int blowup_step;         /* out: first step whose max |v| > 1e25 (or -1)    */
  int threads;             /* OpenMP threads (0 = runtime default)            */
  int split;               /* elastic2 only: 1 = split C-PML variant (below)  */
  const double* msrc[6];   /* elastic: stress-rate series (xx, yy, zz, xy, xz,
                              yz), rate * dt added at the source cell after
                              the stress update; NULL = none                  */
} or_run_t;

/* Number of doubles of psi workspace a routine needs (kept across segments). */
long long or_psi_len(int kind, int nx, int ny, int nz);

/* kind: 0 elastic2, 1 elastic25, 2 em2, 3 em25.
 * Elastic2  f[] = vx, vz, sxx, szz, sxz                   (coef: 21 Cij + rho)
 * Elastic25 f[] = vx, vy, vz, sxx, syy, szz, sxy, sxz, syz
 * EM2       f[] = Ex, Ez, Hy    coef[0..5] = e11,e13,e33,s11,s13,s33 (relative eps, S/m)
 * EM25      f[] = Ex, Ey, Ez, Hx, Hy, Hz
 *           coef[0..11] = e11,e12,e13,e22,e23,e33,s11,s12,s13,s22,s23,s33
 * Traces record the velocity / E components after each completed step. */
int or_run(int kind, or_run_t* r);

#ifdef __cplusplus
}
#endif
#endif
