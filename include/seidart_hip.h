/*
 * seidart_hip.h -- C ABI of libseidart_hip.so, the MI355X (gfx950) FDTD
 * time-stepping core behind SeidarT's run path.
 *
 * What it replaces.  The reference crosses a process boundary at
 *     subprocess.call(['seidartfdtd', json, 'seismic=..', 'density_method=..'])
 * in src/seidart/routines/classes.py:878-924 (Model.run), and the solver
 * behind it is the Komatitsch/Martin C-PML staggered-grid scheme whose
 * readable source is legacy/fortran/seismicfdtd.f95 (seismic2 :12-385,
 * seismic25 :388-949) and legacy/fortran/electromagfdtd.f95 (electromag2
 * :10-326, electromag25 :619-1092).  The receiver gather that the reference
 * does in Python from per-step snapshot files (src/seidart/routines/
 * arraybuild.py:199-295) happens on the device here.  The entry points below
 * are what a ctypes / cgo / JNI binding of that boundary would call; the
 * Python binding is seidart_amd/_hip.py.
 *
 * Conventions.
 *   - All grids are the padded grid of the reference (nx = domain.nx + 2*cpml
 *     and so on, classes.py:608-679), x fastest (Fortran column-major), with
 *     every field of a 2-D run stored as nx*nz and of a 2.5-D run as nx*ny*nz
 *     (cells outside a staggered field's Fortran extent stay zero).
 *   - Every array pointer inside sdt_model / sdt_shot is DEVICE memory owned
 *     by the caller, in the plan's precision (double for SDT_FP64, float for
 *     SDT_FP32) except rec (int32).  The library never frees them.
 *   - The library owns only its descriptor copies, work list and per-shot
 *     instability flags (allocated in sdt_plan_create, freed in
 *     sdt_plan_destroy).
 *   - Errors: every call returns 0 on success or an SDT_E* code; the message
 *     of the last failure on the calling thread is in sdt_last_error().
 *     Nothing calls exit() (the reference's solver stops the process and its
 *     caller ignores the status, classes.py:910-924).
 *   - Threading: a plan is used by one thread at a time; plans are
 *     independent.  sdt_plan_run only enqueues on the given stream.
 */
#ifndef SEIDART_HIP_H
#define SEIDART_HIP_H
#ifdef __cplusplus
extern "C" {
#endif

/* physics kinds: which reference subroutine's scheme runs */
#define SDT_ELASTIC2  0 /* seismic2     seismicfdtd.f95:12-385       */
#define SDT_ELASTIC25 1 /* seismic25    seismicfdtd.f95:388-949      */
#define SDT_EM2       2 /* electromag2  electromagfdtd.f95:10-326    */
#define SDT_EM25      3 /* electromag25 electromagfdtd.f95:619-1092  */

/* Flag OR'ed into an ELASTIC2 kind (sdt_plan_create, sdt_psi_elems,
 * sdt_state_elems, sdt_adjoint_work_elems): the split C-PML variant.  The
 * shear-stress loop keeps its own four velocity-derivative memory arrays
 * instead of sharing the normal-stress loop's (seismicfdtd.f95:255-262 and
 * :291-298 update the same arrays), and the normal loop's memory_dvx_dx
 * decays with b_x_half(i) instead of b_x_half(j) (:255).  Not in the
 * reference: with both legacy quirks the x and top C-PML grow without bound
 * in long runs (a 6000-step VTI ice model with a snow surface blows up at
 * step 4414 in the legacy arithmetic).  psi then holds 6 arrays per slab
 * (x: dvx_dx dvz_dx dsxx_dx dsxz_dx, then the shear loop's dvx_dx dvz_dx;
 * z: dvz_dz dvx_dz dszz_dz dsxz_dz, then dvz_dz dvx_dz) and aux is unused. */
#define SDT_SPLIT_CPML 0x100

/* arithmetic precision of fields, coefficients and traces */
#define SDT_FP64 0
#define SDT_FP32 1

/* error codes */
#define SDT_OK        0
#define SDT_EINVAL    1 /* bad argument (sizes, indices, null pointers)  */
#define SDT_EDEVICE   2 /* HIP runtime failure                           */
#define SDT_EUNSTABLE 3 /* a field exceeded 1e25 or became non-finite    */

/* A model: one padded grid and the kernel coefficient arrays built from its
 * project (FP64 host setup, then stored in the plan's precision).
 * coef[] slots per kind:
 *   ELASTIC2 (12): n11 n13 n15 n33 n35 | s15 s35 s55 | dtr_x dtr_z | om_x om_z
 *       n* = (c(i+1,j)+2c(i,j)+c(i,j-1))/4 at the normal-stress node,
 *       s* = (c(i,j+1)+2c(i,j)+c(i-1,j))/4 at the shear node
 *       (seismicfdtd.f95:269-278, :307-310); dtr_x = dt/rho(i,j),
 *       dtr_z = dt/((2rho(i,j)+rho(i+1,j)+rho(i,j+1))/4) (:334, :342, :354);
 *       om = 1 - gamma (:334, :354).  Slots 2, 4, 5, 6 (the c15 / c35
 *       couplings) may be null together, meaning zero, and slots 10, 11 may
 *       be null together, meaning no attenuation; the kernel then skips them.
 *   ELASTIC25 (39): the 4-point Cij averages of the four stress loops
 *       (seismicfdtd.f95:651-673 [15], :716-721 [6], :761-767 [6],
 *       :806-811 [6]), dt/deltarho for vx, vy, vz (:826, :853, :878),
 *       1 - gamma for vx, vy, vz.  All 2-D (nx*nz), y-invariant.
 *   EM2 (4): caEx cbEx caEz cbEz, already 2-point averaged
 *       (electromagfdtd.f95:171-177, :281-282, :293-294).
 *   EM25 (6): caEx cbEx, caEy cbEy, caEz cbEz; Ey/Ez already (4,1,1)/6
 *       averaged (electromagfdtd.f95:770-781, :985-1026).  2-D, y-invariant.
 * prof[axis][q]: 1-D C-PML profiles per axis (x: nx, y: ny, z: nz entries),
 *   q = 0 1/kappa, 1 a, 2 b, 3 1/kappa_half, 4 a_half, 5 b_half
 *   (the values the legacy loadcpml streams in, readwrite_routines.f95:26-37;
 *   kappa is passed as its reciprocal).  Unused axes may be NULL.
 * aux: ELASTIC2 only, nz entries: b_x_half indexed by the z row, the decay
 *   the legacy normal-stress loop applies to memory_dvx_dx
 *   (seismicfdtd.f95:255); NULL elsewhere.
 * lx, ly, lz: psi slab widths.  Cells with i < lx or i >= nx - lx carry x
 *   memory variables (likewise y, z); every other cell must have a = 0 and
 *   kappa = 1 on that axis.  0 <= l <= (n+1)/2.
 * scal[0]: EM only, dt/mu0 (electromagfdtd.f95:179-181).
 */
typedef struct sdt_model {
  int nx, ny, nz;
  int lx, ly, lz;
  double dx, dy, dz, dt;
  double scal[4];
  const void* coef[40];
  const void* prof[3][6];
  const void* aux;
} sdt_model;

/* A shot: its state, C-PML memory, source and receivers, on one model.
 * field[]: ELASTIC2 vx vz sxx szz sxz | ELASTIC25 vx vy vz sxx syy szz sxy sxz syz
 *          EM2 Ex Ez Hy | EM25 Ex Ey Ez Hx Hy Hz.   In/out; initial values are
 *          the initial conditions (classes.py:888-898).
 * psi: sdt_psi_elems() elements, zero before the first step, kept between
 *      calls of sdt_plan_run (a run may be split into segments).
 * field_b, psi_b: ELASTIC2 and EM2 (NULL elsewhere): a second state buffer
 *      of the same sizes.  Their single-pass step kernels read one buffer and
 *      write the other; at the end of every sdt_plan_run the state is back
 *      in field / psi, with field_b / psi_b holding scratch or a copy.  A
 *      caller that rewrites field / psi between runs writes the same values
 *      into field_b / psi_b (the plan may reuse B's copy for a one-step run
 *      that starts at an odd step).
 * isrc, jsrc, ksrc: 0-based padded source cell (jsrc ignored in 2-D).
 * src[c]: nt per-step increments added to component c at the source cell
 *      after the update (seismicfdtd.f95:359-361: s(it)*dt/rho(src);
 *      electromagfdtd.f95:300-301: s(it)*dt/eps_r(src)); NULL = none.
 *      Components are (x, y, z); 2-D kinds use x and z.
 * msrc[c]: ELASTIC2 / ELASTIC25 only (NULL elsewhere): nt per-step stress
 *      increments (stress rate x dt) added at the source cell after the
 *      stress update, components (xx, yy, zz, xy, xz, yz); NULL = none.  The
 *      moment-tensor series of sourcefunction.py:296-372; the reference's
 *      solver that consumes them is not in its sources, so this injection
 *      is this library's definition.  2-D kinds use xx, zz and xz.
 * rec: nrec triples (i, j, k), 0-based padded; trace: nt * ncomp * nrec,
 *      trace[it][c][r] = component c at receiver r after step it, with
 *      ncomp = 2 (ELASTIC2: vx vz, EM2: Ex Ez) or 3 (vx vy vz / Ex Ey Ez).
 */
typedef struct sdt_shot {
  int model;
  int isrc, jsrc, ksrc;
  void* field[9];
  void* psi;
  void* field_b[9];
  void* psi_b;
  const void* src[3];
  int nrec;
  const int* rec;
  void* trace;
  const void* msrc[6];
} sdt_shot;

/* Elements of one shot's psi workspace for a kind and grid (slab storage:
 * each memory variable lives only on the slabs of its own axis). */
int sdt_psi_elems(int kind, int nx, int ny, int nz, int lx, int ly, int lz,
                  long long* out_elems);

/* Build a plan over nmodels models and nshots shots (each shot names its
 * model).  One sdt_plan_run step advances every shot by one time step; the
 * shots of all models run in the same kernel launches over one
 * (shot, tile) work list.  nt bounds the step index.  Replaces one
 * `seidartfdtd` process per shot (classes.py:910-924,
 * src/seidart/simulations/common_offset.py:143-283). */
int sdt_plan_create(int kind, int precision, const sdt_model* models, int nmodels,
                    const sdt_shot* shots, int nshots, int nt, void** out_plan);

/* Enqueue steps it0 .. it0+nsteps-1 on `stream` (a hipStream_t, NULL =
 * default stream).  graph_block > 1 captures that many steps per hipGraph
 * and replays it; 0 or 1 launches kernels directly. */
int sdt_plan_run(void* plan, int it0, int nsteps, int graph_block, void* stream);

/* ELASTIC2 / EM2 tile shape of the single-pass step kernel (no counterpart
 * in the reference;
 * for an autotuner, seidart_amd.solver.Simulation.autotune): a block of
 * waves_per_block (1, 2 or 4) wavefronts side by side, each owning 61
 * columns (56 in the two-step ELASTIC2 kernel, 60 in the two-step EM2
 * one), walks `rows` rows (1..256).
 * Default: 4 wavefronts and the tallest of 32, 16, 8, 4 rows that gives the
 * plan at least 1024 blocks (else 2).  Results do not depend on the shape.
 * Synchronises the device when the shape changes. */
int sdt_plan_set_tiles(void* plan, int waves_per_block, int rows);
int sdt_plan_get_tiles(void* plan, int* waves_per_block, int* rows);

/* ELASTIC2 / EM2 order of the step kernels' work lists (no counterpart in
 * the reference; the autotuner's third knob beside the tile shape and the
 * steps per launch).  0 (the default) deals a model's tiles to the eight
 * XCDs in turn, so every XCD covers the whole model; 1 lists the tiles
 * column by column and gives each XCD a contiguous run of them, so a tile
 * and the one below it, which read the same halo rows, run on one XCD one
 * after the other, and each XCD keeps to one part of the model from step
 * to step (small single-shot grids gain; the large batched plans do not).
 * Either way all shots of a tile run on one XCD.  Results do not depend on
 * the order.  Synchronises the device when the order changes. */
int sdt_plan_set_tile_order(void* plan, int order);
int sdt_plan_get_tile_order(void* plan, int* order);

/* ELASTIC2 / EM2 steps per launch of the step kernel (no counterpart in the
 * reference): 2 (the default for plans of at least 4M cells per step,
 * summed over shots; smaller plans default to 1) advances two time steps in
 * one launch, keeping the state of the first in registers, so the state
 * crosses HBM once per two steps; 1 is the one-step kernel.  Results are
 * identical bit for bit.  A run of an odd number of steps ends with one
 * single step; with hipGraph blocks, two-step launches need an even
 * graph_block (odd blocks fall back to one step per launch).  The adjoint's
 * trail recomputation always steps one at a time.  In two-step mode
 * sdt_plan_profile's out_ms[0] is the mean per two-step launch and
 * out_ms[2] is 0 (the kernel writes the receiver samples itself). */
int sdt_plan_set_steps_per_launch(void* plan, int steps);
int sdt_plan_get_steps_per_launch(void* plan, int* steps);

/* Synchronise the plan's last stream and report instability: *first_step
 * gets the first step at which any shot's field exceeded 1e25 in magnitude
 * or became non-finite (seismicfdtd.f95:378-379), or -1.  Returns
 * SDT_EUNSTABLE when such a step exists.  per_shot (optional, nshots ints)
 * gets the step per shot (-1 = stable). */
int sdt_plan_check(void* plan, int* first_step, int* per_shot);

/* Allocation-free run.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  One call does what
 * sdt_plan_create + sdt_plan_run(it0, nsteps, graph_block 0) +
 * sdt_plan_check + sdt_plan_destroy do, with the same kernels and results,
 * on the caller's buffers (fields, psi, field_b / psi_b, sources, receivers,
 * traces), but calls no hipMalloc / hipFree: the run's descriptors, work
 * lists, receiver index, per-shot flags and step counters live in a static
 * device array of sdt_noalloc_scratch_bytes() bytes per GPU (8 MiB), held
 * until the process ends.  Inputs whose descriptors and work lists need more
 * are rejected with SDT_EINVAL before any device work.  Every call on one
 * GPU shares that array and resets what it uses; calls that overlap on one
 * GPU (other threads, other streams) race.  Kernels run on `stream`; the
 * call synchronises it before returning, and reports instability like
 * sdt_plan_check (first_step, per_shot optional).  No hipGraph is used;
 * two steps per launch as in sdt_plan_create's default. */
int sdt_run_noalloc(int kind, int precision, const sdt_model* models, int nmodels,
                    const sdt_shot* shots, int nshots, int nt, int it0, int nsteps, void* stream,
                    int* first_step, int* per_shot);
long long sdt_noalloc_scratch_bytes(void);

/* Measurement hook (no counterpart in the reference): run steps it0 ..
 * it0+nsteps-1 like sdt_plan_run (no graph), bracketing every kernel with
 * HIP events on `stream`, and return the mean duration in ms of
 * out_ms[0] the first field update (stress / H), out_ms[1] the second
 * (velocity / E) and out_ms[2] the receiver gather. */
int sdt_plan_profile(void* plan, int it0, int nsteps, void* stream, double* out_ms);

/* Measurement hook (no counterpart in the reference): enqueue on `stream` a
 * copy of nfloat floats from src to dst (device memory), one 4-byte element
 * per lane, the access width of the stencil kernels.  Its traffic is known
 * (4 * nfloat bytes read and written), so the HBM byte counters a profiler
 * reads for it calibrate the counters read for the stencil kernels. */
int sdt_hbm_probe(const void* src, void* dst, long long nfloat, void* stream);

/* ------------------------------------------------------------------------
 * Adjoint (gradients of a trace misfit).  No counterpart in the reference,
 * which has no adjoint or gradient code; these entry points serve an FWI
 * caller (a torch.autograd.Function in seidart_amd/adjoint.py).  ELASTIC2
 * and EM2.
 *
 * A "state" of one shot is one contiguous array of sdt_state_elems()
 * elements: the kind's fields (field order of sdt_shot, nx*nz each) followed
 * by its psi workspace.  The caller checkpoints forward states, copies one
 * into slot 0 of a trail, recomputes the following states into the trail
 * with sdt_plan_run_trail, and runs the adjoint steps over them with
 * sdt_adjoint_run, last segment first.
 * ------------------------------------------------------------------------ */

/* Per-shot adjoint buffers (device memory, plan precision unless noted).
 * trail[b]: trail_len states each; slot k at trail[b] + k * sdt_state_elems().
 *          trail[1] is optional (NULL for every shot, or set for every
 *          shot): with two buffers the recomputation of one segment can run
 *          on another stream while the adjoint walks the previous one.
 * lam[2]:  adjoint state ping-pong, one state each; lambda_t (dJ/du_t) lives
 *          in lam[t & 1].  Zero both before the first (last-in-time) call.
 * work:    sdt_adjoint_work_elems() elements of scratch.
 * resid:   dJ/dtrace, laid out like sdt_shot.trace ([nt][ncomp][nrec]);
 *          NULL = no residual.
 * grad:    FP64 [ncoef][nz][nx], ncoef = 12 (ELASTIC2) or 4 (EM2):
 *          dJ/d(kernel coefficient slot q at cell),
 *          accumulated (+=) over steps; shots of one model may share it
 *          (atomic accumulation).
 * gsrc:    FP64 [nt][3] or NULL: dJ/d(source increment of component c at
 *          step t), written per step. */
typedef struct sdt_adj_shot {
  void* trail[2];
  int trail_len;
  void* lam[2];
  void* work;
  const void* resid;
  double* grad;
  double* gsrc;
} sdt_adj_shot;

/* Elements of one shot's state (fields then psi). */
int sdt_state_elems(int kind, int nx, int ny, int nz, int lx, int ly, int lz,
                    long long* out_elems);

/* Elements of one shot's adjoint scratch. */
int sdt_adjoint_work_elems(int kind, int nx, int ny, int nz, int lx, int ly, int lz,
                           long long* out_elems);

/* Hand the plan one sdt_adj_shot per shot (synchronises the device; call
 * again only to change buffers or options).  grad_mask: bit q set =
 * accumulate coefficient slot q (0xfff = all 12 of ELASTIC2, 0xf = the 4 of
 * EM2; slots nobody differentiates cost nothing when left out).  flags: 0, or SDT_ADJ_REFERENCE to run the
 * unfused three-kernel adjoint (ELASTIC2 and EM2: one thread per cell per
 * phase, scratch planes in `work`), kept as a cross-check of the fused
 * kernels. */
#define SDT_ADJ_REFERENCE 1
int sdt_adjoint_setup(void* plan, const sdt_adj_shot* adj, int nshots, unsigned grad_mask,
                      int flags);

/* Forward recomputation: with slot 0 of trail buffer `buf` holding u_{it0},
 * run steps it0 .. it0+nsteps-1 writing u_{it0+k} into slot k (no receiver
 * gather).  Same kernels and arithmetic as sdt_plan_run.  Segments of two or
 * more steps run as a captured hipGraph. */
int sdt_plan_run_trail(void* plan, int it0, int nsteps, int buf, void* stream);

/* Adjoint steps t = t_hi-1 down to t_lo, with slot k of trail buffer `buf`
 * holding u_{trail_t0 + k} for trail_t0 + k in [t_lo, t_hi].  Step t injects
 * resid[t] at the receivers into lambda_{t+1}, then writes lambda_t and
 * accumulates grad (and gsrc).  grad is complete when the call's work has
 * run (FP32 plans stage it in FP32 within one call).  A recomputation
 * (sdt_plan_run_trail) into the other buffer may run concurrently on
 * another stream; calls of the same kind must not overlap. */
int sdt_adjoint_run(void* plan, int t_hi, int t_lo, int trail_t0, int buf, void* stream);

/* Release the plan's own device memory and graphs. */
void sdt_plan_destroy(void* plan);

/* Message of the last failure on this thread ("" if none). */
const char* sdt_last_error(void);

/* Library version string. */
const char* sdt_version(void);

#ifdef __cplusplus
}
#endif
#endif
