/*
 * tc_abi.h -- C-ABI of the MI355X Taylor-Couette RK3 time-step library
 * (libtcamd.so).
 *
 * The reference exposes its hot path as argument-less Fortran subroutines
 * that mutate module globals (param.F90:4-102).  This ABI keeps those routine
 * names and their call order; the module state becomes an opaque device
 * context.  A Fortran host binds each entry with ISO_C_BINDING.
 *
 * Conventions
 *  - Every function returns 0 on success and a negative code on error; the
 *    message is available from tc_last_error() (thread-local).
 *  - The context owns all device memory.  Host arrays passed in are never
 *    retained after the call returns.
 *  - One host thread per context.  Routine entries are asynchronous on the
 *    context's stream; entries that return host values synchronise.
 *  - Floating point is FP64 throughout.
 */
#ifndef TC_ABI_H
#define TC_ABI_H

#ifdef __cplusplus
extern "C" {
#endif

/* bou.in contents, in the order ReadInputFile.F90:17-37 reads them. */
typedef struct tc_params {
  int nthm, nzm, nrm, nsst, idtv, nread;
  int ntst;
  double tout, tmax;
  int ireset;
  double walltimemax;
  double alx3, rint, rext;
  int istr;
  double str, lamb;
  double virotint, virotext, sizepert, omegapert;
  double ren, romeg, resid, limitcfl;
  double tsta;
  int starea;
  double dtmin, dtmax, limitvel, cfllim;
  int nslabr, dpthcut, dpstave;
  double tframe2d;
  int dpfullf;
  double tframe3d;
  int minstrt, minstrz, multmin;
} tc_params;

typedef struct tc_ctx tc_ctx;

/* field ids: the local_arrays module (param.F90:97-100) */
enum {
  TC_QT = 0, TC_QR = 1, TC_QZ = 2, TC_PR = 3, TC_DPH = 4, TC_RHS = 5,
  TC_DQ = 6, TC_QCAP = 7, TC_RU1 = 8, TC_RU2 = 9, TC_RU3 = 10, TC_NFIELDS = 11
};

/* host array layouts for upload/download */
enum {
  TC_LAYOUT_FORTRAN = 0, /* (1:nr,1:nzm,1:nthm) column-major, r fastest */
  TC_LAYOUT_DEVICE = 1   /* [r][z][theta], theta fastest (device order)  */
};

/* grid-vector ids for tc_grid_vector (out[q] = v(q+1), Fortran 1-based) */
enum {
  TC_V_RC = 0, TC_V_RM, TC_V_G2RC, TC_V_G2RM, TC_V_USRC, TC_V_USRM,
  TC_V_UDX2RM, TC_V_UDX2M, TC_V_UDX2C, TC_V_RUDX2C, TC_V_UDX1VIS1,
  TC_V_UDX1VIS2, TC_V_UDX1QM, TC_V_AP1J, TC_V_AM1J, TC_V_AC1J, TC_V_AP2J,
  TC_V_AM2J, TC_V_AC2J, TC_V_AP2JE, TC_V_AM2JE, TC_V_AC2JE, TC_V_AP3J,
  TC_V_AM3J, TC_V_AC3J, TC_V_APPH, TC_V_AMPH, TC_V_ACPH, TC_V_AK1, TC_V_AK3,
  TC_V_FS, TC_V_PVOL, TC_NVEC
};

/* ---------------- lifecycle and state ---------------- */

/* InitTimeMarchScheme + InitVariables + CreateGrid(+SlipBCs) +
 * InitPressureSolver: main.F90:70-94 (InitTimeMarchScheme.F90:16-48,
 * InitVariables.F90:26-57, CreateGrid.F90:28-323, InitPressureSolver.F90:22-65).
 * Allocates the device state; all fields start at zero
 * (AuxiliaryRoutines.F90:123). */
int tc_create(const tc_params *p, tc_ctx **out);

/* Kernel schedule knobs of one context (no counterpart in the reference,
 * whose schedule is fixed by its MPI decomposition and FFTW plans).  Zero
 * fields take the TC_* environment override, else the default. */
typedef struct tc_tuning {
  int helm_lpt;   /* radial lines per thread of the fused Helmholtz kernel: 1, 2 (0: by occupancy) */
  int fft_t_lds;  /* LDS bytes of a theta-FFT block (default 32768) */
  int fft_z_lds;  /* LDS bytes of a z-FFT block (default 65536) */
  int fft_z_nt;   /* threads of a z-FFT block: 256 or 512 (default) */
  int fft_pad;    /* LDS padding of the FFT tiles: one slot per 8 points (1, default),
                     per 4 (2) or per 16 (3); -1 off */
  int poisson_pq; /* rows per thread of the radial Poisson solve: 13 (default) or 26 */
  int defer;      /* RK3 with even nthm: apply the pressure correction of substeps 1 and 2
                     inside the next substep's Helmholtz kernel: 2 always, -1 never,
                     1 from 8 Mi grid points up (default; TC_DEFER_CORRECTION 2/0/1).
                     Changes results at the rounding level only. */
} tc_tuning;

/* tc_create with an explicit schedule; tc_set_tuning re-plans an existing
 * context (fields untouched); tc_get_tuning returns the effective values. */
int tc_create_tuned(const tc_params *p, const tc_tuning *t, tc_ctx **out);
int tc_set_tuning(tc_ctx *ctx, const tc_tuning *t);
int tc_get_tuning(tc_ctx *ctx, tc_tuning *t);

/* Autotuner (no counterpart in the reference): times `steps` fused RK3
 * steps of this grid shape per candidate schedule (coordinate descent over
 * the knobs) and returns the fastest.  Results are cached per shape for the
 * life of the process. */
int tc_autotune(const tc_params *p, int steps, tc_tuning *best);
int tc_destroy(tc_ctx *ctx);
const char *tc_last_error(void);

/* Host-only CreateGrid/InitPressureSolver vectors (no device needed).
 * Returns the number of values written (<= n) or a negative error. */
int tc_grid_vector(const tc_params *p, int id, double *out, int n);

/* Copy a field in or out; replaces the reference's direct access to the
 * local_arrays module and the restart read (ReadFlowField.F90:135-141).
 * Halos do not exist on the device: z and theta are wrapped periodically. */
int tc_upload_field(tc_ctx *ctx, int field, const double *host, int layout);
int tc_download_field(tc_ctx *ctx, int field, double *host, int layout);

/* icylsl/ocylsl masks (nzm,nthm) column-major (param.F90:59). */
int tc_set_masks(tc_ctx *ctx, const int *icylsl, const int *ocylsl);
int tc_get_masks(tc_ctx *ctx, int *icylsl, int *ocylsl);

/* The scalars time and dt of module param (param.F90:15,29). */
int tc_set_time(tc_ctx *ctx, double time, double dt);
/* out = time dt beta al ga ro vinner_t dvinner_t */
int tc_get_scalars(tc_ctx *ctx, double *out8);

/* TimeMarcher.F90:20-23: beta=dt/(2 Re), InnerCylVelocity, SlipBCs. */
int tc_begin_step(tc_ctx *ctx);
/* TimeMarcher.F90:27-29: al, ga, ro of substep ns (1-based). */
int tc_set_substep(tc_ctx *ctx, int ns);

/* ---------------- the hot-path routines ---------------- */

int tc_explicit_terms_qt(tc_ctx *ctx);       /* ExplicitTermsQT.F90:11-133      */
int tc_explicit_terms_qr(tc_ctx *ctx);       /* ExplicitTermsQR.F90:13-153      */
int tc_explicit_terms_qz(tc_ctx *ctx);       /* ExplicitTermsQZ.F90:11-110      */
int tc_implicit_and_update_qt(tc_ctx *ctx);  /* ImplicitAndUpdateQT.F90:14-124 +
                                                SolveImpEqnUpdate_QT.F90:12-84 */
int tc_implicit_and_update_qr(tc_ctx *ctx);  /* ImplicitAndUpdateQR.F90:15-74  +
                                                SolveImpEqnUpdate_QR.F90:12-72 */
int tc_implicit_and_update_qz(tc_ctx *ctx);  /* ImplicitAndUpdateQZ.F90:14-123 +
                                                SolveImpEqnUpdate_QZ.F90:11-74 */
int tc_calc_local_divergence(tc_ctx *ctx);   /* CalcLocalDivergence.F90:11-44   */
int tc_solve_pressure_correction(tc_ctx *ctx);/* SolvePressureCorrection.F90:12-173 */
int tc_correct_velocity(tc_ctx *ctx);        /* CorrectVelocity.F90:12-50       */
int tc_correct_pressure(tc_ctx *ctx);        /* CorrectPressure.F90:11-52       */

/* TimeMarcher.F90:12-57: one RK3 (or AB2) step at the current dt,
 * including tc_begin_step.  Uses fused kernels unless tc_set_fused(ctx,0).
 * The fused RK3 step (even nthm) applies the pressure correction of
 * substeps 1 and 2 inside the next substep's Helmholtz kernel; the fields
 * are corrected when the call returns, equal to the substep-by-substep
 * schedule to rounding (TC_DEFER_CORRECTION=0 at tc_create selects that
 * schedule). */
int tc_time_marcher(tc_ctx *ctx);
int tc_set_fused(tc_ctx *ctx, int fused);

/* ---------------- reductions (synchronise) ---------------- */

int tc_calc_max_cfl(tc_ctx *ctx, double *cflm);        /* CalcMaxCFL.F90:12-49     */
int tc_check_divergence(tc_ctx *ctx, double *qmax);    /* CheckDivergence.F90:11-46 */

/* tc_check_divergence without any allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops).  No
 * counterpart in the reference.  The reduction's per-block partials live in
 * a fixed 512 KiB device array of the library (one per GPU, kept until the
 * process ends) instead of the context's buffer, and the result, max|div|,
 * is written to qmax_dev, a double in device memory, asynchronously on the
 * context's stream (read it after tc_synchronize).  Planes of more than
 * 16,777,216 points (nzm*nthm) are rejected.  Calls that overlap on one GPU
 * share that array and race. */
int tc_check_divergence_noalloc(tc_ctx *ctx, double *qmax_dev);
int tc_calc_torque(tc_ctx *ctx, double *inner, double *outer); /* CalcTorque.F90:12-62 */
int tc_calc_velocity_stats(tc_ctx *ctx, double *vmax3, double *rms3); /* CalcVelocityStats.F90:11-68 */

/* ---------------- driver helpers ---------------- */

/* One iteration of the main.F90:152-259 loop body: CFL, dt rule (fixed-dt
 * mode uses dt = DTMAX), TimeMarcher, time advance, logging-step
 * diagnostics and the error codes 165/166/169/266/333/334/555.
 * out9 = cfl dt time dmax torque_in torque_out vmax1 vmax2 vmax3.
 * Returns the error code (>= 0) or a negative ABI error. */
int tc_step(tc_ctx *ctx, int ntime, double walltime, double *out9);

/* The main loop from ntime0 for at most nsteps iterations of tc_step, ending
 * at the first nonzero error code (main.F90:152-261 without the host
 * between steps).  walltime_start is the wall-clock time already elapsed
 * when the call starts (seconds); the elapsed time is read after each
 * iteration's work, as main.F90:208,220 does, and above WALLTIMEMAX gives
 * code 334 (555 still wins).  On return *ntime_last is the last iteration
 * run and out9 its diagnostics; the result is that iteration's error code,
 * or a negative ABI error. */
int tc_run_steps(tc_ctx *ctx, int ntime0, int nsteps, double walltime_start, int *ntime_last,
                 double *out9);

/* tc_run_steps that also keeps what the reference writes at every logging
 * iteration (torquenu.out, vmax.out, the divergence print, main.F90:180-216):
 * row r of log10 (10 doubles) = ntime, then that iteration's out9.  At most
 * log_cap rows are stored; *nlog = the number of logging iterations run
 * (may exceed log_cap).  No counterpart in the reference: its loop writes
 * these lines itself between steps. */
int tc_run_steps_logged(tc_ctx *ctx, int ntime0, int nsteps, double walltime_start,
                        int *ntime_last, double *out9, double *log10, int log_cap, int *nlog);

/* Divergence-free projection of the current velocity with al = dt = 1
 * (CalcLocalDivergence + SolvePressureCorrection + CorrectVelocity), used
 * to build starting fields; pr is left untouched. */
int tc_project(tc_ctx *ctx);

/* Seeded starting field in Fortran layout (nr,nzm,nthm): laminar Couette
 * qt = A rm + B/rm (CreateInitialConditions.F90:21-30) plus i.i.d.
 * U(-amp,amp) on qt (k=1..nrm), qr (k=2..nrm) and qz (k=1..nrm) from a
 * splitmix64 stream; the k=nr plane and the qr walls are zero. */
int tc_init_perturbed_couette(const tc_params *p, unsigned long long seed,
                              double amp, double *qt, double *qr, double *qz);

int tc_synchronize(tc_ctx *ctx);

/* Per-kernel timing with HIP events on the context stream.  When enabled,
 * each launch of a timed kernel is bracketed by two events; read returns the
 * summed milliseconds and launch count since the last reset. */
int tc_kernel_timing(tc_ctx *ctx, int enable);
/* Time only the launches named kernel (NULL: timing off); fewer events in
 * the stream when one kernel's duration is wanted inside a timed region. */
int tc_kernel_timing_only(tc_ctx *ctx, const char *kernel);
int tc_kernel_timing_read(tc_ctx *ctx, const char *kernel, double *total_ms,
                          long long *count, int reset);

/* Device-to-device copy of one whole field (nr planes) into another, with a
 * plain streaming kernel; used as the reference point when converting
 * hardware byte counters into bytes per launch. */
int tc_copy_field(tc_ctx *ctx, int dst, int src);

/* STREAM triad a = b + s*c over three whole fields: the achievable
 * two-read-one-write bandwidth of the box, reported beside the copy. */
int tc_stream_triad(tc_ctx *ctx, int a, int b, int c, double s);

/* Device bytes held by the context (fields + scratch). */
int tc_device_bytes(tc_ctx *ctx, unsigned long long *bytes);

/* ---------------- checkpoint / restart ---------------- */

/* Whole-state checkpoint in one raw file: WriteFlowField.F90:11-53 writes
 * qt, qr, qz and the master scalars (time, grid sizes and geometry); this
 * also keeps pr, the RK histories ru1..ru3, ntime and the CFL of the last
 * logging pass, so that a restarted run continues bit for bit (the
 * reference's restart, ReadFlowField.F90:135-144, restarts with pr = 0).
 * Layout: 8-byte magic "TCAFIDR1", int32 nthm nzm nrm nsst ntime cfl_valid
 * istr pad, float64 time dt cfl rint rext alx3 lamb str, then the fields
 * qt qr qz pr ru1 ru2 ru3 in device order ([r][z][theta], nr planes each).
 * tc_read_restart rejects a file of another grid or geometry. */
int tc_write_restart(tc_ctx *ctx, const char *path, int ntime);
int tc_read_restart(tc_ctx *ctx, const char *path, int *ntime);

/* Restart onto another grid (ReadFlowField.F90:63-131, interp.F90:11-303):
 * interpolates one field (TC_QT, TC_QR or TC_QZ) of a continuation file
 * written on an old grid into this context.  old is (nro, nzo-1, ntho-1) in
 * Fortran order, where ntho, nzo, nro are the old nth, nz, nr (nthm+1, ...);
 * istro, stro are the old grid's ISTR and STR.  As in the reference: the new
 * resolution may be at most twice the old one in each direction; the radial
 * grids come from interp.F90's gridnew, which builds ISTR 0 and 6 only; the
 * result is bilinear in theta and z and takes the lower bracketing radial
 * plane.  Coarsening by 2 or more in theta or z is rejected (the reference
 * reads outside its buffer there). */
int tc_regrid_field(tc_ctx *ctx, int field, const double *old, int ntho, int nzo, int nro,
                    int istro, double stro);

/* ---------------- statistics ---------------- */

/* The stat_arrays accumulators of StatRoutines.F90 and CalcDissipation.F90,
 * in WriteStats order (StatRoutines.F90:97-152; the file stems are the
 * names); each is an (nrm, nzm) array in Fortran order (r fastest). */
enum {
  TC_ST_VT_MEAN = 0, TC_ST_VR_MEAN, TC_ST_VZ_MEAN, TC_ST_WT_MEAN, TC_ST_WR_MEAN,
  TC_ST_WZ_MEAN, TC_ST_VT_RMS, TC_ST_VR_RMS, TC_ST_VZ_RMS, TC_ST_WT_RMS, TC_ST_WR_RMS,
  TC_ST_WZ_RMS, TC_ST_VTVR_MEAN, TC_ST_VTVZ_MEAN, TC_ST_VRVZ_MEAN, TC_ST_PR_MEAN,
  TC_ST_PR_RMS, TC_ST_DISSTE, TC_NSTATS
};

/* CalcStats (StatRoutines.F90:12-71) with the vorticity of CalcVorticity
 * (CalcVorticity.F90:11-88): adds one sample of the theta-means to the
 * accumulators and counts it (nstatsamples). */
int tc_calc_stats(tc_ctx *ctx);
/* CalcDissipation (CalcDissipation.F90:11-153): adds dissipation/Re/nthm
 * to disste and returns nute, the volume-mean dissipation over the laminar
 * Couette value (written to dissipnu.out by the reference, :150). */
int tc_calc_dissipation(tc_ctx *ctx, double *nute);
/* The accumulators, TC_NSTATS x nrm x nzm doubles, and the sample count
 * (the data WriteStats reduces and writes, StatRoutines.F90:73-171). */
int tc_get_stats(tc_ctx *ctx, double *out, int *nsamples);
/* Replace the accumulators and the count: continuing statistics from
 * earlier files (StatReadReduceWrite.F90:27-31 adds the old ones). */
int tc_set_stats(tc_ctx *ctx, const double *in, int nsamples);
/* nute of the last CalcDissipation (tc_calc_dissipation or the statistics
 * of tc_step once time > TSTA, main.F90:188-190); NaN before the first. */
int tc_last_dissipation(tc_ctx *ctx, double *nute);

/* ---------------- slab dumps ---------------- */

/* DumpRCuts.F90:28-36 at r-plane kslab (1..nrm): qt, the cell-centre mean
 * of the radial velocity (qr/r on the two faces), qz; each (nzm, nthm) in
 * Fortran order.  InitDumpRCuts.F90 picks kslab from rcuts.in; that
 * selection is host work. */
int tc_dump_rcut(tc_ctx *ctx, int kslab, double *vt, double *vr, double *vz);
/* DumpThCuts.F90:33-44: the theta index 1 cut of qt, qr/rc, qz; each
 * (nrm, nzm) in Fortran order. */
int tc_dump_thcut(tc_ctx *ctx, double *vt, double *vr, double *vz);
/* DumpStreamwiseAvg.F90:33-44: theta-means of qt, qr/rc, qz; each
 * (nrm, nzm) in Fortran order. */
int tc_dump_streamwise_avg(tc_ctx *ctx, double *vt, double *vr, double *vz);

/* ---------------- batched sweep (no counterpart in the reference) ---------------- */

/* Many independent instances, each a tc_ctx of its own shape and
 * parameters, advanced together: instances of one grid shape (and schedule)
 * form a group, and every kernel of the fused RK3 step runs as ONE grouped
 * launch per group (blockIdx selects the instance through a device table of
 * per-instance descriptors).  A step of all groups can be captured once as a
 * hipGraph and replayed.  Instances advance at a fixed dt (IDTV=0, dt =
 * DTMAX) with the main.F90 loop semantics: CFL guard 165 every step,
 * logging-step diagnostics with 169/266, 333 past TMAX, 555 after NTST.
 * Device-detected codes are read back every `sync_every` steps; the report
 * names the first failing step, as the reference would exit there. */
typedef struct tc_batch tc_batch;

int tc_batch_create(tc_batch **out);
int tc_batch_destroy(tc_batch *b);
/* Adds an instance (tc_create_tuned); t may be NULL.  *id identifies it. */
int tc_batch_add(tc_batch *b, const tc_params *p, const tc_tuning *t, int *id);
/* Removes an instance and frees its device memory. */
int tc_batch_remove(tc_batch *b, int id);
/* The instance's context, for uploads, downloads, projection and the
 * single-instance reductions; waits for the batch's queued work first.
 * Work queued on the context itself is waited for by the next advance. */
int tc_batch_ctx(tc_batch *b, int id, tc_ctx **ctx);
/* Advances every unfinished instance by up to nsteps steps (each stops at
 * its own end); use_graph: replay one captured hipGraph per step.  Returns
 * the number of unfinished instances, or a negative error. */
int tc_batch_advance(tc_batch *b, int nsteps, int use_graph, int sync_every);
/* Progress of one instance: steps taken, error code (0 while running; the
 * main.F90 codes once finished), and out9 = cfl dt time dmax torque_in
 * torque_out vmax1 vmax2 vmax3 of its last logging step. */
int tc_batch_status(tc_batch *b, int id, int *steps, int *code, double *out9);
/* Number of shape groups and grouped launches per step of the current
 * membership (for reporting). */
int tc_batch_groups(tc_batch *b, int *ngroups, int *launches_per_step);
/* WriteFlowField.F90:11-53 / ReadFlowField.F90 for one batch instance: its
 * whole state (tc_write_restart's format) at its current step, and the
 * load of such a checkpoint into an instance that has not stepped yet, which
 * then continues from the checkpoint's step and time bit for bit.  The load
 * is refused, with the instance left as it was, when the checkpoint's dt is
 * not the instance's fixed dt or its step is past the instance's NTST. */
int tc_batch_write_restart(tc_batch *b, int id, const char *path);
int tc_batch_read_restart(tc_batch *b, int id, const char *path, int *ntime);

#ifdef __cplusplus
}
#endif
#endif /* TC_ABI_H */
