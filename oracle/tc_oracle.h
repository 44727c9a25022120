/*
 * tc_oracle.h -- CPU restatement of the Taylor-Couette AFiD RK3 time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity reference for the
 * MI355X implementation.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product path never does.
 *
 * Layout follows the reference exactly (Fortran column-major, r fastest):
 *   halo'd fields qt,qr,qz,pr,dph : (1:nr, 0:nzm+1, 0:nthm+1)  InitVariables.F90:26-35
 *   plain fields  rhs,dq,qcap,ru1..3: (1:nr, 1:nzm, 1:nthm)     InitVariables.F90:40-51
 * Every routine cites the reference file:line it restates.
 */
#ifndef TC_ORACLE_H
#define TC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* bou.in contents, in the order ReadInputFile.F90:17-37 reads them. */
typedef struct or_params {
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
} or_params;

typedef struct or_state or_state;

/* field ids (same numbering as the product C-ABI) */
enum { OR_QT = 0, OR_QR, OR_QZ, OR_PR, OR_DPH, OR_RHS, OR_DQ, OR_QCAP,
       OR_RU1, OR_RU2, OR_RU3, OR_NFIELDS };

/* grid-vector ids for or_get_vec (1-based Fortran vectors, returned as
 * out[0..n-1] = v(1..n)) */
enum { OR_V_RC = 0, OR_V_RM, OR_V_G2RC, OR_V_G2RM, OR_V_USRC, OR_V_USRM,
       OR_V_UDX2RM, OR_V_UDX2M, OR_V_UDX2C, OR_V_RUDX2C, OR_V_UDX1VIS1,
       OR_V_UDX1VIS2, OR_V_UDX1QM, OR_V_AP1J, OR_V_AM1J, OR_V_AC1J,
       OR_V_AP2J, OR_V_AM2J, OR_V_AC2J, OR_V_AP2JE, OR_V_AM2JE, OR_V_AC2JE,
       OR_V_AP3J, OR_V_AM3J, OR_V_AC3J, OR_V_APPH, OR_V_AMPH, OR_V_ACPH,
       OR_V_AK1, OR_V_AK3, OR_V_FS, OR_V_PVOL, OR_NVEC };

or_state *or_create(const or_params *p);
void or_destroy(or_state *s);
const char *or_last_error(void);

int or_get_vec(or_state *s, int id, double *out, int n);
/* fields as (nr, nzm, nthm) Fortran order, no halos */
int or_get_field(or_state *s, int id, double *out);
int or_set_field(or_state *s, int id, const double *in);
void or_get_masks(or_state *s, int *icyl, int *ocyl);
void or_set_masks(or_state *s, const int *icyl, const int *ocyl);

void or_set_time(or_state *s, double time, double dt);
void or_get_scalars(or_state *s, double *out8); /* time dt beta al ga ro vinner dvinner */
void or_begin_step(or_state *s);          /* TimeMarcher.F90:20-23 */
void or_set_substep(or_state *s, int ns); /* TimeMarcher.F90:27-29 */

void or_explicit_terms_qt(or_state *s);
void or_explicit_terms_qr(or_state *s);
void or_explicit_terms_qz(or_state *s);
void or_implicit_and_update_qt(or_state *s);
void or_implicit_and_update_qr(or_state *s);
void or_implicit_and_update_qz(or_state *s);
void or_update_halo(or_state *s, int id);
void or_calc_local_divergence(or_state *s);
void or_solve_pressure_correction(or_state *s);
void or_correct_velocity(or_state *s);
void or_correct_pressure(or_state *s);
void or_time_marcher(or_state *s);
double or_calc_max_cfl(or_state *s);
double or_check_divergence(or_state *s);
void or_calc_torque(or_state *s, double *inner, double *outer);
void or_calc_velocity_stats(or_state *s, double *vmax3, double *rms3);

/* stat_arrays of StatRoutines.F90 / CalcDissipation.F90, in WriteStats
 * order; each is (nrm, nzm), Fortran order. */
enum { OR_ST_VT_ME = 0, OR_ST_VR_ME, OR_ST_VZ_ME, OR_ST_WT_ME, OR_ST_WR_ME, OR_ST_WZ_ME,
       OR_ST_VT_RMS, OR_ST_VR_RMS, OR_ST_VZ_RMS, OR_ST_WT_RMS, OR_ST_WR_RMS, OR_ST_WZ_RMS,
       OR_ST_VTVR_ME, OR_ST_VTVZ_ME, OR_ST_VRVZ_ME, OR_ST_PR_ME, OR_ST_PR_RMS, OR_ST_DISSTE,
       OR_NSTATS };
void or_calc_stats(or_state *s);          /* CalcStats incl. CalcVorticity */
double or_calc_dissipation(or_state *s);  /* CalcDissipation; returns nute */
int or_get_stats(or_state *s, double *out, int *nsamples);
double or_last_nute(or_state *s);         /* NaN before the first call */
/* restart onto another grid (ReadFlowField.F90:63-131, interp.F90);
 * intvar 1 qt, 2 qr, 3 qz; old (nro, nzo-1, ntho-1), out (nr, nzm, nthm);
 * ntho, nzo, nro are the old nth, nz, nr. */
int or_regrid(or_state *s, int intvar, const double *old, int ntho, int nzo, int nro, int istro,
              double stro, double *out);
/* slab dumps (DumpRCuts, DumpThCuts, DumpStreamwiseAvg) */
void or_dump_rcut(or_state *s, int kslab, double *vt, double *vr, double *vz);
void or_dump_thcut(or_state *s, double *vt, double *vr, double *vz);
void or_dump_streamwise_avg(or_state *s, double *vt, double *vr, double *vz);

/* Initial-field projection used to build divergence-free starting fields:
 * CalcLocalDivergence + SolvePressureCorrection + CorrectVelocity with
 * al = dt = 1, pr untouched. */
void or_project(or_state *s);

/* One iteration of the main.F90:152-259 loop; returns the error code
 * (0 = continue).  outs: cfl, dt, time, dmax, torque_in, torque_out,
 * vmax1..3. */
int or_step(or_state *s, int ntime, double walltime, double *outs9);

/* host STREAM triad bandwidth, GB/s (benchmark context only) */
double or_stream_triad(long long n, int reps);

/* LAPACK restatements (partial pivoting), 0-based arrays. */
int or_dgttrf(int n, double *dl, double *d, double *du, double *du2, int *ipiv);
void or_dgttrs(int n, const double *dl, const double *d, const double *du,
               const double *du2, const int *ipiv, double *b);
int or_zgttrf(int n, double *dl, double *d, double *du, double *du2, int *ipiv);
void or_zgttrs(int n, const double *dl, const double *d, const double *du,
               const double *du2, const int *ipiv, double *b);
/* complex arrays above are interleaved (re, im) */

/* FFT restatement: unnormalised DFT, sign -1 forward / +1 backward. */
void or_fft_c2c(int n, double *x, int sign);
void or_fft_r2c(int n, const double *in, double *out);
void or_fft_c2r(int n, const double *in, double *out);

#ifdef __cplusplus
}
#endif
#endif
