/*
 * mdc_abi.h -- drop-in C ABI of libmdgpu for mdc's force path.
 *
 * mdc's plugin interface is link-time replacement of translation units:
 * every routine is `void name(void)` and communicates through the global
 * objects `sys`, `ion`, `ctl` and `hm` (mdc/program/md.h:197-202,
 * potential.h:28).  libmdgpu exports the force-path routines with the same
 * names and signatures, computing on the GPU and leaving the results in
 * the same global objects.  A program built from mdc's sources links
 * libmdgpu in place of real.o, rcprcl.o, rv.o, pt.o and msd.o.
 *
 * The struct layouts below are field-for-field the layouts of md.h and
 * potential.h, so that the globals of an mdc executable (which interpose
 * on libmdgpu's own definitions) are read and written in place.  When the
 * library is loaded without such an executable (ctypes, a new C driver),
 * its own definitions of the globals are used.
 *
 * Semantics per call: the host arrays are authoritative.  Each routine
 * copies the host arrays it reads to the device, computes there, and
 * copies the arrays and scalars it writes back before returning, so host
 * code between calls (thermostats, I/O, analysis) sees current values.
 * Errors follow mdc's convention for fatal errors (files.c:9-24): the
 * first failure is printed to stderr and the process exits with
 * EXIT_FAILURE, because mdc's unmodified main.c checks no status and would
 * otherwise step on with stale arrays.  A host that checks status opts out
 * with mdg_dropin_set_error_mode(MDG_DROPIN_STATUS) or by running with
 * MDGPU_DROPIN_ERRORS=status in the environment: then a failed routine
 * leaves its outputs as they were and the error stays in
 * mdg_dropin_status() until cleared.
 */
#ifndef MDC_ABI_H
#define MDC_ABI_H

#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* mdc/program/md.h:23-51 */
typedef struct {
  double delta_time_fs;
  int    calc_max;
  double temp;
  int    t_control_step;
  int    p_control_step;
  double set_press_GPa_X, set_press_GPa_Y, set_press_GPa_Z;
  double press_X, press_Y, press_Z;
  int    kinds_of_ions;
  int    natoms_in_mol_unit;
  int    natoms_in_unit_cell;
  int    kinds_of_interaction;
} calc_control;

/* mdc/program/md.h:53-139 */
typedef struct {
  double *rx, *ry, *rz;
  double *vx, *vy, *vz;
  double *fx0, *fy0, *fz0;
  double *fx, *fy, *fz;
  short  *ion;
  short   N;
  short   nx, ny, nz;
  double  Ax, Ay, Az;
  double  Lx, Ly, Lz;
  double  kB;
  double  kk;
  double  perMol;
  double  e2t, e2tX, e2tY, e2tZ;
  double  pp2gpa;
  int     step;
  double  dt;
  double  dt2;
  double  a1;
  double  p2a2;
  double  a3;
  short   hm;
  short   hm_sqrt;
  double  radius;
  double  radius2;
  double *rcsi, *rsni;
  int     table_division;
  double *pe1r1;
  double *fe1r1;
  int    *lookup;
  double  virX, virY, virZ;
  double  presX, presY, presZ, pres;
  double  pres_scalling_factor;
  double  pot;
  double  kin, kinX, kinY, kinZ;
} system_property;

/* mdc/program/md.h:141-144 */
typedef struct {
  double *m;
  double *Z;
} ion_property;

/* mdc/program/potential.h:2-28 */
typedef struct {
  double b;
  double c_na_na, c_na_cl, c_cl_cl;
  double sigma_na, sigma_cl;
  double rho;
  double C_na_na, C_na_cl, C_cl_cl;
  double D_na_na, D_na_cl, D_cl_cl;
} huggins_mayer_potential_set;

/* mdc/program/md.h:152-172 (calc_msd's state; init.c:143-151 allocates
 * the arrays: N + 1 doubles each, kinds_of_ions entries for value and
 * number_of_ion) */
typedef struct {
  int    *number_of_ion;
  double *rx_old, *ry_old, *rz_old;
  double *dx, *dy, *dz;
  double *value;
  double  Lx2, Ly2, Lz2;
} msd_calculation_variables;

/* mdc/program/md.h:197-202 and potential.h:28 */
extern system_property sys;
extern ion_property ion;
extern calc_control ctl;
extern huggins_mayer_potential_set hm;
extern msd_calculation_variables msd;

/* ---- force path, mdc/program/prototypes.h (same names, same meaning) ---- */
void   calc_foc(void);          /* prototypes.h:53  rv.c:12-18        */
void   clear_foc(void);         /* prototypes.h:34  init.c:179-190    */
void   real_space(void);        /* prototypes.h:49  real.c:15-84      */
void   reciprocal_space(void);  /* prototypes.h:45  rcprcl.c:11-97    */
void   reciprocal_space3(void); /* prototypes.h:46  rcprcl.c:108-133  */
void   next_rv_verlet(void);    /* prototypes.h:54  rv.c:26-31        */
void   next_r(void);            /* prototypes.h:55  rv.c:46-66        */
void   next_v(void);            /* prototypes.h:56  rv.c:33-44        */
void   calc_kin(void);          /* prototypes.h:57  rv.c:68-86        */
void   moment_correction(void); /* prototypes.h:58  rv.c:88-116       */
void   calc_press(void);        /* prototypes.h:40  pt.c:8-21         */
void   mk_table(void);          /* prototypes.h:14  control.c:183-248 (host; NaCl/KCl form) */
double erfcc(double x);         /* prototypes.h:18  ext.c:73-104      */

/* ---- the rest of mdc's step loop (main.c:44-60) ---- */
void   calc_msd(void);          /* prototypes.h:62  msd.c:20-81       */
void   control_temp(int d_step, double temp);  /* prototypes.h:42  pt.c:85-112 */

/* ---- additions of the drop-in (no counterpart in mdc) ---- */
#ifndef MDG_DROPIN_EXIT
#define MDG_DROPIN_EXIT   0   /* print, then exit(EXIT_FAILURE) (default) */
#define MDG_DROPIN_STATUS 1   /* print once, keep a sticky status          */
#endif
/* Chooses what a failing routine does; returns 0 or MDG_E_ARG (3). */
int    mdg_dropin_set_error_mode(int mode);
/* 0 = ok, otherwise an MDG_E* code of mdgpu.h; sticky until cleared */
int    mdg_dropin_status(void);
void   mdg_dropin_clear_status(void);
/* Free the device mirrors of the globals (they are rebuilt on next use). */
void   mdg_dropin_release(void);

#ifdef __cplusplus
}
#endif
#endif
