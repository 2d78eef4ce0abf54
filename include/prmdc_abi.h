/*
 * prmdc_abi.h -- the C ABI of libmdgpu_pr.so: pr-mdc's force routines
 * (pr-mdc, the Parrinello-Rahman constant-pressure variant of mdc) computed
 * on the GPU over pr-mdc's own global objects.
 *
 * A maintainer links pr-mdc without real.c and ewald.c and adds
 * -lmdgpu_pr; pr-mdc's rv.c (calc_foc), cell.c (calc_sigma, set_cutoff,
 * cell_force, cyclic_bc_dr), pt.c (control_press_PR) and the rest stay its
 * own.  The structs below are laid out field for field as pr-mdc's
 * (pr-mdc/program/md.h:25-163); pr-mdc's md.h defines the objects, and
 * ELF interposition binds this library's references to them.
 *
 * This header and mdc_abi.h name different structs with the same names
 * (mdc and pr-mdc are different programs): include one of them, never both.
 */
#ifndef PRMDC_ABI_H
#define PRMDC_ABI_H

#ifdef __cplusplus
extern "C" {
#endif

/* pr-mdc/program/md.h:25-43 */
typedef struct {
  int skip_ewald;
  int calc_max;
  double delta_time_fs;
  int kinds_of_ions;
  int natoms_in_mol_unit;
  int natoms_in_unit_cell;
  int kinds_of_interaction;
} calc_control;

/* pr-mdc/program/md.h:45-85.  The 3x3 matrices are pr-mdc's dmat2d
 * row-pointer arrays (init.c:240-253): m[i][j]. */
typedef struct {
  int nx, ny, nz;
  double vol;
  double *rx, *ry, *rz;
  double *sx, *sy, *sz;
  double *vx, *vy, *vz;
  double *fx0, *fy0, *fz0;
  double *fx, *fy, *fz;
  double **hu, **hur;
  double **hm, **hmr;
  double **hmt, **hmtr;
  double **hmv, **hmvt;
  double **hmf;
  double **g, **sigma;
  double Lx, Ly, Lz;
  double angle_ab, angle_bc, angle_ca;
} cell_parameters;

/* pr-mdc/program/md.h:87-140 */
typedef struct {
  int *ion;
  int N;
  int step;
  double dt, dt2;
  double kB, kk;
  double pot, kin;
  double perMol;
  double e2t, pp2gpa;
  double a1;
  int hm, hm_sqrt;
  double radius, radius2;
  double cutoff, cutoff2;
  double *rcsi, *rsni;
  int table_division;
  double *pe1r1;
  double *fe1r1;
  int *lookup;
} system_property;

/* pr-mdc/program/md.h:142-158 */
typedef struct {
  int t_control_step;
  int p_control_step;
  double pres_scalling_factor;
  double temp_set;
  double temp_set_initial;
  double pres_trace;
  double **pres_set;
  double **pres_tensor;
  double **virial;
  double **kin_tensor;
} pressure_temperature_set;

/* pr-mdc/program/md.h:160-163 */
typedef struct {
  double *m;
  double *Z;
} ion_property;

/* pr-mdc/program/md.h:192-198 */
typedef struct {
  double mass;
  double dt;
  double **v;
  double **f;
  double **f0;
} Parrinello_Rahman_set;

/* pr-mdc/program/md.h:224-234: the objects the routines below read and
 * write (pr-mdc's md.h defines them; so does this library, for callers
 * without pr-mdc) */
extern calc_control ctl;
extern cell_parameters cell;
extern system_property sys;
extern pressure_temperature_set pt;
extern ion_property ion;
extern Parrinello_Rahman_set pr;

/*
 * The force routines.  Each reads the positions cell.r*, the species
 * sys.ion, the charges ion.Z, the cell matrices cell.hm and cell.hmr (and
 * cell.g, cell.vol for ewald2) as pr-mdc has set them, and adds to the
 * forces cell.f* what pr-mdc's routine adds.
 *
 * real_space: Born-Mayer-Huggins pairs from the table sys.pe1r1/fe1r1
 *   within sys.cutoff under the triclinic minimum image; sets sys.pot and
 *   the virial tensor pt.virial (pr-mdc/program/real.c:11-77).
 * ewald1: the Ewald real-space term, Mori's erfc per pair; adds to sys.pot
 *   and pt.virial (pr-mdc/program/ewald.c:11-68).
 * ewald2: the reciprocal-space sum over the half sphere |h|^2 <= sys.hm of
 *   the lattice with reciprocal vectors cell.g; adds to sys.pot
 *   (ewald.c:75-159).  sys.rcsi and sys.rsni, pr-mdc's per-vector scratch
 *   arrays, are not written.
 * ewald3: the self term, adds -sum Z^2 a1/sqrt(pi) to sys.pot
 *   (ewald.c:162-173).
 * calc_alpha: pr-mdc's alpha scan (ewald.c:178-193): 200 passes of
 *   ewald1-3 from alpha = 0.01 in steps of 0.01, printing alpha and the
 *   potential, then exit(0) as pr-mdc does.
 *
 * A routine that fails (no device, an allocation, a pair inside the
 * table's start) prints to stderr and exits with EXIT_FAILURE, as pr-mdc
 * does on its fatal errors (cell.c:218-221), since pr-mdc's main.c checks
 * no status.  After mdg_pr_set_error_mode(MDG_DROPIN_STATUS), or with
 * MDGPU_DROPIN_ERRORS=status in the environment, it instead prints once,
 * leaves the outputs unchanged and sets the status mdg_pr_status returns.
 */
void real_space(void);
void ewald1(void);
void ewald2(void);
void ewald3(void);
void calc_alpha(void);

/* No counterpart in pr-mdc: what a failing routine does, MDG_DROPIN_EXIT
 * (0, the default: print, then exit(EXIT_FAILURE)) or MDG_DROPIN_STATUS
 * (1: print once, keep a sticky status).  Returns 0 or MDG_E_ARG. */
#ifndef MDG_DROPIN_EXIT
#define MDG_DROPIN_EXIT   0
#define MDG_DROPIN_STATUS 1
#endif
int mdg_pr_set_error_mode(int mode);

/* No counterpart in pr-mdc: 0 while every routine succeeded, else the
 * first error code (the MDG_E_* codes of mdgpu.h: 1 no device, 2 HIP
 * error, 3 bad argument, 4 device allocation, 7 a pair inside the table's
 * 0.5 A start or a cutoff past the table's end). */
int mdg_pr_status(void);
/* No counterpart in pr-mdc: clears that status (and re-arms the one
 * message per failure). */
void mdg_pr_clear_status(void);

/* No counterpart in pr-mdc: how many partner splits the pair routines and
 * how many reciprocal-vector splits ewald2's back-projection spread each
 * 64-ion chunk over (0 = chosen from the size, the default).  With 1, each
 * ion's force is summed in pr-mdc's order.  Returns 0 or MDG_E_ARG. */
int mdg_pr_set_splits(int pair, int kspace);

/*
 * No counterpart in pr-mdc: times the device work of one force evaluation
 * (real_space, ewald1, ewald2, ewald3) on the state the last calls left on
 * the device, `reps` times, inputs resident (no host copies).  Writes the
 * average milliseconds per evaluation to *ms_eval and, when `ms_kernel` is
 * not NULL, the average per routine in the order above (4 values).  The
 * device force buffer holds garbage afterwards; the next routine call
 * uploads the forces again.  Returns 0 or an error code.
 */
int mdg_pr_profile(int reps, double *ms_eval, double *ms_kernel);

/*
 * No counterpart as one routine: `nsteps` steps of pr-mdc's step loop
 * (pr-mdc/program/main.c:100-111) with the state resident on the device --
 * next_rv_verlet (rv.c:12-17: clear_foc, calc_sigma, set_cutoff,
 * real_space, cell_force, ewald1-3, next_v, next_r with cyclic_bc),
 * control_press_PR (pt.c:27-63, Parrinello-Rahman cell dynamics and
 * calc_cell), control_temp(pt.t_control_step, pt.temp_set) (pt.c:124-150),
 * calc_kin and calc_press -- reading pr-mdc's globals at the start
 * (including pr, the cell's mass, time step, velocity and forces) and
 * writing back at the end what those routines write: positions, scaled
 * positions, velocities, forces and old forces, the cell matrices, volume,
 * edge lengths and angles, pr's velocity and forces, the pressure, virial
 * and kinetic tensors, pot, kin, a1, the cutoff and sys.step.  MSD and file
 * output, which pr-mdc's loop also runs, are not part of it; the
 * thermostat's running average starts afresh at each call.  When `record`
 * is not NULL it receives per step {pot, kin, pres_trace, vol, hm[3][3]}
 * (13 doubles).  Returns 0, MDG_E_RANGE (a pair inside the table's start),
 * MDG_E_STATE (a singular cell matrix, where pr-mdc exits) or another code.
 */
int mdg_pr_run(int nsteps, double *record);

#ifdef __cplusplus
}
#endif
#endif
