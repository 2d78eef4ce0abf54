/*
 * ref_harness.c -- drives the unmodified mdc sources as a shared library.
 *
 * TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together
 * with mdc/program/{ext,pt,control,init,real,rcprcl,rv,msd,files}.c exactly as they
 * lie in the reference checkout (include path pointed there, -fcommon
 * because md.h defines its globals) into oracle/_ref/libmdcref.so.  It is
 * used to pin the oracle restatement bit for bit and to write golden
 * fixtures; nothing in the product links it.
 *
 * Only the reference's own routines compute anything here: this file
 * sets parameters, copies arrays in and out, and sequences calls the way
 * mdc/program/main.c:14-33 does (minus thermostat and I/O).
 */
#include <stdlib.h>
#include <string.h>
#include "md.h"
#include "prototypes.h"

/* reciprocal_space3 caches -a3*sum Z^2 in statics and refreshes the cache
 * only when N differs from the N it saw at step 1 (rcprcl.c:108-133).  A
 * second setup in the same process would reuse a stale value, so make
 * the cache see a different N once, then the true one. */
static void refresh_self_term_cache(void)
{
  short n = sys.N;
  int st = sys.step;
  double pot = sys.pot;
  sys.step = 1; sys.N = (short)(n + 1); reciprocal_space3();
  sys.step = 2; sys.N = n;              reciprocal_space3();
  sys.pot = pot; sys.step = st;
}

/* Configure an nx*ny*nz rock-salt NaCl box with lattice constant a [A],
 * Ewald alpha [1/A], |h|^2 cutoff hm, real-space cutoff radius [A] and
 * time step dt_fs, through mdc's own init chain (init.c:10-20 without the
 * time-seeded velocities) and mk_table.  Returns N. */
int ref_setup(int nx, int ny, int nz, double a, double alpha, int hm,
              double radius, double dt_fs)
{
  get_control_param();
  sys.Ax = sys.Ay = sys.Az = a;
  sys.nx = (short)nx; sys.ny = (short)ny; sys.nz = (short)nz;
  sys.a1 = alpha;
  sys.hm = (short)hm;
  sys.radius = radius;
  ctl.delta_time_fs = dt_fs;
  ctl.temp = 0.0;
  unit_converter();
  init_mem();
  set_potential();
  identify_ion();
  set_loc();
  mk_table();
  sys.step = 1;
  refresh_self_term_cache();
  return sys.N;
}

void ref_set_positions(const double *rx, const double *ry, const double *rz)
{
  memcpy(sys.rx, rx, sizeof(double) * sys.N);
  memcpy(sys.ry, ry, sizeof(double) * sys.N);
  memcpy(sys.rz, rz, sizeof(double) * sys.N);
}

void ref_set_velocities(const double *vx, const double *vy, const double *vz)
{
  memcpy(sys.vx, vx, sizeof(double) * sys.N);
  memcpy(sys.vy, vy, sizeof(double) * sys.N);
  memcpy(sys.vz, vz, sizeof(double) * sys.N);
}

void ref_set_forces(const double *fx, const double *fy, const double *fz)
{
  memcpy(sys.fx, fx, sizeof(double) * sys.N);
  memcpy(sys.fy, fy, sizeof(double) * sys.N);
  memcpy(sys.fz, fz, sizeof(double) * sys.N);
}

/* out: 12 arrays of N doubles in the order r, v, f, f0 (x, y, z each) */
void ref_get_state(double *out)
{
  double *a[12];
  int k;
  a[0] = sys.rx;  a[1] = sys.ry;  a[2] = sys.rz;
  a[3] = sys.vx;  a[4] = sys.vy;  a[5] = sys.vz;
  a[6] = sys.fx;  a[7] = sys.fy;  a[8] = sys.fz;
  a[9] = sys.fx0; a[10] = sys.fy0; a[11] = sys.fz0;
  for (k = 0; k < 12; k++) memcpy(out + (size_t)k * sys.N, a[k], sizeof(double) * sys.N);
}

void ref_get_species(int *out)
{
  int i;
  for (i = 0; i < sys.N; i++) out[i] = sys.ion[i];
}

/* out: pot, virX, virY, virZ, kin, kinX, kinY, kinZ, presX, presY, presZ, pres */
void ref_get_scalars(double *out)
{
  out[0] = sys.pot;  out[1] = sys.virX; out[2] = sys.virY; out[3] = sys.virZ;
  out[4] = sys.kin;  out[5] = sys.kinX; out[6] = sys.kinY; out[7] = sys.kinZ;
  out[8] = sys.presX; out[9] = sys.presY; out[10] = sys.presZ; out[11] = sys.pres;
}

int ref_table_division(void) { return sys.table_division; }

/* the material's species data as its set_potential and mk_table made it */
int ref_kinds(void) { return ctl.kinds_of_ions; }
int ref_natoms_in_mol_unit(void) { return ctl.natoms_in_mol_unit; }
void ref_get_ions(double *Z, double *m, int *lookup, double *L)
{
  int k;
  for (k = 0; k < ctl.kinds_of_ions; k++) {
    Z[k] = ion.Z[k];
    m[k] = ion.m[k];
  }
  for (k = 0; k < ctl.kinds_of_ions * ctl.kinds_of_ions; k++) lookup[k] = sys.lookup[k];
  L[0] = sys.Lx; L[1] = sys.Ly; L[2] = sys.Lz;
}

void ref_get_table(double *pe, double *fe)
{
  size_t n = (size_t)3 * (size_t)(sys.table_division + 1);
  memcpy(pe, sys.pe1r1, sizeof(double) * n);
  memcpy(fe, sys.fe1r1, sizeof(double) * n);
}

double ref_erfcc(double x) { return erfcc(x); }

void ref_calc_foc(void) { calc_foc(); }
void ref_clear_foc(void) { clear_foc(); }
void ref_real_space(void) { real_space(); }
void ref_reciprocal_space(void) { reciprocal_space(); }
void ref_reciprocal_space3(void) { reciprocal_space3(); }
void ref_calc_kin(void) { calc_kin(); }
void ref_calc_press(void) { calc_press(); }
void ref_moment_correction(void) { moment_correction(); }

/* ---- calc_msd (msd.c:20-81) and control_temp (pt.c:85-112) ---- */
void ref_calc_msd(void) { calc_msd(); }
void ref_control_temp(int d_step, double temp) { control_temp(d_step, temp); }

/* out: rx_old, ry_old, rz_old, dx, dy, dz (N each); value, number_of_ion
 * (kinds_of_ions each) */
void ref_get_msd(double *arrays, double *value, int *number)
{
  double *a[6];
  int k;
  a[0] = msd.rx_old; a[1] = msd.ry_old; a[2] = msd.rz_old;
  a[3] = msd.dx;     a[4] = msd.dy;     a[5] = msd.dz;
  for (k = 0; k < 6; k++) memcpy(arrays + (size_t)k * sys.N, a[k], sizeof(double) * sys.N);
  for (k = 0; k < ctl.kinds_of_ions; k++) {
    value[k] = msd.value[k];
    number[k] = msd.number_of_ion[k];
  }
}

/* control_temp keeps its running sums in statics; a call at a multiple of
 * d_step restarts them.  Make one on a copy of the velocities and restore
 * them, so each test starts from mdc's initial (zero) statics. */
void ref_reset_control_temp(void)
{
  int n = sys.N, st = sys.step;
  double kin = sys.kin;
  double *v = (double *)malloc(sizeof(double) * 3 * n);
  memcpy(v, sys.vx, sizeof(double) * n);
  memcpy(v + n, sys.vy, sizeof(double) * n);
  memcpy(v + 2 * n, sys.vz, sizeof(double) * n);
  sys.step = 1;
  sys.kin = 1.0;
  control_temp(1, 1.0);
  memcpy(sys.vx, v, sizeof(double) * n);
  memcpy(sys.vy, v + n, sizeof(double) * n);
  memcpy(sys.vz, v + 2 * n, sizeof(double) * n);
  free(v);
  sys.kin = kin;
  sys.step = st;
}

/* mdc/program/main.c:26-63 without I/O: next_rv_verlet, calc_kin,
 * calc_press, calc_msd (msd_on), control_temp while step < t_until
 * (t_control_step > 0), moment_correction every moment_every steps */
void ref_newton(int first_step, int nsteps, int t_control_step, double temp, int t_until,
                int moment_every, int msd_on)
{
  int s;
  for (s = 0; s < nsteps; s++) {
    sys.step = first_step + s;
    next_rv_verlet();
    calc_kin();
    calc_press();
    if (msd_on) calc_msd();
    if (t_control_step > 0 && sys.step < t_until) control_temp(t_control_step, temp);
    if (moment_every > 0 && sys.step % moment_every == 0) moment_correction();
  }
}

/* ---- files.c's writers, pointed at `path` (mdc opens ../files/...) ---- */
static int to_file(FILE **fp, const char *path, void (*writer)(void))
{
  *fp = fopen(path, "w");
  if (!*fp) return -1;
  writer();
  fclose(*fp);
  *fp = NULL;
  return 0;
}
int ref_print_to_file(const char *path) { return to_file(&fpout, path, print_to_file); }
int ref_r_position(const char *path) { return to_file(&fpout_positions, path, r_position); }
int ref_r_velocities(const char *path) { return to_file(&fpout_velocities, path, r_velocities); }

/* steps of mdc/program/main.c:26-33: next_rv_verlet, calc_kin, calc_press */
void ref_run(int first_step, int nsteps)
{
  int s;
  for (s = 0; s < nsteps; s++) {
    sys.step = first_step + s;
    next_rv_verlet();
    calc_kin();
    calc_press();
  }
}
