/*
 * dropin_main.c -- a fixed mdc driver, linked two ways by oracle/Makefile.
 *
 * TEST INFRASTRUCTURE ONLY.  Together with ref_harness.c and mdc's own
 * sources it makes two executables in oracle/_ref:
 *   mdc_cpu  every mdc routine from mdc/program (the reference's CPU path);
 *   mdc_gpu  the same program with real.c, rcprcl.c, rv.c, pt.c and msd.c
 *            left out and libmdgpu.so linked in their place, which is the
 *            link-time replacement a maintainer makes to use the GPU path.
 * The GPU test runs both on the same input and compares what they print.
 *
 * The step loop is mdc's (mdc/program/main.c:26-63) without file output:
 * next_rv_verlet, calc_kin, calc_press; then, when asked for, calc_msd,
 * control_temp(t_step, temp) while step < 10000 and moment_correction
 * every `moment` steps.
 *
 * usage: mdc_{cpu,gpu}[_material] n steps seed [t_step temp msd moment [a [vtemp]]]
 *        (n x n x n cells of the material that control.c sets up, lattice
 *        constant a (default 5.64 A), alpha = 5.6 / L, hm = 27, cutoff
 *        L / 2, dt = 2 fs; t_step 0 = no thermostat, msd 0/1, moment 0 =
 *        no moment correction; vtemp > 0: Maxwell-Boltzmann velocities at
 *        vtemp K from the seed (Box-Muller on the seeded generator) in place
 *        of the small uniform ones)
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "md.h"

/* mdc's prototypes.h declares main(void); these are the routines used */
void next_rv_verlet(void);
void calc_kin(void);
void calc_press(void);
void moment_correction(void);
void calc_msd(void);
void control_temp(int d_step, double temp);
int ref_setup(int nx, int ny, int nz, double a, double alpha, int hm, double radius,
              double dt_fs);

static unsigned long long lcg_state;

static double uniform(void) /* in [0, 1) */
{
  lcg_state = lcg_state * 6364136223846793005ull + 1442695040888963407ull;
  return (double)(lcg_state >> 11) * (1.0 / 9007199254740992.0);
}

static double gauss(void) /* standard normal, Box-Muller */
{
  double u1 = uniform(), u2 = uniform();
  if (u1 < 1e-300) u1 = 1e-300;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

int main(int argc, char **argv)
{
  int n, steps, i, t_step = 0, msd_on = 0, moment = 0, k;
  double temp = 0.0, vtemp = 0.0;
  double a = 5.64, L, sx = 0, sy = 0, sz = 0;

  if (argc != 4 && argc != 8 && argc != 9 && argc != 10) {
    fprintf(stderr, "usage: %s n steps seed [t_step temp msd moment [a [vtemp]]]\n", argv[0]);
    return 2;
  }
  if (argc >= 9) a = atof(argv[8]);
  if (argc == 10) vtemp = atof(argv[9]);
  if (argc >= 8) {
    t_step = atoi(argv[4]);
    temp = atof(argv[5]);
    msd_on = atoi(argv[6]);
    moment = atoi(argv[7]);
  }
  n = atoi(argv[1]);
  steps = atoi(argv[2]);
  lcg_state = (unsigned long long)atoll(argv[3]);
  L = n * a;
  ref_setup(n, n, n, a, 5.6 / L, 27, L / 2.0, 2.0);

  /* lattice jitter of +-0.1 A and small random velocities */
  for (i = 0; i < sys.N; i++) {
    sys.rx[i] += 0.2 * (uniform() - 0.5);
    sys.ry[i] += 0.2 * (uniform() - 0.5);
    sys.rz[i] += 0.2 * (uniform() - 0.5);
    if (sys.rx[i] < 0.0) sys.rx[i] += L;
    if (sys.ry[i] < 0.0) sys.ry[i] += L;
    if (sys.rz[i] < 0.0) sys.rz[i] += L;
    if (vtemp > 0.0) { /* sqrt(kB T / m) per component, kB as init.c:66 */
      double s = sqrt(1.380658e-5 * vtemp / ion.m[sys.ion[i]]);
      sys.vx[i] = s * gauss();
      sys.vy[i] = s * gauss();
      sys.vz[i] = s * gauss();
    } else {
      sys.vx[i] = 0.05 * (uniform() - 0.5);
      sys.vy[i] = 0.05 * (uniform() - 0.5);
      sys.vz[i] = 0.05 * (uniform() - 0.5);
    }
  }
  moment_correction();

  for (sys.step = 1; sys.step <= steps; sys.step++) {
    next_rv_verlet();
    calc_kin();
    calc_press();
    printf("%d %.17g %.17g %.17g %.17g %.17g %.17g\n", sys.step, sys.pot, sys.kin,
           sys.virX, sys.virY, sys.virZ, sys.pres);
    if (msd_on) {
      calc_msd();
      printf("msd %d", sys.step);
      for (k = 0; k < ctl.kinds_of_ions; k++) printf(" %.17g", msd.value[k]);
      printf("\n");
    }
    if (t_step > 0 && sys.step < 10000) control_temp(t_step, temp);
    if (moment > 0 && sys.step % moment == 0) moment_correction();
  }
  for (i = 0; i < sys.N; i++) {
    sx += sys.rx[i];
    sy += sys.ry[i];
    sz += sys.rz[i];
  }
  printf("final %.17g %.17g %.17g %.17g %.17g\n", sx, sy, sz, sys.rx[0], sys.vx[sys.N - 1]);
  return 0;
}
