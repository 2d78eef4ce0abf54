/*
 * pr_dropin_main.c -- a fixed pr-mdc driver, linked two ways by
 * oracle/Makefile.  TEST INFRASTRUCTURE ONLY.
 *   prmdc_cpu  every routine from pr-mdc/program (the reference's CPU path);
 *   prmdc_gpu  the same program with real.c and ewald.c left out and
 *              libmdgpu_pr.so linked in their place, the link-time
 *              replacement a maintainer makes to use the GPU routines.
 * The GPU test runs both on the same input and compares what they print.
 *
 * The step loop is pr-mdc's (pr-mdc/program/main.c:27-127) without file
 * output: next_rv_verlet, control_press_PR, control_temp, calc_kin,
 * calc_press, printing per step the potential, kinetic energy, pressure,
 * volume and cell matrix, and at the end the sums of the positions and
 * velocities.
 *
 * usage: prmdc_{cpu,gpu} n steps seed s01 s02 s12
 *        (n x n x n NaCl cells of 5.63 A sheared by s01, s02, s12; see
 *        prref_setup in pr_ref_harness.c)
 */
#include <stdio.h>
#include <stdlib.h>

#include "md.h"

/* pr-mdc's prototypes.h declares main(void); these are the routines used */
void next_rv_verlet(void);
void control_press_PR(void);
void control_temp(int d_step, double temp);
void calc_kin(void);
void calc_press(void);
int prref_setup(int n, double a, double s01, double s02, double s12, unsigned long long seed,
                double jitter, double vmax);

int main(int argc, char **argv)
{
  int n, steps, i, j;
  if (argc != 7) {
    fprintf(stderr, "usage: %s n steps seed s01 s02 s12\n", argv[0]);
    return 2;
  }
  n = atoi(argv[1]);
  steps = atoi(argv[2]);
  prref_setup(n, 5.63, atof(argv[4]), atof(argv[5]), atof(argv[6]),
              (unsigned long long)atoll(argv[3]), 0.2, 0.05);
  for (sys.step = 1; sys.step <= steps; sys.step++) {
    next_rv_verlet();
    control_press_PR();
    control_temp(pt.t_control_step, pt.temp_set);
    calc_kin();
    calc_press();
    printf("%d %.17g %.17g %.17g %.17g", sys.step, sys.pot, sys.kin, pt.pres_trace, cell.vol);
    for (i = 0; i < 3; i++)
      for (j = 0; j < 3; j++) printf(" %.17g", cell.hm[i][j]);
    printf("\n");
  }
  {
    double sr[3] = {0, 0, 0}, sv[3] = {0, 0, 0};
    for (i = 0; i < sys.N; i++) {
      sr[0] += cell.rx[i]; sr[1] += cell.ry[i]; sr[2] += cell.rz[i];
      sv[0] += cell.vx[i]; sv[1] += cell.vy[i]; sv[2] += cell.vz[i];
    }
    printf("final %.17g %.17g %.17g %.17g %.17g %.17g\n", sr[0], sr[1], sr[2], sv[0], sv[1], sv[2]);
  }
  return 0;
}
