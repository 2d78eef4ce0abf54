/*
 * pr_ref_harness.c -- TEST INFRASTRUCTURE ONLY.  Compiled together with
 * pr-mdc's own sources (pr-mdc/program, all but main.c) into
 * oracle/_ref/libprmdcref.so and into the pr-mdc driver programs; never part
 * of the product.
 *
 * prref_setup builds a reproducible pr-mdc system with pr-mdc's own
 * initialisation (pr-mdc/program/init.c:10-29 without set_vel, whose seed
 * is the clock): n x n x n rocksalt cells of edge a, the cell matrix
 * sheared into a triclinic one (b += s01 a, c += s02 a + s12 b, ions kept
 * at their scaled coordinates), positions jittered by up to +-jitter/2 and
 * velocities drawn in +-vmax/2 from a seeded LCG, then mk_table,
 * calc_sigma and set_cutoff as pr-mdc's calc_foc runs them.
 */
#include <stdio.h>
#include <stdlib.h>

#include "md.h"
#include "prototypes.h"

static unsigned long long lcg_state;

static double uniform(void) /* in [0, 1) */
{
  lcg_state = lcg_state * 6364136223846793005ull + 1442695040888963407ull;
  return (double)(lcg_state >> 11) * (1.0 / 9007199254740992.0);
}

int prref_setup(int n, double a, double s01, double s02, double s12, unsigned long long seed,
                double jitter, double vmax)
{
  int i;
  double L;

  init_mem1();
  get_control_param();
  cell.nx = cell.ny = cell.nz = n;
  cell.hu[0][0] = cell.hu[1][1] = cell.hu[2][2] = a;
  unit_converter();
  init_mem2();
  set_potential();
  identify_ion();
  set_loc();
  evaluate_PR_param();

  L = cell.hm[0][0];
  for (i = 0; i < sys.N; i++) {
    cell.sx[i] = cell.rx[i] / L;
    cell.sy[i] = cell.ry[i] / L;
    cell.sz[i] = cell.rz[i] / L;
  }
  cell.hm[1][0] += s01 * cell.hm[0][0];
  cell.hm[2][0] += s02 * cell.hm[0][0];
  cell.hm[2][1] += s12 * cell.hm[1][1];
  calc_cell();
  lcg_state = seed;
  for (i = 0; i < sys.N; i++) {
    s2r(i);
    cell.rx[i] += jitter * (uniform() - 0.5);
    cell.ry[i] += jitter * (uniform() - 0.5);
    cell.rz[i] += jitter * (uniform() - 0.5);
    cyclic_bc(i);
    cell.vx[i] = vmax * (uniform() - 0.5);
    cell.vy[i] = vmax * (uniform() - 0.5);
    cell.vz[i] = vmax * (uniform() - 0.5);
  }
  mk_table();
  calc_sigma();
  set_cutoff();
  return sys.N;
}
