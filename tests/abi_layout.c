/* Prints the size and field offsets of mdc's global structs.  Built twice by
 * tests/test_abi.py: against include/mdc_abi.h, and (where the reference
 * checkout exists) against mdc's own md.h with -DUSE_MDC_HEADERS; the two
 * outputs must be identical for the drop-in to read mdc's globals. */
#include <stddef.h>
#include <stdio.h>
#ifdef USE_MDC_HEADERS
#include "md.h"
#else
#include "mdc_abi.h"
#endif

#define F(T, f) printf(#T "." #f " %zu\n", offsetof(T, f))

int main(void)
{
  printf("system_property %zu\n", sizeof(system_property));
  F(system_property, rx); F(system_property, ry); F(system_property, rz);
  F(system_property, vx); F(system_property, vy); F(system_property, vz);
  F(system_property, fx0); F(system_property, fy0); F(system_property, fz0);
  F(system_property, fx); F(system_property, fy); F(system_property, fz);
  F(system_property, ion); F(system_property, N); F(system_property, nx);
  F(system_property, ny); F(system_property, nz); F(system_property, Ax);
  F(system_property, Ay); F(system_property, Az); F(system_property, Lx);
  F(system_property, Ly); F(system_property, Lz); F(system_property, kB);
  F(system_property, kk); F(system_property, perMol); F(system_property, e2t);
  F(system_property, e2tX); F(system_property, e2tY); F(system_property, e2tZ);
  F(system_property, pp2gpa); F(system_property, step); F(system_property, dt);
  F(system_property, dt2); F(system_property, a1); F(system_property, p2a2);
  F(system_property, a3); F(system_property, hm); F(system_property, hm_sqrt);
  F(system_property, radius); F(system_property, radius2); F(system_property, rcsi);
  F(system_property, rsni); F(system_property, table_division);
  F(system_property, pe1r1); F(system_property, fe1r1); F(system_property, lookup);
  F(system_property, virX); F(system_property, virY); F(system_property, virZ);
  F(system_property, presX); F(system_property, presY); F(system_property, presZ);
  F(system_property, pres); F(system_property, pres_scalling_factor);
  F(system_property, pot); F(system_property, kin); F(system_property, kinX);
  F(system_property, kinY); F(system_property, kinZ);
  printf("calc_control %zu\n", sizeof(calc_control));
  F(calc_control, delta_time_fs); F(calc_control, calc_max); F(calc_control, temp);
  F(calc_control, t_control_step); F(calc_control, p_control_step);
  F(calc_control, set_press_GPa_X); F(calc_control, set_press_GPa_Y);
  F(calc_control, set_press_GPa_Z); F(calc_control, press_X); F(calc_control, press_Y);
  F(calc_control, press_Z); F(calc_control, kinds_of_ions);
  F(calc_control, natoms_in_mol_unit); F(calc_control, natoms_in_unit_cell);
  F(calc_control, kinds_of_interaction);
  printf("ion_property %zu\n", sizeof(ion_property));
  F(ion_property, m); F(ion_property, Z);
  printf("huggins_mayer_potential_set %zu\n", sizeof(huggins_mayer_potential_set));
  F(huggins_mayer_potential_set, b); F(huggins_mayer_potential_set, c_na_na);
  F(huggins_mayer_potential_set, c_na_cl); F(huggins_mayer_potential_set, c_cl_cl);
  F(huggins_mayer_potential_set, sigma_na); F(huggins_mayer_potential_set, sigma_cl);
  F(huggins_mayer_potential_set, rho); F(huggins_mayer_potential_set, C_na_na);
  F(huggins_mayer_potential_set, C_na_cl); F(huggins_mayer_potential_set, C_cl_cl);
  F(huggins_mayer_potential_set, D_na_na); F(huggins_mayer_potential_set, D_na_cl);
  F(huggins_mayer_potential_set, D_cl_cl);
  printf("msd_calculation_variables %zu\n", sizeof(msd_calculation_variables));
  F(msd_calculation_variables, number_of_ion); F(msd_calculation_variables, rx_old);
  F(msd_calculation_variables, ry_old); F(msd_calculation_variables, rz_old);
  F(msd_calculation_variables, dx); F(msd_calculation_variables, dy);
  F(msd_calculation_variables, dz); F(msd_calculation_variables, value);
  F(msd_calculation_variables, Lx2); F(msd_calculation_variables, Ly2);
  F(msd_calculation_variables, Lz2);
  return 0;
}
