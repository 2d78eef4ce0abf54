/* Prints the size and field offsets of pr-mdc's global structs, from
 * include/prmdc_abi.h or (-DUSE_PRMDC_HEADERS) from pr-mdc's own md.h;
 * tests/test_abi.py compares the two. */
#include <stddef.h>
#include <stdio.h>
#ifdef USE_PRMDC_HEADERS
#include "md.h"
#else
#include "prmdc_abi.h"
#endif

#define S(t) printf(#t " %zu\n", sizeof(t))
#define F(t, f) printf(#t "." #f " %zu\n", offsetof(t, f))

int main(void)
{
  S(calc_control);
  F(calc_control, skip_ewald); F(calc_control, delta_time_fs); F(calc_control, kinds_of_ions);
  F(calc_control, kinds_of_interaction);
  S(cell_parameters);
  F(cell_parameters, vol); F(cell_parameters, rx); F(cell_parameters, fx);
  F(cell_parameters, hm); F(cell_parameters, hmr); F(cell_parameters, g);
  F(cell_parameters, sigma); F(cell_parameters, Lx); F(cell_parameters, angle_ca);
  S(system_property);
  F(system_property, ion); F(system_property, N); F(system_property, step);
  F(system_property, pot); F(system_property, a1); F(system_property, hm);
  F(system_property, hm_sqrt); F(system_property, cutoff); F(system_property, cutoff2);
  F(system_property, rcsi); F(system_property, table_division); F(system_property, pe1r1);
  F(system_property, fe1r1); F(system_property, lookup);
  S(pressure_temperature_set);
  F(pressure_temperature_set, pres_trace); F(pressure_temperature_set, virial);
  F(pressure_temperature_set, kin_tensor);
  S(ion_property);
  F(ion_property, m); F(ion_property, Z);
  S(Parrinello_Rahman_set);
  F(Parrinello_Rahman_set, dt); F(Parrinello_Rahman_set, v); F(Parrinello_Rahman_set, f0);
  return 0;
}
