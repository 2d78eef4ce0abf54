// host_setup.h -- host-side setup helpers of libmdgpu.
#pragma once

#include "mdgpu.h"

namespace mdg {

double host_erfcc(double x);
void host_mk_table(const mdg_potential& p, double alpha, double a3, int nb, double* pe,
                   double* fe);

}  // namespace mdg
