// mori_erfc.h -- Mori's erfc in FP64 on the device, shared by libmdgpu's
// analytic real space and libmdgpu_pr's per-pair Ewald term.
#pragma once

#include <hip/hip_runtime.h>

// Mori's erfc (step h = 1/2), the form mdc tabulates (mdc/program/ext.c:73-104):
//   erfc(x) ~ (x/pi) e^{-x^2} [1/(2x^2) + sum_{n=1..13} e^{-n^2/4}/(n^2/4 + x^2)]
//             - 2/(e^{4 pi x} - 1)
// Weights e^{-n^2/4} at the 16 digits mdc uses.
__device__ __forceinline__ double mori_erfc(double x) {
  const double w[13] = {
      0.7788007830714049,    0.3678794411714423,    0.1053992245618643,
      0.01831563888873417,   0.001930454136227709,  0.0001234098040866795,
      4.785117392129008e-6,  1.125351747192591e-7,  1.605228055185611e-9,
      1.388794386496401e-11, 7.287724095819693e-14, 2.319522830243569e-16,
      4.4777324417183e-19};
  double x2 = x * x;
  double s = 0.5 / x2;
#pragma unroll
  for (int n = 1; n <= 13; ++n) s += w[n - 1] / (0.25 * (double)(n * n) + x2);
  double v = -2.0 / (-1.0 + exp(12.56637061435917 * x)) + (0.3183098861837907 * x * s) / exp(x2);
  return x >= 0.0 ? v : 2.0 - v;
}
