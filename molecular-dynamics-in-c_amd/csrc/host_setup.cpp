// host_setup.cpp -- host-side setup of libmdgpu: salt parameters, mdc's
// erfc + BMH pair table, rock-salt lattice, seeded Maxwell velocities and
// the Ewald split helper.  Setup only (runs once per box); every per-step
// computation runs on the GPU.
#include <math.h>

#include <algorithm>
#include <string.h>

#include "host_setup.h"
#include "mdgpu.h"

namespace mdg {

// mdc's erfc (mdc/program/ext.c:73-104): Mori's formula with step 1/2.
double host_erfcc(double x) {
  static const double w[13] = {
      0.7788007830714049,    0.3678794411714423,    0.1053992245618643,
      0.01831563888873417,   0.001930454136227709,  0.0001234098040866795,
      4.785117392129008e-6,  1.125351747192591e-7,  1.605228055185611e-9,
      1.388794386496401e-11, 7.287724095819693e-14, 2.319522830243569e-16,
      4.4777324417183e-19};
  double x2 = x * x;
  double s = 0.5 / x2;
  for (int n = 1; n <= 13; ++n) s += w[n - 1] / (0.25 * (double)(n * n) + x2);
  double v = -2.0 / (-1.0 + exp(12.56637061435917 * x)) + (0.3183098861837907 * x * s) / exp(x2);
  return x >= 0.0 ? v : 2.0 - v;
}

// mdc's mk_table (mdc/program/control.c:183-248) for a two-species salt:
// bins d = 0..nb-1 at r = d/1000 + 0.5 A, pair types in mdc's order
// (0,0), (0,1), (1,1).  pe[t*nb + d], fe[t*nb + d].
void host_mk_table(const mdg_potential& p, double alpha, double a3, int nb, double* pe,
                   double* fe) {
  int t = 0;
  for (int ti = 0; ti < 2; ++ti)
    for (int tj = ti; tj < 2; ++tj, ++t) {
      const double ZZ = p.Z[ti] * p.Z[tj];
      for (int d = 0; d < nb; ++d) {
        const double dr = ((double)d) / 1000.0 + 0.5;
        const double dr6 = dr * dr * dr * dr * dr * dr;
        const double dr8 = dr6 * dr * dr;
        const double adr = alpha * dr;
        const double erfc_per_dr = host_erfcc(adr) / dr;
        const double p_e1 = ZZ * erfc_per_dr;
        const double p_r1 = p.c[t] * p.b * exp((p.sigma[t] - dr) / p.rho) - p.C[t] / dr6 -
                            p.D[t] / dr8;
        const double f_e1 = ZZ * (2.0 * a3 * exp(-adr * adr) + erfc_per_dr) / (dr * dr);
        const double f_r1 = (p.c[t] * p.b * exp((p.sigma[t] - dr) / p.rho) / (p.rho) -
                             6.0 * p.C[t] / (dr6 * dr) - 8.0 * p.D[t] / (dr8 * dr)) /
                            dr;
        pe[t * nb + d] = p_e1 + p_r1;
        fe[t * nb + d] = f_e1 + f_r1;
      }
    }
}

// splitmix64 (S. Vigna) -> uniform doubles in (0, 1]
static inline unsigned long long splitmix64(unsigned long long& s) {
  unsigned long long z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
static inline double uniform01(unsigned long long& s) {
  return ((double)(splitmix64(s) >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}

}  // namespace mdg

extern "C" {

int mdg_potential_nacl(mdg_potential* p) {
  if (!p) return MDG_E_ARG;
  const double kk = 2.307079556, mu = 6.0221367 * 1.6605402 / 10.0;
  memset(p, 0, sizeof(*p));
  // mdc/program/control.c:85-181 (Tosi-Fumi NaCl)
  p->Z[0] = 1.0 * sqrt(kk);
  p->Z[1] = -1.0 * sqrt(kk);
  p->m[0] = 22.9898 / mu;
  p->m[1] = 35.4528 / mu;
  p->b = 3.38 / 100.0;
  p->rho = 0.317;
  p->c[0] = 1.25; p->c[1] = 1.0; p->c[2] = 0.75;
  p->sigma[0] = 1.170 + 1.170; p->sigma[1] = 1.170 + 1.585; p->sigma[2] = 1.585 + 1.585;
  p->C[0] = 1.68 / 10.0; p->C[1] = 11.2 / 10.0; p->C[2] = 116.0 / 10.0;
  p->D[0] = 0.8 / 10.0; p->D[1] = 13.9 / 10.0; p->D[2] = 233.0 / 10.0;
  return MDG_OK;
}

int mdg_potential_kcl(mdg_potential* p) {
  if (!p) return MDG_E_ARG;
  const double kk = 2.307079556, mu = 6.0221367 * 1.6605402 / 10.0;
  memset(p, 0, sizeof(*p));
  // Tosi & Fumi, J. Phys. Chem. Solids 25, 45 (1964), BMH form of
  // mdc/program/control.c:85-181: sigma_K 1.463 A, sigma_Cl 1.585 A,
  // rho 0.337 A, C 24.3/48/124.5 (1e-60 erg cm^6), D 24/73/250 (1e-76 erg cm^8).
  p->Z[0] = 1.0 * sqrt(kk);
  p->Z[1] = -1.0 * sqrt(kk);
  p->m[0] = 39.0983 / mu;
  p->m[1] = 35.4528 / mu;
  p->b = 3.38 / 100.0;
  p->rho = 0.337;
  p->c[0] = 1.25; p->c[1] = 1.0; p->c[2] = 0.75;
  p->sigma[0] = 1.463 + 1.463; p->sigma[1] = 1.463 + 1.585; p->sigma[2] = 1.585 + 1.585;
  p->C[0] = 24.3 / 10.0; p->C[1] = 48.0 / 10.0; p->C[2] = 124.5 / 10.0;
  p->D[0] = 24.0 / 10.0; p->D[1] = 73.0 / 10.0; p->D[2] = 250.0 / 10.0;
  return MDG_OK;
}

// Positions and species in the order of mdc's identify_ion / set_loc
// (mdc/program/control.c:46-82).
int mdg_rocksalt(int nx, int ny, int nz, double ax, double ay, double az, double* r,
                 int* species) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || !r) return MDG_E_ARG;
  const double Lx = ax * nx, Ly = ay * ny, Lz = az * nz;
  const int gx = 2 * nx, gy = 2 * ny, gz = 2 * nz;
  const long n = (long)gx * gy * gz;
  const double dgx = gx, dgy = gy, dgz = gz;
  const double sx = (Lx / dgx) / 2.0, sy = (Ly / dgy) / 2.0, sz = (Lz / dgz) / 2.0;
  long i = 0;
  for (int z = 0; z < gz; ++z)
    for (int y = 0; y < gy; ++y)
      for (int x = 0; x < gx; ++x, ++i) {
        if (species) species[i] = (x % 2 + y % 2 + z % 2) % 2;
        r[i] = Lx * ((double)x / dgx) + sx;
        r[n + i] = Ly * ((double)y / dgy) + sy;
        r[2 * n + i] = Lz * ((double)z / dgz) + sz;
      }
  return MDG_OK;
}

// Maxwell-Boltzmann components N(0, kB T / m) from splitmix64 + Box-Muller,
// then mdc's moment_correction (mdc/program/rv.c:88-116).
int mdg_maxwell(unsigned long long seed, double temp_k, int n, const int* species,
                const mdg_potential* p, double* v) {
  if (n <= 0 || !species || !p || !v || temp_k < 0) return MDG_E_ARG;
  const double kB = 1.380658e-5;
  unsigned long long s = seed ^ 0x6A09E667F3BCC909ULL;
  for (int d = 0; d < 3; ++d)
    for (int i = 0; i < n; i += 2) {
      const double u1 = mdg::uniform01(s), u2 = mdg::uniform01(s);
      const double rad = sqrt(-2.0 * log(u1)), th = 2.0 * M_PI * u2;
      const double g[2] = {rad * cos(th), rad * sin(th)};
      for (int q = 0; q < 2 && i + q < n; ++q) {
        const double m = p->m[species[i + q]];
        v[(long)d * n + i + q] = g[q] * sqrt(kB * temp_k / m);
      }
    }
  for (int d = 0; d < 3; ++d) {
    double sum = 0.0;
    for (int i = 0; i < n; ++i) sum += p->m[species[i]] * v[(long)d * n + i];
    const double mean = sum / (double)n;
    for (int i = 0; i < n; ++i) v[(long)d * n + i] -= mean / p->m[species[i]];
  }
  return MDG_OK;
}

// Ewald split for a relative error target eps: s = sqrt(-ln eps), alpha = s / rc,
// and the integer reciprocal cutoff covering |m| <= s alpha / pi on the
// longest axis.
int mdg_ewald_params(double rc, double l_max, double eps, double* alpha, int* hm) {
  if (!(rc > 0) || !(l_max > 0) || !(eps > 0 && eps < 1) || !alpha || !hm) return MDG_E_ARG;
  const double s = sqrt(-log(eps));
  *alpha = s / rc;
  const double hmax = s * (*alpha) * l_max / M_PI;
  *hm = (int)floor(hmax * hmax);
  return MDG_OK;
}

int mdg_ewald_rc_rule(int n_ions, double l_min, double rc_cap, double* rc) {
  if (n_ions <= 0 || !(l_min > 0) || !(rc_cap > 0) || !rc) return MDG_E_ARG;
  // the real-space cutoff that balances the pair sum against the k sum at
  // a fixed Ewald accuracy grows as N^(1/6) at fixed density; 16.5 A at
  // 32768 ions is where rock-salt NaCl boxes of 4096-32768 ions, run in
  // chip-filling batches on MI355X, step fastest
  double r = 16.5 * pow((double)n_ions / 32768.0, 1.0 / 6.0);
  r = std::max(r, 10.0);
  r = std::min(r, rc_cap);
  r = std::min(r, 0.5 * l_min);
  *rc = r;
  return MDG_OK;
}

}  // extern "C"
