/*
 * mdc_oracle.c -- CPU restatement of mdc's force path (parity checker).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
 * the CPU-baseline leg of bench.py, never by the product library.
 *
 * Every function keeps the floating-point operation order of the mdc
 * routine it restates (cited per function, paths relative to the
 * reference repository), so that built with the reference's flags
 * (-O2 -ffp-contract=off, SSE2 doubles) it reproduces mdc bit for bit on
 * the same inputs.  oracle/_ref (the reference sources compiled with a
 * small harness) checks exactly that.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "mdc_oracle.h"

/* Constants as mdc spells them (mdc/program/md.h:8-16).  The decimal
 * values are kept as written there: PI2 is one ulp above the double
 * nearest 2*pi, and parity at the last bit needs the same value. */
#define ORC_PI        3.1415926535897932385
#define ORC_PI2       6.283185307179587
#define ORC_PIPI      9.869604401089358619
#define ORC_SQRTPI    1.77245385090551602731
#define ORC_SECD_2_FS 4.074972637944947

/* Mori's double-exponential-type erfc with step h = 1/2:
 *   erfc(x) ~ (2xh/pi) e^{-x^2} [1/(2x^2) + sum_{n=1..13} e^{-n^2 h^2}/(n^2 h^2 + x^2)]
 *             - 2/(e^{2 pi x/h} - 1)
 * mdc/program/ext.c:73-104 (derivation in mdc/manual/manual.IEMD:727-793).
 * The weights e^{-n^2/4} are kept at the 16 digits mdc uses. */
static const double mori_w[13] = {
  0.7788007830714049,   0.3678794411714423,    0.1053992245618643,
  0.01831563888873417,  0.001930454136227709,  0.0001234098040866795,
  4.785117392129008e-6, 1.125351747192591e-7,  1.605228055185611e-9,
  1.388794386496401e-11, 7.287724095819693e-14, 2.319522830243569e-16,
  4.4777324417183e-19 };

double orc_erfcc(double x)
{
  const double four_pi = 12.56637061435917;   /* 4*pi as in ext.c:75 */
  const double inv_pi  = 0.3183098861837907;  /* 1/pi                */
  double x2 = x * x;
  double s = 0.5 / x2;
  int n;
  for (n = 1; n <= 13; n++) {
    double nh2 = 0.25 * (double)(n * n);      /* exact: n^2/4 */
    s += mori_w[n - 1] / (nh2 + x2);
  }
  {
    double v = -2.0 / (-1.0 + exp(four_pi * x)) + (inv_pi * x * s) / exp(x2);
    return x >= 0.0 ? v : 2.0 - v;
  }
}

/* Tosi-Fumi NaCl, mdc/program/control.c:85-181: energy unit 1e-18 J,
 * length A, mass in m_u (M / (N_A * u / 10)). */
void orc_potential_nacl(double kk, orc_potential *pp)
{
  double mu = 6.0221367 * 1.6605402 / 10.0;
  memset(pp, 0, sizeof(*pp));
  pp->Z[0] = 1.0 * sqrt(kk);
  pp->Z[1] = -1.0 * sqrt(kk);
  pp->m[0] = 22.9898 / mu;
  pp->m[1] = 35.4528 / mu;
  pp->b = 3.38 / 100.0;
  pp->rho = 0.317;
  pp->c[0] = 1.25;  pp->c[1] = 1.0;  pp->c[2] = 0.75;
  pp->sigma[0] = 1.170 + 1.170;
  pp->sigma[1] = 1.170 + 1.585;
  pp->sigma[2] = 1.585 + 1.585;
  pp->C[0] = 1.68 / 10.0;  pp->C[1] = 11.2 / 10.0;  pp->C[2] = 116.0 / 10.0;
  pp->D[0] = 0.8 / 10.0;   pp->D[1] = 13.9 / 10.0;  pp->D[2] = 233.0 / 10.0;
}

/* Tosi-Fumi KCl (not present in mdc).  Same BMH form and unit handling as
 * control.c:85-181; parameters from M. P. Tosi and F. G. Fumi,
 * J. Phys. Chem. Solids 25, 45 (1964) with the Mayer dispersion
 * constants: sigma_K = 1.463 A, sigma_Cl = 1.585 A, rho = 0.337 A,
 * C = 24.3 / 48 / 124.5 (1e-60 erg cm^6), D = 24 / 73 / 250
 * (1e-76 erg cm^8).  The same values are used by the product library. */
void orc_potential_kcl(double kk, orc_potential *pp)
{
  double mu = 6.0221367 * 1.6605402 / 10.0;
  memset(pp, 0, sizeof(*pp));
  pp->Z[0] = 1.0 * sqrt(kk);
  pp->Z[1] = -1.0 * sqrt(kk);
  pp->m[0] = 39.0983 / mu;
  pp->m[1] = 35.4528 / mu;
  pp->b = 3.38 / 100.0;
  pp->rho = 0.337;
  pp->c[0] = 1.25;  pp->c[1] = 1.0;  pp->c[2] = 0.75;
  pp->sigma[0] = 1.463 + 1.463;
  pp->sigma[1] = 1.463 + 1.585;
  pp->sigma[2] = 1.585 + 1.585;
  pp->C[0] = 24.3 / 10.0;  pp->C[1] = 48.0 / 10.0;  pp->C[2] = 124.5 / 10.0;
  pp->D[0] = 24.0 / 10.0;  pp->D[1] = 73.0 / 10.0;  pp->D[2] = 250.0 / 10.0;
}

/* mdc/program/init.c:25-80 (the parts the force path reads). */
void orc_unit_converter(orc_box *b)
{
  b->Lx = b->Ax * (double)b->nx;
  b->Ly = b->Ay * (double)b->ny;
  b->Lz = b->Az * (double)b->nz;
  if (b->a1 > 0) {
    b->p2a2 = ORC_PIPI / (b->a1 * b->a1);
    b->a3 = b->a1 / ORC_SQRTPI;
  }
  b->radius2 = b->radius * b->radius;
  b->hm_sqrt = (int)sqrtf((float)b->hm);
  b->N = 8 * b->nx * b->ny * b->nz;
  b->dt = b->delta_time_fs / ORC_SECD_2_FS;
  b->dt2 = b->dt * b->dt;
  b->table_division = (int)((b->radius - 0.5) * 1000.0) + 1;
  b->kB = 1.380658e-5;
  b->kk = 2.307079556;
  b->e2t = 2.0 / (3.0 * ((double)b->N * b->kB));
  b->lookup[0] = 0; b->lookup[1] = 1; b->lookup[2] = 1; b->lookup[3] = 2;
}

/* identify_ion + set_loc, mdc/program/control.c:46-82: a 2nx x 2ny x 2nz
 * simple-cubic grid, x fastest, species (x+y+z) mod 2, offset by half a
 * grid spacing. */
void orc_set_loc(orc_box *b)
{
  int gx = 2 * b->nx, gy = 2 * b->ny, gz = 2 * b->nz;
  double dgx = (double)gx, dgy = (double)gy, dgz = (double)gz;
  double sx = (b->Lx / dgx) / 2.0, sy = (b->Ly / dgy) / 2.0, sz = (b->Lz / dgz) / 2.0;
  int x, y, z, i = 0;
  for (z = 0; z < gz; z++)
    for (y = 0; y < gy; y++)
      for (x = 0; x < gx; x++) {
        b->ion[i] = (x % 2 + y % 2 + z % 2) % 2;
        b->rx[i] = b->Lx * ((double)x / dgx) + sx;
        b->ry[i] = b->Ly * ((double)y / dgy) + sy;
        b->rz[i] = b->Lz * ((double)z / dgz) + sz;
        i++;
      }
}

/* Pair potential and -U'(r)/r of erfc-screened Coulomb + BMH at distance
 * r, in the operation order of mdc/program/control.c:219-240. */
static void pair_terms(const orc_box *b, int t, double ZZ, double dr,
                       double *p, double *f)
{
  const orc_potential *pp = &b->pp;
  double dr6 = dr * dr * dr * dr * dr * dr;
  double dr8 = dr6 * dr * dr;
  double adr = b->a1 * dr;
  double erfc_adr_per_dr = orc_erfcc(adr) / dr;
  double p_e1 = ZZ * erfc_adr_per_dr;
  double p_r1 = pp->c[t] * pp->b * exp((pp->sigma[t] - dr) / pp->rho)
                - pp->C[t] / dr6 - pp->D[t] / dr8;
  double f_e1 = ZZ * (2.0 * b->a3 * exp(-adr * adr) + erfc_adr_per_dr) / (dr * dr);
  double f_r1 = (pp->c[t] * pp->b * exp((pp->sigma[t] - dr) / pp->rho) / (pp->rho)
                 - 6.0 * pp->C[t] / (dr6 * dr) - 8.0 * pp->D[t] / (dr8 * dr)) / dr;
  *p = p_e1 + p_r1;
  *f = f_e1 + f_r1;
}

/* mdc/program/control.c:183-248: bins of 0.001 A from 0.5 A. */
void orc_mk_table(orc_box *b)
{
  int ti, tj, d, t = 0;
  int nb = b->table_division + 1;
  for (ti = 0; ti < 2; ti++)
    for (tj = ti; tj < 2; tj++) {
      double ZZ = b->pp.Z[ti] * b->pp.Z[tj];
      for (d = 0; d < nb; d++) {
        double dr = ((double)d) / 1000.0 + 0.5;
        pair_terms(b, t, ZZ, dr, &b->pe1r1[t * nb + d], &b->fe1r1[t * nb + d]);
      }
      b->lookup[2 * ti + tj] = t;
      b->lookup[2 * tj + ti] = t;
      t++;
    }
}

/* mdc/program/init.c:179-190 */
void orc_clear_foc(orc_box *b)
{
  int i;
  for (i = 0; i < b->N; i++) {
    b->fx0[i] = b->fx[i]; b->fy0[i] = b->fy[i]; b->fz0[i] = b->fz[i];
    b->fx[i] = 0.0;       b->fy[i] = 0.0;       b->fz[i] = 0.0;
  }
}

/* mdc/program/real.c:15-84: all pairs i<j, table lookup with a
 * truncating 0.001 A bin; resets pot and the diagonal virial. */
void orc_real_space(orc_box *b)
{
  int i, j, nb = b->table_division + 1;
  double hx = b->Lx * 0.5, hy = b->Ly * 0.5, hz = b->Lz * 0.5;
  b->pot = 0.0;
  b->virX = b->virY = b->virZ = 0.0;
  for (i = 0; i < b->N; i++) {
    double xi = b->rx[i], yi = b->ry[i], zi = b->rz[i];
    for (j = i + 1; j < b->N; j++) {
      double dx = xi - b->rx[j], dy = yi - b->ry[j], dz = zi - b->rz[j];
      double dr, p, f;
      int t, bin;
      /* mdc folds x, y, z with '>' first, then all with '<' */
      if (dx > hx) dx -= b->Lx;
      if (dy > hy) dy -= b->Ly;
      if (dz > hz) dz -= b->Lz;
      if (dx < -hx) dx += b->Lx;
      if (dy < -hy) dy += b->Ly;
      if (dz < -hz) dz += b->Lz;
      dr = dx * dx + dy * dy + dz * dz;
      if (dr > b->radius2) continue;
      t = b->lookup[b->ion[i] * 2 + b->ion[j]];
      bin = (int)((sqrt(dr) - 0.5) * 1000.0);
      p = b->pe1r1[t * nb + bin];
      f = b->fe1r1[t * nb + bin];
      b->pot += p;
      b->fx[i] += f * dx; b->fy[i] += f * dy; b->fz[i] += f * dz;
      b->fx[j] -= f * dx; b->fy[j] -= f * dy; b->fz[j] -= f * dz;
      b->virX += f * dx * dx;
      b->virY += f * dy * dy;
      b->virZ += f * dz * dz;
    }
  }
  b->pot_real = b->pot;
}

/* Analytic real space: the pair loop of mdc/program/real.c:15-84 with
 * the table replaced by the exact-r evaluation that pr-mdc's ewald1 uses
 * for the erfc term (pr-mdc/program/ewald.c:11-68) and mk_table's BMH
 * formula for the short-range term. */
void orc_real_space_analytic(orc_box *b)
{
  int i, j;
  double hx = b->Lx * 0.5, hy = b->Ly * 0.5, hz = b->Lz * 0.5;
  b->pot = 0.0;
  b->virX = b->virY = b->virZ = 0.0;
  for (i = 0; i < b->N; i++) {
    double xi = b->rx[i], yi = b->ry[i], zi = b->rz[i];
    for (j = i + 1; j < b->N; j++) {
      double dx = xi - b->rx[j], dy = yi - b->ry[j], dz = zi - b->rz[j];
      double dr, p, f;
      int t;
      if (dx > hx) dx -= b->Lx;
      if (dy > hy) dy -= b->Ly;
      if (dz > hz) dz -= b->Lz;
      if (dx < -hx) dx += b->Lx;
      if (dy < -hy) dy += b->Ly;
      if (dz < -hz) dz += b->Lz;
      dr = dx * dx + dy * dy + dz * dz;
      if (dr > b->radius2) continue;
      t = b->lookup[b->ion[i] * 2 + b->ion[j]];
      pair_terms(b, t, b->pp.Z[b->ion[i]] * b->pp.Z[b->ion[j]], sqrt(dr), &p, &f);
      b->pot += p;
      b->fx[i] += f * dx; b->fy[i] += f * dy; b->fz[i] += f * dz;
      b->fx[j] -= f * dx; b->fy[j] -= f * dy; b->fz[j] -= f * dz;
      b->virX += f * dx * dx;
      b->virY += f * dy * dy;
      b->virZ += f * dz * dz;
    }
  }
  b->pot_real = b->pot;
}

static int ewald_skipped(const orc_box *b)
{
  return b->a1 == 0.0 || b->a1 > 10.0;   /* rcprcl.c:23,115 */
}

/* mdc/program/rcprcl.c:11-97: direct structure-factor sum over the
 * half-sphere 0 < |h|^2 <= hm (hz >= 0; at hz = 0 only hy > 0, or hy = 0
 * and hx > 0).  Both S sums are doubled before use, so the energy added
 * is 4x the energy whose gradient the force is. */
void orc_reciprocal_space(orc_box *b)
{
  int hx, hy, hz, i, j, s = b->hm_sqrt;
  double vv, v2, vp, pot = 0.0;
  double *cz, *sz;   /* Z cos / Z sin per ion: mdc's rcsi/rsni[0..N-1] */
  b->pot_recip = 0.0;
  if (ewald_skipped(b)) return;
  cz = (double *)malloc(sizeof(double) * (size_t)(b->N > 0 ? b->N : 1));
  sz = (double *)malloc(sizeof(double) * (size_t)(b->N > 0 ? b->N : 1));
  vv = b->Lx * b->Ly * b->Lz;
  v2 = 2.0 / vv;
  vp = 1.0 / (ORC_PI * vv);
  for (hx = -s; hx <= s; hx++)
    for (hy = -s; hy <= s; hy++)
      for (hz = 0; hz <= s; hz++) {
        int h2;
        double mx, my, mz, m2, kx, ky, kz, ex, ce, cf, Sc = 0.0, Ss = 0.0;
        if (hz == 0 && (hy < 0 || (hy == 0 && hx <= 0))) continue;
        h2 = hx * hx + hy * hy + hz * hz;
        if (h2 == 0 || h2 > b->hm) continue;
        mx = ((double)hx) / b->Lx;
        my = ((double)hy) / b->Ly;
        mz = ((double)hz) / b->Lz;
        m2 = mx * mx + my * my + mz * mz;
        kx = ORC_PI2 * mx; ky = ORC_PI2 * my; kz = ORC_PI2 * mz;
        ex = exp(-m2 * b->p2a2);
        ce = vp * (ex / m2);
        cf = v2 * (ex / m2);
        for (j = 0; j < b->N; j++) {
          double th = kx * b->rx[j] + ky * b->ry[j] + kz * b->rz[j];
          double zq = b->pp.Z[b->ion[j]];
          cz[j] = zq * cos(th);
          sz[j] = zq * sin(th);
          Sc += cz[j];
          Ss += sz[j];
        }
        Sc *= 2.0;
        Ss *= 2.0;
        pot += ce * (Sc * Sc + Ss * Ss);
        for (i = 0; i < b->N; i++) {
          double foc = -cf * (cz[i] * Ss - sz[i] * Sc);
          b->fx[i] += foc * mx;
          b->fy[i] += foc * my;
          b->fz[i] += foc * mz;
        }
      }
  free(cz);
  free(sz);
  b->pot_recip = pot;
  b->pot += pot;
}

/* mdc/program/rcprcl.c:108-133: self term -alpha/sqrt(pi) sum Z^2. */
void orc_reciprocal_space3(orc_box *b)
{
  int i;
  double z2 = 0.0;
  b->pot_self = 0.0;
  if (ewald_skipped(b)) return;
  for (i = 0; i < b->N; i++) {
    double zq = b->pp.Z[b->ion[i]];
    z2 += zq * zq;
  }
  b->pot_self = -z2 * b->a3;
  b->pot += b->pot_self;
}

/* mdc/program/rv.c:12-18 */
void orc_calc_foc(orc_box *b)
{
  orc_clear_foc(b);
  if (b->real_mode == 1) orc_real_space_analytic(b);
  else orc_real_space(b);
  orc_reciprocal_space(b);
  orc_reciprocal_space3(b);
}

/* mdc/program/rv.c:33-44 */
void orc_next_v(orc_box *b)
{
  int i;
  for (i = 0; i < b->N; i++) {
    double m2 = 2.0 * b->pp.m[b->ion[i]];
    b->vx[i] += b->dt * (b->fx[i] + b->fx0[i]) / m2;
    b->vy[i] += b->dt * (b->fy[i] + b->fy0[i]) / m2;
    b->vz[i] += b->dt * (b->fz[i] + b->fz0[i]) / m2;
  }
}

/* mdc/program/rv.c:46-66: one fold per axis, strict comparisons, so a
 * coordinate equal to L stays L. */
void orc_next_r(orc_box *b)
{
  int i;
  for (i = 0; i < b->N; i++) {
    double m2 = 2.0 * b->pp.m[b->ion[i]];
    b->rx[i] += b->dt * b->vx[i] + b->dt2 * b->fx[i] / m2;
    b->ry[i] += b->dt * b->vy[i] + b->dt2 * b->fy[i] / m2;
    b->rz[i] += b->dt * b->vz[i] + b->dt2 * b->fz[i] / m2;
    if (b->rx[i] > b->Lx) b->rx[i] = b->rx[i] - b->Lx;
    if (b->ry[i] > b->Ly) b->ry[i] = b->ry[i] - b->Ly;
    if (b->rz[i] > b->Lz) b->rz[i] = b->rz[i] - b->Lz;
    if (b->rx[i] < 0.0) b->rx[i] = b->Lx + b->rx[i];
    if (b->ry[i] < 0.0) b->ry[i] = b->Ly + b->ry[i];
    if (b->rz[i] < 0.0) b->rz[i] = b->Lz + b->rz[i];
  }
}

/* mdc/program/rv.c:26-31 */
void orc_next_rv_verlet(orc_box *b)
{
  orc_calc_foc(b);
  orc_next_v(b);
  orc_next_r(b);
}

/* mdc/program/rv.c:68-86 */
void orc_calc_kin(orc_box *b)
{
  int i;
  double kx = 0.0, ky = 0.0, kz = 0.0;
  for (i = 0; i < b->N; i++) {
    double m = b->pp.m[b->ion[i]];
    kx += m * b->vx[i] * b->vx[i];
    ky += m * b->vy[i] * b->vy[i];
    kz += m * b->vz[i] * b->vz[i];
  }
  b->kinX = kx / 2.0;
  b->kinY = ky / 2.0;
  b->kinZ = kz / 2.0;
  b->kin = b->kinX + b->kinY + b->kinZ;
}

/* mdc/program/rv.c:88-116 */
void orc_moment_correction(orc_box *b)
{
  int i;
  double sx = 0.0, sy = 0.0, sz = 0.0, mx, my, mz;
  for (i = 0; i < b->N; i++) {
    double m = b->pp.m[b->ion[i]];
    sx += m * b->vx[i];
    sy += m * b->vy[i];
    sz += m * b->vz[i];
  }
  mx = sx / ((double)b->N);
  my = sy / ((double)b->N);
  mz = sz / ((double)b->N);
  for (i = 0; i < b->N; i++) {
    double m = b->pp.m[b->ion[i]];
    b->vx[i] -= mx / m;
    b->vy[i] -= my / m;
    b->vz[i] -= mz / m;
  }
}

/* mdc/program/pt.c:8-21 (virial is real-space only, as in mdc) */
void orc_calc_press(orc_box *b)
{
  double vol = b->Lx * b->Ly * b->Lz;
  b->presX = (2.0 * b->kinX + b->virX) / vol;
  b->presY = (2.0 * b->kinY + b->virY) / vol;
  b->presZ = (2.0 * b->kinZ + b->virZ) / vol;
  b->pres = (b->presX + b->presY + b->presZ) / 3.0;
}

/* mdc/program/main.c:26-33 with control_temp, moment_correction every
 * 1000 steps and all I/O removed (NVE). */
void orc_nve(orc_box *b, int steps, double *etot)
{
  int s;
  for (s = 1; s <= steps; s++) {
    b->step = s;
    orc_next_rv_verlet(b);
    orc_calc_kin(b);
    orc_calc_press(b);
    if (etot)
      etot[s - 1] = b->kin + b->pot_real + 0.25 * b->pot_recip + b->pot_self;
  }
}

/* mdc/program/msd.c:20-81.  At step 1 the positions are recorded and
 * nothing else changes; afterwards each ion's displacement since the last
 * call, folded once by the strict half-box tests (msd.c:62-67), is added to
 * its running sum, and value[k] = mean |sum|^2 over the ions of species k,
 * summed in ion order. */
void orc_calc_msd(orc_box *b, orc_observer *o)
{
  int i, k;
  if (b->step == 1) {
    for (i = 0; i < b->N; i++) {
      o->rx_old[i] = b->rx[i];
      o->ry_old[i] = b->ry[i];
      o->rz_old[i] = b->rz[i];
    }
    return;
  }
  o->Lx2 = b->Lx / 2.0;
  o->Ly2 = b->Ly / 2.0;
  o->Lz2 = b->Lz / 2.0;
  for (k = 0; k < o->kinds; k++) {
    o->value[k] = 0.0;
    o->number_of_ion[k] = 0;
  }
  for (i = 0; i < b->N; i++) {
    double x = b->rx[i], y = b->ry[i], z = b->rz[i];
    double ex = o->rx_old[i] - x, ey = o->ry_old[i] - y, ez = o->rz_old[i] - z;
    o->rx_old[i] = x;
    o->ry_old[i] = y;
    o->rz_old[i] = z;
    if (ex > o->Lx2) ex -= b->Lx;
    if (ex < -o->Lx2) ex += b->Lx;
    if (ey > o->Ly2) ey -= b->Ly;
    if (ey < -o->Ly2) ey += b->Ly;
    if (ez > o->Lz2) ez -= b->Lz;
    if (ez < -o->Lz2) ez += b->Lz;
    o->dx[i] += ex;
    o->dy[i] += ey;
    o->dz[i] += ez;
    o->value[b->ion[i]] += o->dx[i] * o->dx[i] + o->dy[i] * o->dy[i] + o->dz[i] * o->dz[i];
    o->number_of_ion[b->ion[i]]++;
  }
  for (k = 0; k < o->kinds; k++) o->value[k] /= (double)o->number_of_ion[k];
}

/* mdc/program/pt.c:85-112: kin is summed every call; every d_step steps
 * the velocities are scaled by sqrt((temp / e2t) / mean kin) and the sum
 * restarts.  Returns 1 when it scaled.  (mdc keeps oldk and count in
 * statics, count as a short.) */
int orc_control_temp(orc_box *b, orc_observer *o, int d_step, double temp)
{
  int i;
  double averaged, tempK, scale;
  o->count++;
  o->oldk += b->kin;
  if (b->step % d_step != 0) return 0;
  averaged = o->oldk / (double)o->count;
  o->oldk = 0.0;
  o->count = 0;
  tempK = temp / b->e2t;
  scale = sqrt(tempK / averaged);
  for (i = 0; i < b->N; i++) {
    b->vx[i] *= scale;
    b->vy[i] *= scale;
    b->vz[i] *= scale;
  }
  return 1;
}

/* mdc/program/main.c:26-63 without I/O: next_rv_verlet, calc_kin,
 * calc_press, then calc_msd (if o->msd_on), control_temp while
 * step < t_until (if t_control_step > 0), moment_correction every
 * moment_every steps (if > 0).  Steps first_step .. first_step+steps-1. */
void orc_newton(orc_box *b, orc_observer *o, int first_step, int steps, int t_control_step,
                double temp, int t_until, int moment_every)
{
  int s;
  for (s = first_step; s < first_step + steps; s++) {
    b->step = s;
    orc_next_rv_verlet(b);
    orc_calc_kin(b);
    orc_calc_press(b);
    if (o->msd_on) orc_calc_msd(b, o);
    if (t_control_step > 0 && s < t_until) o->scalings += orc_control_temp(b, o, t_control_step, temp);
    if (moment_every > 0 && s % moment_every == 0) orc_moment_correction(b);
  }
}

int orc_count_kvectors(int hm)
{
  int s = (int)sqrtf((float)hm), hx, hy, hz, n = 0;
  for (hx = -s; hx <= s; hx++)
    for (hy = -s; hy <= s; hy++)
      for (hz = 0; hz <= s; hz++) {
        int h2;
        if (hz == 0 && (hy < 0 || (hy == 0 && hx <= 0))) continue;
        h2 = hx * hx + hy * hy + hz * hz;
        if (h2 == 0 || h2 > hm) continue;
        n++;
      }
  return n;
}
