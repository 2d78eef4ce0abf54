/*
 * pr_oracle.c -- CPU restatement of pr-mdc's force routines (parity checker).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/ and the CPU-baseline leg of
 * bench.py, never by the product library.
 *
 * pr-mdc is the reference's constant-pressure program: a triclinic MD cell
 * integrated by the Parrinello-Rahman method.  Its force path is the BMH
 * table pairs (real.c), the Ewald real-space term with Mori's erfc per pair
 * (ewald1), the reciprocal sum over the lattice spanned by the face vectors
 * g (ewald2) and the self term (ewald3).  Every function keeps the
 * operation order of the routine it restates (cited per function, paths
 * relative to the reference repository), so that built with the reference's
 * flags plus -ffp-contract=off it reproduces pr-mdc bit for bit;
 * oracle/_ref/libprmdcref.so (pr-mdc's own sources with a small harness)
 * checks exactly that.
 */
#include <math.h>
#include "mdc_oracle.h"
#include "pr_oracle.h"

/* pr-mdc/program/md.h:7-15, digits as written there */
#define PR_PI     3.1415926535897932385
#define PR_PI2    6.283185307179587
#define PR_PIPI   9.869604401089358619
#define PR_SQRTPI 1.77245385090551602731

/* cell.c:214-252: volume |det H| and H^-1 by cofactors; -1 when det = 0
 * (pr-mdc prints and exits there) */
int orc_pr_calc_cell(orc_pr_box *b)
{
  double (*H)[3] = b->hm, (*R)[3] = b->hmr;
  double det = H[0][0] * (H[1][1] * H[2][2] - H[2][1] * H[1][2]) +
               H[1][0] * (H[2][1] * H[0][2] - H[0][1] * H[2][2]) +
               H[2][0] * (H[0][1] * H[1][2] - H[1][1] * H[0][2]);
  if (det == 0.0) return -1;
  b->vol = fabs(det);
  R[0][0] = (-(H[1][2] * H[2][1]) + H[1][1] * H[2][2]) / det;
  R[1][0] = (H[1][2] * H[2][0] - H[1][0] * H[2][2]) / det;
  R[2][0] = (-(H[1][1] * H[2][0]) + H[1][0] * H[2][1]) / det;
  R[0][1] = (H[0][2] * H[2][1] - H[0][1] * H[2][2]) / det;
  R[1][1] = (-(H[0][2] * H[2][0]) + H[0][0] * H[2][2]) / det;
  R[2][1] = (H[0][1] * H[2][0] - H[0][0] * H[2][1]) / det;
  R[0][2] = (-(H[0][2] * H[1][1]) + H[0][1] * H[1][2]) / det;
  R[1][2] = (H[0][2] * H[1][0] - H[0][0] * H[1][2]) / det;
  R[2][2] = (-(H[0][1] * H[1][0]) + H[0][0] * H[1][1]) / det;
  return 0;
}

/* cell.c:148-175: sigma = {b x c, c x a, a x b}, g = sigma / volume */
void orc_pr_calc_sigma(orc_pr_box *b)
{
  double (*H)[3] = b->hm, (*s)[3] = b->sigma;
  int i, j;
  s[0][0] = (H[1][1] * H[2][2] - H[1][2] * H[2][1]);
  s[0][1] = (H[1][2] * H[2][0] - H[1][0] * H[2][2]);
  s[0][2] = (H[1][0] * H[2][1] - H[1][1] * H[2][0]);
  s[1][0] = (H[2][1] * H[0][2] - H[2][2] * H[0][1]);
  s[1][1] = (H[2][2] * H[0][0] - H[2][0] * H[0][2]);
  s[1][2] = (H[2][0] * H[0][1] - H[2][1] * H[0][0]);
  s[2][0] = (H[0][1] * H[1][2] - H[0][2] * H[1][1]);
  s[2][1] = (H[0][2] * H[1][0] - H[0][0] * H[1][2]);
  s[2][2] = (H[0][0] * H[1][1] - H[0][1] * H[1][0]);
  for (i = 0; i < 3; i++)
    for (j = 0; j < 3; j++) b->g[i][j] = s[i][j] / b->vol;
}

/* cell.c:27-51: cutoff = half the smallest face distance 1/|g_k|, and
 * alpha = 3.8 / that distance; both frozen after step 10 */
void orc_pr_set_cutoff(orc_pr_box *b, int step, int skip_ewald)
{
  double (*g)[3] = b->g, h1, h2, h3, mn = 0.0;
  h1 = 1.0 / sqrt(g[0][0] * g[0][0] + g[0][1] * g[0][1] + g[0][2] * g[0][2]);
  h2 = 1.0 / sqrt(g[1][0] * g[1][0] + g[1][1] * g[1][1] + g[1][2] * g[1][2]);
  h3 = 1.0 / sqrt(g[2][0] * g[2][0] + g[2][1] * g[2][1] + g[2][2] * g[2][2]);
  if ((h1 <= h2) && (h1 <= h3)) mn = h1;
  if ((h2 <= h1) && (h2 <= h3)) mn = h2;
  if ((h3 <= h1) && (h3 <= h2)) mn = h3;
  if (step > 10) return;
  b->cutoff = mn / 2.0;
  b->cutoff2 = b->cutoff * b->cutoff;
  if (skip_ewald == 1) return;
  b->a1 = 3.8 / mn;
}

/* cell.c:305-329: minimum image in scaled coordinates */
void orc_pr_min_image(const orc_pr_box *b, double *dx, double *dy, double *dz)
{
  const double (*H)[3] = (const double (*)[3])b->hm, (*R)[3] = (const double (*)[3])b->hmr;
  double sdx = R[0][0] * (*dx) + R[1][0] * (*dy) + R[2][0] * (*dz);
  double sdy = R[0][1] * (*dx) + R[1][1] * (*dy) + R[2][1] * (*dz);
  double sdz = R[0][2] * (*dx) + R[1][2] * (*dy) + R[2][2] * (*dz);
  if (sdx >= 0.5) sdx -= 1.0;
  if (sdy >= 0.5) sdy -= 1.0;
  if (sdz >= 0.5) sdz -= 1.0;
  if (sdx < -0.5) sdx += 1.0;
  if (sdy < -0.5) sdy += 1.0;
  if (sdz < -0.5) sdz += 1.0;
  *dx = H[0][0] * sdx + H[1][0] * sdy + H[2][0] * sdz;
  *dy = H[0][1] * sdx + H[1][1] * sdy + H[2][1] * sdz;
  *dz = H[0][2] * sdx + H[1][2] * sdy + H[2][2] * sdz;
}

static void add_pair(orc_pr_box *b, int i, int j, double fij, double dx, double dy, double dz)
{
  b->fx[i] += fij * dx;
  b->fy[i] += fij * dy;
  b->fz[i] += fij * dz;
  b->fx[j] -= fij * dx;
  b->fy[j] -= fij * dy;
  b->fz[j] -= fij * dz;
  b->virial[0][0] += fij * dx * dx;
  b->virial[0][1] += fij * dx * dy;
  b->virial[0][2] += fij * dx * dz;
  b->virial[1][1] += fij * dy * dy;
  b->virial[1][2] += fij * dy * dz;
  b->virial[2][2] += fij * dz * dz;
}

static void symmetrise(orc_pr_box *b)
{
  b->virial[1][0] = b->virial[0][1];
  b->virial[2][0] = b->virial[0][2];
  b->virial[2][1] = b->virial[1][2];
}

/* real.c:11-77: BMH from the table at 0.001 A bins from 0.5 A; sets pot and
 * the virial tensor, adds forces */
void orc_pr_real_space(orc_pr_box *b)
{
  int i, j, a, c;
  b->pot = 0.0;
  for (a = 0; a < 3; a++)
    for (c = 0; c < 3; c++) b->virial[a][c] = 0.0;
  for (i = 0; i < b->N; i++) {
    double rxi = b->rx[i], ryi = b->ry[i], rzi = b->rz[i];
    for (j = i + 1; j < b->N; j++) {
      double dx = rxi - b->rx[j], dy = ryi - b->ry[j], dz = rzi - b->rz[j], dr;
      int ij, ddr;
      orc_pr_min_image(b, &dx, &dy, &dz);
      dr = dx * dx + dy * dy + dz * dz;
      if (dr > b->cutoff2) continue;
      ij = b->lookup[b->typ[i] * b->kinds + b->typ[j]];
      ddr = (int)((sqrt(dr) - 0.5) * 1000.0);
      b->pot += b->pe[ij * (b->td + 1) + ddr];
      add_pair(b, i, j, b->fe[ij * (b->td + 1) + ddr], dx, dy, dz);
    }
  }
  symmetrise(b);
}

/* ewald.c:11-68: Z_i Z_j erfc(a r)/r per pair within the cutoff */
void orc_pr_ewald1(orc_pr_box *b)
{
  int i, j;
  for (i = 0; i < b->N; i++) {
    double rxi = b->rx[i], ryi = b->ry[i], rzi = b->rz[i];
    for (j = i + 1; j < b->N; j++) {
      double dx = rxi - b->rx[j], dy = ryi - b->ry[j], dz = rzi - b->rz[j];
      double dr, ZZ, adr, e, fij;
      orc_pr_min_image(b, &dx, &dy, &dz);
      dr = dx * dx + dy * dy + dz * dz;
      if (dr > b->cutoff2) continue;
      dr = sqrt(dr);
      ZZ = b->Z[b->typ[i]] * b->Z[b->typ[j]];
      adr = b->a1 * dr;
      e = orc_erfcc(adr) / dr;
      b->pot += ZZ * e;
      fij = ZZ * (2.0 * (b->a1 / PR_SQRTPI) * exp(-adr * adr) + e) / (dr * dr);
      add_pair(b, i, j, fij, dx, dy, dz);
    }
  }
  symmetrise(b);
}

/* ewald.c:96-105: the half sphere of integer (hx, hy, hz), 0 < |h|^2 <= hm */
static int half_sphere(int hx, int hy, int hz, int hmax)
{
  int h2;
  if (hz == 0 && (hy < 0 || (hy == 0 && hx <= 0))) return 0;
  h2 = hx * hx + hy * hy + hz * hz;
  return !(h2 == 0 || h2 > hmax);
}

int orc_pr_count_kvectors(int hmax, int hm_sqrt)
{
  int n = 0, hx, hy, hz;
  for (hx = -hm_sqrt; hx <= hm_sqrt; hx++)
    for (hy = -hm_sqrt; hy <= hm_sqrt; hy++)
      for (hz = 0; hz <= hm_sqrt; hz++) n += half_sphere(hx, hy, hz, hmax);
  return n;
}

/* ewald.c:75-159: for each half-sphere h, hh = h . g (rows of g are the
 * reciprocal basis), S = sum_j Z_j exp(2 pi i hh . r_j) doubled; energy
 * vp e/|hh|^2 |2S|^2 with vp = 1/(pi V) -- four times the textbook
 * reciprocal energy, as pr-mdc writes it -- and force
 * -(2/V) e/|hh|^2 (c_i 2S_im - s_i 2S_re) hh on ion i */
void orc_pr_ewald2(orc_pr_box *b)
{
  const int N = b->N, hs = b->hm_sqrt;
  double (*g)[3] = b->g;
  double v2 = 2.0 / b->vol, vp = 1.0 / (PR_PI * b->vol), pot = 0.0;
  int hx, hy, hz, i, j;
  for (hx = -hs; hx <= hs; hx++)
    for (hy = -hs; hy <= hs; hy++)
      for (hz = 0; hz <= hs; hz++) {
        double hhx, hhy, hhz, hh2, hhx1, hhy1, hhz1, ex, vpe, v2e, C = 0.0, S = 0.0;
        if (!half_sphere(hx, hy, hz, b->hmax)) continue;
        hhx = hx * g[0][0] + hy * g[1][0] + hz * g[2][0];
        hhy = hx * g[0][1] + hy * g[1][1] + hz * g[2][1];
        hhz = hx * g[0][2] + hy * g[1][2] + hz * g[2][2];
        hh2 = hhx * hhx + hhy * hhy + hhz * hhz;
        hhx1 = PR_PI2 * hhx;
        hhy1 = PR_PI2 * hhy;
        hhz1 = PR_PI2 * hhz;
        ex = exp(-hh2 * (PR_PIPI / (b->a1 * b->a1)));
        vpe = vp * (ex / hh2);
        v2e = v2 * (ex / hh2);
        for (j = 0; j < N; j++) {
          double ti = hhx1 * b->rx[j] + hhy1 * b->ry[j] + hhz1 * b->rz[j];
          C += b->Z[b->typ[j]] * cos(ti);
          S += b->Z[b->typ[j]] * sin(ti);
        }
        C *= 2.0;
        S *= 2.0;
        pot += vpe * (C * C + S * S);
        for (i = 0; i < N; i++) {
          double ti = hhx1 * b->rx[i] + hhy1 * b->ry[i] + hhz1 * b->rz[i];
          double ci = b->Z[b->typ[i]] * cos(ti), si = b->Z[b->typ[i]] * sin(ti);
          double foc = -v2e * (ci * S - si * C);
          b->fx[i] += foc * hhx;
          b->fy[i] += foc * hhy;
          b->fz[i] += foc * hhz;
        }
      }
  b->pot += pot;
}

/* ewald.c:162-173 */
void orc_pr_ewald3(orc_pr_box *b)
{
  double Z2 = 0.0;
  int i;
  for (i = 0; i < b->N; i++) Z2 += b->Z[b->typ[i]] * b->Z[b->typ[i]];
  b->pot += -Z2 * (b->a1 / PR_SQRTPI);
}

/* pr-mdc/program/control.c:192-240: BMH only (pr-mdc evaluates the Ewald
 * term per pair); c, C, D per pair type 0 = (0,0), 1 = (0,1), 2 = (1,1),
 * sig per species */
void orc_pr_mk_table(double b, double rho, const double c[3], const double sig[2],
                     const double C[3], const double D[3], int td, double *pe, double *fe,
                     int *lookup)
{
  int ion_i, ion_j, ddr, ij = 0;
  for (ion_i = 0; ion_i < 2; ion_i++)
    for (ion_j = ion_i; ion_j < 2; ion_j++) {
      int t = ion_i + ion_j;
      double cij = c[t], sij = sig[ion_i] + sig[ion_j], Cij = C[t], Dij = D[t];
      for (ddr = 0; ddr < td + 1; ddr++) {
        double dr = ((double)ddr) / 1000.0 + 0.5;
        double dr6 = dr * dr * dr * dr * dr * dr, dr8 = dr6 * dr * dr;
        pe[ij * (td + 1) + ddr] = cij * b * exp((sij - dr) / rho) - Cij / dr6 - Dij / dr8;
        fe[ij * (td + 1) + ddr] =
            (cij * b * exp((sij - dr) / rho) / (rho) - 6.0 * Cij / (dr6 * dr) -
             8.0 * Dij / (dr8 * dr)) / dr;
      }
      lookup[2 * ion_i + ion_j] = ij;
      lookup[2 * ion_j + ion_i] = ij;
      ij++;
    }
}
