/*
 * pr_oracle.h -- TEST INFRASTRUCTURE ONLY: a CPU restatement of pr-mdc's
 * force routines (pr-mdc/program/real.c, ewald.c, cell.c), the checker the
 * GPU library libmdgpu_pr.so is compared with.  Never linked by the product.
 *
 * One box in plain arrays instead of pr-mdc's globals.  Matrices are
 * pr-mdc's m[i][j]: the rows of hm are the cell edges a, b, c.
 */
#ifndef PR_ORACLE_H
#define PR_ORACLE_H

typedef struct {
  int N, kinds;
  const int *typ;           /* species of each ion                     */
  const double *Z;          /* charge of each species                  */
  const double *rx, *ry, *rz;
  double *fx, *fy, *fz;     /* forces, added to                        */
  double hm[3][3];          /* cell matrix                             */
  double hmr[3][3];         /* its inverse (orc_pr_calc_cell)          */
  double sigma[3][3];       /* face vectors (orc_pr_calc_sigma)        */
  double g[3][3];           /* sigma / volume                          */
  double vol;
  double cutoff, cutoff2, a1;
  int hmax, hm_sqrt;        /* pr-mdc's sys.hm, sys.hm_sqrt            */
  int td;                   /* sys.table_division                      */
  const double *pe, *fe;    /* BMH table, kinds_of_interaction rows    */
  const int *lookup;        /* kinds x kinds -> table row              */
  double pot;
  double virial[3][3];
} orc_pr_box;

int orc_pr_calc_cell(orc_pr_box *b);
void orc_pr_calc_sigma(orc_pr_box *b);
void orc_pr_set_cutoff(orc_pr_box *b, int step, int skip_ewald);
void orc_pr_min_image(const orc_pr_box *b, double *dx, double *dy, double *dz);
void orc_pr_real_space(orc_pr_box *b);
void orc_pr_ewald1(orc_pr_box *b);
void orc_pr_ewald2(orc_pr_box *b);
void orc_pr_ewald3(orc_pr_box *b);
int orc_pr_count_kvectors(int hmax, int hm_sqrt);
void orc_pr_mk_table(double b, double rho, const double c[3], const double sig[2],
                     const double C[3], const double D[3], int td, double *pe, double *fe,
                     int *lookup);

#endif
