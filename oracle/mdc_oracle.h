/*
 * mdc_oracle.h -- CPU restatement of mdc's classical-MD force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the
 * MI355X engine: only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it.  No product code links or calls it.
 *
 * It restates, function by function, the single-threaded ANSI-C path of
 * mdc (KazumeNishidate/Molecular-dynamics-in-C, mdc/program):
 *
 *   orc_erfcc              ext.c:73-104       Mori's erfc
 *   orc_unit_converter     init.c:25-80       derived parameters
 *   orc_potential_nacl     control.c:85-181   Tosi-Fumi NaCl
 *   orc_potential_kcl      (not in mdc)       Tosi-Fumi KCl, same functional form
 *   orc_set_loc            control.c:46-82    rock-salt lattice + species
 *   orc_mk_table           control.c:183-248  erfc+BMH pair table
 *   orc_clear_foc          init.c:179-190
 *   orc_real_space         real.c:15-84       table mode, all pairs i<j
 *   orc_real_space_analytic pr-mdc/program/ewald.c:11-68 (+ BMH of control.c:219-240)
 *   orc_reciprocal_space   rcprcl.c:11-97
 *   orc_reciprocal_space3  rcprcl.c:108-133
 *   orc_calc_foc           rv.c:12-18
 *   orc_next_v / next_r    rv.c:33-66
 *   orc_calc_kin           rv.c:68-86
 *   orc_moment_correction  rv.c:88-116
 *   orc_calc_press         pt.c:8-21
 *   orc_nve                main.c:26-33 without thermostat and I/O
 *   orc_calc_msd           msd.c:20-81
 *   orc_control_temp       pt.c:85-112
 *   orc_newton             main.c:26-63 without I/O (MSD, thermostat and
 *                          moment-correction cadence included)
 *
 * Differences from the reference that do not change results on the same
 * inputs: `int` ion counts (mdc uses `short`, which caps N at 32767), an
 * explicit box struct instead of globals, and the reciprocal-space energy
 * also kept separately so that a consistent (reciprocal/4) total can be
 * reported next to the reference's own `pot`.
 *
 * Parity status per function (tests/test_oracle.py, tests/test_abi.py):
 *   erfcc            compiled mdc (bit-equal), math.erfc (|err| < 1e-14)
 *   unit_converter   compiled mdc (tables, trajectory and scalars bit-equal)
 *   potential_nacl   compiled mdc tables bit-equal; closed-form table entries
 *   potential_kcl    mdc has no KCl (tests/test_kcl_constants.py): closed
 *                    forms (Pauling factors of closed 8-electron shells,
 *                    sigma_ij = sigma_i + sigma_j, Cl- radius, b and the
 *                    C/D unit conversion shared with mdc's NaCl set); the
 *                    static crystal's lattice constant, lattice energy and
 *                    bulk modulus against measured data, NaCl as control,
 *                    each plausible constant mistake failing; the crystal
 *                    energy equal to a closed-form lattice sum to 1e-9
 *   set_loc          compiled mdc (positions and species bit-equal)
 *   mk_table         compiled mdc (bit-equal); closed form; -U'/r by
 *                    centred differences
 *   clear_foc, calc_foc, next_v/next_r, calc_kin, calc_press
 *                    compiled mdc 60-step trajectory with every scalar
 *                    bit-equal; golden fixture tests/golden/mdc_nacl216.npz
 *   real_space       compiled mdc forces/pot/virial bit-equal; third law;
 *                    zero force on the perfect lattice; golden fixture
 *   real_space_analytic (pr-mdc form, not compiled here): -dE/dx = F by
 *                    finite differences; Madelung constant of rock salt
 *   reciprocal_space compiled mdc bit-equal; S(k) = 0 on the lattice;
 *                    -d(E/4)/dx = F; Madelung constant
 *   reciprocal_space3 compiled mdc (pot bit-equal)
 *   moment_correction compiled mdc (bit-equal); zero total momentum
 *   nve              compiled mdc trajectory; golden fixture
 *   calc_msd         compiled mdc (msd arrays and values bit-equal over a
 *                    trajectory); uniform translation across the box edge
 *                    gives MSD = |n t d|^2 (closed form)
 *   control_temp     compiled mdc (velocities bit-equal, incl. statics over
 *                    several periods); after a scaling with constant kin
 *                    the kinetic energy is the target (closed form)
 *   newton           compiled mdc loop of main.c:26-63 (no I/O), bit-equal
 */
#ifndef MDC_ORACLE_H
#define MDC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Born-Mayer-Huggins parameters for a two-species salt.
 * Pair type index t: 0 = (0,0), 1 = (0,1), 2 = (1,1). */
typedef struct {
  double Z[2];      /* effective charges, sqrt(kk) units          */
  double m[2];      /* masses [m_u]                                */
  double b;         /* BMH prefactor [1e-18 J]                     */
  double rho;       /* BMH softness [A]                            */
  double c[3];      /* Pauling factors per pair type               */
  double sigma[3];  /* sigma_i + sigma_j per pair type [A]         */
  double C[3];      /* dipole-dipole dispersion [1e-18 J A^6]      */
  double D[3];      /* dipole-quadrupole dispersion [1e-18 J A^8]  */
} orc_potential;

typedef struct {
  /* ---- parameters (orc_unit_converter fills the derived ones) ---- */
  int     N;
  int     nx, ny, nz;       /* conventional cells                   */
  double  Ax, Ay, Az;       /* lattice constants [A]                */
  double  Lx, Ly, Lz;       /* box [A]                              */
  double  a1;               /* Ewald alpha [1/A]; 0 or >10 skips    */
  double  p2a2, a3;         /* pi^2/alpha^2, alpha/sqrt(pi)         */
  int     hm, hm_sqrt;      /* integer |h|^2 cutoff and its root    */
  double  radius, radius2;  /* real-space cutoff and its square     */
  double  delta_time_fs;
  double  dt, dt2;
  int     table_division;
  double  kB, kk, e2t;
  int     real_mode;        /* 0 = table (mdc), 1 = analytic (pr-mdc) */
  orc_potential pp;
  int     lookup[4];
  /* ---- arrays (caller-owned) ---- */
  double *pe1r1, *fe1r1;    /* [3*(table_division+1)]               */
  double *rx, *ry, *rz, *vx, *vy, *vz;
  double *fx, *fy, *fz, *fx0, *fy0, *fz0;
  int    *ion;              /* species 0/1                          */
  /* ---- scalars ---- */
  int     step;
  double  pot;              /* as mdc accumulates it                */
  double  pot_real, pot_recip, pot_self;  /* pieces of pot          */
  double  virX, virY, virZ;
  double  kin, kinX, kinY, kinZ;
  double  presX, presY, presZ, pres;
} orc_box;

/* calc_msd's state (mdc/program/md.h:152-172) and control_temp's statics
 * (pt.c:88-89) for one box; the six arrays are caller-owned [N]. */
typedef struct {
  int     kinds;
  int     msd_on;
  double *rx_old, *ry_old, *rz_old, *dx, *dy, *dz;
  double  value[4];
  int     number_of_ion[4];
  double  Lx2, Ly2, Lz2;
  double  oldk;
  int     count;
  int     scalings;         /* velocity scalings done (not in mdc) */
} orc_observer;

double orc_erfcc(double x);
void   orc_potential_nacl(double kk, orc_potential *pp);
void   orc_potential_kcl(double kk, orc_potential *pp);
void   orc_unit_converter(orc_box *b);
void   orc_set_loc(orc_box *b);
void   orc_mk_table(orc_box *b);
void   orc_clear_foc(orc_box *b);
void   orc_real_space(orc_box *b);
void   orc_real_space_analytic(orc_box *b);
void   orc_reciprocal_space(orc_box *b);
void   orc_reciprocal_space3(orc_box *b);
void   orc_calc_foc(orc_box *b);
void   orc_next_v(orc_box *b);
void   orc_next_r(orc_box *b);
void   orc_next_rv_verlet(orc_box *b);
void   orc_calc_kin(orc_box *b);
void   orc_moment_correction(orc_box *b);
void   orc_calc_press(orc_box *b);
/* NVE driver: `steps` iterations of (next_rv_verlet, calc_kin, calc_press).
 * If etot != NULL, etot[s] = kin + pot_real + pot_recip/4 + pot_self after
 * step s (the energy whose gradient the forces are). */
void   orc_nve(orc_box *b, int steps, double *etot);
void   orc_calc_msd(orc_box *b, orc_observer *o);
int    orc_control_temp(orc_box *b, orc_observer *o, int d_step, double temp);
/* main.c:26-63 without I/O (see mdc_oracle.c) */
void   orc_newton(orc_box *b, orc_observer *o, int first_step, int steps, int t_control_step,
                  double temp, int t_until, int moment_every);
/* number of half-sphere k vectors that reciprocal_space visits */
int    orc_count_kvectors(int hm);

#ifdef __cplusplus
}
#endif
#endif
