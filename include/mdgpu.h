/*
 * mdgpu.h -- handle-based C ABI of libmdgpu: device-resident boxes,
 * batched stepping of many boxes per launch, and a serving loop.
 *
 * mdc itself has one global system per process (mdc/program/md.h:197-202)
 * and no API beyond `void f(void)` routines; mdc_abi.h keeps that surface.
 * This header is the surface mdc's singleton cannot express: any number of
 * independent boxes whose state stays in HBM between steps.  Each step
 * performs, per box, what mdc's newton() loop does per step
 * (mdc/program/main.c:26-33): calc_foc (rv.c:12-18) -> next_v -> next_r
 * (rv.c:33-66) -> calc_kin (rv.c:68-86) -> calc_press (pt.c:8-21).
 *
 * Units are mdc's (mdc/manual/units.IEMD): length A, energy 1e-18 J,
 * mass m_u, time 4.074972637944947 fs.  Arrays are SoA: r = x[N], y[N], z[N].
 * Every function returns MDG_OK or an MDG_E* code; none exits the process.
 */
#ifndef MDGPU_H
#define MDGPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MDG_OK          0
#define MDG_E_NODEVICE  1   /* no HIP device / library built without one  */
#define MDG_E_HIP       2   /* a HIP runtime call failed                   */
#define MDG_E_ARG       3   /* invalid argument                            */
#define MDG_E_NOMEM     4   /* device allocation failed                    */
#define MDG_E_BUDGET    5   /* request exceeds the HBM budget              */
#define MDG_E_STATE     6   /* non-finite or out-of-box state detected     */
#define MDG_E_RANGE     7   /* pair inside the 0.5 A table start, or rc > L/2 */
#define MDG_E_TIMEOUT   8   /* a bounded device wait expired               */

#define MDG_FP64     0      /* plain path: all arithmetic FP64             */
#define MDG_MIXED    1      /* FP32 pair/k-space arithmetic, FP64 state and sums */

#define MDG_TABLE    0      /* mdc's tabulated erfc+BMH (0.001 A bins)     */
#define MDG_ANALYTIC 1      /* per-pair evaluation (pr-mdc ewald1 style)   */

#define MDG_CELLS_AUTO  -1
#define MDG_ALL_PAIRS    0
#define MDG_CELL_LIST    1

/* force terms (mdg_batch_force_terms) */
#define MDG_TERM_REAL  1
#define MDG_TERM_RECIP 2

/* Born-Mayer-Huggins salt: pair type 0 = (0,0), 1 = (0,1), 2 = (1,1). */
typedef struct {
  double Z[2];       /* effective charges (sqrt(kk) = 1 e)          */
  double m[2];       /* masses [m_u]                                 */
  double b, rho;     /* BMH prefactor [1e-18 J], softness [A]        */
  double c[3];       /* Pauling factors                              */
  double sigma[3];   /* sigma_i + sigma_j [A]                        */
  double C[3], D[3]; /* dispersion [1e-18 J A^6], [1e-18 J A^8]      */
} mdg_potential;

typedef struct {
  int    nx, ny, nz;    /* conventional rock-salt cells, N = 8 nx ny nz   */
  double ax, ay, az;    /* lattice constants [A]; L = a * n              */
  double alpha;         /* Ewald alpha [1/A]; 0 or > 10 turns Ewald off  */
  int    hm;            /* reciprocal cutoff: 0 < hx^2+hy^2+hz^2 <= hm   */
  double rc;            /* real-space cutoff [A]                         */
  double dt_fs;         /* time step [fs]                                */
  int    real_mode;     /* MDG_TABLE or MDG_ANALYTIC                     */
  int    cells;         /* MDG_CELLS_AUTO, MDG_ALL_PAIRS or MDG_CELL_LIST */
  mdg_potential pot;
} mdg_box_params;

typedef struct {
  double pot;           /* as mdc accumulates sys.pot: real + recip + self */
  double pot_real;      /* real-space erfc + BMH                           */
  double pot_recip;     /* reciprocal term as mdc writes it (4x, see below) */
  double pot_self;      /* -alpha/sqrt(pi) sum Z^2                         */
  double etot;          /* kin + pot_real + pot_recip/4 + pot_self: the
                           energy the forces conserve (mdc doubles both
                           structure-factor sums before squaring them,
                           rcprcl.c:77-82)                                  */
  double vir[3];        /* diagonal real-space virial (pt.c uses it)       */
  double kin, kin_xyz[3];
  double pres, pres_xyz[3];
  long long step;       /* completed steps                                 */
} mdg_scalars;

typedef struct mdg_batch mdg_batch;

/* ---------------------------------------------------------------- setup */
int  mdg_device_count(int *n);
/* Tosi-Fumi NaCl (mdc/program/control.c:85-181) and KCl (same form). */
int  mdg_potential_nacl(mdg_potential *p);
int  mdg_potential_kcl(mdg_potential *p);
/* Rock-salt positions and species in the order of mdc's set_loc /
 * identify_ion (control.c:46-82).  r: 3N, species: N. */
int  mdg_rocksalt(int nx, int ny, int nz, double ax, double ay, double az,
                  double *r, int *species);
/* Seeded Maxwell-Boltzmann velocities at temp_k with the centre-of-mass
 * momentum removed as mdc's moment_correction does (rv.c:88-116). */
int  mdg_maxwell(unsigned long long seed, double temp_k, int n, const int *species,
                 const mdg_potential *p, double *v);
/* Ewald split for a target relative error eps: alpha = s/rc and
 * hm = floor((s alpha L_max / pi)^2) with s = sqrt(-ln eps). */
int  mdg_ewald_params(double rc, double l_max, double eps, double *alpha, int *hm);
/* No counterpart in mdc (mdc fixes rc per run): a real-space cutoff for a
 * box of n_ions rock-salt ions whose Ewald split mdg_ewald_params then
 * makes, chosen by measurement on MI355X: rc = 16.5 A (n_ions / 32768)^(1/6)
 * (the pair sum and the k sum balance as N^(1/6) at fixed density), at
 * least 10 A, at most rc_cap and L_min / 2. */
int  mdg_ewald_rc_rule(int n_ions, double l_min, double rc_cap, double *rc);

/* ---------------------------------------------------------------- batch */
/* nbox independent boxes in one precision mode.  Each box's arithmetic
 * depends on that box alone (tiling, reduction order), so a box steps to
 * identical bits alone or in any mix. */
int  mdg_batch_create(int nbox, const mdg_box_params *params, int precision,
                      mdg_batch **out);
/* Boxes of any of mdc's tabulated materials (mdc/examples/<m>: the tables,
 * lookup, charges and masses that the material's set_potential and mk_table
 * make, control.c), in table mode: an orthorhombic box of n_ions ions whose
 * positions and species come from mdg_batch_set_state.  The tables are
 * copied at creation. */
typedef struct {
  int           n_ions;
  double        L[3];                /* box [A]                                  */
  int           cells[3];            /* unit cells per axis (sys.nx, ny, nz)     */
  int           natoms_in_mol_unit;  /* ctl.natoms_in_mol_unit                   */
  int           kinds;               /* ctl.kinds_of_ions, <= 4                  */
  double        Z[4], m[4];          /* ion.Z, ion.m per species                 */
  int           lookup[16];          /* sys.lookup: pair type at ti * kinds + tj */
  int           ntypes;              /* kinds (kinds + 1) / 2                    */
  int           table_division;      /* sys.table_division (td + 1 bins)         */
  const double *pe, *fe;             /* sys.pe1r1, sys.fe1r1: ntypes x (td + 1)  */
  double        alpha;               /* sys.a1; 0 or > 10 turns Ewald off        */
  int           hm;                  /* sys.hm                                   */
  double        rc;                  /* sys.radius                               */
  double        dt_fs;               /* ctl.delta_time_fs                        */
  int           cells_mode;          /* MDG_CELLS_AUTO, MDG_ALL_PAIRS, MDG_CELL_LIST */
} mdg_table_box;
int  mdg_batch_create_tabulated(int nbox, const mdg_table_box *boxes, int precision,
                                mdg_batch **out);
int  mdg_batch_destroy(mdg_batch *b);
int  mdg_batch_n_ions(const mdg_batch *b, int box);
/* device bytes held by box `box` (all its arrays), for HBM budgeting */
size_t mdg_batch_box_bytes(const mdg_batch *b, int box);
/* r: 3N SoA [A], v: 3N or NULL (keeps), species: N or NULL (keeps) */
int  mdg_batch_set_state(mdg_batch *b, int box, const double *r, const double *v,
                         const int *species);
/* f0 is the previous step's force (mdc's fx0..); any pointer may be NULL */
int  mdg_batch_get_state(mdg_batch *b, int box, double *r, double *v, double *f,
                         double *f0);
/* mdc's calc_foc for every box: f0 <- f, f <- real + reciprocal forces,
 * pot/vir/self as mdc writes them (no integration). */
int  mdg_batch_forces(mdg_batch *b);
/* One force term alone, as mdc's real_space (mdc/program/real.c:15-84) and
 * reciprocal_space (mdc/program/rcprcl.c:11-97) each add theirs: f0 <- f,
 * f <- the selected terms.  terms: MDG_TERM_REAL | MDG_TERM_RECIP.  The
 * energies are left as they were. */
int  mdg_batch_force_terms(mdg_batch *b, int terms);
/* nsteps velocity-Verlet steps (next_rv_verlet + calc_kin + calc_press) on
 * every box; the step sequence is captured once in a hipGraph and replayed */
int  mdg_batch_step(mdg_batch *b, int nsteps);
/* No counterpart in mdc: captures, instantiates and uploads the 8-step
 * graph that mdg_batch_step replays, executing no step, so that a timed
 * or latency-bound stretch of steps that follows contains no capture.
 * Membership or state changes that rebuild the work lists discard it. */
int  mdg_batch_prepare(mdg_batch *b);
/* No counterpart in mdc: times nsteps steps of the batch under each of a
 * few splits of its boxes between the step's two concurrent groups and
 * keeps the fastest, each timed twice, its faster time counting (the
 * choice moves a 256-box step by several percent and depends on which
 * boxes share a group).  Every box's device state is
 * saved first and restored afterwards, so the trajectory, and every bit
 * of it, is as if nothing had run; the step graph is then prepared.
 * Needs device memory for one copy of the batch's boxes.  *share_out (may
 * be NULL) receives the chosen share of the group that starts its
 * k-space work at once. */
int  mdg_batch_tune_schedule(mdg_batch *b, int nsteps, double *share_out);
int  mdg_batch_scalars(mdg_batch *b, int box, mdg_scalars *out);
int  mdg_batch_sync(mdg_batch *b);

/* The rest of mdc's step loop (mdc/program/main.c:44-60), run on the device
 * inside every step for the boxes that arm it, in mdc's order after the
 * integration and calc_kin/calc_press:
 *   calc_msd           mdc/program/msd.c:20-81, every step;
 *   control_temp       mdc/program/pt.c:85-112, while step < t_control_until;
 *   moment_correction  mdc/program/rv.c:88-116, every moment_every steps.
 * The step is the box's own count of completed steps (mdc's sys.step).
 * Arming resets the state that mdc keeps in msd and in control_temp's
 * statics: the MSD origin is the positions after the next step, and the
 * thermostat's kinetic-energy sum restarts.  All fields 0 disarms. */
typedef struct {
  int       msd;              /* 1: calc_msd every step                  */
  int       t_control_step;   /* ctl.t_control_step; 0: no thermostat    */
  double    temp_K;           /* ctl.temp                                */
  long long t_control_until;  /* thermostat while step < this (mdc 10000) */
  int       moment_every;     /* mdc: 1000; 0: off                       */
} mdg_observe;
int  mdg_batch_set_observe(mdg_batch *b, int box, const mdg_observe *o);
/* msd[4]: MSD per species [A^2] (msd.value) after the last step;
 * scalings: velocity scalings done since arming.  Either may be NULL. */
int  mdg_batch_observed(mdg_batch *b, int box, double *msd, long long *scalings);
/* MSD state: old = positions of the last sample (msd.rx_old, ry_old,
 * rz_old), d = summed displacements (msd.dx, dy, dz); 3N SoA each, or NULL */
int  mdg_batch_msd_state(mdg_batch *b, int box, double *old, double *d);

/* ------------------------------------------------------- state read-back */
/* One box's state at a recorded step, in pinned host memory that is valid
 * for the duration of the callback. */
typedef struct {
  int                box, n_ions;
  long long          step;         /* the box's completed steps          */
  const double      *r, *v;        /* 3N SoA each: x block, y, z [A], [A/time unit] */
  const int         *species;      /* N                                   */
  mdg_scalars        scalars;      /* as mdg_batch_scalars                */
  double             msd[4];       /* as mdg_batch_observed (0 unarmed)   */
} mdg_frame;
typedef void (*mdg_frame_fn)(const mdg_frame *f, void *user);
/* Steps every box nsteps (a multiple of every) and after each `every`
 * steps hands every box's state to fn on the calling thread, while the
 * device goes on stepping: at that point of the step stream one kernel
 * gathers all boxes' r, v, scalars and MSD into a device staging slot,
 * a second stream copies the slot to pinned host memory, and frame k is
 * delivered while the steps of frame k + 1 run (two slots).  fn may be
 * NULL (copy only).  The stepping itself is mdg_batch_step's, so the
 * trajectory is the same bits as without recording. */
int  mdg_batch_step_recorded(mdg_batch *b, int nsteps, int every, mdg_frame_fn fn, void *user);

/* mdc's output files (mdc/program/files.c), written from the device-resident
 * boxes through mdg_batch_step_recorded: per box, dir/box<k>/out with one
 * print_to_file line per frame (files.c:44-65), and, by the bits of
 * `what`, dir/box<k>/positions (MDG_WRITE_POSITIONS, r_position,
 * files.c:142-158) and velocities (MDG_WRITE_VELOCITIES, r_velocities,
 * files.c:161-181).  Each file starts with mdc's single blank
 * (files.c:29-32).  dir/box<k> must exist. */
#define MDG_WRITE_POSITIONS  1
#define MDG_WRITE_VELOCITIES 2
int  mdg_batch_write_mdc_files(mdg_batch *b, int nsteps, int every, int what, const char *dir);

/* The text of mdc's formats, snprintf-like: writes at most cap bytes
 * (NUL-terminated when cap > 0) and returns the length the whole text
 * needs, or a negative MDG_E_* code.
 *   mdg_format_out:        one print_to_file line (files.c:44-65), with
 *                          perMol, e2t and pp2gpa derived from N as
 *                          unit_converter does (init.c:48, 75, 78);
 *   mdg_format_positions:  one r_position record (files.c:142-158);
 *   mdg_format_velocities: one r_velocities record (files.c:161-181). */
typedef struct {
  int    step;
  double kin, pot, pres;        /* sys.kin, sys.pot, sys.pres          */
  int    n_ions, natoms_in_mol_unit;
  double L[3];                  /* sys.Lx, Ly, Lz                      */
  int    cells[3];              /* sys.nx, ny, nz                      */
  int    kinds;                 /* ctl.kinds_of_ions (<= 4)            */
  double msd[4];                /* msd.value                           */
} mdg_out_fields;
int  mdg_format_out(char *buf, size_t cap, const mdg_out_fields *f);
int  mdg_format_positions(char *buf, size_t cap, int n, const int *species, const double *r);
int  mdg_format_velocities(char *buf, size_t cap, int n, const int *species, const double *v);
/* Kernel timing over the next `nsteps` steps, launched eagerly in the
 * step's own schedule (box groups on several streams); every launch
 * carries a start and a stop hipEvent stamped by its own dispatch
 * (hipExtLaunchKernel).  Per kernel name:
 * ms_step = ms per step summed over its launches, launches = launches per
 * step (ms_step / launches is its average launch duration).  Up to `cap`
 * entries; the count is returned in *n.  A negative nsteps times -nsteps
 * steps with every kernel launched alone over all boxes, one after the
 * other, which gives each kernel's duration without the others beside it. */
int  mdg_batch_profile(mdg_batch *b, int nsteps, int cap, char names[][32],
                       double *ms_step, int *launches, int *n);
/* Algorithmic work of one step of box `box`: accepted real-space pairs
 * (unordered), k vectors, and ions; pairs are counted on the device at the
 * last force evaluation. */
int  mdg_batch_work(mdg_batch *b, int box, long long *pairs, int *nk, int *n_ions);

const char *mdg_error_string(int code);

/* No counterpart in mdc.  A plain device copy of `bytes` bytes (read
 * `bytes`, write `bytes`), run reps times after one warm-up: the kernel of
 * known traffic that scales hardware byte counters, and an HBM bandwidth
 * yardstick.  *ms_avg: average time of one copy. */
int  mdg_copy_probe(size_t bytes, int reps, double *ms_avg);

/* ------------------------------------------------------ allocation-free */
/* No counterpart in mdc.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): mdc's calc_foc (rv.c:12-18) for one box, as mdg_batch_forces
 * computes it for a batch of that box alone, on the caller's device
 * buffers, allocating nothing.
 *   r:        3N positions [A], SoA (x block, y, z), N = 8 nx ny nz;
 *   species:  N ints (0 = cation, 1 = anion);
 *   f:        3N forces, SoA (written);
 *   energies: 8 doubles (written): pot, pot_real, pot_recip, pot_self,
 *             pot_real + pot_recip / 4 + pot_self, vir[3], as mdg_scalars;
 *   stream:   a hipStream_t, or NULL for the null stream.
 * Every other device buffer sits in one static device array of
 * MDG_NOALLOC_SCRATCH_BYTES per device, held until the process ends; a box
 * that needs more is refused with MDG_E_BUDGET (a 4096-ion box at
 * Ewald 1e-6, rc = 10 A needs about 18 MiB).  All calls on a device share
 * that array and it keeps the last call's box: calls with the same
 * parameters and precision reuse its uploaded plan.  Calls that overlap
 * on one device race (calls are serialised within a process).  Returns
 * after the results are written: MDG_OK, MDG_E_ARG (a species out of
 * range), MDG_E_STATE (a non-finite coordinate), MDG_E_RANGE, MDG_E_BUDGET. */
#define MDG_NOALLOC_SCRATCH_BYTES (32u << 20)
int  mdg_forces_noalloc(const mdg_box_params *p, int precision, const double *r,
                        const int *species, double *f, double *energies, void *stream);

/* -------------------------------------------------------------- serving */
typedef struct {
  double arrival_s;     /* seconds after the start of the stream          */
  int    nx, ny, nz;    /* NaCl box in conventional cells                  */
  double temp_k;
  int    steps;
  unsigned long long seed;
} mdg_job;

typedef struct {
  size_t hbm_budget;    /* bytes the resident jobs may hold               */
  int    max_slots;     /* jobs stepped together at most                   */
  int    precision;     /* MDG_FP64 or MDG_MIXED                           */
  int    real_mode;
  double eps;           /* Ewald target relative error                     */
  double rc_max;        /* real-space cutoff cap [A]                       */
  double dt_fs;
  int    one_at_a_time; /* 1: run jobs alone in arrival order (baseline)   */
  int    realtime;      /* 1: jobs become visible at their arrival times    */
  int    steps_per_tick;/* steps per batched launch between admissions     */
  int    auto_split;    /* 1: rc from mdg_ewald_rc_rule(N, L_min, rc_max)
                           instead of min(rc_max, L_min / 2)               */
} mdg_serve_config;

typedef struct {
  double admit_s, finish_s, latency_s;   /* latency = finish - arrival     */
  double etot_first, etot_last;          /* energy after first / last step */
  double final_checksum;                 /* sum of final positions          */
  int    status;                         /* MDG_OK or the job's error       */
} mdg_job_result;

typedef struct {
  double wall_s;
  double ion_steps;
  double ion_steps_per_s;
  double p50_latency_s, p99_latency_s, mean_latency_s;
  int    max_resident, ticks, rejected;
  size_t peak_bytes;
} mdg_serve_stats;

int  mdg_serve(const mdg_job *jobs, int njobs, const mdg_serve_config *cfg,
               mdg_job_result *results, mdg_serve_stats *stats);
/* the box parameters mdg_serve gives a job (NaCl, a = 5.64 A, Ewald split
 * for cfg->eps with rc = min(cfg->rc_max, L_min / 2), or by
 * mdg_ewald_rc_rule when cfg->auto_split), e.g. to tune them */
int  mdg_serve_job_params(const mdg_job *job, const mdg_serve_config *cfg,
                          mdg_box_params *out);

/* ------------------------------------------------------------- autotune */
/* No counterpart in mdc (one box, one serial loop).  Tile and cell sizes
 * of a box: real-space cell list or all pairs, cell width, partner splits;
 * structure-factor ion chunks; back-projection k groups. */
typedef struct {
  int    cells;      /* 1: cell list, 0: all pairs                        */
  double cell_w;     /* cell width / rc (>= 1)                             */
  int    rsplit;     /* real-space partner splits per home unit            */
  int    nchunk;     /* structure-factor ion chunks                        */
  int    ngroup;     /* back-projection k groups                           */
  int    tuned;      /* 1: from the tuning table, 0: built-in rule         */
} mdg_tiling;

typedef struct {
  int    classes;    /* box classes measured                               */
  int    runs;       /* replicated-batch runs timed                        */
  double seconds;    /* wall time of the tuning                            */
} mdg_tune_report;

/* Measures candidate tilings for every box class among params (classes:
 * precision, real-space mode, cell list possible, sqrt(2)-wide buckets of
 * N and of the k-vector count).  Each class is timed in a batch of copies
 * of its first box holding about context_ions ions (0: 131072), with
 * per-kernel HIP events; the real-space, structure-factor and
 * back-projection choices are made separately.  Results go to the
 * process-wide tuning table, which boxes created afterwards consult.  A
 * box's arithmetic depends on its tiling, so a fixed table keeps results
 * identical alone and in any mix; save/load fix it across processes. */
int  mdg_tune(int nbox, const mdg_box_params *params, int precision,
              long long context_ions, mdg_tune_report *rep);
int  mdg_tune_clear(void);
int  mdg_tune_save(const char *path);
int  mdg_tune_load(const char *path);
/* the tiling a box of these parameters would be created with */
int  mdg_box_tiling(const mdg_box_params *p, int precision, mdg_tiling *out);

#ifdef __cplusplus
}
#endif
#endif
