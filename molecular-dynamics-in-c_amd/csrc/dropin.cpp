// dropin.cpp -- mdc's force-path routines with mdc's signatures
// (include/mdc_abi.h), computing on the GPU over mdc's global objects.
//
// Per call the host arrays are authoritative: the routine uploads the
// arrays it reads, runs the kernels, and writes back the arrays and
// scalars mdc's routine writes.  The device box is rebuilt when the system
// parameters or the pair table change.  Real space runs as an all-pairs
// loop visiting partners in ascending index up to N = 8192, which
// reproduces mdc's force accumulation order bit for bit.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "engine.h"
#include "host_setup.h"
#include "mdc_abi.h"
#include "mdgpu.h"

using namespace mdg;

// The globals of mdc/program/md.h:197-202 and potential.h:28.  An mdc
// executable defines its own (md.h defines them in the header); ELF
// interposition then binds these references to the executable's objects.
system_property sys;
ion_property ion;
calc_control ctl;
huggins_mayer_potential_set hm;
msd_calculation_variables msd;

namespace {

struct DropIn {
  Box* box = nullptr;
  BoxSet* set = nullptr;
  hipStream_t stream = nullptr;
  int status = MDG_OK;
  bool reported = false;
  // what the device box was built from
  int N = 0, hmv = 0, kinds = 0, td = 0;
  double L[3] = {0, 0, 0}, a1 = 0, a3 = 0, p2a2 = 0, radius2 = 0, dt = 0, dt2 = 0;
  double Z[4] = {0, 0, 0, 0}, m[4] = {0, 0, 0, 0};
  int lookup[16] = {0};
  std::vector<double> pe, fe;
  std::vector<int> species;
  std::vector<double> buf;
};

DropIn g;

// The error mode: mdc's convention is exit(EXIT_FAILURE) on a fatal error
// (files.c:9-24), and an unmodified mdc main.c checks no status, so the
// default is to exit; MDG_DROPIN_STATUS (or MDGPU_DROPIN_ERRORS=status in
// the environment) keeps the process alive with a sticky status instead.
int error_mode = -1;  // -1: not chosen yet (the environment decides)

int current_error_mode() {
  if (error_mode < 0) {
    const char* e = getenv("MDGPU_DROPIN_ERRORS");
    error_mode = (e && strcmp(e, "status") == 0) ? MDG_DROPIN_STATUS : MDG_DROPIN_EXIT;
  }
  return error_mode;
}

void fail(int code, const char* what) {
  if (g.status == MDG_OK) g.status = code;
  if (!g.reported) {
    fprintf(stderr, "libmdgpu drop-in: %s failed: %s\n", what, mdg_error_string(code));
    g.reported = true;
  }
  if (current_error_mode() == MDG_DROPIN_EXIT) {
    fflush(stdout);
    exit(EXIT_FAILURE);
  }
}

void release() {
  delete g.set;
  g.set = nullptr;
  box_destroy(g.box);
  g.box = nullptr;
}

bool same_params() {
  if (!g.box) return false;
  const int kinds = ctl.kinds_of_ions;
  const int td = sys.table_division;
  if (g.N != sys.N || g.hmv != sys.hm || g.kinds != kinds || g.td != td) return false;
  if (g.L[0] != sys.Lx || g.L[1] != sys.Ly || g.L[2] != sys.Lz || g.a1 != sys.a1 ||
      g.a3 != sys.a3 || g.p2a2 != sys.p2a2 || g.radius2 != sys.radius2 || g.dt != sys.dt ||
      g.dt2 != sys.dt2)
    return false;
  for (int k = 0; k < kinds; ++k)
    if (g.Z[k] != ion.Z[k] || g.m[k] != ion.m[k]) return false;
  for (int k = 0; k < kinds * kinds; ++k)
    if (g.lookup[k] != sys.lookup[k]) return false;
  const size_t nt = (size_t)(kinds * (kinds + 1) / 2) * (size_t)(td + 1);
  if (g.pe.size() != nt || memcmp(g.pe.data(), sys.pe1r1, nt * sizeof(double)) != 0 ||
      memcmp(g.fe.data(), sys.fe1r1, nt * sizeof(double)) != 0)
    return false;
  return true;
}

// Make the device box match the globals; upload species every call.
bool ensure(const char* who) {
  if (g.status != MDG_OK) return false;
  if (!g.stream) {
    int n = 0;
    if (mdg_device_count(&n) != MDG_OK) {
      fail(MDG_E_NODEVICE, who);
      return false;
    }
    if (hipStreamCreateWithFlags(&g.stream, hipStreamNonBlocking) != hipSuccess) {
      fail(MDG_E_HIP, who);
      return false;
    }
  }
  const int kinds = ctl.kinds_of_ions;
  if (sys.N <= 0 || kinds < 1 || kinds > 4 || !sys.rx || !sys.pe1r1 || !sys.fe1r1 ||
      !sys.lookup || !ion.Z || !ion.m || !sys.ion) {
    fail(MDG_E_ARG, who);
    return false;
  }
  if (!same_params()) {
    release();
    g.N = sys.N; g.hmv = sys.hm; g.kinds = kinds; g.td = sys.table_division;
    g.L[0] = sys.Lx; g.L[1] = sys.Ly; g.L[2] = sys.Lz;
    g.a1 = sys.a1; g.a3 = sys.a3; g.p2a2 = sys.p2a2; g.radius2 = sys.radius2;
    g.dt = sys.dt; g.dt2 = sys.dt2;
    for (int k = 0; k < kinds; ++k) { g.Z[k] = ion.Z[k]; g.m[k] = ion.m[k]; }
    for (int k = 0; k < kinds * kinds; ++k) g.lookup[k] = sys.lookup[k];
    const size_t nt = (size_t)(kinds * (kinds + 1) / 2) * (size_t)(g.td + 1);
    g.pe.assign(sys.pe1r1, sys.pe1r1 + nt);
    g.fe.assign(sys.fe1r1, sys.fe1r1 + nt);
    BoxSpec s;
    s.N = sys.N;
    s.L[0] = sys.Lx; s.L[1] = sys.Ly; s.L[2] = sys.Lz;
    s.alpha = sys.a1; s.a3 = sys.a3; s.p2a2 = sys.p2a2;
    s.hm = sys.hm;
    s.rc = sqrt(sys.radius2);
    s.rc2 = sys.radius2;
    s.dt = sys.dt; s.dt2 = sys.dt2;
    s.real_mode = MDG_TABLE;
    s.cells = sys.N <= 8192 ? MDG_ALL_PAIRS : MDG_CELLS_AUTO;
    s.ordered = true;
    s.kinds = kinds;
    for (int k = 0; k < kinds; ++k) { s.Z[k] = ion.Z[k]; s.m[k] = ion.m[k]; }
    for (int k = 0; k < 16; ++k) s.lookup[k] = k < kinds * kinds ? sys.lookup[k] : 0;
    s.pe = g.pe.data();
    s.fe = g.fe.data();
    s.table_nb = g.td + 1;
    s.table_ntypes = kinds * (kinds + 1) / 2;
    std::string why;
    int rc = box_create(s, MDG_FP64, &g.box, &why);
    if (rc != MDG_OK) {
      if (!why.empty()) fprintf(stderr, "libmdgpu drop-in: %s\n", why.c_str());
      fail(rc, who);
      return false;
    }
    g.set = new BoxSet(MDG_FP64, MDG_TABLE);
    rc = g.set->set_boxes({g.box});
    if (rc != MDG_OK) {
      fail(rc, who);
      release();
      return false;
    }
  }
  g.species.resize(sys.N);
  for (int i = 0; i < sys.N; ++i) g.species[i] = sys.ion[i];
  int rc = box_upload_state(g.box, nullptr, nullptr, g.species.data(), g.stream);
  if (rc == MDG_OK) rc = g.set->sync_desc(g.stream);
  if (rc != MDG_OK) {
    fail(rc, who);
    return false;
  }
  return true;
}

// host <-> device copies of mdc's SoA arrays
enum { kR = 1, kV = 2, kF = 4, kF0 = 8 };

bool put(int what, const char* who) {
  const int N = sys.N;
  const BoxDesc& D = g.box->desc;
  const double* h[4][3] = {{sys.rx, sys.ry, sys.rz}, {sys.vx, sys.vy, sys.vz},
                           {sys.fx, sys.fy, sys.fz}, {sys.fx0, sys.fy0, sys.fz0}};
  double* d[4][3] = {{D.rx, D.ry, D.rz}, {D.vx, D.vy, D.vz}, {D.fx, D.fy, D.fz},
                     {D.fx0, D.fy0, D.fz0}};
  for (int a = 0; a < 4; ++a) {
    if (!(what & (1 << a))) continue;
    for (int c = 0; c < 3; ++c)
      if (hipMemcpyAsync(d[a][c], h[a][c], sizeof(double) * N, hipMemcpyHostToDevice,
                         g.stream) != hipSuccess) {
        fail(MDG_E_HIP, who);
        return false;
      }
  }
  return true;
}

bool get(int what, const char* who) {
  const int N = sys.N;
  const BoxDesc& D = g.box->desc;
  double* h[4][3] = {{sys.rx, sys.ry, sys.rz}, {sys.vx, sys.vy, sys.vz},
                     {sys.fx, sys.fy, sys.fz}, {sys.fx0, sys.fy0, sys.fz0}};
  const double* d[4][3] = {{D.rx, D.ry, D.rz}, {D.vx, D.vy, D.vz}, {D.fx, D.fy, D.fz},
                           {D.fx0, D.fy0, D.fz0}};
  for (int a = 0; a < 4; ++a) {
    if (!(what & (1 << a))) continue;
    for (int c = 0; c < 3; ++c)
      if (hipMemcpyAsync(h[a][c], d[a][c], sizeof(double) * N, hipMemcpyDeviceToHost,
                         g.stream) != hipSuccess) {
        fail(MDG_E_HIP, who);
        return false;
      }
  }
  if (hipStreamSynchronize(g.stream) != hipSuccess) {
    fail(MDG_E_HIP, who);
    return false;
  }
  return true;
}

bool scalars(double* S, const char* who) {
  int rc = box_read_scalars(g.box, S, g.stream);
  int e = 0;
  if (rc == MDG_OK) rc = box_read_error(g.box, &e, g.stream);
  if (rc == MDG_OK && e) rc = (e & kErrState) ? MDG_E_STATE : MDG_E_RANGE;
  if (rc != MDG_OK) {
    fail(rc, who);
    return false;
  }
  return true;
}

bool ewald_on() { return !(sys.a1 == 0.0 || sys.a1 > 10.0); }  // rcprcl.c:23

}  // namespace

extern "C" {

int mdg_dropin_set_error_mode(int mode) {
  if (mode != MDG_DROPIN_EXIT && mode != MDG_DROPIN_STATUS) return MDG_E_ARG;
  error_mode = mode;
  return MDG_OK;
}
int mdg_dropin_status(void) { return g.status; }
void mdg_dropin_clear_status(void) {
  g.status = MDG_OK;
  g.reported = false;
}
void mdg_dropin_release(void) { release(); }

// rv.c:12-18: clear_foc, real_space, reciprocal_space, reciprocal_space3
void calc_foc(void) {
  if (!ensure("calc_foc") || !put(kR | kF, "calc_foc")) return;
  g.set->enqueue_step(g.stream, false);
  double S[kScalCount];
  if (!get(kF | kF0, "calc_foc") || !scalars(S, "calc_foc")) return;
  sys.pot = S[kScalPot];
  sys.virX = S[kScalVirX]; sys.virY = S[kScalVirY]; sys.virZ = S[kScalVirZ];
}

// init.c:179-190
void clear_foc(void) {
  if (!ensure("clear_foc") || !put(kF, "clear_foc")) return;
  g.set->enqueue_clear_foc(g.stream);
  get(kF | kF0, "clear_foc");
}

// real.c:15-84: adds to f, resets pot and the virial
void real_space(void) {
  if (!ensure("real_space") || !put(kR | kF, "real_space")) return;
  g.set->enqueue_forces_parts(g.stream, true, 1, false);
  g.set->enqueue_scalars(g.stream, false);
  double S[kScalCount];
  if (!get(kF, "real_space") || !scalars(S, "real_space")) return;
  sys.pot = S[kScalPotReal];
  sys.virX = S[kScalVirX]; sys.virY = S[kScalVirY]; sys.virZ = S[kScalVirZ];
}

// rcprcl.c:11-97: adds reciprocal forces and energy
void reciprocal_space(void) {
  if (!ewald_on()) return;
  if (!ensure("reciprocal_space") || !put(kR | kF, "reciprocal_space")) return;
  g.set->enqueue_forces_parts(g.stream, false, 1, true);
  g.set->enqueue_scalars(g.stream, false);
  double S[kScalCount];
  if (!get(kF, "reciprocal_space") || !scalars(S, "reciprocal_space")) return;
  sys.pot += S[kScalPotRecip];
}

// rcprcl.c:108-133: pot += -a3 sum Z^2
void reciprocal_space3(void) {
  if (!ewald_on()) return;
  if (!ensure("reciprocal_space3")) return;
  sys.pot += g.box->desc.self_pot;
}

// rv.c:26-31
void next_rv_verlet(void) {
  if (!ensure("next_rv_verlet") || !put(kR | kV | kF, "next_rv_verlet")) return;
  g.set->enqueue_step(g.stream, true);
  double S[kScalCount];
  if (!get(kR | kV | kF | kF0, "next_rv_verlet") || !scalars(S, "next_rv_verlet")) return;
  sys.pot = S[kScalPot];
  sys.virX = S[kScalVirX]; sys.virY = S[kScalVirY]; sys.virZ = S[kScalVirZ];
}

// rv.c:33-44
void next_v(void) {
  if (!ensure("next_v") || !put(kV | kF | kF0, "next_v")) return;
  g.set->enqueue_next_v(g.stream);
  get(kV, "next_v");
}

// rv.c:46-66
void next_r(void) {
  if (!ensure("next_r") || !put(kR | kV | kF, "next_r")) return;
  g.set->enqueue_next_r(g.stream);
  get(kR, "next_r");
}

// rv.c:68-86
void calc_kin(void) {
  if (!ensure("calc_kin") || !put(kV, "calc_kin")) return;
  g.set->enqueue_scalars(g.stream, false);
  double S[kScalCount];
  if (!scalars(S, "calc_kin")) return;
  sys.kinX = S[kScalKinX]; sys.kinY = S[kScalKinY]; sys.kinZ = S[kScalKinZ];
  sys.kin = sys.kinX + sys.kinY + sys.kinZ;
}

// rv.c:88-116
void moment_correction(void) {
  if (!ensure("moment_correction") || !put(kV, "moment_correction")) return;
  g.set->enqueue_moment_correction(g.stream);
  get(kV, "moment_correction");
}

// msd.c:20-81.  The device box keeps no MSD state between calls: msd's
// arrays are uploaded, updated on the device and written back.
void calc_msd(void) {
  if (!ensure("calc_msd")) return;
  const int N = sys.N;
  double* host[6] = {msd.rx_old, msd.ry_old, msd.rz_old, msd.dx, msd.dy, msd.dz};
  for (double* p : host)
    if (!p) return fail(MDG_E_ARG, "calc_msd");
  if (!msd.value || !msd.number_of_ion) return fail(MDG_E_ARG, "calc_msd");
  BoxDesc& D = g.box->desc;
  if (!D.msd_on) {
    D.msd_on = 1;
    g.box->desc_dirty = true;
  }
  if (g.set->sync_desc(g.stream) != MDG_OK || !put(kR, "calc_msd")) return;
  double* dev[6] = {D.msd_old, D.msd_old + N, D.msd_old + 2 * N,
                    D.msd_d,   D.msd_d + N,   D.msd_d + 2 * N};
  for (int k = 0; k < 6; ++k)
    if (hipMemcpyAsync(dev[k], host[k], sizeof(double) * N, hipMemcpyHostToDevice,
                       g.stream) != hipSuccess)
      return fail(MDG_E_HIP, "calc_msd");
  // bit 0: MSD only; bit 3: the origin is recorded at step 1 only, as mdc
  g.set->enqueue_observe(g.stream, sys.step, 1 | 8);
  for (int k = 0; k < 6; ++k)
    if (hipMemcpyAsync(host[k], dev[k], sizeof(double) * N, hipMemcpyDeviceToHost,
                       g.stream) != hipSuccess)
      return fail(MDG_E_HIP, "calc_msd");
  double O[kObsLen];
  if (box_read_observe(g.box, O, nullptr, nullptr, g.stream) != MDG_OK)
    return fail(MDG_E_HIP, "calc_msd");
  if (sys.step == 1) return;  // origin recorded; nothing else changes
  msd.Lx2 = sys.Lx / 2.0;
  msd.Ly2 = sys.Ly / 2.0;
  msd.Lz2 = sys.Lz / 2.0;
  for (int k = 0; k < ctl.kinds_of_ions && k < 4; ++k) {
    msd.value[k] = O[kObsMsd + k];
    msd.number_of_ion[k] = (int)O[kObsMsdIons + k];
  }
}

// pt.c:85-112.  The averaging over steps and the scale factor are scalar
// work on the host, kept in statics as mdc keeps them; the velocities are
// scaled on the device.
void control_temp(int d_step, double temp) {
  static double oldk = 0.0;
  static short count = 0;
  if (d_step <= 0) return fail(MDG_E_ARG, "control_temp");
  count++;
  oldk += sys.kin;
  if (sys.step % d_step != 0) return;
  const double averaged_K = oldk / (double)count;
  oldk = 0.0;
  count = 0;
  const double tempK = temp / sys.e2t;
  const double scale = sqrt(tempK / averaged_K);
  if (!ensure("control_temp") || !put(kV, "control_temp")) return;
  g.set->enqueue_scale_v(g.stream, scale);
  if (!get(kV, "control_temp")) return;
  printf("---------------------------------------------------------------\n");
  printf(">> Temperature scaling\n");
}

// pt.c:8-21 (six scalar operations: evaluated on the host)
void calc_press(void) {
  const double vol = sys.Lx * sys.Ly * sys.Lz;
  sys.presX = (2.0 * sys.kinX + sys.virX) / vol;
  sys.presY = (2.0 * sys.kinY + sys.virY) / vol;
  sys.presZ = (2.0 * sys.kinZ + sys.virZ) / vol;
  sys.pres = (sys.presX + sys.presY + sys.presZ) / 3.0;
}

// control.c:183-248 for the two-species BMH plugin (NaCl form of
// potential.h): fills sys.pe1r1 / sys.fe1r1 / sys.lookup on the host.
void mk_table(void) {
  if (!sys.pe1r1 || !sys.fe1r1 || !sys.lookup || !ion.Z || ctl.kinds_of_ions != 2) {
    fail(MDG_E_ARG, "mk_table");
    return;
  }
  mdg_potential p;
  memset(&p, 0, sizeof(p));
  p.Z[0] = ion.Z[0]; p.Z[1] = ion.Z[1];
  p.b = hm.b;
  p.rho = hm.rho;
  p.c[0] = hm.c_na_na; p.c[1] = hm.c_na_cl; p.c[2] = hm.c_cl_cl;
  p.sigma[0] = hm.sigma_na + hm.sigma_na;
  p.sigma[1] = hm.sigma_na + hm.sigma_cl;
  p.sigma[2] = hm.sigma_cl + hm.sigma_cl;
  p.C[0] = hm.C_na_na; p.C[1] = hm.C_na_cl; p.C[2] = hm.C_cl_cl;
  p.D[0] = hm.D_na_na; p.D[1] = hm.D_na_cl; p.D[2] = hm.D_cl_cl;
  host_mk_table(p, sys.a1, sys.a3, sys.table_division + 1, sys.pe1r1, sys.fe1r1);
  sys.lookup[0] = 0; sys.lookup[1] = 1; sys.lookup[2] = 1; sys.lookup[3] = 2;
}

// ext.c:73-104
double erfcc(double x) { return host_erfcc(x); }

}  // extern "C"
