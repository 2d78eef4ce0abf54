// api.cpp -- the handle-based C ABI of libmdgpu (include/mdgpu.h).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <vector>

#include "batch.h"
#include "engine.h"
#include "host_setup.h"
#include "launch.h"
#include "mdgpu.h"

using namespace mdg;

// mdc's constants (mdc/program/md.h:8-16)
static const double kSqrtPi = 1.77245385090551602731;
static const double kPiPi = 9.869604401089358619;
static const double kSecd2Fs = 4.074972637944947;


namespace mdg {

// BoxSpec of a rock-salt box described by public parameters, with mdc's
// unit conversion (mdc/program/init.c:25-80).  The table, when needed, is
// built into `pe`/`fe`.
int spec_from_params(const mdg_box_params& p, BoxSpec* s, std::vector<double>* pe,
                     std::vector<double>* fe) {
  if (p.nx <= 0 || p.ny <= 0 || p.nz <= 0 || !(p.ax > 0) || !(p.ay > 0) || !(p.az > 0) ||
      !(p.rc > 0) || !(p.dt_fs > 0) || p.hm < 0)
    return MDG_E_ARG;
  if (p.real_mode != MDG_TABLE && p.real_mode != MDG_ANALYTIC) return MDG_E_ARG;
  *s = BoxSpec();
  s->N = 8 * p.nx * p.ny * p.nz;
  s->L[0] = p.ax * (double)p.nx;
  s->L[1] = p.ay * (double)p.ny;
  s->L[2] = p.az * (double)p.nz;
  s->alpha = p.alpha;
  if (p.alpha > 0) {
    s->p2a2 = kPiPi / (p.alpha * p.alpha);
    s->a3 = p.alpha / kSqrtPi;
  }
  s->hm = p.hm;
  s->rc = p.rc;
  s->rc2 = p.rc * p.rc;
  s->dt = p.dt_fs / kSecd2Fs;
  s->dt2 = s->dt * s->dt;
  s->real_mode = p.real_mode;
  s->cells = p.cells;
  s->kinds = 2;
  for (int k = 0; k < 2; ++k) {
    s->Z[k] = p.pot.Z[k];
    s->m[k] = p.pot.m[k];
  }
  s->pot = p.pot;
  if (p.real_mode == MDG_TABLE) {
    if (p.rc <= 0.5) return MDG_E_ARG;
    const int td = (int)((p.rc - 0.5) * 1000.0) + 1;  // init.c:65
    const int nb = td + 1;
    pe->assign(3 * (size_t)nb, 0.0);
    fe->assign(3 * (size_t)nb, 0.0);
    host_mk_table(p.pot, p.alpha, s->a3, nb, pe->data(), fe->data());
    s->pe = pe->data();
    s->fe = fe->data();
    s->table_nb = nb;
    s->table_ntypes = 3;
  }
  return MDG_OK;
}

}  // namespace mdg

extern "C" {

const char* mdg_error_string(int code) {
  switch (code) {
    case MDG_OK: return "ok";
    case MDG_E_NODEVICE: return "no HIP device";
    case MDG_E_HIP: return "HIP runtime error";
    case MDG_E_ARG: return "invalid argument";
    case MDG_E_NOMEM: return "device allocation failed";
    case MDG_E_BUDGET: return "exceeds HBM budget";
    case MDG_E_STATE: return "non-finite or out-of-box state";
    case MDG_E_RANGE: return "pair inside table start or cutoff/box mismatch";
    case MDG_E_TIMEOUT: return "bounded device wait expired";
    default: return "unknown error";
  }
}

int mdg_device_count(int* n) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) {
    (void)hipGetLastError();
    if (n) *n = 0;
    return MDG_E_NODEVICE;
  }
  if (n) *n = c;
  return c > 0 ? MDG_OK : MDG_E_NODEVICE;
}

int mdg_copy_probe(size_t bytes, int reps, double* ms_avg) {
  if (bytes < 16 || reps < 1 || !ms_avg) return MDG_E_ARG;
  int ndev = 0;
  if (mdg_device_count(&ndev) != MDG_OK) return MDG_E_NODEVICE;
  const size_t n16 = bytes / 16;
  void *src = nullptr, *dst = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = MDG_OK;
  if (hipMalloc(&src, n16 * 16) != hipSuccess || hipMalloc(&dst, n16 * 16) != hipSuccess) {
    (void)hipGetLastError();
    rc = MDG_E_NOMEM;
  }
  if (rc == MDG_OK &&
      (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
       hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
       hipMemsetAsync(src, 0, n16 * 16, s) != hipSuccess))
    rc = MDG_E_HIP;
  if (rc == MDG_OK) {
    launch_copy_probe(src, dst, n16, s);  // warm-up
    (void)hipEventRecord(e0, s);
    for (int k = 0; k < reps; ++k) launch_copy_probe(src, dst, n16, s);
    (void)hipEventRecord(e1, s);
    float ms = 0.f;
    if (hipStreamSynchronize(s) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      rc = MDG_E_HIP;
    *ms_avg = ms / reps;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (src) (void)hipFree(src);
  if (dst) (void)hipFree(dst);
  return rc;
}

int mdg_batch_destroy(mdg_batch* b) {
  if (!b) return MDG_OK;
  delete b->set;
  for (Box* x : b->boxes) box_destroy(x);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return MDG_OK;
}

int mdg_batch_create(int nbox, const mdg_box_params* params, int precision, mdg_batch** out) {
  if (!out || nbox <= 0 || !params) return MDG_E_ARG;
  if (precision != MDG_FP64 && precision != MDG_MIXED) return MDG_E_ARG;
  *out = nullptr;
  int ndev = 0;
  if (mdg_device_count(&ndev) != MDG_OK) {
    fprintf(stderr, "libmdgpu: no HIP device available\n");
    return MDG_E_NODEVICE;
  }
  const int rmode = params[0].real_mode;
  for (int k = 1; k < nbox; ++k)
    if (params[k].real_mode != rmode) return MDG_E_ARG;  // one pair kernel per launch
  mdg_batch* b = new mdg_batch();
  b->precision = precision;
  b->real_mode = rmode;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return MDG_E_HIP;
  }
  for (int k = 0; k < nbox; ++k) {
    BoxSpec spec;
    std::vector<double> pe, fe;
    int rc = spec_from_params(params[k], &spec, &pe, &fe);
    Box* box = nullptr;
    std::string why;
    if (rc == MDG_OK) rc = box_create(spec, precision, &box, &why);
    if (rc != MDG_OK) {
      if (!why.empty()) fprintf(stderr, "libmdgpu: box %d: %s\n", k, why.c_str());
      mdg_batch_destroy(b);
      return rc;
    }
    b->boxes.push_back(box);
    b->cells.push_back({params[k].nx, params[k].ny, params[k].nz});
    b->mol_unit.push_back(2);  // rock salt: two ions per formula unit
    // start from mdc's lattice at rest
    const int N = spec.N;
    std::vector<double> r(3 * (size_t)N);
    std::vector<int> sp(N);
    mdg_rocksalt(params[k].nx, params[k].ny, params[k].nz, params[k].ax, params[k].ay,
                 params[k].az, r.data(), sp.data());
    rc = box_upload_state(box, r.data(), nullptr, sp.data(), b->stream);
    if (rc != MDG_OK) {
      mdg_batch_destroy(b);
      return rc;
    }
  }
  b->set = new BoxSet(precision, rmode);
  int rc = b->set->set_boxes(b->boxes);
  if (rc != MDG_OK) {
    mdg_batch_destroy(b);
    return rc;
  }
  *out = b;
  return MDG_OK;
}

int mdg_batch_create_tabulated(int nbox, const mdg_table_box* tb, int precision,
                               mdg_batch** out) {
  if (!out || nbox <= 0 || !tb) return MDG_E_ARG;
  if (precision != MDG_FP64 && precision != MDG_MIXED) return MDG_E_ARG;
  *out = nullptr;
  int ndev = 0;
  if (mdg_device_count(&ndev) != MDG_OK) {
    fprintf(stderr, "libmdgpu: no HIP device available\n");
    return MDG_E_NODEVICE;
  }
  std::vector<BoxSpec> specs(nbox);
  for (int k = 0; k < nbox; ++k) {
    const mdg_table_box& t = tb[k];
    const int nt = t.kinds * (t.kinds + 1) / 2;
    if (t.n_ions <= 0 || t.kinds < 1 || t.kinds > 4 || t.ntypes != nt || !t.pe || !t.fe ||
        !(t.L[0] > 0) || !(t.L[1] > 0) || !(t.L[2] > 0) || !(t.rc > 0.5) || !(t.dt_fs > 0) ||
        t.hm < 0 || t.cells[0] <= 0 || t.cells[1] <= 0 || t.cells[2] <= 0 ||
        t.natoms_in_mol_unit <= 0 || t.table_division < (int)((t.rc - 0.5) * 1000.0))
      return MDG_E_ARG;
    for (int q = 0; q < t.kinds * t.kinds; ++q)
      if (t.lookup[q] < 0 || t.lookup[q] >= nt) return MDG_E_ARG;
    BoxSpec& s = specs[k];
    s.N = t.n_ions;
    for (int d = 0; d < 3; ++d) s.L[d] = t.L[d];
    s.alpha = t.alpha;
    if (t.alpha > 0) {  // init.c:38-41
      s.p2a2 = kPiPi / (t.alpha * t.alpha);
      s.a3 = t.alpha / kSqrtPi;
    }
    s.hm = t.hm;
    s.rc = t.rc;
    s.rc2 = t.rc * t.rc;
    s.dt = t.dt_fs / kSecd2Fs;
    s.dt2 = s.dt * s.dt;
    s.real_mode = MDG_TABLE;
    s.cells = t.cells_mode;
    s.kinds = t.kinds;
    for (int q = 0; q < t.kinds; ++q) {
      s.Z[q] = t.Z[q];
      s.m[q] = t.m[q];
    }
    for (int q = 0; q < 16; ++q) s.lookup[q] = q < t.kinds * t.kinds ? t.lookup[q] : 0;
    s.pe = t.pe;
    s.fe = t.fe;
    s.table_nb = t.table_division + 1;
    s.table_ntypes = nt;
  }
  mdg_batch* b = new mdg_batch();
  b->precision = precision;
  b->real_mode = MDG_TABLE;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return MDG_E_HIP;
  }
  for (int k = 0; k < nbox; ++k) {
    Box* box = nullptr;
    std::string why;
    const int rc = box_create(specs[k], precision, &box, &why);  // copies the tables
    if (rc != MDG_OK) {
      if (!why.empty()) fprintf(stderr, "libmdgpu: box %d: %s\n", k, why.c_str());
      mdg_batch_destroy(b);
      return rc;
    }
    b->boxes.push_back(box);
    b->cells.push_back({tb[k].cells[0], tb[k].cells[1], tb[k].cells[2]});
    b->mol_unit.push_back(tb[k].natoms_in_mol_unit);
  }
  b->set = new BoxSet(precision, MDG_TABLE);
  const int rc = b->set->set_boxes(b->boxes);
  if (rc != MDG_OK) {
    mdg_batch_destroy(b);
    return rc;
  }
  *out = b;
  return MDG_OK;
}

int mdg_batch_n_ions(const mdg_batch* b, int box) {
  if (!b || box < 0 || box >= (int)b->boxes.size()) return -1;
  return b->boxes[box]->desc.N;
}

size_t mdg_batch_box_bytes(const mdg_batch* b, int box) {
  if (!b || box < 0 || box >= (int)b->boxes.size()) return 0;
  return b->boxes[box]->bytes;
}

int mdg_batch_set_state(mdg_batch* b, int box, const double* r, const double* v,
                        const int* species) {
  if (!b || box < 0 || box >= (int)b->boxes.size()) return MDG_E_ARG;
  return box_upload_state(b->boxes[box], r, v, species, b->stream);
}

int mdg_batch_get_state(mdg_batch* b, int box, double* r, double* v, double* f, double* f0) {
  if (!b || box < 0 || box >= (int)b->boxes.size()) return MDG_E_ARG;
  return box_download_state(b->boxes[box], r, v, f, f0, b->stream);
}

int mdg_batch_forces(mdg_batch* b) {
  if (!b) return MDG_E_ARG;
  int rc = b->set->sync_desc(b->stream);
  if (rc != MDG_OK) return rc;
  b->set->enqueue_step(b->stream, false);
  MDG_HIP(hipGetLastError());
  MDG_HIP(hipStreamSynchronize(b->stream));
  return MDG_OK;
}

int mdg_batch_force_terms(mdg_batch* b, int terms) {
  if (!b || terms < 1 || terms > 3) return MDG_E_ARG;
  int rc = b->set->sync_desc(b->stream);
  if (rc != MDG_OK) return rc;
  b->set->enqueue_force_terms(b->stream, terms);
  MDG_HIP(hipGetLastError());
  MDG_HIP(hipStreamSynchronize(b->stream));
  return MDG_OK;
}

int mdg_batch_step(mdg_batch* b, int nsteps) {
  if (!b || nsteps < 0) return MDG_E_ARG;
  return b->set->step(nsteps, b->stream);
}

int mdg_batch_prepare(mdg_batch* b) {
  if (!b) return MDG_E_ARG;
  return b->set->prepare(b->stream);
}

int mdg_batch_tune_schedule(mdg_batch* b, int nsteps, double* share_out) {
  if (!b || nsteps <= 0) return MDG_E_ARG;
  for (Box* x : b->boxes)
    if (!x->dmem) return MDG_E_ARG;
  // every box's device memory is saved and put back, so the batch's state
  // (and so its trajectory) is the same as if nothing had run
  size_t total = 0;
  for (Box* x : b->boxes) total += x->bytes;
  void* save = nullptr;
  if (hipMalloc(&save, std::max<size_t>(total, 1)) != hipSuccess) {
    (void)hipGetLastError();
    return MDG_E_NOMEM;
  }
  auto copy_all = [&](bool back) {
    size_t off = 0;
    for (Box* x : b->boxes) {
      char* keep = (char*)save + off;
      if (hipMemcpyAsync(back ? x->dmem : keep, back ? keep : x->dmem, x->bytes,
                         hipMemcpyDeviceToDevice, b->stream) != hipSuccess)
        return MDG_E_HIP;
      off += x->bytes;
    }
    return hipStreamSynchronize(b->stream) == hipSuccess ? MDG_OK : MDG_E_HIP;
  };
  int rc = copy_all(false);
  // two passes over the candidates, each candidate's faster time counting:
  // a single window of a few tens of steps can be disturbed
  const double cand[] = {0.5, 0.3, 0.35, 0.4, 0.45, 0.55, 0.6, 0.65, 0.7};
  constexpr int kCand = sizeof(cand) / sizeof(cand[0]);
  double cand_ms[kCand];
  for (double& m : cand_ms) m = 1e30;
  for (int pass = 0; pass < 2 && rc == MDG_OK; ++pass)
    for (int ci = 0; ci < kCand; ++ci) {
      if (rc != MDG_OK) break;
      b->set->set_group_share(cand[ci]);
      rc = b->set->set_boxes(b->boxes);
      if (rc == MDG_OK) rc = b->set->prepare(b->stream);
      if (rc == MDG_OK) rc = b->set->step(8, b->stream);  // warm
      if (rc == MDG_OK && hipStreamSynchronize(b->stream) != hipSuccess) rc = MDG_E_HIP;
      if (rc != MDG_OK) break;
      const auto t0 = std::chrono::steady_clock::now();
      rc = b->set->step(nsteps, b->stream);
      if (rc == MDG_OK && hipStreamSynchronize(b->stream) != hipSuccess) rc = MDG_E_HIP;
      const double ms =
          std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      if (rc == MDG_OK) cand_ms[ci] = std::min(cand_ms[ci], ms);
    }
  double best = b->set->group_share(), best_ms = 1e30;
  for (int ci = 0; ci < kCand; ++ci)
    if (cand_ms[ci] < best_ms) {
      best_ms = cand_ms[ci];
      best = cand[ci];
    }
  if (rc == MDG_OK) rc = copy_all(true);
  (void)hipFree(save);
  b->set->set_group_share(best);
  const int rc2 = b->set->set_boxes(b->boxes);
  if (rc == MDG_OK) rc = rc2;
  if (rc == MDG_OK) rc = b->set->prepare(b->stream);
  if (rc == MDG_OK && share_out) *share_out = best;
  return rc;
}

int mdg_batch_set_observe(mdg_batch* b, int box, const mdg_observe* o) {
  if (!b || !o || box < 0 || box >= (int)b->boxes.size()) return MDG_E_ARG;
  ObserveOpts x;
  x.msd = o->msd;
  x.t_control_step = o->t_control_step;
  x.temp_K = o->temp_K;
  x.t_control_until = o->t_control_until;
  x.moment_every = o->moment_every;
  int rc = box_set_observe(b->boxes[box], x, b->stream);
  if (rc != MDG_OK) return rc;
  return b->set->set_boxes(b->boxes);  // the observe list changes
}

int mdg_batch_observed(mdg_batch* b, int box, double* msd, long long* scalings) {
  if (!b || box < 0 || box >= (int)b->boxes.size()) return MDG_E_ARG;
  double O[kObsLen];
  int rc = box_read_observe(b->boxes[box], O, nullptr, nullptr, b->stream);
  if (rc != MDG_OK) return rc;
  if (msd)
    for (int k = 0; k < 4; ++k) msd[k] = O[kObsMsd + k];
  if (scalings) *scalings = (long long)O[kObsScalings];
  return MDG_OK;
}

int mdg_batch_msd_state(mdg_batch* b, int box, double* old, double* d) {
  if (!b || box < 0 || box >= (int)b->boxes.size()) return MDG_E_ARG;
  return box_read_observe(b->boxes[box], nullptr, old, d, b->stream);
}

int mdg_batch_sync(mdg_batch* b) {
  if (!b) return MDG_E_ARG;
  MDG_HIP(hipStreamSynchronize(b->stream));
  return MDG_OK;
}

static int err_code(int e) {
  if (e & kErrState) return MDG_E_STATE;
  if (e & kErrRange) return MDG_E_RANGE;
  return MDG_OK;
}

int mdg_batch_scalars(mdg_batch* b, int box, mdg_scalars* out) {
  if (!b || !out || box < 0 || box >= (int)b->boxes.size()) return MDG_E_ARG;
  double S[kScalCount];
  int rc = box_read_scalars(b->boxes[box], S, b->stream);
  if (rc != MDG_OK) return rc;
  int e = 0;
  rc = box_read_error(b->boxes[box], &e, b->stream);
  if (rc != MDG_OK) return rc;
  fill_scalars(S, out);
  return err_code(e);
}

int mdg_batch_profile(mdg_batch* b, int nsteps, int cap, char names[][32], double* ms_step,
                      int* launches, int* n) {
  if (!b || nsteps == 0 || !names || !ms_step || !launches || !n) return MDG_E_ARG;
  std::vector<std::string> nm;
  std::vector<double> t;
  std::vector<int> c;
  // nsteps < 0: -nsteps steps with every kernel alone over all boxes
  const bool as_step = nsteps > 0;
  int rc = b->set->profile(as_step ? nsteps : -nsteps, b->stream, as_step, &nm, &t, &c);
  if (rc != MDG_OK) return rc;
  int k = 0;
  for (; k < (int)nm.size() && k < cap; ++k) {
    strncpy(names[k], nm[k].c_str(), 31);
    names[k][31] = 0;
    ms_step[k] = t[k];
    launches[k] = c[k];
  }
  *n = k;
  return MDG_OK;
}

int mdg_batch_work(mdg_batch* b, int box, long long* pairs, int* nk, int* n_ions) {
  if (!b || box < 0 || box >= (int)b->boxes.size()) return MDG_E_ARG;
  double S[kScalCount];
  int rc = box_read_scalars(b->boxes[box], S, b->stream);
  if (rc != MDG_OK) return rc;
  if (pairs) *pairs = (long long)S[kScalPairs];
  if (nk) *nk = b->boxes[box]->plan.kspace ? b->boxes[box]->plan.nk : 0;
  if (n_ions) *n_ions = b->boxes[box]->desc.N;
  return MDG_OK;
}

}  // extern "C"
