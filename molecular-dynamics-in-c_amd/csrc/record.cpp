// record.cpp -- mdc's output formats (mdc/program/files.c) and the read-back
// of device-resident boxes that feeds them, overlapped with stepping.
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "batch.h"
#include "engine.h"
#include "mdgpu.h"

using namespace mdg;

namespace {

// mdc's constants (mdc/program/md.h:14-16) and unit factors (init.c:67)
const double kAvogadro = 6.0221367e5;
const double kSecd2Fs = 4.074972637944947;
const double kBoltzmann = 1.380658e-5;

// snprintf-like appender: keeps counting past the capacity
struct Text {
  char* buf;
  size_t cap, len = 0;
  Text(char* b, size_t c) : buf(b), cap(c) {
    if (cap) buf[0] = 0;
  }
  template <class... A>
  void add(const char* fmt, A... a) {
    const size_t room = len < cap ? cap - len : 0;
    const int n = snprintf(room ? buf + len : nullptr, room, fmt, a...);
    if (n > 0) len += (size_t)n;
  }
  int result() const { return len > 0x7fffffff ? MDG_E_RANGE : (int)len; }
};

// staging layout of one box in a frame: [r 3N][v 3N][scalars][observables]
size_t frame_doubles(int n) { return 6 * (size_t)n + kScalCount + kObsLen; }

}  // namespace

extern "C" {

// files.c:44-65 (print_to_file), with unit_converter's factors
// (init.c:48, 75, 78)
int mdg_format_out(char* buf, size_t cap, const mdg_out_fields* f) {
  if (!f || (!buf && cap) || f->n_ions <= 0 || f->natoms_in_mol_unit <= 0 || f->kinds < 0 ||
      f->kinds > 4 || f->cells[0] <= 0 || f->cells[1] <= 0 || f->cells[2] <= 0)
    return MDG_E_ARG;
  const double perMol = kAvogadro / ((double)(f->n_ions / f->natoms_in_mol_unit));
  const double e2t = 2.0 / (3.0 * ((double)(f->n_ions) * kBoltzmann));
  const double pp2gpa = (1.6605402 * 10000.0) / ((kSecd2Fs) * (kSecd2Fs));
  Text t(buf, cap);
  t.add("%d  %.2f %.2f %.2f  %.2f %.2f  %.3f %.3f %.3f ", f->step, f->kin * perMol,
        f->pot * perMol, (f->kin + f->pot) * perMol, f->kin * e2t, f->pres * pp2gpa,
        f->L[0] / f->cells[0], f->L[1] / f->cells[1], f->L[2] / f->cells[2]);
  for (int k = 0; k < f->kinds; ++k) t.add(" %.3f ", f->msd[k]);
  t.add("%s", "\n");
  return t.result();
}

// files.c:142-158 (r_position): four ions per line [A]
int mdg_format_positions(char* buf, size_t cap, int n, const int* species, const double* r) {
  if (n < 0 || !species || !r || (!buf && cap)) return MDG_E_ARG;
  Text t(buf, cap);
  int cnt = 0;
  for (int i = 0; i < n; ++i) {
    t.add("%d %.5f %.5f %.5f ", species[i], r[i], r[(size_t)n + i], r[2 * (size_t)n + i]);
    if (++cnt == 4) {
      t.add("%s", " \n");
      cnt = 0;
    }
  }
  t.add("%s", "\n");
  return t.result();
}

// files.c:161-181 (r_velocities): [A/fs], four ions per line
int mdg_format_velocities(char* buf, size_t cap, int n, const int* species, const double* v) {
  if (n < 0 || !species || !v || (!buf && cap)) return MDG_E_ARG;
  Text t(buf, cap);
  int cnt = 0;
  for (int i = 0; i < n; ++i) {
    const double vx = v[i] / kSecd2Fs;
    const double vy = v[(size_t)n + i] / kSecd2Fs;
    const double vz = v[2 * (size_t)n + i] / kSecd2Fs;
    t.add("%d %.8e %.8e %.8e ", species[i], vx, vy, vz);
    if (++cnt == 4) {
      t.add("%s", " \n");
      cnt = 0;
    }
  }
  t.add("%s", "\n");
  return t.result();
}

int mdg_batch_step_recorded(mdg_batch* b, int nsteps, int every, mdg_frame_fn fn, void* user) {
  if (!b || every <= 0 || nsteps < 0 || nsteps % every) return MDG_E_ARG;
  const int nb = (int)b->boxes.size();
  const int frames = nsteps / every;
  if (frames == 0) return MDG_OK;
  std::vector<long long> off(nb);
  size_t total = 0;
  for (int k = 0; k < nb; ++k) {
    off[k] = (long long)total;
    total += frame_doubles(b->boxes[k]->desc.N);
  }
  const size_t bytes = sizeof(double) * total;
  struct Res {  // released on every path out
    long long* d_off = nullptr;
    double* dev[2] = {nullptr, nullptr};
    double* host[2] = {nullptr, nullptr};
    hipStream_t copy = nullptr;
    hipEvent_t snap[2] = {nullptr, nullptr}, copied[2] = {nullptr, nullptr};
    ~Res() {
      if (copy) (void)hipStreamSynchronize(copy);
      for (int k = 0; k < 2; ++k) {
        if (dev[k]) (void)hipFree(dev[k]);
        if (host[k]) (void)hipHostFree(host[k]);
        if (snap[k]) (void)hipEventDestroy(snap[k]);
        if (copied[k]) (void)hipEventDestroy(copied[k]);
      }
      if (d_off) (void)hipFree(d_off);
      if (copy) (void)hipStreamDestroy(copy);
    }
  } R;
  MDG_HIP(hipStreamCreateWithFlags(&R.copy, hipStreamNonBlocking));
  MDG_HIP(hipMalloc(&R.d_off, sizeof(long long) * nb));
  for (int k = 0; k < 2; ++k) {
    MDG_HIP(hipMalloc(&R.dev[k], bytes));
    MDG_HIP(hipHostMalloc((void**)&R.host[k], bytes, hipHostMallocDefault));
    MDG_HIP(hipEventCreateWithFlags(&R.snap[k], hipEventDisableTiming));
    MDG_HIP(hipEventCreateWithFlags(&R.copied[k], hipEventDisableTiming));
    MDG_HIP(hipEventRecord(R.copied[k], R.copy));  // both slots start free
  }
  MDG_HIP(hipMemcpyAsync(R.d_off, off.data(), sizeof(long long) * nb, hipMemcpyHostToDevice,
                         b->stream));
  std::vector<mdg_frame> fr(nb);
  auto deliver = [&](int slot) -> int {
    MDG_HIP(hipEventSynchronize(R.copied[slot]));
    if (!fn) return MDG_OK;
    for (int k = 0; k < nb; ++k) {
      const Box* B = b->boxes[k];
      const int N = B->desc.N;
      const double* p = R.host[slot] + off[k];
      mdg_frame& f = fr[k];
      f.box = k;
      f.n_ions = N;
      f.r = p;
      f.v = p + 3 * (size_t)N;
      f.species = B->species.data();
      fill_scalars(p + 6 * (size_t)N, &f.scalars);
      f.step = f.scalars.step;
      for (int q = 0; q < 4; ++q) f.msd[q] = p[6 * (size_t)N + kScalCount + kObsMsd + q];
      fn(&f, user);
    }
    return MDG_OK;
  };
  for (int k = 0; k < frames; ++k) {
    const int slot = k & 1;
    int rc = b->set->step(every, b->stream);
    if (rc != MDG_OK) return rc;
    // the slot's last copy to the host must be done before the gather reuses it
    MDG_HIP(hipStreamWaitEvent(b->stream, R.copied[slot], 0));
    b->set->enqueue_snapshot(b->stream, R.d_off, R.dev[slot]);
    MDG_HIP(hipGetLastError());
    MDG_HIP(hipEventRecord(R.snap[slot], b->stream));
    MDG_HIP(hipStreamWaitEvent(R.copy, R.snap[slot], 0));
    MDG_HIP(hipMemcpyAsync(R.host[slot], R.dev[slot], bytes, hipMemcpyDeviceToHost, R.copy));
    MDG_HIP(hipEventRecord(R.copied[slot], R.copy));
    if (k > 0) {  // the previous frame, while this one's steps run
      rc = deliver(slot ^ 1);
      if (rc != MDG_OK) return rc;
    }
  }
  int rc = deliver((frames - 1) & 1);
  if (rc != MDG_OK) return rc;
  MDG_HIP(hipStreamSynchronize(b->stream));
  return MDG_OK;
}

}  // extern "C"

namespace {

// per-box output files of mdg_batch_write_mdc_files
struct MdcFiles {
  std::vector<FILE*> out, pos, vel;
  std::vector<std::array<int, 3>> cells;
  std::vector<int> mol_unit, kinds;
  std::vector<std::array<double, 3>> L;
  std::vector<char> text;
  int what = 0, status = MDG_OK;
  ~MdcFiles() {
    for (auto* v : {&out, &pos, &vel})
      for (FILE* fp : *v)
        if (fp) fclose(fp);
  }
};

template <class F>
void emit(MdcFiles& W, FILE* fp, F&& fmt) {
  const int need = fmt((char*)nullptr, (size_t)0);
  if (need < 0) {
    W.status = need;
    return;
  }
  if (W.text.size() < (size_t)need + 1) W.text.resize((size_t)need + 1);
  fmt(W.text.data(), W.text.size());
  if (fwrite(W.text.data(), 1, (size_t)need, fp) != (size_t)need) W.status = MDG_E_ARG;
}

void write_frame(const mdg_frame* f, void* user) {
  MdcFiles& W = *(MdcFiles*)user;
  const int k = f->box;
  mdg_out_fields o{};
  o.step = (int)f->step;
  o.kin = f->scalars.kin;
  o.pot = f->scalars.pot;
  o.pres = f->scalars.pres;
  o.n_ions = f->n_ions;
  o.natoms_in_mol_unit = W.mol_unit[k];
  for (int d = 0; d < 3; ++d) {
    o.cells[d] = W.cells[k][d];
    o.L[d] = W.L[k][d];
  }
  o.kinds = W.kinds[k];
  for (int q = 0; q < 4; ++q) o.msd[q] = f->msd[q];
  emit(W, W.out[k], [&](char* b, size_t c) { return mdg_format_out(b, c, &o); });
  if (W.what & MDG_WRITE_POSITIONS)
    emit(W, W.pos[k], [&](char* b, size_t c) {
      return mdg_format_positions(b, c, f->n_ions, f->species, f->r);
    });
  if (W.what & MDG_WRITE_VELOCITIES)
    emit(W, W.vel[k], [&](char* b, size_t c) {
      return mdg_format_velocities(b, c, f->n_ions, f->species, f->v);
    });
}

}  // namespace

extern "C" int mdg_batch_write_mdc_files(mdg_batch* b, int nsteps, int every, int what,
                                         const char* dir) {
  if (!b || !dir || what < 0 || what > 3) return MDG_E_ARG;
  const int nb = (int)b->boxes.size();
  MdcFiles W;
  W.what = what;
  W.cells = b->cells;
  W.mol_unit = b->mol_unit;
  for (Box* x : b->boxes) W.kinds.push_back(x->desc.kinds);
  W.out.assign(nb, nullptr);
  W.pos.assign(nb, nullptr);
  W.vel.assign(nb, nullptr);
  for (int k = 0; k < nb; ++k) {
    const BoxSpec& sp = b->boxes[k]->plan.spec;
    W.L.push_back({sp.L[0], sp.L[1], sp.L[2]});
    const std::string base = std::string(dir) + "/box" + std::to_string(k) + "/";
    auto open = [&](FILE** fp, const char* name) {
      *fp = fopen((base + name).c_str(), "w");
      if (*fp) fputc(' ', *fp);  // files.c:29-32 starts every file with a blank
      return *fp != nullptr;
    };
    if (!open(&W.out[k], "out") ||
        ((what & MDG_WRITE_POSITIONS) && !open(&W.pos[k], "positions")) ||
        ((what & MDG_WRITE_VELOCITIES) && !open(&W.vel[k], "velocities"))) {
      fprintf(stderr, "libmdgpu: cannot open the output files under %s\n", base.c_str());
      return MDG_E_ARG;
    }
  }
  int rc = mdg_batch_step_recorded(b, nsteps, every, write_frame, &W);
  return rc != MDG_OK ? rc : W.status;
}
