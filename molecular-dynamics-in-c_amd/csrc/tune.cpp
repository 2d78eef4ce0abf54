// tune.cpp -- the tiling autotuner (mdg_tune) and its table's persistence.
//
// For each box class the tuner times candidate tilings of the class's first
// box in a batch of copies large enough to fill the chip (the batched
// setting the tiling is for).  One run times every kernel of a step with HIP
// events, so a run can carry one real-space, one structure-factor and one
// back-projection candidate at once: the three kernel families touch
// disjoint work and are chosen independently.  Each family's first
// candidate is the built-in rule, and a candidate must beat the best so far
// by 2% to replace it, so tuning never trades a measured loss for noise.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <string>
#include <vector>

#include "engine.h"
#include "host_setup.h"
#include "mdgpu.h"

using namespace mdg;

namespace mdg {
int spec_from_params(const mdg_box_params& p, BoxSpec* s, std::vector<double>* pe,
                     std::vector<double>* fe);
}

namespace {

unsigned long long mix64(unsigned long long& s) {
  unsigned long long z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Run {
  double real_ms = 0, sk_ms = 0, kf_ms = 0;
};

// Times one candidate: R copies of the box, untimed warm-up steps, then 3
// timed steps.
int time_candidate(const BoxSpec& base, const mdg_box_params& prm, int precision,
                   const Tiling& t, int R, hipStream_t s, Run* out) {
  BoxSpec sp = base;
  sp.tile_override = &t;
  std::vector<Box*> boxes;
  std::string why;
  int rc = MDG_OK;
  const int N = sp.N;
  std::vector<double> r(3 * (size_t)N);
  std::vector<int> species(N);
  mdg_rocksalt(prm.nx, prm.ny, prm.nz, prm.ax, prm.ay, prm.az, r.data(), species.data());
  // lattice with a small deterministic jitter (a liquid-like pair count)
  unsigned long long seed = 0x5EEDull + (unsigned long long)N;
  for (int d = 0; d < 3; ++d)
    for (int i = 0; i < N; ++i) {
      double x = r[(size_t)d * N + i] + 0.3 * ((double)(mix64(seed) >> 11) * 0x1.0p-53 - 0.5);
      const double L = sp.L[d];
      if (x < 0) x += L;
      if (x >= L) x -= L;
      r[(size_t)d * N + i] = x;
    }
  for (int k = 0; k < R && rc == MDG_OK; ++k) {
    Box* b = nullptr;
    rc = box_create(sp, precision, &b, &why);
    if (rc == MDG_OK) {
      boxes.push_back(b);
      rc = box_upload_state(b, r.data(), nullptr, species.data(), s);
    }
  }
  if (rc == MDG_OK) {
    BoxSet set(precision, sp.real_mode);
    rc = set.set_boxes(boxes);
    std::vector<std::string> names;
    std::vector<double> ms;
    std::vector<int> launches;
    // untimed steps for at least 20 ms first: a device left idle while the
    // host built the boxes runs its first kernels at a lower clock (times
    // measured straight away came out two to three times too long, and
    // ranked the candidates wrongly)
    if (rc == MDG_OK) {
      const auto w0 = std::chrono::steady_clock::now();
      do {
        rc = set.step(4, s, false);
        if (rc == MDG_OK && hipStreamSynchronize(s) != hipSuccess) rc = MDG_E_HIP;
      } while (rc == MDG_OK &&
               std::chrono::steady_clock::now() - w0 < std::chrono::milliseconds(20));
    }
    // each family once over all copies, back to back (families are timed apart)
    if (rc == MDG_OK) rc = set.profile(3, s, false, &names, &ms, &launches);
    if (rc == MDG_OK) {
      *out = Run();
      for (size_t q = 0; q < names.size(); ++q) {
        const std::string& n = names[q];
        if (n == "cell_build" || n == "real_space") out->real_ms += ms[q];
        else if (n == "phase_tables" || n == "sk_partial" || n == "sk_finalize")
          out->sk_ms += ms[q];
        else if (n == "kforce" || n == "kf_pack") out->kf_ms += ms[q];
      }
      if (getenv("MDGPU_TUNE_TRACE") && getenv("MDGPU_TUNE_TRACE")[0] == '2')
        for (size_t q = 0; q < names.size(); ++q)
          fprintf(stderr, "  %s %.4f ms (%d)\n", names[q].c_str(), ms[q], launches[q]);
      // k_force_vv reads every group's (and in FP64 every split's) partial
      // forces: charge those bytes (at 5 TB/s) to the families that create them
      const double per = 24.0 * N * R / 5e12 * 1e3;
      if (precision == MDG_FP64) out->real_ms += (t.rsplit > 1 ? t.rsplit : 0) * per;
      out->kf_ms += t.ngroup * per;
    }
    set.set_boxes({});
  }
  (void)hipStreamSynchronize(s);
  for (Box* b : boxes) box_destroy(b);
  return rc;
}

template <class T>
void push_unique(std::vector<T>& v, const T& x) {
  if (std::find(v.begin(), v.end(), x) == v.end()) v.push_back(x);
}

struct RealCand {
  bool cells;
  double w;
  int rsplit;
  bool operator==(const RealCand& o) const {
    return cells == o.cells && w == o.w && rsplit == o.rsplit;
  }
};

int tune_class(const BoxSpec& spec, const mdg_box_params& prm, int precision, long long ctx,
               hipStream_t s, int* runs) {
  BoxPlan P;
  BoxSpec sp0 = spec;
  sp0.tile_override = nullptr;
  int rc = make_plan(sp0, precision, &P, nullptr);
  if (rc != MDG_OK) return rc;
  const Tiling def = default_tiling(P, P.cells_ok);
  const int N = spec.N;
  // candidates (the built-in rule first)
  std::vector<RealCand> rc_list;
  rc_list.push_back({def.cells, 1.0, def.rsplit});
  if (!spec.ordered) {
    if (P.cells_ok && spec.cells != MDG_ALL_PAIRS)
      for (double w : {1.0, 1.25})
        for (int r : {1, 3, 9, 27}) push_unique(rc_list, RealCand{true, w, r});
    if (spec.cells != MDG_CELL_LIST && N <= 8192)
      for (int r : {1, 2, 4, 8, 16})
        if (r <= std::max(1, N / 32)) push_unique(rc_list, RealCand{false, 1.0, r});
  }
  std::vector<int> ch_list, gr_list;
  ch_list.push_back(def.nchunk);
  gr_list.push_back(def.ngroup);
  if (P.kspace) {
    for (int c : {1, 2, 4, 8, 16, 32, 64})
      if (c * kSkTile <= N + kSkTile - 1) push_unique(ch_list, c);
    for (int g : {1, 2, 4, 8, 16, 32, 64})
      if (g <= P.ncol) push_unique(gr_list, g);
  }
  // copies that fill the chip, within 8 GiB of device memory
  const size_t bytes = box_bytes_for(sp0, precision);
  if (bytes == 0) return MDG_E_ARG;
  long long R = std::max(1ll, (ctx + N - 1) / N);
  R = std::min(R, 512ll);
  R = std::min(R, std::max(1ll, (long long)((8ull << 30) / bytes)));
  const size_t nrun = std::max(rc_list.size(), std::max(ch_list.size(), gr_list.size()));
  // MDGPU_TUNE_TRACE=1: every candidate's times on stderr
  const char* tr = getenv("MDGPU_TUNE_TRACE");
  const bool trace = tr && (tr[0] == '1' || tr[0] == '2');
  std::vector<double> best_real(rc_list.size(), 1e30), best_ch(ch_list.size(), 1e30),
      best_gr(gr_list.size(), 1e30);
  for (size_t q = 0; q < nrun; ++q) {
    const size_t ir = q % rc_list.size(), ic = q % ch_list.size(), ig = q % gr_list.size();
    Tiling t = def;
    t.cells = rc_list[ir].cells;
    t.cell_w = rc_list[ir].w;
    t.rsplit = rc_list[ir].rsplit;
    t.nchunk = ch_list[ic];
    t.ngroup = gr_list[ig];
    Run run;
    rc = time_candidate(spec, prm, precision, t, (int)R, s, &run);
    if (rc != MDG_OK) return rc;
    ++*runs;
    if (trace)
      fprintf(stderr,
              "mdg_tune N=%d nk=%d: cells %d w %.2f rsplit %d | nchunk %d | ngroup %d -> "
              "real %.4f sk %.4f kf %.4f ms\n",
              N, P.nk, t.cells ? 1 : 0, t.cell_w, t.rsplit, t.nchunk, t.ngroup, run.real_ms,
              run.sk_ms, run.kf_ms);
    best_real[ir] = std::min(best_real[ir], run.real_ms);
    best_ch[ic] = std::min(best_ch[ic], run.sk_ms);
    best_gr[ig] = std::min(best_gr[ig], run.kf_ms);
  }
  auto pick = [](const std::vector<double>& v) {
    size_t b = 0;
    for (size_t k = 1; k < v.size(); ++k)
      if (v[k] < 0.98 * v[b]) b = k;
    return b;
  };
  Tiling t = def;
  const RealCand& r = rc_list[pick(best_real)];
  t.cells = r.cells;
  t.cell_w = r.w;
  t.rsplit = r.rsplit;
  t.nchunk = ch_list[pick(best_ch)];
  t.ngroup = gr_list[pick(best_gr)];
  t.tuned = true;
  tune_put(tune_key(sp0, precision, P.nk, P.cells_ok), t);
  return MDG_OK;
}

}  // namespace

extern "C" {

int mdg_tune(int nbox, const mdg_box_params* params, int precision, long long context_ions,
             mdg_tune_report* rep) {
  if (nbox <= 0 || !params) return MDG_E_ARG;
  if (precision != MDG_FP64 && precision != MDG_MIXED) return MDG_E_ARG;
  int ndev = 0;
  if (mdg_device_count(&ndev) != MDG_OK || ndev < 1) return MDG_E_NODEVICE;
  const long long ctx = context_ions > 0 ? context_ions : 131072;
  const auto t0 = std::chrono::steady_clock::now();
  // one representative per class, in order of first appearance
  std::map<TuneKey, int> first;
  std::vector<int> reps;
  for (int k = 0; k < nbox; ++k) {
    BoxSpec s;
    std::vector<double> pe, fe;
    int rc = spec_from_params(params[k], &s, &pe, &fe);
    if (rc != MDG_OK) return rc;
    BoxPlan P;
    rc = make_plan(s, precision, &P, nullptr);
    if (rc != MDG_OK) return rc;
    const TuneKey key = tune_key(s, precision, P.nk, P.cells_ok);
    if (first.emplace(key, k).second) reps.push_back(k);
  }
  hipStream_t stream = nullptr;
  if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) return MDG_E_HIP;
  int runs = 0, rc = MDG_OK;
  for (int k : reps) {
    BoxSpec s;
    std::vector<double> pe, fe;
    rc = spec_from_params(params[k], &s, &pe, &fe);
    if (rc == MDG_OK) rc = tune_class(s, params[k], precision, ctx, stream, &runs);
    if (rc != MDG_OK) break;
  }
  (void)hipStreamDestroy(stream);
  if (rep) {
    rep->classes = (int)reps.size();
    rep->runs = runs;
    rep->seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  return rc;
}

int mdg_tune_clear(void) {
  tune_clear();
  return MDG_OK;
}

// One line per class: the key, then the tiling.
int mdg_tune_save(const char* path) {
  if (!path) return MDG_E_ARG;
  FILE* f = fopen(path, "w");
  if (!f) return MDG_E_ARG;
  fprintf(f, "# libmdgpu tuning table: precision real_mode cells_ok nbucket kbucket  "
             "cells cell_w rsplit nchunk ngroup\n");
  for (const auto& e : tune_entries()) {
    const TuneKey& k = e.first;
    const Tiling& t = e.second;
    fprintf(f, "%d %d %d %d %d  %d %.17g %d %d %d\n", k.precision, k.real_mode, k.cells_ok,
            k.nbucket, k.kbucket, t.cells ? 1 : 0, t.cell_w, t.rsplit, t.nchunk, t.ngroup);
  }
  const bool ok = fclose(f) == 0;
  return ok ? MDG_OK : MDG_E_ARG;
}

int mdg_tune_load(const char* path) {
  if (!path) return MDG_E_ARG;
  FILE* f = fopen(path, "r");
  if (!f) return MDG_E_ARG;
  char line[512];
  int rc = MDG_OK;
  std::vector<std::pair<TuneKey, Tiling>> got;
  while (fgets(line, sizeof(line), f)) {
    if (line[0] == '#' || line[0] == '\n') continue;
    TuneKey k;
    Tiling t;
    int cells = 0;
    if (sscanf(line, "%d %d %d %d %d %d %lf %d %d %d", &k.precision, &k.real_mode,
               &k.cells_ok, &k.nbucket, &k.kbucket, &cells, &t.cell_w, &t.rsplit, &t.nchunk,
               &t.ngroup) != 10 ||
        t.rsplit < 1 || t.nchunk < 1 || t.ngroup < 1 || !(t.cell_w >= 1.0)) {
      rc = MDG_E_ARG;
      break;
    }
    t.cells = cells != 0;
    t.tuned = true;
    got.push_back({k, t});
  }
  fclose(f);
  if (rc == MDG_OK)
    for (const auto& e : got) tune_put(e.first, e.second);
  return rc;
}

int mdg_box_tiling(const mdg_box_params* p, int precision, mdg_tiling* out) {
  if (!p || !out) return MDG_E_ARG;
  if (precision != MDG_FP64 && precision != MDG_MIXED) return MDG_E_ARG;
  BoxSpec s;
  std::vector<double> pe, fe;
  int rc = spec_from_params(*p, &s, &pe, &fe);
  if (rc != MDG_OK) return rc;
  BoxPlan P;
  rc = make_plan(s, precision, &P, nullptr);
  if (rc != MDG_OK) return rc;
  out->cells = P.tile.cells ? 1 : 0;
  out->cell_w = P.tile.cell_w;
  out->rsplit = P.tile.rsplit;
  out->nchunk = P.tile.nchunk;
  out->ngroup = P.tile.ngroup;
  out->tuned = P.tile.tuned ? 1 : 0;
  return MDG_OK;
}

}  // extern "C"
