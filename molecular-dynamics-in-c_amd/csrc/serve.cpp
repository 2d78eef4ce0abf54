// serve.cpp -- continuously batched serving of MD jobs (mdg_serve).
//
// Jobs arrive over time (their arrival_s, measured from the start of the
// call when cfg->realtime is set).  Waiting jobs are admitted in arrival
// order whenever a slot is free and the job's device footprint fits in the
// HBM budget next to the resident jobs; a job that cannot fit even alone
// is rejected.  Resident jobs keep their state on the device and advance
// together, a tick of K steps at a time (one graph replay per 8 steps of
// all resident boxes), K = min(steps_per_tick, smallest remaining step
// count).  Between ticks finished jobs leave (their memory is freed) and
// new ones join, so work of jobs that arrived at different times overlaps
// on the device.  Because every box's arithmetic depends on that box alone,
// a job's result is identical whether it ran alone or in any mix.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <thread>
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

// A membership that lasts at least this many steps runs from a captured
// step graph; shorter ones launch their steps directly.
constexpr int kGraphHorizon = 32;

struct Running {
  int job;
  Box* box;
  int left;
  bool first_done;
  long ions;
};

double now_s(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// NaCl box of the job (a = 5.64 A), Ewald split for cfg->eps with
// rc = min(cfg->rc_max, L_min/2), or mdg_ewald_rc_rule's with auto_split.
int job_params(const mdg_job& j, const mdg_serve_config& c, mdg_box_params* p) {
  memset(p, 0, sizeof(*p));
  if (j.nx <= 0 || j.ny <= 0 || j.nz <= 0 || j.steps <= 0 || !(j.temp_k >= 0)) return MDG_E_ARG;
  p->nx = j.nx; p->ny = j.ny; p->nz = j.nz;
  p->ax = p->ay = p->az = 5.64;
  const double Lmin = 5.64 * std::min(j.nx, std::min(j.ny, j.nz));
  const double Lmax = 5.64 * std::max(j.nx, std::max(j.ny, j.nz));
  p->rc = std::min(c.rc_max, 0.5 * Lmin);
  if (c.auto_split) {
    const int rc0 = mdg_ewald_rc_rule(8 * j.nx * j.ny * j.nz, Lmin, c.rc_max, &p->rc);
    if (rc0 != MDG_OK) return rc0;
  }
  int rc = mdg_ewald_params(p->rc, Lmax, c.eps, &p->alpha, &p->hm);
  if (rc != MDG_OK) return rc;
  p->dt_fs = c.dt_fs;
  p->real_mode = c.real_mode;
  p->cells = MDG_CELLS_AUTO;
  return mdg_potential_nacl(&p->pot);
}

double checksum(const std::vector<double>& r) {
  double s = 0.0;
  for (double x : r) s += x;
  return s;
}

}  // namespace

extern "C" int mdg_serve_job_params(const mdg_job* job, const mdg_serve_config* cfg,
                                    mdg_box_params* out) {
  if (!job || !cfg || !out || !(cfg->eps > 0) || !(cfg->rc_max > 0) || !(cfg->dt_fs > 0))
    return MDG_E_ARG;
  return job_params(*job, *cfg, out);
}

extern "C" int mdg_serve(const mdg_job* jobs, int njobs, const mdg_serve_config* cfg,
                         mdg_job_result* results, mdg_serve_stats* stats) {
  if (!jobs || njobs <= 0 || !cfg || !results || !stats) return MDG_E_ARG;
  if (cfg->precision != MDG_FP64 && cfg->precision != MDG_MIXED) return MDG_E_ARG;
  if (cfg->max_slots <= 0 || cfg->hbm_budget == 0 || !(cfg->eps > 0) || !(cfg->rc_max > 0) ||
      !(cfg->dt_fs > 0))
    return MDG_E_ARG;
  int ndev = 0;
  if (mdg_device_count(&ndev) != MDG_OK) return MDG_E_NODEVICE;
  const int slots = cfg->one_at_a_time ? 1 : cfg->max_slots;
  const int per_tick = std::max(1, cfg->steps_per_tick);
  memset(results, 0, sizeof(mdg_job_result) * njobs);
  memset(stats, 0, sizeof(*stats));

  // jobs in arrival order; footprints known before admission
  std::vector<int> order(njobs);
  for (int k = 0; k < njobs; ++k) order[k] = k;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return jobs[a].arrival_s < jobs[b].arrival_s; });
  std::vector<mdg_box_params> prm(njobs);
  std::vector<size_t> need(njobs, 0);
  for (int k = 0; k < njobs; ++k) {
    int rc = job_params(jobs[k], *cfg, &prm[k]);
    if (rc == MDG_OK) {
      BoxSpec s;
      std::vector<double> pe, fe;
      rc = spec_from_params(prm[k], &s, &pe, &fe);
      if (rc == MDG_OK) need[k] = box_bytes_for(s, cfg->precision);
      if (rc == MDG_OK && need[k] == 0) rc = MDG_E_ARG;
    }
    results[k].status = rc;
  }

  hipStream_t stream = nullptr;
  if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) return MDG_E_HIP;
  // the end of the last two ticks (the host runs at most two ahead)
  hipEvent_t ev[2] = {nullptr, nullptr};
  for (auto& e : ev)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      for (auto& f : ev)
        if (f) (void)hipEventDestroy(f);
      (void)hipStreamDestroy(stream);
      return MDG_E_HIP;
    }
  long long ticks_enq = 0;
  BoxSet set(cfg->precision, cfg->real_mode);
  std::vector<Running> run;
  std::deque<int> waiting;
  size_t resident = 0;
  int next = 0, done = 0;
  double ion_steps = 0.0;
  int rc_all = MDG_OK;
  const auto t0 = std::chrono::steady_clock::now();

  while (done < njobs && rc_all == MDG_OK) {
    const double t = now_s(t0);
    // arrivals (all at once when not real-time)
    while (next < njobs && (!cfg->realtime || jobs[order[next]].arrival_s <= t)) {
      const int j = order[next++];
      if (results[j].status != MDG_OK) {
        ++done;
        ++stats->rejected;
        continue;
      }
      if (need[j] > cfg->hbm_budget) {
        results[j].status = MDG_E_BUDGET;
        ++done;
        ++stats->rejected;
        continue;
      }
      waiting.push_back(j);
    }
    // admission in arrival order while a slot and the budget allow; it
    // rewrites the work lists, so the device must have finished the ticks
    // in flight first
    bool changed = false;
    if (!waiting.empty() && (int)run.size() < slots &&
        resident + need[waiting.front()] <= cfg->hbm_budget &&
        hipStreamSynchronize(stream) != hipSuccess) {
      rc_all = MDG_E_HIP;
      break;
    }
    while (!waiting.empty() && (int)run.size() < slots &&
           resident + need[waiting.front()] <= cfg->hbm_budget) {
      const int j = waiting.front();
      waiting.pop_front();
      BoxSpec s;
      std::vector<double> pe, fe;
      Box* box = nullptr;
      std::string why;
      int rc = spec_from_params(prm[j], &s, &pe, &fe);
      if (rc == MDG_OK) rc = box_create(s, cfg->precision, &box, &why);
      std::vector<double> r, v;
      std::vector<int> sp;
      if (rc == MDG_OK) {
        r.resize(3 * (size_t)s.N);
        v.resize(3 * (size_t)s.N);
        sp.resize(s.N);
        mdg_rocksalt(prm[j].nx, prm[j].ny, prm[j].nz, prm[j].ax, prm[j].ay, prm[j].az, r.data(),
                     sp.data());
        rc = mdg_maxwell(jobs[j].seed, jobs[j].temp_k, s.N, sp.data(), &prm[j].pot, v.data());
        if (rc == MDG_OK) rc = box_upload_state(box, r.data(), v.data(), sp.data(), stream);
      }
      if (rc != MDG_OK) {
        box_destroy(box);
        results[j].status = rc;
        ++done;
        continue;
      }
      results[j].admit_s = now_s(t0);
      resident += box->bytes;
      stats->peak_bytes = std::max(stats->peak_bytes, resident);
      run.push_back({j, box, jobs[j].steps, false, (long)s.N});
      changed = true;
    }
    stats->max_resident = std::max(stats->max_resident, (int)run.size());
    if (run.empty()) {
      if (next >= njobs && waiting.empty()) break;
      if (cfg->realtime && next < njobs) {
        const double wait = jobs[order[next]].arrival_s - now_s(t0);
        if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
      }
      continue;
    }
    if (changed) {
      std::vector<Box*> bs;
      for (auto& x : run) bs.push_back(x.box);
      int rc = set.set_boxes(bs);
      if (rc != MDG_OK) { rc_all = rc; break; }
    }
    // a tick right after an admission is one step long, so every job's first
    // energy is taken after exactly one step whatever the mix
    int K = changed ? 1 : per_tick;
    int horizon = 1 << 30;  // steps until the first resident job leaves
    for (auto& x : run) {
      K = std::min(K, x.left);
      horizon = std::min(horizon, x.left);
    }
    // a captured step graph pays for its capture when the membership lasts;
    // shorter stretches run eagerly
    int rc = set.step(K, stream, horizon >= kGraphHorizon);
    if (rc == MDG_OK && hipEventRecord(ev[ticks_enq % 2], stream) != hipSuccess) rc = MDG_E_HIP;
    if (rc != MDG_OK) { rc_all = rc; break; }
    ++ticks_enq;
    ++stats->ticks;
    bool need_results = false;
    for (auto& x : run) {
      x.left -= K;
      ion_steps += (double)x.ions * K;
      need_results |= !x.first_done || x.left == 0;
    }
    if (!need_results) {
      // nothing to read back: keep at most two ticks in flight
      if (ticks_enq >= 2 && hipEventSynchronize(ev[ticks_enq % 2]) != hipSuccess) {
        rc_all = MDG_E_HIP;
        break;
      }
      continue;
    }
    if (hipStreamSynchronize(stream) != hipSuccess) { rc_all = MDG_E_HIP; break; }
    // bookkeeping: first energies, finished jobs leave
    std::vector<Running> keep;
    bool left_any = false;
    for (auto& x : run) {
      double S[kScalCount];
      mdg_job_result& R = results[x.job];
      if (!x.first_done || x.left == 0) {
        int e = 0;
        if (box_read_scalars(x.box, S, stream) != MDG_OK ||
            box_read_error(x.box, &e, stream) != MDG_OK) {
          rc_all = MDG_E_HIP;
          keep.push_back(x);  // released after the loop
          continue;
        }
        if (!x.first_done) { R.etot_first = S[kScalEtot]; x.first_done = true; }
        if (e) R.status = (e & kErrState) ? MDG_E_STATE : MDG_E_RANGE;
        if (x.left == 0) R.etot_last = S[kScalEtot];
      }
      if (x.left == 0) {
        std::vector<double> r(3 * (size_t)x.box->desc.N);
        const int drc = box_download_state(x.box, r.data(), nullptr, nullptr, nullptr, stream);
        if (drc != MDG_OK) {
          R.status = drc;
          rc_all = drc;
        } else {
          R.final_checksum = checksum(r);
        }
        R.finish_s = now_s(t0);
        R.latency_s = R.finish_s - (cfg->realtime ? jobs[x.job].arrival_s : 0.0);
        resident -= x.box->bytes;
        box_destroy(x.box);
        ++done;
        left_any = true;
      } else {
        keep.push_back(x);
      }
    }
    run.swap(keep);
    if (left_any && !run.empty()) {
      std::vector<Box*> bs;
      for (auto& x : run) bs.push_back(x.box);
      int r2 = set.set_boxes(bs);
      if (r2 != MDG_OK) rc_all = r2;
    } else if (run.empty()) {
      set.set_boxes({});
    }
  }
  (void)hipStreamSynchronize(stream);
  for (auto& x : run) box_destroy(x.box);
  for (auto& e : ev) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(stream);

  stats->wall_s = now_s(t0);
  stats->ion_steps = ion_steps;
  stats->ion_steps_per_s = stats->wall_s > 0 ? ion_steps / stats->wall_s : 0.0;
  std::vector<double> lat;
  for (int k = 0; k < njobs; ++k)
    if (results[k].status == MDG_OK && results[k].finish_s > 0) lat.push_back(results[k].latency_s);
  if (!lat.empty()) {
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double p) {
      const double x = p * (lat.size() - 1);
      const size_t a = (size_t)floor(x), b = std::min(lat.size() - 1, a + 1);
      return lat[a] + (x - a) * (lat[b] - lat[a]);
    };
    stats->p50_latency_s = pct(0.50);
    stats->p99_latency_s = pct(0.99);
    double s = 0.0;
    for (double x : lat) s += x;
    stats->mean_latency_s = s / lat.size();
  }
  return rc_all;
}
