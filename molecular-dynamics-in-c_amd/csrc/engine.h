// engine.h -- host orchestration of libmdgpu: box plans, device-resident
// boxes, and box sets (the work lists of one launch over many boxes).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "mdg_internal.h"
#include "mdgpu.h"

namespace mdg {

// Everything that defines one box's computation.
struct BoxSpec {
  int N = 0;
  double L[3] = {0, 0, 0};
  double alpha = 0;       // Ewald alpha; 0 or > 10 skips Ewald
  double a3 = 0;          // alpha / sqrt(pi) as mdc computes it
  double p2a2 = 0;        // pi^2 / alpha^2
  int hm = 0;
  double rc = 0;
  double rc2 = 0;         // cutoff squared; rc * rc when 0
  double dt = 0, dt2 = 0; // mdc time units
  int real_mode = MDG_TABLE;
  int cells = MDG_CELLS_AUTO;
  bool ordered = false;   // real space in mdc's partner order (no splitting)
  int kinds = 2;
  double Z[4] = {0, 0, 0, 0}, m[4] = {0, 0, 0, 0};
  int lookup[16] = {0, 1, 1, 2};
  mdg_potential pot{};    // BMH parameters (analytic mode)
  // pair table (table mode): ntypes * nb values each, mdc's layout
  const double* pe = nullptr;
  const double* fe = nullptr;
  int table_nb = 0, table_ntypes = 3;
  const struct Tiling* tile_override = nullptr;  // the tuner's candidates
};

// Tile sizes of one box; a function of the box alone (and its precision),
// so a box computes the same bits alone and in any batch.
struct Tiling {
  int rsplit = 1;        // real-space partner splits per home unit
  int nchunk = 1;        // structure-factor ion chunks
  int chunk = 0;         // ions per chunk (multiple of kSkTile)
  int ngroup = 1;        // back-projection column groups
  bool cells = false;    // cell list (else all pairs)
  double cell_w = 1.0;   // cell width / rc (>= 1)
  bool tuned = false;    // taken from the tuning table
};

// The tuning table: measured tilings per box class (precision, real-space
// mode, whether a cell list fits, and sqrt(2)-wide buckets of N and of the
// k-vector count).  Boxes created while an entry exists for their class use
// it; the others use the built-in rule.
struct TuneKey {
  int precision, real_mode, cells_ok, nbucket, kbucket;
  bool operator<(const TuneKey& o) const;
};
TuneKey tune_key(const BoxSpec& spec, int precision, int nk, bool cells_ok);
bool tune_get(const TuneKey& k, Tiling* t);
void tune_put(const TuneKey& k, const Tiling& t);
void tune_clear();
std::vector<std::pair<TuneKey, Tiling>> tune_entries();

struct BoxPlan {
  BoxSpec spec;
  bool ewald = false;               // alpha rule of rcprcl.c:23,115 (gates the self term)
  bool kspace = false;              // ewald and at least one k vector (k-space work)
  bool mixed = false;               // planned for the mixed precision path
  int hm_sqrt = 0, hmax = 0, nk = 0, ncol = 0, nseg = 0;
  std::vector<KCol> cols;
  std::vector<KSeg> segs;
  std::vector<KQuad> quads;
  std::vector<int> khz, kslot;
  int nquad = 0, nslot = 0;
  std::vector<double> kce, kcf;
  bool all_pairs = true;
  bool cells_ok = false;            // >= 3 cells of width rc per axis
  int nc[3] = {0, 0, 0}, ncell = 0;
  int cell_hb = 1;                  // home blocks of 64 per cell, one item each (strided)
  int nreal = 0;
  Tiling tile;
  // back-projection groups: first quad (FP64) or first MFMA tile (mixed) of
  // each group, plus the end
  std::vector<int> kf_group_start;
  std::vector<KfTile> kftiles;      // mixed: MFMA tiles of 8 columns
  int kbk = 0;                      // mixed: K blocks (8 hz each) covering 0..hmax
};

int make_plan(const BoxSpec& spec, int precision, BoxPlan* out, std::string* why);
// The built-in tiling rule of a plan (before tuning-table lookup).
Tiling default_tiling(const BoxPlan& P, bool cells_ok);

struct Box {
  BoxPlan plan;
  int precision = MDG_FP64;
  BoxDesc desc{};
  void* dmem = nullptr;
  size_t bytes = 0;
  std::vector<int> species;   // host copy (self term, masses)
  bool desc_dirty = true;
};

// Device bytes a box of this spec would hold (for admission control).
size_t box_bytes_for(const BoxSpec& spec, int precision);
int box_create(const BoxSpec& spec, int precision, Box** out, std::string* why);
// The same box in `cap` bytes of device memory at `base` that the caller
// owns (nothing is allocated; box_destroy frees nothing of it).  The block
// is zeroed on stream s.  MDG_E_BUDGET when the box needs more than cap.
int box_create_at(const BoxSpec& spec, int precision, void* base, size_t cap, hipStream_t s,
                  Box** out, std::string* why);
void box_destroy(Box* b);
int box_upload_state(Box* b, const double* r, const double* v, const int* species,
                     hipStream_t s);
int box_download_state(Box* b, double* r, double* v, double* f, double* f0, hipStream_t s);
int box_read_scalars(Box* b, double* scal /*kScalCount*/, hipStream_t s);
int box_read_error(Box* b, int* err, hipStream_t s);

// The per-step observables and thermostat of mdc's step loop for one box
// (main.c:44-60).  Arming resets their state: the MSD origin is recorded
// at the next step, and the thermostat's running sums restart.
struct ObserveOpts {
  int msd = 0;                       // calc_msd every step
  int t_control_step = 0;            // control_temp period; 0: off
  double temp_K = 0;                 // control_temp's target temperature
  long long t_control_until = 10000; // control_temp only while step < this
  int moment_every = 0;              // moment_correction period; 0: off
};
int box_set_observe(Box* b, const ObserveOpts& o, hipStream_t s);
// obs: kObsLen values; old, d: 3N values each (x, y, z blocks) or null
int box_read_observe(Box* b, double* obs, double* old, double* d, hipStream_t s);
// upload msd_old, msd_d (3N values each, or null) and the observable state
int box_write_observe(Box* b, const double* obs, const double* old, const double* d,
                      hipStream_t s);

// Box groups of a batch whose step chains run concurrently (16+ boxes).
constexpr int kStepGroups = 2;
// default share of the work in the group that starts k-space first
// (mdg_batch_tune_schedule picks one per batch by timing)
constexpr double kGroup0Share = 0.5;

// Steps per captured step graph (BoxSet::step replays it).
constexpr int kGraphSteps = 8;

// A set of boxes stepped together (one launch per kernel per step).
class BoxSet {
 public:
  explicit BoxSet(int precision, int real_mode);
  ~BoxSet();
  int set_boxes(const std::vector<Box*>& boxes);  // rebuild work lists
  // Group 0's share of the work in the two-group step schedule (takes
  // effect at the next set_boxes).
  void set_group_share(double share) { group0_share_ = share; }
  double group_share() const { return group0_share_; }
  // Device memory for the set's descriptors and work lists, owned by the
  // caller: set_boxes then allocates nothing (MDG_E_BUDGET when they do
  // not fit), and steps run on the caller's stream alone (no side streams
  // or events are created).
  void set_arena(void* base, size_t bytes);
  const std::vector<Box*>& boxes() const { return boxes_; }
  BoxDesc* device_desc() const { return d_desc_; }  // the boxes' descriptors on the device
  // mdc's calc_foc on every box (f0 <- f, f <- forces); vv: also next_v and
  // next_r and count the step.
  void enqueue_step(hipStream_t s, bool vv);
  void enqueue_forces_parts(hipStream_t s, bool real, int real_write_mode, bool recip);
  // f0 <- f, f <- the selected terms (bit 0 real space, bit 1 reciprocal
  // space); energies are not updated.
  void enqueue_force_terms(hipStream_t s, int terms);
  void enqueue_scalars(hipStream_t s, bool count_step);
  void enqueue_clear_foc(hipStream_t s);
  void enqueue_next_v(hipStream_t s);
  void enqueue_next_r(hipStream_t s);
  void enqueue_moment_correction(hipStream_t s);
  // k_observe on every box (step >= 0: mdc's sys.step; see launch_observe)
  void enqueue_observe(hipStream_t s, long long step, int phases);
  void enqueue_scale_v(hipStream_t s, double scale);
  // launch_snapshot over every box (off: device array, one entry per box)
  void enqueue_snapshot(hipStream_t s, const long long* off, double* dst);
  int step(int nsteps, hipStream_t s, bool use_graph = true);  // graph replay
  // Captures and uploads the step graph that step() replays, running nothing.
  int prepare(hipStream_t s);
  // Kernel timing over `nsteps` eager steps: every launch carries a start
  // and a stop hipEvent that its own dispatch stamps.  as_step: the step's own schedule
  // (box groups, side streams, kernels overlapping); else every kernel once
  // over all boxes, back to back.  ms: per kernel, ms per step summed over
  // its launches; launches: its launches per step.
  int profile(int nsteps, hipStream_t s, bool as_step, std::vector<std::string>* names,
              std::vector<double>* ms, std::vector<int>* launches);
  int sync_desc(hipStream_t s);                 // upload dirty descriptors
  void invalidate_graph();

 private:
  int precision_, real_mode_;
  std::vector<Box*> boxes_;
  BoxDesc* d_desc_ = nullptr;
  Item* d_items_ = nullptr;
  int* d_ids_ = nullptr;
  size_t items_cap_ = 0, ids_cap_ = 0, desc_cap_ = 0;
  double group0_share_ = kGroup0Share;
  char* arena_ = nullptr;   // caller-owned device memory (set_arena)
  size_t arena_bytes_ = 0;
  // offsets into d_items_ / d_ids_ and counts
  // Work lists: item ranges in d_items_ per kernel, box-id ranges in d_ids_.
  enum { LReal, LTab, LSk, LFin, LKf, LIon, LPack, LCellIon, LCellPlace, kLists };
  enum { ICell, IAll, IObs, kIdLists };
  struct Ranges {
    int off[kLists] = {}, n[kLists] = {};
    int ioff[kIdLists] = {}, in[kIdLists] = {};
  };
  Ranges all_;              // every box (each list is its groups' lists back to back;
                            // the real-space list is all in group 0's range, cost-ordered)
  std::vector<Ranges> grp_; // the box groups whose chains a step runs concurrently
  int skn_ = 1;  // hmax + 1 of the widest box (MFMA structure factor LDS tiles)
  int hs_ = 1, hsw_ = 8, hz_ = 8;  // LDS strides / register row of the k-space kernels
  hipGraphExec_t graph_exec_ = nullptr;
  hipGraph_t graph_ = nullptr;
  int graph_steps_ = 0;
  // side streams of a step's concurrent branches (see enqueue_step)
  std::vector<hipStream_t> side_;
  std::vector<hipEvent_t> ev_join_;
  hipEvent_t ev_fork_ = nullptr;
  bool side_failed_ = false;
  bool serial_ = false;     // profile(): every kernel once, back to back
  // per-launch event timing while profile() runs; null otherwise
  struct LaunchTimer {
    std::vector<hipEvent_t> ev;
    std::vector<int> kernel;
  };
  LaunchTimer* timer_ = nullptr;
  template <class F>
  void timed(int kernel, hipStream_t s, F&& launch);
  bool ensure_side_streams(int count);
  void launch_cells_real(hipStream_t s, int mode, const Ranges& r);
  void launch_recip(hipStream_t s, const Ranges& r);
};

void set_last_hip_error(hipError_t e, const char* what);

}  // namespace mdg

#define MDG_HIP(x)                                       \
  do {                                                   \
    hipError_t e_ = (x);                                 \
    if (e_ != hipSuccess) {                              \
      ::mdg::set_last_hip_error(e_, #x);                 \
      return MDG_E_HIP;                                  \
    }                                                    \
  } while (0)
