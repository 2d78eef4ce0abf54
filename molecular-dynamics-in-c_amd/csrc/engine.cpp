// engine.cpp -- box plans, device-resident boxes and box sets.
#include "engine.h"

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>

#include "host_setup.h"
#include "launch.h"

namespace mdg {

static std::mutex g_err_mu;
static char g_last_hip[256] = "";

void set_last_hip_error(hipError_t e, const char* what) {
  std::lock_guard<std::mutex> lk(g_err_mu);
  snprintf(g_last_hip, sizeof(g_last_hip), "%s: %s", what, hipGetErrorString(e));
  fprintf(stderr, "libmdgpu: %s\n", g_last_hip);
}

// mdc's constant (mdc/program/md.h:8), as spelled there.
static const double kPi = 3.1415926535897932385;

static int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

// ------------------------------------------------------------------ tuning table
static std::mutex g_tune_mu;
static std::map<TuneKey, Tiling> g_tune;

bool TuneKey::operator<(const TuneKey& o) const {
  const int a[5] = {precision, real_mode, cells_ok, nbucket, kbucket};
  const int b[5] = {o.precision, o.real_mode, o.cells_ok, o.nbucket, o.kbucket};
  return std::lexicographical_compare(a, a + 5, b, b + 5);
}

TuneKey tune_key(const BoxSpec& spec, int precision, int nk, bool cells_ok) {
  TuneKey k;
  k.precision = precision;
  k.real_mode = spec.real_mode;
  k.cells_ok = cells_ok ? 1 : 0;
  k.nbucket = (int)lround(2.0 * log2((double)std::max(1, spec.N)));
  k.kbucket = nk > 0 ? (int)lround(2.0 * log2((double)nk)) : -1;
  return k;
}

bool tune_get(const TuneKey& k, Tiling* t) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(k);
  if (it == g_tune.end()) return false;
  *t = it->second;
  return true;
}

void tune_put(const TuneKey& k, const Tiling& t) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[k] = t;
}

void tune_clear() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune.clear();
}

std::vector<std::pair<TuneKey, Tiling>> tune_entries() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  return std::vector<std::pair<TuneKey, Tiling>>(g_tune.begin(), g_tune.end());
}

// ------------------------------------------------------------------ tiling
// Built-in rule: work granularity per box, from the box's own size only.  A
// lone box should still spread over the chip (about 4096 waves, four per
// SIMD, for the large ones: a lone 32768-ion KCl box at rc = 16.5 A ran its
// real space in 0.43 ms with 1944 waves and 0.24 ms with 5832, and its
// back-projection in 0.26 ms with 1024 waves and 0.14 ms with 4096);
// partial-sum buffers stay bounded.
Tiling default_tiling(const BoxPlan& P, bool can_cells) {
  const BoxSpec& sp = P.spec;
  Tiling t;
  const int N = sp.N;
  if (sp.cells == MDG_CELL_LIST) t.cells = true;
  else if (sp.cells == MDG_CELLS_AUTO) t.cells = can_cells && N >= 1500;
  // real space
  if (!sp.ordered) {
    if (t.cells) {
      // items: cells x home blocks per cell x splits
      const int ncell = P.nc[0] * P.nc[1] * P.nc[2];
      const long homes = (long)ncell * std::max(1, ceil_div(ceil_div(N, std::max(1, ncell)),
                                                           kRealThreads));
      const int opts[4] = {1, 3, 9, 27};
      t.rsplit = 27;
      for (int s : opts)
        if (homes * s >= 4096) { t.rsplit = s; break; }
    } else {
      const int nhome = ceil_div(N, kRealThreads);
      t.rsplit = std::max(1, std::min(ceil_div(64, nhome), N / 128));
    }
  }
  if (P.kspace) {
    // structure factor: ion chunks of >= 128 ions, at most 64 chunks and
    // 64 MB of partial sums
    int nch = std::max(1, std::min(ceil_div(N, 128), 64));
    nch = std::max(1, std::min(nch, (int)((64l << 20) / (16l * std::max(1, P.nk)))));
    t.chunk = ceil_div(ceil_div(N, nch), kSkTile) * kSkTile;
    t.nchunk = ceil_div(N, t.chunk);
    if (!P.kftiles.empty()) {
      // MFMA back-projection (32 ions per wave): ~4096 waves for a lone
      // box, >= 8 tiles per group
      const int nib = ceil_div(N, 32);
      int g = ceil_div(4096, nib);
      g = std::min(g, std::max(1, (int)P.kftiles.size() / 8));
      t.ngroup = std::max(1, std::min(g, 64));
    } else {
      // back-projection: ~2048 waves for a lone box, >= 64 k per group
      const int nib = ceil_div(N, kKfThreads);
      int g = ceil_div(2048, nib);
      g = std::min(g, std::max(1, P.nk / 64));
      g = std::min(g, 64);
      t.ngroup = std::max(1, std::min(g, P.ncol));
    }
  }
  return t;
}

// Brings a tiling from the table or the tuner within what this box allows
// (a table entry covers a class of boxes).
static Tiling finish_tiling(const BoxPlan& P, bool can_cells, Tiling t) {
  const BoxSpec& sp = P.spec;
  const int N = sp.N;
  if (sp.cells == MDG_CELL_LIST) t.cells = true;
  else if (sp.cells == MDG_ALL_PAIRS) t.cells = false;
  else t.cells = t.cells && can_cells;
  if (!t.cells) t.cell_w = 1.0;
  t.cell_w = std::max(1.0, std::min(t.cell_w, 4.0));
  if (t.cells) {
    for (int d = 0; d < 3; ++d)
      if ((int)floor(sp.L[d] / (sp.rc * t.cell_w * (1.0 + 1e-9))) < 3) t.cell_w = 1.0;
    // FP64: the 27 cells around a home cell; mixed: the home cell and the
    // 13 of its half shell (each pair once)
    t.rsplit = std::max(1, std::min(t.rsplit, P.mixed ? 14 : 27));
  } else {
    t.rsplit = std::max(1, std::min(t.rsplit, std::max(1, N / 32)));
  }
  if (sp.ordered) t.rsplit = 1;
  if (P.kspace) {
    int nch = std::max(1, std::min(t.nchunk, 64));
    nch = std::max(1, std::min(nch, ceil_div(N, kSkTile)));
    nch = std::max(1, std::min(nch, (int)((64l << 20) / (16l * std::max(1, P.nk)))));
    // mixed path: its structure-factor partials are FP32 sums over one
    // chunk, so chunks hold at most kSkmChunk ions
    if (!P.kftiles.empty()) nch = std::max(nch, ceil_div(N, kSkmChunk));
    t.chunk = ceil_div(ceil_div(N, nch), kSkTile) * kSkTile;
    t.nchunk = ceil_div(N, t.chunk);
    t.ngroup = std::max(1, std::min(std::min(t.ngroup, 64), P.ncol));
  } else {
    t.nchunk = 1;
    t.chunk = 0;
    t.ngroup = 1;
  }
  return t;
}

// Contiguous runs of units with about equal total weight.
static std::vector<int> partition_runs(const std::vector<long>& weight, int groups) {
  std::vector<int> start;
  const int nq = (int)weight.size();
  long rows = 0;
  for (long w : weight) rows += w;
  groups = std::max(1, std::min(groups, nq));
  start.push_back(0);
  long acc = 0;
  int g = 1;
  for (int c = 0; c < nq && g < groups; ++c) {
    acc += weight[c];
    if (acc * groups >= (long)g * rows) {
      if (c + 1 < nq) start.push_back(c + 1);
      ++g;
    }
  }
  start.push_back(nq);
  return start;
}

// Quads weighted by their weight rows (FP64 back-projection).
static std::vector<int> partition_quads(const std::vector<KQuad>& quads, int groups) {
  std::vector<long> w;
  for (const KQuad& q : quads) w.push_back(q.nrow);
  return partition_runs(w, groups);
}

// MFMA tiles weighted by their K blocks (mixed back-projection).
static std::vector<int> partition_tiles(const std::vector<KfTile>& tiles, int groups) {
  std::vector<long> w;
  for (const KfTile& t : tiles) w.push_back(std::max(1, t.kb_hi - t.kb_lo));
  return partition_runs(w, groups);
}

int make_plan(const BoxSpec& sp, int precision, BoxPlan* P, std::string* why) {
  *P = BoxPlan();
  P->spec = sp;
  P->mixed = precision == MDG_MIXED;
  if (sp.N <= 0 || !(sp.L[0] > 0) || !(sp.L[1] > 0) || !(sp.L[2] > 0) || !(sp.rc > 0)) {
    if (why) *why = "box needs N > 0, L > 0 and rc > 0";
    return MDG_E_ARG;
  }
  if (precision == MDG_MIXED && sp.alpha > 0.0 && sp.alpha <= 10.0 && sp.alpha * sp.rc > 8.0) {
    if (why) *why = "mixed real space: erfc fitted for alpha * rc <= 8 (Ewald accuracy >= 1e-27)";
    return MDG_E_RANGE;
  }
  if (sp.N > kMaxIons) {
    if (why) *why = "at most 2^21 ions per box";
    return MDG_E_RANGE;
  }
  if (sp.kinds < 1 || sp.kinds > 4) {
    if (why) *why = "1..4 species supported";
    return MDG_E_ARG;
  }
  if (sp.real_mode == MDG_TABLE && (!sp.pe || !sp.fe || sp.table_nb <= 0)) {
    if (why) *why = "table mode needs a pair table";
    return MDG_E_ARG;
  }
  if (sp.real_mode == MDG_ANALYTIC && sp.kinds != 2) {
    if (why) *why = "analytic mode supports two species";
    return MDG_E_ARG;
  }
  P->ewald = !(sp.alpha == 0.0 || sp.alpha > 10.0);  // rcprcl.c:23
  if (P->ewald) {
    const int s = (int)sqrtf((float)sp.hm);           // init.c:44
    if (kforce_hz_for(s) < 0) {
      if (why) *why = "reciprocal cutoff too large (hm_sqrt > 63)";
      return MDG_E_ARG;
    }
    P->hm_sqrt = s;
    P->hmax = s;
    const double vv = sp.L[0] * sp.L[1] * sp.L[2];
    const double v2 = 2.0 / vv, vp = 1.0 / (kPi * vv);
    for (int hx = -s; hx <= s; ++hx)
      for (int hy = -s; hy <= s; ++hy) {
        // rcprcl.c:32-38: hz >= 0; at hz = 0 only hy > 0, or hy = 0 and hx > 0
        const int lo = (hy > 0 || (hy == 0 && hx > 0)) ? 0 : 1;
        int hi = -1;
        for (int hz = lo; hz <= s; ++hz)
          if (hx * hx + hy * hy + hz * hz <= sp.hm) hi = hz;
        if (hi < lo) continue;
        KCol c{};
        c.hx = hx; c.hy = hy; c.hz0 = lo; c.len = hi - lo + 1; c.k0 = P->nk;
        P->cols.push_back(c);
        for (int hz = lo; hz <= hi; ++hz) {
          const double mx = ((double)hx) / sp.L[0], my = ((double)hy) / sp.L[1],
                       mz = ((double)hz) / sp.L[2];
          const double m2 = mx * mx + my * my + mz * mz;
          const double ex = exp(-m2 * sp.p2a2);
          P->kce.push_back(vp * (ex / m2));
          P->kcf.push_back(v2 * (ex / m2));
          P->khz.push_back(hz);
        }
        // segments start at multiples of kSegLen in hz (aligned LDS reads);
        // slots below lo in the first one are computed and dropped
        for (int h0 = 0; h0 <= hi; h0 += kSegLen) {
          KSeg g{};
          g.hx = hx; g.hy = hy; g.hz0 = h0;
          g.n = std::min(kSegLen, hi + 1 - h0);
          g.qlo = std::max(0, lo - h0);
          g.k0 = c.k0 + (h0 - lo);
          P->segs.push_back(g);
        }
        P->nk += c.len;
      }
    P->ncol = (int)P->cols.size();
    P->nseg = (int)P->segs.size();
    // with no k vector there is no k-space work, but the self term of
    // reciprocal_space3 (rcprcl.c:115-133) still follows the alpha rule
    P->kspace = P->nk > 0;
    // quads of four consecutive columns over zero-padded weight rows
    P->kslot.assign(P->nk, 0);
    int row = 0;
    for (int c0 = 0; c0 < P->ncol; c0 += 4) {
      KQuad qd{};
      const int nc = std::min(4, P->ncol - c0);
      int lo = 1 << 30, hi = -1;
      for (int c = 0; c < nc; ++c) {
        const KCol& col = P->cols[c0 + c];
        lo = std::min(lo, col.hz0);
        hi = std::max(hi, col.hz0 + col.len);
        qd.hx[c] = col.hx;
        qd.hy[c] = col.hy;
      }
      for (int c = 0; c < nc; ++c) {
        const KCol& col = P->cols[c0 + c];
        for (int j = 0; j < col.len; ++j)
          P->kslot[col.k0 + j] = (row + col.hz0 + j - lo) * 4 + c;
      }
      qd.lo = lo;
      qd.nrow = hi - lo;
      qd.row0 = row;
      qd.ncol = nc;
      row += qd.nrow;
      P->quads.push_back(qd);
    }
    P->nquad = (int)P->quads.size();
    P->nslot = 4 * row;
    if (precision == MDG_MIXED) {
      // MFMA tiles: 8 consecutive columns and the K blocks their hz span
      P->kbk = (P->hmax + 1 + 7) / 8;
      for (int c0 = 0; c0 < P->ncol; c0 += kKfTileCols) {
        KfTile t{};
        t.c0 = c0;
        t.nc = std::min(kKfTileCols, P->ncol - c0);
        int lo = 1 << 30, hi = 0;
        for (int c = c0; c < c0 + t.nc; ++c) {
          lo = std::min(lo, P->cols[c].hz0);
          hi = std::max(hi, P->cols[c].hz0 + P->cols[c].len);
        }
        t.kb_lo = lo / 8;
        t.kb_hi = (hi + 7) / 8;
        P->kftiles.push_back(t);
      }
    }
  }
  // cell grid: cells wider than rc (with a relative margin, so that two ions
  // within rc never bin two cells apart after rounding)
  bool can_cells = true;
  for (int d = 0; d < 3; ++d) {
    P->nc[d] = (int)floor(sp.L[d] / (sp.rc * (1.0 + 1e-9)));
    if (P->nc[d] < 3) can_cells = false;
  }
  if (sp.cells == MDG_CELL_LIST && !can_cells) {
    if (why) *why = "cell list needs at least 3 cells of width >= rc per axis";
    return MDG_E_RANGE;
  }
  P->cells_ok = can_cells;
  Tiling t;
  if (sp.tile_override) {
    t = *sp.tile_override;
  } else if (tune_get(tune_key(sp, precision, P->nk, can_cells), &t)) {
    t.tuned = true;
  } else {
    t = default_tiling(*P, can_cells);
  }
  P->tile = finish_tiling(*P, can_cells, t);
  if (P->tile.cells && P->tile.cell_w > 1.0)
    for (int d = 0; d < 3; ++d)
      P->nc[d] = (int)floor(sp.L[d] / (sp.rc * P->tile.cell_w * (1.0 + 1e-9)));
  P->all_pairs = !P->tile.cells;
  if (!P->all_pairs) P->ncell = P->nc[0] * P->nc[1] * P->nc[2];
  else P->nc[0] = P->nc[1] = P->nc[2] = 0;
  // a cell's home ions in blocks of 64 spread over cell_hb items (block b
  // goes to item b mod cell_hb), from the mean population: large cells
  // (large cutoffs) do not serialise one item
  P->cell_hb = P->all_pairs ? 1 : std::max(1, ceil_div(ceil_div(sp.N, std::max(1, P->ncell)),
                                                       kRealThreads));
  const int homes = P->all_pairs ? ceil_div(sp.N, kRealThreads) : P->ncell * P->cell_hb;
  P->nreal = homes * P->tile.rsplit;
  if (P->kspace)
    P->kf_group_start = precision == MDG_MIXED ? partition_tiles(P->kftiles, P->tile.ngroup)
                                               : partition_quads(P->quads, P->tile.ngroup);
  return MDG_OK;
}

// The mixed real space over all pairs visits a box's ions in the order
// p -> (p s) mod N, s the integer nearest 0.618 N that is coprime with N:
// consecutive positions are then spread over the box evenly (a rock-salt
// lattice in mdc's order has neighbouring indices side by side), so every
// 64-ion home block and every 128-partner window samples the whole box, and
// the lanes' candidate counts in a window are close to one another.  On
// the 256-box batch a count over the windows gives 84% of the pop loop's
// lane slots doing accepted pairs in this order against 73% in index
// order.  Which ion pairs are evaluated, and their forces, do not depend on
// the order; the fixed-point sums are exact.  The real-space force of the
// ion at position p accumulates at fix[p], so a window's partner shares and
// a home block's sums land on consecutive words; the force assembly reads
// ion i's at position i rstride^-1 mod N.
static int real_stride(int n) {
  if (n <= 2) return 1;
  auto gcd = [](long a, long b) {
    while (b) { const long t = a % b; a = b; b = t; }
    return a;
  };
  int s = (int)std::lround(0.6180339887498949 * n);
  while (gcd(s, n) != 1) ++s;
  return s % n;
}

// s^-1 mod n for s coprime with n (extended Euclid); 1 for n <= 1
static int inverse_mod(int s, int n) {
  if (n <= 1) return 1;
  long r0 = n, r1 = s % n, t0 = 0, t1 = 1;
  while (r1) {
    const long q = r0 / r1, r2 = r0 - q * r1, t2 = t0 - q * t1;
    r0 = r1; r1 = r2; t0 = t1; t1 = t2;
  }
  return (int)(((t0 % n) + n) % n);
}

// ------------------------------------------------------------------ layout
struct Layout {
  size_t ion[12], typ, rpf, fix, cell_of, cnt, ctmp, start, sorted, s[3], styp;
  size_t cols, segs, quads, khz, kslot, kce, kcf, kab, kab32, kpot, tx, ty, tw, skp, fp;
  size_t wb, kfa, kft, sb;
  size_t tab, tab32, rpart, rpairs, spart, scal, err, msd_old, msd_d, obs, total;
};

static size_t take(size_t& off, size_t bytes) {
  const size_t o = (off + 255) & ~(size_t)255;
  off = o + bytes;
  return o;
}

static Layout layout(const BoxPlan& P, int precision) {
  Layout L;
  memset(&L, 0, sizeof(L));
  const BoxSpec& sp = P.spec;
  const long N = sp.N;
  const bool mixed = precision == MDG_MIXED;
  const size_t cplx = mixed ? sizeof(float2) : sizeof(double2);
  const int ngroup = P.kspace ? (int)P.kf_group_start.size() - 1 : 0;
  size_t off = 0;
  for (int k = 0; k < 12; ++k) L.ion[k] = take(off, sizeof(double) * N);
  L.typ = take(off, sizeof(int) * N);
  // FP64 split partials (the mixed real space sums into the fixed-point array)
  if (P.tile.rsplit > 1 && !mixed) L.rpf = take(off, sizeof(double) * 3 * (size_t)P.tile.rsplit * N);
  if (mixed) L.fix = take(off, sizeof(unsigned long long) * 3 * (size_t)N);
  if (!P.all_pairs) {
    L.cell_of = take(off, sizeof(int) * N);
    L.cnt = take(off, sizeof(int) * P.ncell);
    L.ctmp = take(off, sizeof(int) * N);
    L.start = take(off, sizeof(int) * (P.ncell + 1));
    L.sorted = take(off, sizeof(int) * N);
    for (int d = 0; d < 3; ++d) L.s[d] = take(off, sizeof(double) * N);
    L.styp = take(off, sizeof(int) * N);
  }
  if (P.kspace) {
    const size_t H = P.hmax + 1;
    L.cols = take(off, sizeof(KCol) * P.ncol);
    L.segs = take(off, sizeof(KSeg) * P.nseg);
    L.quads = take(off, sizeof(KQuad) * P.nquad);
    L.khz = take(off, sizeof(int) * P.nk);
    L.kslot = take(off, sizeof(int) * P.nk);
    L.kce = take(off, sizeof(double) * P.nk);
    L.kcf = take(off, sizeof(double) * P.nk);
    L.kab = take(off, sizeof(double) * 4 * (size_t)P.nslot);
    if (mixed) {
      const size_t nt = P.kftiles.size();
      L.wb = take(off, 32 * (size_t)(2 * P.kbk) * N);
      L.kfa = take(off, 16 * (size_t)kf_tile_stride(P.kbk) * nt);
      L.kft = take(off, sizeof(KfTile) * nt);
      const size_t npad = (size_t)(N + 7) / 8 * 8, hp = 16 * (size_t)((P.hmax + 16) / 16);
      L.sb = take(off, 16 * npad * hp);
    }
    L.kpot = take(off, sizeof(double) * P.nk);
    L.tx = take(off, cplx * H * N);
    L.ty = take(off, cplx * H * N);
    L.tw = take(off, cplx * H * N);
    L.skp = take(off, sizeof(double2) * (size_t)P.tile.nchunk * P.nk);
    L.fp = take(off, sizeof(double) * 3 * (size_t)ngroup * N);
  }
  const size_t ntab = (size_t)sp.table_ntypes * sp.table_nb;
  if (sp.real_mode == MDG_TABLE) {
    if (mixed) L.tab32 = take(off, sizeof(float2) * ntab);
    else L.tab = take(off, sizeof(double2) * ntab);
  }
  L.rpart = take(off, sizeof(double) * 4 * P.nreal);
  L.rpairs = take(off, sizeof(unsigned long long) * P.nreal);
  L.spart = take(off, sizeof(double) * 9 * (size_t)ceil_div((int)N, kIonThreads));
  L.scal = take(off, sizeof(double) * kScalCount);
  L.err = take(off, sizeof(int));
  L.msd_old = take(off, sizeof(double) * 3 * N);
  L.msd_d = take(off, sizeof(double) * 3 * N);
  L.obs = take(off, sizeof(double) * kObsLen);
  L.total = (off + 255) & ~(size_t)255;
  return L;
}

size_t box_bytes_for(const BoxSpec& spec, int precision) {
  BoxPlan P;
  if (make_plan(spec, precision, &P, nullptr) != MDG_OK) return 0;
  return layout(P, precision).total;
}

// ------------------------------------------------------------------ boxes
// Points the box's descriptor at its arrays in the zeroed device block
// `base` (layout(P).total bytes) and uploads the box's constants there.
static int bind_box(Box* b, char* base, std::string* why);

int box_create(const BoxSpec& spec, int precision, Box** out, std::string* why) {
  Box* b = new Box();
  b->precision = precision;
  int rc = make_plan(spec, precision, &b->plan, why);
  if (rc != MDG_OK) {
    delete b;
    return rc;
  }
  b->bytes = layout(b->plan, precision).total;
  if (hipMalloc(&b->dmem, b->bytes) != hipSuccess) {
    (void)hipGetLastError();
    if (why) *why = "device allocation failed";
    delete b;
    return MDG_E_NOMEM;
  }
  // hipMemset runs on the null stream, which the library's non-blocking
  // streams do not wait for: finish it before any upload or kernel
  if (hipMemset(b->dmem, 0, b->bytes) != hipSuccess || hipStreamSynchronize(nullptr) != hipSuccess) {
    (void)hipFree(b->dmem);
    delete b;
    return MDG_E_HIP;
  }
  rc = bind_box(b, (char*)b->dmem, why);
  if (rc != MDG_OK) {
    box_destroy(b);
    return rc;
  }
  *out = b;
  return MDG_OK;
}

int box_create_at(const BoxSpec& spec, int precision, void* base, size_t cap, hipStream_t s,
                  Box** out, std::string* why) {
  Box* b = new Box();
  b->precision = precision;
  int rc = make_plan(spec, precision, &b->plan, why);
  if (rc != MDG_OK) {
    delete b;
    return rc;
  }
  b->bytes = layout(b->plan, precision).total;
  if (b->bytes > cap) {
    if (why) *why = "box exceeds the device memory it was given";
    delete b;
    return MDG_E_BUDGET;
  }
  if (hipMemsetAsync(base, 0, b->bytes, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    (void)hipGetLastError();
    delete b;
    return MDG_E_HIP;
  }
  rc = bind_box(b, (char*)base, why);  // b->dmem stays null: the caller owns base
  if (rc != MDG_OK) {
    delete b;
    return rc;
  }
  *out = b;
  return MDG_OK;
}

static int bind_box(Box* b, char* base, std::string* why) {
  const BoxPlan& P = b->plan;
  const BoxSpec& spec = P.spec;
  const int precision = b->precision;
  const int N = spec.N;
  const bool mixed = precision == MDG_MIXED;
  const Layout L = layout(P, precision);
  BoxDesc& D = b->desc;
  memset(&D, 0, sizeof(D));
  double** ion_ptrs[12] = {&D.rx, &D.ry, &D.rz, &D.vx, &D.vy, &D.vz,
                           &D.fx, &D.fy, &D.fz, &D.fx0, &D.fy0, &D.fz0};
  for (int k = 0; k < 12; ++k) *ion_ptrs[k] = (double*)(base + L.ion[k]);
  D.typ = (int*)(base + L.typ);
  D.N = N;
  D.ewald = P.ewald ? 1 : 0;
  D.kspace = P.kspace ? 1 : 0;
  D.real_mode = spec.real_mode;
  D.all_pairs = P.all_pairs ? 1 : 0;
  D.rstride = P.mixed && P.all_pairs ? real_stride(N) : 1;
  D.rstride_inv = inverse_mod(D.rstride, N);
  D.kinds = spec.kinds;
  for (int k = 0; k < 16; ++k) D.lookup[k] = spec.lookup[k];
  for (int d = 0; d < 3; ++d) {
    D.L[d] = spec.L[d];
    D.halfL[d] = spec.L[d] * 0.5;  // real.c:24-26
    D.invL[d] = 1.0 / spec.L[d];
  }
  D.rc2 = spec.rc2 > 0 ? spec.rc2 : spec.rc * spec.rc;  // init.c:43
  D.alpha = spec.alpha;
  D.a3 = spec.a3;
  D.dt = spec.dt;
  D.dt2 = spec.dt2;
  for (int k = 0; k < 4; ++k) {
    D.Z[k] = spec.Z[k];
    D.m[k] = spec.m[k];
  }
  D.vol = spec.L[0] * spec.L[1] * spec.L[2];
  for (int t = 0; t < 3; ++t) {
    D.cb[t] = spec.pot.c[t] * spec.pot.b;
    D.sig[t] = spec.pot.sigma[t];
    D.Cd[t] = spec.pot.C[t];
    D.Dd[t] = spec.pot.D[t];
  }
  D.rho = spec.pot.rho;
  D.table_nb = spec.table_nb;
  D.rsplit = P.tile.rsplit;
  D.rpf = L.rpf ? (double*)(base + L.rpf) : nullptr;
  D.fix = mixed ? (unsigned long long*)(base + L.fix) : nullptr;
  if (!P.all_pairs) {
    for (int d = 0; d < 3; ++d) {
      D.nc[d] = P.nc[d];
      D.cellscale[d] = (double)P.nc[d] / spec.L[d];
    }
    D.ncell = P.ncell;
    D.cell_hb = P.cell_hb;
    D.cell_of = (int*)(base + L.cell_of);
    D.cell_count = (int*)(base + L.cnt);
    D.cell_tmp = (int*)(base + L.ctmp);
    D.cell_start = (int*)(base + L.start);
    D.sorted = (int*)(base + L.sorted);
    D.sx = (double*)(base + L.s[0]);
    D.sy = (double*)(base + L.s[1]);
    D.sz = (double*)(base + L.s[2]);
    D.styp = (int*)(base + L.styp);
  }
  int err = MDG_OK;
  auto up = [&](size_t o, const void* src, size_t n) {
    if (n && err == MDG_OK && hipMemcpy(base + o, src, n, hipMemcpyHostToDevice) != hipSuccess)
      err = MDG_E_HIP;
  };
  if (P.kspace) {
    D.nk = P.nk;
    D.ncol = P.ncol;
    D.nseg = P.nseg;
    D.hmax = P.hmax;
    D.cols = (const KCol*)(base + L.cols);
    D.segs = (const KSeg*)(base + L.segs);
    D.nquad = P.nquad;
    D.nslot = P.nslot;
    D.quads = (const KQuad*)(base + L.quads);
    D.kslot = (const int*)(base + L.kslot);
    up(L.quads, P.quads.data(), sizeof(KQuad) * P.nquad);
    up(L.kslot, P.kslot.data(), sizeof(int) * P.nk);
    D.khz = (const int*)(base + L.khz);
    D.kce = (const double*)(base + L.kce);
    D.kcf = (const double*)(base + L.kcf);
    D.kab = (double*)(base + L.kab);
    D.kab32 = nullptr;
    if (mixed) {
      D.wb = base + L.wb;
      D.kfa = base + L.kfa;
      D.kft = (const KfTile*)(base + L.kft);
      D.kbk = P.kbk;
      D.ntile = (int)P.kftiles.size();
      D.sb = base + L.sb;
      D.sknt = (P.hmax + 16) / 16;
      D.npad = (N + 7) / 8 * 8;
      up(L.kft, P.kftiles.data(), sizeof(KfTile) * P.kftiles.size());
    }
    D.kpot = (double*)(base + L.kpot);
    D.tabx = base + L.tx;
    D.taby = base + L.ty;
    D.tabw = base + L.tw;
    D.skpart = (double2*)(base + L.skp);
    D.nchunk = P.tile.nchunk;
    D.chunk = P.tile.chunk;
    D.fpart = (double*)(base + L.fp);
    D.ngroup = (int)P.kf_group_start.size() - 1;
    up(L.cols, P.cols.data(), sizeof(KCol) * P.ncol);
    up(L.segs, P.segs.data(), sizeof(KSeg) * P.nseg);
    up(L.khz, P.khz.data(), sizeof(int) * P.nk);
    up(L.kce, P.kce.data(), sizeof(double) * P.nk);
    up(L.kcf, P.kcf.data(), sizeof(double) * P.nk);
  }
  const int ntab = spec.table_ntypes * spec.table_nb;
  if (spec.real_mode == MDG_TABLE) {
    if (mixed) {
      std::vector<float2> t(ntab);
      for (int k = 0; k < ntab; ++k) t[k] = make_float2((float)spec.pe[k], (float)spec.fe[k]);
      up(L.tab32, t.data(), sizeof(float2) * ntab);
      D.table32 = (const float2*)(base + L.tab32);
    } else {
      std::vector<double2> t(ntab);
      for (int k = 0; k < ntab; ++k) t[k] = make_double2(spec.pe[k], spec.fe[k]);
      up(L.tab, t.data(), sizeof(double2) * ntab);
      D.table = (const double2*)(base + L.tab);
    }
  }
  D.rpart = (double*)(base + L.rpart);
  D.rpairs = (unsigned long long*)(base + L.rpairs);
  D.nreal = P.nreal;
  D.spart = (double*)(base + L.spart);
  D.nspart = ceil_div(N, kIonThreads);
  D.scal = (double*)(base + L.scal);
  D.err = (int*)(base + L.err);
  D.msd_old = (double*)(base + L.msd_old);
  D.msd_d = (double*)(base + L.msd_d);
  D.obs = (double*)(base + L.obs);  // observables off until set_observe()
  if (err != MDG_OK) {
    if (why) *why = "uploading box constants failed";
    return err;
  }
  b->species.assign(N, 0);
  b->desc_dirty = true;
  return MDG_OK;
}

void box_destroy(Box* b) {
  if (!b) return;
  if (b->dmem) (void)hipFree(b->dmem);
  delete b;
}

int box_upload_state(Box* b, const double* r, const double* v, const int* species,
                     hipStream_t s) {
  const int N = b->desc.N;
  BoxDesc& D = b->desc;
  if (species) {
    for (int i = 0; i < N; ++i)
      if (species[i] < 0 || species[i] >= b->plan.spec.kinds) return MDG_E_ARG;
    b->species.assign(species, species + N);
    MDG_HIP(hipMemcpyAsync(D.typ, species, sizeof(int) * N, hipMemcpyHostToDevice, s));
    // mdc's self term, rcprcl.c:119-125: Z2 summed in index order, then -Z2*a3
    double z2 = 0.0;
    for (int i = 0; i < N; ++i) {
      const double z = b->plan.spec.Z[species[i]];
      z2 += z * z;
    }
    D.self_pot = -z2 * b->plan.spec.a3;
    b->desc_dirty = true;
  }
  if (r) {
    for (long k = 0; k < 3L * N; ++k)
      if (!isfinite(r[k])) return MDG_E_STATE;
    MDG_HIP(hipMemcpyAsync(D.rx, r, sizeof(double) * N, hipMemcpyHostToDevice, s));
    MDG_HIP(hipMemcpyAsync(D.ry, r + N, sizeof(double) * N, hipMemcpyHostToDevice, s));
    MDG_HIP(hipMemcpyAsync(D.rz, r + 2 * (long)N, sizeof(double) * N, hipMemcpyHostToDevice, s));
  }
  if (v) {
    MDG_HIP(hipMemcpyAsync(D.vx, v, sizeof(double) * N, hipMemcpyHostToDevice, s));
    MDG_HIP(hipMemcpyAsync(D.vy, v + N, sizeof(double) * N, hipMemcpyHostToDevice, s));
    MDG_HIP(hipMemcpyAsync(D.vz, v + 2 * (long)N, sizeof(double) * N, hipMemcpyHostToDevice, s));
  }
  MDG_HIP(hipStreamSynchronize(s));
  return MDG_OK;
}

int box_download_state(Box* b, double* r, double* v, double* f, double* f0, hipStream_t s) {
  const int N = b->desc.N;
  const BoxDesc& D = b->desc;
  double* dst[4] = {r, v, f, f0};
  const double* src[4][3] = {{D.rx, D.ry, D.rz}, {D.vx, D.vy, D.vz}, {D.fx, D.fy, D.fz},
                             {D.fx0, D.fy0, D.fz0}};
  for (int a = 0; a < 4; ++a) {
    if (!dst[a]) continue;
    for (int d = 0; d < 3; ++d)
      MDG_HIP(hipMemcpyAsync(dst[a] + (long)d * N, src[a][d], sizeof(double) * N,
                             hipMemcpyDeviceToHost, s));
  }
  MDG_HIP(hipStreamSynchronize(s));
  return MDG_OK;
}

int box_read_scalars(Box* b, double* scal, hipStream_t s) {
  MDG_HIP(hipMemcpyAsync(scal, b->desc.scal, sizeof(double) * kScalCount, hipMemcpyDeviceToHost,
                         s));
  MDG_HIP(hipStreamSynchronize(s));
  return MDG_OK;
}

int box_read_error(Box* b, int* err, hipStream_t s) {
  MDG_HIP(hipMemcpyAsync(err, b->desc.err, sizeof(int), hipMemcpyDeviceToHost, s));
  MDG_HIP(hipStreamSynchronize(s));
  return MDG_OK;
}

int box_set_observe(Box* b, const ObserveOpts& o, hipStream_t s) {
  if (!b || o.t_control_step < 0 || o.moment_every < 0 || !(o.temp_K >= 0.0))
    return MDG_E_ARG;
  BoxDesc& D = b->desc;
  const int N = D.N;
  D.msd_on = o.msd ? 1 : 0;
  D.t_control_step = o.t_control_step;
  D.t_control_until = o.t_control_until;
  // tempK = temp / e2t (pt.c:99) with e2t = 2 / (3 N kB) (init.c:67,75)
  const double e2t = 2.0 / (3.0 * ((double)N * 1.380658e-5));
  D.t_control_ke = o.temp_K / e2t;
  D.moment_every = o.moment_every;
  b->desc_dirty = true;
  MDG_HIP(hipMemsetAsync(D.obs, 0, sizeof(double) * kObsLen, s));
  MDG_HIP(hipMemsetAsync(D.msd_old, 0, sizeof(double) * 3 * (size_t)N, s));
  MDG_HIP(hipMemsetAsync(D.msd_d, 0, sizeof(double) * 3 * (size_t)N, s));
  MDG_HIP(hipStreamSynchronize(s));
  return MDG_OK;
}

int box_read_observe(Box* b, double* obs, double* old, double* d, hipStream_t s) {
  const BoxDesc& D = b->desc;
  const size_t n3 = sizeof(double) * 3 * (size_t)D.N;
  if (obs) MDG_HIP(hipMemcpyAsync(obs, D.obs, sizeof(double) * kObsLen, hipMemcpyDeviceToHost, s));
  if (old) MDG_HIP(hipMemcpyAsync(old, D.msd_old, n3, hipMemcpyDeviceToHost, s));
  if (d) MDG_HIP(hipMemcpyAsync(d, D.msd_d, n3, hipMemcpyDeviceToHost, s));
  MDG_HIP(hipStreamSynchronize(s));
  return MDG_OK;
}

int box_write_observe(Box* b, const double* obs, const double* old, const double* d,
                      hipStream_t s) {
  const BoxDesc& D = b->desc;
  const size_t n3 = sizeof(double) * 3 * (size_t)D.N;
  if (obs) MDG_HIP(hipMemcpyAsync(D.obs, obs, sizeof(double) * kObsLen, hipMemcpyHostToDevice, s));
  if (old) MDG_HIP(hipMemcpyAsync(D.msd_old, old, n3, hipMemcpyHostToDevice, s));
  if (d) MDG_HIP(hipMemcpyAsync(D.msd_d, d, n3, hipMemcpyHostToDevice, s));
  MDG_HIP(hipStreamSynchronize(s));
  return MDG_OK;
}

// ------------------------------------------------------------------ sets
BoxSet::BoxSet(int precision, int real_mode) : precision_(precision), real_mode_(real_mode) {}

BoxSet::~BoxSet() {
  invalidate_graph();
  for (hipStream_t st : side_) {
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
  }
  for (hipEvent_t ev : ev_join_) (void)hipEventDestroy(ev);
  if (ev_fork_) (void)hipEventDestroy(ev_fork_);
  if (arena_) return;  // the caller owns the descriptors and work lists
  if (d_desc_) (void)hipFree(d_desc_);
  if (d_items_) (void)hipFree(d_items_);
  if (d_ids_) (void)hipFree(d_ids_);
}

void BoxSet::set_arena(void* base, size_t bytes) {
  arena_ = (char*)base;
  arena_bytes_ = bytes;
  serial_ = true;  // no side streams: creating them is what such callers avoid
}

// Side streams of a step's concurrent branches and their fork/join events,
// created on first use (default priority).  False if they cannot be made;
// the step then runs on one stream.
bool BoxSet::ensure_side_streams(int count) {
  if ((int)side_.size() >= count) return true;
  if (side_failed_) return false;
  if (!ev_fork_ && hipEventCreateWithFlags(&ev_fork_, hipEventDisableTiming) != hipSuccess) {
    ev_fork_ = nullptr;
    (void)hipGetLastError();
    side_failed_ = true;
    return false;
  }
  while ((int)side_.size() < count) {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
      if (st) (void)hipStreamDestroy(st);
      (void)hipGetLastError();
      side_failed_ = true;
      return false;
    }
    side_.push_back(st);
    ev_join_.push_back(ev);
  }
  return true;
}

void BoxSet::invalidate_graph() {
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  graph_exec_ = nullptr;
  graph_ = nullptr;
  graph_steps_ = 0;
}

// The workgroups of one launch go to the chip's XCDs in turn (workgroup w
// to XCD w % kXcds), and each XCD has its own L2.  This reorders a launch's
// items (a permutation: every item's work and output slot stay the same)
// so that the items of one box sit at positions of one residue mod kXcds,
// i.e. run on one XCD and share its L2.  Boxes are dealt, largest first, to
// the XCD with the fewest items; a box with more than 1/16 of all items is
// spread over every XCD instead, so that a lone large box still fills the
// chip.  Positions an XCD has no items left for take the longest queue's.
static void xcd_order(std::vector<Item>& v) {
  constexpr int kXcds = 8;
  if (v.size() < 4 * kXcds) return;
  std::vector<int> first;  // boxes in order of first appearance
  std::map<int, std::vector<Item>> by;
  for (const Item& it : v) {
    std::vector<Item>& q = by[it.box];
    if (q.empty()) first.push_back(it.box);
    q.push_back(it);
  }
  if (first.size() < 2) return;
  std::vector<std::vector<Item>> qx(kXcds);
  size_t next = 0;  // round-robin cursor for spread boxes
  std::vector<int> grouped;
  for (int b : first) {
    const std::vector<Item>& q = by[b];
    if (q.size() * 16 > v.size()) {
      for (const Item& it : q) qx[next++ % kXcds].push_back(it);
    } else {
      grouped.push_back(b);
    }
  }
  std::stable_sort(grouped.begin(), grouped.end(),
                   [&](int a, int b) { return by[a].size() > by[b].size(); });
  for (int b : grouped) {
    int x = 0;
    for (int y = 1; y < kXcds; ++y)
      if (qx[y].size() < qx[x].size()) x = y;
    qx[x].insert(qx[x].end(), by[b].begin(), by[b].end());
  }
  std::vector<Item> out;
  out.reserve(v.size());
  std::vector<size_t> pos(kXcds, 0);
  while (out.size() < v.size()) {
    for (int x = 0; x < kXcds && out.size() < v.size(); ++x) {
      int y = x;
      if (pos[x] == qx[x].size()) {  // borrow from the longest remaining queue
        size_t best = 0;
        for (int z = 0; z < kXcds; ++z)
          if (qx[z].size() - pos[z] > best) { best = qx[z].size() - pos[z]; y = z; }
      }
      out.push_back(qx[y][pos[y]++]);
    }
  }
  v.swap(out);
}

// Relative issue cost of one real-space item, for ordering the launch: a
// fixed part per item and per window of 128 partners (staging, screening,
// the partner flush), and the pair evaluations, about 1.4 popped
// candidates per accepted pair for the lanes' uneven candidate counts.
// The accepted fraction is the cutoff sphere's share of the box (all
// pairs) or of the 27-cell stencil (cells).
static double real_item_cost(const Box* B, const Item& it) {
  const BoxDesc& D = B->desc;
  const BoxPlan& P = B->plan;
  const int N = D.N, rs = std::max(1, P.tile.rsplit);
  const double rc = std::sqrt(D.rc2);
  const double sphere = 4.0 / 3.0 * M_PI * rc * rc * rc;
  const double V = D.L[0] * D.L[1] * D.L[2];
  double partners, dens, blocks = 1.0;
  if (P.all_pairs) {
    const int chunk = (N + rs - 1) / rs;
    partners = (double)std::max(0, std::min(chunk, N - (it.a + it.c * chunk)));
    dens = std::min(1.0, sphere / V);
  } else {
    const double occ = (double)N / std::max(1, P.ncell);
    const int n0 = 13 + it.c * 14 / rs, n1 = 13 + (it.c + 1) * 14 / rs;
    partners = occ * (n1 - n0);
    dens = std::min(1.0, sphere / (27.0 * V / std::max(1, P.ncell)));
    const int nblk = std::max(1, (int)std::ceil(occ / kRealThreads));
    const int hb = std::max(1, P.cell_hb);
    blocks = std::max(1, (nblk - it.d + hb - 1) / hb);
  }
  return blocks * (800.0 + std::ceil(partners / 128.0) * 880.0 + 148.0 * dens * partners);
}

// Relative cost of a mixed structure-factor item (k_sk_mfma): its chunk's
// ions times its hz tiles (one or two), plus a fixed part.
static double sk_item_cost(const Box* B, const Item& it) {
  const BoxPlan& P = B->plan;
  const int N = B->desc.N, chunk = (N + P.tile.nchunk - 1) / std::max(1, P.tile.nchunk);
  const int ions = std::max(0, std::min(chunk, N - it.a * chunk));
  const int nt = (P.hmax + 16) / 16;
  return (ions + 64.0) * std::min(2, nt - it.c);
}

// Relative cost of a mixed back-projection item (k_kforce_mfma): its tiles,
// each three MFMAs per K block of the box and an epilogue, plus a fixed
// part.
static double kf_item_cost(const Box* B, const Item& it) {
  return it.c * (3.0 * B->desc.kbk + 8.0) + 20.0;
}

// Launches over many boxes have uneven items (box sizes, cutoffs and k
// counts differ) and few per wave slot, so the order sets the kernel's
// tail: boxes are dealt to the eight XCDs by cost, heaviest first to the
// least loaded (a box heavier than 1/16 of the launch is spread over all of
// them), each XCD's items run heaviest first, and the launch order takes
// the XCD queues in turn, the order in which workgroups are assigned to
// XCDs.  A dispatch model of the 256-box batch's real space put the slowest
// XCD's end 1.16x above the ideal with the item-count balance of xcd_order
// and 1.03x with this one.  No result depends on the order.
template <typename Cost>
static void xcd_cost_order(std::vector<Item>& v, const std::vector<Box*>& boxes, Cost cost) {
  constexpr int kXcds = 8;
  if (v.size() < 4 * kXcds) return;
  std::vector<int> first;
  std::map<int, std::vector<std::pair<double, Item>>> by;
  double total = 0.0;
  for (const Item& it : v) {
    auto& q = by[it.box];
    if (q.empty()) first.push_back(it.box);
    const double c = cost(boxes[it.box], it);
    q.push_back({c, it});
    total += c;
  }
  std::vector<std::vector<std::pair<double, Item>>> qx(kXcds);
  std::vector<double> load(kXcds, 0.0);
  size_t next = 0;
  std::vector<std::pair<double, int>> grouped;  // (box cost, box)
  for (int b : first) {
    double cb = 0.0;
    for (const auto& e : by[b]) cb += e.first;
    if (cb * 16.0 > total) {
      for (const auto& e : by[b]) {
        const int x = (int)(next++ % kXcds);
        qx[x].push_back(e);
        load[x] += e.first;
      }
    } else {
      grouped.push_back({cb, b});
    }
  }
  std::stable_sort(grouped.begin(), grouped.end(),
                   [](const std::pair<double, int>& a, const std::pair<double, int>& b) {
                     return a.first > b.first;
                   });
  for (const auto& g : grouped) {
    int x = 0;
    for (int y = 1; y < kXcds; ++y)
      if (load[y] < load[x]) x = y;
    qx[x].insert(qx[x].end(), by[g.second].begin(), by[g.second].end());
    load[x] += g.first;
  }
  for (auto& q : qx)
    std::stable_sort(q.begin(), q.end(),
                     [](const std::pair<double, Item>& a, const std::pair<double, Item>& b) {
                       return a.first > b.first;
                     });
  std::vector<Item> out;
  out.reserve(v.size());
  std::vector<size_t> pos(kXcds, 0);
  while (out.size() < v.size()) {
    for (int x = 0; x < kXcds && out.size() < v.size(); ++x) {
      int y = x;
      if (pos[x] == qx[x].size()) {  // borrow from the longest remaining queue
        size_t best = 0;
        for (int z = 0; z < kXcds; ++z)
          if (qx[z].size() - pos[z] > best) { best = qx[z].size() - pos[z]; y = z; }
      }
      out.push_back(qx[y][pos[y]++].second);
    }
  }
  v.swap(out);
}

int BoxSet::set_boxes(const std::vector<Box*>& boxes) {
  invalidate_graph();
  boxes_ = boxes;
  const int nb = (int)boxes_.size();
  std::vector<int> order(nb);
  for (int k = 0; k < nb; ++k) order[k] = k;
  // big boxes' items first, so the longest tiles start earliest
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return boxes_[a]->desc.N > boxes_[b]->desc.N; });
  // Box groups whose step chains run concurrently: with enough boxes, two
  // groups of about equal work (ions x (k vectors + partners)), dealt
  // largest first to the lighter group.
  const int G = nb >= 2 * kStepGroups * 4 ? kStepGroups : 1;
  std::vector<int> group_of(nb, 0);
  {
    // group 0 starts its reciprocal chain at once, the others after their
    // real space; group0_share_ is group 0's share of the work
    std::vector<double> load(G, 0.0), share(G, G > 1 ? (1.0 - group0_share_) / (G - 1) : 1.0);
    share[0] = G > 1 ? group0_share_ : 1.0;
    for (int bi : order) {
      const BoxPlan& P = boxes_[bi]->plan;
      const double N = (double)boxes_[bi]->desc.N;
      const double w = N * ((P.kspace ? (double)P.nk : 0.0) + 2.0 * std::min(N, 200.0));
      int g = 0;
      for (int h = 1; h < G; ++h)
        if (load[h] / share[h] < load[g] / share[g]) g = h;
      group_of[bi] = g;
      load[g] += w;
    }
  }
  std::vector<std::vector<Item>> L[kLists];
  std::vector<std::vector<int>> I[kIdLists];
  for (auto& v : L) v.assign(G, {});
  for (auto& v : I) v.assign(G, {});
  skn_ = 1;
  int hmax = 0;
  const bool mixed = precision_ == MDG_MIXED;
  // the widest box fixes the back-projection's K blocks, hence its ion blocks
  int hmax_all = 0;
  for (const Box* B : boxes_)
    if (B->plan.kspace) hmax_all = std::max(hmax_all, B->plan.hmax);
  const int kf_ions = kforce_mfma_ions(kforce_hz_for(hmax_all) / 8);
  for (int bi : order) {
    const Box* B = boxes_[bi];
    const BoxPlan& P = B->plan;
    const int N = B->desc.N;
    const int rs = P.tile.rsplit;
    const int g = group_of[bi];
    std::vector<Item>& real = L[LReal][g];
    I[IAll][g].push_back(bi);
    const BoxDesc& D = B->desc;
    if (D.msd_on || D.t_control_step > 0 || D.moment_every > 0) I[IObs][g].push_back(bi);
    if (P.all_pairs) {
      int slot = 0;
      if (mixed) {
        // each pair once: a home block meets the ions from its own first
        // on, in chunks of ceil(N / rs) (k_real_win), so every item has
        // the same partner count at most
        const int chunk = ceil_div(N, rs);
        for (int i0 = 0; i0 < N; i0 += kRealThreads)
          for (int s = 0; s < rs && i0 + s * chunk < N; ++s) real.push_back({bi, i0, slot++, s, 0});
      } else {
        for (int i0 = 0; i0 < N; i0 += kRealThreads)
          for (int s = 0; s < rs; ++s) real.push_back({bi, i0, slot++, s, 0});
      }
    } else {
      I[ICell][g].push_back(bi);
      if (!cell_fused(N, P.ncell)) {  // built in four launches: per-ion and per-cell items
        for (int i0 = 0; i0 < N; i0 += kIonThreads) L[LCellIon][g].push_back({bi, i0, 0, 0, 0});
        for (int c = 0; c < P.ncell; ++c) L[LCellPlace][g].push_back({bi, c, 0, 0, 0});
      }
      int slot = 0;
      for (int c = 0; c < P.ncell; ++c)
        for (int hb = 0; hb < P.cell_hb; ++hb)
          for (int s = 0; s < rs; ++s) real.push_back({bi, c, slot++, s, hb});
    }
    for (int i0 = 0; i0 < N; i0 += kIonThreads) L[LIon][g].push_back({bi, i0, 0, 0, 0});
    if (P.kspace) {
      hmax = std::max(hmax, P.hmax);
      for (int i0 = 0; i0 < N; i0 += kIonThreads) L[LTab][g].push_back({bi, i0, 0, 0, 0});
      if (mixed) {
        // MFMA structure factor: 4 waves x 32 columns per item
        // (two hz tiles of 16 per item)
        const int cw = 32 * kSkmWaves;
        const int nt = (P.hmax + 16) / 16;
        for (int ch = 0; ch < P.tile.nchunk; ++ch)
          for (int c0 = 0; c0 < P.ncol; c0 += cw)
            for (int t0 = 0; t0 < nt; t0 += 2) L[LSk][g].push_back({bi, ch, c0, t0, 0});
        skn_ = std::max(skn_, P.hmax + 1);
      } else {
        for (int ch = 0; ch < P.tile.nchunk; ++ch)
          for (int s0 = 0; s0 < P.nseg; s0 += kSkThreads)
            L[LSk][g].push_back({bi, ch, s0, std::min(kSkThreads, P.nseg - s0), 0});
      }
      for (int k0 = 0; k0 < P.nk; k0 += kIonThreads) L[LFin][g].push_back({bi, k0, 0, 0, 0});
      const int gk = (int)P.kf_group_start.size() - 1;
      const int ion_step = mixed ? kf_ions : kKfThreads;
      for (int i0 = 0; i0 < N; i0 += ion_step)
        for (int q = 0; q < gk; ++q)
          L[LKf][g].push_back({bi, i0, P.kf_group_start[q],
                               P.kf_group_start[q + 1] - P.kf_group_start[q], q});
      if (mixed)
        for (int t = 0; t < (int)P.kftiles.size(); ++t)
          L[LPack][g].push_back({bi, t, 0, 0, 0});
    }
  }
  const int H = hmax + 1;
  hs_ = H | 1;                      // odd strides spread lanes over LDS banks
  hsw_ = (H + kSegLen - 1) / kSegLen * kSegLen;  // segment reads stay 16-byte aligned
  hz_ = kforce_hz_for(hmax);
  // each list: its groups' lists back to back (all_ spans them, grp_ each)
  std::vector<Item> items;
  std::vector<int> ids;
  all_ = Ranges();
  grp_.assign(G, Ranges());
  {
    // the real space runs as one launch over every box (launch_cells_real
    // with all_): one list, in cost order, held by group 0's range
    std::vector<Item>& r0 = L[LReal][0];
    for (int g = 1; g < G; ++g) {
      r0.insert(r0.end(), L[LReal][g].begin(), L[LReal][g].end());
      L[LReal][g].clear();
    }
    xcd_cost_order(r0, boxes_, real_item_cost);
  }
  for (int l = 0; l < kLists; ++l) {
    all_.off[l] = (int)items.size();
    for (int g = 0; g < G; ++g) {
      if (mixed && l == LSk)
        xcd_cost_order(L[l][g], boxes_, sk_item_cost);
      else if (mixed && l == LKf)
        xcd_cost_order(L[l][g], boxes_, kf_item_cost);
      else if (l != LReal)
        xcd_order(L[l][g]);
      grp_[g].off[l] = (int)items.size();
      grp_[g].n[l] = (int)L[l][g].size();
      items.insert(items.end(), L[l][g].begin(), L[l][g].end());
    }
    all_.n[l] = (int)items.size() - all_.off[l];
  }
  for (int l = 0; l < kIdLists; ++l) {
    all_.ioff[l] = (int)ids.size();
    for (int g = 0; g < G; ++g) {
      grp_[g].ioff[l] = (int)ids.size();
      grp_[g].in[l] = (int)I[l][g].size();
      ids.insert(ids.end(), I[l][g].begin(), I[l][g].end());
    }
    all_.in[l] = (int)ids.size() - all_.ioff[l];
  }
  if (arena_) {
    // descriptors, items and box ids back to back in the caller's memory
    size_t off = 0;
    const size_t o_desc = take(off, sizeof(BoxDesc) * std::max(nb, 1));
    const size_t o_items = take(off, sizeof(Item) * std::max<size_t>(items.size(), 1));
    const size_t o_ids = take(off, sizeof(int) * std::max<size_t>(ids.size(), 1));
    if (off > arena_bytes_) return MDG_E_BUDGET;
    d_desc_ = (BoxDesc*)(arena_ + o_desc);
    d_items_ = (Item*)(arena_ + o_items);
    d_ids_ = (int*)(arena_ + o_ids);
    desc_cap_ = std::max(nb, 1);
    items_cap_ = std::max<size_t>(items.size(), 1);
    ids_cap_ = std::max<size_t>(ids.size(), 1);
  }
  if (items.size() > items_cap_) {
    if (d_items_) (void)hipFree(d_items_);
    d_items_ = nullptr;
    items_cap_ = std::max<size_t>(items.size(), 2 * items_cap_);
    MDG_HIP(hipMalloc(&d_items_, sizeof(Item) * items_cap_));
  }
  if (ids.size() > ids_cap_) {
    if (d_ids_) (void)hipFree(d_ids_);
    d_ids_ = nullptr;
    ids_cap_ = std::max<size_t>(ids.size(), 2 * ids_cap_);
    MDG_HIP(hipMalloc(&d_ids_, sizeof(int) * ids_cap_));
  }
  if ((size_t)nb > desc_cap_) {
    if (d_desc_) (void)hipFree(d_desc_);
    d_desc_ = nullptr;
    desc_cap_ = std::max<size_t>(nb, 2 * desc_cap_);
    MDG_HIP(hipMalloc(&d_desc_, sizeof(BoxDesc) * desc_cap_));
  }
  if (!items.empty())
    MDG_HIP(hipMemcpy(d_items_, items.data(), sizeof(Item) * items.size(), hipMemcpyHostToDevice));
  if (!ids.empty())
    MDG_HIP(hipMemcpy(d_ids_, ids.data(), sizeof(int) * ids.size(), hipMemcpyHostToDevice));
  for (Box* b : boxes_) b->desc_dirty = true;
  return MDG_OK;
}

int BoxSet::sync_desc(hipStream_t s) {
  bool any = false;
  for (Box* b : boxes_) any |= b->desc_dirty;
  if (!any) return MDG_OK;
  std::vector<BoxDesc> d(boxes_.size());
  for (size_t k = 0; k < boxes_.size(); ++k) {
    d[k] = boxes_[k]->desc;
    boxes_[k]->desc_dirty = false;
  }
  MDG_HIP(hipMemcpyAsync(d_desc_, d.data(), sizeof(BoxDesc) * d.size(), hipMemcpyHostToDevice, s));
  MDG_HIP(hipStreamSynchronize(s));
  return MDG_OK;
}

// cell lists (cell-list boxes) and real-space forces of the boxes in r
// Kernels as profile() names them.  Mixed path: "phase_tables" is
// k_mixed_tables (the phase tables and both passes' MFMA B operands).
enum {
  KCell, KTables, KReal, KSk, KFin, KKf, KVv, KScal, KPack, KObs, kKernelNames
};
static const char* const kKernelName[kKernelNames] = {
    "cell_build", "phase_tables", "real_space", "sk_partial", "sk_finalize",
    "kforce",     "force_vv",     "box_scalars", "kf_pack",   "observe"};

thread_local LaunchEvents g_launch_events;

// Under profile(), the launch carries a start and a stop event that the
// kernel's own dispatch stamps (the same timestamps a kernel trace reads).
template <class F>
void BoxSet::timed(int kernel, hipStream_t, F&& launch) {
  if (!timer_) {
    launch();
    return;
  }
  hipEvent_t a = nullptr, b = nullptr;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
    if (a) (void)hipEventDestroy(a);
    launch();
    return;
  }
  g_launch_events = LaunchEvents{a, b, false};
  launch();
  const bool used = g_launch_events.used;
  g_launch_events = LaunchEvents{};
  if (!used) {  // nothing to launch (an empty work list)
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return;
  }
  timer_->ev.push_back(a);
  timer_->ev.push_back(b);
  timer_->kernel.push_back(kernel);
}

void BoxSet::launch_cells_real(hipStream_t s, int mode, const Ranges& r) {
  const bool analytic = real_mode_ == MDG_ANALYTIC;
  if (r.in[ICell])
    timed(KCell, s, [&] {
      launch_cell_build(d_desc_, d_ids_ + r.ioff[ICell], r.in[ICell],
                        d_items_ + r.off[LCellIon], r.n[LCellIon], d_items_ + r.off[LCellPlace],
                        r.n[LCellPlace], s);
    });
  timed(KReal, s, [&] {
    if (precision_ == MDG_MIXED)
      launch_real32(analytic, d_desc_, d_items_ + r.off[LReal], r.n[LReal], mode, s);
    else
      launch_real64(analytic, d_desc_, d_items_ + r.off[LReal], r.n[LReal], mode, s);
  });
}

// phase tables, partial structure factors, finalisation, back-projection
void BoxSet::launch_recip(hipStream_t s, const Ranges& r) {
  const bool mixed = precision_ == MDG_MIXED;
  const Item* it = d_items_;
  if (mixed) {
    timed(KTables, s, [&] { launch_mixed_tables(d_desc_, it + r.off[LTab], r.n[LTab], s); });
    timed(KSk, s, [&] { launch_sk_mfma(skn_, d_desc_, it + r.off[LSk], r.n[LSk], s); });
    timed(KFin, s, [&] { launch_sk_finalize(d_desc_, it + r.off[LFin], r.n[LFin], s); });
    timed(KPack, s, [&] { launch_kf_pack(d_desc_, it + r.off[LPack], r.n[LPack], s); });
    timed(KKf, s,
          [&] { launch_kforce_mfma(hz_ / 8, d_desc_, it + r.off[LKf], r.n[LKf], s); });
  } else {
    timed(KTables, s,
          [&] { launch_phase_tables(false, d_desc_, it + r.off[LTab], r.n[LTab], s); });
    timed(KSk, s, [&] { launch_sk64(d_desc_, it + r.off[LSk], r.n[LSk], hs_, hsw_, s); });
    timed(KFin, s, [&] { launch_sk_finalize(d_desc_, it + r.off[LFin], r.n[LFin], s); });
    timed(KKf, s, [&] { launch_kforce64(hz_, d_desc_, it + r.off[LKf], r.n[LKf], s); });
  }
}

void BoxSet::enqueue_forces_parts(hipStream_t s, bool real, int real_write_mode, bool recip) {
  if (real) launch_cells_real(s, real_write_mode, all_);
  if (recip && all_.n[LSk] > 0) {
    launch_recip(s, all_);
    launch_force_vv(false, d_desc_, d_items_ + all_.off[LIon], all_.n[LIon], s);
  }
}

void BoxSet::enqueue_force_terms(hipStream_t s, int terms) {
  if (terms & 1) launch_cells_real(s, 0, all_);
  if ((terms & 2) && all_.n[LSk] > 0) launch_recip(s, all_);
  launch_force_vv(false, d_desc_, d_items_ + all_.off[LIon], all_.n[LIon], s, terms);
}

// Within a step, every box's real-space branch (cell lists, pair forces)
// and reciprocal chain read the same positions and write disjoint buffers
// (f/f0 or the split partials; tables, S(k), the back-projection
// partials), and different boxes share nothing.  So they run on several
// streams between a fork event and join events: the kernels are
// latency-bound and share the chip better than they run back to back.
// The real space of every box runs on side stream 0, group 0's reciprocal
// chain on the caller's stream from the start, and group 1's reciprocal
// chain on side stream 1 once the real space is done, so that different
// kinds of kernel overlap.  Measured on the 256-box batch with one tuning
// table: 1.22 ms per step with each group's real space beside or before
// its own chain, 1.135 ms with this order, and worse with any split other
// than an equal one.  Captured into the step graph, the fork and the joins
// become graph edges.
void BoxSet::enqueue_step(hipStream_t s, bool vv) {
  const int G = (int)grp_.size();
  if (!serial_ && all_.n[LSk] > 0 && ensure_side_streams(G)) {
    (void)hipEventRecord(ev_fork_, s);
    for (int k = 0; k < G; ++k) (void)hipStreamWaitEvent(side_[k], ev_fork_, 0);
    launch_cells_real(side_[0], 0, all_);
    launch_recip(s, grp_[0]);
    if (G > 1) {  // the other groups' chains start when the real space is done
      (void)hipEventRecord(ev_join_[0], side_[0]);
      for (int g = 1; g < G; ++g) {
        (void)hipStreamWaitEvent(side_[g], ev_join_[0], 0);
        launch_recip(side_[g], grp_[g]);
      }
    }
    for (int k = 0; k < G; ++k) {
      (void)hipEventRecord(ev_join_[k], side_[k]);
      (void)hipStreamWaitEvent(s, ev_join_[k], 0);
    }
  } else {
    launch_cells_real(s, 0, all_);
    if (all_.n[LSk] > 0) launch_recip(s, all_);
  }
  // the force assembly also forms the scalars' per-item partials
  timed(KVv, s, [&] {
    launch_force_vv(vv, d_desc_, d_items_ + all_.off[LIon], all_.n[LIon], s, 3, true);
  });
  timed(KScal, s, [&] {
    launch_box_scalars(d_desc_, d_ids_ + all_.ioff[IAll], all_.in[IAll], d_items_ + all_.off[LIon],
                       all_.n[LIon], vv ? 1 : 0, s, false);
  });
  // mdc's calc_msd, control_temp and moment_correction of the boxes that
  // have them armed (main.c:44-60)
  if (vv && all_.in[IObs])
    timed(KObs, s, [&] {
      launch_observe(d_desc_, d_ids_ + all_.ioff[IObs], all_.in[IObs], -1, 7, s);
    });
}

void BoxSet::enqueue_scalars(hipStream_t s, bool count_step) {
  launch_box_scalars(d_desc_, d_ids_ + all_.ioff[IAll], all_.in[IAll], d_items_ + all_.off[LIon],
                     all_.n[LIon], count_step ? 1 : 0, s);
}
void BoxSet::enqueue_clear_foc(hipStream_t s) {
  launch_clear_foc(d_desc_, d_items_ + all_.off[LIon], all_.n[LIon], s);
}
void BoxSet::enqueue_next_v(hipStream_t s) {
  launch_next_v(d_desc_, d_items_ + all_.off[LIon], all_.n[LIon], s);
}
void BoxSet::enqueue_next_r(hipStream_t s) {
  launch_next_r(d_desc_, d_items_ + all_.off[LIon], all_.n[LIon], s);
}
void BoxSet::enqueue_moment_correction(hipStream_t s) {
  launch_moment_correction(d_desc_, d_ids_ + all_.ioff[IAll], all_.in[IAll], s);
}
void BoxSet::enqueue_observe(hipStream_t s, long long step, int phases) {
  launch_observe(d_desc_, d_ids_ + all_.ioff[IAll], all_.in[IAll], step, phases, s);
}
void BoxSet::enqueue_snapshot(hipStream_t s, const long long* off, double* dst) {
  launch_snapshot(d_desc_, d_items_ + all_.off[LIon], all_.n[LIon], off, dst, s);
}
void BoxSet::enqueue_scale_v(hipStream_t s, double scale) {
  launch_scale_v(d_desc_, d_items_ + all_.off[LIon], all_.n[LIon], scale, s);
}

// Captures the kGraphSteps-step graph, instantiates it and uploads it to
// the device; no step is executed (the kernels only run when it is launched).
int BoxSet::prepare(hipStream_t s) {
  if (boxes_.empty()) return MDG_OK;
  int rc = sync_desc(s);
  if (rc != MDG_OK) return rc;
  if (graph_exec_ && graph_steps_ == kGraphSteps) return MDG_OK;
  invalidate_graph();
  MDG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  for (int k = 0; k < kGraphSteps; ++k) enqueue_step(s, true);
  MDG_HIP(hipStreamEndCapture(s, &graph_));
  MDG_HIP(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
  graph_steps_ = kGraphSteps;
  MDG_HIP(hipGraphUpload(graph_exec_, s));
  MDG_HIP(hipStreamSynchronize(s));
  return MDG_OK;
}

int BoxSet::step(int nsteps, hipStream_t s, bool use_graph) {
  if (nsteps <= 0 || boxes_.empty()) return MDG_OK;
  int rc = sync_desc(s);
  if (rc != MDG_OK) return rc;
  // Whole groups of 8 steps replay one captured graph; a shorter remainder
  // (or a call of fewer than 8 steps) is launched directly, so callers that
  // step by varying amounts do not re-capture.  use_graph = false launches
  // every step directly (no capture for a set that is about to change).
  const int G = kGraphSteps;
  if (nsteps < G || (!use_graph && !graph_exec_)) {
    for (int k = 0; k < nsteps; ++k) enqueue_step(s, true);
    MDG_HIP(hipGetLastError());
    return MDG_OK;
  }
  if (!graph_exec_ || graph_steps_ != G) {
    rc = prepare(s);
    if (rc != MDG_OK) return rc;
  }
  int left = nsteps;
  while (left >= G) {
    MDG_HIP(hipGraphLaunch(graph_exec_, s));
    left -= G;
  }
  for (; left > 0; --left) enqueue_step(s, true);
  MDG_HIP(hipGetLastError());
  return MDG_OK;
}

int BoxSet::profile(int nsteps, hipStream_t s, bool as_step, std::vector<std::string>* names,
                    std::vector<double>* ms, std::vector<int>* launches) {
  if (nsteps <= 0) return MDG_E_ARG;
  int rc = sync_desc(s);
  if (rc != MDG_OK) return rc;
  LaunchTimer t;
  timer_ = &t;
  const bool was_serial = serial_;
  serial_ = was_serial || !as_step;
  for (int st = 0; st < nsteps; ++st) enqueue_step(s, true);
  timer_ = nullptr;
  serial_ = was_serial;
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  std::vector<double> tot(kKernelNames, 0.0);
  std::vector<int> cnt(kKernelNames, 0);
  for (size_t q = 0; q < t.kernel.size() && e == hipSuccess; ++q) {
    float x = 0.f;
    e = hipEventElapsedTime(&x, t.ev[2 * q], t.ev[2 * q + 1]);
    tot[t.kernel[q]] += x;
    cnt[t.kernel[q]] += 1;
  }
  for (auto& x : t.ev) (void)hipEventDestroy(x);
  MDG_HIP(e);
  names->clear();
  ms->clear();
  launches->clear();
  for (int k = 0; k < kKernelNames; ++k)
    if (cnt[k]) {
      names->push_back(kKernelName[k]);
      ms->push_back(tot[k] / nsteps);
      launches->push_back(cnt[k] / nsteps);
    }
  return MDG_OK;
}

}  // namespace mdg
