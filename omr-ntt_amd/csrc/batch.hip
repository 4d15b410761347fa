// batch.hip -- batched transforms of vectors of uneven power-of-two sizes.
//
// A plan splits every vector into Stockham passes of radix 2^4 .. 2^12 and
// lists the tiles (8192 or 4096 elements) of pass i of all vectors as the tasks of
// launch i of the persistent kernel k_batch_tiles (batch_kernel.h).  Pass
// boundaries are kernel boundaries: a pass reads what other workgroups (on
// other XCDs, with their own L2) wrote in the previous pass, which coarse-
// grained device memory makes visible only across launches.  The launches
// are captured once into a hipGraph and replayed.  The plain form (one
// launch sequence per vector) is kept for comparison.
#include <algorithm>
#include <new>
#include <cstdlib>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "batch_kernel.h"
#include "dispatch.h"
#include "internal.h"

using namespace omr;

struct omr_batch_plan {
  omr_ctx *ctx = nullptr;
  int field = 0;
  std::vector<omr_batch_item> items;
  std::vector<std::vector<int>> sched;  // per item pass radices
  uint32_t nlaunch = 0;
  std::vector<uint32_t> ntasks, grid;
  std::vector<void *> d_tasks, d_vecs;
  void *d_tsets = nullptr;
  // twiddle tables whose device pointers d_tsets holds: pinned in the
  // context's cache for the plan's lifetime
  std::vector<std::tuple<int, uint32_t, uint64_t>> pinned;
  DevBuf scratch[2];
  uint64_t total_elems = 0;
  // graph cache
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const void *g_in = nullptr;
  void *g_out = nullptr;
  // grouped form: the vectors of one (size, direction), at most 65535 a
  // group, run through the pass engine with per-vector offset tables
  struct Group {
    uint32_t log_n, inverse, count;
    uint64_t root;         // the transform's root (inverse roots for inverse groups)
    uint64_t *d_off;       // count input offsets, then count output offsets (elements)
    uint32_t npass;        // passes of the group's schedule
    uint32_t lane;         // stream the group runs on
    omr::Schedule sch;     // fixed when the plan is made (the lanes' scratch is sized for it)
  };
  std::vector<Group> groups;
  // The groups are independent: they run on up to kLanes (default 2) streams at once
  // (forked from and joined back to the caller's stream), each lane with
  // the dense intermediates of its largest multi-pass group.
  static constexpr uint32_t kLanes = 4;
  uint32_t nlanes = 0;
  omr::DevBuf lscratch[kLanes][2];
  hipStream_t lane_stream[kLanes] = {};
  hipEvent_t fork_ev = nullptr;
  hipEvent_t join_ev[kLanes] = {};
  hipGraph_t ggraph = nullptr;
  hipGraphExec_t gexec = nullptr;
  const void *gg_in = nullptr;
  void *gg_out = nullptr;
};

namespace omr {

static hipError_t launch_batch(int f, dim3 grid, hipStream_t s, const BatchTask *t, uint32_t nt,
                               const BatchVec *v, const BatchTables *tb, BatchBases b) {
  switch (f) {
    case kBabyBear:
      hipLaunchKernelGGL(k_batch_tiles<BabyBear>, grid, dim3(batch_tile_threads(f)), 0, s, t, nt, v, tb, b);
      break;
    case kKoalaBear:
      hipLaunchKernelGGL(k_batch_tiles<KoalaBear>, grid, dim3(batch_tile_threads(f)), 0, s, t, nt, v, tb, b);
      break;
    case kGoldilocks:
      hipLaunchKernelGGL(k_batch_tiles<Goldilocks>, grid, dim3(batch_tile_threads(f)), 0, s, t, nt, v, tb, b);
      break;
    case kTeddyBear:
      hipLaunchKernelGGL(k_batch_tiles<TeddyBear>, grid, dim3(batch_tile_threads(f)), 0, s, t, nt, v, tb, b);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Largest pass radix the batch planner uses for multi-pass vectors; 0 =
// automatic.  Goldilocks 2^10 (8 columns of its 8192-element tiles, 64-byte
// runs; vectors up to 2^20 take two passes), BabyBear / KoalaBear 2^9 (16
// columns, 64-byte runs), TeddyBear 2^9 (8 columns of its 4096-element
// tiles).  Measured on 4096 vectors of log n ~ U{10..20}: Goldilocks 11.7 ms
// at 2^9 against 10.5 ms at 2^10, BabyBear 8.4 ms at 2^9 against 8.9 ms at
// 2^10.
static int g_batch_max_log_radix = 0;
static int batch_cap(int field) {
  return g_batch_max_log_radix ? g_batch_max_log_radix : field == kGoldilocks ? 10 : 9;
}

// Pass radices of one vector of size 2^logn inside a batch, for a radix cap
// of 2^cap: one pass up to 2^12, else ceil(logn / cap) balanced passes.
// k_batch_tiles runs a pass of radix 2^4 .. 2^12 (register radix 16), and
// with cap >= 5 every balanced part of a log n >= 13 is >= 4: with
// parts = ceil(logn / 5) <= (logn + 4) / 5, logn / parts >= 5 logn / (logn + 4) >= 4
// for logn >= 16, and logn = 13..15 split into three parts of 4 or 5.
static bool batch_schedule(uint32_t logn, int cap, std::vector<int> &out) {
  out.clear();
  if (logn <= 12) {
    out.push_back((int)logn);
    return logn >= 4;
  }
  if (cap < 5 || cap > 12) return false;
  const int parts = ((int)logn + cap - 1) / cap;
  const int base = (int)logn / parts, rem = (int)logn % parts;
  for (int i = 0; i < parts; i++) out.push_back(base + (i < rem ? 1 : 0));
  for (int r : out)
    if (r < 4 || r > 12) return false;
  return true;
}

}  // namespace omr

extern "C" omr_status omr_batch_set_max_log_radix(int log_radix) try {
  // below 2^5 a balanced split can leave parts of 2^3, which the batch tile
  // kernel (register radix 16) cannot run; 0 restores the automatic cap
  if (log_radix != 0 && (log_radix < 5 || log_radix > 12)) return OMR_E_SCHEDULE;
  g_batch_max_log_radix = log_radix;
  return OMR_OK;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

extern "C" omr_status omr_batch_vector_schedule(uint32_t log_n, int max_log_radix, int *log_radix,
                                                int *npass) try {
  if (!log_radix || !npass) return OMR_E_NULL;
  if (log_n < 4 || log_n > 36) return OMR_E_SIZE;
  std::vector<int> sc;
  if (!batch_schedule(log_n, max_log_radix == 0 ? batch_cap(kGoldilocks) : max_log_radix, sc) || sc.size() > 9)
    return OMR_E_SCHEDULE;
  *npass = (int)sc.size();
  for (size_t i = 0; i < sc.size(); i++) log_radix[i] = sc[i];
  return OMR_OK;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

extern "C" omr_status omr_batch_plan_destroy(omr_batch_plan *p) try {
  if (!p) return OMR_E_NULL;
  (void)hipDeviceSynchronize();
  if (p->ctx && !p->pinned.empty()) {
    std::lock_guard<std::mutex> lk(p->ctx->mu);
    for (auto &k : p->pinned) pin_tables(p->ctx, std::get<0>(k), std::get<1>(k), std::get<2>(k), -1);
  }
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
  if (p->ggraph) (void)hipGraphDestroy(p->ggraph);
  if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
  for (auto &g : p->groups) (void)hipFree(g.d_off);
  for (uint32_t l = 0; l < omr_batch_plan::kLanes; l++) {
    free_buf(p->lscratch[l][0]);
    free_buf(p->lscratch[l][1]);
    if (p->lane_stream[l]) (void)hipStreamDestroy(p->lane_stream[l]);
    if (p->join_ev[l]) (void)hipEventDestroy(p->join_ev[l]);
  }
  if (p->fork_ev) (void)hipEventDestroy(p->fork_ev);
  for (void *q : p->d_tasks) (void)hipFree(q);
  for (void *q : p->d_vecs) (void)hipFree(q);
  if (p->d_tsets) (void)hipFree(p->d_tsets);
  free_buf(p->scratch[0]);
  free_buf(p->scratch[1]);
  delete p;
  return OMR_OK;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

// Schedule of a group: a group of 32 MB or more takes the tuned table's
// batched slot or, if that is empty, the default (coalescing-first)
// schedule, never the single-vector entry, whose small cache-resident passes
// lose on such batches (4096 Goldilocks vectors, log n ~ U{10..20}: 10.5 ms against 13.5 ms).
static void group_schedule(int f, uint32_t log_n, uint32_t count, Schedule &s) {
  if (!tuned_schedule_for(f, log_n, count, s)) default_schedule(f, log_n, s);
}

// Spreads the groups over the lanes, largest work (elements x passes) first
// onto the least-loaded lane, and gives each lane its scratch and stream.
static omr_status assign_lanes(omr_batch_plan *p, int es) {
  auto &gs = p->groups;
  if (gs.empty()) return OMR_OK;
  // 2 lanes by default: batches of 4096 vectors with log n ~ U{10..20} gain 3 %
  // over one lane, and 4 lanes no more; OMR_BATCH_LANES=1..4 overrides
  // (read per plan)
  uint32_t lanes = 2;
  if (const char *e = std::getenv("OMR_BATCH_LANES"))
    if (e[0] >= '1' && e[0] <= '0' + (int)omr_batch_plan::kLanes && e[1] == 0) lanes = (uint32_t)(e[0] - '0');
  p->nlanes = (uint32_t)std::min<size_t>(lanes, gs.size());
  std::vector<size_t> order(gs.size());
  for (size_t i = 0; i < order.size(); i++) order[i] = i;
  auto work = [&](size_t i) { return ((uint64_t)gs[i].count << gs[i].log_n) * gs[i].npass; };
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return work(a) > work(b); });
  uint64_t load[omr_batch_plan::kLanes] = {}, smax[omr_batch_plan::kLanes] = {};
  for (size_t i : order) {
    uint32_t l = 0;
    for (uint32_t k = 1; k < p->nlanes; k++)
      if (load[k] < load[l]) l = k;
    gs[i].lane = l;
    load[l] += work(i);
    if (gs[i].npass >= 2) smax[l] = std::max<uint64_t>(smax[l], (uint64_t)gs[i].count << gs[i].log_n);
  }
  for (uint32_t l = 0; l < p->nlanes; l++) {
    omr_status st;
    if (smax[l] && ((st = ensure_buf(p->lscratch[l][0], smax[l] * es)) != OMR_OK ||
                    (st = ensure_buf(p->lscratch[l][1], smax[l] * es)) != OMR_OK))
      return st;
    if (p->nlanes > 1 && (hipStreamCreateWithFlags(&p->lane_stream[l], hipStreamNonBlocking) != hipSuccess ||
                          hipEventCreateWithFlags(&p->join_ev[l], hipEventDisableTiming) != hipSuccess))
      return OMR_E_DEVICE;
  }
  if (p->nlanes > 1 && hipEventCreateWithFlags(&p->fork_ev, hipEventDisableTiming) != hipSuccess)
    return OMR_E_DEVICE;
  return OMR_OK;
}

extern "C" omr_status omr_batch_plan_create(omr_ctx *ctx, int field, const omr_batch_item *items,
                                            uint32_t count, omr_batch_plan **out) try {
  if (!field_ok(field)) return OMR_E_FIELD;
  if (!items || !out) return OMR_E_NULL;
  if (count == 0) return OMR_E_ARG;
  int two_adicity = 0;
  omr_field_info(field, nullptr, nullptr, &two_adicity);
  for (uint32_t i = 0; i < count; i++) {
    if (items[i].log_n < 4 || (int)items[i].log_n > two_adicity || items[i].log_n > 36)
      return items[i].log_n > 36 || (int)items[i].log_n > two_adicity ? OMR_E_TWO_ADICITY : OMR_E_SIZE;
    if (items[i].inverse > 1) return OMR_E_ARG;
  }
  if (!ctx) return OMR_E_NULL;
  *out = nullptr;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess) return OMR_E_DEVICE;
  omr_batch_plan *p = new (std::nothrow) omr_batch_plan();
  if (!p) return OMR_E_ALLOC;
  p->ctx = ctx;
  p->field = field;
  p->items.assign(items, items + count);
  p->sched.resize(count);
  const int es = field_bytes(field);

  // twiddle-table sets, one per (log n, direction)
  std::map<std::pair<uint32_t, uint32_t>, uint16_t> tset_of;
  std::vector<BatchTables> tsets;
  uint64_t scratch_elems = 0;
  std::vector<uint64_t> scratch_off(count, 0);
  uint32_t maxpass = 0;
  omr_status st = OMR_OK;
  const int cap = batch_cap(field);
  for (uint32_t i = 0; i < count; i++) {
    const uint32_t logn = items[i].log_n;
    if (!batch_schedule(logn, cap, p->sched[i])) {
      st = OMR_E_SCHEDULE;
      break;
    }
    maxpass = std::max<uint32_t>(maxpass, (uint32_t)p->sched[i].size());
    p->total_elems += 1ull << logn;
    if (p->sched[i].size() >= 2) {
      scratch_off[i] = scratch_elems;
      scratch_elems += 1ull << logn;
    }
    auto key = std::make_pair(logn, items[i].inverse);
    if (!tset_of.count(key)) {
      uint64_t w;
      omr_root_of_unity(field, 1ull << logn, &w);
      BatchTables bt{};
      if (items[i].inverse) {
        w = finv_host(field, w);
        bt.scale_w = to_w_host(field, finv_host(field, (1ull << logn) % field_modulus(field)));
      }
      const TwTables *tab;
      if ((st = get_tables(ctx, field, logn, w, &tab)) != OMR_OK) break;
      pin_tables(ctx, field, logn, w, +1);
      p->pinned.emplace_back(field, logn, w);
      bt.lo = tab->lo;
      bt.hi = tab->hi;
      bt.pyr = tab->pyr;
      bt.lo_bits = tab->lo_bits;
      bt.log_rmax = tab->log_rmax;
      bt.mid = tab->mid;
      bt.log_mid = tab->log_mid;  // = log_rmax when there is no mid table
      bt.unit = (uint32_t)tab->unit();
      tset_of[key] = (uint16_t)tsets.size();
      tsets.push_back(bt);
    }
  }
  if (st == OMR_OK && tsets.size() > 65535) st = OMR_E_ARG;
  if (st == OMR_OK && scratch_elems) {
    if ((st = ensure_buf(p->scratch[0], scratch_elems * es)) == OMR_OK)
      st = ensure_buf(p->scratch[1], scratch_elems * es);
  }
  if (st == OMR_OK && hipMalloc(&p->d_tsets, tsets.size() * sizeof(BatchTables)) != hipSuccess) st = OMR_E_ALLOC;
  if (st == OMR_OK && hipMemcpy(p->d_tsets, tsets.data(), tsets.size() * sizeof(BatchTables),
                                hipMemcpyHostToDevice) != hipSuccess)
    st = OMR_E_DEVICE;

  // vectors ordered by size, largest first (the tasks inherit this order)
  std::vector<uint32_t> order(count);
  for (uint32_t i = 0; i < count; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](uint32_t a, uint32_t b) { return items[a].log_n > items[b].log_n; });

  int cus = 256;
  {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess) cus = prop.multiProcessorCount;
  }
  int occ = 0;
  const void *kern = field == kBabyBear    ? (const void *)k_batch_tiles<BabyBear>
                     : field == kKoalaBear ? (const void *)k_batch_tiles<KoalaBear>
                     : field == kTeddyBear ? (const void *)k_batch_tiles<TeddyBear>
                                           : (const void *)k_batch_tiles<Goldilocks>;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, batch_tile_threads(field), 0) != hipSuccess ||
      occ < 1)
    occ = 1;

  for (uint32_t pass = 0; st == OMR_OK && pass < maxpass; pass++) {
    std::vector<BatchTask> tasks;
    std::vector<BatchVec> vecs;
    // vectors taking part in this pass, grouped by (log n, direction) so that
    // small whole vectors can share a tile
    uint32_t i0 = 0;
    while (i0 < count) {
      uint32_t i1 = i0;
      const uint32_t ln = items[order[i0]].log_n;
      while (i1 < count && items[order[i1]].log_n == ln) i1++;
      for (uint32_t dir = 0; dir < 2; dir++) {
        std::vector<uint32_t> grp;
        for (uint32_t k = i0; k < i1; k++)
          if (items[order[k]].inverse == dir && p->sched[order[k]].size() > pass) grp.push_back(order[k]);
        if (grp.empty()) continue;
        const std::vector<int> &sc = p->sched[grp[0]];
        const int np = (int)sc.size();
        int consumed = 0;
        for (uint32_t q = 0; q <= pass; q++) consumed += sc[q];
        const int logR = sc[pass];
        const int logS = (int)ln - consumed;
        const int logC = batch_log_tile(field) - logR;
        const int logCols = (int)ln - logR;
        const uint16_t ts = tset_of[std::make_pair(ln, dir)];
        uint8_t flags = 0;
        if (pass > 0) flags |= kTwiddle;
        if (pass == (uint32_t)np - 1 && dir) flags |= kScale;
        const uint32_t vbase = (uint32_t)vecs.size();
        for (uint32_t v : grp) {
          BatchVec bv{};
          if (pass == 0) {
            bv.src_buf = 0;
            bv.src_off = items[v].in_offset;
          } else {
            bv.src_buf = (pass % 2 == 1) ? 2 : 3;
            bv.src_off = scratch_off[v];
          }
          if (pass == (uint32_t)np - 1) {
            bv.dst_buf = 1;
            bv.dst_off = items[v].out_offset;
          } else {
            bv.dst_buf = (pass % 2 == 0) ? 2 : 3;
            bv.dst_off = scratch_off[v];
          }
          vecs.push_back(bv);
        }
        if (logCols >= logC) {  // tiles within one vector
          const uint32_t ntile = 1u << (logCols - logC);
          for (uint32_t gi = 0; gi < grp.size(); gi++)
            for (uint32_t tl = 0; tl < ntile; tl++) {
              BatchTask t{};
              t.vec0 = vbase + gi;
              t.tile = tl;
              t.logN = (uint8_t)ln;
              t.logR = (uint8_t)logR;
              t.logS = (uint8_t)logS;
              t.flags = flags;
              t.tset = ts;
              t.nvec = 1;
              tasks.push_back(t);
            }
        } else {  // several whole vectors per tile
          const uint32_t per = 1u << (logC - logCols);
          for (uint32_t gi = 0; gi < grp.size(); gi += per) {
            BatchTask t{};
            t.vec0 = vbase + gi;
            t.tile = 0;
            t.logN = (uint8_t)ln;
            t.logR = (uint8_t)logR;
            t.logS = (uint8_t)logS;
            t.flags = flags;
            t.tset = ts;
            t.nvec = (uint16_t)std::min<uint32_t>(per, (uint32_t)grp.size() - gi);
            tasks.push_back(t);
          }
        }
      }
      i0 = i1;
    }
    void *dt = nullptr, *dv = nullptr;
    if (hipMalloc(&dt, std::max<size_t>(1, tasks.size()) * sizeof(BatchTask)) != hipSuccess ||
        hipMalloc(&dv, std::max<size_t>(1, vecs.size()) * sizeof(BatchVec)) != hipSuccess) {
      st = OMR_E_ALLOC;
      if (dt) (void)hipFree(dt);
      break;
    }
    if (hipMemcpy(dt, tasks.data(), tasks.size() * sizeof(BatchTask), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dv, vecs.data(), vecs.size() * sizeof(BatchVec), hipMemcpyHostToDevice) != hipSuccess)
      st = OMR_E_DEVICE;
    p->d_tasks.push_back(dt);
    p->d_vecs.push_back(dv);
    p->ntasks.push_back((uint32_t)tasks.size());
    p->grid.push_back((uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(tasks.size(), (uint64_t)cus * occ)));
  }
  p->nlaunch = (uint32_t)p->ntasks.size();

  // grouped form: one group per (size, direction), chunks of 65535 vectors
  if (st == OMR_OK) {
    std::map<std::pair<uint32_t, uint32_t>, std::vector<uint32_t>> by_shape;
    for (uint32_t i = 0; i < count; i++) by_shape[{items[i].log_n, items[i].inverse}].push_back(i);
    for (auto it = by_shape.rbegin(); it != by_shape.rend() && st == OMR_OK; ++it) {  // largest first
      const std::vector<uint32_t> &idx = it->second;
      const uint32_t logn = it->first.first, inv = it->first.second;
      uint64_t w;
      omr_root_of_unity(field, 1ull << logn, &w);
      if (inv) w = finv_host(field, w);
      for (size_t c0 = 0; c0 < idx.size() && st == OMR_OK; c0 += 65535) {
        const uint32_t cnt = (uint32_t)std::min<size_t>(65535, idx.size() - c0);
        Schedule sch;
        group_schedule(field, logn, cnt, sch);
        std::vector<uint64_t> off(2 * (size_t)cnt);
        for (uint32_t j = 0; j < cnt; j++) {
          off[j] = items[idx[c0 + j]].in_offset;
          off[cnt + j] = items[idx[c0 + j]].out_offset;
        }
        omr_batch_plan::Group g{logn, inv, cnt, w, nullptr, (uint32_t)sch.logR.size(), 0, sch};
        if (hipMalloc((void **)&g.d_off, off.size() * sizeof(uint64_t)) != hipSuccess) {
          (void)hipGetLastError();
          st = OMR_E_ALLOC;
          break;
        }
        p->groups.push_back(g);
        if (hipMemcpy(g.d_off, off.data(), off.size() * sizeof(uint64_t), hipMemcpyHostToDevice) != hipSuccess)
          st = OMR_E_DEVICE;
      }
    }
    if (st == OMR_OK) st = assign_lanes(p, es);
  }
  if (st != OMR_OK) {
    // the context lock is held here: unpin directly, then free the rest
    for (auto &k : p->pinned) pin_tables(ctx, std::get<0>(k), std::get<1>(k), std::get<2>(k), -1);
    p->pinned.clear();
    omr_batch_plan_destroy(p);
    return st;
  }
  *out = p;
  return OMR_OK;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

extern "C" omr_status omr_batch_plan_info(const omr_batch_plan *p, uint64_t *total_elems,
                                          uint32_t *nlaunch, uint32_t *ntasks) try {
  if (!p) return OMR_E_NULL;
  if (total_elems) *total_elems = p->total_elems;
  if (nlaunch) *nlaunch = p->nlaunch;
  if (ntasks) {
    uint32_t s = 0;
    for (uint32_t n : p->ntasks) s += n;
    *ntasks = s;
  }
  return OMR_OK;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

static omr_status batch_launches(omr_batch_plan *p, const void *d_in, void *d_out, hipStream_t s) {
  BatchBases b;
  b.buf[0] = d_in;
  b.buf[1] = d_out;
  b.buf[2] = p->scratch[0].p;
  b.buf[3] = p->scratch[1].p;
  for (uint32_t i = 0; i < p->nlaunch; i++) {
    if (p->ntasks[i] == 0) continue;
    if (launch_batch(p->field, dim3(p->grid[i]), s, (const BatchTask *)p->d_tasks[i], p->ntasks[i],
                     (const BatchVec *)p->d_vecs[i], (const BatchTables *)p->d_tsets, b) != hipSuccess)
      return OMR_E_DEVICE;
  }
  return OMR_OK;
}

extern "C" omr_status omr_batch_execute(omr_batch_plan *p, const void *d_in, void *d_out, void *stream) try {
  if (!p || !d_in || !d_out) return OMR_E_NULL;
  if (hipSetDevice(p->ctx->device) != hipSuccess) return OMR_E_DEVICE;
  return batch_launches(p, d_in, d_out, (hipStream_t)stream);
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

extern "C" omr_status omr_batch_execute_graph(omr_batch_plan *p, const void *d_in, void *d_out,
                                              void *stream) try {
  if (!p || !d_in || !d_out) return OMR_E_NULL;
  if (hipSetDevice(p->ctx->device) != hipSuccess) return OMR_E_DEVICE;
  if (!p->exec || p->g_in != d_in || p->g_out != d_out) {
    if (p->exec) {
      (void)hipGraphExecDestroy(p->exec);
      p->exec = nullptr;
    }
    if (p->graph) {
      (void)hipGraphDestroy(p->graph);
      p->graph = nullptr;
    }
    if (!p->cap_stream && hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking) != hipSuccess)
      return OMR_E_DEVICE;
    if (hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return OMR_E_DEVICE;
    omr_status st = batch_launches(p, d_in, d_out, p->cap_stream);
    hipGraph_t g = nullptr;
    if (hipStreamEndCapture(p->cap_stream, &g) != hipSuccess || st != OMR_OK) {
      if (g) (void)hipGraphDestroy(g);
      return st != OMR_OK ? st : OMR_E_DEVICE;
    }
    p->graph = g;
    if (hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0) != hipSuccess) return OMR_E_DEVICE;
    p->g_in = d_in;
    p->g_out = d_out;
  }
  return hipGraphLaunch(p->exec, (hipStream_t)stream) == hipSuccess ? OMR_OK : OMR_E_DEVICE;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

// Baseline form: one launch sequence per vector (the engine's plain
// single-vector transform, one call per vector).
extern "C" omr_status omr_batch_execute_plain(omr_batch_plan *p, const void *d_in, void *d_out,
                                              void *stream) try {
  if (!p || !d_in || !d_out) return OMR_E_NULL;
  omr_ctx *ctx = p->ctx;
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (hipSetDevice(ctx->device) != hipSuccess) return OMR_E_DEVICE;
  StreamOrder ord(ctx, (hipStream_t)stream);  // the plain form runs through the context's scratch
  const int es = field_bytes(p->field);
  for (const omr_batch_item &it : p->items) {
    uint64_t w;
    omr_root_of_unity(p->field, 1ull << it.log_n, &w);
    Schedule s;
    if (!tuned_schedule(p->field, it.log_n, s)) default_schedule(p->field, it.log_n, s);
    XformOpts o;
    if (it.inverse) {
      w = finv_host(p->field, w);
      o.scale = true;
      o.scale_w = to_w_host(p->field, finv_host(p->field, (1ull << it.log_n) % field_modulus(p->field)));
    }
    const uint64_t n = 1ull << it.log_n;
    omr_status st = run_transform(ctx, p->field, (const char *)d_in + it.in_offset * es,
                                  (char *)d_out + it.out_offset * es, it.log_n, w, s, 1, n, n, o,
                                  (hipStream_t)stream);
    if (st != OMR_OK) return st;
  }
  return OMR_OK;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

// Grouped form: each group's vectors as one batched transform of the pass
// engine (the tuned schedule of their size), the first pass reading and the
// last pass writing through the group's offset tables, the intermediates
// dense in the plan's own buffers.
static omr_status grouped_launches(omr_batch_plan *p, const void *d_in, void *d_out, hipStream_t s) {
  const bool fork = p->nlanes > 1;
  if (fork) {  // every lane starts after the work already queued on s
    if (hipEventRecord(p->fork_ev, s) != hipSuccess) return OMR_E_DEVICE;
    for (uint32_t l = 0; l < p->nlanes; l++)
      if (hipStreamWaitEvent(p->lane_stream[l], p->fork_ev, 0) != hipSuccess) return OMR_E_DEVICE;
  }
  omr_status st = OMR_OK;
  for (const auto &g : p->groups) {
    XformOpts o;
    o.off_in = g.d_off;
    o.off_out = g.d_off + g.count;
    const DevBuf *ls = p->lscratch[g.lane];
    if (g.npass >= 2) {  // the lane's own intermediates (never the context's, which lanes would share)
      if (!ls[0].p || (ls[0].bytes >> g.log_n) / field_bytes(p->field) < g.count) {
        st = OMR_E_DEVICE;
        break;
      }
      o.scratch[0] = ls[0].p;
      o.scratch[1] = ls[1].p;
    }
    if (g.inverse) {
      o.scale = true;
      o.scale_w = to_w_host(p->field, finv_host(p->field, (1ull << g.log_n) % field_modulus(p->field)));
    }
    st = run_transform(p->ctx, p->field, d_in, d_out, g.log_n, g.root, g.sch, g.count, 0, 0, o,
                       fork ? p->lane_stream[g.lane] : s);
    if (st != OMR_OK) break;
  }
  if (fork) {  // s continues after every lane, also after an error (a capture must end joined)
    for (uint32_t l = 0; l < p->nlanes; l++)
      if (hipEventRecord(p->join_ev[l], p->lane_stream[l]) != hipSuccess ||
          hipStreamWaitEvent(s, p->join_ev[l], 0) != hipSuccess)
        st = OMR_E_DEVICE;
  }
  return st;
}

extern "C" omr_status omr_batch_execute_grouped(omr_batch_plan *p, const void *d_in, void *d_out,
                                                void *stream) try {
  if (!p || !d_in || !d_out) return OMR_E_NULL;
  std::lock_guard<std::mutex> lk(p->ctx->mu);
  if (hipSetDevice(p->ctx->device) != hipSuccess) return OMR_E_DEVICE;
  return grouped_launches(p, d_in, d_out, (hipStream_t)stream);
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}

// The grouped launches captured once per (d_in, d_out) and replayed.  The
// call that captures first runs the launches eagerly, which creates any
// table or buffer they need outside the capture.
extern "C" omr_status omr_batch_execute_grouped_graph(omr_batch_plan *p, const void *d_in, void *d_out,
                                                      void *stream) try {
  if (!p || !d_in || !d_out) return OMR_E_NULL;
  std::lock_guard<std::mutex> lk(p->ctx->mu);
  if (hipSetDevice(p->ctx->device) != hipSuccess) return OMR_E_DEVICE;
  if (p->gexec && p->gg_in == d_in && p->gg_out == d_out)
    return hipGraphLaunch(p->gexec, (hipStream_t)stream) == hipSuccess ? OMR_OK : OMR_E_DEVICE;
  omr_status st = grouped_launches(p, d_in, d_out, (hipStream_t)stream);
  if (st != OMR_OK) return st;
  if (p->gexec) (void)hipGraphExecDestroy(p->gexec), p->gexec = nullptr;
  if (p->ggraph) (void)hipGraphDestroy(p->ggraph), p->ggraph = nullptr;
  if (!p->cap_stream && hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking) != hipSuccess)
    return OMR_E_DEVICE;
  if (hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return OMR_E_DEVICE;
  st = grouped_launches(p, d_in, d_out, p->cap_stream);
  hipGraph_t g = nullptr;
  if (hipStreamEndCapture(p->cap_stream, &g) != hipSuccess || st != OMR_OK) {
    if (g) (void)hipGraphDestroy(g);
    return st != OMR_OK ? st : OMR_E_DEVICE;
  }
  p->ggraph = g;
  if (hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0) != hipSuccess) return OMR_E_DEVICE;
  p->gg_in = d_in;
  p->gg_out = d_out;
  return OMR_OK;
} catch (const std::bad_alloc &) {
  return OMR_E_ALLOC;
} catch (...) {
  return OMR_E_DEVICE;
}
