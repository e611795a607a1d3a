// LDS slot plan of the direction-resident LDS sweep.
//
// While a direction is swept level by level, a cell's flux is needed again only
// by its downstream neighbours, which sit in later levels.  The LDS sweep keeps
// such traces in LDS slots ([slot][node][group] doubles, one workgroup per
// direction).  The plan fixes, per direction and before any sweep, which slot
// each cell writes and where each upstream trace is read from:
//
// * Ring slots, assigned level by level when the producing cell is solved.  A
//   producer holds its slot until its last consumer has read it.  When no slot
//   is free, the live producer whose next use lies furthest ahead gives its slot
//   up if that use is later than the new cell's first use; otherwise the new
//   cell gets none.  This is Belady's MIN replacement rule, optimal for a known
//   reference string (L. A. Belady, IBM Systems J. 5 (1966) 78-101); a sweep's
//   string of trace uses is known once the schedule is.  A consumer reads its
//   producer's slot if the producer still holds it at the consumer's level.
// * Staging slots, two sets of `stage` used by alternate levels.  A trace that
//   misses the ring and whose producer lies two or more levels back is copied
//   from global memory into a staging slot of the consumer's level parity
//   during the level before, so that the copy overlaps that level's work and the
//   consumer still reads LDS.  Distinct producers per level beyond `stage`, and
//   misses whose producer lies one level back (its global store is ordered only
//   by the barrier that ends that level), are read from global memory.
//
// A slot written during level l is read at later levels only, and every write
// of level l goes to a slot whose previous holder's readers all ran before l,
// so the barrier between levels is the only ordering the kernel needs.
#include <omp.h>

#include <algorithm>
#include <climits>
#include <cstring>

#include "lb_internal.hpp"

namespace lb {

namespace {

struct Scratch {
  std::vector<int32_t> pos, lev, start, cnt, clev, cursor, slot, evict, stamp, sidx, freel, live, live_at;
  std::vector<uint8_t> state;   // 0 no slot yet, 1 holding, 2 released or evicted
  std::vector<std::pair<int32_t, int32_t>> heap, order;
};

// One direction.  tk: task words of the direction in level order, off: its level
// offsets (nl + 1).  Writes uslot (4 per position), wslot (1 per position) and
// appends the staging entries (LDS slot, element) level by level, with
// soff[l] = first entry of consumer level l (nl + 1 values).  gprod (optional,
// 1 per position): 1 for a cell whose flux some consumer reads from global
// memory: through a staging copy, or directly by a consumer with any upstream
// trace outside LDS (such a consumer reads all its traces from global memory);
// 0 for one read from LDS only.
void plan_direction(int64_t NE, int nl, const int32_t* off, const uint32_t* tk, const int32_t* nb, int ring,
                    int stage, int32_t* uslot, int32_t* wslot, int32_t* soff, std::vector<int32_t>& ent,
                    int64_t* stats, Scratch& s, uint8_t* gprod) {
  s.pos.resize(NE);
  s.lev.resize(NE);
  s.start.assign(NE + 1, 0);
  s.cnt.assign(NE, 0);
  s.cursor.resize(NE);
  s.slot.assign(NE, -1);
  s.evict.assign(NE, INT_MAX);
  s.state.assign(NE, 0);
  s.stamp.assign(NE, -1);
  s.sidx.resize(NE);
  s.live_at.resize(NE);
  for (int64_t p = 0; p < NE; ++p) s.pos[tk[p] & 0xFFFFFFu] = (int32_t)p;
  for (int l = 0; l < nl; ++l)
    for (int32_t p = off[l]; p < off[l + 1]; ++p) s.lev[p] = l;
  auto producer = [&](int64_t p, int f) { return s.pos[nb[(int64_t)(tk[p] & 0xFFFFFFu) * 4 + f]]; };
  // consumer levels of every producer, ascending (consumers are visited in level order)
  for (int64_t p = 0; p < NE; ++p) {
    const unsigned up = (tk[p] >> 24) & 0xFu;
    for (int f = 0; f < 4; ++f)
      if (up >> f & 1u) ++s.cnt[producer(p, f)];
  }
  for (int64_t q = 0; q < NE; ++q) s.start[q + 1] = s.start[q] + s.cnt[q];
  s.clev.resize(s.start[NE]);
  for (int64_t q = 0; q < NE; ++q) s.cursor[q] = s.start[q];
  for (int64_t p = 0; p < NE; ++p) {
    const unsigned up = (tk[p] >> 24) & 0xFu;
    for (int f = 0; f < 4; ++f)
      if (up >> f & 1u) {
        const int32_t q = producer(p, f);
        s.clev[s.cursor[q]++] = s.lev[p];
      }
  }
  for (int64_t q = 0; q < NE; ++q) s.cursor[q] = s.start[q];

  // ring slots by the MIN rule.  The contest needs the live producers ordered by
  // next use only while the ring is full; the heap (lazy deletion: an entry is
  // current while its producer holds a slot and its key is that producer's
  // next use) is built from the live list when a level first needs it and
  // dropped once a level fits the free slots again.
  s.freel.clear();
  for (int k = ring - 1; k >= 0; --k) s.freel.push_back(k);
  s.live.clear();
  s.heap.clear();
  bool heap_on = false;
  auto push = [&](int32_t key, int32_t q) {
    s.heap.emplace_back(key, q);
    std::push_heap(s.heap.begin(), s.heap.end());
  };
  auto hold = [&](int32_t p, int32_t slot) {
    s.slot[p] = slot;
    s.state[p] = 1;
    s.live_at[p] = (int32_t)s.live.size();
    s.live.push_back(p);
    if (heap_on) push(s.clev[s.cursor[p]], p);
  };
  auto drop = [&](int32_t q) {   // q leaves the live list
    s.state[q] = 2;
    const int32_t k = s.live_at[q], last = s.live.back();
    s.live[k] = last;
    s.live_at[last] = k;
    s.live.pop_back();
  };
  for (int l = 0; l < nl; ++l) {
    if (l > 0)   // the readers of level l - 1 are done
      for (int32_t p = off[l - 1]; p < off[l]; ++p) {
        const unsigned up = (tk[p] >> 24) & 0xFu;
        for (int f = 0; f < 4; ++f) {
          if (!(up >> f & 1u)) continue;
          const int32_t q = producer(p, f);
          if (s.state[q] != 1 || s.clev[s.cursor[q]] > l - 1) continue;
          while (s.cursor[q] < s.start[q + 1] && s.clev[s.cursor[q]] <= l - 1) ++s.cursor[q];
          if (s.cursor[q] == s.start[q + 1]) {
            drop(q);
            s.freel.push_back(s.slot[q]);
          } else if (heap_on) {
            push(s.clev[s.cursor[q]], q);
          }
        }
      }
    s.order.clear();
    for (int32_t p = off[l]; p < off[l + 1]; ++p)
      if (s.cnt[p] > 0) s.order.emplace_back(s.clev[s.start[p]], p);
    if (s.order.size() <= s.freel.size()) {   // the level fits the free slots
      heap_on = false;
      s.heap.clear();
      for (const auto& op : s.order) {
        const int32_t slot = s.freel.back();
        s.freel.pop_back();
        hold(op.second, slot);
      }
      continue;
    }
    if (!heap_on) {
      heap_on = true;
      s.heap.clear();
      for (int32_t q : s.live) s.heap.emplace_back(s.clev[s.cursor[q]], q);
      std::make_heap(s.heap.begin(), s.heap.end());
    }
    // this level's producers, most urgent first: a producer that loses the
    // contest is then never one written in this level (its next use is at most
    // the new cell's), so no two cells of a level are given the same slot
    std::sort(s.order.begin(), s.order.end());
    for (const auto& op : s.order) {
      const int32_t p = op.second, next = op.first;
      if (!s.freel.empty()) {
        const int32_t slot = s.freel.back();
        s.freel.pop_back();
        hold(p, slot);
        continue;
      }
      while (!s.heap.empty()) {
        const auto top = s.heap.front();
        const int32_t q = top.second;
        if (s.state[q] == 1 && s.clev[s.cursor[q]] == top.first) break;
        std::pop_heap(s.heap.begin(), s.heap.end());
        s.heap.pop_back();
      }
      if (s.heap.empty() || s.heap.front().first <= next) continue;   // no slot for p
      const int32_t q = s.heap.front().second;
      std::pop_heap(s.heap.begin(), s.heap.end());
      s.heap.pop_back();
      drop(q);
      s.evict[q] = l;
      hold(p, s.slot[q]);
    }
  }

  // where each upstream trace is read, and the staging lists
  for (int l = 0; l < nl; ++l) {
    soff[l] = (int32_t)(ent.size() / 2);
    int used = 0;
    for (int32_t p = off[l]; p < off[l + 1]; ++p) {
      const unsigned up = (tk[p] >> 24) & 0xFu;
      bool miss = false;
      for (int f = 0; f < 4; ++f) {
        int32_t v = -1;
        if (up >> f & 1u) {
          const int32_t q = producer(p, f);
          if (s.slot[q] >= 0 && l < s.evict[q]) {
            v = s.slot[q];
            ++stats[0];
          } else if (s.lev[q] <= l - 2) {
            if (s.stamp[q] != l) {
              s.stamp[q] = l;
              s.sidx[q] = used < stage ? used++ : -1;
              if (s.sidx[q] >= 0) {
                ent.push_back(ring + (l & 1) * stage + s.sidx[q]);
                ent.push_back((int32_t)(tk[q] & 0xFFFFFFu));
              }
            }
            if (s.sidx[q] >= 0) {
              v = ring + (l & 1) * stage + s.sidx[q];
              ++stats[1];
            } else {
              ++stats[3];
            }
          } else {
            ++stats[2];
          }
        }
        uslot[(int64_t)p * 4 + f] = v;
        if (up >> f & 1u) {
          miss |= v < 0;
          if (gprod && v >= ring) gprod[producer(p, f)] = 1;   // its staging copy is read from global memory
        }
      }
      // a cell with any trace outside LDS reads all its upstream traces from global memory
      if (gprod && miss)
        for (int f = 0; f < 4; ++f)
          if (up >> f & 1u) gprod[producer(p, f)] = 1;
      wslot[p] = s.slot[p];
    }
  }
  soff[nl] = (int32_t)(ent.size() / 2);
}

}  // namespace

int trace_plan_sizes(int lds_bytes, int G, int* ring, int* stage) {
  const int per_slot = 4 * G * (int)sizeof(double);
  const int total = lds_bytes / per_slot;
  const int st = std::min(64 / G, total / 8);
  if (total - 2 * st < 1) {
    set_error("sweep: the LDS cannot hold one trace of this group batch");
    return LB_E_ARG;
  }
  *ring = total - 2 * st;
  *stage = st;
  return LB_OK;
}

int build_trace_plan(int64_t NI, int64_t NE, int maxlev, const int32_t* nlev, const int32_t* lvl_off,
                     const uint32_t* tasks, const int32_t* nbr, int ring, int stage, int32_t* uslot, int32_t* wslot,
                     int32_t* stage_off, std::vector<int32_t>& entries, int64_t stats[4], uint8_t* gprod) {
  if (gprod) std::memset(gprod, 0, (size_t)(NI * NE));
  const int64_t stride = maxlev + 2;
  std::vector<std::vector<int32_t>> ent(NI);
  std::vector<int64_t> st(NI * 4, 0);
#pragma omp parallel
  {
    Scratch s;
#pragma omp for schedule(dynamic, 1)
    for (int64_t i = 0; i < NI; ++i) {
      int32_t* so = stage_off + i * stride;
      plan_direction(NE, nlev[i], lvl_off + i * (maxlev + 1), tasks + i * NE, nbr, ring, stage, uslot + i * NE * 4,
                     wslot + i * NE, so, ent[i], &st[i * 4], s, gprod ? gprod + i * NE : nullptr);
      for (int l = nlev[i] + 1; l < stride; ++l) so[l] = so[nlev[i]];
    }
  }
  // entries of all directions in one array: offsets become global (int32)
  int64_t total = 0;
  for (int64_t i = 0; i < NI; ++i) total += (int64_t)ent[i].size() / 2;
  if (total > INT_MAX) {
    set_error("sweep: LDS slot plan has more staging entries than a 32-bit offset holds");
    return LB_E_ARG;
  }
  entries.clear();
  entries.reserve(total * 2);
  for (int64_t i = 0; i < NI; ++i) {
    const int32_t base = (int32_t)(entries.size() / 2);
    for (int64_t l = 0; l < stride; ++l) stage_off[i * stride + l] += base;
    entries.insert(entries.end(), ent[i].begin(), ent[i].end());
  }
  for (int k = 0; k < 4; ++k) {
    stats[k] = 0;
    for (int64_t i = 0; i < NI; ++i) stats[k] += st[i * 4 + k];
  }
  return LB_OK;
}

}  // namespace lb
