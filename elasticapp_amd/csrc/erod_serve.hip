// Continuous batching of rod-timestep requests on one device.
//
// A server owns a pool batch of rod slots in size classes (slot capacities in
// elements).  A request names rods of a template batch (the library) and a
// number of steps.  Each tick
//   1. admits queued requests in order: a request whose rods all find free
//      slots (each the smallest class that holds it) is admitted; one device
//      copy puts the template rods into their slots and each slot gets the
//      request's steps as its step budget.  A server created with a budget
//      only (erod_server_create_budgeted) first grows the pool for the
//      earliest waiting request, a class at a time by doubling, as far as
//      the HBM budget allows.  The earliest request that still waits reserves
//      every class it needs: later requests are admitted past it only into
//      slots of the other classes, so it waits for nothing but slots of its
//      own classes to be released;
//   2. advances every slot with budget left by up to `steps_per_tick` fused
//      steps (rods whose budget runs out stop mid-launch and keep their state);
//   3. waits for the device and reports the requests whose rods have all spent
//      their budgets.
// A finished request keeps its slots until it is released, so its results
// can be read.  Rods never interact, and the fused kernels' arithmetic does not
// depend on the tile a rod lands in, so a request's result does not depend on
// what else is in the pool.
//
// No counterpart in the reference: its closest analogue is the first-fit
// admission of rods into blocks at emplace time
// (elastica/Simulator/Configuration/ModuleConfiguration/BlockingPolicy.hpp:202-222).
#include <algorithm>
#include <cstring>
#include <deque>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "erod_common.hpp"

namespace erod {

  namespace {
    struct CopyFields {
      double* dst[EROD_N_FIELDS];
      const double* src[EROD_N_FIELDS];
      uint64_t c32_off[EROD_N_FIELDS];  // ~0: no FP32 copy of the field
      uint64_t dst_slots[3], src_slots[3];
      int8_t ncomp[EROD_N_FIELDS], stagger[EROD_N_FIELDS];
      float* c32;
    };

    // one workgroup per copied rod: every line of every field, the rod clock,
    // and the FP32 constant copy of the destination where it is current
    __global__ void k_copy_rods(CopyFields F, const CopyPair* __restrict__ pairs, const double* __restrict__ src_time,
                                double* __restrict__ dst_time) {
      const CopyPair p = pairs[blockIdx.x];
      for (int f = 0; f < EROD_N_FIELDS; ++f) {
        const int st = F.stagger[f];
        const uint64_t cnt = st == EROD_NODE ? p.n + 1 : (st == EROD_ELEMENT ? p.n : p.n - 1);
        const uint64_t total = cnt * uint64_t(F.ncomp[f]);
        const uint64_t ss = F.src_slots[st], ds = F.dst_slots[st];
        for (uint64_t i = threadIdx.x; i < total; i += blockDim.x) {
          const uint64_t c = i / cnt, k = i - c * cnt;
          const double v = F.src[f][c * ss + p.src_start + k];
          F.dst[f][c * ds + p.dst_start + k] = v;
          if (F.c32 && F.c32_off[f] != ~uint64_t(0)) F.c32[F.c32_off[f] + c * ds + p.dst_start + k] = static_cast<float>(v);
        }
      }
      if (threadIdx.x == 0) dst_time[p.dst_rod] = src_time[p.src_rod];
    }
  }  // namespace

  int serve_copy_rods(erod_batch* dst, const erod_batch* src, const std::vector<CopyPair>& pairs) {
    if (pairs.empty()) return EROD_OK;
    if (dst->copy_capacity < pairs.size()) {
      EROD_HIP(hipStreamSynchronize(dst->stream));
      if (dst->d_copy_pairs) EROD_HIP(hipFree(dst->d_copy_pairs));
      dst->d_copy_pairs = nullptr;
      dst->copy_capacity = 0;
      const uint64_t cap = std::max<uint64_t>(pairs.size(), 256);
      EROD_HIP(hipMalloc(&dst->d_copy_pairs, cap * sizeof(CopyPair)));
      dst->copy_capacity = cap;
    }
    CopyFields F;
    std::memset(&F, 0, sizeof(F));
    const bool c32 = dst->d_const32 && !dst->const32_dirty;
    for (int f = 0; f < EROD_N_FIELDS; ++f) {
      F.dst[f] = dst->d_field[f];
      F.src[f] = src->d_field[f];
      F.ncomp[f] = static_cast<int8_t>(field_info(f).ncomp);
      F.stagger[f] = static_cast<int8_t>(field_info(f).stagger);
      F.c32_off[f] = c32 ? fused_const32_offset(dst, f) : ~uint64_t(0);
    }
    for (int st = 0; st < 3; ++st) {
      F.dst_slots[st] = dst->slots(st);
      F.src_slots[st] = src->slots(st);
    }
    F.c32 = c32 ? dst->d_const32 : nullptr;
    // the source's last writes are on its own stream
    EROD_HIP(hipStreamSynchronize(src->stream));
    EROD_HIP(hipMemcpyAsync(dst->d_copy_pairs, pairs.data(), pairs.size() * sizeof(CopyPair), hipMemcpyHostToDevice,
                            dst->stream));
    hipLaunchKernelGGL(k_copy_rods, dim3(static_cast<unsigned>(pairs.size())), dim3(256), 0, dst->stream, F,
                       dst->d_copy_pairs, src->d_time, dst->d_time);
    EROD_HIP(hipGetLastError());
    // the staging buffer is reused by the next call
    EROD_HIP(hipStreamSynchronize(dst->stream));
    return EROD_OK;
  }

}  // namespace erod

using namespace erod;

struct erod_server {
  erod_batch* library = nullptr;
  erod_batch* pool = nullptr;
  int precision = EROD_FP32_BULK;
  int steps_per_tick = 100;
  uint64_t pool_bytes = 0;
  double hbm_budget = 0.0;
  bool grow = false;      // the pool grows on demand within hbm_budget
  uint64_t growths = 0;   // pool rebuilds so far
  std::vector<uint64_t> class_capacity, class_total;
  std::vector<std::vector<uint64_t>> free_slots;  // per class, a stack of free slot ids
  std::vector<int> slot_class;
  struct Request {
    uint64_t id = 0, steps = 0, elements = 0;
    std::vector<uint64_t> lib_rods, slots;
    bool running = false, done = false;
  };
  std::unordered_map<uint64_t, std::unique_ptr<Request>> requests;
  std::deque<Request*> queue;
  std::vector<Request*> running;
  std::deque<uint64_t> finished;  // done, not yet reported
  uint64_t active_elements = 0;
  uint64_t ticks = 0, element_steps = 0;  // launched ticks and the element-steps they advanced
};

namespace {
  // HBM of a pool of slots with these capacities: every field over every
  // slot, the FP32 constant copy, rod records
  double pool_bytes_for(const std::vector<uint64_t>& n_elems) {
    uint64_t slots = 0;
    for (auto n : n_elems) slots += n + 2;
    double bytes = 0.0;
    for (int f = 0; f < EROD_N_FIELDS; ++f) bytes += 8.0 * field_info(f).ncomp * double(slots);
    return bytes + 4.0 * 22.0 * double(slots) + double(sizeof(RodInfo) + 64) * double(n_elems.size());
  }

  // the pool's slots as a fresh batch: every slot empty with a zero budget,
  // in the library's environment, stepping steps_per_tick steps per launch
  int make_pool(erod_server* s, const std::vector<uint64_t>& n_elems, erod_batch** out);

  int check_server(const erod_server* s) {
    if (!s) {
      set_error("null server");
      return EROD_ERR_INVALID;
    }
    return EROD_OK;
  }

  void give_back(erod_server* s, uint64_t slot) { s->free_slots[s->slot_class[slot]].push_back(slot); }

  // the smallest class holding n elements (the last class if none does)
  size_t class_of(const erod_server* s, uint64_t n) {
    size_t c = 0;
    while (c + 1 < s->class_capacity.size() && s->class_capacity[c] < n) ++c;
    return c;
  }

  // take slots for every rod of a request, each of the smallest class that
  // holds it and is not reserved; all or nothing
  bool place(erod_server* s, const std::vector<uint64_t>& lib_rods, const std::vector<char>& reserved,
             std::vector<uint64_t>& got) {
    got.clear();
    for (uint64_t r : lib_rods) {
      const uint64_t n = s->library->n_active[r];
      int64_t slot = -1;
      for (size_t c = 0; c < s->class_capacity.size() && slot < 0; ++c)
        if (!reserved[c] && s->class_capacity[c] >= n && !s->free_slots[c].empty()) {
          slot = static_cast<int64_t>(s->free_slots[c].back());
          s->free_slots[c].pop_back();
        }
      if (slot < 0) {
        for (uint64_t g : got) give_back(s, g);
        got.clear();
        return false;
      }
      got.push_back(static_cast<uint64_t>(slot));
    }
    return true;
  }
  // empty a slot without touching the pool's damping bookkeeping (the
  // emptied slot is skipped by it)
  void empty_slot(erod_batch* b, uint64_t slot) {
    b->n_active[slot] = 0;
    b->rods[slot].n = 0;
    b->rods[slot].flags = 0;
    b->steps_left[slot] = 0;
    mark_rod(b, slot);
    b->budgets_dirty = true;
    b->plan.valid = false;
    b->graph_precision = -1;
  }

  int make_pool(erod_server* s, const std::vector<uint64_t>& n_elems, erod_batch** out) {
    erod_batch* p = nullptr;
    int rc = erod_batch_create(n_elems.data(), n_elems.size(), s->library->damped ? 1 : 0, &p);
    if (rc) return rc;
    if ((rc = erod_set_rod_step_budget(p, 0, 0))) {
      erod_batch_destroy(p);
      return rc;
    }
    for (uint64_t r = 0; r < p->n_rods; ++r) empty_slot(p, r);
    p->all_full = false;
    p->damping_dirty = false;  // empty slots carry no damping; copied rods bring their rates
    p->env = s->library->env;
    p->plan.steps_per_launch = s->steps_per_tick;
    *out = p;
    return EROD_OK;
  }

  // grow the pool by add[c] slots of each class: a new pool batch holding the
  // old slots under the same ids (the rods in them copied bit for bit, with
  // their clocks, records and budgets) and the new slots after them
  int grow_pool(erod_server* s, const std::vector<uint64_t>& add) {
    erod_batch* old = s->pool;
    std::vector<uint64_t> n_elems = old->n_elems;
    std::vector<int> cls;
    for (size_t c = 0; c < add.size(); ++c)
      for (uint64_t i = 0; i < add[c]; ++i) {
        n_elems.push_back(s->class_capacity[c]);
        cls.push_back(static_cast<int>(c));
      }
    erod_batch* p = nullptr;
    int rc = make_pool(s, n_elems, &p);
    if (rc) return rc;
    // the tile configuration the caller set on the pool carries over
    p->plan.lanes_per_tile = old->plan.lanes_per_tile;
    p->plan.elements_per_thread = old->plan.elements_per_thread;
    p->plan.steps_per_launch = old->plan.steps_per_launch;
    p->plan.split_lanes = old->plan.split_lanes;
    p->plan.split_epl = old->plan.split_epl;
    p->plan.seg_lanes = old->plan.seg_lanes;
    p->plan.seg_epl = old->plan.seg_epl;
    p->plan.halo = old->plan.halo;
    p->plan.tuned_shape = old->plan.tuned_shape;
    // and so does the re-orthonormalisation schedule, at its phase
    p->reortho_every = old->reortho_every;
    p->reortho_phase = old->reortho_phase;
    // (requests admitted in this tick but not yet copied hold slots whose
    // rods arrive after the growth)
    std::vector<uint64_t> used;
    for (auto& kv : s->requests)
      if (kv.second->running || kv.second->done)
        for (uint64_t slot : kv.second->slots) used.push_back(slot);
    if (!used.empty()) {
      if ((rc = erod_copy_rods(p, used.data(), old, used.data(), used.size()))) {
        erod_batch_destroy(p);
        return rc;
      }
      for (uint64_t slot : used) p->steps_left[slot] = old->steps_left[slot];
      p->budgets_dirty = true;
    }
    erod_batch_destroy(old);
    s->pool = p;
    const uint64_t first = s->slot_class.size();
    for (size_t i = 0; i < cls.size(); ++i) {
      s->slot_class.push_back(cls[i]);
      s->free_slots[cls[i]].push_back(first + i);
    }
    for (size_t c = 0; c < add.size(); ++c) {
      s->class_total[c] += add[c];
      // (lowest ids on top of the stack)
      std::sort(s->free_slots[c].rbegin(), s->free_slots[c].rend());
    }
    s->pool_bytes = static_cast<uint64_t>(pool_bytes_for(n_elems));
    ++s->growths;
    return EROD_OK;
  }

  // grow the pool so that the request fits it, within the HBM budget: each
  // class that lacks slots doubles, or grows by what it lacks if that is
  // more, or, when doubling would pass the budget, by what it lacks alone
  int grow_for(erod_server* s, const std::vector<uint64_t>& lib_rods, bool* grown) {
    *grown = false;
    std::vector<uint64_t> need(s->class_capacity.size(), 0), add(need.size(), 0);
    for (uint64_t r : lib_rods) ++need[class_of(s, s->library->n_active[r])];
    bool short_of = false;
    for (size_t c = 0; c < need.size(); ++c)
      if (need[c] > s->free_slots[c].size()) {
        const uint64_t lack = need[c] - s->free_slots[c].size();
        add[c] = std::max(lack, s->class_total[c]);
        short_of = true;
      }
    if (!short_of) return EROD_OK;
    auto bytes_with = [&](const std::vector<uint64_t>& a) {
      std::vector<uint64_t> n = s->pool->n_elems;
      for (size_t c = 0; c < a.size(); ++c) n.insert(n.end(), a[c], s->class_capacity[c]);
      return pool_bytes_for(n);
    };
    if (bytes_with(add) > s->hbm_budget) {
      for (size_t c = 0; c < need.size(); ++c)
        add[c] = need[c] > s->free_slots[c].size() ? need[c] - s->free_slots[c].size() : 0;
      if (bytes_with(add) > s->hbm_budget) return EROD_OK;  // waits for slots to be released
    }
    int rc = grow_pool(s, add);
    if (rc == EROD_OK) *grown = true;
    return rc;
  }

}  // namespace

extern "C" {

// the server of both create calls; class_slots: the initial slot count of each class
static int server_create(erod_batch* library, const uint64_t* class_capacity, const uint64_t* class_slots, int n_classes,
                         int precision, int steps_per_tick, double hbm_budget_bytes, bool grow, erod_server** out) {
  if (!library || !class_capacity || n_classes <= 0 || steps_per_tick <= 0 || !out ||
      (precision != EROD_FP64 && precision != EROD_FP32_BULK) || !(hbm_budget_bytes > 0.0)) {
    set_error("erod_server_create: bad argument");
    return EROD_ERR_INVALID;
  }
  *out = nullptr;
  std::vector<std::pair<uint64_t, uint64_t>> cls;
  for (int c = 0; c < n_classes; ++c) {
    if (class_capacity[c] < 2 || class_slots[c] == 0) {
      set_error("erod_server_create: a class needs capacity >= 2 and at least one slot");
      return EROD_ERR_INVALID;
    }
    cls.emplace_back(class_capacity[c], class_slots[c]);
  }
  std::sort(cls.begin(), cls.end());
  for (size_t c = 1; c < cls.size(); ++c)
    if (cls[c].first == cls[c - 1].first) {
      set_error("erod_server_create: two classes of one capacity");
      return EROD_ERR_INVALID;
    }
  std::vector<uint64_t> n_elems;
  for (auto& [cap, cnt] : cls) n_elems.insert(n_elems.end(), cnt, cap);
  const double bytes = pool_bytes_for(n_elems);
  if (bytes > hbm_budget_bytes) {
    set_error("erod_server_create: the pool (" + std::to_string(bytes / 1e9) + " GB) exceeds the HBM budget");
    return EROD_ERR_INVALID;
  }
  auto s = std::make_unique<erod_server>();
  s->library = library;
  s->precision = precision;
  s->steps_per_tick = steps_per_tick;
  s->pool_bytes = static_cast<uint64_t>(bytes);
  s->hbm_budget = hbm_budget_bytes;
  s->grow = grow;
  int rc = make_pool(s.get(), n_elems, &s->pool);
  if (rc) return rc;
  uint64_t slot = 0;
  for (size_t c = 0; c < cls.size(); ++c) {
    s->class_capacity.push_back(cls[c].first);
    s->class_total.push_back(cls[c].second);
    s->free_slots.emplace_back();
    for (uint64_t i = 0; i < cls[c].second; ++i, ++slot) {
      s->slot_class.push_back(static_cast<int>(c));
      s->free_slots[c].push_back(slot);
    }
    std::reverse(s->free_slots[c].begin(), s->free_slots[c].end());
  }
  *out = s.release();
  return EROD_OK;
}

int erod_server_create(erod_batch* library, const uint64_t* class_capacity, const uint64_t* class_slots, int n_classes,
                       int precision, int steps_per_tick, double hbm_budget_bytes, erod_server** out) {
  if (!class_slots) {
    set_error("erod_server_create: bad argument");
    return EROD_ERR_INVALID;
  }
  return server_create(library, class_capacity, class_slots, n_classes, precision, steps_per_tick, hbm_budget_bytes,
                       false, out);
}

int erod_server_create_budgeted(erod_batch* library, const uint64_t* class_capacity, int n_classes, int precision,
                                int steps_per_tick, double hbm_budget_bytes, erod_server** out) {
  if (n_classes <= 0) {
    set_error("erod_server_create_budgeted: bad argument");
    return EROD_ERR_INVALID;
  }
  // a few slots of each class to start with; the pool grows as requests come
  const std::vector<uint64_t> initial(static_cast<size_t>(n_classes), 16);
  return server_create(library, class_capacity, initial.data(), n_classes, precision, steps_per_tick, hbm_budget_bytes,
                       true, out);
}

int erod_server_destroy(erod_server* s) {
  if (check_server(s)) return EROD_ERR_INVALID;
  erod_batch_destroy(s->pool);
  delete s;
  return EROD_OK;
}

int erod_server_pool(erod_server* s, erod_batch** pool) {
  if (check_server(s) || !pool) return EROD_ERR_INVALID;
  *pool = s->pool;
  return EROD_OK;
}

int erod_server_submit(erod_server* s, uint64_t id, const uint64_t* library_rods, uint64_t n_rods, uint64_t steps) {
  if (check_server(s) || !library_rods || n_rods == 0 || steps == 0 || steps >= kNoBudget) {
    set_error("erod_server_submit: bad argument (rods and steps must be positive, steps < 2^32 - 1)");
    return EROD_ERR_INVALID;
  }
  if (s->requests.count(id)) {
    set_error("erod_server_submit: request id already in use");
    return EROD_ERR_INVALID;
  }
  const uint64_t maxcap = s->class_capacity.back();
  auto rq = std::make_unique<erod_server::Request>();
  rq->id = id;
  rq->steps = steps;
  for (uint64_t i = 0; i < n_rods; ++i) {
    const uint64_t r = library_rods[i];
    if (r >= s->library->n_rods || s->library->n_active[r] == 0 || s->library->n_active[r] > maxcap) {
      set_error("erod_server_submit: library rod out of range, empty, or longer than the largest slot");
      return EROD_ERR_INVALID;
    }
    rq->lib_rods.push_back(r);
    rq->elements += s->library->n_active[r];
  }
  // a request must fit the empty pool (one that may grow: grown as far as the
  // budget allows), or it would wait forever
  std::vector<uint64_t> need(rq->lib_rods.size());
  for (size_t i = 0; i < need.size(); ++i) need[i] = s->library->n_active[rq->lib_rods[i]];
  std::sort(need.rbegin(), need.rend());
  std::vector<uint64_t> left(s->class_capacity.size());
  for (size_t c = 0; c < left.size(); ++c) left[c] = s->class_total[c];
  std::vector<uint64_t> grown = s->pool->n_elems;
  for (uint64_t n : need) {
    size_t c = 0;
    while (c < left.size() && (s->class_capacity[c] < n || left[c] == 0)) ++c;
    if (c < left.size()) {
      --left[c];
    } else if (s->grow) {
      grown.push_back(s->class_capacity[class_of(s, n)]);
    } else {
      set_error("erod_server_submit: the request does not fit the pool");
      return EROD_ERR_INVALID;
    }
  }
  if (s->grow && pool_bytes_for(grown) > s->hbm_budget) {
    set_error("erod_server_submit: the request does not fit the pool within the HBM budget");
    return EROD_ERR_INVALID;
  }
  s->queue.push_back(rq.get());
  s->requests[id] = std::move(rq);
  return EROD_OK;
}

int erod_server_tick(erod_server* s, uint64_t* done_ids, uint64_t capacity, uint64_t* n_done) {
  if (check_server(s) || !n_done || (capacity && !done_ids)) return EROD_ERR_INVALID;
  // 1. admission, in queue order: the earliest request that cannot be placed
  // (after growing the pool, if it may grow) reserves the classes it needs;
  // up to kLookahead later requests are tried past it, outside those classes
  constexpr size_t kLookahead = 1024;
  std::vector<uint64_t> dst, src;
  std::vector<erod_server::Request*> admitted;
  std::vector<char> reserved(s->class_capacity.size(), 0);
  bool blocked = false;
  size_t past = 0;
  for (auto it = s->queue.begin(); it != s->queue.end() && past < kLookahead;) {
    erod_server::Request* rq = *it;
    std::vector<uint64_t> got;
    bool ok = place(s, rq->lib_rods, reserved, got);
    if (!ok && !blocked && s->grow) {
      bool grown = false;
      if (int grc = grow_for(s, rq->lib_rods, &grown)) {
        for (erod_server::Request* a : admitted) {
          for (uint64_t slot : a->slots) give_back(s, slot);
          a->slots.clear();
        }
        return grc;
      }
      if (grown) ok = place(s, rq->lib_rods, reserved, got);
    }
    if (ok) {
      rq->slots = got;
      dst.insert(dst.end(), got.begin(), got.end());
      src.insert(src.end(), rq->lib_rods.begin(), rq->lib_rods.end());
      admitted.push_back(rq);
      it = s->queue.erase(it);
    } else {
      if (!blocked) {
        blocked = true;
        for (uint64_t r : rq->lib_rods) reserved[class_of(s, s->library->n_active[r])] = 1;
      }
      ++it;
    }
    if (blocked) ++past;
  }
  erod_batch* p = s->pool;
  int rc = EROD_OK;
  if (!dst.empty()) {
    rc = erod_copy_rods(p, dst.data(), s->library, src.data(), dst.size());
    for (erod_server::Request* rq : admitted)
      for (uint64_t slot : rq->slots)
        if (rc == EROD_OK) rc = erod_set_rod_step_budget(p, slot, rq->steps);
    if (rc != EROD_OK) {
      // put the requests back at the head of the queue, in order, and free their slots
      const std::string msg = erod_last_error();
      for (auto it = admitted.rbegin(); it != admitted.rend(); ++it) {
        for (uint64_t slot : (*it)->slots) {
          empty_slot(p, slot);
          give_back(s, slot);
        }
        (*it)->slots.clear();
        s->queue.push_front(*it);
      }
      set_error(msg);
      return rc;
    }
    for (erod_server::Request* rq : admitted) {
      rq->running = true;
      s->running.push_back(rq);
      s->active_elements += rq->elements;
    }
  }
  // 2. one launch over every slot with budget left, in the library's current
  // environment (gravity, contact), as the same rods would step there
  p->env = s->library->env;
  if (!s->running.empty()) {
    uint64_t work = 0;
    for (erod_server::Request* rq : s->running)
      for (uint64_t slot : rq->slots)
        work += std::min<uint64_t>(uint64_t(s->steps_per_tick), p->steps_left[slot]) * p->n_active[slot];
    if ((rc = erod_step_fused(p, s->precision, static_cast<uint64_t>(s->steps_per_tick)))) return rc;
    EROD_HIP(hipStreamSynchronize(p->stream));
    ++s->ticks;
    s->element_steps += work;
  }
  // 3. completions
  std::vector<erod_server::Request*> still;
  for (erod_server::Request* rq : s->running) {
    bool done = true;
    for (uint64_t slot : rq->slots) done &= p->steps_left[slot] == 0;
    if (done) {
      rq->done = true;
      rq->running = false;
      s->active_elements -= rq->elements;
      s->finished.push_back(rq->id);
    } else {
      still.push_back(rq);
    }
  }
  s->running.swap(still);
  uint64_t k = 0;
  while (k < capacity && !s->finished.empty()) {
    done_ids[k++] = s->finished.front();
    s->finished.pop_front();
  }
  *n_done = k;
  return EROD_OK;
}

int erod_server_read(erod_server* s, uint64_t id, uint64_t rod, int field, double* host) {
  if (check_server(s)) return EROD_ERR_INVALID;
  auto it = s->requests.find(id);
  if (it == s->requests.end() || !it->second->done || rod >= it->second->slots.size()) {
    set_error("erod_server_read: unknown or unfinished request, or rod out of range");
    return EROD_ERR_INVALID;
  }
  return erod_download_rod(s->pool, field, it->second->slots[rod], host);
}

int erod_server_release(erod_server* s, uint64_t id) {
  if (check_server(s)) return EROD_ERR_INVALID;
  auto it = s->requests.find(id);
  if (it == s->requests.end() || !it->second->done) {
    set_error("erod_server_release: unknown or unfinished request");
    return EROD_ERR_INVALID;
  }
  for (uint64_t slot : it->second->slots) {
    empty_slot(s->pool, slot);
    give_back(s, slot);
  }
  s->requests.erase(it);
  return EROD_OK;
}

int erod_server_status(const erod_server* s, uint64_t* queued, uint64_t* running, uint64_t* active_elements,
                       uint64_t* pool_bytes) {
  if (check_server(s)) return EROD_ERR_INVALID;
  if (queued) *queued = s->queue.size();
  if (running) *running = s->running.size();
  if (active_elements) *active_elements = s->active_elements;
  if (pool_bytes) *pool_bytes = s->pool_bytes;
  return EROD_OK;
}

int erod_server_pool_info(const erod_server* s, uint64_t* slots, uint64_t* growths, double* hbm_budget_bytes) {
  if (check_server(s)) return EROD_ERR_INVALID;
  if (slots) *slots = s->slot_class.size();
  if (growths) *growths = s->growths;
  if (hbm_budget_bytes) *hbm_budget_bytes = s->hbm_budget;
  return EROD_OK;
}

int erod_server_stats(const erod_server* s, uint64_t* ticks, uint64_t* element_steps) {
  if (check_server(s)) return EROD_ERR_INVALID;
  if (ticks) *ticks = s->ticks;
  if (element_steps) *element_steps = s->element_steps;
  return EROD_OK;
}

}  // extern "C"
