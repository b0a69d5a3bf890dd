// Continuous batching of rod-timestep requests on one device.
//
// A server owns a pool batch of rod slots in size classes (slot capacities in
// elements).  A request names rods of a template batch (the library) and a
// number of steps.  Each tick
//   1. admits queued requests first-come first-served while every rod of the
//      head request finds a free slot (the smallest class that holds it):
//      one device copy puts the template rods into their slots and each slot
//      gets the request's steps as its step budget;
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
  int check_server(const erod_server* s) {
    if (!s) {
      set_error("null server");
      return EROD_ERR_INVALID;
    }
    return EROD_OK;
  }

  // take a slot of the smallest class holding n elements; -1 if none is free
  int64_t take_slot(erod_server* s, uint64_t n) {
    for (size_t c = 0; c < s->class_capacity.size(); ++c)
      if (s->class_capacity[c] >= n && !s->free_slots[c].empty()) {
        const uint64_t slot = s->free_slots[c].back();
        s->free_slots[c].pop_back();
        return static_cast<int64_t>(slot);
      }
    return -1;
  }

  void give_back(erod_server* s, uint64_t slot) { s->free_slots[s->slot_class[slot]].push_back(slot); }

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
}  // namespace

extern "C" {

int erod_server_create(erod_batch* library, const uint64_t* class_capacity, const uint64_t* class_slots, int n_classes,
                       int precision, int steps_per_tick, double hbm_budget_bytes, erod_server** out) {
  if (!library || !class_capacity || !class_slots || n_classes <= 0 || steps_per_tick <= 0 || !out ||
      (precision != EROD_FP64 && precision != EROD_FP32_BULK)) {
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
  std::vector<uint64_t> n_elems;
  for (auto& [cap, cnt] : cls) n_elems.insert(n_elems.end(), cnt, cap);
  // HBM of the pool: every field over every slot, the FP32 constant copy, rod records
  uint64_t slots = 0;
  for (auto n : n_elems) slots += n + 2;
  double bytes = 0.0;
  for (int f = 0; f < EROD_N_FIELDS; ++f) bytes += 8.0 * field_info(f).ncomp * double(slots);
  bytes += 4.0 * 22.0 * double(slots) + double(sizeof(RodInfo) + 64) * double(n_elems.size());
  if (bytes > hbm_budget_bytes) {
    set_error("erod_server_create: the pool (" + std::to_string(bytes / 1e9) + " GB) exceeds the HBM budget");
    return EROD_ERR_INVALID;
  }
  auto s = std::make_unique<erod_server>();
  s->library = library;
  s->precision = precision;
  s->steps_per_tick = steps_per_tick;
  s->pool_bytes = static_cast<uint64_t>(bytes);
  int rc = erod_batch_create(n_elems.data(), n_elems.size(), library->damped ? 1 : 0, &s->pool);
  if (rc) return rc;
  erod_batch* p = s->pool;
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
  // every slot starts empty, with a zero budget; each tick takes the library's environment
  if ((rc = erod_set_rod_step_budget(p, 0, 0))) {
    erod_batch_destroy(p);
    return rc;
  }
  for (uint64_t r = 0; r < p->n_rods; ++r) empty_slot(p, r);
  p->all_full = false;
  p->damping_dirty = false;  // empty slots carry no damping; copied rods bring their rates
  p->env = library->env;
  p->plan.steps_per_launch = steps_per_tick;
  *out = s.release();
  return EROD_OK;
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
  // a request must fit the empty pool, or it would block the queue forever
  std::vector<uint64_t> need(rq->lib_rods.size());
  for (size_t i = 0; i < need.size(); ++i) need[i] = s->library->n_active[rq->lib_rods[i]];
  std::sort(need.rbegin(), need.rend());
  std::vector<uint64_t> left(s->class_capacity.size());
  for (size_t c = 0; c < left.size(); ++c) left[c] = s->class_total[c];
  for (uint64_t n : need) {
    size_t c = 0;
    while (c < left.size() && (s->class_capacity[c] < n || left[c] == 0)) ++c;
    if (c == left.size()) {
      set_error("erod_server_submit: the request does not fit the pool");
      return EROD_ERR_INVALID;
    }
    --left[c];
  }
  s->queue.push_back(rq.get());
  s->requests[id] = std::move(rq);
  return EROD_OK;
}

int erod_server_tick(erod_server* s, uint64_t* done_ids, uint64_t capacity, uint64_t* n_done) {
  if (check_server(s) || !n_done || (capacity && !done_ids)) return EROD_ERR_INVALID;
  erod_batch* p = s->pool;
  // 1. admission, first come first served
  std::vector<uint64_t> dst, src;
  std::vector<erod_server::Request*> admitted;
  while (!s->queue.empty()) {
    erod_server::Request* rq = s->queue.front();
    std::vector<uint64_t> got;
    for (uint64_t r : rq->lib_rods) {
      const int64_t slot = take_slot(s, s->library->n_active[r]);
      if (slot < 0) break;
      got.push_back(static_cast<uint64_t>(slot));
    }
    if (got.size() < rq->lib_rods.size()) {
      for (uint64_t slot : got) give_back(s, slot);
      break;
    }
    rq->slots = got;
    dst.insert(dst.end(), got.begin(), got.end());
    src.insert(src.end(), rq->lib_rods.begin(), rq->lib_rods.end());
    admitted.push_back(rq);
    s->queue.pop_front();
  }
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

int erod_server_stats(const erod_server* s, uint64_t* ticks, uint64_t* element_steps) {
  if (check_server(s)) return EROD_ERR_INVALID;
  if (ticks) *ticks = s->ticks;
  if (element_steps) *element_steps = s->element_steps;
  return EROD_OK;
}

}  // extern "C"
