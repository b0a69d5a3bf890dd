// C ABI of the device-resident Cosserat-rod block (include/erod.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "erod_common.hpp"

namespace erod {

  static thread_local std::string g_last_error;
  void set_error(const std::string& msg) { g_last_error = msg; }

  namespace {
    __global__ void k_scatter_rods(const RodInfo* __restrict__ stage, const uint32_t* __restrict__ idx, uint64_t count,
                                   RodInfo* __restrict__ rods) {
      const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
      if (i < count) rods[idx[i]] = stage[i];
    }
  }  // namespace

  // Upload the rod records the host changed: all of them, or, when only a few
  // changed (a serving pool admits and releases a handful of slots per tick),
  // just those, staged and scattered by a small kernel.
  int sync_rods(erod_batch* b) {
    if (!b->rods_dirty) {
      std::vector<uint32_t>& d = b->dirty_rods;
      if (d.empty()) return EROD_OK;
      std::sort(d.begin(), d.end());
      d.erase(std::unique(d.begin(), d.end()), d.end());
      if (d.size() * 4 <= b->n_rods) {
        if (b->rod_stage_capacity < d.size()) {
          EROD_HIP(hipStreamSynchronize(b->stream));
          if (b->d_rod_stage) EROD_HIP(hipFree(b->d_rod_stage));
          if (b->d_rod_stage_idx) EROD_HIP(hipFree(b->d_rod_stage_idx));
          b->d_rod_stage = nullptr;
          b->d_rod_stage_idx = nullptr;
          b->rod_stage_capacity = 0;
          const uint64_t cap = std::max<uint64_t>(d.size(), 256);
          EROD_HIP(hipMalloc(&b->d_rod_stage, cap * sizeof(RodInfo)));
          EROD_HIP(hipMalloc(&b->d_rod_stage_idx, cap * sizeof(uint32_t)));
          b->rod_stage_capacity = cap;
        }
        std::vector<RodInfo> stage(d.size());
        for (size_t i = 0; i < d.size(); ++i) stage[i] = b->rods[d[i]];
        EROD_HIP(hipMemcpyAsync(b->d_rod_stage, stage.data(), stage.size() * sizeof(RodInfo), hipMemcpyHostToDevice,
                                b->stream));
        EROD_HIP(hipMemcpyAsync(b->d_rod_stage_idx, d.data(), d.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                                b->stream));
        hipLaunchKernelGGL(k_scatter_rods, dim3(static_cast<unsigned>((d.size() + 255) / 256)), dim3(256), 0, b->stream,
                           b->d_rod_stage, b->d_rod_stage_idx, uint64_t(d.size()), b->d_rods);
        EROD_HIP(hipGetLastError());
        // the host staging vectors are released on return
        EROD_HIP(hipStreamSynchronize(b->stream));
        d.clear();
        return EROD_OK;
      }
    }
    EROD_HIP(hipMemcpyAsync(b->d_rods, b->rods.data(), sizeof(RodInfo) * b->n_rods, hipMemcpyHostToDevice, b->stream));
    EROD_HIP(hipStreamSynchronize(b->stream));
    b->rods_dirty = false;
    b->dirty_rods.clear();
    return EROD_OK;
  }

  // per-rod damping rates for the fused path: the reference stores one rate per
  // node / element (ExplicitDampingAdapterFacade.hpp:185-198); a rod whose
  // rates are the same along it carries them in its record, any other is
  // flagged (bit 2) and the fused kernels read its rates per node and element
  static int refresh_rod_damping(erod_batch* b) {
    const uint64_t N = b->E + 1, E = b->E;
    std::vector<double> gf(N), gt(E);
    EROD_HIP(hipMemcpyAsync(gf.data(), b->d_field[EROD_FORCE_DAMPING_RATE], N * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    EROD_HIP(hipMemcpyAsync(gt.data(), b->d_field[EROD_TORQUE_DAMPING_RATE], E * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    EROD_HIP(hipStreamSynchronize(b->stream));
    for (uint64_t r = 0; r < b->n_rods; ++r) {
      const uint64_t s = b->starts[r], n = b->n_active[r];
      double f = b->damped && n ? gf[s] : 0.0, t = b->damped && n ? gt[s] : 0.0;
      bool uniform = true;
      if (b->damped && n) {
        for (uint64_t i = 0; i <= n && uniform; ++i) uniform = gf[s + i] == f;
        for (uint64_t i = 0; i < n && uniform; ++i) uniform = gt[s + i] == t;
      }
      b->rods[r].gamma_f = uniform ? f : 0.0;
      b->rods[r].gamma_t = uniform ? t : 0.0;
      b->rods[r].flags = uniform ? (b->rods[r].flags & ~4u) : (b->rods[r].flags | 4u);
    }
    b->rods_dirty = true;
    return sync_rods(b);
  }

}  // namespace erod

using namespace erod;

namespace {
  bool is_fp32_const(int f) {
    return f == EROD_MASS || f == EROD_INV_MASS || f == EROD_REFERENCE_ELEMENT_LENGTH ||
           f == EROD_REFERENCE_SHEAR_STRETCH_STRAIN || f == EROD_SHEAR_STRETCH_RIGIDITY ||
           f == EROD_MASS_SECOND_MOMENT_OF_INERTIA || f == EROD_INV_MASS_SECOND_MOMENT_OF_INERTIA ||
           f == EROD_REFERENCE_VORONOI_LENGTH || f == EROD_REFERENCE_CURVATURE || f == EROD_BENDING_TWIST_RIGIDITY ||
           f == EROD_FORCE_DAMPING_RATE || f == EROD_TORQUE_DAMPING_RATE;
  }

  int check_batch(const erod_batch* b) {
    if (!b) {
      set_error("null batch");
      return EROD_ERR_INVALID;
    }
    return EROD_OK;
  }

  int check_field(int f) {
    if (f < 0 || f >= EROD_N_FIELDS) {
      set_error("field id out of range");
      return EROD_ERR_INVALID;
    }
    return EROD_OK;
  }

  // the op-by-op kernels run over the whole ghost-padded block
  int check_plain(erod_batch* b) {
    if (check_batch(b)) return EROD_ERR_INVALID;
    if (!b->all_full) {
      set_error("the op-by-op path needs every rod at its full length (see erod_set_rod_active_length)");
      return EROD_ERR_UNSUPPORTED;
    }
    return EROD_OK;
  }

  void mark_field_written(erod_batch* b, int f) {
    if (is_fp32_const(f)) b->const32_dirty = true;
    if (f == EROD_FORCE_DAMPING_RATE || f == EROD_TORQUE_DAMPING_RATE) b->damping_dirty = true;
  }
}  // namespace

// damping rates are re-read before the next fused call after they were written
static int ensure_damping(erod_batch* b) {
  if (!b->damping_dirty) return EROD_OK;
  b->damping_dirty = false;
  return refresh_rod_damping(b);
}

extern "C" {

const char* erod_last_error(void) { return g_last_error.c_str(); }

int erod_field_shape(int field, int* ncomp, int* stagger) {
  if (check_field(field)) return EROD_ERR_INVALID;
  if (ncomp) *ncomp = field_info(field).ncomp;
  if (stagger) *stagger = field_info(field).stagger;
  return EROD_OK;
}

int erod_batch_create(const uint64_t* n_elems, uint64_t n_rods, int damped, erod_batch** out) {
  if (!out || !n_elems || n_rods == 0 || n_rods > UINT32_MAX) {
    // rod ids travel as 32-bit values in the tile tables and record uploads
    set_error("erod_batch_create: need between 1 and 2^32 - 1 rods");
    return EROD_ERR_INVALID;
  }
  *out = nullptr;
  for (uint64_t r = 0; r < n_rods; ++r)
    if (n_elems[r] < 2 || n_elems[r] > (uint64_t(1) << 31)) {
      set_error("Cannot have rods with < 2 elements! (or more than 2^31)");
      return EROD_ERR_INVALID;
    }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_error("no HIP device visible");
    return EROD_ERR_NO_DEVICE;
  }
  int device = 0;
  EROD_HIP(hipGetDevice(&device));
  auto* b = new erod_batch();
  b->device = device;
  b->n_rods = n_rods;
  b->damped = damped != 0;
  b->n_elems.assign(n_elems, n_elems + n_rods);
  uint64_t s = 0;
  for (uint64_t r = 0; r < n_rods; ++r) {
    b->starts.push_back(s);
    if (r + 1 < n_rods) b->ghost_nodes.push_back(s + n_elems[r] + 1);
    s += n_elems[r] + 2;
  }
  b->E = s - 2;
  int rc = EROD_OK;
  auto fail = [&](int code) {
    erod_batch_destroy(b);
    return code;
  };
  if ((rc = hip_check(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate"))) return fail(rc);
  for (int f = 0; f < EROD_N_FIELDS; ++f) {
    const size_t bytes = sizeof(double) * field_info(f).ncomp * b->field_slots(f);
    if ((rc = hip_check(hipMalloc(&b->d_field[f], bytes), "hipMalloc field"))) return fail(rc);
    if ((rc = hip_check(hipMemsetAsync(b->d_field[f], 0, bytes, b->stream), "hipMemset field"))) return fail(rc);
  }
  if (!b->ghost_nodes.empty()) {
    if ((rc = hip_check(hipMalloc(&b->d_ghost_nodes, sizeof(uint64_t) * b->ghost_nodes.size()), "hipMalloc ghosts"))) return fail(rc);
    if ((rc = hip_check(hipMemcpyAsync(b->d_ghost_nodes, b->ghost_nodes.data(), sizeof(uint64_t) * b->ghost_nodes.size(),
                                       hipMemcpyHostToDevice, b->stream), "copy ghosts"))) return fail(rc);
  }
  if ((rc = hip_check(hipMalloc(&b->d_node_is_ghost, b->E + 1), "hipMalloc ghost flags"))) return fail(rc);
  if ((rc = ops_mark_ghost_nodes(b))) return fail(rc);
  b->n_active = b->n_elems;
  b->steps_left.assign(n_rods, kNoBudget);
  b->rods.resize(n_rods);
  for (uint64_t r = 0; r < n_rods; ++r) {
    RodInfo& ri = b->rods[r];
    std::memset(&ri, 0, sizeof(ri));
    ri.start = b->starts[r];
    ri.n = static_cast<uint32_t>(n_elems[r]);
  }
  if ((rc = hip_check(hipMalloc(&b->d_rods, sizeof(RodInfo) * n_rods), "hipMalloc rods"))) return fail(rc);
  if ((rc = hip_check(hipMalloc(&b->d_time, sizeof(double) * (n_rods + 1)), "hipMalloc time"))) return fail(rc);
  if ((rc = hip_check(hipMemsetAsync(b->d_time, 0, sizeof(double) * (n_rods + 1), b->stream), "hipMemset time"))) return fail(rc);
  if ((rc = hip_check(hipMalloc(&b->d_flags, sizeof(int) * n_rods), "hipMalloc flags"))) return fail(rc);
  if ((rc = hip_check(hipEventCreate(&b->ev0), "hipEventCreate"))) return fail(rc);
  if ((rc = hip_check(hipEventCreate(&b->ev1), "hipEventCreate"))) return fail(rc);
  if ((rc = hip_check(hipStreamCreateWithFlags(&b->stream2, hipStreamNonBlocking), "hipStreamCreate"))) return fail(rc);
  if ((rc = hip_check(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming), "hipEventCreate"))) return fail(rc);
  if ((rc = hip_check(hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming), "hipEventCreate"))) return fail(rc);
  if ((rc = erod_fill_ghosts(b))) return fail(rc);
  if ((rc = hip_check(hipStreamSynchronize(b->stream), "sync"))) return fail(rc);
  *out = b;
  return EROD_OK;
}

int erod_batch_destroy(erod_batch* b) {
  if (!b) return EROD_OK;
  if (b->stream) hipStreamSynchronize(b->stream);
  if (b->stream2) hipStreamSynchronize(b->stream2);
  if (b->graph_exec) hipGraphExecDestroy(b->graph_exec);
  for (int f = 0; f < EROD_N_FIELDS; ++f) hipFree(b->d_field[f]);
  hipFree(b->d_ghost_nodes);
  hipFree(b->d_node_is_ghost);
  hipFree(b->d_steps_left);
  hipFree(b->d_copy_pairs);
  hipFree(b->d_rods);
  hipFree(b->d_rod_stage);
  hipFree(b->d_rod_stage_idx);
  hipFree(b->d_time);
  hipFree(b->d_flags);
  hipFree(b->d_const32);
  for (TileGroup& g : b->plan.g) {
    hipFree(g.d_tiles);
    hipFree(g.d_tile_rods);
    hipFree(g.d_tile_lane0);
    hipFree(g.d_segs);
    hipFree(g.d_seg_rods);
  }
  hipFree(b->d_time_next);
  for (double* p : b->d_shadow) hipFree(p);
  if (b->ev0) hipEventDestroy(b->ev0);
  if (b->ev1) hipEventDestroy(b->ev1);
  if (b->ev_fork) hipEventDestroy(b->ev_fork);
  if (b->ev_join) hipEventDestroy(b->ev_join);
  if (b->stream2) hipStreamDestroy(b->stream2);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return EROD_OK;
}

int erod_batch_device(const erod_batch* b, int* device) {
  if (check_batch(b) || !device) return EROD_ERR_INVALID;
  *device = b->device;
  return EROD_OK;
}

int erod_batch_slots(const erod_batch* b, uint64_t* element_slots) {
  if (check_batch(b) || !element_slots) return EROD_ERR_INVALID;
  *element_slots = b->E;
  return EROD_OK;
}

int erod_batch_rods(const erod_batch* b, uint64_t* n_rods, uint64_t* capacity) {
  if (check_batch(b) || !n_rods) return EROD_ERR_INVALID;
  *n_rods = b->n_rods;
  if (capacity)
    for (uint64_t r = 0; r < b->n_rods; ++r) capacity[r] = b->n_elems[r];
  return EROD_OK;
}

int erod_rod_start(const erod_batch* b, uint64_t rod, uint64_t* start) {
  if (check_batch(b) || !start) return EROD_ERR_INVALID;
  if (rod >= b->n_rods) {
    set_error("rod index out of range");
    return EROD_ERR_INVALID;
  }
  *start = b->starts[rod];
  return EROD_OK;
}

int erod_upload_field(erod_batch* b, int field, const double* host) {
  if (check_batch(b) || check_field(field) || !host) return EROD_ERR_INVALID;
  const size_t bytes = sizeof(double) * field_info(field).ncomp * b->field_slots(field);
  EROD_HIP(hipMemcpyAsync(b->d_field[field], host, bytes, hipMemcpyHostToDevice, b->stream));
  EROD_HIP(hipStreamSynchronize(b->stream));
  mark_field_written(b, field);
  return EROD_OK;
}

int erod_download_field(erod_batch* b, int field, double* host) {
  if (check_batch(b) || check_field(field) || !host) return EROD_ERR_INVALID;
  const size_t bytes = sizeof(double) * field_info(field).ncomp * b->field_slots(field);
  EROD_HIP(hipMemcpyAsync(host, b->d_field[field], bytes, hipMemcpyDeviceToHost, b->stream));
  EROD_HIP(hipStreamSynchronize(b->stream));
  return EROD_OK;
}

static uint64_t rod_count(const erod_batch* b, int field, uint64_t rod) {
  const int st = field_info(field).stagger;
  const uint64_t n = b->n_active[rod];
  return st == EROD_NODE ? n + 1 : (st == EROD_ELEMENT ? n : n - 1);
}

int erod_upload_rod(erod_batch* b, int field, uint64_t rod, const double* host) {
  if (check_batch(b) || check_field(field) || !host || rod >= b->n_rods || b->n_active[rod] == 0) {
    set_error("erod_upload_rod: bad argument (or an empty rod slot)");
    return EROD_ERR_INVALID;
  }
  const uint64_t cnt = rod_count(b, field, rod), slots = b->field_slots(field), s = b->starts[rod];
  const int nc = field_info(field).ncomp;
  // 2D copy: nc lines of cnt values, pitch = slots
  EROD_HIP(hipMemcpy2DAsync(b->d_field[field] + s, slots * sizeof(double), host, cnt * sizeof(double), cnt * sizeof(double),
                            nc, hipMemcpyHostToDevice, b->stream));
  EROD_HIP(hipStreamSynchronize(b->stream));
  mark_field_written(b, field);
  return EROD_OK;
}

int erod_download_rod(erod_batch* b, int field, uint64_t rod, double* host) {
  if (check_batch(b) || check_field(field) || !host || rod >= b->n_rods || b->n_active[rod] == 0) {
    set_error("erod_download_rod: bad argument (or an empty rod slot)");
    return EROD_ERR_INVALID;
  }
  const uint64_t cnt = rod_count(b, field, rod), slots = b->field_slots(field), s = b->starts[rod];
  const int nc = field_info(field).ncomp;
  EROD_HIP(hipMemcpy2DAsync(host, cnt * sizeof(double), b->d_field[field] + s, slots * sizeof(double), cnt * sizeof(double),
                            nc, hipMemcpyDeviceToHost, b->stream));
  EROD_HIP(hipStreamSynchronize(b->stream));
  return EROD_OK;
}

int erod_fill_ghosts(erod_batch* b) {
  if (check_batch(b)) return EROD_ERR_INVALID;
  for (int f = 0; f < EROD_N_FIELDS; ++f) {
    const int rc = ops_fill_ghosts(b, f);
    if (rc) return rc;
  }
  return EROD_OK;
}

int erod_kinematic_update(erod_batch* b, double dt) {
  if (int rc = check_plain(b)) return rc;
  return ops_kinematic_update(b, dt);
}

int erod_kinematic_update_by(erod_batch* b, double c) {
  if (int rc = check_plain(b)) return rc;
  return ops_kinematic_update_by(b, c);
}

int erod_fill_field(erod_batch* b, int field, double value) {
  if (check_batch(b) || check_field(field)) return EROD_ERR_INVALID;
  int rc = ops_fill_field(b, field, value);
  if (rc == EROD_OK) mark_field_written(b, field);
  return rc;
}

int erod_compute_internal_dynamics(erod_batch* b, double time) {
  (void)time;
  if (int rc = check_plain(b)) return rc;
  return ops_compute_internal_dynamics(b);
}

int erod_dynamic_update(erod_batch* b, double dt) {
  if (int rc = check_plain(b)) return rc;
  return ops_dynamic_update(b, dt);
}

int erod_add_clamp(erod_batch* b, uint64_t rod, uint64_t index, const double x[3], const double Q[9]) {
  if (check_batch(b) || rod >= b->n_rods || !x || !Q) {
    set_error("erod_add_clamp: bad argument");
    return EROD_ERR_INVALID;
  }
  if (index > b->n_active[rod]) {
    set_error("erod_add_clamp: index past the rod's last node");
    return EROD_ERR_INVALID;
  }
  RodInfo& ri = b->rods[rod];
  const uint64_t n = b->n_active[rod];
  // clamps at n - 1 and n would both hold element n - 1
  for (int c = 0; c < 2; ++c)
    if ((ri.flags & clamp_bit(c)) && ((index == n && ri.clamp_idx[c] + 1 == n) || (index + 1 == n && ri.clamp_idx[c] == n))) {
      set_error("erod_add_clamp: clamps at n - 1 and n would both hold the rod's last element");
      return EROD_ERR_INVALID;
    }
  // a clamp at an index that has one replaces it; otherwise it takes a free one of the two
  int k = -1;
  for (int c = 0; c < 2 && k < 0; ++c)
    if ((ri.flags & clamp_bit(c)) && ri.clamp_idx[c] == index) k = c;
  for (int c = 0; c < 2 && k < 0; ++c)
    if (!(ri.flags & clamp_bit(c))) k = c;
  if (k < 0) {
    set_error("erod_add_clamp: a rod takes at most two clamps");
    return EROD_ERR_UNSUPPORTED;
  }
  ri.flags |= clamp_bit(k);
  ri.clamp_idx[k] = static_cast<uint32_t>(index);
  std::memcpy(ri.clamp[k].x, x, sizeof(ri.clamp[k].x));
  std::memcpy(ri.clamp[k].q, Q, sizeof(ri.clamp[k].q));
  b->has_clamps = true;
  mark_rod(b, rod);
  b->graph_precision = -1;
  return EROD_OK;
}

int erod_add_end_force(erod_batch* b, uint64_t rod, const double F[3], int ramp, double start, double interval) {
  if (check_batch(b) || rod >= b->n_rods || !F || ramp < 0 || ramp > 2) {
    set_error("erod_add_end_force: bad argument");
    return EROD_ERR_INVALID;
  }
  // a ramp divides by its interval: an empty one would give 0/0 at t = start
  if (ramp != EROD_RAMP_NONE && !(interval > 0.0 && std::isfinite(interval) && std::isfinite(start))) {
    set_error("erod_add_end_force: a ramped force needs a positive, finite interval and a finite start");
    return EROD_ERR_INVALID;
  }
  RodInfo& ri = b->rods[rod];
  ri.flags |= 2u;
  std::memcpy(ri.force, F, sizeof(ri.force));
  ri.ramp = ramp;
  ri.ramp_start = start;
  ri.ramp_interval = interval;
  b->has_forces = true;
  mark_rod(b, rod);
  b->graph_precision = -1;
  return EROD_OK;
}

int erod_set_gravity(erod_batch* b, const double g[3]) {
  if (check_batch(b) || !g) return EROD_ERR_INVALID;
  b->env.gravity = 1;
  std::memcpy(b->env.g, g, sizeof(b->env.g));
  b->graph_precision = -1;
  return EROD_OK;
}

int erod_set_plane_contact(erod_batch* b, double k, double c, double mu, double ct, double z0, double offset) {
  if (check_batch(b)) return EROD_ERR_INVALID;
  b->env.contact = 1;
  b->env.k = k;
  b->env.c = c;
  b->env.mu = mu;
  b->env.ct = ct;
  b->env.z0 = z0;
  b->env.offset = offset;
  b->graph_precision = -1;
  return EROD_OK;
}

int erod_set_plane_normal(erod_batch* b, const double n[3]) {
  if (check_batch(b) || !n) return EROD_ERR_INVALID;
  // scaled to unit length; (0, 0, 1) and the other axes stay exact
  const double len = std::sqrt((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]);
  if (!(len > 0.0) || !std::isfinite(len)) {
    set_error("erod_set_plane_normal: the normal must be finite and non-zero");
    return EROD_ERR_INVALID;
  }
  for (int c = 0; c < 3; ++c) b->env.n[c] = n[c] / len;
  b->graph_precision = -1;
  return EROD_OK;
}

int erod_apply_post_kinematic(erod_batch* b, double time) {
  (void)time;
  if (int rc = check_plain(b)) return rc;
  int rc = sync_rods(b);
  return rc ? rc : ops_apply_post_kinematic(b);
}

int erod_apply_pre_dynamic(erod_batch* b, double time) {
  if (int rc = check_plain(b)) return rc;
  int rc = sync_rods(b);
  return rc ? rc : ops_apply_pre_dynamic(b, time);
}

int erod_apply_post_dynamic(erod_batch* b, double time) {
  (void)time;
  if (int rc = check_plain(b)) return rc;
  int rc = sync_rods(b);
  return rc ? rc : ops_apply_post_dynamic(b);
}

int erod_step(erod_batch* b, double* time, double dt, uint64_t n) {
  if (!time) {
    set_error("erod_step: null time");
    return EROD_ERR_INVALID;
  }
  if (int rc = check_plain(b)) return rc;
  int rc = sync_rods(b);
  if (rc) return rc;
  double t = *time;
  for (uint64_t i = 0; i < n; ++i) {
    if ((rc = ops_kinematic_update(b, dt))) return rc;
    t += 0.5 * dt;
    if ((rc = ops_apply_post_kinematic(b))) return rc;
    if ((rc = ops_compute_internal_dynamics(b))) return rc;
    if ((rc = ops_apply_pre_dynamic(b, t))) return rc;
    if ((rc = ops_dynamic_update(b, dt))) return rc;
    if ((rc = ops_apply_post_dynamic(b))) return rc;
    t += 0.5 * dt;
    if ((rc = ops_kinematic_update(b, dt))) return rc;
  }
  *time = t;
  // every rod clock follows the caller's clock, so checkpoints carry it
  return ops_set_rod_clocks(b, t);
}

int erod_set_rod_dt(erod_batch* b, const double* dt_per_rod) {
  if (check_batch(b) || !dt_per_rod) return EROD_ERR_INVALID;
  // all or nothing: a rejected call leaves every rod's dt as it was
  for (uint64_t r = 0; r < b->n_rods; ++r)
    if (!(dt_per_rod[r] > 0.0) || !std::isfinite(dt_per_rod[r])) {
      set_error("erod_set_rod_dt: dt must be positive and finite");
      return EROD_ERR_INVALID;
    }
  for (uint64_t r = 0; r < b->n_rods; ++r) b->rods[r].dt = dt_per_rod[r];
  b->rods_dirty = true;
  return EROD_OK;
}

int erod_get_rod_time(erod_batch* b, double* t) {
  if (check_batch(b) || !t) return EROD_ERR_INVALID;
  EROD_HIP(hipMemcpyAsync(t, b->d_time, sizeof(double) * b->n_rods, hipMemcpyDeviceToHost, b->stream));
  EROD_HIP(hipStreamSynchronize(b->stream));
  return EROD_OK;
}

int erod_set_rod_time(erod_batch* b, const double* t) {
  if (check_batch(b) || !t) return EROD_ERR_INVALID;
  EROD_HIP(hipMemcpyAsync(b->d_time, t, sizeof(double) * b->n_rods, hipMemcpyHostToDevice, b->stream));
  EROD_HIP(hipStreamSynchronize(b->stream));
  return EROD_OK;
}

static int fused_common(erod_batch* b, int precision) {
  if (precision != EROD_FP64 && precision != EROD_FP32_BULK) {
    set_error("unknown precision");
    return EROD_ERR_INVALID;
  }
  for (uint64_t r = 0; r < b->n_rods; ++r)
    if (b->n_active[r] && !(b->rods[r].dt > 0.0)) {
      set_error("fused path: set a positive dt for every rod with erod_set_rod_dt");
      return EROD_ERR_INVALID;
    }
  int rc = ensure_damping(b);
  if (rc) return rc;
  if (b->budgets && b->budgets_dirty) {
    EROD_HIP(hipMemcpyAsync(b->d_steps_left, b->steps_left.data(), sizeof(uint32_t) * b->n_rods, hipMemcpyHostToDevice,
                            b->stream));
    // the host mirror is spent as soon as the launches are queued: the copy
    // must have read it before
    EROD_HIP(hipStreamSynchronize(b->stream));
    b->budgets_dirty = false;
  }
  return fused_prepare(b, precision);
}

// The kernel spends min(budget, s) steps of each planned rod's budget per
// launch of s steps; the host copy follows the same rule.  A rod that runs
// out leaves the plan at the next call.
static void spend_budgets(erod_batch* b, uint64_t s) {
  if (!b->budgets) return;
  bool emptied = false;
  for (uint64_t r = 0; r < b->n_rods; ++r) {
    uint32_t& l = b->steps_left[r];
    if (!b->n_active[r] || l == kNoBudget || l == 0) continue;
    l = l > s ? uint32_t(l - s) : 0u;
    emptied |= l == 0;
  }
  if (emptied) b->plan.valid = false;
}

int erod_set_rod_active_length(erod_batch* b, uint64_t rod, uint64_t n) {
  if (check_batch(b) || rod >= b->n_rods || n == 1 || n > b->n_elems[rod]) {
    set_error("erod_set_rod_active_length: rod out of range, or n not 0 and not in [2, capacity]");
    return EROD_ERR_INVALID;
  }
  b->n_active[rod] = n;
  b->rods[rod].n = static_cast<uint32_t>(n);
  mark_rod(b, rod);
  b->damping_dirty = true;
  b->all_full = b->n_active == b->n_elems;
  b->plan.valid = false;
  b->graph_precision = -1;
  return EROD_OK;
}

int erod_get_rod_active_length(const erod_batch* b, uint64_t rod, uint64_t* n) {
  if (check_batch(b) || rod >= b->n_rods || !n) return EROD_ERR_INVALID;
  *n = b->n_active[rod];
  return EROD_OK;
}

static int ensure_budget_buffer(erod_batch* b) {
  if (b->d_steps_left) return EROD_OK;
  EROD_HIP(hipMalloc(&b->d_steps_left, sizeof(uint32_t) * b->n_rods));
  b->budgets_dirty = true;
  return EROD_OK;
}

int erod_set_rod_step_budget(erod_batch* b, uint64_t rod, uint64_t steps) {
  if (check_batch(b) || rod >= b->n_rods) return EROD_ERR_INVALID;
  if (int rc = ensure_budget_buffer(b)) return rc;
  const uint32_t v = steps >= kNoBudget ? kNoBudget : static_cast<uint32_t>(steps);
  if ((v == 0) != (b->steps_left[rod] == 0)) {
    b->plan.valid = false;
    b->graph_precision = -1;
  }
  b->steps_left[rod] = v;
  b->budgets = true;
  b->budgets_dirty = true;
  return EROD_OK;
}

int erod_get_rod_step_budget(const erod_batch* b, uint64_t rod, uint64_t* steps) {
  if (check_batch(b) || rod >= b->n_rods || !steps) return EROD_ERR_INVALID;
  *steps = b->steps_left[rod] == kNoBudget ? UINT64_MAX : b->steps_left[rod];
  return EROD_OK;
}

int erod_copy_rods(erod_batch* dst, const uint64_t* dst_rods, erod_batch* src, const uint64_t* src_rods, uint64_t count) {
  if (check_batch(dst) || check_batch(src) || (count && (!dst_rods || !src_rods)) || dst == src) {
    set_error("erod_copy_rods: bad argument (two distinct batches needed)");
    return EROD_ERR_INVALID;
  }
  if (dst->device != src->device || dst->damped != src->damped) {
    set_error("erod_copy_rods: batches on different devices or of different damping models");
    return EROD_ERR_INVALID;
  }
  {
    std::vector<uint64_t> d(dst_rods, dst_rods + count);
    std::sort(d.begin(), d.end());
    if (std::adjacent_find(d.begin(), d.end()) != d.end()) {
      set_error("erod_copy_rods: a destination slot appears twice");
      return EROD_ERR_INVALID;
    }
  }
  std::vector<CopyPair> pairs(count);
  for (uint64_t i = 0; i < count; ++i) {
    const uint64_t d = dst_rods[i], r = src_rods[i];
    if (d >= dst->n_rods || r >= src->n_rods || src->n_active[r] == 0 || src->n_active[r] > dst->n_elems[d]) {
      set_error("erod_copy_rods: rod out of range, empty source, or source longer than the destination slot");
      return EROD_ERR_INVALID;
    }
    pairs[i] = CopyPair{src->starts[r], dst->starts[d], static_cast<uint32_t>(src->n_active[r]), static_cast<uint32_t>(r),
                        static_cast<uint32_t>(d), 0u};
  }
  // the per-rod damping rates travel in the rod records
  if (int rc = ensure_damping(src)) return rc;
  if (int rc = serve_copy_rods(dst, src, pairs)) return rc;
  for (const CopyPair& p : pairs) {
    RodInfo ri = src->rods[p.src_rod];
    ri.start = dst->starts[p.dst_rod];
    dst->rods[p.dst_rod] = ri;
    dst->n_active[p.dst_rod] = p.n;
    dst->steps_left[p.dst_rod] = kNoBudget;
    dst->has_clamps |= (ri.flags & kClampBits) != 0;
    dst->has_forces |= (ri.flags & 2u) != 0;
    mark_rod(dst, p.dst_rod);
  }
  dst->budgets_dirty = true;
  dst->all_full = dst->n_active == dst->n_elems;
  dst->plan.valid = false;
  dst->graph_precision = -1;
  return EROD_OK;
}

int erod_set_fused_config(erod_batch* b, int lanes_per_tile, int elements_per_thread, int steps_per_launch) {
  if (check_batch(b) || steps_per_launch < 0 || lanes_per_tile < 0 || elements_per_thread < 0) {
    set_error("erod_set_fused_config: negative argument");
    return EROD_ERR_INVALID;
  }
  b->plan.lanes_per_tile = lanes_per_tile;
  b->plan.elements_per_thread = elements_per_thread;
  b->plan.steps_per_launch = steps_per_launch;
  b->plan.valid = false;
  b->graph_precision = -1;
  return EROD_OK;
}

int erod_set_fused_split(erod_batch* b, int lanes_per_tile, int elements_per_thread) {
  if (check_batch(b) || lanes_per_tile < 0 || elements_per_thread < 0) {
    set_error("erod_set_fused_split: negative argument");
    return EROD_ERR_INVALID;
  }
  b->plan.split_lanes = lanes_per_tile;
  b->plan.split_epl = elements_per_thread;
  b->plan.valid = false;
  b->graph_precision = -1;
  return EROD_OK;
}

int erod_get_fused_split(const erod_batch* b, int* lanes, int* elements_per_thread, uint64_t* tiles) {
  if (check_batch(b)) return EROD_ERR_INVALID;
  const int si = b->plan.valid ? b->plan.g[1].shape : -1;
  if (lanes) *lanes = si >= 0 ? fused_shape_lanes(si) : 0;
  if (elements_per_thread) *elements_per_thread = si >= 0 ? fused_shape_epl(si) : 0;
  if (tiles) *tiles = si >= 0 ? b->plan.g[1].n_tiles : 0;
  return EROD_OK;
}

int erod_set_fused_pieces(erod_batch* b, int lanes_per_tile, int elements_per_thread, int halo) {
  if (check_batch(b) || lanes_per_tile < 0 || elements_per_thread < 0 || halo < 0 || halo == 1) {
    set_error("erod_set_fused_pieces: negative argument, or a halo of 1");
    return EROD_ERR_INVALID;
  }
  b->plan.seg_lanes = lanes_per_tile;
  b->plan.seg_epl = elements_per_thread;
  b->plan.halo = halo ? halo : 16;
  b->plan.valid = false;
  b->graph_precision = -1;
  return EROD_OK;
}

int erod_get_fused_pieces(const erod_batch* b, int* lanes, int* elements_per_thread, int* halo, uint64_t* tiles,
                          uint64_t* pieces) {
  if (check_batch(b)) return EROD_ERR_INVALID;
  const TileGroup& g = b->plan.g[2];
  const int si = b->plan.valid ? g.shape : -1;
  if (lanes) *lanes = si >= 0 ? fused_shape_lanes(si) : 0;
  if (elements_per_thread) *elements_per_thread = si >= 0 ? fused_shape_epl(si) : 0;
  if (halo) *halo = b->plan.halo;
  if (tiles) *tiles = si >= 0 ? g.n_tiles : 0;
  if (pieces) *pieces = si >= 0 ? g.n_segs : 0;
  return EROD_OK;
}

int erod_get_fused_config(const erod_batch* b, int* lanes, int* elements_per_thread, int* steps) {
  if (check_batch(b)) return EROD_ERR_INVALID;
  const int si = b->plan.valid ? b->plan.g[0].shape : -1;
  if (lanes) *lanes = si >= 0 ? fused_shape_lanes(si) : 0;
  if (elements_per_thread) *elements_per_thread = si >= 0 ? fused_shape_epl(si) : 0;
  if (steps) *steps = b->plan.steps_per_launch > 0 ? b->plan.steps_per_launch : 100;
  return EROD_OK;
}

int erod_autotune_fused(erod_batch* b, int precision, int trial_steps, double* best_ms_per_step) {
  if (check_batch(b) || trial_steps <= 0) return EROD_ERR_INVALID;
  int rc = fused_common(b, precision);
  if (rc) return rc;
  return fused_autotune(b, precision, trial_steps, best_ms_per_step);
}

// the scheduled re-orthonormalisation after a launch of s steps: when the
// batch's step count reaches a multiple of the interval during the launch
static int after_launch(erod_batch* b, uint64_t s) {
  if (b->reortho_every == 0) return EROD_OK;
  b->reortho_phase += s;
  if (b->reortho_phase < b->reortho_every) return EROD_OK;
  b->reortho_phase %= b->reortho_every;
  return fused_reorthonormalize_live(b);
}

int erod_step_fused(erod_batch* b, int precision, uint64_t n) {
  if (check_batch(b)) return EROD_ERR_INVALID;
  int rc = fused_common(b, precision);
  if (rc) return rc;
  const uint64_t spl = b->plan.steps_per_launch > 0 ? uint64_t(b->plan.steps_per_launch) : 100;
  uint64_t done = 0;
  b->last_launches = 0;
  EROD_HIP(hipEventRecord(b->ev0, b->stream));
  while (done < n) {
    const uint64_t s = std::min<uint64_t>(spl, n - done);
    if ((rc = fused_launch(b, precision, static_cast<int>(s), false))) return rc;
    spend_budgets(b, s);
    if ((rc = after_launch(b, s))) return rc;
    ++b->last_launches;
    done += s;
  }
  EROD_HIP(hipEventRecord(b->ev1, b->stream));
  return EROD_OK;
}

// capture (or keep) the graph of one chunk of `chunk` fused steps, with the
// scheduled re-orthonormalisations that fall into it from the current phase
// (a chunk that is a multiple of the interval starts every replay at the
// same phase; otherwise a replay at another phase captures again)
static int ensure_graph(erod_batch* b, int precision, uint64_t chunk) {
  const uint64_t spl = b->plan.steps_per_launch > 0 ? uint64_t(b->plan.steps_per_launch) : 100;
  const bool same_schedule =
      b->graph_reortho[0] == b->reortho_every && (b->reortho_every == 0 || b->graph_reortho[1] == b->reortho_phase);
  if (b->graph_precision != precision || b->graph_chunk != chunk || !b->graph_exec || !same_schedule) {
    if (b->graph_exec) {
      // (an earlier replay may still run on the stream)
      EROD_HIP(hipStreamSynchronize(b->stream));
      hipGraphExecDestroy(b->graph_exec);
      b->graph_exec = nullptr;
    }
    hipGraph_t g = nullptr;
    const uint64_t phase0 = b->reortho_phase;
    EROD_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    uint64_t done = 0, launches = 0;
    int lrc = EROD_OK;
    while (done < chunk && lrc == EROD_OK) {
      const uint64_t s = std::min<uint64_t>(spl, chunk - done);
      lrc = fused_launch(b, precision, static_cast<int>(s), false);
      if (lrc == EROD_OK) lrc = after_launch(b, s);
      done += s;
      ++launches;
    }
    // the capture ran nothing: the phase moves when the graph is replayed
    const uint64_t phase1 = b->reortho_phase;
    b->reortho_phase = phase0;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (lrc || e != hipSuccess) {
      if (g) hipGraphDestroy(g);
      return lrc ? lrc : hip_check(e, "hipStreamEndCapture");
    }
    size_t nodes = 0;
    e = hipGraphGetNodes(g, nullptr, &nodes);
    if (e == hipSuccess) e = hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
      b->graph_exec = nullptr;
      return hip_check(e, "hipGraphGetNodes / hipGraphInstantiate");
    }
    b->graph_precision = precision;
    b->graph_chunk = chunk;
    b->graph_launches = launches;
    b->graph_one_node = nodes == 1;
    b->graph_reortho[0] = b->reortho_every;
    b->graph_reortho[1] = phase0;
    b->graph_phase_after = phase1;
  }
  return EROD_OK;
}

int erod_capture_fused_graph(erod_batch* b, int precision, uint64_t chunk) {
  if (check_batch(b) || chunk == 0) {
    set_error("erod_capture_fused_graph: chunk must be positive");
    return EROD_ERR_INVALID;
  }
  int rc = fused_common(b, precision);
  return rc ? rc : ensure_graph(b, precision, chunk);
}

int erod_step_fused_graph(erod_batch* b, int precision, uint64_t n, uint64_t chunk) {
  if (check_batch(b) || chunk == 0) {
    set_error("erod_step_fused_graph: chunk must be positive");
    return EROD_ERR_INVALID;
  }
  int rc = fused_common(b, precision);
  if (rc) return rc;
  const uint64_t spl = b->plan.steps_per_launch > 0 ? uint64_t(b->plan.steps_per_launch) : 100;
  const uint64_t whole = n / chunk, rest = n % chunk;
  if (whole > 0 && (rc = ensure_graph(b, precision, chunk))) return rc;
  EROD_HIP(hipEventRecord(b->ev0, b->stream));
  for (uint64_t i = 0; i < whole; ++i) {
    if (i > 0 && b->reortho_every && (rc = ensure_graph(b, precision, chunk))) return rc;
    if (b->graph_one_node) {
      // a chunk of one kernel launch goes out as that launch: there is no
      // launch work for a graph to save, and the replay of a one-node graph
      // takes the device longer than the launch (7 us more on a 0.32 ms
      // launch of 4,096 rods)
      if ((rc = fused_launch(b, precision, static_cast<int>(chunk), false))) return rc;
      spend_budgets(b, chunk);
      if ((rc = after_launch(b, chunk))) return rc;
      continue;
    }
    EROD_HIP(hipGraphLaunch(b->graph_exec, b->stream));
    for (uint64_t d = 0; d < chunk; d += spl) spend_budgets(b, std::min<uint64_t>(spl, chunk - d));
    b->reortho_phase = b->graph_phase_after;
  }
  uint64_t launches = whole * b->graph_launches;
  // the steps that do not fill a chunk: direct launches of the same length
  // a chunk's launches have (spl, the last one shorter), so the launch
  // segmentation equals that of erod_step_fused over n steps whenever chunk
  // is a multiple of spl
  for (uint64_t done = 0; done < rest;) {
    const uint64_t s = std::min<uint64_t>(spl, rest - done);
    if ((rc = fused_launch(b, precision, static_cast<int>(s), false))) return rc;
    spend_budgets(b, s);
    if ((rc = after_launch(b, s))) return rc;
    ++launches;
    done += s;
  }
  EROD_HIP(hipEventRecord(b->ev1, b->stream));
  b->last_launches = launches;
  return EROD_OK;
}

int erod_reorthonormalize(erod_batch* b) {
  if (check_batch(b)) return EROD_ERR_INVALID;
  return fused_reorthonormalize(b);
}

int erod_set_reortho_interval(erod_batch* b, uint64_t every) {
  if (check_batch(b)) return EROD_ERR_INVALID;
  b->reortho_every = every;
  b->reortho_phase = 0;
  return EROD_OK;
}

int erod_get_reortho_interval(const erod_batch* b, uint64_t* every, uint64_t* phase) {
  if (check_batch(b)) return EROD_ERR_INVALID;
  if (every) *every = b->reortho_every;
  if (phase) *phase = b->reortho_phase;
  return EROD_OK;
}

int erod_last_kernel_ms(const erod_batch* b, double* ms, uint64_t* launches) {
  if (check_batch(b)) return EROD_ERR_INVALID;
  if (b->last_launches == 0) {
    set_error("erod_last_kernel_ms: no fused call has run on this batch");
    return EROD_ERR_INVALID;
  }
  float f = 0.0f;
  EROD_HIP(hipEventSynchronize(b->ev1));
  EROD_HIP(hipEventElapsedTime(&f, b->ev0, b->ev1));
  if (ms) *ms = f / static_cast<double>(b->last_launches);
  if (launches) *launches = b->last_launches;
  return EROD_OK;
}

int erod_synchronize(erod_batch* b) {
  if (check_batch(b)) return EROD_ERR_INVALID;
  EROD_HIP(hipStreamSynchronize(b->stream));
  return EROD_OK;
}

int erod_check_finite(erod_batch* b, uint8_t* rod_flags) {
  if (check_batch(b) || !rod_flags) return EROD_ERR_INVALID;
  return fused_check_finite(b, rod_flags);
}

int erod_check_finite_noalloc(erod_batch* b, uint8_t* rod_flags) {
  if (check_batch(b) || !rod_flags) return EROD_ERR_INVALID;
  return fused_check_finite_noalloc(b, rod_flags);
}

}  // extern "C"
