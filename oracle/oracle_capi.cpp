// ORACLE — TEST INFRASTRUCTURE ONLY (see cosserat_oracle.hpp).
//
// A small C interface over the FP64 oracle so that the Python test suite and
// the benchmark's CPU-baseline leg can drive it through ctypes.  A "system" is
// a list of ghost-padded blocks (first-fit blocking like RestrictSize,
// elastica/Simulator/Configuration/ModuleConfiguration/BlockingPolicy.hpp:202-222,
// 291-321) plus hook descriptors; steps can be spread over a std::thread pool,
// one task per block as tbb::parallel_for_each does (elastica/Parallel/for_each.hpp:90-99).
#include <atomic>
#include <cfenv>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <thread>

#include "cosserat_oracle.hpp"

using T = double;
using Sys = oracle::System<T>;

namespace {
  // The reference's test runner enables floating-point exceptions, so an
  // invalid operation, a division by zero or an overflow on the path aborts
  // its tests (tests/Unit/RunTests.cpp).  Here the calls that compute on the
  // calling thread record the flags they raised instead; the tests read them.
  int g_fp_flags = 0;
  struct FpWatch {
    FpWatch() { std::feclearexcept(FE_ALL_EXCEPT); }
    ~FpWatch() { g_fp_flags |= std::fetestexcept(FE_INVALID | FE_DIVBYZERO | FE_OVERFLOW); }
  };

  // Worker threads that live for a whole stepping call, one task per block
  // per substep (tbb::parallel_for_each keeps its workers the same way); the
  // return of run() is the barrier between substeps.  Workers wait for the
  // next substep by spinning, then yielding, so a substep costs no thread
  // creation.
  class Pool {
   public:
    explicit Pool(int nthreads) : nt_(nthreads > 1 ? nthreads : 1) {
      for (int t = 1; t < nt_; ++t) th_.emplace_back([this] { work(); });
    }
    ~Pool() {
      stop_.store(true, std::memory_order_release);
      gen_.fetch_add(1, std::memory_order_acq_rel);
      for (auto& t : th_) t.join();
    }
    Pool(const Pool&) = delete;
    Pool& operator=(const Pool&) = delete;
    int threads() const { return nt_; }

    // f(i) for i in [0, n), spread over the workers and the calling thread
    template <typename F>
    void run(std::size_t n, F&& f) {
      if (nt_ == 1 || n <= 1) {
        for (std::size_t i = 0; i < n; ++i) f(i);
        return;
      }
      std::function<void(std::size_t)> job(std::ref(f));
      job_ = &job;
      n_ = n;
      next_.store(0, std::memory_order_relaxed);
      busy_.store(nt_ - 1, std::memory_order_relaxed);
      gen_.fetch_add(1, std::memory_order_acq_rel);  // release the workers
      drain();
      while (busy_.load(std::memory_order_acquire) != 0) std::this_thread::yield();
      job_ = nullptr;
    }

   private:
    void drain() {
      for (;;) {
        const std::size_t i = next_.fetch_add(1, std::memory_order_relaxed);
        if (i >= n_) break;
        (*job_)(i);
      }
    }
    void work() {
      uint64_t seen = 0;
      for (;;) {
        uint64_t g;
        int spins = 0;
        while ((g = gen_.load(std::memory_order_acquire)) == seen) {
          if (++spins > 2000) std::this_thread::yield();
        }
        seen = g;
        if (stop_.load(std::memory_order_acquire)) return;
        drain();
        busy_.fetch_sub(1, std::memory_order_acq_rel);
      }
    }
    int nt_;
    std::vector<std::thread> th_;
    std::atomic<uint64_t> gen_{0};
    std::atomic<bool> stop_{false};
    std::atomic<std::size_t> next_{0};
    std::atomic<int> busy_{0};
    std::function<void(std::size_t)>* job_ = nullptr;
    std::size_t n_ = 0;
  };
}  // namespace

extern "C" {

// Create a system of rods.  max_block_elems = 0 puts every rod into its own
// block; otherwise rods are packed first-fit into blocks holding at most that
// many elements (the reference default is 1024, elastica/Configuration/Systems.hpp:67).
void* oracle_system_create(const uint64_t* n_elems, uint64_t n_rods, uint64_t max_block_elems, int damped) {
  try {
    auto* s = new Sys();
    std::vector<std::vector<std::size_t>> groups;
    std::vector<std::size_t> fill;
    s->where.resize(n_rods);
    for (uint64_t r = 0; r < n_rods; ++r) {
      const std::size_t n = n_elems[r];
      std::size_t bi = groups.size();
      if (max_block_elems) {
        for (std::size_t g = 0; g < groups.size(); ++g)
          if (fill[g] + n <= max_block_elems) { bi = g; break; }
      }
      if (bi == groups.size()) { groups.emplace_back(); fill.push_back(0); }
      s->where[r] = {bi, groups[bi].size()};
      groups[bi].push_back(n);
      fill[bi] += n;
    }
    for (auto const& g : groups) s->blocks.emplace_back(g, damped != 0);
    return s;
  } catch (...) {
    return nullptr;
  }
}

void oracle_system_destroy(void* h) { delete static_cast<Sys*>(h); }

uint64_t oracle_system_n_blocks(void* h) { return static_cast<Sys*>(h)->blocks.size(); }

// Where a rod lives: block index, start slot inside that block.
void oracle_rod_location(void* h, uint64_t rod, uint64_t* block, uint64_t* start) {
  auto* s = static_cast<Sys*>(h);
  auto [bi, ri] = s->where[rod];
  *block = bi;
  *start = s->blocks[bi].starts[ri];
}

// Direct pointer to a field of one block: ncomp lines of `slots` values.
double* oracle_block_field(void* h, uint64_t block, const char* name, uint64_t* ncomp, uint64_t* slots) {
  auto* s = static_cast<Sys*>(h);
  const int f = oracle::field_id(name);
  if (f < 0 || block >= s->blocks.size()) return nullptr;
  auto& b = s->blocks[block];
  *ncomp = oracle::field_table()[f].ncomp;
  *slots = b.slots(f);
  return b.get(f);
}

int oracle_block_fill_ghosts(void* h, uint64_t block) {
  auto* s = static_cast<Sys*>(h);
  s->blocks[block].fill_all_ghosts();
  return 0;
}

int oracle_initialize_straight_rod(void* h, uint64_t rod, const double* params16) {
  // params16: radius, length, origin[3], direction[3], normal[3], density, youngs,
  // shear, force_damping, torque_damping
  auto* s = static_cast<Sys*>(h);
  oracle::StraightRodParams p{};
  p.radius = params16[0];
  p.length = params16[1];
  for (int i = 0; i < 3; ++i) {
    p.origin[i] = params16[2 + i];
    p.direction[i] = params16[5 + i];
    p.normal[i] = params16[8 + i];
  }
  p.density = params16[11];
  p.youngs = params16[12];
  p.shear = params16[13];
  p.force_damping = params16[14];
  p.torque_damping = params16[15];
  auto [bi, ri] = s->where[rod];
  oracle::initialize_straight_rod(s->blocks[bi], ri, p);
  return 0;
}

// Single component calls on every block (the FullRodTest InitializeStraightRod path).
int oracle_compute_internal_forces(void* h) {
  FpWatch w;
  for (auto& b : static_cast<Sys*>(h)->blocks) oracle::compute_internal_forces(b);
  return 0;
}
int oracle_compute_internal_torques(void* h) {
  FpWatch w;
  for (auto& b : static_cast<Sys*>(h)->blocks) oracle::compute_internal_torques(b);
  return 0;
}
int oracle_update_acceleration(void* h) {
  FpWatch w;
  for (auto& b : static_cast<Sys*>(h)->blocks) oracle::update_acceleration(b);
  return 0;
}

// the floating-point exception flags (FE_INVALID | FE_DIVBYZERO | FE_OVERFLOW)
// raised by the watched calls since the last read; steps on one thread only
int oracle_fp_flags(int clear) {
  const int f = g_fp_flags;
  if (clear) g_fp_flags = 0;
  return f;
}
int oracle_fe_invalid(void) { return FE_INVALID; }
int oracle_fe_divbyzero(void) { return FE_DIVBYZERO; }
int oracle_fe_overflow(void) { return FE_OVERFLOW; }
int oracle_kinematic_update(void* h, double dt) {
  for (auto& b : static_cast<Sys*>(h)->blocks) oracle::kinematic_update(b, dt);
  return 0;
}

int oracle_add_clamp(void* h, uint64_t rod, uint64_t index, const double* x3, const double* q9) {
  auto* s = static_cast<Sys*>(h);
  if (rod >= s->where.size()) return -1;
  auto [bi, ri] = s->where[rod];
  if (index > s->blocks[bi].n_elems[ri]) return -1;  // past the rod's last node
  oracle::Clamp c{};
  c.rod = rod;
  c.index = index;
  std::memcpy(c.x, x3, sizeof(c.x));
  std::memcpy(c.Q, q9, sizeof(c.Q));
  s->clamps.push_back(c);
  return 0;
}

int oracle_add_end_force(void* h, uint64_t rod, const double* F3, int ramp, double start, double interval) {
  auto* s = static_cast<Sys*>(h);
  if (rod >= s->where.size()) return -1;
  oracle::EndForce f{};
  f.rod = rod;
  std::memcpy(f.F, F3, sizeof(f.F));
  f.ramp = ramp;
  f.start = start;
  f.interval = interval;
  s->forces.push_back(f);
  return 0;
}

int oracle_set_gravity(void* h, const double* g3) {
  auto* s = static_cast<Sys*>(h);
  s->gravity = true;
  std::memcpy(s->g, g3, sizeof(s->g));
  return 0;
}

int oracle_set_plane_contact(void* h, double k, double c, double mu, double ct, double z0, double offset) {
  auto* s = static_cast<Sys*>(h);
  s->contact = true;
  // the plane's normal stays as set (oracle_set_plane_normal)
  s->plane.k = k;
  s->plane.c = c;
  s->plane.mu = mu;
  s->plane.ct = ct;
  s->plane.z0 = z0;
  s->plane.offset = offset;
  return 0;
}

// unit normal of the contact plane n . x = z0 (default (0, 0, 1)), scaled to
// unit length as the device library does; 1 if n is zero or not finite
int oracle_set_plane_normal(void* h, const double* n3) {
  auto* s = static_cast<Sys*>(h);
  const double len = std::sqrt((n3[0] * n3[0] + n3[1] * n3[1]) + n3[2] * n3[2]);
  if (!(len > 0.0) || !std::isfinite(len)) return 1;
  for (int c = 0; c < 3; ++c) s->plane.n[c] = n3[c] / len;
  return 0;
}

// the ramp of an end force at time t (0 none, 1 linear, 2 sin)
double oracle_ramp_value(int kind, double t, double start, double interval) {
  return oracle::ramp_value(kind, t, start, interval);
}

double oracle_time(void* h) { return static_cast<Sys*>(h)->time; }
void oracle_set_time(void* h, double t) { static_cast<Sys*>(h)->time = t; }

// n position-Verlet steps with a uniform dt (SymplecticSteppers.hpp:133-217).
int oracle_step(void* h, double dt, uint64_t n, int nthreads) {
  auto* s = static_cast<Sys*>(h);
  Pool pool(std::min<int>(nthreads, int(s->blocks.size())));
  auto fe = [&](auto&& fn) { pool.run(s->blocks.size(), [&](std::size_t i) { fn(s->blocks[i]); }); };
  FpWatch w;
  for (uint64_t i = 0; i < n; ++i) oracle::step(*s, dt, fe);
  return 0;
}

// n steps where every block advances with its own dt (rods are independent, so
// a per-rod dt is a per-block dt once rods with different dt live in different
// blocks).  Clamps and constant forces need no clock; ramped forces are refused.
int oracle_step_per_block_dt(void* h, const double* block_dt, uint64_t n, int nthreads) {
  auto* s = static_cast<Sys*>(h);
  if (!s->clamps.empty() || !s->forces.empty()) {
    // tip forces with a time ramp need one clock per rod; only constant ones here
    for (auto const& f : s->forces)
      if (f.ramp != 0) return -1;
  }
  Pool pool(std::min<int>(nthreads, int(s->blocks.size())));
  auto run = [&](auto&& fn) { pool.run(s->blocks.size(), [&](std::size_t i) { fn(s->blocks[i], block_dt[i]); }); };
  for (uint64_t it = 0; it < n; ++it) {
    run([](oracle::Block<T>& b, double dt) { oracle::kinematic_update(b, dt); });
    oracle::apply_post_kinematic(*s);
    run([](oracle::Block<T>& b, double) { oracle::compute_internal_dynamics(b); });
    oracle::apply_pre_dynamic(*s, s->time);
    run([](oracle::Block<T>& b, double dt) { oracle::dynamic_update(b, dt); });
    oracle::apply_post_dynamic(*s);
    run([](oracle::Block<T>& b, double dt) { oracle::kinematic_update(b, dt); });
  }
  return 0;
}

// Primitive kernels exposed for the closed-form tests.
void oracle_two_point_difference(const double* in, uint64_t nin, double* out) {
  oracle::two_point_difference(in, nin, out, nin + 1);
}
void oracle_quadrature(const double* in, uint64_t nin, double* out) {
  oracle::quadrature(in, nin, out, nin + 1);
}
void oracle_inv_rotate_divide(const double* Q, uint64_t ne, const double* D0, double* kappa) {
  oracle::inv_rotate_divide(Q, ne, D0, kappa, ne - 1);
}
void oracle_so3_add_assign(double* Q, uint64_t ne, const double* u) {
  for (uint64_t k = 0; k < ne; ++k) {
    const double uk[3] = {u[k], u[ne + k], u[2 * ne + k]};
    oracle::so3_add_assign_one(Q, ne, k, uk);
  }
}

}  // extern "C"
