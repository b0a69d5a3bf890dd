// C++20 host layer over include/erod.h: a device-resident Cosserat-rod block
// with the per-block interface the reference's position-Verlet stepper
// drives, so a block marked device-resident drops in where a TBB-dispatched
// CPU block runs today.
//
//   reference                                       here
//   blocks::Block<CosseratRod> + emplace            erod::DeviceBlock(n_elems)
//   blocks::get<tags::Position>(slice) = X          block.set(tags::Position, rod, X)
//   KinematicSubStep (PositionVerlet.hpp:60-69)     block.kinematic_update(dt)
//   compute_internal_dynamics(t) with zeroed        block.compute_internal_dynamics(t)
//     external loads (SymplecticSteppers.hpp:178-188)
//   DynamicSubStep (PositionVerlet.hpp:71-77)        block.dynamic_update(dt)
//   SymplecticStepper<PositionVerlet>::step          erod::PositionVerlet::step(blocks, t, dt)
//     (SymplecticSteppers.hpp:133-217)
//
// Errors: every failed C call throws erod::Error with erod_last_error().
#pragma once

#include <cmath>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/erod.h"

namespace erod {

  struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& what) : std::runtime_error(what), code(c) {}
  };

  inline void check(int rc) {
    if (rc != EROD_OK) throw Error(rc, std::string("erod: ") + erod_last_error());
  }

  // tag names of the reference, mapped to field ids
  namespace tags {
    inline constexpr int Position = EROD_POSITION, Velocity = EROD_VELOCITY, Director = EROD_DIRECTOR,
                         AngularVelocity = EROD_ANGULAR_VELOCITY, Mass = EROD_MASS, InvMass = EROD_INV_MASS,
                         ExternalLoads = EROD_EXTERNAL_LOADS, ExternalTorques = EROD_EXTERNAL_TORQUES,
                         InternalLoads = EROD_INTERNAL_LOADS, InternalTorques = EROD_INTERNAL_TORQUES,
                         ReferenceElementLength = EROD_REFERENCE_ELEMENT_LENGTH,
                         ReferenceVoronoiLength = EROD_REFERENCE_VORONOI_LENGTH,
                         ReferenceCurvature = EROD_REFERENCE_CURVATURE,
                         ReferenceShearStretchStrain = EROD_REFERENCE_SHEAR_STRETCH_STRAIN,
                         ShearStretchRigidityMatrix = EROD_SHEAR_STRETCH_RIGIDITY,
                         BendingTwistRigidityMatrix = EROD_BENDING_TWIST_RIGIDITY,
                         MassSecondMomentOfInertia = EROD_MASS_SECOND_MOMENT_OF_INERTIA,
                         InvMassSecondMomentOfInertia = EROD_INV_MASS_SECOND_MOMENT_OF_INERTIA,
                         ForceDampingRate = EROD_FORCE_DAMPING_RATE, TorqueDampingRate = EROD_TORQUE_DAMPING_RATE,
                         ElementVolume = EROD_ELEMENT_VOLUME, ElementDimension = EROD_ELEMENT_DIMENSION;
  }  // namespace tags

  class DeviceBlock {
   public:
    explicit DeviceBlock(std::vector<std::uint64_t> n_elems, bool damped = true) : n_(std::move(n_elems)) {
      check(erod_batch_create(n_.data(), n_.size(), damped ? 1 : 0, &b_));
    }
    // a batch continuing from a checkpoint written by save()
    static DeviceBlock load(const std::string& path) {
      erod_batch* b = nullptr;
      check(erod_load(path.c_str(), &b));
      std::uint64_t slots = 0;
      check(erod_batch_slots(b, &slots));
      std::vector<std::uint64_t> n;
      std::uint64_t s0 = 0, s1 = 0;
      for (std::uint64_t r = 0; erod_rod_start(b, r, &s0) == EROD_OK; ++r) {
        const std::uint64_t next = erod_rod_start(b, r + 1, &s1) == EROD_OK ? s1 : slots + 2;
        n.push_back(next - s0 - 2);
      }
      return DeviceBlock(b, std::move(n));
    }
    DeviceBlock(const DeviceBlock&) = delete;
    DeviceBlock& operator=(const DeviceBlock&) = delete;
    DeviceBlock(DeviceBlock&& o) noexcept : b_(o.b_), n_(std::move(o.n_)) { o.b_ = nullptr; }
    ~DeviceBlock() { erod_batch_destroy(b_); }

    std::size_t n_units() const { return n_.size(); }
    std::uint64_t n_elements(std::size_t rod) const { return n_.at(rod); }
    erod_batch* handle() const { return b_; }

    // number of values per line of a field for one rod (n + 1 / n / n - 1)
    std::uint64_t rod_count(int field, std::size_t rod) const {
      int nc = 0, st = 0;
      check(erod_field_shape(field, &nc, &st));
      const std::uint64_t n = n_.at(rod);
      return st == EROD_NODE ? n + 1 : (st == EROD_ELEMENT ? n : n - 1);
    }
    int ncomp(int field) const {
      int nc = 0, st = 0;
      check(erod_field_shape(field, &nc, &st));
      return nc;
    }

    // blocks::get<tag>(slice) read / write, component-major host arrays
    std::vector<double> get(int field, std::size_t rod) const {
      std::vector<double> out(std::size_t(ncomp(field)) * rod_count(field, rod));
      check(erod_download_rod(b_, field, rod, out.data()));
      return out;
    }
    void set(int field, std::size_t rod, const std::vector<double>& v) {
      if (v.size() != std::size_t(ncomp(field)) * rod_count(field, rod)) throw Error(EROD_ERR_INVALID, "erod: size mismatch");
      check(erod_upload_rod(b_, field, rod, v.data()));
    }

    // the three per-block calls of the stepper
    void kinematic_update(double dt) { check(erod_kinematic_update(b_, dt)); }
    void compute_internal_dynamics(double t) { check(erod_compute_internal_dynamics(b_, t)); }
    void dynamic_update(double dt) { check(erod_dynamic_update(b_, dt)); }

    // module hooks as device descriptors
    // node `index` (and element `index`, below n) held at x and Q
    void clamp(std::size_t rod, const double x[3], const double Q[9], std::size_t index = 0) {
      check(erod_add_clamp(b_, rod, index, x, Q));
    }
    void end_force(std::size_t rod, const double F[3], erod_ramp ramp = EROD_RAMP_NONE, double start = 0.0,
                   double interval = 1.0) {
      check(erod_add_end_force(b_, rod, F, ramp, start, interval));
    }
    void gravity(const double g[3]) { check(erod_set_gravity(b_, g)); }
    // per-node contact with the plane n . x = z0 (n = (0, 0, 1) unless given)
    void plane_contact(double k, double c, double mu, double ct, double z0 = 0.0, double offset = 0.0,
                       const double* normal = nullptr) {
      check(erod_set_plane_contact(b_, k, c, mu, ct, z0, offset));
      if (normal) check(erod_set_plane_normal(b_, normal));
    }
    void post_kinematic(double t) { check(erod_apply_post_kinematic(b_, t)); }
    void pre_dynamic(double t) { check(erod_apply_pre_dynamic(b_, t)); }
    void post_dynamic(double t) { check(erod_apply_post_dynamic(b_, t)); }

    // whole steps on the fused persistent kernels
    void set_dt(double dt) { check(erod_set_rod_dt(b_, std::vector<double>(n_.size(), dt).data())); }
    void step_fused(std::uint64_t n, erod_precision p = EROD_FP64) { check(erod_step_fused(b_, p, n)); }
    // tile shape chosen by measurement (and a second group of smaller tiles when faster)
    double autotune(erod_precision p = EROD_FP64, int trial_steps = 100) {
      double ms = 0.0;
      check(erod_autotune_fused(b_, p, trial_steps, &ms));
      return ms;
    }
    void set_split(int lanes_per_tile, int elements_per_thread) {
      check(erod_set_fused_split(b_, lanes_per_tile, elements_per_thread));
    }
    void synchronize() { check(erod_synchronize(b_)); }

    // slots and budgets (rods of different lengths and horizons in one batch)
    void set_active_length(std::size_t rod, std::uint64_t n) { check(erod_set_rod_active_length(b_, rod, n)); }
    void set_step_budget(std::size_t rod, std::uint64_t steps) { check(erod_set_rod_step_budget(b_, rod, steps)); }

    // checkpoints (Block/Serialize.hpp's archive of the block, device side)
    void save(const std::string& path) { check(erod_save(b_, path.c_str())); }

   private:
    DeviceBlock(erod_batch* b, std::vector<std::uint64_t> n) : b_(b), n_(std::move(n)) {}
    erod_batch* b_ = nullptr;
    std::vector<std::uint64_t> n_;
  };

  // Continuous batching of requests (erod_server_*): rods of a library batch
  // are admitted into a pool of slots in size classes and advanced in ticks.
  class Server {
   public:
    Server(DeviceBlock& library, const std::vector<std::pair<std::uint64_t, std::uint64_t>>& classes,
           erod_precision p = EROD_FP32_BULK, int steps_per_tick = 100, double hbm_budget_bytes = 200e9) {
      std::vector<std::uint64_t> cap, cnt;
      for (auto& [c, k] : classes) {
        cap.push_back(c);
        cnt.push_back(k);
      }
      check(erod_server_create(library.handle(), cap.data(), cnt.data(), int(classes.size()), p, steps_per_tick,
                               hbm_budget_bytes, &s_));
    }
    // a pool of the given slot capacities that grows on demand while it fits
    // the HBM budget (erod_server_create_budgeted)
    Server(DeviceBlock& library, const std::vector<std::uint64_t>& capacities, erod_precision p, int steps_per_tick,
           double hbm_budget_bytes) {
      check(erod_server_create_budgeted(library.handle(), capacities.data(), int(capacities.size()), p, steps_per_tick,
                                        hbm_budget_bytes, &s_));
    }
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;
    ~Server() { erod_server_destroy(s_); }

    void submit(std::uint64_t id, const std::vector<std::uint64_t>& library_rods, std::uint64_t steps) {
      check(erod_server_submit(s_, id, library_rods.data(), library_rods.size(), steps));
    }
    // one tick; the ids of requests that finished
    std::vector<std::uint64_t> tick() {
      std::vector<std::uint64_t> done(4096);
      std::uint64_t n = 0;
      check(erod_server_tick(s_, done.data(), done.size(), &n));
      done.resize(n);
      return done;
    }
    void read(std::uint64_t id, std::uint64_t rod, int field, double* host) {
      check(erod_server_read(s_, id, rod, field, host));
    }
    void release(std::uint64_t id) { check(erod_server_release(s_, id)); }
    // (ticks launched, element-steps advanced) since creation
    std::pair<std::uint64_t, std::uint64_t> stats() const {
      std::uint64_t t = 0, e = 0;
      check(erod_server_stats(s_, &t, &e));
      return {t, e};
    }

   private:
    erod_server* s_ = nullptr;
  };

  // Straight-rod initializer (the reference's StraightCosseratRodInitializer:
  // Initializers/detail/Geometry.hpp:110-214, Orthogonal.cpp:33-116,
  // CosseratRodSpanwiseGeometry.hpp:525-545, CosseratRodCrossSectionInterface.hpp:266-307,
  // CircleCrossSectionOperations.hpp:83-132, LinearHyperElasticityFacade.hpp:260-354,
  // WithRodKinematics.hpp:330-436), in the reference's operation order.
  // Returns field id -> component-major per-rod array.
  struct StraightRod {
    std::uint64_t n;
    double radius, length, origin[3], direction[3], normal[3], density, youngs, shear, force_damping, torque_damping;
  };

  inline std::map<int, std::vector<double>> straight_rod_state(const StraightRod& p) {
    auto normalize = [](const double a[3], double o[3]) {
      const double inv = 1.0 / std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
      for (int i = 0; i < 3; ++i) o[i] = a[i] * inv;
    };
    auto cross = [](const double a[3], const double b[3], double o[3]) {
      o[0] = a[1] * b[2] - a[2] * b[1];
      o[1] = a[2] * b[0] - a[0] * b[2];
      o[2] = a[0] * b[1] - a[1] * b[0];
    };
    const std::uint64_t n = p.n;
    double d[3], nr[3], v1[3], v2[3], v3[3], t3[3], d2[3], d3[3];
    normalize(p.direction, d);
    normalize(p.normal, nr);
    normalize(nr, v1);
    cross(d, nr, v2);
    cross(v1, v2, v3);
    cross(v3, v1, t3);
    normalize(t3, d2);
    normalize(v3, d3);
    const double rows[9] = {v1[0], v1[1], v1[2], d2[0], d2[1], d2[2], d3[0], d3[1], d3[2]};
    std::map<int, std::vector<double>> f;
    auto& x = f[tags::Position] = std::vector<double>(3 * (n + 1));
    double step[3];
    for (int c = 0; c < 3; ++c) step[c] = ((p.origin[c] + p.length * d[c]) - p.origin[c]) / double(n);
    for (std::uint64_t i = 0; i <= n; ++i)
      for (int c = 0; c < 3; ++c) x[c * (n + 1) + i] = p.origin[c] + double(i) * step[c];
    auto& Q = f[tags::Director] = std::vector<double>(9 * n);
    for (std::uint64_t k = 0; k < n; ++k)
      for (int m = 0; m < 9; ++m) Q[m * n + k] = rows[m];
    auto& l0 = f[tags::ReferenceElementLength] = std::vector<double>(n);
    std::vector<double> l(n);
    for (std::uint64_t k = 0; k < n; ++k) {
      double dd[3];
      for (int c = 0; c < 3; ++c) dd[c] = x[c * (n + 1) + k + 1] - x[c * (n + 1) + k];
      l[k] = std::sqrt(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]);
      l0[k] = l[k] / 1.0;
    }
    auto& D0 = f[tags::ReferenceVoronoiLength] = std::vector<double>(n - 1);
    for (std::uint64_t k = 0; k + 1 < n; ++k) D0[k] = 0.5 * (l0[k] + l0[k + 1]);
    auto& vol = f[tags::ElementVolume] = std::vector<double>(n);
    auto& r = f[tags::ElementDimension] = std::vector<double>(n);
    for (std::uint64_t k = 0; k < n; ++k) {
      vol[k] = (M_PI * (p.radius * p.radius)) * l0[k];
      r[k] = std::sqrt(M_1_PI * vol[k] / l[k]);
    }
    auto moi = [&](std::uint64_t k, double I[3]) {
      const double A = M_PI * (r[k] * r[k]);
      const double i1 = (0.25 * M_1_PI) * (A * A);
      I[0] = i1;
      I[1] = i1;
      I[2] = 2.0 * i1;
    };
    const double EEG[3] = {p.youngs, p.youngs, p.shear};
    auto& B = f[tags::BendingTwistRigidityMatrix] = std::vector<double>(3 * (n - 1));
    for (std::uint64_t k = 0; k + 1 < n; ++k) {
      double I0[3], I1[3];
      moi(k, I0);
      moi(k + 1, I1);
      for (int c = 0; c < 3; ++c)
        B[c * (n - 1) + k] = ((l0[k] * EEG[c]) * I0[c] + (l0[k + 1] * EEG[c]) * I1[c]) / (l0[k] + l0[k + 1]);
    }
    const double alpha = 13.5 / 14.0;
    auto& S = f[tags::ShearStretchRigidityMatrix] = std::vector<double>(3 * n);
    auto& J = f[tags::MassSecondMomentOfInertia] = std::vector<double>(3 * n);
    auto& iJ = f[tags::InvMassSecondMomentOfInertia] = std::vector<double>(3 * n);
    for (std::uint64_t k = 0; k < n; ++k) {
      const double A = M_PI * (r[k] * r[k]);
      S[k] = p.shear * (A * alpha);
      S[n + k] = p.shear * (A * alpha);
      S[2 * n + k] = p.youngs * A;
      double I[3];
      moi(k, I);
      for (int c = 0; c < 3; ++c) {
        J[c * n + k] = (I[c] * p.density) * l0[k];
        iJ[c * n + k] = 1.0 / ((I[c] * p.density) * l0[k]);
      }
    }
    auto& m = f[tags::Mass] = std::vector<double>(n + 1, 0.0);
    for (std::uint64_t k = 0; k < n; ++k) {
      const double me = p.density * vol[k];
      m[k] += 0.5 * me;
      m[k + 1] += 0.5 * me;
    }
    auto& im = f[tags::InvMass] = std::vector<double>(n + 1);
    for (std::uint64_t i = 0; i <= n; ++i) im[i] = 1.0 / m[i];
    f[tags::ForceDampingRate] = std::vector<double>(n + 1, p.force_damping);
    f[tags::TorqueDampingRate] = std::vector<double>(n, p.torque_damping);
    f[EROD_ELEMENT_LENGTH] = l;
    f[EROD_ELEMENT_DILATATION] = std::vector<double>(n, 1.0);
    return f;
  }

  inline void load_rod(DeviceBlock& blk, std::size_t rod, const std::map<int, std::vector<double>>& fields) {
    for (auto const& [field, v] : fields) blk.set(field, rod, v);
  }

  // SymplecticStepper<PositionVerlet>::step over a collection of blocks, in the
  // reference's substep order {Kinematic, TimeUpdate, Dynamic, TimeUpdate,
  // Kinematic} (PositionVerlet.hpp:84-86, SymplecticSteppers.hpp:152-216).
  struct PositionVerlet {
    template <typename Blocks>
    static void step(Blocks& blocks, double& time, double dt) {
      for (auto& b : blocks) b.kinematic_update(dt);
      time += 0.5 * dt;
      for (auto& b : blocks) b.post_kinematic(time);
      for (auto& b : blocks) b.compute_internal_dynamics(time);
      for (auto& b : blocks) b.pre_dynamic(time);
      for (auto& b : blocks) b.dynamic_update(dt);
      for (auto& b : blocks) b.post_dynamic(time);
      time += 0.5 * dt;
      for (auto& b : blocks) b.kinematic_update(dt);
    }
  };

}  // namespace erod
