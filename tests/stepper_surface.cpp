// Test harness: a restatement of the call sites through which the reference's
// symplectic stepper drives a system, applied to
// erod_elastica::DeviceCosseratRodBlock (include/erod_elastica.hpp).  Built by
// tests/test_stepper_surface.py against a recording stub of the C ABI (CPU)
// and against liberod.so (GPU).
//
// Restated from the reference:
//   PositionVerlet substeps        elastica/Time/SymplecticSteppers/Algorithms/PositionVerlet.hpp:49-88
//   VelocityVerlet substeps        elastica/Time/SymplecticSteppers/Algorithms/VelocityVerlet.hpp:50-90
//   the step body                  elastica/Simulator/Interface/TimeSteppers/SymplecticSteppers.hpp:152-216
//   time and time delta            elastica/Time/SimulationTime.hpp (value(), delta arithmetic)
//   accessors of loads             elastica/Systems/CosseratRods/Access.hpp:229-247 (templates
//                                  over the rod plugin, brought in by using-declarations)
#include "../include/erod_elastica.hpp"

#include <vector>

namespace elastica {

  class TimeDelta {
   public:
    explicit TimeDelta(double v) : v_(v) {}
    double value() const { return v_; }

   private:
    double v_;
  };
  inline TimeDelta operator*(double a, TimeDelta d) { return TimeDelta(a * d.value()); }

  class Time {
   public:
    explicit Time(double v) : v_(v) {}
    double value() const { return v_; }
    Time& operator+=(TimeDelta d) {
      v_ += d.value();
      return *this;
    }

   private:
    double v_;
  };

  template <typename Plugin>
  struct CosseratRodPlugin {};
  template <typename P>
  double& external_loads(CosseratRodPlugin<P>&);
  template <typename P>
  double& external_torques(CosseratRodPlugin<P>&);

  namespace timestepper {
    namespace symplectic {
      struct PositionVerlet {
        struct TimeUpdateSubStep {
          static constexpr double temporal_coefficient() { return 0.5; }
        };
        struct KinematicSubStep {
          template <typename System>
          void operator()(System& sys, Time const time_v, TimeDelta const dt) const noexcept {
            sys.kinematic_states() += TimeUpdateSubStep::temporal_coefficient() * dt * sys.kinematic_rates(time_v);
          }
        };
        struct DynamicSubStep {
          template <typename System>
          void operator()(System& sys, Time const time_v, TimeDelta const dt) const noexcept {
            sys.dynamic_states() += dt * sys.dynamic_rates(time_v);
          }
        };
      };
      struct VelocityVerlet {
        struct DynamicSubStep {
          template <typename System>
          void operator()(System& sys, Time const time_v, TimeDelta const dt) const noexcept {
            sys.dynamic_states() += (0.5 * dt) * sys.dynamic_rates(time_v);
          }
        };
        struct KinematicSubStep {
          template <typename System>
          void operator()(System& sys, Time const time_v, TimeDelta const dt) const noexcept {
            sys.kinematic_states() += dt * sys.kinematic_rates(time_v);
          }
        };
      };
    }  // namespace symplectic
  }  // namespace timestepper

  // one position-Verlet step over the blocks (SymplecticSteppers.hpp:152-216;
  // the hooks of apply_*_updates are the blocks' device descriptors)
  template <typename Blocks>
  void position_verlet_step(Blocks& blocks, Time& time_v, TimeDelta dt) {
    using PV = timestepper::symplectic::PositionVerlet;
    const PV::KinematicSubStep kinematic{};
    const PV::DynamicSubStep dynamic{};
    for (auto* system : blocks) kinematic(*system, time_v, dt);
    time_v += PV::TimeUpdateSubStep::temporal_coefficient() * dt;
    for (auto* system : blocks) system->apply_post_kinematic(time_v);
    for (auto* system : blocks) {
      using ::elastica::external_loads;
      using ::elastica::external_torques;
      external_loads(*system) = 0.0;
      external_torques(*system) = 0.0;
      system->compute_internal_dynamics(time_v);
    }
    for (auto* system : blocks) system->apply_pre_dynamic(time_v);
    for (auto* system : blocks) dynamic(*system, time_v, dt);
    for (auto* system : blocks) system->apply_post_dynamic(time_v);
    time_v += PV::TimeUpdateSubStep::temporal_coefficient() * dt;
    for (auto* system : blocks) kinematic(*system, time_v, dt);
  }

}  // namespace elastica

extern "C" {

// n position-Verlet steps of dt over nb batches from time t0; the final time
// goes to *t_out.  0 on success, -1 if a call failed.
int surface_position_verlet(erod_batch** batches, int nb, double t0, double dt, int n, double* t_out) {
  try {
    std::vector<erod_elastica::DeviceCosseratRodBlock*> blocks;
    for (int i = 0; i < nb; ++i) blocks.push_back(new erod_elastica::DeviceCosseratRodBlock(batches[i]));
    elastica::Time t(t0);
    for (int s = 0; s < n; ++s) elastica::position_verlet_step(blocks, t, elastica::TimeDelta(dt));
    for (auto* b : blocks) delete b;
    *t_out = t.value();
    return 0;
  } catch (...) {
    return -1;
  }
}

// the velocity-Verlet substeps once each on one batch (their coefficients)
int surface_velocity_verlet_substeps(erod_batch* batch, double t, double dt) {
  try {
    erod_elastica::DeviceCosseratRodBlock block(batch);
    using VV = elastica::timestepper::symplectic::VelocityVerlet;
    VV::DynamicSubStep{}(block, elastica::Time(t), elastica::TimeDelta(dt));
    VV::KinematicSubStep{}(block, elastica::Time(t), elastica::TimeDelta(dt));
    return 0;
  } catch (...) {
    return -1;
  }
}
}
