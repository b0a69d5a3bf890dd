/*
 * erod_elastica.hpp — binding of a device-resident rod block to the surface
 * through which Elastica++'s symplectic steppers drive a system (C++17,
 * header only, over the C ABI of erod.h).
 *
 * Per step the reference's stepper does, for each block `sys`
 * (elastica/Simulator/Interface/TimeSteppers/SymplecticSteppers.hpp:152-216):
 *
 *   kinematic substep   sys.kinematic_states() += c * sys.kinematic_rates(t)
 *                       (PositionVerlet.hpp:60-69: c = (1/2) dt;
 *                        VelocityVerlet.hpp:74: c = dt)
 *   internal dynamics   using ::elastica::external_loads; using ::elastica::external_torques;
 *                       external_loads(sys) = 0.0; external_torques(sys) = 0.0;
 *                       sys.compute_internal_dynamics(t)          (:178-188)
 *   dynamic substep     sys.dynamic_states() += c * sys.dynamic_rates(t)
 *                       (PositionVerlet.hpp:71-77: c = dt; dynamic_rates runs
 *                        update_acceleration, SymplecticStepperAdapter.hpp:357-366)
 *
 * DeviceCosseratRodBlock answers each of these with a proxy: the rates are
 * handles, a coefficient times a rate records the coefficient, and `+=` of
 * the result onto the states issues one device call with it
 * (erod_kinematic_update_by, erod_dynamic_update), so every symplectic
 * algorithm's coefficients pass through unchanged.  external_loads and
 * external_torques are found by argument-dependent lookup next to the
 * reference's own overloads (elastica/Systems/CosseratRods/Access.hpp:229-247)
 * and return a proxy whose `= value` fills the device field
 * (erod_fill_field).  Times and time deltas may be plain numbers or the
 * reference's SimulationTime / SimulationTimeDelta (anything with value()).
 *
 * Errors: a failed call throws std::runtime_error with erod_last_error();
 * inside the stepper's noexcept lambdas that ends the program, as a failed
 * assertion of the reference does.
 */
#ifndef EROD_ELASTICA_HPP
#define EROD_ELASTICA_HPP

#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>

#include "erod.h"

namespace erod_elastica {

  inline void check(int rc) {
    if (rc != EROD_OK) throw std::runtime_error(std::string("erod: ") + erod_last_error());
  }

  // the value of a time or time delta: t.value() if it has one, else t
  template <typename T, typename = void>
  struct has_value : std::false_type {};
  template <typename T>
  struct has_value<T, std::void_t<decltype(std::declval<const T&>().value())>> : std::true_type {};
  template <typename T>
  inline double time_value(const T& t) {
    if constexpr (has_value<T>::value) return static_cast<double>(t.value());
    else return static_cast<double>(t);
  }

  class DeviceCosseratRodBlock;

  // rates of the kinematic states (v and w) and of the dynamic states (a and
  // alpha, computed by the update that consumes them)
  struct KinematicRates {
    DeviceCosseratRodBlock* block;
  };
  struct DynamicRates {
    DeviceCosseratRodBlock* block;
  };
  // coefficient * rates
  template <typename Rates>
  struct Scaled {
    Rates rates;
    double c;
  };

  template <typename T>
  using if_scalar_t = std::enable_if_t<std::is_arithmetic_v<T> || has_value<T>::value, int>;

  template <typename T, if_scalar_t<T> = 0>
  inline Scaled<KinematicRates> operator*(const T& c, KinematicRates r) { return {r, time_value(c)}; }
  template <typename T, if_scalar_t<T> = 0>
  inline Scaled<KinematicRates> operator*(KinematicRates r, const T& c) { return {r, time_value(c)}; }
  template <typename T, if_scalar_t<T> = 0>
  inline Scaled<DynamicRates> operator*(const T& c, DynamicRates r) { return {r, time_value(c)}; }
  template <typename T, if_scalar_t<T> = 0>
  inline Scaled<DynamicRates> operator*(DynamicRates r, const T& c) { return {r, time_value(c)}; }
  template <typename R, typename T, if_scalar_t<T> = 0>
  inline Scaled<R> operator*(const T& c, Scaled<R> s) { return {s.rates, time_value(c) * s.c}; }
  template <typename R, typename T, if_scalar_t<T> = 0>
  inline Scaled<R> operator*(Scaled<R> s, const T& c) { return {s.rates, s.c * time_value(c)}; }

  // sys.kinematic_states() and sys.dynamic_states()
  struct KinematicStates {
    DeviceCosseratRodBlock* block;
    KinematicStates& operator+=(Scaled<KinematicRates> s);
  };
  struct DynamicStates {
    DeviceCosseratRodBlock* block;
    DynamicStates& operator+=(Scaled<DynamicRates> s);
  };

  // external_loads(sys) / external_torques(sys): `= value` fills the field
  struct FieldRef {
    erod_batch* batch;
    int field;
    FieldRef& operator=(double v) {
      check(erod_fill_field(batch, field, v));
      return *this;
    }
  };

  // A device-resident block of Cosserat rods (a non-owning view of an
  // erod_batch) with the stepper-facing surface of the reference's
  // SymplecticPolicy (SymplecticStepperAdapter.hpp:324-366).
  class DeviceCosseratRodBlock {
   public:
    explicit DeviceCosseratRodBlock(erod_batch* b) noexcept : b_(b), ks_{this}, ds_{this} {}
    DeviceCosseratRodBlock(const DeviceCosseratRodBlock&) = delete;
    DeviceCosseratRodBlock& operator=(const DeviceCosseratRodBlock&) = delete;

    KinematicStates& kinematic_states() & noexcept { return ks_; }
    template <typename TimeT>
    KinematicRates kinematic_rates(const TimeT& /*time*/) noexcept {
      return {this};
    }
    DynamicStates& dynamic_states() & noexcept { return ds_; }
    template <typename TimeT>
    DynamicRates dynamic_rates(const TimeT& /*time*/) noexcept {
      return {this};
    }
    // compute_internal_dynamics(t) (ElasticityInterface.hpp:360-365); the C
    // call also zeroes the external loads, which the stepper has just done
    template <typename TimeT>
    void compute_internal_dynamics(const TimeT& time) {
      check(erod_compute_internal_dynamics(b_, time_value(time)));
    }

    // the module hooks of the block as device descriptors, for the
    // simulator's apply_post_kinematic_updates / apply_pre_dynamic_updates /
    // apply_post_dynamic_updates (Simulator.hpp:753-807)
    template <typename TimeT>
    void apply_post_kinematic(const TimeT& time) {
      check(erod_apply_post_kinematic(b_, time_value(time)));
    }
    template <typename TimeT>
    void apply_pre_dynamic(const TimeT& time) {
      check(erod_apply_pre_dynamic(b_, time_value(time)));
    }
    template <typename TimeT>
    void apply_post_dynamic(const TimeT& time) {
      check(erod_apply_post_dynamic(b_, time_value(time)));
    }

    erod_batch* handle() const noexcept { return b_; }

   private:
    erod_batch* b_;
    KinematicStates ks_;
    DynamicStates ds_;
  };

  inline KinematicStates& KinematicStates::operator+=(Scaled<KinematicRates> s) {
    check(erod_kinematic_update_by(s.rates.block->handle(), s.c));
    return *this;
  }
  inline DynamicStates& DynamicStates::operator+=(Scaled<DynamicRates> s) {
    check(erod_dynamic_update(s.rates.block->handle(), s.c));
    return *this;
  }

  // found by argument-dependent lookup from the stepper's
  // `using ::elastica::external_loads; external_loads(system) = 0.0;`
  inline FieldRef external_loads(DeviceCosseratRodBlock& sys) { return {sys.handle(), EROD_EXTERNAL_LOADS}; }
  inline FieldRef external_torques(DeviceCosseratRodBlock& sys) { return {sys.handle(), EROD_EXTERNAL_TORQUES}; }

}  // namespace erod_elastica

#endif  // EROD_ELASTICA_HPP
