// Diagnostics (reference: src/model/shallow_water_diagnostics.hpp:6-13,
// shallow_water_diagnostics.cpp:5-89): deterministic reductions on the
// device, returned to the host.
#pragma once

#include "shallow_water_model.hpp"
#include "shallow_water_state.hpp"

namespace omega {

inline Real energy_integral(const ShallowWaterState &state, const ShallowWaterModel &model) {
  Real out = 0;
  check(omega_energy_integral(model.handle(), state.handle(), &out, model.stream()));
  return out;
}

inline Real mass_integral(const ShallowWaterState &state, const ShallowWaterModel &model) {
  Real out = 0;
  check(omega_mass_integral(model.handle(), state.handle(), &out, model.stream()));
  return out;
}

inline Real circulation_integral(const ShallowWaterState &state,
                                 const ShallowWaterModel &model) {
  Real out = 0;
  check(omega_circulation_integral(model.handle(), state.handle(), &out, model.stream()));
  return out;
}

} // namespace omega
