// Discrete adjoint of the C++ drop-in API (no reference counterpart: the
// reference has no adjoint).  Cotangents are ShallowWaterState objects;
// gradients with respect to the mixing coefficients are returned in the
// order (visc_del2, visc_del4, eddy_diff2, eddy_diff4).
#pragma once

#include "../time_integration/time_stepper.hpp"

namespace omega {

using CoefficientGradient = std::array<Real, 4>;

// state_cot += J(state)^T tend_cot
inline CoefficientGradient tendency_vjp(const ShallowWaterModel &model,
                                        const ShallowWaterState &state,
                                        const ShallowWaterState &tend_cot,
                                        const ShallowWaterState &state_cot) {
  model.settle_hook(state);
  CoefficientGradient g{};
  check(omega_model_tendency_vjp(model.handle(), state.handle(), tend_cot.handle(),
                                 state_cot.handle(), g.data(), model.stream()));
  return g;
}

// cot <- (d y_{n+1} / d y_n)^T cot for one step from state
inline CoefficientGradient step_vjp(const TimeStepper &stepper, Real t, Real dt,
                                    const ShallowWaterState &state,
                                    const ShallowWaterState &cot) {
  stepper.m_shallow_water->settle_hook(state);
  CoefficientGradient g{};
  check(omega_stepper_step_vjp(stepper.handle(), t, dt, state.handle(), cot.handle(), g.data(),
                               stepper.m_shallow_water->stream()));
  return g;
}

// Binomial checkpointing over nsteps steps (n_snapshots stored states, the
// initial one included; byte_cap > 0 bounds their memory).
class Tape {
public:
  Tape(const TimeStepper &stepper, Int nsteps, Int n_snapshots, int64_t byte_cap = 0)
      : m_stepper(&stepper) {
    check(omega_tape_create(stepper.handle(), nsteps, n_snapshots, byte_cap, &m_handle));
  }
  Tape(const Tape &) = delete;
  Tape &operator=(const Tape &) = delete;
  ~Tape() {
    if (m_handle) omega_tape_destroy(m_handle);
  }
  // state: y_0 on entry, y_nsteps on exit
  void forward(Real t0, Real dt, const ShallowWaterState &state) const {
    m_stepper->m_shallow_water->settle_hook(state);
    check(omega_tape_forward(m_handle, t0, dt, state.handle(), stream()));
  }
  // cot: dJ/dy_nsteps on entry, dJ/dy_0 on exit
  CoefficientGradient reverse(const ShallowWaterState &cot) const {
    CoefficientGradient g{};
    check(omega_tape_reverse(m_handle, cot.handle(), g.data(), stream()));
    return g;
  }
  int64_t forward_steps() const {
    int64_t n = 0;
    check(omega_tape_info(m_handle, nullptr, &n, nullptr));
    return n;
  }

private:
  const TimeStepper *m_stepper;
  omega_tape_t m_handle = nullptr;
  void *stream() const { return m_stepper->m_shallow_water->stream(); }
};

// J = sum_f w_f/2 sum_i area_i sum_k (y - obs)^2, cot = dJ/dy
inline Real misfit(const ShallowWaterModel &model, const ShallowWaterState &state,
                   const ShallowWaterState &obs, const ShallowWaterState &cot,
                   const std::array<Real, 3> &weights = {1.0, 0.0, 1.0}) {
  Real J = 0;
  check(omega_misfit(model.handle(), state.handle(), obs.handle(), weights.data(), cot.handle(),
                     &J, model.stream()));
  return J;
}

// Misfit after nsteps steps from state (left unchanged) and its gradient:
// grad = dJ/dy_0 (overwritten); binomial checkpointing with n_snapshots
// stored states and an optional byte cap.
struct AdjointResult {
  Real J = 0;
  CoefficientGradient coefficients{};
};

inline AdjointResult adjoint_run(const TimeStepper &stepper, Real t0, Real dt, Int nsteps,
                                 const ShallowWaterState &state, const ShallowWaterState &obs,
                                 const ShallowWaterState &grad,
                                 const std::array<Real, 3> &weights = {1.0, 0.0, 1.0},
                                 Int n_snapshots = 10, int64_t byte_cap = 0) {
  stepper.m_shallow_water->settle_hook(state);
  AdjointResult r;
  check(omega_adjoint_run(stepper.handle(), t0, dt, nsteps, n_snapshots, byte_cap, state.handle(),
                          obs.handle(), weights.data(), &r.J, grad.handle(),
                          r.coefficients.data(), stepper.m_shallow_water->stream()));
  return r;
}

} // namespace omega
