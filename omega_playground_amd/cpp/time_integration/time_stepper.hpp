// Time steppers (reference: src/time_integration/time_stepper.hpp:8-38,
// rk4.cpp:8-104, lsrk.cpp:5-89).  Without an additional-tendency override a
// stage is one diagnostics launch and one fused tendency + update launch;
// enable_graph() captures the step in a HIP graph.
#pragma once

#include "../model/shallow_water_model.hpp"

namespace omega {

struct TimeStepper {
  ShallowWaterModel *m_shallow_water;

  explicit TimeStepper(ShallowWaterModel &shallow_water) : m_shallow_water(&shallow_water) {}
  TimeStepper(const TimeStepper &) = delete;
  TimeStepper &operator=(const TimeStepper &) = delete;
  virtual ~TimeStepper() {
    if (m_handle) omega_stepper_destroy(m_handle);
  }

  virtual void do_step(Real t, Real dt, const ShallowWaterState &state) const {
    check(omega_stepper_do_step(m_handle, t, dt, state.handle(), m_shallow_water->stream()));
    m_shallow_water->drop_noop_hook();
  }

  // nsteps calls of do_step with t = t0 + (i + t_offset_steps) * dt, without
  // a host round trip per step.
  void run(Real t0, Real dt, Int nsteps, const ShallowWaterState &state,
           Int t_offset_steps = 0) const {
    if (nsteps <= 0) return;
    do_step(t0 + t_offset_steps * dt, dt, state);
    check(omega_stepper_run(m_handle, t0, dt, nsteps - 1, t_offset_steps + 1, state.handle(),
                            m_shallow_water->stream()));
  }

  void enable_graph(bool on = true) const { check(omega_stepper_set_graph(m_handle, on)); }

  omega_stepper_t handle() const { return m_handle; }

protected:
  omega_stepper_t m_handle = nullptr;

  void create(Int kind, Int nstages, std::vector<Real> &a, std::vector<Real> &b,
              std::vector<Real> &c) {
    check(omega_stepper_create(m_shallow_water->handle(), kind, nstages, &m_handle));
    int n = 0;
    check(omega_stepper_coeffs(m_handle, &n, nullptr, nullptr, nullptr));
    a.assign(n, 0.0);
    b.assign(n, 0.0);
    c.assign(n, 0.0);
    check(omega_stepper_coeffs(m_handle, &n, a.data(), b.data(), c.data()));
  }
};

struct LSRKStepper : TimeStepper {
  Int m_nstages;
  std::vector<Real> m_rka;
  std::vector<Real> m_rkb;
  std::vector<Real> m_rkc;

  LSRKStepper(ShallowWaterModel &shallow_water, Int nstages = 5)
      : TimeStepper(shallow_water), m_nstages(nstages) {
    create(OMEGA_STEPPER_LSRK, nstages, m_rka, m_rkb, m_rkc);
  }
};

struct RK4Stepper : TimeStepper {
  static constexpr Int nstages = 4;
  std::vector<Real> m_rka;
  std::vector<Real> m_rkb;
  std::vector<Real> m_rkc;

  explicit RK4Stepper(ShallowWaterModel &shallow_water) : TimeStepper(shallow_water) {
    create(OMEGA_STEPPER_RK4, nstages, m_rka, m_rkb, m_rkc);
    m_rka.resize(nstages - 1);
    m_rkc.resize(nstages - 1);
  }
};

} // namespace omega
