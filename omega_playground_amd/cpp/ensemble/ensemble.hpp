// Ensembles of independent members (no reference counterpart: the
// reference steps one model at a time).  A member is a model, its state and
// its dt; an Ensemble step advances every member by one step of the
// ensemble's stepper kind with one launch per stage phase for all members,
// and each member's state ends bit-identical to stepping it alone.  The
// Ensemble keeps references to the members' models and states: keep them
// alive.
#pragma once

#include <vector>

#include "../adjoint/adjoint.hpp"

namespace omega {

class Ensemble {
public:
  // stepper kind of RK4Stepper (nstages ignored) or LSRKStepper(nstages)
  enum class Kind { rk4, lsrk };

  explicit Ensemble(Kind kind = Kind::lsrk, Int nstages = 5) {
    check(omega_batch_create(kind == Kind::rk4 ? OMEGA_STEPPER_RK4 : OMEGA_STEPPER_LSRK, nstages,
                             &m_handle));
  }
  Ensemble(const Ensemble &) = delete;
  Ensemble &operator=(const Ensemble &) = delete;
  ~Ensemble() {
    if (m_handle) omega_batch_destroy(m_handle);
  }

  // returns the member's index
  Int add(const ShallowWaterModel &model, const ShallowWaterState &state, Real dt) {
    model.settle_hook(state);
    int idx = 0;
    check(omega_batch_add(m_handle, model.handle(), state.handle(), dt, &idx));
    return idx;
  }
  void step(void *stream = nullptr) const { check(omega_batch_step(m_handle, stream)); }
  void run(Int nsteps, void *stream = nullptr) const {
    check(omega_batch_run(m_handle, nsteps, stream));
  }
  void enable_graph(bool on = true) const { check(omega_batch_set_graph(m_handle, on)); }
  // times every admissible launch shape over reps steps, keeps the fastest,
  // restores the members' states; returns that shape's ms per step
  Real autotune(Int reps = 3, void *stream = nullptr) const {
    double ms = 0;
    check(omega_batch_autotune(m_handle, reps, &ms, stream));
    return ms;
  }

  // Ensemble-mean misfit after nsteps steps and its gradient: grads[m] =
  // dJ/d(member m's initial state); members must share one mesh, layer and
  // tracer counts, parameters and dt, and are left unchanged.
  AdjointResult adjoint_run(Real t0, Int nsteps, const std::vector<const ShallowWaterState *> &obs,
                            const std::vector<const ShallowWaterState *> &grads,
                            const std::array<Real, 3> &weights = {1.0, 0.0, 1.0},
                            Int n_snapshots = 10, int64_t byte_cap = 0,
                            void *stream = nullptr) const {
    std::vector<omega_state_t> o, g;
    for (auto *s : obs) o.push_back(s->handle());
    for (auto *s : grads) g.push_back(s->handle());
    AdjointResult r;
    check(omega_batch_adjoint_run(m_handle, t0, nsteps, n_snapshots, byte_cap, o.data(),
                                  weights.data(), &r.J, g.data(), r.coefficients.data(), stream));
    return r;
  }

  omega_batch_t handle() const { return m_handle; }

private:
  omega_batch_t m_handle = nullptr;
};

} // namespace omega
