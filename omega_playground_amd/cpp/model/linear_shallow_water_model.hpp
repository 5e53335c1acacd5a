// LinearShallowWaterModel (reference: src/shallow_water_hierarchical.cpp:676-773,
// which the reference does not compile): the shallow-water equations
// linearised about a resting layer of thickness h0,
//   h_t  = -h0 div(vn)                                  (:680-707)
//   vn_t = f_edge * (TRiSK tangential velocity) - g grad(h)  (:709-741)
// with f_edge = f0 and no tracers.  Steppers, additional_tendency, increment
// mode and the integrals (energy with h0*K, :743-773) work as for
// ShallowWaterModel; adjoint_run and Ensemble reject it.
#pragma once

#include "shallow_water_model.hpp"

namespace omega {

struct LinearShallowWaterParams : ShallowWaterParams {
  Real m_h0 = 1;
};

struct LinearShallowWaterModel : ShallowWaterModel {
  Real m_h0;

  LinearShallowWaterModel(MPASMesh *mesh, const LinearShallowWaterParams &params)
      : ShallowWaterModel(mesh, params, &params.m_h0), m_h0(params.m_h0) {}
};

} // namespace omega
