// ShallowWaterParams (reference: src/model/shallow_water_params.hpp:7-18),
// same fields, defaults and order so aggregate initialisation
// {f0, grav, drag} keeps its meaning.
#pragma once

#include "../common.hpp"

namespace omega {

struct ShallowWaterParams {
  Real m_f0;
  Real m_grav = 9.81;
  Real m_drag_coeff = 0;
  Real m_visc_del2 = 0;
  Real m_visc_del4 = 0;
  Real m_eddy_diff2 = 0;
  Real m_eddy_diff4 = 0;
  Int m_ntracers = 0;
  bool m_disable_h_tendency = false;
  bool m_disable_vn_tendency = false;
};

inline omega_params to_c(const ShallowWaterParams &p) {
  omega_params c{};
  c.f0 = p.m_f0;
  c.grav = p.m_grav;
  c.drag_coeff = p.m_drag_coeff;
  c.visc_del2 = p.m_visc_del2;
  c.visc_del4 = p.m_visc_del4;
  c.eddy_diff2 = p.m_eddy_diff2;
  c.eddy_diff4 = p.m_eddy_diff4;
  c.ntracers = p.m_ntracers;
  c.disable_h_tendency = p.m_disable_h_tendency;
  c.disable_vn_tendency = p.m_disable_vn_tendency;
  return c;
}

} // namespace omega
