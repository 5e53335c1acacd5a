// The additions of the C++ API (discrete adjoint and ensembles), exercised
// the way a driver written against omega.hpp would use them.
//
//   api_tests adjoint            adjoint_run == Tape forward + misfit + reverse
//   api_tests ensemble           Ensemble steps == each member stepped alone
//   api_tests ensemble_adjoint   Ensemble::adjoint_run == per-member adjoint_run
//   api_tests linear             LinearShallowWaterModel: RK4 keeps mass and
//                                (to its own O(dt^5) dissipation) energy
//   api_tests fields             compute_aux + compute_{h,vn,tr}_tendency on
//                                device arrays == compute_tendency
//
// Results are compared bit for bit (a gradient of the shared coefficients
// to 1e-12, since it sums members in another order).  Exit status 0 on
// success.
#include <omega.hpp>

#include <cmath>
#include <cstring>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

using namespace omega;

namespace {

void require(bool ok, const std::string &what) {
  if (!ok) throw std::runtime_error("check failed: " + what);
}

// smooth initial state from the mesh coordinates; seed shifts the pattern
void fill(const PlanarHexagonalMesh &mesh, const ShallowWaterState &s, Real seed) {
  const Real lx = mesh.m_period_x, ly = mesh.m_period_y;
  for (Int c = 0; c < mesh.m_ncells; ++c)
    for (Int k = 0; k < mesh.m_nlayers; ++k) {
      const Real x = 2 * M_PI * mesh.m_x_cell(c) / lx, y = 2 * M_PI * mesh.m_y_cell(c) / ly;
      s.m_h_cell(c, k) = 1000 + (1 + 0.01 * k) * std::sin(x + seed) * std::cos(y);
      for (Int l = 0; l < s.ntracers(); ++l)
        s.m_tr_cell(l, c, k) = s.m_h_cell(c, k) * (1 + 0.1 * std::cos(x - y + l + seed));
    }
  for (Int e = 0; e < mesh.m_nedges; ++e)
    for (Int k = 0; k < mesh.m_nlayers; ++k) {
      const Real x = 2 * M_PI * mesh.m_x_edge(e) / lx, y = 2 * M_PI * mesh.m_y_edge(e) / ly;
      s.m_vn_edge(e, k) = 0.1 * std::cos(x + 2 * y + seed) * std::cos(mesh.m_angle_edge(e));
    }
  s.copy_to_device();
}

bool same(const ShallowWaterState &a, const ShallowWaterState &b) {
  a.copy_to_host();
  b.copy_to_host();
  auto eq = [](const auto &x, const auto &y) {
    return x.size() == y.size() && std::memcmp(x.data(), y.data(), x.size() * sizeof(Real)) == 0;
  };
  return eq(a.m_h_cell, b.m_h_cell) && eq(a.m_vn_edge, b.m_vn_edge) &&
         eq(a.m_tr_cell, b.m_tr_cell);
}

ShallowWaterParams params(Int ntr) {
  ShallowWaterParams p{1e-4, 9.81, 0};
  p.m_visc_del2 = 1e3;
  p.m_eddy_diff4 = 1e9;
  p.m_ntracers = ntr;
  return p;
}

void adjoint() {
  auto mesh = std::make_unique<PlanarHexagonalMesh>(10, 10, 10e3, 8);
  const auto p = params(2);
  ShallowWaterModel model(mesh.get(), p);
  RK4Stepper stepper(model);
  ShallowWaterState y0(mesh.get(), p), obs(mesh.get(), p), g1(mesh.get(), p), g2(mesh.get(), p);
  fill(*mesh, y0, 0.0);
  fill(*mesh, obs, 0.3);
  const Real dt = 40;
  const Int n = 8;
  const AdjointResult r = adjoint_run(stepper, 0, dt, n, y0, obs, g1, {1, 0, 1}, 3);
  // the same by hand
  ShallowWaterState y(mesh.get(), p);
  y.copy_from(y0);
  Tape tape(stepper, n, 3);
  tape.forward(0, dt, y);
  const Real J = misfit(model, y, obs, g2);
  const CoefficientGradient cg = tape.reverse(g2);
  std::cout << "adjoint J " << r.J << " " << J << " dJ/dvisc_del2 " << r.coefficients[0] << " "
            << cg[0] << std::endl;
  require(r.J == J, "adjoint_run misfit equals the hand-made one");
  require(r.coefficients == cg, "coefficient gradients equal");
  require(same(g1, g2), "state gradients equal");
}

void ensemble() {
  struct M {
    std::unique_ptr<PlanarHexagonalMesh> mesh;
    std::unique_ptr<ShallowWaterModel> model;
    std::unique_ptr<ShallowWaterState> a, b;
    Real dt;
  };
  std::vector<M> ms;
  const Int shapes[3][3] = {{8, 4, 1}, {10, 8, 2}, {12, 64, 3}}; // nx, layers, tracers
  Ensemble ens(Ensemble::Kind::lsrk, 5);
  for (int i = 0; i < 3; ++i) {
    M m;
    m.mesh = std::make_unique<PlanarHexagonalMesh>(shapes[i][0], shapes[i][0], 10e3, shapes[i][1]);
    const auto p = params(shapes[i][2]);
    m.model = std::make_unique<ShallowWaterModel>(m.mesh.get(), p);
    m.a = std::make_unique<ShallowWaterState>(m.mesh.get(), p);
    m.b = std::make_unique<ShallowWaterState>(m.mesh.get(), p);
    fill(*m.mesh, *m.a, i);
    fill(*m.mesh, *m.b, i);
    m.dt = 20 + 5 * i;
    require(ens.add(*m.model, *m.a, m.dt) == i, "member index");
    ms.push_back(std::move(m));
  }
  ens.enable_graph();
  ens.run(6);
  for (auto &m : ms) {
    LSRKStepper st(*m.model, 5);
    st.run(0, m.dt, 6, *m.b);
    require(same(*m.a, *m.b), "ensemble member equals its solo run");
  }
  std::cout << "ensemble: 3 members bit-identical to solo steps" << std::endl;
}

void ensemble_adjoint() {
  auto mesh = std::make_unique<PlanarHexagonalMesh>(10, 10, 10e3, 8);
  const auto p = params(2);
  const Int M = 3, n = 8;
  const Real dt = 40;
  std::vector<std::unique_ptr<ShallowWaterModel>> models;
  std::vector<std::unique_ptr<ShallowWaterState>> y0, obs, g, gs;
  Ensemble ens(Ensemble::Kind::rk4);
  for (Int m = 0; m < M; ++m) {
    models.push_back(std::make_unique<ShallowWaterModel>(mesh.get(), p));
    y0.push_back(std::make_unique<ShallowWaterState>(mesh.get(), p));
    obs.push_back(std::make_unique<ShallowWaterState>(mesh.get(), p));
    g.push_back(std::make_unique<ShallowWaterState>(mesh.get(), p));
    gs.push_back(std::make_unique<ShallowWaterState>(mesh.get(), p));
    fill(*mesh, *y0[m], m);
    fill(*mesh, *obs[m], m + 0.2);
    ens.add(*models[m], *y0[m], dt);
  }
  std::vector<const ShallowWaterState *> po, pg;
  for (Int m = 0; m < M; ++m) {
    po.push_back(obs[m].get());
    pg.push_back(g[m].get());
  }
  const AdjointResult r = ens.adjoint_run(0, n, po, pg, {1, 0, 1}, 4);
  Real J = 0;
  CoefficientGradient cg{};
  for (Int m = 0; m < M; ++m) {
    RK4Stepper st(*models[m]);
    const AdjointResult rm =
        adjoint_run(st, 0, dt, n, *y0[m], *obs[m], *gs[m], {1.0 / M, 0, 1.0 / M}, 4);
    J += rm.J;
    for (int i = 0; i < 4; ++i) cg[i] += rm.coefficients[i];
    require(same(*g[m], *gs[m]), "member gradient equals its own adjoint run");
  }
  std::cout << "ensemble adjoint J " << r.J << " " << J << " dJ/dvisc_del2 "
            << r.coefficients[0] << " " << cg[0] << std::endl;
  require(std::abs(r.J - J) <= 1e-12 * std::abs(J), "ensemble misfit");
  for (int i = 0; i < 4; ++i)
    require(std::abs(r.coefficients[i] - cg[i]) <= 1e-12 * std::abs(cg[i]) + 1e-300,
            "shared coefficient gradient");
}

void linear() {
  auto mesh = std::make_unique<PlanarHexagonalMesh>(16, 16, 10e3, 4);
  LinearShallowWaterParams p;
  p.m_f0 = 1e-4;
  p.m_grav = 9.81;
  p.m_h0 = 1000;
  LinearShallowWaterModel model(mesh.get(), p);
  ShallowWaterState s(mesh.get(), p), rest(mesh.get(), p);
  fill(*mesh, s, 0.5);
  for (Int c = 0; c < mesh->m_ncells; ++c)
    for (Int k = 0; k < mesh->m_nlayers; ++k) rest.m_h_cell(c, k) = p.m_h0;
  rest.copy_to_device();
  const Real e0 = energy_integral(s, model), m0 = mass_integral(s, model);
  const Real er = energy_integral(rest, model);
  RK4Stepper st(model);
  st.enable_graph();
  for (Int n = 0; n < 100; ++n) st.do_step(n * 2.0, 2.0, s);
  const Real e1 = energy_integral(s, model), m1 = mass_integral(s, model);
  std::cout << "linear energy change " << (e1 - e0) / (e0 - er) << " mass change "
            << (m1 - m0) / m0 << std::endl;
  require(std::abs(e1 - e0) <= 1e-8 * std::abs(e0 - er), "linear energy conserved");
  require(std::abs(m1 - m0) <= 1e-13 * std::abs(m0), "linear mass conserved");
}

// The reference's per-field calls (shallow_water_model.hpp:35-40) on device
// arrays after compute_aux give the fields of compute_tendency, in replace
// and increment mode.
void fields() {
  auto mesh = std::make_unique<PlanarHexagonalMesh>(12, 12, 10e3, 64);
  auto p = params(2);
  p.m_eddy_diff2 = 1e2;
  ShallowWaterModel model(mesh.get(), p);
  ShallowWaterState s(mesh.get(), p), full(mesh.get(), p), part(mesh.get(), p);
  fill(*mesh, s, 0.7);
  model.compute_tendency(full, s, 0.0);
  model.compute_aux(s);
  model.compute_h_tendency(part.h_cell_device(), s.h_cell_device(), s.vn_edge_device(),
                           AddMode::replace);
  model.compute_vn_tendency(part.vn_edge_device(), s.h_cell_device(), s.vn_edge_device(),
                            AddMode::replace);
  model.compute_tr_tendency(part.tr_cell_device(), s.tr_cell_device(), s.vn_edge_device(),
                            AddMode::replace);
  require(same(full, part), "per-field tendencies equal compute_tendency");
  // increment mode: the terms accumulate onto what the tendency holds, as
  // compute_tendency(..., AddMode::increment) does
  ShallowWaterState inc(mesh.get(), p);
  inc.copy_from(full);
  model.compute_tendency(inc, s, 0.0, AddMode::increment);
  model.compute_aux(s);
  model.compute_h_tendency(part.h_cell_device(), s.h_cell_device(), s.vn_edge_device(),
                           AddMode::increment);
  model.compute_vn_tendency(part.vn_edge_device(), s.h_cell_device(), s.vn_edge_device(),
                            AddMode::increment);
  model.compute_tr_tendency(part.tr_cell_device(), s.tr_cell_device(), s.vn_edge_device(),
                            AddMode::increment);
  require(same(inc, part), "increment-mode field tendencies equal compute_tendency's");
  // compute_vn_tendency with a thickness other than the auxiliary state's
  bool threw = false;
  try {
    model.compute_vn_tendency(part.vn_edge_device(), part.h_cell_device(), s.vn_edge_device(),
                              AddMode::replace);
  } catch (const std::runtime_error &) {
    threw = true;
  }
  require(threw, "vn tendency with another thickness array is rejected");
  // the numbering of the device arrays: a permutation of the host fields'
  for (Int kind = 0; kind < 3; ++kind) {
    const std::vector<Int> o = mesh->device_to_original(kind);
    std::vector<Int> seen(o.size(), 0);
    for (Int i : o) {
      require(i >= 0 && i < (Int)o.size() && !seen[i], "device_to_original is a permutation");
      seen[i] = 1;
    }
  }
  require(mesh->device_geometry("x_cell") != nullptr && mesh->device_geometry("angle_edge"),
          "device geometry available");
  std::cout << "fields: per-field device-array tendencies equal compute_tendency" << std::endl;
}

} // namespace

int main(int argc, char **argv) {
  const std::string what = argc > 1 ? argv[1] : "";
  try {
    if (what == "adjoint") adjoint();
    else if (what == "ensemble") ensemble();
    else if (what == "ensemble_adjoint") ensemble_adjoint();
    else if (what == "linear") linear();
    else if (what == "fields") fields();
    else {
      std::cerr << "usage: api_tests adjoint|ensemble|ensemble_adjoint|linear|fields"
                << std::endl;
      return 2;
    }
  } catch (const std::exception &e) {
    std::cerr << e.what() << std::endl;
    return 1;
  }
  return 0;
}
