// The reference's self-checking test programs (test/*.cpp of
// omega_playground), restated against the C++ drop-in API: every run goes
// through PlanarHexagonalMesh / ShallowWaterModel / RK4Stepper /
// LSRKStepper and the library's HIP kernels.  Initial conditions are
// written into the state's host mirrors with plain loops (the reference
// uses Kokkos parallel_for on the device views).
//
//   reference_tests mesh            host-only checks of the mesh API
//   reference_tests double_vortex   test/double_vortex.cpp:79-168
//   reference_tests nonlinear_swe   test/nonlinear_swe_convergence.cpp:115-212
//   reference_tests tracer adv|diff|hyper
//                                   test/tracer_{advection,diffusion,
//                                   hyperdiffusion}_convergence.cpp
//
// Exit status 0 on success; a failed check throws std::runtime_error, as the
// reference's tests do.
#include <omega.hpp>

#include <cmath>
#include <cstdio>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

using namespace omega;

namespace {

void require(bool ok, const std::string &what) {
  if (!ok) throw std::runtime_error("check failed: " + what);
}

Real rate_of(Real e0, Real e1) { return std::log2(e0 / e1); }

bool rate_ok(Real rate, Real expected, Real atol) {
  return std::abs(rate - expected) < atol && !std::isnan(rate);
}

// ---------------------------------------------------------------------------
// Mesh API (no device work)
// ---------------------------------------------------------------------------
void mesh_checks() {
  const Int nx = 6, ny = 4, L = 3;
  const Real dc = 2.5e3;
  PlanarHexagonalMesh mesh(nx, ny, dc, L);
  require(mesh.m_ncells == nx * ny && mesh.m_nedges == 3 * nx * ny &&
              mesh.m_nvertices == 2 * nx * ny,
          "entity counts");
  require(mesh.m_nlayers == L && mesh.m_nlayers_vec == L, "layer counts");
  require(std::abs(mesh.m_period_x - nx * dc) < 1e-9, "period_x");
  require(std::abs(mesh.m_period_y - ny * dc * std::sqrt(3.0) / 2) < 1e-6, "period_y");
  const Real area = std::sqrt(3.0) / 2 * dc * dc;
  for (Int c = 0; c < mesh.m_ncells; ++c) {
    require(mesh.m_nedges_on_cell(c) == 6, "six edges per cell");
    require(std::abs(mesh.m_area_cell(c) - area) < 1e-6 * area, "cell area");
    Real sum_sign = 0;
    for (Int j = 0; j < 6; ++j) {
      const Int e = mesh.m_edges_on_cell(c, j);
      const Real s = mesh.m_edge_sign_on_cell(c, j);
      require(s == (mesh.m_cells_on_edge(e, 0) == c ? 1.0 : -1.0), "edge_sign_on_cell");
      sum_sign += s;
    }
    (void)sum_sign;
  }
  for (Int e = 0; e < mesh.m_nedges; ++e) {
    require(std::abs(mesh.m_dc_edge(e) - dc) < 1e-9 * dc, "dc_edge");
    require(std::abs(mesh.m_dv_edge(e) - dc / std::sqrt(3.0)) < 1e-9 * dc, "dv_edge");
    require(mesh.m_nedges_on_edge(e) == 10, "ten edges on edge");
    for (Int k = 0; k < L; ++k) require(mesh.m_edge_mask(e, k) == 1.0, "edge_mask");
  }
  for (Int v = 0; v < mesh.m_nvertices; ++v)
    for (Int j = 0; j < 3; ++j) {
      const Int e = mesh.m_edges_on_vertex(v, j);
      require(mesh.m_edge_sign_on_vertex(v, j) == (mesh.m_vertices_on_edge(e, 0) == v ? -1.0 : 1.0),
              "edge_sign_on_vertex");
    }
  ShallowWaterParams p{1e-4, 9.81, 0};
  require(p.m_f0 == 1e-4 && p.m_grav == 9.81 && p.m_drag_coeff == 0 && p.m_ntracers == 0,
          "ShallowWaterParams aggregate order");
  std::cout << "mesh checks passed" << std::endl;
}

// ---------------------------------------------------------------------------
// test/double_vortex.cpp: mass, circulation and energy conservation
// ---------------------------------------------------------------------------
struct DoubleVortex {
  Real g = 9.80616, lx = 5000000, ly = 5000000, coriolis = 0.00006147, h0 = 750;
  Real ox = 0.1, oy = 0.1, dh = 75;
  Real sx = 3. / 40. * lx, sy = 3. / 40. * ly;
  Real xc1 = (0.5 - ox) * lx, yc1 = (0.5 - oy) * ly, xc2 = (0.5 + ox) * lx,
       yc2 = (0.5 + oy) * ly;

  // periodic Gaussian bumps in the stretched coordinates of the test
  Real p(Real d, Real len, Real s) const { return len / (pi * s) * std::sin(pi / len * d); }
  Real q(Real d, Real len, Real s) const {
    return len / (2 * pi * s) * std::sin(2 * pi / len * d);
  }
  Real bump(Real x, Real y, Real xc, Real yc) const {
    const Real a = p(x - xc, lx, sx), b = p(y - yc, ly, sy);
    return std::exp(-0.5 * (a * a + b * b));
  }
  Real h(Real x, Real y) const {
    return h0 - dh * (bump(x, y, xc1, yc1) + bump(x, y, xc2, yc2) -
                      4. * pi * sx * sy / lx / ly);
  }
  Real vx(Real x, Real y) const {
    return -g * dh / coriolis / sy *
           (q(y - yc1, ly, sy) * bump(x, y, xc1, yc1) + q(y - yc2, ly, sy) * bump(x, y, xc2, yc2));
  }
  Real vy(Real x, Real y) const {
    return g * dh / coriolis / sx *
           (q(x - xc1, lx, sx) * bump(x, y, xc1, yc1) + q(x - xc2, lx, sx) * bump(x, y, xc2, yc2));
  }
};

void double_vortex(Int nx, Real cfl) {
  DoubleVortex dv;
  const Real dc = dv.lx / nx;
  const Int ny = std::ceil(dv.ly / (dc * std::sqrt(3) / 2));
  auto mesh = std::make_unique<PlanarHexagonalMesh>(nx, ny, dc);
  ShallowWaterParams params;
  params.m_f0 = dv.coriolis;
  params.m_grav = dv.g;
  ShallowWaterModel model(mesh.get(), params);
  ShallowWaterState state(mesh.get(), params);
  LSRKStepper stepper(model);

  const Real timeend = 200 * 1000;
  Real dt = cfl * mesh->m_dc / std::sqrt(params.m_grav * dv.h0);
  const Int nsteps = std::ceil(timeend / dt);
  dt = timeend / nsteps;

  for (Int c = 0; c < mesh->m_ncells; ++c)
    for (Int k = 0; k < mesh->m_nlayers; ++k)
      state.m_h_cell(c, k) = dv.h(mesh->m_x_cell(c), mesh->m_y_cell(c));
  for (Int e = 0; e < mesh->m_nedges; ++e)
    for (Int k = 0; k < mesh->m_nlayers; ++k) {
      const Real x = mesh->m_x_edge(e), y = mesh->m_y_edge(e), a = mesh->m_angle_edge(e);
      state.m_vn_edge(e, k) = std::cos(a) * dv.vx(x, y) + std::sin(a) * dv.vy(x, y);
    }
  state.copy_to_device();

  const Real mass0 = mass_integral(state, model);
  const Real cir0 = circulation_integral(state, model);
  const Real en0 = energy_integral(state, model);
  for (Int step = 0; step < nsteps; ++step) stepper.do_step(step * dt, dt, state);
  const Real mass_change = (mass_integral(state, model) - mass0) / mass0;
  const Real cir_change = (circulation_integral(state, model) - cir0) / cir0;
  const Real en_change = (energy_integral(state, model) - en0) / en0;
  std::cout << "Mass change: " << mass_change << "\nCirculation change: " << cir_change
            << "\nEnergy change: " << en_change << std::endl;
  require(std::abs(mass_change) <= 5e-14, "mass conservation");
  require(std::abs(cir_change) <= 5e-14, "circulation conservation");
  require(std::abs(en_change) <= 1e-10, "energy conservation");
}

// ---------------------------------------------------------------------------
// test/nonlinear_swe_convergence.cpp: manufactured solution, RK4, rate 2
// ---------------------------------------------------------------------------
struct Manufactured {
  Real grav = 9.80616, f0 = 1e-4, lx = 10000 * 1e3, ly = std::sqrt(3) / 2 * lx, eta0 = 1,
       h0 = 1000;
  Real kx = 2 * (2 * pi / lx), ky = 2 * (2 * pi / ly);
  Real om = std::sqrt(grav * h0 * (kx * kx + ky * ky));

  Real phase(Real x, Real y, Real t) const { return kx * x + ky * y - om * t; }
  Real h(Real x, Real y, Real t) const { return h0 + eta0 * std::sin(phase(x, y, t)); }
  Real v(Real x, Real y, Real t) const { return eta0 * std::cos(phase(x, y, t)); }
  Real h_tend(Real x, Real y, Real t) const {
    const Real f = phase(x, y, t);
    return eta0 * (-h0 * (kx + ky) * std::sin(f) - om * std::cos(f) +
                   eta0 * (kx + ky) * std::cos(2 * f));
  }
  Real vx_tend(Real x, Real y, Real t) const {
    const Real f = phase(x, y, t);
    return eta0 * ((-f0 + grav * kx) * std::cos(f) + om * std::sin(f) -
                   eta0 * (kx + ky) * std::sin(2 * f) / 2);
  }
  Real vy_tend(Real x, Real y, Real t) const {
    const Real f = phase(x, y, t);
    return eta0 * ((f0 + grav * ky) * std::cos(f) + om * std::sin(f) -
                   eta0 * (kx + ky) * std::sin(2 * f) / 2);
  }
};

// The forcing is evaluated on the host in the original numbering and added
// to the device tendencies with add_host_field.
struct ManufacturedModel : ShallowWaterModel {
  Manufactured ms;
  mutable std::vector<Real> fh, fv;
  ManufacturedModel(MPASMesh *mesh, const Manufactured &m)
      : ShallowWaterModel(mesh, ShallowWaterParams{m.f0, m.grav, 0}), ms(m),
        fh(mesh->m_ncells), fv(mesh->m_nedges) {}

  void additional_tendency(DeviceArray2d h_tend_cell, DeviceArray2d vn_tend_edge,
                           ConstDeviceArray2d, ConstDeviceArray2d, Real t) const override {
    for (Int c = 0; c < m_mesh->m_ncells; ++c)
      fh[c] = ms.h_tend(m_mesh->m_x_cell(c), m_mesh->m_y_cell(c), t);
    for (Int e = 0; e < m_mesh->m_nedges; ++e) {
      const Real x = m_mesh->m_x_edge(e), y = m_mesh->m_y_edge(e), a = m_mesh->m_angle_edge(e);
      fv[e] = std::cos(a) * ms.vx_tend(x, y, t) + std::sin(a) * ms.vy_tend(x, y, t);
    }
    add_host_field(*m_mesh, 0, h_tend_cell, fh.data(), 1, 1.0, stream());
    add_host_field(*m_mesh, 1, vn_tend_edge, fv.data(), 1, 1.0, stream());
  }
};

Real nonlinear_swe_error(Int n) {
  Manufactured ms;
  auto mesh = std::make_unique<PlanarHexagonalMesh>(n, n, ms.lx / n);
  ManufacturedModel model(mesh.get(), ms);
  ShallowWaterState state(mesh.get(), 0);
  RK4Stepper stepper(model);
  const Real timeend = 10 * 60 * 60;
  Real dt = 3 * mesh->m_dc / 1e3;
  const Int nsteps = std::ceil(timeend / dt);
  dt = timeend / nsteps;
  for (Int c = 0; c < mesh->m_ncells; ++c)
    state.m_h_cell(c, 0) = ms.h(mesh->m_x_cell(c), mesh->m_y_cell(c), 0);
  for (Int e = 0; e < mesh->m_nedges; ++e) {
    const Real x = mesh->m_x_edge(e), y = mesh->m_y_edge(e), a = mesh->m_angle_edge(e);
    state.m_vn_edge(e, 0) = std::cos(a) * ms.v(x, y, 0) + std::sin(a) * ms.v(x, y, 0);
  }
  state.copy_to_device();
  for (Int step = 0; step < nsteps; ++step) stepper.do_step(step * dt, dt, state);
  state.copy_to_host();
  Real err = 0;
  for (Int c = 0; c < mesh->m_ncells; ++c) {
    const Real d = ms.h(mesh->m_x_cell(c), mesh->m_y_cell(c), timeend) - state.m_h_cell(c, 0);
    err += d * d;
  }
  return std::sqrt(err / (n * n));
}

void nonlinear_swe() {
  const Real e0 = nonlinear_swe_error(50), e1 = nonlinear_swe_error(100);
  const Real r = rate_of(e0, e1);
  std::cout << "errors " << e0 << " " << e1 << " rate " << r << std::endl;
  require(rate_ok(r, 2, 0.05), "nonlinear SWE convergence rate 2");
}

// ---------------------------------------------------------------------------
// test/tracer_{advection,diffusion,hyperdiffusion}_convergence.cpp
// ---------------------------------------------------------------------------
Real tracer_error(const std::string &kind, Int nx) {
  const Real lx = kind == "adv" ? 1.0 : 10.0, ly = std::sqrt(3) / 2 * lx;
  auto mesh = std::make_unique<PlanarHexagonalMesh>(nx, nx, lx / nx);
  ShallowWaterParams params{0, 9.81, 0};
  params.m_ntracers = 1;
  params.m_disable_h_tendency = true;
  params.m_disable_vn_tendency = true;
  if (kind == "diff") params.m_eddy_diff2 = 1.0;
  if (kind == "hyper") params.m_eddy_diff4 = 1.0;
  ShallowWaterModel model(mesh.get(), params);
  ShallowWaterState state(mesh.get(), params);
  LSRKStepper stepper(model, 5);
  const Real dc = mesh->m_dc;
  Real T, dt, vel;
  Real decay = 0;
  if (kind == "adv") {
    T = 1.0;
    dt = 0.5 * dc;
    vel = 1.0;
  } else {
    T = 2.0;
    const Real c = 4 * pi * pi * (1 / (lx * lx) + 1 / (ly * ly));
    decay = kind == "diff" ? c : c * c;
    dt = kind == "diff" ? dc * dc / 4 : dc * dc * dc * dc / 8;
    vel = 0.0;
  }
  auto exact = [&](Real x, Real y, Real t) {
    if (kind == "adv")
      return std::sin(2 * pi * (x - t) / lx) * std::sin(2 * pi * (y - t) / ly);
    return std::sin(2 * pi * x / lx) * std::sin(2 * pi * y / ly) * std::exp(-decay * t);
  };
  const Int nsteps = std::ceil(T / dt);
  dt = T / nsteps;
  for (Int c = 0; c < mesh->m_ncells; ++c) {
    const Real x = mesh->m_x_cell(c), y = mesh->m_y_cell(c);
    state.m_h_cell(c, 0) = kind == "adv" ? std::exp(std::cos(2 * pi * x / lx)) : 1.0;
    state.m_tr_cell(0, c, 0) = exact(x, y, 0);
  }
  for (Int e = 0; e < mesh->m_nedges; ++e) {
    const Real a = mesh->m_angle_edge(e);
    state.m_vn_edge(e, 0) = std::cos(a) * vel + std::sin(a) * vel;
  }
  state.copy_to_device();
  for (Int step = 0; step < nsteps; ++step) stepper.do_step(step * dt, dt, state);
  state.copy_to_host();
  Real err = 0;
  for (Int c = 0; c < mesh->m_ncells; ++c) {
    const Real d = exact(mesh->m_x_cell(c), mesh->m_y_cell(c), T) - state.m_tr_cell(0, c, 0);
    err += d * d;
  }
  return std::sqrt(err / (nx * nx));
}

void tracer(const std::string &kind) {
  const Int nx = kind == "adv" ? 50 : 16;
  const Real e0 = tracer_error(kind, nx), e1 = tracer_error(kind, 2 * nx);
  const Real r = rate_of(e0, e1);
  std::cout << "tracer " << kind << " errors " << e0 << " " << e1 << " rate " << r << std::endl;
  require(rate_ok(r, 2, 0.05), "tracer " + kind + " convergence rate 2");
}

} // namespace

int main(int argc, char **argv) {
  const std::string what = argc > 1 ? argv[1] : "";
  try {
    if (what == "mesh") mesh_checks();
    else if (what == "double_vortex") double_vortex(argc > 2 ? std::stoi(argv[2]) : 25, 0.1);
    else if (what == "nonlinear_swe") nonlinear_swe();
    else if (what == "tracer" && argc > 2) tracer(argv[2]);
    else {
      std::cerr << "usage: reference_tests mesh|double_vortex|nonlinear_swe|tracer adv|diff|hyper"
                << std::endl;
      return 2;
    }
  } catch (const std::exception &e) {
    std::cerr << e.what() << std::endl;
    return 1;
  }
  return 0;
}
