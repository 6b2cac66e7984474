// unit_edge_kernels.cpp -- the reference's edge-kernel unit tests, driven
// through the C++ adapter (nalu-wind_amd/host/nalu_adapter.h) the way
// unit_tests/edge_kernels/UnitTest{Continuity,Momentum,Scalar}*Edge.C drive
// the reference algorithms: one Hex8 ("generated:1x1x1"), the trigonometric
// test fields, helper objects that create the algorithm and execute it.
// Prints one JSON object per test; tests/test_gpu_adapter.py compares the
// numbers with the reference golds.  Given a mesh file and an output path it
// instead runs a realm read from that Exodus II file (create_mesh, device
// geometry, boundary-condition node and side sets, a continuity solve,
// results output), which the same test module checks against the oracle.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "nalu_adapter.h"
#include "nalu_amd_mesh.h"

using namespace nalu_amd;

namespace {

const double kPi = std::acos(-1.0);

struct Hex8Mesh {
  std::vector<double> coords;     // [8][3]
  std::vector<int32_t> conn;      // [1][8]
  std::vector<int32_t> edges;     // [12][2]
  std::vector<double> areav;      // [12][3]
  std::vector<double> dualvol;    // [8]
};

// TestKernelHex8Mesh::fill_mesh_and_init_fields
// (unit_tests/kernels/UnitTestKernelUtils.h:315-342)
Hex8Mesh fill_hex8_mesh()
{
  Hex8Mesh m;
  m.coords.resize(24);
  m.conn.resize(8);
  check(nw_hex_box(1, 1, 1, 1.0, 0.0, 0, m.coords.data(), m.conn.data()),
        "nw_hex_box");
  // stk::mesh::create_edges on one element: edges in element edge-ordinal
  // order, node[0] = the smaller id (the bucket order the golds rely on)
  static const int hex_edge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0},
                                      {4, 5}, {5, 6}, {6, 7}, {7, 4},
                                      {0, 4}, {1, 5}, {2, 6}, {3, 7}};
  std::vector<int32_t> ee(12);
  m.edges.resize(24);
  for (int k = 0; k < 12; ++k) {
    const int32_t a = m.conn[hex_edge[k][0]], b = m.conn[hex_edge[k][1]];
    m.edges[2 * k] = a < b ? a : b;
    m.edges[2 * k + 1] = a < b ? b : a;
    ee[k] = k;
  }
  m.areav.resize(36);
  check(nw_hex_edge_area_vector(1, m.conn.data(), m.coords.data(), ee.data(),
                                m.edges.data(), 12, m.areav.data()),
        "area");
  m.dualvol.resize(8);
  check(nw_hex_dual_volume(1, m.conn.data(), m.coords.data(), 8,
                           m.dualvol.data()),
        "dual volume");
  return m;
}

// TrigFieldFunction, a = 0.3 (unit_tests/kernels/UnitTestKernelUtils.C:31-367)
const double kA = 0.3;
std::vector<double> velocity(const Hex8Mesh& m)
{
  std::vector<double> v(24, 0.0);
  for (int n = 0; n < 8; ++n) {
    const double x = m.coords[3 * n], y = m.coords[3 * n + 1];
    v[3 * n] = -std::cos(kA * kPi * x) * std::sin(kA * kPi * y);
    v[3 * n + 1] = std::sin(kA * kPi * x) * std::cos(kA * kPi * y);
  }
  return v;
}
std::vector<double> dudx(const Hex8Mesh& m)
{
  std::vector<double> g(72, 0.0);
  const double ap = kA * kPi;
  for (int n = 0; n < 8; ++n) {
    const double x = m.coords[3 * n], y = m.coords[3 * n + 1];
    g[9 * n + 0] = ap * std::sin(ap * x) * std::sin(ap * y);
    g[9 * n + 1] = -ap * std::cos(ap * x) * std::cos(ap * y);
    g[9 * n + 3] = ap * std::cos(ap * x) * std::cos(ap * y);
    g[9 * n + 4] = -ap * std::sin(ap * x) * std::sin(ap * y);
  }
  return g;
}
std::vector<double> pressure(const Hex8Mesh& m)
{
  std::vector<double> p(8);
  for (int n = 0; n < 8; ++n)
    p[n] = -0.25 * (std::cos(2 * kA * kPi * m.coords[3 * n]) +
                    std::cos(2 * kA * kPi * m.coords[3 * n + 1]));
  return p;
}
std::vector<double> dpdx(const Hex8Mesh& m)
{
  std::vector<double> g(24, 0.0);
  for (int n = 0; n < 8; ++n) {
    g[3 * n] = 0.5 * kA * kPi * std::sin(2 * kA * kPi * m.coords[3 * n]);
    g[3 * n + 1] = 0.5 * kA * kPi * std::sin(2 * kA * kPi * m.coords[3 * n + 1]);
  }
  return g;
}

// calc_mass_flow_rate into the fixtures' 3-component "mass_flow_rate" edge
// field: edge ie's value lands at flat index ie, the algorithms read
// component 0 of each edge (flat 3e) -- the layout the golds were made with
std::vector<double> fixture_mdot(const Hex8Mesh& m, const std::vector<double>& rho,
                                 const std::vector<double>& vel)
{
  std::vector<double> md(12), flat(36, 0.0), used(12);
  check(nw_mdot_edge(12, m.edges.data(), m.areav.data(), rho.data(), vel.data(),
                     md.data()),
        "mdot");
  for (int e = 0; e < 12; ++e) flat[e] = md[e];
  for (int e = 0; e < 12; ++e) used[e] = flat[3 * e];
  return used;
}

void print_system(const char* name, const GpuLinearSystem& ls, bool last)
{
  std::vector<int64_t> ro;
  std::vector<int32_t> cols;
  std::vector<double> vals, rhs;
  ls.download(ro, cols, vals, rhs);
  std::printf("\"%s\": {\"rowOffsets\": [", name);
  for (size_t i = 0; i < ro.size(); ++i) std::printf("%s%lld", i ? "," : "", (long long)ro[i]);
  std::printf("], \"cols\": [");
  for (size_t i = 0; i < cols.size(); ++i) std::printf("%s%d", i ? "," : "", cols[i]);
  std::printf("], \"vals\": [");
  for (size_t i = 0; i < vals.size(); ++i) std::printf("%s%.17g", i ? "," : "", vals[i]);
  std::printf("], \"rhs\": [");
  for (size_t i = 0; i < rhs.size(); ++i) std::printf("%s%.17g", i ? "," : "", rhs[i]);
  std::printf("]}%s\n", last ? "" : ",");
}

// UnitTestContinuityAdvEdge.C:109-141
void test_continuity_advection()
{
  Hex8Mesh m = fill_hex8_mesh();
  Realm realm(0);
  realm.set_mesh(8, m.edges);
  realm.set_field(NW_F_COORDINATES, m.coords);
  realm.set_field(NW_E_AREA_VECTOR, m.areav);
  realm.set_field(NW_F_VELOCITY, velocity(m));
  realm.set_field(NW_F_PRESSURE, pressure(m));
  realm.set_field(NW_F_DPDX, dpdx(m));
  realm.set_field(NW_F_DENSITY, std::vector<double>(8, 1.0));
  realm.set_field(NW_F_UDIAG, std::vector<double>(8, 1.0));
  realm.timeIntegrator.gamma1 = 1.0;
  realm.timeIntegrator.timeStepN = 1.0;
  EquationSystem eq(realm, 1);
  eq.add_algorithm(std::unique_ptr<SolverAlgorithm>(new ContinuityEdgeSolverAlg(realm, &eq)));
  eq.initialize();
  eq.assemble_and_solve(nullptr);
  print_system("continuity_advection", eq.linsys(), false);
}

// UnitTestMomentumAdvDiffEdge.C:232-266
void test_momentum_advection_diffusion()
{
  Hex8Mesh m = fill_hex8_mesh();
  Realm realm(0);
  realm.set_mesh(8, m.edges);
  const auto vel = velocity(m);
  const std::vector<double> rho(8, 1.0);
  realm.set_field(NW_F_COORDINATES, m.coords);
  realm.set_field(NW_E_AREA_VECTOR, m.areav);
  realm.set_field(NW_F_VELOCITY, vel);
  realm.set_field(NW_F_DUDX, dudx(m));
  realm.set_field(NW_F_DENSITY, rho);
  realm.set_field(NW_F_VISCOSITY, std::vector<double>(8, 0.1));
  realm.set_field(NW_F_WALL_MASK, std::vector<double>(8, 1.0));
  realm.set_field(NW_E_MASS_FLOW_RATE, fixture_mdot(m, rho, vel));
  realm.set_field(NW_E_PECLET_FACTOR, std::vector<double>(12, 0.0));
  EquationSystem eq(realm, 3);
  eq.add_algorithm(std::unique_ptr<SolverAlgorithm>(new MomentumEdgeSolverAlg(realm, &eq)));
  eq.initialize();
  eq.assemble_and_solve(nullptr);
  print_system("momentum_advection_diffusion", eq.linsys(), false);
}

// MixtureFractionKernelHex8Mesh (UnitTestKernelUtils.h:1357-1464) and
// UnitTestScalarAdvDiffEdge.C:159-381
void setup_mixture_fraction(Realm& realm, const Hex8Mesh& m,
                            std::vector<double>& rho)
{
  std::vector<double> z(8), visc(8);
  rho.resize(8);
  for (int n = 0; n < 8; ++n) {
    // the fixture passes (amf_, znot_) = (2, 1) into (znot, amf)
    z[n] = 2.0 * std::cos(kPi * m.coords[3 * n]) *
           std::cos(kPi * m.coords[3 * n + 1]) * std::cos(kPi * m.coords[3 * n + 2]);
    rho[n] = 1.0 / (z[n] / 0.163 + (1.0 - z[n]) / 1.18);
    visc[n] = 1.967e-5 * z[n] + 1.85e-5 * (1.0 - z[n]);
  }
  const auto vel = velocity(m);
  realm.set_mesh(8, m.edges);
  realm.set_field(NW_F_COORDINATES, m.coords);
  realm.set_field(NW_E_AREA_VECTOR, m.areav);
  realm.set_field(NW_F_VELOCITY, vel);
  realm.set_field(NW_F_SCALAR_Q, z);
  realm.set_field(NW_F_DQDX, std::vector<double>(24, 0.0));
  realm.set_field(NW_F_DENSITY, rho);
  realm.set_field(NW_F_DIFF_FLUX_COEFF, visc);
  realm.set_field(NW_E_MASS_FLOW_RATE, fixture_mdot(m, rho, vel));
  realm.solutionOptions.alphaMap["mixture_fraction"] = 0.0;
  realm.solutionOptions.alphaUpwMap["mixture_fraction"] = 0.0;
  realm.solutionOptions.upwMap["mixture_fraction"] = 0.0;
}

void test_scalar_adv_diff()
{
  Hex8Mesh m = fill_hex8_mesh();
  std::vector<double> rho;
  {
    Realm realm(0);
    setup_mixture_fraction(realm, m, rho);
    EquationSystem eq(realm, 1);
    eq.add_algorithm(std::unique_ptr<SolverAlgorithm>(
      new ScalarEdgeSolverAlg(realm, &eq, "mixture_fraction")));
    eq.initialize();
    eq.assemble_and_solve(nullptr);
    print_system("scalar_adv_diff", eq.linsys(), false);
  }
  {
    // fix pressure at STK node id 1 to 1.0; "pressure" is the density here
    Realm realm(0);
    setup_mixture_fraction(realm, m, rho);
    realm.set_field(NW_F_PRESSURE, rho);
    EquationSystem eq(realm, 1);
    eq.add_algorithm(std::unique_ptr<SolverAlgorithm>(
      new ScalarEdgeSolverAlg(realm, &eq, "mixture_fraction")));
    eq.add_algorithm(std::unique_ptr<SolverAlgorithm>(
      new FixPressureAtNodeAlgorithm(realm, &eq, 0, 1.0)));
    eq.initialize();
    eq.assemble_and_solve(nullptr);
    print_system("scalar_adv_diff_fix_pressure", eq.linsys(), false);
  }
  {
    // applyDirichletBCs: solution 2 everywhere, bc 0 except node id 1 -> 1
    Realm realm(0);
    setup_mixture_fraction(realm, m, rho);
    EquationSystem eq(realm, 1);
    eq.add_algorithm(std::unique_ptr<SolverAlgorithm>(
      new ScalarEdgeSolverAlg(realm, &eq, "mixture_fraction")));
    eq.initialize();
    eq.assemble_and_solve(nullptr);
    realm.set_field(NW_F_SCALAR_Q, std::vector<double>(8, 2.0));
    std::vector<double> bc(8, 0.0);
    bc[0] = 1.0;
    realm.set_field(NW_F_DIFF_FLUX_COEFF, bc);
    std::vector<int32_t> all = {0, 1, 2, 3, 4, 5, 6, 7};
    eq.linsys().applyDirichletBCs(NW_F_SCALAR_Q, NW_F_DIFF_FLUX_COEFF, all, 0, 1);
    print_system("scalar_adv_diff_dirichlet", eq.linsys(), false);
  }
}

// UnitTestMdotAlg.C:21-78 (NGP_mdot_calc_edge): density 1, velocity 10,
// pressure and its gradient 0, rho*U interpolated together
void test_mdot_calc_edge()
{
  Hex8Mesh m = fill_hex8_mesh();
  Realm realm(0);
  realm.set_mesh(8, m.edges);
  realm.set_field(NW_F_COORDINATES, m.coords);
  realm.set_field(NW_E_AREA_VECTOR, m.areav);
  realm.set_field(NW_F_DENSITY, std::vector<double>(8, 1.0));
  realm.set_field(NW_F_VELOCITY, std::vector<double>(24, 10.0));
  realm.set_field(NW_F_PRESSURE, std::vector<double>(8, 0.0));
  realm.set_field(NW_F_DPDX, std::vector<double>(24, 0.0));
  realm.set_field(NW_F_UDIAG, std::vector<double>(8, 1.0));
  realm.solutionOptions.mdotInterpRhoUTogether = true;
  MdotEdgeAlg(realm).execute();
  const std::vector<double> mdot = realm.get_field(NW_E_MASS_FLOW_RATE);
  std::printf("\"mdot_calc_edge\": [");
  for (size_t i = 0; i < mdot.size(); ++i) std::printf("%s%.17g", i ? "," : "", mdot[i]);
  std::printf("],\n");
}

// UnitTestGeometryAlg.C:20-100 through GeometryInteriorAlg on the device
void test_geometry_unit_hex()
{
  Hex8Mesh m = fill_hex8_mesh();
  Realm realm(0);
  realm.set_mesh(8, m.edges);
  realm.set_field(NW_F_COORDINATES, m.coords);
  GeometryInteriorAlg geom(realm, 1, nullptr, m.conn.data());
  geom.execute();
  const std::vector<double> dv = realm.get_field(NW_F_DUAL_VOLUME);
  const std::vector<double> av = realm.get_field(NW_E_AREA_VECTOR);
  std::printf("\"geometry_unit_hex\": {\"dual_volume\": [");
  for (size_t i = 0; i < dv.size(); ++i) std::printf("%s%.17g", i ? "," : "", dv[i]);
  std::printf("], \"area_vector\": [");
  for (size_t i = 0; i < av.size(); ++i) std::printf("%s%.17g", i ? "," : "", av[i]);
  std::printf("]},\n");
}

void print_array(const char* name, const std::vector<double>& a, bool comma = true)
{
  std::printf("\"%s\": [", name);
  for (size_t i = 0; i < a.size(); ++i) std::printf("%s%.17g", i ? "," : "", a[i]);
  std::printf("]%s\n", comma ? "," : "");
}

void print_solve(const char* name, const GpuLinearSystem& ls, const nw_solve_stats& st,
                 double timerSolve)
{
  std::printf("\"%s\": {\"iters\": %d, \"converged\": %d, \"final_rel\": %.17g, "
              "\"final_abs\": %.17g, \"bnorm\": %.17g, \"linearResidual\": %.17g, "
              "\"nonLinearResidual\": %.17g, \"linearSolveIterations\": %d, "
              "\"timerSolve\": %.17g},\n",
              name, st.iters, st.converged, st.final_rel, st.final_abs, st.bnorm,
              ls.linearResidual(), ls.nonLinearResidual(), ls.linearSolveIterations(),
              timerSolve);
}

// Two equation systems on one Realm (continuity with a Dirichlet pressure
// on the top face, solved through EquationSystem::assemble_and_solve with a
// GMRES + Jacobi configuration; and a scalar adv-diff system), which take
// turns on the Realm's context, then the continuity system again on the
// hypre backend.  Box 5 x 4 x 6 stretched 1.1, jitter 0.15, seed 3; the
// fields are printed for the host-side checks.
void test_two_systems_one_realm()
{
  const int nx = 5, ny = 4, nz = 6;
  int64_t nn = 0, nel = 0;
  check(nw_hex_box_sizes(nx, ny, nz, &nn, &nel), "sizes");
  std::vector<double> coords(3 * nn);
  std::vector<int32_t> conn(8 * nel);
  check(nw_hex_box(nx, ny, nz, 1.1, 0.15, 3, coords.data(), conn.data()), "box");
  const int64_t cap = 3 * nn + 12 * nel;
  std::vector<int32_t> edges(2 * cap), ee(12 * nel);
  int64_t ne = 0;
  check(nw_hex_edges(nel, conn.data(), nn, edges.data(), cap, &ne, ee.data()), "edges");
  edges.resize(2 * ne);
  std::vector<double> areav(3 * ne);
  check(nw_hex_edge_area_vector(nel, conn.data(), coords.data(), ee.data(),
                                edges.data(), ne, areav.data()),
        "area");
  std::vector<double> vel(3 * nn), p(nn), dp(3 * nn), q(nn), rho(nn), gam(nn);
  for (int64_t n = 0; n < nn; ++n) {
    const double x = coords[3 * n], y = coords[3 * n + 1], z = coords[3 * n + 2];
    vel[3 * n] = std::sin(1.3 * x + 0.2) * std::cos(0.7 * y);
    vel[3 * n + 1] = std::cos(0.9 * x) * std::sin(1.1 * z + 0.4);
    vel[3 * n + 2] = 0.3 * std::sin(2.0 * y + z);
    p[n] = std::cos(1.7 * x + 0.5 * y) * (1.0 + z);
    dp[3 * n] = -1.7 * std::sin(1.7 * x + 0.5 * y) * (1.0 + z);
    dp[3 * n + 1] = -0.5 * std::sin(1.7 * x + 0.5 * y) * (1.0 + z);
    dp[3 * n + 2] = std::cos(1.7 * x + 0.5 * y);
    q[n] = std::sin(3.0 * x) * std::cos(2.0 * y) + z;
    rho[n] = 1.0 + 0.2 * std::sin(x + 2.0 * z);
    gam[n] = 1.0e-2 * (1.0 + x * y);
  }
  std::vector<double> mdot(ne);
  check(nw_mdot_edge(ne, edges.data(), areav.data(), rho.data(), vel.data(),
                     mdot.data()),
        "mdot");
  std::vector<int32_t> top;
  for (int64_t n = 0; n < nn; ++n)
    if (coords[3 * n + 2] >= 1.0 - 1e-12) top.push_back((int32_t)n);
  Realm realm(0);
  realm.set_mesh(nn, edges);
  realm.set_field(NW_F_COORDINATES, coords);
  realm.set_field(NW_E_AREA_VECTOR, areav);
  realm.set_field(NW_F_VELOCITY, vel);
  realm.set_field(NW_F_PRESSURE, p);
  realm.set_field(NW_F_DPDX, dp);
  realm.set_field(NW_F_DENSITY, rho);
  realm.set_field(NW_F_UDIAG, std::vector<double>(nn, 2.0));
  realm.set_field(NW_F_PRESSURE_BC, std::vector<double>(nn, 0.0));
  realm.set_field(NW_F_SCALAR_Q, q);
  realm.set_field(NW_F_DQDX, std::vector<double>(3 * nn, 0.0));
  realm.set_field(NW_F_DIFF_FLUX_COEFF, gam);
  realm.set_field(NW_E_MASS_FLOW_RATE, mdot);
  realm.timeIntegrator.timeStepN = 0.5;
  realm.timeIntegrator.gamma1 = 1.0;
  realm.solutionOptions.hybridMap["mixture_fraction"] = 1.0;
  realm.solutionOptions.alphaMap["mixture_fraction"] = 0.3;

  auto make_continuity = [&](EquationSystem& eq) {
    eq.add_algorithm(std::unique_ptr<SolverAlgorithm>(new ContinuityEdgeSolverAlg(realm, &eq)));
    eq.add_algorithm(std::unique_ptr<SolverAlgorithm>(
      new DirichletBC(realm, &eq, NW_F_PRESSURE, NW_F_PRESSURE_BC, top, 0, 1)));
  };
  EquationSystem cont(realm, 1);
  make_continuity(cont);
  EquationSystem scal(realm, 1);
  scal.add_algorithm(std::unique_ptr<SolverAlgorithm>(
    new ScalarEdgeSolverAlg(realm, &scal, "mixture_fraction")));
  cont.initialize();
  scal.initialize();  // now the scalar system holds the context

  nw_solver_cfg cfg{};
  cfg.kspace = 50;
  cfg.max_iters = 2000;
  cfg.tol = 1e-10;
  cfg.precond = NW_PREC_JACOBI;
  cfg.mode = NW_SOLVE_FP64;
  cfg.x0 = NW_X0_ZERO;

  std::printf("\"two_systems\": {\n");
  print_array("coords", coords);
  print_array("velocity", vel);
  print_array("pressure", p);
  print_array("dpdx", dp);
  print_array("density", rho);
  print_array("scalar_q", q);
  print_array("diff_flux_coeff", gam);
  print_array("mdot", mdot);
  std::printf("\"edges\": [");
  for (size_t i = 0; i < edges.size(); ++i) std::printf("%s%d", i ? "," : "", edges[i]);
  std::printf("],\n\"top\": [");
  for (size_t i = 0; i < top.size(); ++i) std::printf("%s%d", i ? "," : "", top[i]);
  std::printf("],\n");
  // the continuity system takes the context back (zeroSystem), assembles,
  // applies its Dirichlet rows and solves
  cont.assemble_and_solve(&cfg);
  print_system("continuity", cont.linsys(), false);
  print_array("continuity_solution", cont.linsys().solution());
  print_solve("continuity_solve", cont.linsys(), cont.linsys().lastStats(),
              cont.timerSolve_);
  // the scalar system next; the continuity system is then inactive
  scal.assemble_and_solve(nullptr);
  print_system("scalar", scal.linsys(), false);
  int inactive_throws = 0;
  try {
    cont.linsys().solve(cfg);
  } catch (const std::runtime_error&) {
    inactive_throws = 1;
  }
  std::printf("\"inactive_throws\": %d,\n", inactive_throws);
  // the same continuity equation system on the hypre backend
  EquationSystem hyp(realm, 1, LinearSystemType::Hypre);
  make_continuity(hyp);
  hyp.initialize();
  hyp.assemble_and_solve(&cfg);
  print_system("hypre_continuity", hyp.linsys(), false);
  print_array("hypre_solution", hyp.linsys().solution());
  print_solve("hypre_solve", hyp.linsys(), hyp.linsys().lastStats(), hyp.timerSolve_);
  std::printf("\"timers\": [%.17g, %.17g, %.17g]\n", cont.timerAssemble_,
              cont.timerLoadComplete_, cont.timerSolve_);
  std::printf("}\n");
}

// A realm read from an Exodus II file (Realm::create_mesh): the device
// geometry of its elements, a continuity system with a Dirichlet pressure on
// the node set "top" solved through EquationSystem::assemble_and_solve, the
// nodes of the side set "bottom", and the results written to `out`.
void test_exodus_realm(const char* mesh, const char* out)
{
  Realm realm(0);
  realm.create_mesh(mesh);
  GeometryInteriorAlg geom(realm);
  geom.execute();
  const int64_t nn = realm.nnode();
  const std::vector<double> coords = realm.get_field(NW_F_COORDINATES);
  std::vector<double> vel(3 * nn), p(nn), dp(3 * nn), rho(nn);
  for (int64_t n = 0; n < nn; ++n) {
    const double x = coords[3 * n], y = coords[3 * n + 1], z = coords[3 * n + 2];
    vel[3 * n] = std::sin(1.3 * x + 0.2) * std::cos(0.7 * y);
    vel[3 * n + 1] = std::cos(0.9 * x) * std::sin(1.1 * z + 0.4);
    vel[3 * n + 2] = 0.3 * std::sin(2.0 * y + z);
    p[n] = std::cos(1.7 * x + 0.5 * y) * (1.0 + z);
    dp[3 * n] = -1.7 * std::sin(1.7 * x + 0.5 * y) * (1.0 + z);
    dp[3 * n + 1] = -0.5 * std::sin(1.7 * x + 0.5 * y) * (1.0 + z);
    dp[3 * n + 2] = std::cos(1.7 * x + 0.5 * y);
    rho[n] = 1.0 + 0.2 * std::sin(x + 2.0 * z);
  }
  realm.set_field(NW_F_VELOCITY, vel);
  realm.set_field(NW_F_PRESSURE, p);
  realm.set_field(NW_F_DPDX, dp);
  realm.set_field(NW_F_DENSITY, rho);
  realm.set_field(NW_F_UDIAG, std::vector<double>(nn, 2.0));
  realm.set_field(NW_F_PRESSURE_BC, std::vector<double>(nn, 0.0));
  realm.timeIntegrator.timeStepN = 0.5;
  realm.timeIntegrator.gamma1 = 1.0;
  const std::vector<int32_t> top = realm.bc_nodes("top"), bottom = realm.bc_nodes("bottom");
  EquationSystem cont(realm, 1);
  cont.add_algorithm(std::unique_ptr<SolverAlgorithm>(new ContinuityEdgeSolverAlg(realm, &cont)));
  cont.add_algorithm(std::unique_ptr<SolverAlgorithm>(
    new DirichletBC(realm, &cont, NW_F_PRESSURE, NW_F_PRESSURE_BC, top, 0, 1)));
  cont.initialize();
  nw_solver_cfg cfg{};
  cfg.kspace = 50;
  cfg.max_iters = 2000;
  cfg.tol = 1e-10;
  cfg.precond = NW_PREC_AMG;
  cfg.mode = NW_SOLVE_FP64;
  cfg.x0 = NW_X0_ZERO;
  cont.assemble_and_solve(&cfg);
  realm.write_output(out, {{"pressure", NW_F_PRESSURE}, {"velocity", NW_F_VELOCITY},
                           {"dual_nodal_volume", NW_F_DUAL_VOLUME}}, 2.5);
  std::printf("\"exodus_realm\": {\n");
  std::printf("\"nnode\": %lld, \"nelem\": %lld, \"nedge\": %lld,\n", (long long)nn,
              (long long)realm.nelem(), (long long)realm.nedge());
  print_array("velocity", vel);
  print_array("pressure", p);
  print_array("dpdx", dp);
  print_array("density", rho);
  print_array("area_vector", realm.get_field(NW_E_AREA_VECTOR));
  print_array("dual_volume", realm.get_field(NW_F_DUAL_VOLUME));
  std::printf("\"top\": [");
  for (size_t i = 0; i < top.size(); ++i) std::printf("%s%d", i ? "," : "", top[i]);
  std::printf("],\n\"bottom\": [");
  for (size_t i = 0; i < bottom.size(); ++i) std::printf("%s%d", i ? "," : "", bottom[i]);
  std::printf("],\n");
  print_system("continuity", cont.linsys(), false);
  print_array("solution", cont.linsys().solution());
  print_solve("solve", cont.linsys(), cont.linsys().lastStats(), cont.timerSolve_);
  int missing_throws = 0;
  try {
    realm.bc_nodes("no_such_surface");
  } catch (const std::runtime_error&) {
    missing_throws = 1;
  }
  std::printf("\"missing_throws\": %d\n}\n", missing_throws);
}

}  // namespace

// With no arguments: the reference unit-test fixtures.  With a mesh file
// and an output path: the realm read from that Exodus II file.
int main(int argc, char** argv)
{
  try {
    if (argc == 3) {
      std::printf("{\n");
      test_exodus_realm(argv[1], argv[2]);
      std::printf("}\n");
      return 0;
    }
    std::printf("{\n");
    test_continuity_advection();
    test_momentum_advection_diffusion();
    test_scalar_adv_diff();
    test_mdot_calc_edge();
    test_geometry_unit_hex();
    test_two_systems_one_realm();
    std::printf("}\n");
  } catch (const std::exception& e) {
    std::fprintf(stderr, "FAILED: %s\n", e.what());
    return 1;
  }
  return 0;
}
