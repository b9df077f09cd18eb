// Drives every entry of the CPU oracle (oracle/sdx_oracle.cpp) once or more,
// for a build with AddressSanitizer and UndefinedBehaviorSanitizer: scenes in
// both precisions with pins, per-element materials, plane and sphere contact,
// friction, stiffness and mass damping, every line search and PSD mode,
// Eisenstat-Walker forcing terms, the gradient-descent fallback, boundary
// conditions set and cleared, contact queries, assembly and objective; the
// synthetic Newton problems; the CSR PCG; the element and material routines.
//
// Input (argv[1]): "nV nT", 3 nV rest positions, 4 nT tet indices, then the
// count and indices of the nodes to pin.  Prints "OK" (exit 0) when every
// call returns its success code and every result is finite.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <vector>

#include "sdx_fem.h"

extern "C" {
void* sdxo_scene_create(int fp64);
void sdxo_scene_destroy(void* s);
int sdxo_set_threads(int n);
int sdxo_add_soft_tet_mesh(void* sv, const double* X, int nV, const int32_t* tets, int nT, const sdx_soft_params* sp);
int sdxo_set_element_materials(void* sv, int actor, const double* E, const double* nu, int nT);
int sdxo_pin_nodes(void* sv, int actor, const int32_t* nodes, int n);
int sdxo_add_collider(void* sv, int kind, const double* p, const double* nrm, double r, int actor);
int sdxo_set_collider_position(void* sv, int id, const double* p);
int sdxo_step(void* sv, double dt, const sdx_solver_params* sp, sdx_step_stats* stats, int nStats);
int sdxo_get_state(void* sv, int actor, int which, double* out);
int sdxo_set_state(void* sv, int actor, int which, const double* in);
int sdxo_actor_info(void* sv, int actor, int64_t* out);
int sdxo_assemble(void* sv, int actor, double dt, const sdx_solver_params* sp, const double* u0in, const double* v0in,
                  double* res, int32_t* ptr, int32_t* col, double* val);
int sdxo_objective(void* sv, int actor, double dt, const sdx_solver_params* sp, const double* u0in,
                   const double* v0in, double* out);
int sdxo_newton_synthetic(int kind, int n, const double* x0, const sdx_solver_params* sp, double* x,
                          sdx_step_stats* stats, double* merit);
double sdxo_forcing_term(int strategy, int k, double resNorm, double prevResNorm, double linResNorm, double prevEta,
                         double relTol);
int sdxo_set_contact_query(void* sv, int on);
int sdxo_get_contact_wrench(void* sv, int actor, int collider, double* out);
int sdxo_set_boundary_conditions(void* sv, int actor, const int32_t* dofs, const double* values, int n, int clear);
int sdxo_boundary(void* sv, int actor, int32_t* tris, double* area);
void sdxo_smith_nh(int fp64, double mu, double lambda, const double* F, double* energy, double* P, double* C,
                   int project);
void sdxo_polyrelu(double x, double t, double shift, double* out3);
void sdxo_stress_element(const double* X, const double* u, double mu, double lambda, int project, double* res,
                         double* H);
double sdxo_stress_damping_element(const double* X, const double* u, const double* us, double kappa, double mu,
                                   double lambda, int geo, int project, double* res, double* H);
int sdxo_pcg_csr(int fp64, int n, const int32_t* ptr, const int32_t* col, const double* val, const double* b,
                 double* x, int bs, const sdx_krylov_params* kp, int* iters, int* conv, double* resNorm);
}

namespace {

int g_fail = 0;

void expect(bool ok, const char* what) {
  if (!ok && g_fail++ < 10) std::printf("FAIL %s\n", what);
}

bool finite(const std::vector<double>& v) {
  for (double x : v)
    if (!std::isfinite(x)) return false;
  return true;
}

// The reference's defaults (nonlinear_solver_params.h, linear_solver_params.h).
sdx_solver_params defaults() {
  sdx_solver_params p{};
  p.newton_max_iter = 4;
  p.psd_proj_mode = SDX_PSD_ALWAYS;
  p.newton_abs_tol = 1e-3;
  p.newton_rel_tol = 1e-6;
  p.rel_step_tol = 10 * 1.1920928955078125e-07;
  p.explosion_control = 1;
  p.line_search_type = SDX_LS_RESIDUAL_NORM;
  p.abs_div_tol = 1e9;
  p.rel_div_tol = 1e4;
  p.line_search_max_iter = 4;
  p.linear_tolerance_strategy = SDX_LINTOL_EW2;
  p.line_search_alpha = 0.5;
  p.line_search_wolfe1 = 1e-4;
  p.line_search_wolfe2 = 0.9;
  p.linear.abs_tol = 1e-9;
  p.linear.rel_tol = 1e-5;
  p.linear.rel_div_tol = 1e10;
  p.linear.max_iter = 500;
  p.linear.norm_type = SDX_NORM_PRECONDITIONED_RESIDUAL_L2;
  p.linear.preconditioner = SDX_PREC_BLOCK_JACOBI;
  p.linear.initial_guess_zero = 1;
  p.gravity[1] = -9.8;
  return p;
}

sdx_soft_params material(double coulomb, double beta) {
  sdx_soft_params s{};
  s.youngs_modulus = 1e5;
  s.poisson_ratio = 0.4;
  s.density = 1000;
  s.mass_damping = 0.5;
  s.stiffness_damping = beta;
  s.penalty_coefficient = 1e6;
  s.penalty_smoothing_half_distance = 0.005;
  s.penalty_threshold = 0.001;
  s.coulomb_friction_coefficient = coulomb;
  s.viscous_friction_coefficient = coulomb > 0 ? 0.1 : 0.0;
  s.friction_falloff_vel = 0.01;
  s.normal_viscous_damping_coefficient = coulomb > 0 ? 0.1 : 0.0;
  s.stiffness_damping_geometric_term = beta > 0 ? 1 : 0;
  return s;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  std::ifstream in(argv[1]);
  int nV = 0, nT = 0, nPin = 0;
  in >> nV >> nT;
  std::vector<double> X(size_t(3) * nV);
  std::vector<int32_t> tets(size_t(4) * nT);
  for (auto& x : X) in >> x;
  for (auto& t : tets) in >> t;
  in >> nPin;
  std::vector<int32_t> pins(static_cast<size_t>(nPin));
  for (auto& p : pins) in >> p;
  if (!in) return 2;
  sdxo_set_threads(2);
  double ytop = 0;
  for (int i = 0; i < nV; ++i) ytop = std::max(ytop, X[size_t(3 * i + 1)]);

  for (int fp64 = 0; fp64 < 2; ++fp64) {
    void* sc = sdxo_scene_create(fp64);
    // actor 0: frictionless, per-element materials; actor 1: friction and stiffness damping
    for (int a = 0; a < 2; ++a) {
      std::vector<double> Xa = X;
      for (int i = 0; i < nV; ++i) Xa[size_t(3 * i)] += 0.2 * a;
      const sdx_soft_params sp = material(a ? 0.5 : 0.0, a ? 0.01 : 0.0);
      expect(sdxo_add_soft_tet_mesh(sc, Xa.data(), nV, tets.data(), nT, &sp) == a, "add mesh");  // its index
      expect(sdxo_pin_nodes(sc, a, pins.data(), nPin) == 0, "pin");
    }
    std::vector<double> E(static_cast<size_t>(nT)), nu(static_cast<size_t>(nT));
    for (int t = 0; t < nT; ++t) {
      E[size_t(t)] = 1e4 * std::pow(100.0, double(t % 7) / 6.0);
      nu[size_t(t)] = 0.3 + 0.02 * (t % 8);
    }
    expect(sdxo_set_element_materials(sc, 0, E.data(), nu.data(), nT) == 0, "materials");
    const double up[3] = {0, 1, 0};
    const double floor_p[3] = {0, -0.0005, 0};
    int ids[4];
    for (int a = 0; a < 2; ++a) {
      const double c[3] = {0.02 + 0.2 * a, ytop + 0.02 - 0.001, 0.02};
      ids[2 * a] = sdxo_add_collider(sc, 1, c, nullptr, 0.02, a);
      ids[2 * a + 1] = sdxo_add_collider(sc, 0, floor_p, up, 0.0, a);
      expect(ids[2 * a] >= 0 && ids[2 * a + 1] >= 0, "collider");
    }
    expect(sdxo_set_contact_query(sc, 1) == 0, "contact query");
    int64_t info[4];
    expect(sdxo_actor_info(sc, 0, info) == 0, "info");
    std::vector<int32_t> tris(size_t(3 * info[3]));
    std::vector<double> area(static_cast<size_t>(info[3]));
    expect(sdxo_boundary(sc, 0, tris.data(), area.data()) == 0 && finite(area), "boundary");

    std::vector<sdx_solver_params> runs;
    sdx_solver_params p = defaults();
    runs.push_back(p);
    for (int ls : {SDX_LS_NONE, SDX_LS_SIMPLE, SDX_LS_ARMIJO, SDX_LS_WOLFE_WEAK, SDX_LS_WOLFE_STRONG,
                   SDX_LS_ARMIJO_OR_RESIDUAL_NORM}) {
      sdx_solver_params q = defaults();
      q.line_search_type = ls;
      runs.push_back(q);
    }
    for (int psd : {SDX_PSD_NEVER, SDX_PSD_IF_FAIL_RETRY, SDX_PSD_IF_FAIL_ALWAYS}) {
      sdx_solver_params q = defaults();
      q.psd_proj_mode = psd;
      q.gradient_descent_fallback = 1;
      q.linear.max_iter = 5;  // unconverged linear solves: retries and the fallback run
      runs.push_back(q);
    }
    for (int lt : {SDX_LINTOL_CONSTANT, SDX_LINTOL_EW1, SDX_LINTOL_EW3}) {
      sdx_solver_params q = defaults();
      q.linear_tolerance_strategy = lt;
      q.linear.norm_type = SDX_NORM_RESIDUAL_L2;
      runs.push_back(q);
    }
    const double dt = 1e-3;
    for (size_t r = 0; r < runs.size(); ++r) {
      for (int a = 0; a < 2; ++a) {
        const double c[3] = {0.02 + 0.2 * a, ytop + 0.02 - 0.001 - 2e-4 * double(r + 1), 0.02};
        expect(sdxo_set_collider_position(sc, ids[2 * a], c) == 0, "move");
      }
      sdx_step_stats st[2];
      expect(sdxo_step(sc, dt, &runs[r], st, 2) == 0, "step");
      for (const auto& s : st) expect(std::isfinite(s.res_norm) && std::isfinite(s.res_norm0), "stats");
      double w[7];
      expect(sdxo_get_contact_wrench(sc, 0, ids[0], w) == 0 && std::isfinite(w[0] + w[3] + w[6]), "wrench");
    }
    // boundary conditions: move the pinned nodes, then clear them
    std::vector<int32_t> dofs;
    std::vector<double> vals;
    for (int k = 0; k < std::min(nPin, 4); ++k) {
      dofs.push_back(3 * pins[size_t(k)]);
      vals.push_back(X[size_t(3 * pins[size_t(k)])] + 1e-4);
    }
    expect(sdxo_set_boundary_conditions(sc, 0, dofs.data(), vals.data(), int(dofs.size()), 0) == 0, "bc set");
    sdx_step_stats st[2];
    expect(sdxo_step(sc, dt, &runs[0], st, 2) == 0, "step bc");
    expect(sdxo_set_boundary_conditions(sc, 0, nullptr, nullptr, 0, 1) == 0, "bc clear");
    expect(sdxo_step(sc, dt, &runs[0], st, 2) == 0, "step cleared");
    // state, assembly, objective
    std::vector<double> u(size_t(3) * nV), v(size_t(3) * nV), res(size_t(3) * nV);
    expect(sdxo_get_state(sc, 1, 0, u.data()) == 0 && sdxo_get_state(sc, 1, 1, v.data()) == 0 && finite(u) &&
               finite(v),
           "state");
    expect(sdxo_set_state(sc, 1, 1, v.data()) == 0, "set state");
    std::vector<int32_t> ptr(size_t(nV) + 1), col(static_cast<size_t>(info[2]));
    std::vector<double> val(size_t(9) * size_t(info[2]));
    expect(sdxo_assemble(sc, 1, dt, &runs[0], nullptr, v.data(), res.data(), ptr.data(), col.data(), val.data()) ==
                   0 &&
               finite(res) && finite(val),
           "assemble");
    double obj = 0;
    expect(sdxo_objective(sc, 1, dt, &runs[0], u.data(), v.data(), &obj) == 0 && std::isfinite(obj), "objective");
    sdxo_scene_destroy(sc);
  }

  // synthetic Newton problems, every line search
  for (int ls = SDX_LS_NONE; ls <= SDX_LS_ARMIJO_OR_RESIDUAL_NORM; ++ls) {
    sdx_solver_params p = defaults();
    p.line_search_type = ls;
    p.linear.norm_type = SDX_NORM_RESIDUAL_L2;
    p.newton_max_iter = 20;
    const double x1 = 3.0;
    double xo = 0, merit = 0;
    sdx_step_stats st{};
    expect(sdxo_newton_synthetic(0, 1, &x1, &p, &xo, &st, &merit) == 0 && std::isfinite(xo), "newton scalar");
    std::vector<double> x0(10, -1.0), x(10);
    expect(sdxo_newton_synthetic(1, 10, x0.data(), &p, x.data(), &st, &merit) == 0 && finite(x), "newton broyden");
  }
  for (int s = SDX_LINTOL_CONSTANT; s <= SDX_LINTOL_EW3; ++s)
    for (int k = 0; k < 3; ++k) expect(std::isfinite(sdxo_forcing_term(s, k, 0.5, 1.0, 0.1, 0.01, 1e-6)), "forcing");

  // CSR PCG on a tridiagonal SPD system, both precisions, scalar and 3x3 blocks
  const int n = 30;
  std::vector<int32_t> ptr(1, 0), col;
  std::vector<double> val, b(static_cast<size_t>(n), 1.0), x(static_cast<size_t>(n));
  for (int i = 0; i < n; ++i) {
    for (int j = std::max(0, i - 1); j <= std::min(n - 1, i + 1); ++j) {
      col.push_back(j);
      val.push_back(i == j ? 4.0 : -1.0);
    }
    ptr.push_back(int32_t(col.size()));
  }
  sdx_krylov_params kp = defaults().linear;
  kp.rel_tol = 1e-8;
  for (int fp64 = 0; fp64 < 2; ++fp64)
    for (int bs : {1, 3}) {
      int iters = 0, conv = 0;
      double rn = 0;
      expect(sdxo_pcg_csr(fp64, n, ptr.data(), col.data(), val.data(), b.data(), x.data(), bs, &kp, &iters, &conv,
                          &rn) == 0 &&
                 finite(x) && conv == SDX_CONV_CONVERGED,
             "pcg");
    }

  // element and material routines
  const double Xt[12] = {0, 0, 0, 0.02, 0.001, 0, 0.002, 0.018, 0.001, 0.001, 0.002, 0.021};
  const double ut[12] = {1e-4, -2e-4, 0, 3e-4, 0, 1e-4, -1e-4, 2e-4, 0, 0, 1e-4, -3e-4};
  const double zero[12] = {0};
  std::vector<double> r12(12), H(144), P(9), Cm(81);
  for (int pr = 0; pr < 2; ++pr) {
    sdxo_stress_element(Xt, ut, 1.5e5, 4e5, pr, r12.data(), H.data());
    expect(finite(r12) && finite(H), "stress element");
    const double e = sdxo_stress_damping_element(Xt, ut, zero, 3.0, 1.5e5, 4e5, 1, pr, r12.data(), H.data());
    expect(std::isfinite(e) && finite(r12) && finite(H), "damping element");
  }
  const double F[9] = {1.1, 0.05, 0, -0.02, 0.95, 0.01, 0, 0.03, 1.02};
  for (int fp64 = 0; fp64 < 2; ++fp64) {
    double en = 0;
    sdxo_smith_nh(fp64, 1.5e5, 4e5, F, &en, P.data(), Cm.data(), 1);
    expect(std::isfinite(en) && finite(P) && finite(Cm), "smith");
  }
  double pr3[3];
  sdxo_polyrelu(0.003, 0.005, 0.004, pr3);
  expect(std::isfinite(pr3[0] + pr3[1] + pr3[2]), "polyrelu");

  if (g_fail) return 1;
  std::printf("OK\n");
  return 0;
}
