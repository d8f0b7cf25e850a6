// The reference's end-to-end tests (ie_solver/ie_solver_test.cpp), run
// through the C++ facade on the GPU path.  Same cases and tolerances:
// backward error <= 50 id_tol (:29-96) and analytic agreement <= 10 id_tol
// (:378-562), Laplace Neumann included.  The perturbation/update tests are
// out of scope for the GPU path.  Prints one line per case; exit status 0
// iff all pass.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "ie_solver_gpu.hpp"

using namespace ie_solver;

static int g_failures = 0;

static void expect_le(const std::string& name, double v, double bound) {
  const bool ok = v <= bound;
  std::printf("%-44s %s  value %.3e  bound %.3e\n", name.c_str(), ok ? "PASS" : "FAIL", v, bound);
  if (!ok) g_failures++;
}

// ie_solver_test.cpp:29-96
static double solve_err(const ie_solver_config& config, Boundary* boundary) {
  QuadTree quadtree;
  quadtree.initialize_tree(boundary, std::vector<double>(), config.solution_dimension,
                           config.domain_dimension);
  SkelFactorization sf(config.id_tol, config.is_strong_admissibility, config.num_threads);
  Kernel kernel(config.solution_dimension, config.domain_dimension, config.pde, boundary,
                std::vector<double>());
  std::vector<int> all;
  for (int i = 0; i < config.solution_dimension * (int)boundary->weights.size(); i++)
    all.push_back(i);
  if (!boundary->holes.empty() && config.pde != ie_solver_config::LAPLACE_NEUMANN) {
    ie_Mat U = initialize_U_mat(config.pde, boundary->holes, boundary->points);
    ie_Mat Psi = initialize_Psi_mat(config.pde, boundary->holes, boundary);
    quadtree.U = U;
    quadtree.Psi = Psi;
    sf.skeletonize(kernel, &quadtree);
    ie_Mat mu, alpha;
    linear_solve(sf, quadtree, boundary->boundary_values, &mu, &alpha);
    ie_Mat K = kernel(all, all);
    const int n = (int)all.size(), h = alpha.height();
    double err = 0;
    for (int i = 0; i < n; i++) {
      double acc = -boundary->boundary_values.get(i, 0);
      for (int j = 0; j < n; j++) acc += K.get(i, j) * mu.get(j, 0);
      for (int k = 0; k < h; k++) acc += U.get(i, k) * alpha.get(k, 0);
      err = std::fmax(err, std::fabs(acc));
    }
    for (int k = 0; k < h; k++) {
      double acc = -alpha.get(k, 0);
      for (int j = 0; j < n; j++) acc += Psi.get(k, j) * mu.get(j, 0);
      err = std::fmax(err, std::fabs(acc));
    }
    return err;
  }
  sf.skeletonize(kernel, &quadtree);
  ie_Mat mu;
  linear_solve(sf, quadtree, boundary->boundary_values, &mu);
  ie_Mat err = (kernel(all, all) * mu) - boundary->boundary_values;
  return err.max_entry_magnitude();
}

// ie_solver_test.cpp:99-109
static ie_Mat get_dom_sol(const ie_solver_config& config, std::vector<double>* dom,
                          Boundary* boundary) {
  QuadTree quadtree;
  quadtree.initialize_tree(boundary, std::vector<double>(), config.solution_dimension,
                           config.domain_dimension);
  get_domain_points(config.domain_size, dom, quadtree.min, quadtree.max, quadtree.min,
                    quadtree.max);
  return boundary_integral_solve(config, &quadtree, *dom);
}

// ie_solver_test.cpp:378-401
static double laplace_error(const ie_Mat& sol, const std::vector<double>& dom, Boundary* b) {
  double dn = 0, nt = 0;
  for (size_t i = 0; i < dom.size(); i += 2) {
    const Vec2 x(dom[i], dom[i + 1]);
    if (!b->is_in_domain(x)) continue;
    const double pot = std::log(std::sqrt(std::pow(x.a[0] + 3, 2) + std::pow(x.a[1] + 2, 2))) / (2 * M_PI);
    dn += std::pow(pot - sol.get(i / 2, 0), 2);
    nt += pot * pot;
  }
  return std::sqrt(dn) / std::sqrt(nt);
}

// ie_solver_test.cpp:433-471: the field minus log r, less its mean.
static double laplace_neumann_error(const ie_Mat& sol, const std::vector<double>& dom,
                                    Boundary* b) {
  std::vector<double> d;
  for (size_t i = 0; i < dom.size(); i += 2) {
    const Vec2 x(dom[i], dom[i + 1]);
    if (!b->is_in_domain(x) || std::isnan(sol.get(i / 2, 0))) continue;
    const double r = std::sqrt(std::pow(x.a[0] - 0.5, 2) + std::pow(x.a[1] - 0.5, 2));
    const double v = sol.get(i / 2, 0) - std::log(r);
    if (v != 0) d.push_back(v);
  }
  double avg = 0;
  for (double v : d) avg += v;
  avg /= d.size();
  double mx = 0;
  for (double v : d) mx = std::fmax(mx, v - avg);
  return mx;
}

// ie_solver_test.cpp:474-507
static double stokes_error(const ie_Mat& sol, const std::vector<double>& dom, Boundary* b) {
  double mx = 0;
  for (size_t i = 0; i < dom.size(); i += 2) {
    const Vec2 x(dom[i], dom[i + 1]);
    if (!b->is_in_domain(x)) continue;
    const Vec2 r = x - Vec2(0.5, 0.5);
    Vec2 truth(-r.a[1], r.a[0]);
    if (b->boundary_shape == Boundary::CIRCLE) {
      truth = truth * (r.norm() / truth.norm()) * 4.;
    } else {
      const double p = (-1. / r.norm()) + 2 * r.norm();
      truth = truth * (1. / truth.norm()) * p;
    }
    const Vec2 s(sol.get(i, 0), sol.get(i + 1, 0));
    mx = std::fmax(mx, (truth - s).norm());
  }
  return mx;
}

template <class B>
static void backward_case(const std::string& name, ie_solver_config config) {
  B boundary;
  boundary.initialize(config.num_boundary_points, config.boundary_condition);
  expect_le(name, solve_err(config, &boundary), 50 * config.id_tol);
}

// A circle of radius 0.25 with r raised by 0.25 h (1 - u^2)^3 on the arc
// |u| < 1, u = (t - t0) / 0.5: a local, smooth deformation for the update
// tests (the reference moves parts of its Ex1-Ex3 boundaries instead).
class BumpCircle : public Boundary {
 public:
  double h = 0, t0 = 0;
  void initialize(int N, BoundaryCondition bc) override {
    clear_nodes();
    for (int i = 0; i < N; i++) {
      const double t = i * 2.0 * M_PI / N;
      const double u = std::remainder(t - t0, 2 * M_PI) / 0.5;
      double r = 0.25, dr = 0, ddr = 0;
      if (std::fabs(u) < 1) {
        const double g = 1 - u * u;
        r += 0.25 * h * g * g * g;
        dr += 0.25 * h * (-6 * u * g * g / 0.5);
        ddr += 0.25 * h * (-6 * g * (1 - 5 * u * u) / 0.25);
      }
      const double c = std::cos(t), sn = std::sin(t);
      const double x1 = dr * c - r * sn, y1 = dr * sn + r * c;  // d(x, y)/dt
      const double x2 = ddr * c - 2 * dr * sn - r * c, y2 = ddr * sn + 2 * dr * c - r * sn;
      const double sp = std::sqrt(x1 * x1 + y1 * y1);
      push_node(0.5 + r * c, 0.5 + r * sn, y1 / sp, -x1 / sp,
                (x1 * y2 - y1 * x2) / (sp * sp * sp), sp * 2 * M_PI / N);
    }
    if (bc == DEFAULT) {
      boundary_values = ie_Mat(N, 1);
      apply_boundary_condition(0, N, SINGLE_ELECTRON);
    } else {
      set_boundary_values_size(bc);
      apply_boundary_condition(0, N, bc);
    }
  }
  bool is_in_domain(const Vec2& a) override {
    const double x = a.a[0] - 0.5, y = a.a[1] - 0.5;
    return std::sqrt(x * x + y * y) + 2e-2 < 0.25 * (1 - std::fabs(h));
  }
};

int main() {
  const char* tol_env = std::getenv("IE_TEST_TOL");
  const double tol = tol_env ? std::atof(tol_env) : DEFAULT_ID_TOL;
  // Analytic agreement is also limited by the quadrature: the 20x20 grid of
  // ie_solver_test.cpp:105-107 reaches 0.01 from a 512-node boundary (node
  // spacing 0.012), where the trapezoid rule errs by ~1e-6 (multiply
  // connected cases).  The reference checks these at 10 x 1e-6; below that
  // tolerance they are held to that floor, while the backward-error cases
  // are held to 50 x tol.
  const double atol = 10 * std::fmax(tol, 1e-6);
  ie_solver_config lap;
  lap.id_tol = tol;
  ie_solver_config neu = lap;
  neu.pde = ie_solver_config::LAPLACE_NEUMANN;
  neu.boundary_condition = ALL_ONES;
  ie_solver_config sto = lap;
  sto.solution_dimension = 2;
  sto.pde = ie_solver_config::STOKES;
  sto.boundary_condition = TANGENT_VEC;

  // ie_solver_test.cpp:112-236 (N=512) and :240-372 (N=4096)
  for (int big = 0; big < 2; big++) {
    const std::string p = big ? "Big" : "";
    ie_solver_config l = lap, s = sto, nm = neu;
    if (big) {
      l.num_boundary_points = 4096;
      s.num_boundary_points = 4096;
      nm.num_boundary_points = 4096;
    }
    backward_case<Circle>(p + "LaplaceCircleBackwardError", l);
    backward_case<Circle>(p + "LaplaceNeumannCircleBackwardError", nm);
    backward_case<Circle>(p + "StokesCircleBackwardError", s);
    ie_solver_config na = nm;
    na.boundary_shape = Boundary::ANNULUS;
    backward_case<Annulus>(p + "LaplaceNeumannAnnulusBackwardError", na);
    ie_solver_config la = l;
    la.boundary_shape = Boundary::ANNULUS;
    backward_case<Annulus>(p + "LaplaceAnnulusBackwardError", la);
    ie_solver_config sa = s;
    sa.boundary_shape = Boundary::ANNULUS;
    backward_case<Annulus>(p + "StokesAnnulusBackwardError", sa);
  }

  {  // ie_solver_test.cpp:577-617 (Ex1UpdateLosesNoAcc): perturb the tree,
     // solve, and compare with a fresh tree of the same boundary
    ie_solver_config c = lap;
    c.num_boundary_points = 4096;
    BumpCircle b0;
    b0.initialize(c.num_boundary_points, c.boundary_condition);
    QuadTree qt;
    qt.initialize_tree(&b0, std::vector<double>(), 1, 2);
    std::vector<double> dom;
    get_domain_points(c.domain_size, &dom, qt.min, qt.max, qt.min, qt.max);
    boundary_integral_solve(c, &qt, dom);
    for (int frame = 0; frame < 3; frame++) {
      BumpCircle b1;
      b1.h = 0.004 * (frame + 1);
      b1.t0 = 1.0 + 2.0 * frame;
      b1.initialize(c.num_boundary_points, c.boundary_condition);
      qt.perturb(b1);
      const ie_Mat sol = boundary_integral_solve(c, &qt, dom);
      int n_dirty = 0, incremental = 0;
      iegpu_update_info(qt.ctx, &n_dirty, &incremental, nullptr);
      expect_le("UpdateIsIncremental frame " + std::to_string(frame), incremental ? 0 : 1, 0);
      QuadTree fresh;
      fresh.initialize_tree(&b1, std::vector<double>(), 1, 2);
      const ie_Mat new_sol = boundary_integral_solve(c, &fresh, dom);
      expect_le("UpdateLosesNoAcc frame " + std::to_string(frame),
                (new_sol - sol).max_entry_magnitude(), 100 * c.id_tol);
    }
  }
  {  // ie_solver_test.cpp:404-416
    Circle b;
    b.initialize(lap.num_boundary_points, lap.boundary_condition);
    std::vector<double> dom;
    ie_Mat sol = get_dom_sol(lap, &dom, &b);
    expect_le("LaplaceCircleAnalyticAgreementElectron", laplace_error(sol, dom, &b), atol);
  }
  {  // ie_solver_test.cpp:419-431
    Annulus b;
    b.initialize(lap.num_boundary_points, lap.boundary_condition);
    std::vector<double> dom;
    ie_Mat sol = get_dom_sol(lap, &dom, &b);
    expect_le("LaplaceAnnulusAnalyticAgreementElectron", laplace_error(sol, dom, &b), atol);
  }
  {  // ie_solver_test.cpp:510-525
    ie_solver_config c = neu;
    c.boundary_condition = DEFAULT;
    c.boundary_shape = Boundary::DONUT;
    Donut b;
    b.initialize(c.num_boundary_points, c.boundary_condition);
    std::vector<double> dom;
    ie_Mat sol = get_dom_sol(c, &dom, &b);
    expect_le("LaplaceNeumannDonutAnalyticAgreement", laplace_neumann_error(sol, dom, &b), atol);
  }
  {  // ie_solver_test.cpp:528-543
    Circle b;
    b.initialize(sto.num_boundary_points, sto.boundary_condition);
    std::vector<double> dom;
    ie_Mat sol = get_dom_sol(sto, &dom, &b);
    expect_le("StokesCircleAnalyticAgreementTangent", stokes_error(sol, dom, &b), atol);
  }
  {  // ie_solver_test.cpp:545-562
    ie_solver_config c = sto;
    c.num_boundary_points *= 4;
    c.boundary_shape = Boundary::DONUT;
    Donut b;
    b.initialize(c.num_boundary_points, c.boundary_condition);
    std::vector<double> dom;
    ie_Mat sol = get_dom_sol(c, &dom, &b);
    expect_le("StokesDonutAnalyticAgreementTangent", stokes_error(sol, dom, &b), atol);
  }
  {  // skel_factorization.cpp:388-406,507-592: a multi-column b solves every
     // column; each equals that column solved alone, bit for bit
    ie_solver_config c = lap;
    c.num_boundary_points = 4096;
    Circle b;
    b.initialize(c.num_boundary_points, c.boundary_condition);
    QuadTree q;
    q.initialize_tree(&b, std::vector<double>(), 1, 2);
    SkelFactorization sf(c.id_tol, false, c.num_threads);
    Kernel kernel(1, 2, c.pde, q.boundary, std::vector<double>());
    sf.skeletonize(kernel, &q);
    const int n = c.num_boundary_points;
    ie_Mat B(n, 3), X;
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < n; i++) B.set(i, j, std::cos(0.37 * (i + 1) * (j + 1)));
    sf.solve(q, &X, B);
    double diff = X.width() == 3 ? 0.0 : 1.0;
    for (int j = 0; j < 3; j++) {
      ie_Mat bj(n, 1), xj;
      for (int i = 0; i < n; i++) bj.set(i, 0, B.get(i, j));
      sf.solve(q, &xj, bj);
      for (int i = 0; i < n; i++) diff = std::fmax(diff, std::fabs(xj.get(i, 0) - X.get(i, j)));
    }
    expect_le("MultiColumnSolveEqualsColumns", diff, 0.0);

    // QuadTreeNode::T of a compressed box: its redundant columns of the
    // kernel against every DOF outside the proxy circle (both directions)
    // are the skeleton columns times T (interpolative decomposition).
    int nb = 0;
    iegpu_tree_info(q.ctx, nullptr, &nb, nullptr, nullptr);
    std::vector<int> lev(nb), leaf(nb), rank(nb);
    iegpu_box_info(q.ctx, lev.data(), nullptr, leaf.data(), nullptr, nullptr, rank.data(),
                   nullptr);
    int box = -1;
    for (int id = nb - 1; id >= 0 && box < 0; id--)
      if (leaf[id] && rank[id] > 0) box = id;
    ie_Mat T, L, U, Xi, S;
    q.node_factors(box, &T, &L, &U, &Xi, &S);
    std::vector<int> sk(T.height()), rd(T.width()), orig(4096);
    iegpu_box_list(q.ctx, box, 3, sk.data(), (int)sk.size());
    iegpu_box_list(q.ctx, box, 4, rd.data(), (int)rd.size());
    const int no = iegpu_box_list(q.ctx, box, 0, orig.data(), (int)orig.size());
    double cx = 0, cy = 0, rad = 0;
    for (int t = 0; t < no; t++) {
      cx += q.boundary->points[2 * orig[t]] / no;
      cy += q.boundary->points[2 * orig[t] + 1] / no;
    }
    for (int t = 0; t < no; t++)
      rad = std::fmax(rad, std::hypot(q.boundary->points[2 * orig[t]] - cx,
                                      q.boundary->points[2 * orig[t] + 1] - cy));
    std::vector<int> far;
    for (int i = 0; i < n; i++)
      if (std::hypot(q.boundary->points[2 * i] - cx, q.boundary->points[2 * i + 1] - cy) >
          6 * rad)
        far.push_back(i);
    const ie_Mat As = kernel(far, sk), Ar = kernel(far, rd);
    const ie_Mat Ast = kernel(sk, far), Art = kernel(rd, far);
    double err = 0, nrm = 0;
    for (int i = 0; i < (int)far.size(); i++)
      for (int jr = 0; jr < (int)rd.size(); jr++) {
        double a = Ar.get(i, jr), at = Art.get(jr, i);
        for (int js = 0; js < (int)sk.size(); js++) {
          a -= As.get(i, js) * T.get(js, jr);
          at -= Ast.get(js, i) * T.get(js, jr);
        }
        err = std::fmax(err, std::fmax(std::fabs(a), std::fabs(at)));
        nrm = std::fmax(nrm, std::fmax(std::fabs(Ar.get(i, jr)), std::fabs(Art.get(jr, i))));
      }
    expect_le("NodeFactorsTReconstructsFarField", err / nrm, 1e3 * c.id_tol);
    expect_le("NodeFactorsShapes", (double)(L.height() != T.height() || U.width() != T.height() ||
                                            Xi.height() != T.width() || S.height() != T.height()),
              0.0);
  }
  {  // skeletonize reads the caller's tree->U and tree->Psi
     // (linear_solve.cpp:256-257): a doubled U halves nothing else, and the
     // solution satisfies the caller's bordered system
    ie_solver_config c = lap;
    c.boundary_shape = Boundary::ANNULUS;
    Annulus b;
    b.initialize(c.num_boundary_points, c.boundary_condition);
    QuadTree q;
    q.initialize_tree(&b, std::vector<double>(), 1, 2);
    SkelFactorization sf(c.id_tol, false, c.num_threads);
    Kernel kernel(1, 2, c.pde, q.boundary, std::vector<double>());
    q.U = initialize_U_mat(c.pde, b.holes, b.points);
    q.Psi = initialize_Psi_mat(c.pde, b.holes, &b);
    for (int j = 0; j < q.U.width(); j++)
      for (int i = 0; i < q.U.height(); i++) q.U.set(i, j, 2 * q.U.get(i, j) + 0.25);
    sf.skeletonize(kernel, &q);
    ie_Mat mu, alpha;
    linear_solve(sf, q, b.boundary_values, &mu, &alpha);
    std::vector<int> all;
    for (int i = 0; i < (int)b.weights.size(); i++) all.push_back(i);
    const ie_Mat K = kernel(all, all);
    double err = 0;
    for (int i = 0; i < (int)all.size(); i++) {
      double acc = -b.boundary_values.get(i, 0);
      for (int j = 0; j < (int)all.size(); j++) acc += K.get(i, j) * mu.get(j, 0);
      for (int k = 0; k < alpha.height(); k++) acc += q.U.get(i, k) * alpha.get(k, 0);
      err = std::fmax(err, std::fabs(acc));
    }
    for (int k = 0; k < alpha.height(); k++) {
      double acc = -alpha.get(k, 0);
      for (int j = 0; j < (int)all.size(); j++) acc += q.Psi.get(k, j) * mu.get(j, 0);
      err = std::fmax(err, std::fabs(acc));
    }
    expect_le("CallersBorderIsUsed", err, 50 * c.id_tol);
  }
  {  // bie_time_trial (linear_solve.cpp:156-185) runs and reports times
    ie_solver_config c = lap;
    c.num_boundary_points = 4096;
    Circle b;
    b.initialize(c.num_boundary_points, c.boundary_condition);
    QuadTree q;
    q.initialize_tree(&b, std::vector<double>(), 1, 2);
    double ts = 0, tv = 0;
    bie_time_trial(c, &q, &ts, &tv);
    std::printf("bie_time_trial N=4096: skeletonize %.3f ms, solve %.3f ms\n", ts * 1e3, tv * 1e3);
    expect_le("TimeTrialRuns", ts > 0 && tv > 0 ? 0.0 : 1.0, 0.0);
  }
  std::printf("%d failure(s)\n", g_failures);
  return g_failures == 0 ? 0 : 1;
}
