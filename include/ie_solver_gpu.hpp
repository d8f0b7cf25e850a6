// ie_solver_gpu.hpp -- C++ facade with the reference's API names over the
// C ABI in iegpu.h (header-only).
//
// Mirrors, by name and argument meaning (paths relative to the reference
// root):
//   ie_Mat                         ie_solver/ie_mat.h:19-111
//   Vec2, Hole, Boundary, Circle,  ie_solver/vec2.h, boundaries/boundary.h:11-65,
//   Annulus, Donut                 boundaries/{circle,annulus,donut}.cpp
//   ie_solver_config               ie_solver/io/io.h:18-38
//   QuadTree                       ie_solver/quadtree/quadtree.h:95-139
//   Kernel                         ie_solver/kernel/kernel.h:25-88
//   SkelFactorization              ie_solver/skel_factorization/skel_factorization.h:19-95
//   initialize_U_mat / _Psi_mat,   ie_solver/linear_solve.h:18-46
//   linear_solve, schur_solve,
//   bie_time_trial,
//   boundary_integral_solve,
//   get_domain_points
//
// The factorization lives on the GPU inside the QuadTree (as the
// reference keeps all factor state inside QuadTree).  Errors throw
// ie_solver::GpuError carrying the iegpu_status, where the reference
// asserts or exits.
#ifndef IE_SOLVER_GPU_HPP_
#define IE_SOLVER_GPU_HPP_

#include <chrono>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "iegpu.h"

namespace ie_solver {

#ifndef MAX_LEAF_DOFS
#define MAX_LEAF_DOFS 32  // quadtree.h:10
#endif
#define DEFAULT_NUM_DISCRETIZATION_POINTS 512  // io.h:11
#define DEFAULT_ID_TOL 1e-6                     // io.h:12
#define DEFAULT_DOMAIN_SIZE 20                  // io.h:13
#define DEFAULT_NUM_THREADS 4                   // io.h:14

class GpuError : public std::runtime_error {
 public:
  GpuError(int code, const std::string& msg) : std::runtime_error(msg), code(code) {}
  int code;
};

struct Vec2 {
  double a[2];
  Vec2() : a{0, 0} {}
  Vec2(double m, double n) : a{m, n} {}
  double norm() const { return std::sqrt(a[0] * a[0] + a[1] * a[1]); }
  double dot(const Vec2& o) const { return a[0] * o.a[0] + a[1] * o.a[1]; }
  Vec2 operator-(const Vec2& o) const { return Vec2(a[0] - o.a[0], a[1] - o.a[1]); }
  Vec2 operator*(double d) const { return Vec2(a[0] * d, a[1] * d); }
};

// Column-major dense matrix owning its buffer (lda == height).
struct ie_Mat {
  double* mat = nullptr;
  int lda_ = 0, height_ = 0, width_ = 0;

  ie_Mat() {}
  ie_Mat(int h, int w) : lda_(h), height_(h), width_(w) {
    if (h * w > 0) {
      mat = new double[(size_t)h * w];
      std::memset(mat, 0, sizeof(double) * (size_t)h * w);
    }
  }
  ie_Mat(const ie_Mat& o) : ie_Mat(o.height_, o.width_) {
    if (mat) std::memcpy(mat, o.mat, sizeof(double) * (size_t)height_ * width_);
  }
  ie_Mat(ie_Mat&& o) noexcept
      : mat(o.mat), lda_(o.lda_), height_(o.height_), width_(o.width_) {
    o.mat = nullptr;
  }
  ie_Mat& operator=(const ie_Mat& o) {
    if (this != &o) {
      ie_Mat t(o);
      *this = std::move(t);
    }
    return *this;
  }
  ie_Mat& operator=(ie_Mat&& o) noexcept {
    std::swap(mat, o.mat);
    lda_ = o.lda_;
    height_ = o.height_;
    width_ = o.width_;
    return *this;
  }
  ~ie_Mat() { delete[] mat; }

  int height() const { return height_; }
  int width() const { return width_; }
  double get(int i, int j) const { return mat[i + (size_t)lda_ * j]; }
  void set(int i, int j, double v) { mat[i + (size_t)lda_ * j] = v; }
  void addset(int i, int j, double v) { mat[i + (size_t)lda_ * j] += v; }
  void eye(int n) {
    *this = ie_Mat(n, n);
    for (int i = 0; i < n; i++) set(i, i, 1.0);
  }
  ie_Mat operator-() const {
    ie_Mat r(height_, width_);
    for (size_t i = 0; i < (size_t)height_ * width_; i++) r.mat[i] = -mat[i];
    return r;
  }
  ie_Mat operator-(const ie_Mat& o) const {
    ie_Mat r(height_, width_);
    for (size_t i = 0; i < (size_t)height_ * width_; i++) r.mat[i] = mat[i] - o.mat[i];
    return r;
  }
  ie_Mat operator+(const ie_Mat& o) const {
    ie_Mat r(height_, width_);
    for (size_t i = 0; i < (size_t)height_ * width_; i++) r.mat[i] = mat[i] + o.mat[i];
    return r;
  }
  // Host product (used by callers for residual checks, not on the GPU path).
  ie_Mat operator*(const ie_Mat& o) const {
    ie_Mat r(height_, o.width_);
    for (int j = 0; j < o.width_; j++)
      for (int l = 0; l < width_; l++) {
        const double b = o.get(l, j);
        if (b == 0) continue;
        for (int i = 0; i < height_; i++) r.mat[i + (size_t)height_ * j] += get(i, l) * b;
      }
    return r;
  }
  ie_Mat operator()(int row_s, int row_e, int col_s, int col_e) const {
    ie_Mat r(row_e - row_s, col_e - col_s);
    for (int j = 0; j < col_e - col_s; j++)
      for (int i = 0; i < row_e - row_s; i++) r.set(i, j, get(i + row_s, j + col_s));
    return r;
  }
  void set_submatrix(int row_s, int row_e, int col_s, int col_e, const ie_Mat& A) {
    for (int j = 0; j < col_e - col_s; j++)
      for (int i = 0; i < row_e - row_s; i++) set(i + row_s, j + col_s, A.get(i, j));
  }
  double max_entry_magnitude() const {
    double m = 0;
    for (size_t i = 0; i < (size_t)height_ * width_; i++) m = std::fmax(m, std::fabs(mat[i]));
    return m;
  }
  double frob_norm() const {
    double s = 0;
    for (size_t i = 0; i < (size_t)height_ * width_; i++) s += mat[i] * mat[i];
    return std::sqrt(s);
  }
};

struct Hole {
  Vec2 center;
  double radius = 0;
  int num_nodes = 0;
};

enum BoundaryCondition {
  SINGLE_ELECTRON, ALL_ONES, ALL_NEG_ONES, ALL_ZEROS, TANGENT_VEC,
  REVERSE_TANGENT_VEC, NORMAL_VEC, REVERSE_NORMAL_VEC, LEFT_TO_RIGHT_FLOW,
  NO_SLIP, HORIZONTAL_VEC, EX3A, EX3B, DEFAULT
};

class Boundary {
 public:
  enum BoundaryShape { CIRCLE, ROUNDED_SQUARE, ANNULUS, DONUT, CUBIC_SPLINE, EX1, EX2, EX3 };
  int num_outer_nodes = -1;
  std::vector<double> perturbation_parameters;
  std::vector<double> points, normals, curvatures, weights;
  std::vector<Hole> holes;
  ie_Mat boundary_values;
  BoundaryShape boundary_shape = CIRCLE;
  virtual ~Boundary() {}
  virtual void initialize(int n, BoundaryCondition bc) = 0;
  virtual bool is_in_domain(const Vec2& a) = 0;

  // boundary.cpp:11-40
  void set_boundary_values_size(BoundaryCondition bc) {
    const int np = (int)weights.size();
    const bool vec = bc == TANGENT_VEC || bc == REVERSE_TANGENT_VEC || bc == NORMAL_VEC ||
                     bc == REVERSE_NORMAL_VEC || bc == LEFT_TO_RIGHT_FLOW || bc == EX3B ||
                     bc == HORIZONTAL_VEC || bc == NO_SLIP;
    boundary_values = ie_Mat(vec ? 2 * np : np, 1);
  }
  // boundary.cpp:43-120
  void apply_boundary_condition(int start, int end, BoundaryCondition bc) {
    for (int p = start; p < end; p++) {
      const double x = points[2 * p], y = points[2 * p + 1];
      const double n0 = normals[2 * p], n1 = normals[2 * p + 1];
      switch (bc) {
        case SINGLE_ELECTRON:
          boundary_values.set(p, 0, std::log(std::sqrt(std::pow(x + 3, 2) + std::pow(y + 2, 2))) / (2 * M_PI));
          break;
        case ALL_ONES: boundary_values.set(p, 0, 1.); break;
        case ALL_NEG_ONES: boundary_values.set(p, 0, -1.); break;
        case ALL_ZEROS: boundary_values.set(p, 0, 0.); break;
        case TANGENT_VEC: boundary_values.set(2 * p, 0, -n1); boundary_values.set(2 * p + 1, 0, n0); break;
        case REVERSE_TANGENT_VEC: boundary_values.set(2 * p, 0, n1); boundary_values.set(2 * p + 1, 0, -n0); break;
        case NORMAL_VEC: boundary_values.set(2 * p, 0, n0); boundary_values.set(2 * p + 1, 0, n1); break;
        case REVERSE_NORMAL_VEC: boundary_values.set(2 * p, 0, -n0); boundary_values.set(2 * p + 1, 0, -n1); break;
        case HORIZONTAL_VEC: boundary_values.set(2 * p, 0, 1.); boundary_values.set(2 * p + 1, 0, 0.); break;
        case LEFT_TO_RIGHT_FLOW:
          boundary_values.set(2 * p, 0, (x < -0.99 || x > 1.99) ? 1. : 0.);
          boundary_values.set(2 * p + 1, 0, 0.);
          break;
        case NO_SLIP: boundary_values.set(2 * p, 0, 0.); boundary_values.set(2 * p + 1, 0, 0.); break;
        default: throw std::invalid_argument("boundary condition not applicable here");
      }
    }
  }

 protected:
  void push_node(double x, double y, double n0, double n1, double curv, double w) {
    points.push_back(x);
    points.push_back(y);
    normals.push_back(n0);
    normals.push_back(n1);
    curvatures.push_back(curv);
    weights.push_back(w);
  }
  void clear_nodes() {
    points.clear();
    normals.clear();
    curvatures.clear();
    weights.clear();
  }
};

// circle.cpp:9-44: centre (0.5,0.5), radius 0.25.
class Circle : public Boundary {
 public:
  void initialize(int N, BoundaryCondition bc) override {
    boundary_shape = CIRCLE;
    clear_nodes();
    for (int i = 0; i < N; i++) {
      const double ang = i * 2.0 * M_PI / N;
      push_node(0.5 + 0.25 * std::cos(ang), 0.5 + 0.25 * std::sin(ang), std::cos(ang),
                std::sin(ang), 4, M_PI / (N * 2));
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
    return !(std::sqrt(x * x + y * y) + 1e-2 > 0.25);
  }
};

// Unit circle about (0.5,0.5) with circular holes, N/5 nodes per hole
// (annulus.cpp:9-66, donut.cpp:9-66).
class HoledCircle : public Boundary {
 protected:
  void build(int N) {
    clear_nodes();
    const int hn = N / 5;
    for (int i = 0; i < N; i++) {
      const double ang = i * 2.0 * M_PI / N;
      push_node(0.5 + std::cos(ang), 0.5 + std::sin(ang), std::cos(ang), std::sin(ang), 1,
                2 * M_PI / N);
    }
    for (size_t h = 0; h < holes.size(); h++) {
      const Hole& ho = holes[h];
      for (int i = 0; i < hn; i++) {
        const double ang = i * 2.0 * M_PI / hn;
        push_node(ho.center.a[0] + ho.radius * std::cos(ang),
                  ho.center.a[1] + ho.radius * std::sin(ang), -std::cos(ang), -std::sin(ang),
                  -1.0 / ho.radius, (2 * ho.radius * M_PI) / hn);
      }
    }
  }

 public:
  bool is_in_domain(const Vec2& a) override {
    const double x = a.a[0] - 0.5, y = a.a[1] - 0.5;
    if (std::sqrt(x * x + y * y) + 1e-2 > 1) return false;
    for (const Hole& h : holes)
      if ((a - h.center).norm() - 1e-2 < h.radius) return false;
    return true;
  }
};

class Annulus : public HoledCircle {
 public:
  void initialize(int N, BoundaryCondition bc) override {
    boundary_shape = ANNULUS;
    if (holes.empty()) {
      Hole h;
      h.center = Vec2(0.9, 0.5);
      h.radius = 0.2;
      holes.push_back(h);
      h.center = Vec2(0.1, 0.5);
      holes.push_back(h);
    }
    build(N);
    const int np = (int)weights.size(), hn = N / 5;
    if (bc == DEFAULT) {
      boundary_values = ie_Mat(np, 1);
      apply_boundary_condition(0, N, ALL_ZEROS);
      apply_boundary_condition(N, N + hn, ALL_ONES);
      apply_boundary_condition(N + hn, N + 2 * hn, ALL_NEG_ONES);
    } else {
      set_boundary_values_size(bc);
      apply_boundary_condition(0, np, bc);
    }
  }
};

class Donut : public HoledCircle {
 public:
  void initialize(int N, BoundaryCondition bc) override {
    boundary_shape = DONUT;
    if (holes.empty()) {
      Hole h;
      h.center = Vec2(0.5, 0.5);
      h.radius = 0.5;
      holes.push_back(h);
    }
    build(N);
    const int np = (int)weights.size();
    if (bc == DEFAULT) {
      boundary_values = ie_Mat(np, 1);
      for (int i = 0; i < N; i++) boundary_values.set(i, 0, 1);
      for (int i = N; i < np; i++) boundary_values.set(i, 0, -1. / holes[0].radius);
    } else {
      set_boundary_values_size(bc);
      apply_boundary_condition(0, np, bc);
    }
  }
};

struct ie_solver_config {
  enum Pde { LAPLACE, LAPLACE_NEUMANN, STOKES };
  int num_boundary_points = DEFAULT_NUM_DISCRETIZATION_POINTS;
  int domain_size = DEFAULT_DOMAIN_SIZE;
  int domain_dimension = 2;
  int solution_dimension = 1;
  double id_tol = DEFAULT_ID_TOL;
  Pde pde = LAPLACE;
  bool is_strong_admissibility = false;
  BoundaryCondition boundary_condition = SINGLE_ELECTRON;
  Boundary::BoundaryShape boundary_shape = Boundary::CIRCLE;
  bool scaling = false;
  bool animation = false;
  int num_threads = DEFAULT_NUM_THREADS;
};

namespace detail {
inline void check(iegpu_ctx* c, int st) {
  if (st != IEGPU_OK) throw GpuError(st, c ? iegpu_last_error(c) : "iegpu call failed");
}
}  // namespace detail

// Owns the boundary copy and the GPU context that holds the tree and the
// factorization (quadtree.h:95-139).
class QuadTree {
 public:
  int solution_dimension = 1, domain_dimension = 2;
  int leaf_dofs = MAX_LEAF_DOFS;
  double min = 0, max = 0;
  Boundary* boundary = nullptr;
  std::vector<double> domain_points;
  ie_Mat U, Psi;
  iegpu_ctx* ctx = nullptr;
  int device = 0;

  QuadTree() {}
  QuadTree(const QuadTree&) = delete;
  QuadTree& operator=(const QuadTree&) = delete;
  ~QuadTree() {
    if (ctx) iegpu_destroy(ctx);
    delete owned_;
  }

  // quadtree.cpp:39-180.  The boundary arrays are copied; the tree itself
  // is built by the library when it is first needed.
  void initialize_tree(Boundary* b, const std::vector<double>& dom,
                       int sol_dim, int dom_dim) {
    delete owned_;
    owned_ = clone_shape(*b);
    boundary = owned_;
    domain_points = dom;
    solution_dimension = sol_dim;
    domain_dimension = dom_dim;
    min = max = b->points.empty() ? 0 : b->points[0];
    for (double p : b->points) {
      if (p < min) min = p;
      if (p > max) max = p;
    }
    if (!ctx) detail::check(nullptr, iegpu_create(&ctx, device));
    pde_ = -1;
  }
  void reset() { pde_ = -1; }

  // quadtree.cpp:520-801: the boundary moved.  With the same node count
  // the library puts the nodes into the kept cells and the next
  // skeletonize re-factors only the boxes the reference marks
  // (iegpu_update_geometry; a tree it cannot keep is rebuilt).  With
  // another count the next skeletonize builds a new tree.  min / max keep
  // the first tree's bounds, as the reference's do.
  // The new arrays are copied into the tree's own boundary, as the
  // reference does (quadtree.cpp:718-725): the caller may free b after.
  void perturb(const Boundary& b) {
    if (!owned_) throw std::invalid_argument("perturb before initialize_tree");
    const bool same_n = b.weights.size() == owned_->weights.size();
    copy_arrays(b, owned_);
    if (pde_ >= 0 && same_n) {
      int kept = 0;
      detail::check(ctx, iegpu_update_geometry(ctx, (int)owned_->weights.size(),
                                               owned_->points.data(), owned_->normals.data(),
                                               owned_->curvatures.data(),
                                               owned_->weights.data(), &kept));
    } else {
      pde_ = -1;
    }
  }

  // Uploads geometry and builds the device tree for `pde` (done once).
  void prepare(int pde) {
    if (pde_ == pde) return;
    const int n = (int)boundary->weights.size();
    detail::check(ctx, iegpu_set_geometry(ctx, pde, n, boundary->points.data(),
                                          boundary->normals.data(),
                                          boundary->curvatures.data(),
                                          boundary->weights.data()));
    detail::check(ctx, iegpu_build_tree(ctx, leaf_dofs));
    pde_ = pde;
  }
  int num_levels() const {
    int nl = 0;
    if (ctx) iegpu_tree_info(ctx, &nl, nullptr, nullptr, nullptr);
    return nl;
  }
  // The factors skeletonize left on node `id` (QuadTreeNode::T, L, U,
  // X_rr (inverted here) and schur_update, quadtree.h:55-66), read back
  // from the device; all empty for an uncompressed node.
  void node_factors(int id, ie_Mat* T, ie_Mat* L, ie_Mat* U, ie_Mat* X_rr_inv,
                    ie_Mat* schur) const {
    int k = 0, r = 0;
    detail::check(ctx, iegpu_box_factors(ctx, id, &k, &r, nullptr, nullptr, nullptr, nullptr,
                                         nullptr));
    *T = ie_Mat(k * (r > 0), r);
    *L = ie_Mat(k * (r > 0), r);
    *U = ie_Mat(r, k * (r > 0));
    *X_rr_inv = ie_Mat(r, r);
    *schur = ie_Mat(k * (r > 0), k * (r > 0));
    if (r == 0) return;
    detail::check(ctx, iegpu_box_factors(ctx, id, nullptr, nullptr, T->mat, L->mat, U->mat,
                                         X_rr_inv->mat, schur->mat));
  }

 private:
  // quadtree.cpp:718-725 (and the copies of quadtree.cpp:78-86).
  static void copy_arrays(const Boundary& b, Boundary* to) {
    to->points = b.points;
    to->normals = b.normals;
    to->weights = b.weights;
    to->curvatures = b.curvatures;
    to->boundary_values = b.boundary_values;
    to->perturbation_parameters = b.perturbation_parameters;
    to->holes = b.holes;
    to->num_outer_nodes = b.num_outer_nodes;
    to->boundary_shape = b.boundary_shape;
  }
  // A boundary of b's shape holding b's arrays, which the tree owns
  // (quadtree.cpp:47-86).  Shapes this header does not define keep b's
  // domain test by reference, so b must outlive the tree for them.
  class Copy : public Boundary {
   public:
    explicit Copy(const Boundary& b) : inner_(&b) {}
    void initialize(int, BoundaryCondition) override {}
    bool is_in_domain(const Vec2& a) override {
      return const_cast<Boundary*>(inner_)->is_in_domain(a);
    }

   private:
    const Boundary* inner_;
  };
  static Boundary* clone_shape(const Boundary& b) {
    Boundary* c;
    switch (b.boundary_shape) {
      case Boundary::CIRCLE: c = new Circle(); break;
      case Boundary::ANNULUS: c = new Annulus(); break;
      case Boundary::DONUT: c = new Donut(); break;
      default: c = new Copy(b); break;
    }
    copy_arrays(b, c);
    return c;
  }
  Boundary* owned_ = nullptr;
  int pde_ = -1;
};

// kernel.h:25-88; entries evaluated on the host (for residual checks).
struct Kernel {
  int solution_dimension, domain_dimension;
  ie_solver_config::Pde pde;
  Boundary* boundary;
  std::vector<double> domain_points;
  Kernel(int sd, int dd, ie_solver_config::Pde p, Boundary* b, std::vector<double> dom)
      : solution_dimension(sd), domain_dimension(dd), pde(p), boundary(b),
        domain_points(std::move(dom)) {}

  // kernel.cpp:123-165 (Laplace), :168-209 (Laplace Neumann) and :212-290
  // (Stokes).
  ie_Mat operator()(const std::vector<int>& I, const std::vector<int>& J) const {
    ie_Mat K((int)I.size(), (int)J.size());
    const auto& P = boundary->points;
    const auto& N = boundary->normals;
    const auto& W = boundary->weights;
    const auto& C = boundary->curvatures;
    for (int j = 0; j < (int)J.size(); j++)
      for (int i = 0; i < (int)I.size(); i++) {
        double v;
        if (pde == ie_solver_config::STOKES) {
          const int t = I[i], s = J[j], tp = t / 2, sp = s / 2, a = t % 2, b = s % 2;
          const double r0 = P[2 * tp] - P[2 * sp], r1 = P[2 * tp + 1] - P[2 * sp + 1];
          const double sn[2] = {N[2 * sp], N[2 * sp + 1]}, tn[2] = {N[2 * tp], N[2 * tp + 1]};
          if (r0 == 0 && r1 == 0) {
            const double pot = -0.5 * C[sp] * W[sp] / M_PI;
            const double tt[2] = {-sn[1], sn[0]};
            v = (a == b ? -0.5 : 0.0) + pot * tt[a] * tt[b] + W[sp] * tn[a] * sn[b];
          } else {
            const double d2 = r0 * r0 + r1 * r1, r[2] = {r0, r1};
            const double pot = W[sp] / M_PI * (r0 * sn[0] + r1 * sn[1]) / (d2 * d2);
            v = pot * r[a] * r[b] + W[sp] * tn[a] * sn[b];
          }
        } else if (pde == ie_solver_config::LAPLACE_NEUMANN) {  // kernel.cpp:168-209
          const int t = I[i], s = J[j];
          const double r0 = P[2 * t] - P[2 * s], r1 = P[2 * t + 1] - P[2 * s + 1];
          if (r0 == 0 && r1 == 0) v = W[s] - 0.5 + 0.5 * C[s] * W[s] / (2 * M_PI);
          else v = W[s] + W[s] / (2 * M_PI) * (r0 * N[2 * t] + r1 * N[2 * t + 1]) / (r0 * r0 + r1 * r1);
        } else {
          const int t = I[i], s = J[j];
          const double r0 = P[2 * t] - P[2 * s], r1 = P[2 * t + 1] - P[2 * s + 1];
          if (r0 == 0 && r1 == 0) v = 0.5 + 0.5 * C[s] * W[s] / (2 * M_PI);
          else v = -W[s] / (2 * M_PI) * (r0 * N[2 * s] + r1 * N[2 * s + 1]) / (r0 * r0 + r1 * r1);
        }
        K.set(i, j, v);
      }
    return K;
  }
};

inline int gpu_pde(ie_solver_config::Pde p) {
  if (p == ie_solver_config::LAPLACE) return IEGPU_PDE_LAPLACE;
  if (p == ie_solver_config::LAPLACE_NEUMANN) return IEGPU_PDE_LAPLACE_NEUMANN;
  return IEGPU_PDE_STOKES;
}

// skel_factorization.h:19-95.
class SkelFactorization {
 public:
  double id_tol = DEFAULT_ID_TOL;
  int num_threads = DEFAULT_NUM_THREADS;  // unused: the GPU path is one device
  bool strong_admissibility = false;
  int precision = IEGPU_FP64;

  SkelFactorization() {}
  SkelFactorization(double tol, bool strong, int nt)
      : id_tol(tol), num_threads(nt), strong_admissibility(strong) {
    if (!(tol > 0)) throw GpuError(IEGPU_ERR_INVALID, "id_tol must be > 0");
    if (strong) throw GpuError(IEGPU_ERR_INVALID, "strong admissibility is not implemented");
  }

  // skel_factorization.cpp:140-310.  The border uses the caller's tree->U
  // and tree->Psi (linear_solve.cpp:256-257); the hole centres stay with
  // the library for interior evaluation.
  void skeletonize(const Kernel& K, QuadTree* tree) const {
    tree->prepare(gpu_pde(K.pde));
    if (tree->U.width() > 0) {
      std::vector<double> c, r;
      for (const Hole& h : tree->boundary->holes) {
        c.push_back(h.center.a[0]);
        c.push_back(h.center.a[1]);
        r.push_back(h.radius);
      }
      detail::check(tree->ctx, iegpu_set_holes(tree->ctx, (int)r.size(), c.data(), r.data()));
      int hd = 0, nd = 0;
      iegpu_tree_info(tree->ctx, nullptr, nullptr, &nd, &hd);
      if (hd > 0) {
        if (tree->U.height() != nd || tree->U.width() != hd || tree->Psi.height() != hd ||
            tree->Psi.width() != nd)
          throw GpuError(IEGPU_ERR_INVALID, "tree->U / tree->Psi do not match the holes' border");
        detail::check(tree->ctx, iegpu_set_border(tree->ctx, hd, tree->U.mat, tree->Psi.mat));
      }
    } else {
      detail::check(tree->ctx, iegpu_set_holes(tree->ctx, 0, nullptr, nullptr));
    }
    detail::check(tree->ctx, iegpu_factor(tree->ctx, id_tol, precision));
  }
  // skel_factorization.cpp:507-592: every column of b (apply_sweep_matrix
  // sweeps vec->width() columns, :388-406).
  void solve(const QuadTree& tree, ie_Mat* x, const ie_Mat& b) const {
    if (x->height() != b.height() || x->width() != b.width()) *x = ie_Mat(b.height(), b.width());
    detail::check(tree.ctx, iegpu_solve_nrhs(tree.ctx, b.width(), b.mat, x->mat, nullptr));
  }
  // skel_factorization.cpp:596-777, column by column of b.
  void multiply_connected_solve(const QuadTree& tree, ie_Mat* mu, ie_Mat* alpha,
                                const ie_Mat& b) const {
    if (mu->height() != b.height() || mu->width() != b.width())
      *mu = ie_Mat(b.height(), b.width());
    int hd = 0;
    iegpu_tree_info(tree.ctx, nullptr, nullptr, nullptr, &hd);
    if (alpha->height() != hd || alpha->width() != b.width()) *alpha = ie_Mat(hd, b.width());
    detail::check(tree.ctx, iegpu_solve_nrhs(tree.ctx, b.width(), b.mat, mu->mat, alpha->mat));
  }
};

// linear_solve.cpp:20-67
inline ie_Mat initialize_U_mat(ie_solver_config::Pde pde, const std::vector<Hole>& holes,
                               const std::vector<double>& tgt) {
  if (holes.empty()) return ie_Mat(0, 0);
  const int h = (int)holes.size(), nt = (int)tgt.size() / 2;
  if (pde == ie_solver_config::LAPLACE) {
    ie_Mat U(nt, h);
    for (int i = 0; i < nt; i++)
      for (int k = 0; k < h; k++)
        U.set(i, k, std::log((Vec2(tgt[2 * i], tgt[2 * i + 1]) - holes[k].center).norm()));
    return U;
  }
  if (pde != ie_solver_config::STOKES) return ie_Mat(0, 0);
  ie_Mat U(2 * nt, 3 * h);
  const double s = 1.0 / (4 * M_PI);
  for (int i = 0; i < nt; i++)
    for (int k = 0; k < h; k++) {
      const Vec2 r = Vec2(tgt[2 * i], tgt[2 * i + 1]) - holes[k].center;
      const double n2 = std::pow(r.norm(), 2);
      U.set(2 * i, 3 * k, s * (std::log(1 / r.norm()) + r.a[0] * r.a[0] / n2));
      U.set(2 * i + 1, 3 * k, s * (r.a[1] * r.a[0] / n2));
      U.set(2 * i, 3 * k + 1, s * (r.a[0] * r.a[1] / n2));
      U.set(2 * i + 1, 3 * k + 1, s * (std::log(1 / r.norm()) + r.a[1] * r.a[1] / n2));
      U.set(2 * i, 3 * k + 2, r.a[1] * (s / n2));
      U.set(2 * i + 1, 3 * k + 2, -r.a[0] * (s / n2));
    }
  return U;
}

// linear_solve.cpp:70-112
inline ie_Mat initialize_Psi_mat(ie_solver_config::Pde pde, const std::vector<Hole>& holes,
                                 Boundary* b) {
  if (holes.empty()) return ie_Mat(0, 0);
  const int h = (int)holes.size(), np = (int)b->weights.size();
  const bool st = pde == ie_solver_config::STOKES;
  if (!st && pde != ie_solver_config::LAPLACE) return ie_Mat(0, 0);
  ie_Mat Psi(st ? 3 * h : h, st ? 2 * np : np);
  for (int i = 0; i < np; i++) {
    const Vec2 x(b->points[2 * i], b->points[2 * i + 1]);
    for (int k = 0; k < h; k++) {
      if ((x - holes[k].center).norm() < holes[k].radius + 1e-8) {
        const double w = b->weights[i];
        if (!st) {
          Psi.set(k, i, w);
        } else {
          Psi.set(3 * k, 2 * i, w);
          Psi.set(3 * k + 1, 2 * i + 1, w);
          Psi.set(3 * k + 2, 2 * i, w * x.a[1]);
          Psi.set(3 * k + 2, 2 * i + 1, -w * x.a[0]);
        }
        break;
      }
    }
  }
  return Psi;
}

// linear_solve.cpp:115-129
inline void linear_solve(const SkelFactorization& sf, const QuadTree& tree, const ie_Mat& f,
                         ie_Mat* mu, ie_Mat* alpha = nullptr) {
  *mu = ie_Mat((int)tree.boundary->weights.size() * tree.solution_dimension, 1);
  if (alpha == nullptr) {
    sf.solve(tree, mu, f);
  } else {
    *alpha = ie_Mat(tree.U.width(), 1);
    sf.multiply_connected_solve(tree, mu, alpha, f);
  }
}

// linear_solve.cpp:140-154, with K_domain mu + U_forward alpha evaluated on
// the GPU (iegpu_evaluate).
inline void schur_solve(const SkelFactorization& sf, const QuadTree& tree,
                        const std::vector<double>& domain_points, const ie_Mat& f,
                        ie_Mat* solution) {
  ie_Mat mu, alpha;
  const bool holes = tree.U.width() > 0;
  linear_solve(sf, tree, f, &mu, holes ? &alpha : nullptr);
  const int nt = (int)domain_points.size() / 2;
  std::vector<int> inside(nt);
  for (int t = 0; t < nt; t++)
    inside[t] = tree.boundary->is_in_domain(Vec2(domain_points[2 * t], domain_points[2 * t + 1]));
  *solution = ie_Mat(nt * tree.solution_dimension, 1);
  detail::check(tree.ctx, iegpu_evaluate(tree.ctx, nt, domain_points.data(), inside.data(),
                                         mu.mat, holes ? alpha.mat : nullptr, solution->mat));
}

// linear_solve.cpp:156-185: ten factor+solve runs, averages in seconds.
inline void bie_time_trial(const ie_solver_config& config, QuadTree* tree,
                           double* avg_skel_time, double* avg_solve_time) {
  SkelFactorization sf(config.id_tol, config.is_strong_admissibility, DEFAULT_NUM_THREADS);
  Kernel kernel(config.solution_dimension, config.domain_dimension, config.pde,
                tree->boundary, std::vector<double>());
  const ie_Mat f = tree->boundary->boundary_values;
  ie_Mat mu(f.height(), 1);
  double ts = 0, tv = 0;
  for (int i = 0; i < 10; i++) {
    auto t0 = std::chrono::steady_clock::now();
    sf.skeletonize(kernel, tree);
    auto t1 = std::chrono::steady_clock::now();
    sf.solve(*tree, &mu, f);
    auto t2 = std::chrono::steady_clock::now();
    ts += std::chrono::duration<double>(t1 - t0).count();
    tv += std::chrono::duration<double>(t2 - t1).count();
  }
  *avg_skel_time = ts / 10.0;
  *avg_solve_time = tv / 10.0;
}

// linear_solve.cpp:188-266
inline ie_Mat boundary_integral_solve(const ie_solver_config& config, QuadTree* tree,
                                      const std::vector<double>& domain_points) {
  Boundary* b = tree->boundary;
  SkelFactorization sf(config.id_tol, config.is_strong_admissibility, config.num_threads);
  Kernel kernel(config.solution_dimension, config.domain_dimension, config.pde, b,
                domain_points);
  tree->U = initialize_U_mat(config.pde, b->holes, b->points);
  tree->Psi = initialize_Psi_mat(config.pde, b->holes, b);
  sf.skeletonize(kernel, tree);
  ie_Mat solution;
  schur_solve(sf, *tree, domain_points, b->boundary_values, &solution);
  return solution;
}

// linear_solve.cpp:269-279
inline void get_domain_points(int domain_size, std::vector<double>* points, double x_min,
                              double x_max, double y_min, double y_max) {
  for (int i = 0; i < domain_size; i++) {
    const double x = x_min + ((i + 0.0) / (domain_size - 1)) * (x_max - x_min);
    for (int j = 0; j < domain_size; j++) {
      const double y = y_min + ((j + 0.0) / (domain_size - 1)) * (y_max - y_min);
      points->push_back(x);
      points->push_back(y);
    }
  }
}

}  // namespace ie_solver

#endif  // IE_SOLVER_GPU_HPP_
