// C++ facade over the C ABI (include/itp2d.h) with the class names and
// error behaviour of the reference's hot-path API, so the ITP loop reads as it
// does in the reference:
//   DataLayout + Transformer  datalayout.hpp:30-54, transformer.hpp:30-63
//   Potential                 potential.hpp:31-58
//   StateSet::orthonormalize  stateset.hpp:73 (throws NonPositiveEigenvalue,
//                             NonNormalEigenvalue, EigensolverError)
//   MultiProductSplit         multiproductsplit.hpp (EvolutionOperator,
//                             set_time_step, operators.hpp:65-69)
//   Hamiltonian               H = Kinetic + Potential (itpsystem.cpp:117-119),
//                             mean_and_standard_deviation (operators.cpp:40-50)
//   EigenSolver               eigensolver.hpp:57-84
// Operators act on a whole StateSet at once (the batched form of
// Operator::operate over the OpenMP loop at itpsystem.cpp:247-254).
#pragma once

#include <complex>
#include <cstdint>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/itp2d.h"

namespace itp2d {

typedef std::complex<double> comp;
enum BoundaryType { Periodic = ITP_PERIODIC, Dirichlet = ITP_DIRICHLET };
enum Precision { FP64 = ITP_FP64, FP32 = ITP_FP32 };

// exceptions.hpp:55-86, 237-244
class GeneralError : public std::runtime_error {
 public:
  explicit GeneralError(const std::string& s) : std::runtime_error(s) {}
};
class NonPositiveEigenvalue : public std::runtime_error {
 public:
  NonPositiveEigenvalue(size_t n, double val)
      : std::runtime_error(msg(n, val)), index(n), value(val) {}
  size_t index;
  double value;
 private:
  static std::string msg(size_t n, double v) {
    std::ostringstream o;
    o << "Non-positive eigenvalue " << v << " of the overlap matrix at index " << n;
    return o.str();
  }
};
class NonNormalEigenvalue : public std::runtime_error {
 public:
  NonNormalEigenvalue(size_t n, double val) : std::runtime_error("Non-normal eigenvalue of the overlap matrix"), index(n), value(val) {}
  size_t index;
  double value;
};
class EigensolverError : public std::runtime_error {
 public:
  explicit EigensolverError(const std::string& s) : std::runtime_error("Eigensolver failed: " + s) {}
};

class Context {
 public:
  struct Borrowed {};
  explicit Context(int device = 0) {
    if (itp_ctx_create(device, &c_) != ITP_OK) throw GeneralError("itp_ctx_create failed");
  }
  // wraps a context owned elsewhere
  Context(itp_ctx* existing, Borrowed) : c_(existing), owned_(false) {}
  ~Context() { if (owned_) itp_ctx_destroy(c_); }
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;
  itp_ctx* get() const { return c_; }
  void check(int rc, const char* what) const {
    if (rc == ITP_OK) return;
    std::string m = std::string(what) + ": " + itp_last_error(c_);
    if (rc == ITP_E_EIGENSOLVER) throw EigensolverError(m);
    throw GeneralError(m);
  }
 private:
  itp_ctx* c_ = nullptr;
  bool owned_ = true;
};

// DataLayout and the Transformer's tables in one object.
class Transformer {
 public:
  Transformer(Context& ctx, size_t sx, size_t sy, double dx, BoundaryType bt)
      : sizex(sx), sizey(sy), dx(dx), boundary(bt), ctx_(ctx) {
    ctx.check(itp_grid_create(ctx.get(), sx, sy, dx, bt, &g_), "itp_grid_create");
  }
  ~Transformer() { itp_grid_destroy(g_); }
  Transformer(const Transformer&) = delete;
  Transformer& operator=(const Transformer&) = delete;
  // Transformer::transform(comp*, Transform) on host fields, nfields at once
  void transform(comp* data, int kind, size_t nfields = 1) const {
    ctx_.check(itp_transform_host(g_, reinterpret_cast<double*>(data), nfields, kind), "transform");
  }
  double normalization_factor(int kind) const { return itp_grid_normalization_factor(g_, kind); }
  itp_grid* get() const { return g_; }
  Context& context() const { return ctx_; }
  const size_t sizex, sizey;
  const double dx;
  const BoundaryType boundary;
 private:
  Context& ctx_;
  itp_grid* g_ = nullptr;
};

class Potential {
 public:
  Potential(Transformer& tr, const double* values) : tr_(tr), null_(values == nullptr) {
    tr.context().check(itp_potential_create(tr.get(), values, &p_), "itp_potential_create");
  }
  ~Potential() { itp_potential_destroy(p_); }
  Potential(const Potential&) = delete;
  Potential& operator=(const Potential&) = delete;
  bool is_null() const { return null_; }
  itp_potential* get() const { return p_; }
  Transformer& transformer() const { return tr_; }
 private:
  Transformer& tr_;
  bool null_;
  itp_potential* p_ = nullptr;
};

class StateSet {
 public:
  StateSet(Transformer& tr, size_t N, Precision prec = FP64) : N(N), precision(prec), tr_(tr) {
    tr.context().check(itp_stateset_create(tr.get(), N, prec, &s_), "itp_stateset_create");
  }
  ~StateSet() { itp_stateset_destroy(s_); }
  StateSet(const StateSet&) = delete;
  StateSet& operator=(const StateSet&) = delete;
  void upload(const comp* data) { tr_.context().check(itp_stateset_upload(s_, reinterpret_cast<const double*>(data)), "upload"); }
  void download(comp* data) const { tr_.context().check(itp_stateset_download(s_, reinterpret_cast<double*>(data)), "download"); }
  // StateSet::orthonormalize (stateset.cpp:136-232)
  void orthonormalize() {
    size_t n = 0;
    double v = 0;
    const int rc = itp_orthonormalize(s_, &n, &v);
    if (rc == ITP_E_NONPOSITIVE_EIG) throw NonPositiveEigenvalue(n, v);
    if (rc == ITP_E_NONNORMAL_EIG) throw NonNormalEigenvalue(n, v);
    tr_.context().check(rc, "orthonormalize");
  }
  // Cholesky orthonormalization (the mixed path's CholeskyQR step); returns
  // true when it fell back to the eigendecomposition route
  bool orthonormalize_cholesky() {
    size_t n = 0;
    double v = 0;
    int used = 0;
    const int rc = itp_orthonormalize_cholesky(s_, &used, &n, &v);
    if (rc == ITP_E_NONPOSITIVE_EIG) throw NonPositiveEigenvalue(n, v);
    if (rc == ITP_E_NONNORMAL_EIG) throw NonNormalEigenvalue(n, v);
    tr_.context().check(rc, "orthonormalize_cholesky");
    return used != 0;
  }
  // max R_ii^2 / min R_ii^2 of the last Cholesky factor (0: none)
  double cholesky_condition() const {
    double k = 0;
    tr_.context().check(itp_stateset_cholesky_condition(s_, &k), "cholesky_condition");
    return k;
  }
  // R_ii^2 of the last Cholesky factor; false if it fell back
  bool cholesky_diagonal(std::vector<double>& d) const {
    d.resize(N);
    return itp_stateset_cholesky_diagonal(s_, d.data()) == ITP_OK;
  }
  // Ascending overlap eigenvalues of the last orthonormalization.
  void overlap_eigenvalues(std::vector<double>& lam) const {
    lam.resize(N);
    tr_.context().check(itp_stateset_overlap_eigenvalues(s_, lam.data()), "overlap_eigenvalues");
  }
  // Fields of other (same grid and size), converted to this set's precision.
  void init_noise(uint64_t& rng, double dx) {
    tr_.context().check(itp_stateset_init_noise(s_, &rng, dx), "init_noise");
  }
  void assign(const StateSet& other) { tr_.context().check(itp_stateset_convert(s_, other.get()), "convert"); }
  itp_stateset* get() const { return s_; }
  Transformer& transformer() const { return tr_; }
  const size_t N;
  const Precision precision;
 private:
  Transformer& tr_;
  itp_stateset* s_ = nullptr;
};

// EvolutionOperator interface (operators.hpp:65-69) for the batched loop.
class EvolutionOperator {
 public:
  virtual ~EvolutionOperator() {}
  virtual void operate(StateSet& states) = 0;
  virtual void set_time_step(double e) = 0;
};

class MultiProductSplit : public EvolutionOperator {
 public:
  MultiProductSplit(int halforder, Potential& pot, double time_step, double B) : pot_(pot) {
    pot.transformer().context().check(itp_split_create(pot.get(), halforder, time_step, B, &sp_), "itp_split_create");
    this->halforder = itp_split_halforder(sp_);
  }
  ~MultiProductSplit() { itp_split_destroy(sp_); }
  MultiProductSplit(const MultiProductSplit&) = delete;
  MultiProductSplit& operator=(const MultiProductSplit&) = delete;
  void operate(StateSet& states) override {
    pot_.transformer().context().check(itp_propagate(sp_, states.get()), "propagate");
  }
  void set_time_step(double e) override {
    pot_.transformer().context().check(itp_split_set_time_step(sp_, e), "set_time_step");
  }
  int halforder = 1;
 private:
  Potential& pot_;
  itp_split* sp_ = nullptr;
};

class Hamiltonian {
 public:
  Hamiltonian(Potential& pot, double B) : pot_(pot) {
    pot.transformer().context().check(itp_hamiltonian_create(pot.get(), B, &h_), "itp_hamiltonian_create");
  }
  ~Hamiltonian() { itp_hamiltonian_destroy(h_); }
  Hamiltonian(const Hamiltonian&) = delete;
  Hamiltonian& operator=(const Hamiltonian&) = delete;
  // FP64 Rayleigh quotients of src's states, evaluated on a copy in dst (FP64)
  void rayleigh_quotients(const StateSet& src, StateSet& dst, std::vector<double>& E) {
    E.resize(src.N);
    pot_.transformer().context().check(itp_rayleigh_quotients(h_, src.get(), dst.get(), E.data()),
                                       "rayleigh_quotients");
  }
  // real parts of mean and standard deviation for every state, by slot
  void mean_and_standard_deviation(StateSet& s, std::vector<double>& E, std::vector<double>& sigma) {
    E.resize(s.N);
    sigma.resize(s.N);
    pot_.transformer().context().check(itp_energies(h_, s.get(), E.data(), sigma.data()), "energies");
  }
  // Rayleigh-Ritz on the set: slot n becomes the Ritz vector of E[n]
  // (ascending), sigma[n] its residual norm.
  void rayleigh_ritz(StateSet& s, std::vector<double>& E, std::vector<double>& sigma) {
    E.resize(s.N);
    sigma.resize(s.N);
    pot_.transformer().context().check(itp_rayleigh_ritz(h_, s.get(), E.data(), sigma.data()), "rayleigh_ritz");
  }
 private:
  Potential& pot_;
  itp_hamiltonian* h_ = nullptr;
};

// EigenSolver (eigensolver.hpp:57-84): column-major input, destroyed;
// eigenvectors in columns, eigenvalues ascending.
class EigenSolver {
 public:
  EigenSolver(Context& ctx, size_t N) : ctx_(ctx), size(N), evals(N) {}
  void solve(comp* input_matrix) {
    ctx_.check(itp_eigh_host(ctx_.get(), size, reinterpret_cast<double*>(input_matrix), evals.data()), "eigh");
  }
  double eigenvalue(size_t n) const { return evals[n]; }
  const comp& eigenvector(const comp* m, size_t n, size_t i) const { return m[n * size + i]; }
  void scale_eigenvector(comp* m, size_t n, double v) const {
    for (size_t i = 0; i < size; i++) m[n * size + i] *= v;
  }
 private:
  Context& ctx_;
  size_t size;
  std::vector<double> evals;
};

}  // namespace itp2d
