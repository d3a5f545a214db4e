// CPU emulation of the gfx950 sweep kernels, for tests only.
//
// Runs the product's own line-operation passes (itp2d_amd/csrc/line_ops.hpp)
// and sweep sequencing (itp2d_amd/csrc/engine.hpp) on the host, one emulated
// workgroup at a time, passes in order, threads 0..255 in turn. This checks
// the kernels' indexing and algebra against the oracle without a GPU.
#include <vector>
#include <memory>
#include <omp.h>
#include "../../itp2d_amd/csrc/engine.hpp"
#include "../../itp2d_amd/csrc/block_jacobi.hpp"
#include "../../itp2d_amd/csrc/cluster_eig.hpp"
#include <algorithm>

using namespace itp;

namespace {
struct CpuExec {
  int nthr = 256;
  template <class F> void run(F&& f) { for (int t = 0; t < nthr; t++) f(t, nthr); }
};

// Register engine: per-thread register state kept per emulated thread.
template <class S>
struct CpuRegExec {
  std::vector<S> st;
  CpuRegExec() : st(kRegThreads) {}
  template <class F> void run(F&& f) { for (int t = 0; t < kRegThreads; t++) f(t, st[t]); }
  template <class F> void local(F&& f) { run(f); }
};

bool g_use_reg = true;

struct CpuBackend {
  template <class R>
  void sweep(const SweepArgs<R>& a, int nblk, int nfields) {
    if (a.reg) {
      const int n = a.plan.n;
      const int lines = a.orient == 0 ? a.sy : a.sx;
      if (a.L != reg_lines_per_wg(n) || nblk * a.L != lines || (a.orient == 0 ? a.sx : a.sy) != n)
        throw std::runtime_error("register sweep shape");
      const bool ok = reg_dispatch(n, a.orient, ops_have_xsplit(a.ops, a.nops), [&](auto nc, auto oc, auto xc) {
        constexpr int N = decltype(nc)::value, O = decltype(oc)::value;
        constexpr bool X = decltype(xc)::value;
        const size_t bytes = reg_lds_bytes<R>(N);
#pragma omp parallel
        {
          std::vector<cplx<R>> lds(bytes / sizeof(cplx<R>) + 1);
#pragma omp for collapse(2) schedule(static)
          for (int f = 0; f < nfields; f++)
            for (int b = 0; b < nblk; b++) {
              CpuRegExec<RegState<R, RegShape<N>::Q>> ex;
              reg_sweep_program<R, N, O, X>(ex, a, lds.data(), f, b);
            }
        }
      });
      if (!ok) throw std::runtime_error("register sweep length");
      return;
    }
    const size_t bytes = sweep_lds_bytes<R>(a.L, a.plan.n, 256);
#pragma omp parallel
    {
      std::vector<char> lds(bytes);
#pragma omp for collapse(2) schedule(static)
      for (int f = 0; f < nfields; f++)
        for (int b = 0; b < nblk; b++) {
          SweepLds<R> s = carve_lds<R>(lds.data(), a.L, a.plan.n, 256);
          CpuExec ex;
          sweep_program<R>(ex, a, s, f, b);
        }
    }
  }
};

struct Env {
  GridInfo g;
  LineGeom lg;
  std::vector<cplx<double>> twx, mkx, twy, mky;
  Env(int sx, int sy, double dx, int bt) {
    g = make_grid(sx, sy, dx, bt);
    if (!make_plan(sx, lg.plan_x) || !make_plan(sy, lg.plan_y)) throw std::runtime_error("plan");
    lg.use_reg = g_use_reg;
    twiddle_tables(sx, twx, mkx);
    twiddle_tables(sy, twy, mky);
    lg.tw_x = twx.data(); lg.mk_x = mkx.data(); lg.tw_y = twy.data(); lg.mk_y = mky.data();
  }
};

template <class F> int guard(F f) { try { return f(); } catch (const std::exception&) { return 1; } }

// FP64 or FP32 run of a field program: convert in/out for FP32.
template <class R>
std::vector<cplx<R>> to_r(const double* d, size_t n) {
  std::vector<cplx<R>> v(n);
  for (size_t i = 0; i < n; i++) v[i] = mk<R>(R(d[2 * i]), R(d[2 * i + 1]));
  return v;
}
template <class R>
void from_r(const std::vector<cplx<R>>& v, double* d) {
  for (size_t i = 0; i < v.size(); i++) { d[2 * i] = double(v[i].x); d[2 * i + 1] = double(v[i].y); }
}
}  // namespace

extern "C" {

// 1: sweeps use the register engine where the line length allows it (default);
// 0: always the LDS-resident engine.
void emu_set_reg(int on) { g_use_reg = on != 0; }

int emu_plan(int n, int* radix, int* nst) {
  LinePlan p;
  if (!make_plan(n, p)) return 1;
  *nst = p.nst;
  for (int i = 0; i < p.nst; i++) radix[i] = p.radix[i];
  return 0;
}

int emu_transform(int sx, int sy, double dx, int bt, double* data, int nfields, int kind, int fp32) {
  return guard([&] {
    Env e(sx, sy, dx, bt);
    CpuBackend be;
    auto go = [&](auto tag) {
      using R = decltype(tag);
      auto f = to_r<R>(data, (size_t)nfields * e.g.M);
      std::vector<Sweep<R>> prog;
      if (!transform_program<R>(kind, f.data(), prog)) throw std::runtime_error("kind");
      run_program<R>(be, e.g, e.lg, prog, nfields);
      from_r<R>(f, data);
    };
    if (fp32) go(float()); else go(double());
    return 0;
  });
}

int emu_expkinetic(int sx, int sy, double dx, int bt, double B, double eps, double coeff, double pref,
                   double* data, int nfields) {
  return guard([&] {
    Env e(sx, sy, dx, bt);
    CpuBackend be;
    ExpKinTabs t;
    expkin_tables(e.g, B, eps, coeff, pref, t);
    KinDev k{t.mult.data(), t.xm.data(), t.xm2.data(), t.ym.data()};
    auto f = to_r<double>(data, (size_t)nfields * e.g.M);
    std::vector<cplx<double>> w(f.size());
    std::vector<Sweep<double>> prog;
    expkinetic_program<double>(branch_of(B, bt), k, f.data(), w.data(), prog);
    run_program<double>(be, e.g, e.lg, prog, nfields);
    from_r<double>(f, data);
    return 0;
  });
}

// ExpPotential alone (exppotential.hpp:56-65): one row sweep of elementwise ops.
int emu_exppotential(int sx, int sy, double dx, int bt, const double* V, double eps, double coeff, double pref,
                     double* data, int nfields) {
  return guard([&] {
    Env e(sx, sy, dx, bt);
    CpuBackend be;
    std::vector<double> vals;
    if (V) exppot_values(e.g, V, eps, coeff, pref, vals);
    auto f = to_r<double>(data, (size_t)nfields * e.g.M);
    std::vector<Sweep<double>> prog;
    exppotential_program<double>(V ? vals.data() : nullptr, pref, f.data(), prog);
    run_program<double>(be, e.g, e.lg, prog, nfields);
    from_r<double>(f, data);
    return 0;
  });
}

// One SecondOrderSplit member (secondordersplit.cpp:21-49), or the whole
// MultiProductSplit when halforder > 0 (exponent ignored).
int emu_split(int sx, int sy, double dx, int bt, double B, const double* V, double eps, double prefactor,
              int exponent, int halforder, double* data, int nfields, int fp32) {
  return guard([&] {
    Env e(sx, sy, dx, bt);
    CpuBackend be;
    const Branch br = branch_of(B, bt);
    const bool has_v = V != nullptr;
    struct MemberHost { ExpKinTabs k; std::vector<double> vh, vf, vp; MemberDev d; };
    std::vector<std::unique_ptr<MemberHost>> mem;
    std::vector<double> coefs;
    std::vector<int> exps;
    std::vector<double> steps;
    if (halforder > 0) {
      const int h = has_v ? halforder : 1;
      coefs = mps_coefficients(h);
      for (int t = 1; t <= h; t++) { exps.push_back(t); steps.push_back(eps / t); }
    } else {
      coefs.push_back(prefactor); exps.push_back(exponent); steps.push_back(eps);
    }
    for (size_t i = 0; i < exps.size(); i++) {
      std::unique_ptr<MemberHost> m(new MemberHost);
      const double h = steps[i];
      if (has_v) {
        expkin_tables(e.g, B, h, -1.0, 1.0, m->k);
        exppot_values(e.g, V, h, -0.5, 1.0, m->vh);
        exppot_values(e.g, V, h, -1.0, 1.0, m->vf);
        exppot_values(e.g, V, h, -0.5, coefs[i], m->vp);
      } else {
        expkin_tables(e.g, B, h, -1.0 * exps[i], coefs[i], m->k);
      }
      m->d.kin = KinDev{m->k.mult.data(), m->k.xm.data(), m->k.xm2.data(), m->k.ym.data()};
      m->d.vhalf = m->vh.data(); m->d.vfull = m->vf.data(); m->d.vpref = m->vp.data();
      mem.push_back(std::move(m));
    }
    auto go = [&](auto tag) {
      using R = decltype(tag);
      auto in = to_r<R>(data, (size_t)nfields * e.g.M);
      std::vector<cplx<R>> w(in.size()), out(in.size());
      std::vector<Sweep<R>> prog;
      for (size_t i = 0; i < mem.size(); i++)
        member_program<R>(br, mem[i]->d, has_v ? exps[i] : 1, has_v, in.data(), w.data(), out.data(), i > 0, prog);
      run_program<R>(be, e.g, e.lg, prog, nfields);
      from_r<R>(out, data);
    };
    if (fp32) go(float()); else go(double());
    return 0;
  });
}

// H applied once (out) and the energies of the fields (E, sigma by field).
int emu_hamiltonian(int sx, int sy, double dx, int bt, double B, const double* V, double* data, int nfields,
                    double* E, double* sigma, int fp32) {
  return guard([&] {
    Env e(sx, sy, dx, bt);
    CpuBackend be;
    KinTabs t;
    kin_tables(e.g, B, t);
    KinHDev k{t.mxy.data(), t.mx.data(), t.mx2.data(), t.my.data()};
    const Branch br = branch_of(B, bt);
    auto go = [&](auto tag) {
      using R = decltype(tag);
      auto psi = to_r<R>(data, (size_t)nfields * e.g.M);
      std::vector<cplx<R>> w1(psi.size()), h1(psi.size());
      const int nbx = e.g.sy, nby = e.g.sx;
      std::vector<cplx<double>> p1((size_t)nfields * std::max(nbx, nby)), p2(p1.size());
      std::vector<Sweep<R>> prog;
      const int nb = hamiltonian_program<R>(e.g, e.lg, br, k, V, psi.data(), w1.data(), h1.data(), psi.data(),
                                            p1.data(), prog);
      run_program<R>(be, e.g, e.lg, prog, nfields);
      prog.clear();
      hamiltonian_program<R>(e.g, e.lg, br, k, V, h1.data(), w1.data(), nullptr, psi.data(), p2.data(), prog);
      run_program<R>(be, e.g, e.lg, prog, nfields);
      for (int n = 0; n < nfields; n++)
        energy_from_partials(p1.data() + (size_t)n * nb, p2.data() + (size_t)n * nb, nb, nb, e.g.dx, E + n,
                             sigma + n, nullptr, nullptr);
      from_r<R>(h1, data);
    };
    if (fp32) go(float()); else go(double());
    return 0;
  });
}

// Block Jacobi eigensolver phases (block_jacobi.hpp) run in kernel order on
// the CPU, one emulated workgroup at a time (32 workgroups).
// A: n x n complex column-major (upper triangle read). Outputs ascending
// eigenvalues w and eigenvectors (columns) in A. Returns sweeps done, or -1.
int emu_block_jacobi_dbg(int n, double* Ain, double* w, int max_inner, int max_sweeps, double* Aout, double* Vout);
int emu_block_jacobi(int n, double* Ain, double* w, int max_inner) { return emu_block_jacobi_dbg(n, Ain, w, max_inner, 40, nullptr, nullptr); }
int emu_block_jacobi_dbg(int n, double* Ain, double* w, int max_inner, int max_sweeps, double* Aout, double* Vout) {
  const int np = (n + 63) / 64 * 64;
  std::vector<cplx<double>> A((size_t)np * np, mk<double>(0.0, 0.0)), V(A.size(), mk<double>(0.0, 0.0));
  const cplx<double>* a = reinterpret_cast<const cplx<double>*>(Ain);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) A[(size_t)i * np + j] = (i <= j) ? a[(size_t)j * n + i] : conj(a[(size_t)i * n + j]);
  for (int i = 0; i < n; i++) A[(size_t)i * np + i].y = 0;
  for (int i = 0; i < np; i++) V[(size_t)i * np + i] = mk<double>(1.0, 0.0);
  const int P = np / JS, nb = np / JB, G = 32;
  std::vector<cplx<double>> Q((size_t)P * JS2), D(Q.size());
  std::vector<int> ident(P, 1);
  BlockJacobiArgs args;
  args.A = A.data(); args.V = V.data(); args.Q = Q.data(); args.D = D.data(); args.ident = ident.data();
  args.slots = nullptr; args.info = nullptr;
  args.np = np; args.max_sweeps = max_sweeps; args.max_inner = max_inner; args.tol = 1e-14; args.handoff = 0;
  // one LDS image per emulated workgroup (each keeps its own position table)
  std::vector<std::vector<char>> lds(G, std::vector<char>(bj_lds_bytes(256)));
  std::vector<BJLds> ls;
  for (int g = 0; g < G; g++) ls.push_back(bj_carve(lds[g].data(), 256));
  CpuExec ex;
  const long long tot = (long long)np * np, per = (tot + G - 1) / G;
  int sweeps = 0;
  bool conv = false;
  for (int chk = 0;; chk++) {
    double worst = 0;
    for (int g = 0; g < G; g++) {
      const long long e0 = g * per, e1 = std::min(tot, e0 + per);
      bj_check_rows(ex, args, e0, e1, ls[g]);
      worst = std::max(worst, ls[g].red[256]);
    }
    if (worst < 1.0) { conv = true; break; }
    if (chk == args.max_sweeps) break;
    for (int g = 0; g < G; g++) bj_reset_positions(ex, ls[g], nb);
    for (int r = 0; r < nb - 1; r++) {
      for (int g = 0; g < G; g++)
        for (int k = g; k < P; k += G) bj_inner(ex, args, k, ls[g].pos2blk[2 * k], ls[g].pos2blk[2 * k + 1], ls[g]);
      for (int g = 0; g < G; g++)
        for (int t = g; t < bj_items(np); t += G) bj_apply(ex, args, t, ls[g]);
      for (int g = 0; g < G; g++) bj_next_positions(ex, ls[g], nb);
    }
    sweeps++;
  }
  if (Aout) std::copy((double*)A.data(), (double*)A.data() + 2 * A.size(), Aout);
  if (Vout) std::copy((double*)V.data(), (double*)V.data() + 2 * V.size(), Vout);
  if (!conv) return -1;
  std::vector<int> idx(n);
  for (int i = 0; i < n; i++) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return A[(size_t)x * np + x].x < A[(size_t)y * np + y].x; });
  cplx<double>* out = reinterpret_cast<cplx<double>*>(Ain);
  for (int r = 0; r < n; r++) {
    w[r] = A[(size_t)idx[r] * np + idx[r]].x;
    for (int c = 0; c < n; c++) out[(size_t)r * n + c] = V[(size_t)c * np + idx[r]];
  }
  return sweeps;
}

// Cluster solve (cluster_eig.hpp) of the principal sub-matrix of the n x n
// Hermitian A (row-major, complex) on idx[0..k): X[:, idx] <- X[:, idx] U
// with U the cluster's eigenvectors (X row-major, n x n).
int emu_cluster_solve(int n, const double* A, const int* idx, int k, double* X) {
  if (k < 1 || k > kMaxCluster) return -1;
  std::vector<char> lds(cl_lds_bytes(256));
  CpuExec ex;
  cl_solve(ex, reinterpret_cast<const cplx<double>*>(A), n, idx, k, reinterpret_cast<cplx<double>*>(X), 1e-15, 30,
           lds.data(), n);
  return 0;
}

// Strong pairs of rows 0..N) as the device kernel lists them (one emulated
// workgroup per row); returns the count and fills pairs (up to cap).
int emu_cluster_pairs(int np, int N, const double* A, double tau, int cap, int* pairs, double* maxrho) {
  int cnt = 0;
  double mx = 0;
  for (int i = 0; i < N; i++)
    for (int t = 0; t < 256; t++) {
      const double m = cl_row_pairs(reinterpret_cast<const cplx<double>*>(A), np, N, i, t, 256, tau, [&](int a, int b) {
        if (cnt < cap) { pairs[2 * cnt] = a; pairs[2 * cnt + 1] = b; }
        cnt++;
      });
      mx = std::max(mx, m);
    }
  *maxrho = mx;
  return cnt;
}

}  // extern "C"
