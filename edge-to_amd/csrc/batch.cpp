// Batched SIMP engine: many independent structured problems optimised
// concurrently on one GPU.
//
// Every problem keeps its whole state resident (design, filtered design,
// densities, displacement, sensitivities, filter tables, OC workspace and its
// own GMG-PCG solver with its hierarchy and coarse factor) and owns a stream.
// A pool of host threads takes problems from a work queue (largest first);
// each thread drives its problem's OC iterations on that problem's stream, so
// the kernels of different problems run side by side on the GPU and every
// problem's host-side round trips (the PCG convergence ring, the coarse-factor
// check) overlap with the other problems' work.
//
// Per OC iteration a problem issues one captured graph and one solve.  The
// graph holds everything between two solves: the fused SIMP adjoint with the
// filter transpose, the OC update (device bisection, no host round trip), the
// design change, the filter and the SIMP map of the new design.  It depends on
// the two penalties involved (the previous solve's and the next one's), so a
// problem keeps one graph per pair that occurs -- one without a penalty
// continuation, one per step of it otherwise -- captured on first use and
// replayed.  The solve itself (hierarchy rebuild and PCG with its
// data-dependent iteration count) and the compliance f.u after it are issued
// directly; the compliance feeds the stopping test.
//
// The sequence is CuTopOpt.optimize (TopOpt/_cuda.py:304-382) per problem:
// x = clip(x0, void, 1); initial solve; then per iteration update_design,
// solve_and_process.  Results are those of the single-problem path, kernel for
// kernel.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

#include "blas.hpp"
#include "common.hpp"
#include "solver.hpp"
#include "stiffness.hpp"

namespace et {

struct BatchProblem {
  et_batch_problem d{};
  Grid g{};
  int dof = 0, k0_flags = 0;
  int64_t n_dof = 0, n_el = 0;
  std::vector<double> K0h, fh, x0h, wh, sched;
  std::vector<uint8_t> consh;
  std::vector<int32_t> offh;
  hipStream_t stream = nullptr;
  MGSolver mg;
  bool mg_ok = false;
  double *K0 = nullptr, *f = nullptr, *U = nullptr, *x = nullptr, *xn = nullptr, *xt = nullptr, *E = nullptr,
         *dc = nullptr, *w = nullptr, *norm = nullptr, *scal = nullptr, *partial = nullptr;
  uint8_t* cons = nullptr;
  int32_t* off = nullptr;
  void* ocw = nullptr;
  // captured between-solve graphs, keyed by (adjoint penalty, forward penalty)
  struct Captured {
    double pa, pf;
    hipGraph_t graph;
    hipGraphExec_t exec;
  };
  std::vector<Captured> graphs;
  // results
  std::vector<double> comp, iter_ms, change;
  std::vector<int> pcg;
  int done = 0;
  bool converged = false;
  int status = ET_OK;
  std::string error;
};

struct BatchEngine {
  std::vector<BatchProblem> P;
  int precision = 1;
  double run_ms = 0.0;
  ~BatchEngine() { release(); }
  void release();
};

void BatchEngine::release() {
  for (auto& p : P) {
    for (auto& c : p.graphs) {
      if (c.exec) (void)hipGraphExecDestroy(c.exec);
      if (c.graph) (void)hipGraphDestroy(c.graph);
    }
    p.graphs.clear();
    if (p.mg_ok) p.mg.destroy();
    p.mg_ok = false;
    for (double** v : {&p.K0, &p.f, &p.U, &p.x, &p.xn, &p.xt, &p.E, &p.dc, &p.w, &p.norm, &p.scal, &p.partial})
      if (*v) (void)hipFree(*v), *v = nullptr;
    if (p.cons) (void)hipFree(p.cons), p.cons = nullptr;
    if (p.off) (void)hipFree(p.off), p.off = nullptr;
    if (p.ocw) (void)hipFree(p.ocw), p.ocw = nullptr;
    if (p.stream) (void)hipStreamDestroy(p.stream), p.stream = nullptr;
  }
}

namespace {

__global__ void copy_kernel(int64_t n, const double* __restrict__ a, double* __restrict__ b) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = a[t];
}

int setup_problem(BatchProblem& p, int precision) {
  const et_batch_problem& d = p.d;
  ET_TRY(make_grid(d.dim, d.nel, &p.g));
  p.dof = d.dim;
  p.n_dof = p.g.n_nodes() * p.dof;
  p.n_el = p.g.n_elem();
  const int nd = (d.dim == 3 ? 8 : 4) * d.dim;
  ET_REQUIRE(d.K0 && d.cons && d.f && d.x0 && d.offsets && d.weights, "et_batch_create: null problem array");
  ET_REQUIRE(d.n_off > 0 && d.n_off <= 1024, "et_batch_create: filter offset count out of range (1..1024)");
  ET_REQUIRE(d.A > 0 && d.volfrac > 0 && d.tol > 0 && d.maxiter >= 1 && d.n_level >= 1,
             "et_batch_create: bad problem parameters");
  p.K0h.assign(d.K0, d.K0 + nd * nd);
  p.k0_flags = k0_structure(d.dim, p.K0h.data(), kK0Rtol);
  p.consh.assign(d.cons, d.cons + p.n_dof);
  p.fh.assign(d.f, d.f + p.n_dof);
  p.x0h.assign(d.x0, d.x0 + p.n_el);
  p.offh.assign(d.offsets, d.offsets + (size_t)d.n_off * d.dim);
  p.wh.assign(d.weights, d.weights + d.n_off);
  ET_REQUIRE(d.n_schedule >= 0 && (d.n_schedule == 0 || d.penal_schedule != nullptr),
             "et_batch_create: penalty schedule without values");
  p.sched.assign(d.penal_schedule, d.penal_schedule + d.n_schedule);
  for (const double v : p.sched) ET_REQUIRE(v > 0.0 && std::isfinite(v), "et_batch_create: bad penalty in the schedule");
  ET_CHECK_HIP(hipStreamCreateWithFlags(&p.stream, hipStreamNonBlocking));
  const size_t vb = p.n_dof * sizeof(double), eb = p.n_el * sizeof(double);
  ET_CHECK_HIP(hipMalloc(&p.K0, nd * nd * sizeof(double)));
  ET_CHECK_HIP(hipMalloc(&p.f, vb));
  ET_CHECK_HIP(hipMalloc(&p.U, vb));
  for (double** v : {&p.x, &p.xn, &p.xt, &p.E, &p.dc, &p.norm}) ET_CHECK_HIP(hipMalloc(v, eb));
  ET_CHECK_HIP(hipMalloc(&p.w, d.n_off * sizeof(double)));
  ET_CHECK_HIP(hipMalloc(&p.scal, 8 * sizeof(double)));
  ET_CHECK_HIP(hipMalloc(&p.partial, kRedBlocks * sizeof(double)));
  ET_CHECK_HIP(hipMalloc(&p.cons, p.n_dof));
  ET_CHECK_HIP(hipMalloc(&p.off, p.offh.size() * sizeof(int32_t)));
  ET_CHECK_HIP(hipMalloc(&p.ocw, oc_work_size(p.n_el)));
  hipStream_t s = p.stream;
  ET_CHECK_HIP(hipMemcpyAsync(p.K0, p.K0h.data(), nd * nd * sizeof(double), hipMemcpyHostToDevice, s));
  ET_CHECK_HIP(hipMemcpyAsync(p.f, p.fh.data(), vb, hipMemcpyHostToDevice, s));
  ET_CHECK_HIP(hipMemcpyAsync(p.x, p.x0h.data(), eb, hipMemcpyHostToDevice, s));
  ET_CHECK_HIP(hipMemcpyAsync(p.cons, p.consh.data(), p.n_dof, hipMemcpyHostToDevice, s));
  ET_CHECK_HIP(hipMemcpyAsync(p.off, p.offh.data(), p.offh.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
  ET_CHECK_HIP(hipMemcpyAsync(p.w, p.wh.data(), d.n_off * sizeof(double), hipMemcpyHostToDevice, s));
  ET_TRY(filter_normalizer(p.g, d.n_off, p.off, p.w, p.norm, s));
  ET_CHECK_HIP(hipStreamSynchronize(s));
  et_mg_params mp{};
  mp.dim = d.dim;
  for (int a = 0; a < 3; ++a) mp.nel[a] = a < d.dim ? d.nel[a] : 0;
  mp.n_level = d.n_level;
  mp.n_smooth = d.n_smooth;
  mp.omega = 0.5;
  mp.min_omega = 0.4;
  mp.cycle_w = 0;
  mp.w_level_mask = 2;
  mp.precision = precision;
  const int st = p.mg.create(mp, p.K0h.data(), s);
  if (st != ET_OK) {
    p.mg.destroy();
    return st;
  }
  p.mg_ok = true;
  return ET_OK;
}

// penalty of the densities of OC iteration k (k = 0 also for the initial solve)
double penalty_at(const BatchProblem& p, int k) {
  if (p.sched.empty()) return p.d.penal;
  return p.sched[std::min<size_t>((size_t)k, p.sched.size() - 1)];
}

// the work between two solves, in stream order (captured into a graph): dc = H^T s
// with the previous solve's penalty pa, x_new = OC(x, dc), |x_new - x|^2, x = x_new,
// xt = H x, E = SIMP(xt) with the next solve's penalty pf
int between_solves(BatchProblem& p, double pa, double pf) {
  const et_batch_problem& d = p.d;
  hipStream_t s = p.stream;
  ET_TRY(simp_adjoint(p.g, p.K0, p.U, p.xt, pa, d.void_, d.n_off, p.off, p.w, p.norm, p.k0_flags, p.dc, s));
  ET_TRY(oc_update(p.n_el, p.x, p.dc, d.A, d.A * p.n_el, d.volfrac, d.move, p.xn, p.ocw, nullptr, s));
  ET_TRY(sqdist(p.n_el, p.xn, p.x, p.partial, p.scal + 1, s));
  copy_kernel<<<grid_for(p.n_el), kBlock, 0, s>>>(p.n_el, p.xn, p.x);
  ET_CHECK_LAUNCH();
  ET_TRY(filter_fwd(p.g, d.n_off, p.off, p.w, p.x, p.xt, s));
  ET_TRY(simp_forward(p.n_el, p.xt, pf, d.void_, 0, p.E, s));
  return ET_OK;
}

// the executable graph for the penalty pair (pa, pf), captured on first use
int graph_for(BatchProblem& p, double pa, double pf, hipGraphExec_t* out) {
  for (const auto& c : p.graphs)
    if (c.pa == pa && c.pf == pf) {
      *out = c.exec;
      return ET_OK;
    }
  hipStream_t s = p.stream;
  ET_CHECK_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const int st = between_solves(p, pa, pf);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(s, &g);
  if (st != ET_OK || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    ET_TRY(st);
    ET_CHECK_HIP(e);
  }
  hipGraphExec_t x = nullptr;
  const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    (void)hipGraphDestroy(g);
    ET_CHECK_HIP(ei);
  }
  p.graphs.push_back({pa, pf, g, x});
  *out = x;
  return ET_OK;
}

// compliance f.u of the last solve (synchronises the problem's stream)
int compliance(BatchProblem& p, double* c) {
  hipStream_t s = p.stream;
  ET_TRY(dot(p.n_dof, p.f, p.U, p.partial, p.scal, s));
  ET_CHECK_HIP(hipMemcpyAsync(c, p.scal, sizeof(double), hipMemcpyDeviceToHost, s));
  ET_CHECK_HIP(hipStreamSynchronize(s));
  return ET_OK;
}

int solve(BatchProblem& p) {
  int it = 0;
  double res = 0.0;
  const int st = p.mg.solve(p.E, p.cons, p.f, p.U, p.d.tol, p.d.maxiter, &it, &res);
  p.pcg.push_back(it);
  // maxiter above the tolerance is reported by the residual, as CuMultiGrid does
  return st == ET_ENOCONV ? ET_OK : st;
}

int run_problem(BatchProblem& p, int n_iter) {
  using clock = std::chrono::steady_clock;
  const et_batch_problem& d = p.d;
  hipStream_t s = p.stream;
  const int nt[2] = {d.tile_threads_f32, d.tile_threads_f64};
  const int ns[2] = {d.tile_slices_f32, d.tile_slices_f64};
  set_thread_wht_tiles(nt, ns);
  p.comp.clear();
  p.pcg.clear();
  p.iter_ms.clear();
  p.change.clear();
  p.done = 0;
  p.converged = false;
  ET_CHECK_HIP(hipMemcpyAsync(p.xn, p.x0h.data(), p.n_el * sizeof(double), hipMemcpyHostToDevice, s));
  // x = clip(x0, void, 1); xt = H x; E = SIMP(xt); U = 0
  ET_TRY(simp_forward(p.n_el, p.xn, 1.0, d.void_, 1, p.x, s));
  ET_TRY(filter_fwd(p.g, d.n_off, p.off, p.w, p.x, p.xt, s));
  ET_TRY(simp_forward(p.n_el, p.xt, penalty_at(p, 0), d.void_, 0, p.E, s));
  ET_CHECK_HIP(hipMemsetAsync(p.U, 0, p.n_dof * sizeof(double), s));
  ET_TRY(solve(p));
  double c = 0.0, h = 0.0;
  ET_TRY(compliance(p, &c));
  p.comp.push_back(c);
  for (int k = 0; k < n_iter; ++k) {
    const auto t0 = clock::now();
    // iteration k's update uses the sensitivities of the previous solve (penalty of
    // iteration k-1, or of the initial solve) and builds the densities with iteration k's
    const double pf = penalty_at(p, k);
    hipGraphExec_t exec = nullptr;
    ET_TRY(graph_for(p, penalty_at(p, k > 0 ? k - 1 : 0), pf, &exec));
    ET_CHECK_HIP(hipGraphLaunch(exec, s));
    ET_CHECK_HIP(hipMemcpyAsync(&h, p.scal + 1, sizeof(double), hipMemcpyDeviceToHost, s));
    ET_CHECK_HIP(hipStreamSynchronize(s));
    const double change = std::sqrt(h / (double)p.n_el);
    p.change.push_back(change);
    ET_TRY(solve(p));
    const double c_old = c;
    ET_TRY(compliance(p, &c));
    p.comp.push_back(c);
    p.iter_ms.push_back(std::chrono::duration<double, std::milli>(clock::now() - t0).count());
    p.done = k + 1;
    // CuTopOpt.optimize's test (TopOpt/_cuda.py:366-373); never true with zero tolerances
    if (change < d.ch_tol && std::fabs(c - c_old) / c < d.fun_tol && pf == d.penal) {
      p.converged = true;
      break;
    }
  }
  const int zero[2] = {0, 0};
  set_thread_wht_tiles(zero, zero);
  return ET_OK;
}

}  // namespace

int batch_create(BatchEngine** out, int n, const et_batch_problem* probs, int precision) {
  ET_REQUIRE(out && probs && n >= 1, "et_batch_create: bad arguments");
  ET_REQUIRE(precision == 0 || precision == 1, "et_batch_create: precision must be 0 or 1");
  *out = nullptr;
  BatchEngine* b = new BatchEngine();
  b->precision = precision;
  b->P.resize(n);
  for (int i = 0; i < n; ++i) {
    b->P[i].d = probs[i];
    const int st = setup_problem(b->P[i], precision);
    if (st != ET_OK) {
      const std::string msg = std::string("problem ") + std::to_string(i) + ": " + last_error();
      delete b;
      set_error(msg);
      return st;
    }
  }
  *out = b;
  return ET_OK;
}

int batch_run(BatchEngine* b, int n_iter, int n_workers) {
  ET_REQUIRE(b && n_iter >= 0 && n_workers >= 1 && n_workers <= 64, "et_batch_run: bad arguments");
  const int n = (int)b->P.size();
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return b->P[a].n_dof > b->P[c].n_dof; });
  std::atomic<int> next{0};
  int dev = 0;
  ET_CHECK_HIP(hipGetDevice(&dev));
  const auto t0 = std::chrono::steady_clock::now();
  auto worker = [&]() {
    (void)hipSetDevice(dev);
    for (int i = next.fetch_add(1); i < n; i = next.fetch_add(1)) {
      BatchProblem& p = b->P[order[i]];
      p.status = run_problem(p, n_iter);
      if (p.status != ET_OK) p.error = last_error();
    }
  };
  const int nw = std::min(n_workers, n);
  std::vector<std::thread> pool;
  for (int t = 1; t < nw; ++t) pool.emplace_back(worker);
  worker();
  for (auto& t : pool) t.join();
  b->run_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  for (int i = 0; i < n; ++i)
    if (b->P[i].status != ET_OK) {
      set_error(std::string("problem ") + std::to_string(i) + ": " + b->P[i].error);
      return b->P[i].status;
    }
  return ET_OK;
}

int batch_status(BatchEngine* b, int i, int* n_done, int* converged) {
  ET_REQUIRE(b && i >= 0 && i < (int)b->P.size(), "et_batch_status: bad problem index");
  if (n_done) *n_done = b->P[i].done;
  if (converged) *converged = b->P[i].converged ? 1 : 0;
  return ET_OK;
}

int batch_result(BatchEngine* b, int i, double* comp, int* pcg, double* iter_ms, double* x_host) {
  ET_REQUIRE(b && i >= 0 && i < (int)b->P.size(), "et_batch_result: bad problem index");
  BatchProblem& p = b->P[i];
  if (comp) std::copy(p.comp.begin(), p.comp.end(), comp);
  if (pcg) std::copy(p.pcg.begin(), p.pcg.end(), pcg);
  if (iter_ms) std::copy(p.iter_ms.begin(), p.iter_ms.end(), iter_ms);
  if (x_host) {
    ET_CHECK_HIP(hipMemcpyAsync(x_host, p.x, p.n_el * sizeof(double), hipMemcpyDeviceToHost, p.stream));
    ET_CHECK_HIP(hipStreamSynchronize(p.stream));
  }
  return ET_OK;
}

}  // namespace et

struct et_batch {
  et::BatchEngine* impl;
};

extern "C" {

int et_batch_create(et_batch** handle, int n, const et_batch_problem* probs, int precision) {
  try {
    ET_REQUIRE(handle, "et_batch_create: null handle");
    *handle = nullptr;
    et::BatchEngine* b = nullptr;
    const int st = et::batch_create(&b, n, probs, precision);
    if (st != ET_OK) return st;
    *handle = new et_batch{b};
    return ET_OK;
  } catch (...) {
    et::set_error("et_batch_create: host exception");
    return ET_ENOMEM;
  }
}

int et_batch_run(et_batch* handle, int n_iter, int n_workers, double* wall_ms) {
  try {
    ET_REQUIRE(handle, "et_batch_run: null handle");
    const int st = et::batch_run(handle->impl, n_iter, n_workers);
    if (wall_ms) *wall_ms = handle->impl->run_ms;
    return st;
  } catch (...) {
    et::set_error("et_batch_run: host exception");
    return ET_EINVAL;
  }
}

int et_batch_result(et_batch* handle, int i, double* comp, int* pcg, double* iter_ms, double* x_host) {
  try {
    ET_REQUIRE(handle, "et_batch_result: null handle");
    return et::batch_result(handle->impl, i, comp, pcg, iter_ms, x_host);
  } catch (...) {
    et::set_error("et_batch_result: host exception");
    return ET_EINVAL;
  }
}

int et_batch_status(et_batch* handle, int i, int* n_iterations, int* converged) {
  try {
    ET_REQUIRE(handle, "et_batch_status: null handle");
    return et::batch_status(handle->impl, i, n_iterations, converged);
  } catch (...) {
    et::set_error("et_batch_status: host exception");
    return ET_EINVAL;
  }
}

int et_batch_destroy(et_batch* handle) {
  if (handle) {
    delete handle->impl;
    delete handle;
  }
  return ET_OK;
}

}  // extern "C"
