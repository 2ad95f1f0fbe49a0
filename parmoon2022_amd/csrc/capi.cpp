// extern "C" boundary of libpmcd.so (declarations in include/pmcd.h).
#include "../../include/pmcd.h"
#include "system.hpp"
#include "batch.hpp"
#include "io.hpp"
#include "dropin.hpp"
#include "server.hpp"
#include <cmath>
#include <random>
#include <cstring>
#include <algorithm>
#include <string>

using namespace pmcd;

// Each handle owns a device-byte tally, declared before the object it
// charges so that it outlives that object's buffers.
struct pmcd_system {
  hipStream_t stream = nullptr;
  DevTally tally;
  std::unique_ptr<CDSystem> sys;
};

struct pmcd_batch {
  hipStream_t stream = nullptr;
  DevTally tally;
  std::unique_ptr<BatchSystem> sys;
};

struct pmcd_assembler {
  hipStream_t stream = nullptr;
  DevTally tally;
  std::unique_ptr<CallerAssembler> a;
};

struct pmcd_solver {
  hipStream_t stream = nullptr;
  SolverDesc sd{};
  DevTally tally;
  std::unique_ptr<CDSystem> sys;
};

struct pmcd_server {
  std::unique_ptr<Server> srv;
};

namespace {
thread_local std::string g_err;

template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch(const std::exception& e) {
    g_err = e.what();
  } catch(...) {
    g_err = "unknown error";
  }
  return 1;
}

// guarded() with the device allocations of the call charged to handle h
template <class H, class F>
int guarded_on(H* h, F&& f) {
  TallyScope scope(h ? &h->tally : nullptr);
  return guarded(std::forward<F>(f));
}

Level& level_of(pmcd_system* s, int level) {
  PMCD_CHECK(s && s->sys, "null system");
  const int nl = s->sys->n_levels();
  if(level < 0) level = nl + level;
  PMCD_CHECK(level >= 0 && level < nl, "level ", level, " out of range");
  return s->sys->level(level);
}

ProblemDesc to_desc(const pmcd_problem_desc& d) {
  ProblemDesc p{};
  p.geometry = d.geometry;
  p.n_refine = d.n_refine;
  p.structured_n = d.structured_n;
  p.order = d.order;
  p.supg = d.supg;
  p.linear_quad = d.linear_quad;
  p.mg_levels = d.mg_levels;
  p.ex.kind = d.example;
  p.ex.eps = d.eps;
  for(int k = 0; k < 3; ++k) p.ex.b[k] = d.b[k];
  p.ex.c = d.c;
  p.ex.f0 = d.f0;
  set_example_time(p.ex, 0.0);
  if(d.geometry == PMCD_MESH_FILE) {
    PMCD_CHECK(d.mesh_file != nullptr, "geometry PMCD_MESH_FILE needs mesh_file");
    p.mesh_file = d.mesh_file;
  }
  PMCD_CHECK(d.adapt_steps >= 0 && d.adapt_steps <= 30, "adapt_steps out of range");
  if(d.adapt_steps > 0) {
    PMCD_CHECK(d.indicator != nullptr, "adapt_steps > 0 needs an indicator function");
    p.adapt_steps = d.adapt_steps;
    p.indicator = d.indicator;
    p.indicator_ctx = d.indicator_ctx;
    PMCD_CHECK(d.adapt_conforming == 0 || d.adapt_conforming == 1, "adapt_conforming must be 0 or 1");
    p.adapt_conforming = d.adapt_conforming;
  }
  return p;
}

SolverDesc to_solver(const pmcd_solver_desc& s) {
  SolverDesc d{};
  d.method = s.method;
  d.precond = s.precond;
  d.max_iter = s.max_iter;
  d.min_iter = s.min_iter;
  d.res_tol = s.res_tol;
  d.res_red = s.res_red;
  d.div_factor = s.div_factor > 0 ? s.div_factor : 1e10;
  d.restart = s.restart > 0 ? s.restart : 20;
  d.n_pre = s.n_pre;
  d.n_post = s.n_post;
  d.jacobi_damp = s.jacobi_damp;
  d.mg_damp = s.mg_damp;
  d.mixed = s.mixed;
  d.smoother = s.smoother;
  d.sor_omega = s.sor_omega;
  d.cycle = s.cycle;
  d.coarse_smoother = s.coarse_smoother;
  d.coarse_residual = s.coarse_residual;
  d.coarse_max_iter = s.coarse_max_iter;
  PMCD_CHECK(d.coarse_smoother >= 0 && d.coarse_smoother <= 3, "unknown coarse smoother ", d.coarse_smoother);
  PMCD_CHECK(d.smoother >= 0 && d.smoother <= 2, "unknown multigrid smoother ", d.smoother);
  PMCD_CHECK(d.cycle >= 0 && d.cycle <= 2, "unknown multigrid cycle ", d.cycle);
  return d;
}
}  // namespace

struct pmcd_host {
  std::vector<std::unique_ptr<Level>> levels;
  int n_quad = 0;
};

namespace {
Level& host_level(pmcd_host* h, int level) {
  PMCD_CHECK(h, "null handle");
  const int nl = (int)h->levels.size();
  if(level < 0) level = nl + level;
  PMCD_CHECK(level >= 0 && level < nl, "level ", level, " out of range");
  return *h->levels[level];
}
void fill_info(const Level& L, int n_quad, int n_levels, long long* info) {
  info[0] = L.mesh.n_cells;
  info[1] = L.space.n_dof;
  info[2] = L.space.n_active;
  info[3] = L.st.nnz;
  info[4] = L.space.n_local;
  info[5] = n_quad;
  info[6] = n_levels;
  info[7] = L.space.n_dirichlet;
}
void copy_coords(const Mesh& m, double* out) {
  for(int c = 0; c < m.n_cells; ++c)
    for(int a = 0; a < m.nv; ++a)
      for(int k = 0; k < 3; ++k) out[((size_t)c * m.nv + a) * 3 + k] = m.coords[3 * m.cell(c)[a] + k];
}
}  // namespace

extern "C" {

const char* pmcd_last_error(void) { return g_err.c_str(); }

int pmcd_host_create(const pmcd_problem_desc* desc, pmcd_host** out) {
  return guarded([&] {
    PMCD_CHECK(desc && out, "null argument");
    auto h = std::make_unique<pmcd_host>();
    ProblemDesc d = to_desc(*desc);
    h->levels = build_levels(d);
    const Shape s = h->levels.back()->mesh.shape;
    h->n_quad = quad_formula(cell_rule_for(s, d.order, d.linear_quad != 0)).n();
    *out = h.release();
  });
}

int pmcd_host_destroy(pmcd_host* h) {
  return guarded([&] { delete h; });
}

int pmcd_host_info(pmcd_host* h, int level, long long* info) {
  return guarded([&] { fill_info(host_level(h, level), h->n_quad, (int)h->levels.size(), info); });
}

int pmcd_host_dofmap(pmcd_host* h, int level, int* out) {
  return guarded([&] {
    Level& L = host_level(h, level);
    std::memcpy(out, L.space.dof.data(), L.space.dof.size() * sizeof(int));
  });
}

int pmcd_host_pattern(pmcd_host* h, int level, int* row_ptr, int* col) {
  return guarded([&] {
    Level& L = host_level(h, level);
    std::memcpy(row_ptr, L.st.row_ptr.data(), L.st.row_ptr.size() * sizeof(int));
    std::memcpy(col, L.st.col.data(), L.st.col.size() * sizeof(int));
  });
}

int pmcd_host_sweep_order(pmcd_host* h, int level, int max_exact_classes, int* rows, int* class_ptr,
                          int* n_classes, int* exact) {
  return guarded([&] {
    PMCD_CHECK(rows && n_classes && exact, "null argument");
    Level& L = host_level(h, level);
    SweepOrder o = build_sweep_order(L.st, max_exact_classes);
    std::memcpy(rows, o.rows.data(), o.rows.size() * sizeof(int));
    *n_classes = (int)o.class_ptr.size() - 1;
    if(class_ptr) std::memcpy(class_ptr, o.class_ptr.data(), o.class_ptr.size() * sizeof(int));
    *exact = o.exact ? 1 : 0;
  });
}

int pmcd_host_sell_layout(pmcd_host* h, int level, int* perm, int* len, long long* slice_off, int* n_slots) {
  return guarded([&] {
    PMCD_CHECK(perm && len && slice_off && n_slots, "null argument");
    Level& L = host_level(h, level);
    SellLayout o = build_sell(L.st, L.n_active(), sell_window());
    std::memcpy(perm, o.perm.data(), o.perm.size() * sizeof(int));
    std::memcpy(len, o.len.data(), o.len.size() * sizeof(int));
    std::memcpy(slice_off, o.slice_off.data(), o.slice_off.size() * sizeof(long long));
    *n_slots = o.n_slots();
  });
}

int pmcd_host_sell_footprint(pmcd_host* h, int level, int* fp_off, int* fp, unsigned short* lcol,
                             long long* sizes) {
  return guarded([&] {
    PMCD_CHECK(sizes, "null argument");
    PMCD_CHECK(sell_window() == 4 * kSellRows, "column footprints need 4-slice windows");
    Level& L = host_level(h, level);
    SellLayout o = build_sell(L.st, L.n_active(), sell_window());
    SellFootprint f = build_sell_footprint(L.st, o, 4 * kSellRows, kSellMaxFootprint);
    sizes[0] = (long long)f.fp_off.size();
    sizes[1] = (long long)f.fp.size();
    sizes[2] = (long long)f.lcol.size();
    sizes[3] = f.fp_max;
    if(fp_off) std::memcpy(fp_off, f.fp_off.data(), f.fp_off.size() * sizeof(int));
    if(fp) std::memcpy(fp, f.fp.data(), f.fp.size() * sizeof(int));
    if(lcol) std::memcpy(lcol, f.lcol.data(), f.lcol.size() * sizeof(unsigned short));
  });
}

static int copy_hanging(const Level& L, int* out, double* w, int* n) {
  PMCD_CHECK(n != nullptr, "null argument");
  const FESpace& sp = L.space;
  constexpr int MC = FESpace::kMaxCoupling;
  *n = sp.n_hanging;
  for(int t = 0; t < sp.n_hanging; ++t) {
    const int h = sp.hang_order[t];
    if(out) {
      out[4 * t] = sp.n_active_non_hanging() + h;
      for(int k = 0; k < MC; ++k) out[4 * t + 1 + k] = k < sp.hang_n[h] ? sp.hang_cpl[MC * h + k] : -1;
    }
    if(w)
      for(int k = 0; k < MC; ++k) w[3 * t + k] = k < sp.hang_n[h] ? sp.hang_w[MC * h + k] : 0.0;
  }
  return 0;
}

int pmcd_host_hanging(pmcd_host* h, int level, int* out, double* w, int* n) {
  return guarded([&] { copy_hanging(host_level(h, level), out, w, n); });
}

int pmcd_host_write_mesh(pmcd_host* h, int level, const char* path) {
  return guarded([&] {
    PMCD_CHECK(path != nullptr, "null path");
    write_mesh_file(host_level(h, level).mesh, path);
  });
}

int pmcd_host_cell_coords(pmcd_host* h, int level, double* out) {
  return guarded([&] { copy_coords(host_level(h, level).mesh, out); });
}

int pmcd_device_count(int* n) {
  return guarded([&] { PMCD_HIP(hipGetDeviceCount(n)); });
}

int pmcd_mem_info(unsigned long long* free_bytes, unsigned long long* total_bytes) {
  return guarded([&] {
    PMCD_CHECK(free_bytes && total_bytes, "null argument");
    size_t f = 0, t = 0;
    PMCD_HIP(hipMemGetInfo(&f, &t));
    *free_bytes = f;
    *total_bytes = t;
  });
}

int pmcd_system_set_data(pmcd_system* s, const pmcd_problem_desc* desc) {
  return guarded_on(s, [&] {
    PMCD_CHECK(s && s->sys && desc, "null argument");
    const ProblemDesc& cur = s->sys->desc();
    const ProblemDesc d = to_desc(*desc);
    PMCD_CHECK(d.geometry == cur.geometry && d.n_refine == cur.n_refine &&
                   d.structured_n == cur.structured_n && d.order == cur.order && d.supg == cur.supg &&
                   d.linear_quad == cur.linear_quad && d.mg_levels == cur.mg_levels,
               "pmcd_system_set_data changes only the PDE data (example, eps, b, c, f0)");
    PMCD_CHECK(d.adapt_steps == cur.adapt_steps && d.indicator == cur.indicator &&
                   d.indicator_ctx == cur.indicator_ctx && d.adapt_conforming == cur.adapt_conforming,
               "pmcd_system_set_data changes only the PDE data (example, eps, b, c, f0)");
    s->sys->set_example(d.ex);
  });
}

int pmcd_system_create(const pmcd_problem_desc* desc, pmcd_system** out) {
  return guarded([&] {
    PMCD_CHECK(desc && out, "null argument");
    int ndev = 0;
    if(hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
      PMCD_FAIL("no GPU visible: the HIP path needs an MI355X device");
    auto s = std::make_unique<pmcd_system>();
    PMCD_HIP(hipStreamCreate(&s->stream));
    TallyScope scope(&s->tally);
    try {
      s->sys = std::make_unique<CDSystem>(to_desc(*desc), s->stream);
    } catch(...) {
      (void)hipStreamDestroy(s->stream);
      throw;
    }
    *out = s.release();
  });
}

int pmcd_system_destroy(pmcd_system* s) {
  return guarded([&] {
    if(!s) return;
    s->sys.reset();
    if(s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
  });
}

int pmcd_system_info(pmcd_system* s, int level, long long* info) {
  return guarded_on(s, [&] {
    Level& L = level_of(s, level);
    info[0] = L.mesh.n_cells;
    info[1] = L.n();
    info[2] = L.n_active();
    info[3] = L.st.nnz;
    info[4] = L.space.n_local;
    info[5] = quad_formula(s->sys->cell_rule()).n();
    info[6] = s->sys->n_levels();
    info[7] = L.space.n_dirichlet;
  });
}

int pmcd_get_hanging(pmcd_system* s, int level, int* out, double* w, int* n) {
  return guarded_on(s, [&] { copy_hanging(level_of(s, level), out, w, n); });
}

int pmcd_get_dofmap(pmcd_system* s, int level, int* out) {
  return guarded_on(s, [&] {
    Level& L = level_of(s, level);
    std::memcpy(out, L.space.dof.data(), L.space.dof.size() * sizeof(int));
  });
}

int pmcd_get_pattern(pmcd_system* s, int level, int* row_ptr, int* col) {
  return guarded_on(s, [&] {
    Level& L = level_of(s, level);
    std::memcpy(row_ptr, L.st.row_ptr.data(), L.st.row_ptr.size() * sizeof(int));
    std::memcpy(col, L.st.col.data(), L.st.col.size() * sizeof(int));
  });
}

int pmcd_get_cell_coords(pmcd_system* s, int level, double* out) {
  return guarded_on(s, [&] {
    Level& L = level_of(s, level);
    const Mesh& m = L.mesh;
    for(int c = 0; c < m.n_cells; ++c)
      for(int a = 0; a < m.nv; ++a)
        for(int k = 0; k < 3; ++k) out[((size_t)c * m.nv + a) * 3 + k] = m.coords[3 * m.cell(c)[a] + k];
  });
}

int pmcd_assemble(pmcd_system* s, double* ms) {
  return guarded_on(s, [&] {
    PMCD_CHECK(s && s->sys, "null system");
    s->sys->assemble();
    if(ms) *ms = s->sys->last_assemble_ms;
  });
}

int pmcd_get_matrix(pmcd_system* s, int level, double* val) {
  return guarded_on(s, [&] {
    Level& L = level_of(s, level);
    L.val.download(val, L.val.n, s->stream);
    PMCD_HIP(hipStreamSynchronize(s->stream));
  });
}

int pmcd_get_rhs(pmcd_system* s, int level, double* rhs) {
  return guarded_on(s, [&] {
    Level& L = level_of(s, level);
    L.rhs.download(rhs, L.n(), s->stream);
    PMCD_HIP(hipStreamSynchronize(s->stream));
  });
}

int pmcd_solve(pmcd_system* s, const pmcd_solver_desc* sd, pmcd_solve_stats* stats) {
  return guarded_on(s, [&] {
    PMCD_CHECK(s && s->sys && sd, "null argument");
    SolveStats st = s->sys->solve(to_solver(*sd));
    if(stats) {
      stats->iterations = st.iterations;
      stats->converged = st.converged;
      stats->initial_residual = st.initial_residual;
      stats->final_residual = st.final_residual;
      stats->seconds = st.seconds;
    }
  });
}

int pmcd_get_solution(pmcd_system* s, double* out) {
  return guarded_on(s, [&] {
    Level& L = level_of(s, -1);
    L.sol.download(out, L.n(), s->stream);
    PMCD_HIP(hipStreamSynchronize(s->stream));
  });
}

int pmcd_set_solution(pmcd_system* s, const double* in) {
  return guarded_on(s, [&] {
    Level& L = level_of(s, -1);
    PMCD_HIP(hipMemcpyAsync(L.sol.p, in, sizeof(double) * L.n(), hipMemcpyHostToDevice, s->stream));
    PMCD_HIP(hipStreamSynchronize(s->stream));
  });
}

int pmcd_errors(pmcd_system* s, double* out5) {
  return guarded_on(s, [&] {
    PMCD_CHECK(s && s->sys, "null system");
    s->sys->errors(out5);
  });
}

int pmcd_apply(pmcd_system* s, int level, const double* x, double* y, double a, int mode) {
  return guarded_on(s, [&] {
    Level& L = level_of(s, level);
    DevBuf<double> dx, dy;
    dx.upload(x, L.n(), s->stream);
    if(mode == 1) dy.upload(y, L.n(), s->stream);
    else dy.alloc(L.n());
    if(mode == 1) s->sys->apply_scaled_add(L, dx.p, dy.p, a);
    else s->sys->apply(L, dx.p, dy.p);
    dy.download(y, L.n(), s->stream);
    PMCD_HIP(hipStreamSynchronize(s->stream));
  });
}

int pmcd_jacobi_sweep(pmcd_system* s, int level, const double* b, double* x, double omega) {
  return guarded_on(s, [&] {
    PMCD_CHECK(b && x, "null argument");
    Level& L = level_of(s, level);
    DevBuf<double> db, dx, dy(L.n());
    db.upload(b, L.n(), s->stream);
    dx.upload(x, L.n(), s->stream);
    level_jacobi(L, db.p, dx.p, dy.p, omega, s->stream);
    dy.download(x, L.n(), s->stream);
    PMCD_HIP(hipStreamSynchronize(s->stream));
  });
}

int pmcd_sor_sweep(pmcd_system* s, int level, const double* b, double* x, double omega, int flag,
                   int* exact) {
  return guarded_on(s, [&] {
    PMCD_CHECK(b && x, "null argument");
    Level& L = level_of(s, level);
    DevBuf<double> db, dx;
    db.upload(b, L.n(), s->stream);
    dx.upload(x, L.n(), s->stream);
    s->sys->sor_sweep(L, db.p, dx.p, omega, flag);
    dx.download(x, L.n(), s->stream);
    PMCD_HIP(hipStreamSynchronize(s->stream));
    if(exact) *exact = L.sweep_exact ? 1 : 0;
  });
}

int pmcd_vcycle(pmcd_system* s, const pmcd_solver_desc* sd, const double* v, double* z) {
  return guarded_on(s, [&] {
    PMCD_CHECK(s && s->sys && sd, "null argument");
    Level& F = s->sys->finest();
    s->sys->prepare(to_solver(*sd));
    DevBuf<double> dv, dz(F.n());
    dv.upload(v, F.n(), s->stream);
    s->sys->vcycle(dv.p, dz.p);
    dz.download(z, F.n(), s->stream);
    PMCD_HIP(hipStreamSynchronize(s->stream));
  });
}

int pmcd_time_kernel(pmcd_system* s, int which, int level, int iters, double* avg_ms) {
  return guarded_on(s, [&] {
    PMCD_CHECK(iters > 0, "iters must be positive");
    s->sys->time_kernel(which, level, iters, avg_ms);
  });
}

int pmcd_apply_transfer(pmcd_system* s, int level, int which, const double* x, double* y, double a) {
  return guarded_on(s, [&] {
    PMCD_CHECK(x && y, "null argument");
    s->sys->apply_transfer(level, which, x, y, a);
  });
}

int pmcd_sell_stats(pmcd_system* s, int level, long long* out) {
  return guarded_on(s, [&] {
    PMCD_CHECK(out, "null argument");
    const int nl = s->sys->n_levels();
    const int l = level < 0 ? nl + level : level;
    PMCD_CHECK(l >= 0 && l < nl, "level ", level, " out of range");
    const Level& L = s->sys->level(l);
    const bool fp = L.sell_fp_max > 0;
    out[0] = L.sell_slots;
    out[1] = L.sell_slots >= 0 ? (long long)L.sell_val.n : 0;
    out[2] = fp ? (long long)L.sell_fp_off.n - 1 : 0;
    out[3] = fp ? (long long)L.sell_fp.n : 0;
    out[4] = fp ? L.sell_fp_max : 0;
    out[5] = L.sell_slots >= 0 ? sell_mode_for(L.sell_window, fp ? L.sell_fp_max : 0) : -1;
  });
}

int pmcd_cycle_directions(int type, int n_levels, int* out, int cap, int* n) {
  return guarded([&] {
    PMCD_CHECK(n && (out || cap == 0), "null argument");
    const std::vector<int> d = cycle_directions(type, n_levels);
    for(int i = 0; i < (int)d.size() && i < cap; ++i) out[i] = d[i];
    *n = (int)d.size();
  });
}

int pmcd_set_tuning(int key, int value) {
  return guarded([&] { set_tuning(key, value); });
}

int pmcd_get_tuning(int key, int* value) {
  return guarded([&] {
    PMCD_CHECK(value != nullptr, "null output");
    *value = get_tuning(key);
  });
}

int pmcd_probe_mfma(double* out256) {
  return guarded([&] {
    int ndev = 0;
    if(hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) PMCD_FAIL("no GPU visible");
    DevBuf<double> d(256);
    launch_mfma_probe(d.p, nullptr);
    PMCD_HIP(hipMemcpy(out256, d.p, 256 * sizeof(double), hipMemcpyDeviceToHost));
  });
}

int pmcd_spmv_device(int n, int n_active, const int* row_ptr, const int* col, const double* val,
                     const double* x, const double* z, double* y, double alpha, double beta,
                     void* stream) {
  return guarded([&] {
    launch_spmv(n, n_active, row_ptr, col, val, x, z, y, alpha, beta, (hipStream_t)stream);
  });
}

int pmcd_dot_device(int n, const double* x, const double* y, double* result, void* stream) {
  return guarded([&] {
    DevBuf<double> part(1024 + 8), out(1);
    launch_dot(n, x, y, part.p, out.p, (hipStream_t)stream);
    PMCD_HIP(hipMemcpyAsync(result, out.p, sizeof(double), hipMemcpyDeviceToHost, (hipStream_t)stream));
    PMCD_HIP(hipStreamSynchronize((hipStream_t)stream));
  });
}

int pmcd_dot_device_noalloc(int n, const double* x, const double* y, double* result, void* stream) {
  return guarded([&] {
    PMCD_CHECK(n >= 0 && result && (n == 0 || (x && y)), "invalid argument");
    const int b = reduce_blocks(n);
    // partials in [0, b), the result at kMaxReduceBlocks
    PMCD_CHECK(b <= kMaxReduceBlocks && kMaxReduceBlocks + 1 <= kNoAllocScratchDoubles,
               "reduction of ", n, " entries needs more than the ", kNoAllocScratchDoubles,
               " scratch doubles");
    double* scratch = noalloc_scratch();
    launch_dot(n, x, y, scratch, scratch + kMaxReduceBlocks, (hipStream_t)stream);
    PMCD_HIP(hipMemcpyAsync(result, scratch + kMaxReduceBlocks, sizeof(double), hipMemcpyDeviceToHost,
                            (hipStream_t)stream));
    PMCD_HIP(hipStreamSynchronize((hipStream_t)stream));
  });
}

// ---- batched instances -------------------------------------------------------
int pmcd_batch_create(const pmcd_problem_desc* descs, int n_inst, pmcd_batch** out) {
  return guarded([&] {
    PMCD_CHECK(descs && out && n_inst > 0, "null argument or empty batch");
    int ndev = 0;
    if(hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
      PMCD_FAIL("no GPU visible: the HIP path needs an MI355X device");
    std::vector<ProblemDesc> d(n_inst);
    for(int i = 0; i < n_inst; ++i) {
      d[i] = to_desc(descs[i]);
      PMCD_CHECK(d[i].adapt_steps == 0, "batched instances take regularly refined meshes");
    }
    auto b = std::make_unique<pmcd_batch>();
    PMCD_HIP(hipStreamCreate(&b->stream));
    TallyScope scope(&b->tally);
    try {
      b->sys = std::make_unique<BatchSystem>(d, b->stream);
    } catch(...) {
      (void)hipStreamDestroy(b->stream);
      throw;
    }
    *out = b.release();
  });
}

int pmcd_batch_destroy(pmcd_batch* b) {
  return guarded([&] {
    if(!b) return;
    b->sys.reset();
    if(b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
  });
}

int pmcd_batch_info(pmcd_batch* b, int inst, long long* info) {
  return guarded_on(b, [&] {
    PMCD_CHECK(b && b->sys && info, "null argument");
    BatchSystem& s = *b->sys;
    if(inst < 0) {
      info[0] = s.n_inst(); info[1] = s.n_dof(); info[2] = s.n_active(); info[3] = s.nnz();
      info[4] = s.n_cells();
    } else {
      PMCD_CHECK(inst < s.n_inst(), "instance out of range");
      info[0] = 1; info[1] = s.inst_n_dof(inst); info[2] = s.inst_n_active(inst);
      info[3] = s.inst_nnz(inst); info[4] = 0;
    }
  });
}

int pmcd_batch_assemble(pmcd_batch* b, double* ms) {
  return guarded_on(b, [&] {
    PMCD_CHECK(b && b->sys, "null argument");
    b->sys->assemble();
    if(ms) *ms = b->sys->last_assemble_ms;
  });
}

int pmcd_batch_solve(pmcd_batch* b, const pmcd_solver_desc* sd, pmcd_solve_stats* stats) {
  return guarded_on(b, [&] {
    PMCD_CHECK(b && b->sys && sd, "null argument");
    auto st = b->sys->solve(to_solver(*sd));
    if(stats)
      for(size_t i = 0; i < st.size(); ++i) {
        stats[i].iterations = st[i].iterations;
        stats[i].converged = st[i].converged;
        stats[i].initial_residual = st[i].initial_residual;
        stats[i].final_residual = st[i].final_residual;
        stats[i].seconds = st[i].seconds;
      }
  });
}

int pmcd_batch_get_solution(pmcd_batch* b, int inst, double* out) {
  return guarded_on(b, [&] {
    PMCD_CHECK(b && b->sys && out, "null argument");
    b->sys->get_solution(inst, out);
  });
}

int pmcd_batch_set_solution(pmcd_batch* b, int inst, const double* in) {
  return guarded_on(b, [&] {
    PMCD_CHECK(b && b->sys && in, "null argument");
    b->sys->set_solution(inst, in);
  });
}

int pmcd_batch_get_matrix(pmcd_batch* b, int inst, double* val) {
  return guarded_on(b, [&] {
    PMCD_CHECK(b && b->sys && val, "null argument");
    b->sys->get_matrix(inst, val);
  });
}

int pmcd_batch_get_rhs(pmcd_batch* b, int inst, double* rhs) {
  return guarded_on(b, [&] {
    PMCD_CHECK(b && b->sys && rhs, "null argument");
    b->sys->get_rhs(inst, rhs);
  });
}

int pmcd_tcd_init(pmcd_system* s, const pmcd_tcd_desc* t) {
  return guarded_on(s, [&] {
    PMCD_CHECK(s && t, "null argument");
    PMCD_CHECK(s->sys->desc().adapt_steps == 0, "time-dependent problems take regularly refined meshes");
    TcdDesc d{t->scheme, t->bdf_first, t->tau, t->t0};
    s->sys->tcd_init(d);
  });
}

int pmcd_tcd_step(pmcd_system* s, const pmcd_solver_desc* sd, pmcd_solve_stats* stats, double* time) {
  return guarded_on(s, [&] {
    PMCD_CHECK(s && sd, "null argument");
    const SolveStats st = s->sys->tcd_step(to_solver(*sd));
    if(stats) {
      stats->iterations = st.iterations;
      stats->converged = st.converged;
      stats->initial_residual = st.initial_residual;
      stats->final_residual = st.final_residual;
      stats->seconds = st.seconds;
    }
    if(time) *time = s->sys->tcd_time();
  });
}

int pmcd_afc_solve(pmcd_system* s, const pmcd_solver_desc* sd, const double* prm, double* out,
                   pmcd_solve_stats* stats) {
  return guarded_on(s, [&] {
    PMCD_CHECK(s && sd && prm && out, "null argument");
    PMCD_CHECK(s->sys->desc().adapt_steps == 0,
               "algebraic flux correction takes regularly refined meshes (ModifyMatrixAccordingToCouplingAFC "
               "is not implemented)");
    PMCD_CHECK(prm[0] > 0 && prm[1] >= 1, "afc_nonlinloop_epsilon must be positive, afc_nonlinloop_maxit >= 1");
    PMCD_CHECK(prm[11] == 0.0 || prm[11] == 1.0, "afc limiter must be 0 (kuzmin) or 1 (BJK17)");
    PMCD_CHECK(prm[12] == 0.0 || prm[12] == 1.0 || prm[12] == 2.0,
               "afc iteration scheme must be 0 (fixed_point_rhs), 1 (fixed_point_matrix) or 2 (newton)");
    PMCD_CHECK(prm[16] == 0.0 || prm[16] == 1.0, "afc_nonlinloop_anderson_acc must be 0 or 1");
    PMCD_CHECK(prm[16] == 0.0 || (prm[17] >= 1 && prm[17] <= LinComb::kMax && prm[18] >= 0),
               "afc_nonlinloop_anderson_acc_vec must be in 1..", LinComb::kMax,
               ", afc_nonlinloop_anderson_acc_start >= 0");
    const SolveStats st = s->sys->afc_solve(to_solver(*sd), prm, out);
    if(stats) {
      stats->iterations = st.iterations;
      stats->converged = st.converged;
      stats->initial_residual = st.initial_residual;
      stats->final_residual = st.final_residual;
      stats->seconds = st.seconds;
    }
  });
}

int pmcd_vector_write(const char* path, long long n, const double* v) {
  return guarded([&] {
    PMCD_CHECK(path && (v || n == 0) && n >= 0, "invalid argument");
    write_vector_stream(path, n, v);
  });
}

int pmcd_vector_read(const char* path, long long n, double* v) {
  return guarded([&] {
    PMCD_CHECK(path && (v || n == 0) && n >= 0, "invalid argument");
    read_vector_stream(path, n, v);
  });
}

int pmcd_vector_write_market(const char* path, long long n, const double* v) {
  return guarded([&] {
    PMCD_CHECK(path && (v || n == 0) && n >= 0, "invalid argument");
    write_vector_market(path, n, v);
  });
}

int pmcd_matrix_write_market(const char* path, int n_rows, int n_cols, const int* row_ptr, const int* col,
                             const double* val) {
  return guarded([&] {
    PMCD_CHECK(path && row_ptr && n_rows >= 0 && n_cols >= 0, "invalid argument");
    PMCD_CHECK(row_ptr[n_rows] == 0 || (col && val), "invalid argument");
    write_matrix_market(path, n_rows, n_cols, row_ptr, col, val);
  });
}


// ---- drop-in entries on caller-owned spaces, structures and matrices ----------
static ProblemDesc forms_desc(const pmcd_problem_desc& la) {
  pmcd_problem_desc d = la;
  d.geometry = 0;
  d.mesh_file = nullptr;
  d.adapt_steps = 0;
  return to_desc(d);
}

static void check_visible_gpu() {
  int ndev = 0;
  if(hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    PMCD_FAIL("no GPU visible: the HIP path needs an MI355X device");
}

int pmcd_assembler_create(const pmcd_fe_space* space, const pmcd_structure* st, const pmcd_problem_desc* la,
                          pmcd_assembler** out) {
  return guarded([&] {
    PMCD_CHECK(space && st && la && out, "null argument");
    PMCD_CHECK(st->n_rows == space->n_dof, "structure rows ", st->n_rows, " differ from the space's DOFs ",
               space->n_dof);
    PMCD_CHECK(space->shape >= 0 && space->shape <= 3, "unknown cell shape ", space->shape);
    PMCD_CHECK(space->n_hanging >= 0, "negative hanging DOF count");
    PMCD_CHECK(space->n_hanging == 0 || space->shape <= 1, "hanging DOFs are supported in 2D only");
    check_visible_gpu();
    auto h = std::make_unique<pmcd_assembler>();
    PMCD_HIP(hipStreamCreate(&h->stream));
    TallyScope scope(&h->tally);
    try {
      h->a = std::make_unique<CallerAssembler>((Shape)space->shape, space->order, space->n_cells, space->n_dof,
                                               space->n_active, space->vertices, space->global_numbers,
                                               space->begin_index, st->row_ptr, st->col, forms_desc(*la),
                                               h->stream, space->n_hanging, space->hanging,
                                               space->hanging_weights);
    } catch(...) {
      (void)hipStreamDestroy(h->stream);
      throw;
    }
    *out = h.release();
  });
}

int pmcd_assembler_destroy(pmcd_assembler* h) {
  return guarded([&] {
    if(!h) return;
    h->a.reset();
    if(h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
  });
}

int pmcd_assembler_assemble(pmcd_assembler* h, const double* dirichlet_values, double* entries, double* rhs) {
  return guarded_on(h, [&] {
    PMCD_CHECK(h && h->a, "null assembler");
    h->a->assemble(dirichlet_values, false);
    h->a->add_into_host(entries, rhs);
  });
}

int pmcd_assembler_assemble_device(pmcd_assembler* h, const double* dirichlet_values, double* entries,
                                   double* rhs, void* stream) {
  return guarded_on(h, [&] {
    PMCD_CHECK(h && h->a, "null assembler");
    // the caller's arrays (and device Dirichlet values) are complete on its stream
    PMCD_HIP(hipStreamSynchronize((hipStream_t)stream));
    h->a->assemble(dirichlet_values, true);
    h->a->add_into_device(entries, rhs, (hipStream_t)stream);
  });
}

int pmcd_assemble_into(const pmcd_fe_space* space, const pmcd_structure* st, const pmcd_problem_desc* la,
                       const double* dirichlet_values, double* entries, double* rhs) {
  pmcd_assembler* h = nullptr;
  int rc = pmcd_assembler_create(space, st, la, &h);
  if(rc != 0) return rc;
  rc = pmcd_assembler_assemble(h, dirichlet_values, entries, rhs);
  std::string err = g_err;
  pmcd_assembler_destroy(h);
  g_err = err;
  return rc;
}

int pmcd_solver_create(const pmcd_solver_desc* sd, pmcd_solver** out) {
  return guarded([&] {
    PMCD_CHECK(sd && out, "null argument");
    check_visible_gpu();
    auto s = std::make_unique<pmcd_solver>();
    s->sd = to_solver(*sd);
    PMCD_HIP(hipStreamCreate(&s->stream));
    *out = s.release();
  });
}

int pmcd_solver_destroy(pmcd_solver* s) {
  return guarded([&] {
    if(!s) return;
    s->sys.reset();
    if(s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
  });
}

int pmcd_solver_update_matrix(pmcd_solver* s, int n_levels, const pmcd_csr_level* levels) {
  return guarded_on(s, [&] {
    PMCD_CHECK(s && levels && n_levels >= 1, "null argument or no level");
    const bool mg = s->sd.precond == 2;
    PMCD_CHECK(mg || n_levels == 1, "only a multigrid preconditioner takes several levels");
    // same patterns: only the values change (Solver::update_matrix on a
    // matrix of the same structure)
    bool same = s->sys && s->sys->n_levels() == n_levels;
    for(int l = 0; same && l < n_levels; ++l) {
      const Level& L = s->sys->level(l);
      const pmcd_csr_level& o = levels[l];
      same = o.n == L.n() && o.n_active == L.n_active() && o.row_ptr && o.col &&
             std::equal(o.row_ptr, o.row_ptr + o.n + 1, L.st.row_ptr.begin()) &&
             std::equal(o.col, o.col + o.row_ptr[o.n], L.st.col.begin());
    }
    if(same) {
      for(int l = 0; l < n_levels; ++l) {
        PMCD_CHECK(levels[l].entries, "level ", l, ": null entries");
        s->sys->set_level_values(l, levels[l].entries);
      }
      return;
    }
    std::vector<OperatorLevel> lv(n_levels);
    for(int l = 0; l < n_levels; ++l) {
      const pmcd_csr_level& o = levels[l];
      PMCD_CHECK(o.n > 0 && o.row_ptr && o.col && o.entries, "level ", l, ": null or empty matrix");
      OperatorLevel& d = lv[l];
      d.n = o.n;
      d.n_active = o.n_active;
      d.row_ptr.assign(o.row_ptr, o.row_ptr + o.n + 1);
      const int nnz = o.row_ptr[o.n];
      d.col.assign(o.col, o.col + nnz);
      d.val.assign(o.entries, o.entries + nnz);
      if(l > 0) {
        PMCD_CHECK(o.p_row_ptr && o.p_col && o.p_entries, "level ", l, ": no prolongation");
        d.n_coarse = o.n_coarse;
        d.p_row_ptr.assign(o.p_row_ptr, o.p_row_ptr + o.n + 1);
        const int pn = o.p_row_ptr[o.n];
        d.p_col.assign(o.p_col, o.p_col + pn);
        d.p_val.assign(o.p_entries, o.p_entries + pn);
      }
    }
    s->sys.reset();
    s->sys = std::make_unique<CDSystem>(lv, s->stream);
  });
}

int pmcd_solver_solve(pmcd_solver* s, const double* b, double* x, pmcd_solve_stats* st) {
  return guarded_on(s, [&] {
    PMCD_CHECK(s && b && x, "null argument");
    PMCD_CHECK(s->sys, "pmcd_solver_update_matrix has not been called");
    CDSystem& sys = *s->sys;
    Level& F = sys.finest();
    sys.set_rhs(b);
    F.sol.upload(x, F.sol.n, s->stream);
    const SolveStats r = sys.solve(s->sd);
    F.sol.download(x, F.sol.n, s->stream);
    PMCD_HIP(hipStreamSynchronize(s->stream));
    if(st) {
      st->iterations = r.iterations;
      st->initial_residual = r.initial_residual;
      st->final_residual = r.final_residual;
      st->converged = r.converged;
      st->seconds = r.seconds;
    }
  });
}

int pmcd_solver_is_using_multigrid(const pmcd_solver* s, int* out) {
  return guarded([&] {
    PMCD_CHECK(s && out, "null argument");
    *out = s->sd.precond == 2 ? 1 : 0;
  });
}

int pmcd_get_transfer(pmcd_system* s, int level, int* n_fine, int* n_coarse, int* nnz, int* row_ptr, int* col,
                      double* val) {
  return guarded_on(s, [&] {
    Level& L = level_of(s, level);
    PMCD_CHECK(L.n_coarse > 0, "level ", level, " has no coarser level");
    const int nf = L.n();
    const int pn = (int)L.p_col.n;
    if(n_fine) *n_fine = nf;
    if(n_coarse) *n_coarse = L.n_coarse;
    if(nnz) *nnz = pn;
    hipStream_t st = s->stream;
    if(row_ptr) L.p_rp.download(row_ptr, nf + 1, st);
    if(col) L.p_col.download(col, pn, st);
    if(val) L.p_val.download(val, pn, st);
    PMCD_HIP(hipStreamSynchronize(st));
  });
}

int pmcd_axpy_device(int n, double a, const double* x, double* y, void* stream) {
  return guarded([&] {
    PMCD_CHECK(n >= 0, "negative size");
    launch_axpy(n, a, x, y, (hipStream_t)stream);
    PMCD_HIP(hipGetLastError());
  });
}

int pmcd_norm_device(int n, const double* x, double* result, void* stream) {
  return guarded([&] {
    PMCD_CHECK(result, "null output");
    double d = 0;
    if(pmcd_dot_device(n, x, x, &d, stream) != 0) PMCD_FAIL(g_err);
    *result = std::sqrt(d);
  });
}

int pmcd_jacobi_device(int n, int n_active, const int* row_ptr, const int* col, const double* entries,
                       const double* diag, const double* b, const double* x, double* x_out, double omega,
                       void* stream) {
  return guarded([&] {
    launch_jacobi(n, n_active, row_ptr, col, entries, nullptr, diag, b, x, x_out, omega, (hipStream_t)stream);
    PMCD_HIP(hipGetLastError());
  });
}

int pmcd_prolongate_device(int n_fine, int n_active_fine, const int* p_row_ptr, const int* p_col,
                           const double* p_entries, const double* coarse, double* fine, double a, void* stream) {
  return guarded([&] {
    launch_prolong_add(n_fine, n_active_fine, 0, -1, p_row_ptr, p_col, p_entries, coarse, fine, a,
                       (hipStream_t)stream);
    PMCD_HIP(hipGetLastError());
  });
}

int pmcd_restrict_device(int n_coarse, int n_active_coarse, const int* r_row_ptr, const int* r_col,
                         const double* r_entries, const double* fine, double* coarse, void* stream) {
  return guarded([&] {
    launch_restrict(n_coarse, n_active_coarse, 0, -1, r_row_ptr, r_col, r_entries, fine, coarse,
                    (hipStream_t)stream);
    PMCD_HIP(hipGetLastError());
  });
}


int pmcd_batch_instances(unsigned long long seed, int n_inst, int n_lo, int n_hi, int* n_out, double* eps_out,
                         double* b_out) {
  return guarded([&] {
    PMCD_CHECK(n_inst >= 0 && n_lo >= 1 && n_hi >= n_lo && n_out && eps_out && b_out, "invalid argument");
    std::mt19937_64 gen(seed);
    std::uniform_int_distribution<int> cells(n_lo, n_hi);
    std::uniform_real_distribution<double> expo(-6.0, 0.0);
    std::normal_distribution<double> normal(0.0, 1.0);
    for(int i = 0; i < n_inst; ++i) {
      n_out[i] = cells(gen);
      eps_out[i] = std::pow(10.0, expo(gen));
      double g[3];
      for(double& v : g) v = normal(gen);
      const double len = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
      for(int k = 0; k < 3; ++k) b_out[3 * i + k] = g[k] / len;
    }
  });
}

int pmcd_system_device_bytes(pmcd_system* s, long long* bytes, long long* peak) {
  return guarded([&] {
    PMCD_CHECK(s && s->sys, "null system");
    if(bytes) *bytes = s->tally.bytes.load();
    if(peak) *peak = s->tally.peak.load();
  });
}

int pmcd_estimate_device_bytes(const pmcd_problem_desc* pd, const pmcd_solver_desc* sd, long long* bytes) {
  return guarded([&] {
    PMCD_CHECK(pd && sd && bytes, "null argument");
    *bytes = estimate_device_bytes(to_desc(*pd), to_solver(*sd));
  });
}

int pmcd_estimate_counts(const pmcd_problem_desc* pd, long long* out3) {
  return guarded([&] {
    PMCD_CHECK(pd && out3, "null argument");
    const LevelEstimate e = estimate_levels(to_desc(*pd)).back();
    out3[0] = (long long)e.cells;
    out3[1] = (long long)e.dofs;
    out3[2] = (long long)e.entries;
  });
}

int pmcd_autotune(pmcd_system* s, int level, int iters, int mixed, double* ms_out, int* chosen_out) {
  return guarded_on(s, [&] {
    PMCD_CHECK(s && s->sys && ms_out && chosen_out, "null argument");
    autotune(*s->sys, level, iters, mixed != 0, ms_out, chosen_out);
  });
}

int pmcd_server_create(int max_slots, long long budget_bytes, int keep_solution, pmcd_server** out) {
  return guarded([&] {
    PMCD_CHECK(out, "null argument");
    check_visible_gpu();
    auto h = std::make_unique<pmcd_server>();
    h->srv = std::make_unique<Server>(max_slots, budget_bytes, keep_solution != 0);
    *out = h.release();
  });
}

int pmcd_server_destroy(pmcd_server* s) {
  return guarded([&] {
    if(!s) return;
    s->srv.reset();
    delete s;
  });
}

int pmcd_server_now(pmcd_server* s, double* t) {
  return guarded([&] {
    PMCD_CHECK(s && s->srv && t, "null argument");
    *t = s->srv->now();
  });
}

int pmcd_server_submit(pmcd_server* s, const pmcd_problem_desc* pd, const pmcd_solver_desc* sd, double arrival,
                       long long* id) {
  return guarded([&] {
    PMCD_CHECK(s && s->srv && pd && sd && id, "null argument");
    *id = s->srv->submit(to_desc(*pd), to_solver(*sd), arrival);
  });
}

int pmcd_server_warm(pmcd_server* s, const pmcd_problem_desc* pd, const pmcd_solver_desc* sd, int count) {
  return guarded([&] {
    PMCD_CHECK(s && s->srv && pd && sd, "null argument");
    s->srv->warm(to_desc(*pd), to_solver(*sd), count);
  });
}

int pmcd_server_wait(pmcd_server* s, long long id, double timeout_s, int* done) {
  return guarded([&] {
    PMCD_CHECK(s && s->srv, "null argument");
    const bool d = s->srv->wait(id, timeout_s);
    if(done) *done = d ? 1 : 0;
  });
}

int pmcd_server_result(pmcd_server* s, long long id, pmcd_served_result* out) {
  return guarded([&] {
    PMCD_CHECK(s && s->srv && out, "null argument");
    const ServedRequest r = s->srv->result(id);
    std::memset(out, 0, sizeof(*out));
    out->id = r.id;
    out->status = r.status;
    out->reused = r.reused;
    out->arrival = r.arrival;
    out->admitted = r.admitted;
    out->finished = r.finished;
    out->assemble_ms = r.assemble_ms;
    out->n_dof = r.n_dof;
    out->reserved_bytes = r.reserved;
    out->stats.iterations = r.stats.iterations;
    out->stats.converged = r.stats.converged;
    out->stats.initial_residual = r.stats.initial_residual;
    out->stats.final_residual = r.stats.final_residual;
    out->stats.seconds = r.stats.seconds;
    std::strncpy(out->error, r.error.c_str(), sizeof(out->error) - 1);
  });
}

int pmcd_server_solution(pmcd_server* s, long long id, double* out, long long n) {
  return guarded([&] {
    PMCD_CHECK(s && s->srv && out, "null argument");
    s->srv->solution(id, out, n);
  });
}

int pmcd_server_release(pmcd_server* s, long long id) {
  return guarded([&] {
    PMCD_CHECK(s && s->srv, "null argument");
    s->srv->release(id);
  });
}

int pmcd_server_info(pmcd_server* s, long long* out8) {
  return guarded([&] {
    PMCD_CHECK(s && s->srv && out8, "null argument");
    s->srv->info(out8);
  });
}

}  // extern "C"
