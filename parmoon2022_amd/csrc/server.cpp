// Request loop (continuous batching over device-resident systems) and the
// launch-shape autotuner.
#include "server.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <set>
#include "elements.hpp"
#include "kernels.hpp"
#include "mesh.hpp"

namespace pmcd {

// ---------------------------------------------------------------------------
// device-memory estimate

namespace {

struct Topology {
  double v = 0, e = 0, f = 0, c = 0;  // vertices, edges, faces (3D), cells
};

Topology coarse_topology(const Mesh& m) {
  Topology t;
  t.v = m.n_vertices();
  t.c = m.n_cells;
  const auto& joints = shape_faces(m.shape);
  std::set<std::pair<int, int>> edges;
  std::set<std::array<int, 4>> faces;
  for(int c = 0; c < m.n_cells; ++c) {
    const int* cv = m.cell(c);
    for(const auto& j : joints) {
      const int k = (int)j.size();
      // a joint's boundary cycle gives the cell's edges (every edge of these
      // shapes lies on a face)
      for(int a = 0; a < (k == 2 ? 1 : k); ++a) {
        int x = cv[j[a]], y = cv[j[(a + 1) % k]];
        if(x > y) std::swap(x, y);
        edges.insert({x, y});
      }
      if(k > 2) {
        std::array<int, 4> key{-1, -1, -1, -1};
        for(int a = 0; a < k; ++a) key[a] = cv[j[a]];
        std::sort(key.begin(), key.begin() + k);
        faces.insert(key);
      }
    }
  }
  t.e = (double)edges.size();
  t.f = (double)faces.size();
  return t;
}

// one step of regular refinement (every cell into 4 or 8 children)
Topology refine_counts(const Topology& k, Shape s) {
  Topology r;
  switch(s) {
    case Shape::Triangle:
      r.v = k.v + k.e; r.e = 2 * k.e + 3 * k.c; r.c = 4 * k.c; break;
    case Shape::Quadrangle:
      r.v = k.v + k.e + k.c; r.e = 2 * k.e + 4 * k.c; r.c = 4 * k.c; break;
    case Shape::Tetrahedron:
      r.v = k.v + k.e; r.e = 2 * k.e + 3 * k.f + k.c; r.f = 4 * k.f + 8 * k.c; r.c = 8 * k.c; break;
    default:
      r.v = k.v + k.e + k.f + k.c; r.e = 2 * k.e + 4 * k.f + 6 * k.c; r.f = 4 * k.f + 12 * k.c;
      r.c = 8 * k.c; break;
  }
  return r;
}

double dof_count(const Topology& t, Shape s, int order) {
  if(order == 1) return t.v;
  switch(s) {
    case Shape::Triangle: case Shape::Tetrahedron: return t.v + t.e;
    case Shape::Quadrangle: return t.v + t.e + t.c;
    default: return t.v + t.e + t.f + t.c;
  }
}

// largest row of the matrix on a regular refinement of the built-in meshes
int row_cap(Shape s, int order) {
  switch(s) {
    case Shape::Triangle: return order == 1 ? 7 : 19;
    case Shape::Quadrangle: return order == 1 ? 9 : 25;
    case Shape::Tetrahedron: return order == 1 ? 27 : 65;
    default: return order == 1 ? 27 : 125;
  }
}

Mesh coarse_mesh(const ProblemDesc& d) {
  if(d.geometry == 10 || d.geometry == 11) return structured_cube(d.structured_n, d.geometry == 11);
  if(d.geometry == 12) {
    const std::string& f = d.mesh_file;
    const bool geo = f.size() > 4 && (f.compare(f.size() - 4, 4, ".GEO") == 0 || f.compare(f.size() - 4, 4, ".geo") == 0);
    return geo ? read_geo_file(f) : read_mesh_file(f);
  }
  PMCD_CHECK(d.geometry >= 0 && d.geometry <= 4, "unknown geometry ", d.geometry);
  return default_mesh((DefaultGeometry)d.geometry);
}

// device bytes of one allocation of `count` elements of `size` bytes
double page(double count, double size) {
  const double b = count * size;
  return b <= 0 ? 0 : std::ceil(b / 4096.0) * 4096.0;
}

}  // namespace

std::vector<LevelEstimate> estimate_levels(const ProblemDesc& pd, Shape* shape_out) {
  PMCD_CHECK(pd.order == 1 || pd.order == 2, "ansatz order must be 1 or 2");
  PMCD_CHECK(pd.n_refine >= 0 && pd.n_refine <= 12, "n_refine out of range");
  if(pd.adapt_steps > 0) {
    // no closed form for a refinement by indicator: count the host set-up
    const auto levels = build_levels(pd);
    const Level& L = *levels.back();
    if(shape_out) *shape_out = L.mesh.shape;
    LevelEstimate e;
    e.cells = L.mesh.n_cells;
    e.dofs = L.n();
    e.entries = (double)L.st.nnz;
    return {e};
  }
  const Mesh base = coarse_mesh(pd);
  const Shape shape = base.shape;
  if(shape_out) *shape_out = shape;
  const bool file_mesh = pd.geometry == 12;
  const double nl = element_n_dof(element_for(shape, pd.order));
  std::vector<Topology> topo{coarse_topology(base)};
  for(int i = 0; i < pd.n_refine; ++i) topo.push_back(refine_counts(topo.back(), shape));
  const int n_meshes = (int)topo.size();
  const int keep = std::max(1, std::min(pd.mg_levels, n_meshes));
  std::vector<LevelEstimate> out;
  for(int l = n_meshes - keep; l < n_meshes; ++l) {
    LevelEstimate e;
    e.cells = topo[l].c;
    e.dofs = dof_count(topo[l], shape, pd.order);
    e.entries = e.cells * nl * nl;
    if(!file_mesh) e.entries = std::min(e.entries, e.dofs * row_cap(shape, pd.order));
    if(pd.order == 1) {
      // vertex pairs on a common cell: the diagonal, both directions of every
      // edge, and of every face diagonal and body diagonal of a quadrilateral
      // or hexahedron (each on one face or cell); Dirichlet rows keep at most
      // as many
      const Topology& t = topo[l];
      double pairs = t.e;
      if(shape == Shape::Quadrangle) pairs += 2 * t.c;
      if(shape == Shape::Hexahedron) pairs += 2 * t.f + 4 * t.c;
      e.entries = std::min(e.entries, t.v + 2 * pairs);
    }
    out.push_back(e);
  }
  return out;
}

long long estimate_device_bytes(const ProblemDesc& pd, const SolverDesc& sd) {
  Shape shape;
  const std::vector<LevelEstimate> lv = estimate_levels(pd, &shape);
  const double nl = element_n_dof(element_for(shape, pd.order)), nv = shape_n_vertices(shape);
  const int keep = (int)lv.size();
  const bool mixed = sd.mixed != 0;
  const bool mg = sd.precond == 2 && keep > 1;
  const bool sweeps_all = mg && sd.smoother != 0;
  double total = 0, max_cells = 0, prev_n = 0, n_finest = 0, n_coarsest = 0;
  for(int l = 0; l < keep; ++l) {
    const double C = lv[l].cells, n = lv[l].dofs, nnz = lv[l].entries;
    const bool coarsest = l == 0, finest = l == keep - 1;
    if(coarsest) n_coarsest = n;
    if(finest) n_finest = n;
    max_cells = std::max(max_cells, C);
    double b = 0;
    // mesh, numbering, structure and assembly maps
    b += page(C * nv * 3, 8) + page(C * nl, 4);
    b += page(n + 1, 4) + page(nnz, 4) + page(n, 4) + page(nnz + 1, 4) + page(C * nl * nl, 4) + page(C * nl * nl, 1);
    b += page(n, 8) + page(C * nl + 1, 4) + page(C * nl * 4, 8);  // Dirichlet values, boundary-integral terms
    // matrix, vectors, multigrid vectors
    b += page(nnz, 8) + 3 * page(n, 8) + 5 * page(n, 8);
    if(mixed) b += page(nnz, 4) + page(n, 4) + 5 * page(n, 4);
    // transfer from the coarser level: at most nl coarse DoFs per fine DoF
    if(!coarsest) {
      const double np = n * nl;
      b += page(n + 1, 4) + page(prev_n + 1, 4) + 2 * (page(np, 4) + page(np, 8) + page(np, 4));
      // P's sliced ELL copy (padded entries: columns, FP64 and FP32 values,
      // 16-bit footprint positions; footprints; slot arrays)
      const double ps = 1.25 * np + 64 * 64;
      b += page(ps, 4) + page(ps, 8) + page(ps, 4) + page(ps, 2) + page(0.5 * np + 256, 4) + 3 * page(n + 64, 4) +
           page(n / 64 + 2, 8);
    }
    // sweep-ordered copy (SOR/SSOR smoothing or preconditioning)
    const bool sweep = sweeps_all || (mg && coarsest && sd.coarse_smoother >= 2) ||
                       (finest && (sd.precond == 3 || sd.precond == 4 || sd.method == 4 || sd.method == 5));
    if(sweep) b += 3 * page(n + 1, 4) + page(nnz, 4) + page(nnz, 8) + (mixed ? page(nnz, 4) : 0);
    // sliced ELL copy: rows padded to the longest row of each 64-row slice
    const double slots = 1.25 * nnz + 64 * 64;
    b += 3 * page(n, 4) + page(n / 64 + 2, 8) + page(slots, 4) + page(slots, 8) + (mixed ? page(slots, 4) : 0);
    // its block column footprints (x staged in LDS): a 16-bit position per
    // slot entry, the footprint columns (at most one per entry; a fraction
    // on meshes whose neighbouring rows share columns) and block offsets
    b += page(slots, 2) + page(0.5 * nnz + 256, 4) + page(n / 256 + 2, 4);
    // hanging-DOF coupling lists: at most one item per entry and two sources
    // per item (matrix), one item and two sources per DOF (rhs), the
    // constraint entries
    if(pd.adapt_steps > 0)
      b += 2 * page(nnz + 1, 4) + page(2 * nnz, 4) + page(2 * nnz, 8) + 2 * page(n + 1, 4) + page(2 * n, 4) +
           page(2 * n, 8) + page(nnz, 4) + page(nnz, 8);
    total += b;
    prev_n = n;
  }
  // local matrices and right-hand sides, per-cell errors, reductions;
  // reference tables (basis values at up to 128 points, weights, points,
  // the component-pair tensors)
  total += page(max_cells * nl * nl, 8) + page(max_cells * nl, 8) + page(max_cells * 5, 8);
  total += page(1024 * 5 + 64, 8) + page(64, 8);
  total += page(128 * 10 * nl, 8) + page(128, 8) + page(128 * 3, 8) + page(36 * (std::ceil(nl * nl / 16) * 16), 8);
  total += 64 * 1024;
  // solver
  const double n = n_finest;
  switch(sd.method) {
    case 2: {
      const double m = std::max(1, sd.restart);
      total += (2 * m + 3) * page(n, 8) + page(m + 2, 8) * 3 + page(2 * kMaxReduceBlocks, 8);
      break;
    }
    case 1: total += 9 * page(n, 8) + page(kCgBlocks, 8) + page(2 * kCgUpdateBlocks, 8) + page(64, 8); break;
    default: total += 3 * page(n, 8); break;
  }
  if(sd.method == 0 || (mg && sd.coarse_smoother == 0)) {
    const double nc = sd.method == 0 ? n : n_coarsest;
    total += 2 * page(nc * nc, 8) + 2 * page(2 * nc, 8) + 2 * page(nc, 4) + page(nc, 8);
  }
  return (long long)(total * 1.02);
}

// ---------------------------------------------------------------------------
// server

namespace {

std::string mesh_key(const ProblemDesc& d) {
  std::ostringstream os;
  os << d.geometry << '|' << d.n_refine << '|' << d.structured_n << '|' << d.order << '|' << d.supg << '|'
     << d.linear_quad << '|' << d.mg_levels << '|' << d.mesh_file << '|' << d.adapt_steps << '|'
     << (const void*)d.indicator << '|' << d.indicator_ctx << '|' << d.adapt_conforming;
  return os.str();
}

// solver parameters that decide which buffers a solve allocates
std::string solver_key(const SolverDesc& s) {
  std::ostringstream os;
  os << s.method << '|' << s.precond << '|' << s.restart << '|' << s.mixed << '|' << s.smoother << '|'
     << s.coarse_smoother;
  return os.str();
}

}  // namespace

struct Server::Resident {
  std::string key;
  std::string solved_with;  // solver_key of the last solve
  hipStream_t stream = nullptr;
  DevTally tally;  // outlives sys (declared first)
  std::unique_ptr<CDSystem> sys;
  long long bytes() const { return tally.peak.load(); }
  ~Resident() {
    sys.reset();
    if(stream) (void)hipStreamDestroy(stream);
  }
};

Server::Server(int max_slots, long long budget, bool keep_solution)
    : max_slots_(max_slots), budget_(budget), keep_solution_(keep_solution), t0_(std::chrono::steady_clock::now()) {
  PMCD_CHECK(max_slots >= 1 && max_slots <= 64, "max_slots must be in 1..64, got ", max_slots);
  PMCD_HIP(hipGetDevice(&device_));
  if(budget_ <= 0) {
    size_t f = 0, t = 0;
    PMCD_HIP(hipMemGetInfo(&f, &t));
    budget_ = (long long)(0.9 * (double)f);
  }
  for(int i = 0; i < max_slots_; ++i) workers_.emplace_back([this] { worker(); });
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    closed_ = true;
  }
  cv_.notify_all();
  for(auto& t : workers_) t.join();
  idle_.clear();
}

double Server::now() const {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

long long Server::submit(const ProblemDesc& pd, const SolverDesc& sd, double arrival) {
  auto r = std::make_unique<ServedRequest>();
  r->pd = pd;
  r->sd = sd;
  r->arrival = arrival >= 0 ? arrival : now();
  long long id;
  {
    std::lock_guard<std::mutex> lk(mu_);
    PMCD_CHECK(!closed_, "server is closed");
    id = r->id = next_id_++;
    queue_.push_back(r.get());
    requests_[id] = std::move(r);
  }
  cv_.notify_all();
  return id;
}

void Server::warm(const ProblemDesc& pd, const SolverDesc& sd, int count) {
  const std::string key = mesh_key(pd);
  for(int i = 0; i < count; ++i) {
    const long long need = estimate_device_bytes(pd, sd);
    std::vector<std::unique_ptr<Resident>> evicted;
    {
      std::lock_guard<std::mutex> lk(mu_);
      if(!make_room_locked(need, evicted)) return;
      reserved_ += need;
    }
    evicted.clear();
    auto r = std::make_unique<Resident>();
    r->key = key;
    try {
      PMCD_HIP(hipStreamCreate(&r->stream));
      TallyScope scope(&r->tally);
      r->sys = std::make_unique<CDSystem>(pd, r->stream);
    } catch(...) {
      std::lock_guard<std::mutex> lk(mu_);
      reserved_ -= need;
      throw;
    }
    std::lock_guard<std::mutex> lk(mu_);
    reserved_ -= need;
    ++created_;
    park_locked(std::move(r));
  }
}

bool Server::make_room_locked(long long need, std::vector<std::unique_ptr<Resident>>& evicted) {
  while(reserved_ + idle_bytes_ + need > budget_ && !idle_.empty()) {
    idle_bytes_ -= idle_.front()->bytes();
    evicted.push_back(std::move(idle_.front()));
    idle_.erase(idle_.begin());
  }
  return reserved_ + idle_bytes_ + need <= budget_;
}

std::unique_ptr<Server::Resident> Server::take_idle_locked(const std::string& key) {
  for(size_t i = 0; i < idle_.size(); ++i)
    if(idle_[i]->key == key) {
      std::unique_ptr<Resident> r = std::move(idle_[i]);
      idle_.erase(idle_.begin() + i);
      idle_bytes_ -= r->bytes();
      return r;
    }
  return nullptr;
}

void Server::park_locked(std::unique_ptr<Resident> r) {
  idle_bytes_ += r->bytes();
  idle_.push_back(std::move(r));
}

// Admission of the head request: a resident system for its mesh, or room
// for a new one within the budget (or nothing else running).
bool Server::try_admit_locked(Admission& a, std::vector<std::unique_ptr<Resident>>& evicted) {
  if(queue_.empty() || running_ >= max_slots_) return false;
  ServedRequest* req = queue_.front();
  const std::string key = mesh_key(req->pd);
  std::unique_ptr<Resident> res = take_idle_locked(key);
  long long need;
  if(res) {
    // a system that already solved with these solver settings holds every
    // buffer it needs; otherwise the new solver may add buffers
    need = res->solved_with == solver_key(req->sd) ? res->bytes()
                                                     : std::max(res->bytes(), estimate_device_bytes(req->pd, req->sd));
    // the resident's bytes left the idle tally when it was taken
    if(!make_room_locked(need, evicted) && running_ > 0) {
      park_locked(std::move(res));
      return false;
    }
  } else {
    need = estimate_device_bytes(req->pd, req->sd);
    if(!make_room_locked(need, evicted) && running_ > 0) return false;
  }
  queue_.pop_front();
  ++running_;
  reserved_ += need;
  req->admitted = now();
  req->reserved = need;
  req->reused = res ? 1 : 0;
  a.req = req;
  a.res = std::move(res);
  a.need = need;
  return true;
}

void Server::worker() {
  (void)hipSetDevice(device_);
  for(;;) {
    Admission a;
    std::vector<std::unique_ptr<Resident>> evicted;
    {
      std::unique_lock<std::mutex> lk(mu_);
      for(;;) {
        bool ok = false;
        try {
          ok = try_admit_locked(a, evicted);
        } catch(const std::exception& e) {
          // the estimate failed (e.g. an unreadable mesh file): fail the request
          ServedRequest* req = queue_.front();
          queue_.pop_front();
          req->status = 1;
          req->error = e.what();
          req->finished = now();
          req->done = true;
          ++finished_;
          cv_.notify_all();
          continue;
        }
        if(ok || (closed_ && queue_.empty())) break;
        if(!evicted.empty()) {
          lk.unlock();
          evicted.clear();  // frees device memory outside the lock
          lk.lock();
          continue;
        }
        cv_.wait(lk);
      }
    }
    evicted.clear();
    if(!a.req) return;
    run(a);
  }
}

void Server::run(Admission& a) {
  ServedRequest* req = a.req;
  std::unique_ptr<Resident> res = std::move(a.res);
  try {
    if(!res) {
      res = std::make_unique<Resident>();
      res->key = mesh_key(req->pd);
      PMCD_HIP(hipStreamCreate(&res->stream));
      TallyScope scope(&res->tally);
      res->sys = std::make_unique<CDSystem>(req->pd, res->stream);
      std::lock_guard<std::mutex> lk(mu_);
      ++created_;
    } else {
      TallyScope scope(&res->tally);
      res->sys->set_example(req->pd.ex);
    }
    TallyScope scope(&res->tally);
    auto t = std::chrono::steady_clock::now();
    res->sys->assemble();
    PMCD_HIP(hipStreamSynchronize(res->stream));
    req->assemble_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
    req->stats = res->sys->solve(req->sd);
    res->solved_with = solver_key(req->sd);
    Level& F = res->sys->level(res->sys->n_levels() - 1);
    req->n_dof = F.n();
    if(keep_solution_) {
      req->solution.resize(F.n());
      F.sol.download(req->solution.data(), F.n(), res->stream);
    }
    PMCD_HIP(hipStreamSynchronize(res->stream));
    req->status = 0;
  } catch(const std::exception& e) {
    req->status = 1;
    req->error = e.what();
  }
  std::unique_ptr<Resident> drop;
  {
    std::lock_guard<std::mutex> lk(mu_);
    req->finished = now();
    --running_;
    reserved_ -= a.need;
    if(req->status == 0) park_locked(std::move(res));
    else drop = std::move(res);
    req->done = true;
    ++finished_;
  }
  drop.reset();
  cv_.notify_all();
}

bool Server::wait(long long id, double timeout_s) {
  std::unique_lock<std::mutex> lk(mu_);
  auto it = requests_.find(id);
  PMCD_CHECK(it != requests_.end(), "unknown request ", id);
  ServedRequest* r = it->second.get();
  if(timeout_s < 0) {
    cv_.wait(lk, [&] { return r->done; });
    return true;
  }
  return cv_.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] { return r->done; });
}

ServedRequest Server::result(long long id) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = requests_.find(id);
  PMCD_CHECK(it != requests_.end(), "unknown request ", id);
  PMCD_CHECK(it->second->done, "request ", id, " has not finished");
  ServedRequest out;
  const ServedRequest& r = *it->second;
  out.id = r.id; out.pd = r.pd; out.sd = r.sd;
  out.arrival = r.arrival; out.admitted = r.admitted; out.finished = r.finished;
  out.assemble_ms = r.assemble_ms; out.n_dof = r.n_dof; out.reserved = r.reserved; out.reused = r.reused;
  out.stats = r.stats; out.status = r.status; out.error = r.error; out.done = true;
  return out;
}

void Server::solution(long long id, double* out, long long n) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = requests_.find(id);
  PMCD_CHECK(it != requests_.end(), "unknown request ", id);
  const ServedRequest& r = *it->second;
  PMCD_CHECK(r.done && r.status == 0, "request ", id, " has no solution");
  PMCD_CHECK(keep_solution_, "the server does not keep solutions");
  PMCD_CHECK(n == (long long)r.solution.size(), "solution has ", r.solution.size(), " values, asked for ", n);
  std::memcpy(out, r.solution.data(), sizeof(double) * n);
}

void Server::release(long long id) {
  std::lock_guard<std::mutex> lk(mu_);
  auto it = requests_.find(id);
  PMCD_CHECK(it != requests_.end(), "unknown request ", id);
  PMCD_CHECK(it->second->done, "request ", id, " has not finished");
  requests_.erase(it);
}

void Server::info(long long out[8]) {
  std::lock_guard<std::mutex> lk(mu_);
  out[0] = (long long)queue_.size();
  out[1] = running_;
  out[2] = reserved_;
  out[3] = idle_bytes_;
  out[4] = (long long)idle_.size();
  out[5] = budget_;
  out[6] = finished_;
  out[7] = created_;
}

// ---------------------------------------------------------------------------
// autotuner

void autotune(CDSystem& sys, int level, int iters, bool mixed, double* ms, int* chosen) {
  PMCD_CHECK(iters > 0, "iters must be positive");
  const int nl = sys.n_levels();
  const int finest = level < 0 ? nl + level : level;
  PMCD_CHECK(finest >= 0 && finest < nl, "level ", level, " out of range");
  static const int lanes[3] = {8, 16, 32};
  set_tuning(11, 0);  // sliced ELL on every level while timing
  int csr_best[2][2];  // [kernel] -> (variant, lanes)
  int best[2][2];
  for(int k = 0; k < 2; ++k) {
    // kernel 0: SpMV (timed in FP64), kernel 1: Jacobi (FP32 under `mixed`)
    const int timed = k == 0 ? 0 : (mixed ? 5 : 1);
    double* res = ms + 13 * k;
    for(int v = 0; v < 4; ++v)
      for(int w = 0; w < 3; ++w) {
        set_tuning(k, lanes[w]);
        set_tuning(k + 2, v);
        sys.time_kernel(timed, finest, iters, &res[3 * v + w]);
      }
    set_tuning(k, 16);
    set_tuning(k + 2, 4);
    sys.time_kernel(timed, finest, iters, &res[12]);
    int ib = 0;
    for(int c = 1; c < 12; ++c)
      if(res[c] < res[ib]) ib = c;
    csr_best[k][0] = ib / 3;
    csr_best[k][1] = lanes[ib % 3];
    if(res[12] < res[ib]) {
      // under the sliced ELL variant the lane count is the one of the CSR
      // kernels that small levels fall back to
      best[k][0] = 4;
      best[k][1] = csr_best[k][1];
    } else {
      best[k][0] = csr_best[k][0];
      best[k][1] = csr_best[k][1];
    }
    set_tuning(k, best[k][1]);
    set_tuning(k + 2, best[k][0]);
  }
  // below `finest`, keep the sliced ELL kernels while they stay faster
  int threshold = 0;
  for(int k = 0; k < 2; ++k) {
    if(best[k][0] != 4) continue;
    const int timed = k == 0 ? 0 : (mixed ? 5 : 1);
    for(int l = finest - 1; l >= 0; --l) {
      double t_sell = 0, t_csr = 0;
      set_tuning(k + 2, 4);
      sys.time_kernel(timed, l, iters, &t_sell);
      set_tuning(k + 2, csr_best[k][0]);
      sys.time_kernel(timed, l, iters, &t_csr);
      set_tuning(k + 2, 4);
      if(t_sell > t_csr) {
        threshold = std::max(threshold, sys.level(l + 1).n_active());
        break;
      }
    }
  }
  set_tuning(11, threshold);
  chosen[0] = best[0][0];
  chosen[1] = best[0][1];
  chosen[2] = best[1][0];
  chosen[3] = best[1][1];
  chosen[4] = threshold;
}

}  // namespace pmcd
