// Request loop for convection-diffusion solves of mixed sizes on one GPU,
// and the launch-shape autotuner.
//
// The reference has no request loop: each problem is a new
// ConvectionDiffusion<d> (src/System/ConvectionDiffusion.C:112-452) followed
// by assemble(), solve() and output() (3DPrograms/CD3D_ParMooN.C:40-53,
// 2DPrograms/CD2D_ParMooN.C:19-22). A served request does exactly the
// assemble and solve steps, on a device-resident system that outlives it.
//
// Continuous batching: requests wait in a FIFO queue. A free worker admits
// the head request when a slot is free and either an idle resident system
// for its mesh exists or its estimated device bytes fit the budget (idle
// systems are destroyed, oldest first, to make room; a request larger than
// the whole budget runs once nothing else is running). Every worker owns a
// HIP stream per resident system, so admitted requests overlap on the
// device. A finished request parks its system in the pool keyed by the mesh;
// a later request on the same mesh only replaces the PDE data, reassembles
// and solves.
#pragma once
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>
#include "system.hpp"

namespace pmcd {

// Upper estimate of the device bytes a system for `pd` holds after a solve
// with `sd`: every per-level buffer (mesh and numbering arrays, CSR matrix
// and its FP32, sliced-ELL and sweep-ordered copies, transfer operators,
// multigrid vectors), the per-cell assembly scratch and the solver's
// vectors, with each buffer rounded to a 4 KiB page. DoF counts follow the
// regular-refinement recurrences of the coarse mesh's vertex, edge, face and
// cell counts; matrix entries are bounded by min(cells * local^2, DoF * the
// largest row of a regular refinement).
long long estimate_device_bytes(const ProblemDesc& pd, const SolverDesc& sd);

// Counts behind the estimate, per kept level (coarsest first): cells and
// DoFs exactly (regular refinement of the coarse mesh), matrix entries as
// the bound above.
struct LevelEstimate {
  double cells = 0, dofs = 0, entries = 0;
};
std::vector<LevelEstimate> estimate_levels(const ProblemDesc& pd, Shape* shape_out = nullptr);

struct ServedRequest {
  long long id = 0;
  ProblemDesc pd{};
  SolverDesc sd{};
  double arrival = -1, admitted = -1, finished = -1;  // seconds on the server clock
  double assemble_ms = 0;
  long long n_dof = 0;
  long long reserved = 0;  // device bytes reserved at admission
  int reused = 0;          // 1: the system came from the resident pool
  SolveStats stats{};
  int status = 0;          // 0 solved, 1 failed
  std::string error;
  std::vector<double> solution;  // finest-level solution when the server keeps solutions
  bool done = false;
};

class Server {
 public:
  // budget <= 0: 90% of the device memory free at construction
  Server(int max_slots, long long budget, bool keep_solution);
  ~Server();  // runs every queued request, then destroys the resident systems
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;

  double now() const;
  long long submit(const ProblemDesc& pd, const SolverDesc& sd, double arrival);
  // create and park up to `count` systems for this mesh (set-up outside a measured stream)
  void warm(const ProblemDesc& pd, const SolverDesc& sd, int count);
  // true once request `id` has finished; waits at most timeout_s (< 0: no limit)
  bool wait(long long id, double timeout_s);
  // copy of a finished request's record (solution left out)
  ServedRequest result(long long id);
  void solution(long long id, double* out, long long n);
  void release(long long id);
  // queued, running, reserved bytes, idle bytes, idle systems, budget,
  // finished requests, systems created
  void info(long long out[8]);

 private:
  struct Resident;
  struct Admission {
    ServedRequest* req = nullptr;
    std::unique_ptr<Resident> res;
    long long need = 0;
  };
  bool try_admit_locked(Admission& a, std::vector<std::unique_ptr<Resident>>& evicted);
  bool make_room_locked(long long need, std::vector<std::unique_ptr<Resident>>& evicted);
  std::unique_ptr<Resident> take_idle_locked(const std::string& key);
  void park_locked(std::unique_ptr<Resident> r);
  void worker();
  void run(Admission& a);

  int device_ = 0;
  int max_slots_;
  long long budget_;
  bool keep_solution_;
  std::chrono::steady_clock::time_point t0_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::deque<ServedRequest*> queue_;
  std::map<long long, std::unique_ptr<ServedRequest>> requests_;
  std::vector<std::unique_ptr<Resident>> idle_;  // oldest first
  long long idle_bytes_ = 0, reserved_ = 0;
  int running_ = 0;
  long long next_id_ = 0, finished_ = 0, created_ = 0;
  bool closed_ = false;
  std::vector<std::thread> workers_;
};

// Launch-shape autotuner for the SpMV and Jacobi kernels of one level: times
// every CSR variant (0 plain, 1 streaming, 2 unrolled, 3 unrolled with
// non-temporal loads) at 8, 16 and 32 lanes per row, then the sliced ELL
// kernels, with HIP events, and sets the fastest. When sliced ELL wins, the
// coarser levels on which it loses keep the CSR kernels (sell_min_rows).
// `mixed` times the Jacobi sweep in FP32 (the FP32 copies must exist, i.e.
// the system has solved with the mixed V-cycle once).
// ms[26]: spmv candidates 0..12 then jacobi candidates 0..12, candidate
// 3 * variant + lane index (lanes 8, 16, 32), candidate 12 = sliced ELL.
// chosen[5]: spmv variant, spmv lanes, jacobi variant, jacobi lanes,
// sell_min_rows.
void autotune(CDSystem& sys, int level, int iters, bool mixed, double* ms, int* chosen);

}  // namespace pmcd
