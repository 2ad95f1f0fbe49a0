/*
 * C ABI of the MI355X convection-diffusion assemble-and-solve library
 * (libpmcd.so). Plain C types only; every entry point returns 0 on success
 * and a non-zero status on failure, with the message available from
 * pmcd_last_error() (the reference throws std::runtime_error through ErrThrow,
 * include/General/MooNMD_Io.h:435-447; exceptions never cross this boundary).
 *
 * Each entry point names the reference interface it stands in for.
 */
#ifndef PMCD_H
#define PMCD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pmcd_system pmcd_system;

/* Geometry codes: default coarse meshes of the reference
 * (src/Geometry/Tests.C:95,473,1893,1937), structured N^3 cubes, or a 3D
 * .mesh file; each refined n_refine times. */
enum {
  PMCD_UNIT_SQUARE = 0,
  PMCD_TWO_TRIANGLES = 1,
  PMCD_TWO_TRIANGLES2 = 2,
  PMCD_CUBE_HEXA = 3,
  PMCD_CUBE_TETRA = 4,
  PMCD_STRUCTURED_HEXA = 10,
  PMCD_STRUCTURED_TETRA = 11,
  PMCD_MESH_FILE = 12
};

/* Example codes (PDE data, include/Examples/CD_2D/SineLaplace.h,
 * include/Examples/CD_3D/test_p0.h, test_p1.h, test_p2.h and the custom
 * manufactured/constant-data examples of the benchmark configurations). */
enum {
  PMCD_EX_SINE_LAPLACE_2D = 0,
  PMCD_EX_SINE_2D = 1,
  PMCD_EX_SINE_3D = 2,
  PMCD_EX_TEST_P0_3D = 3,
  PMCD_EX_TEST_P1_3D = 4,
  PMCD_EX_TEST_P2_3D = 5,
  PMCD_EX_CONST_RHS = 6,
  PMCD_EX_TCD_LINEAR_3D = 7, /* include/Examples/TCD_3D/linear_space_time.h */
  PMCD_EX_TCD_QUAD_3D = 8,   /* include/Examples/TCD_3D/quadratic_space_time.h */
  PMCD_EX_TCD_EXP_2D = 9,    /* include/Examples/TCD_2D/exp.h; eps = 1 / Reynolds number */
  PMCD_EX_SHARP_LAYER_2D = 10, /* include/Examples/CD_2D/SharpBoundaryLayer.h */
  PMCD_EX_SMOOTH_3D = 11     /* include/Examples/CD_3D/smooth_solution_3d.h; eps = diffusion_coefficient */
};

typedef struct {
  int geometry;     /* PMCD_* geometry code */
  int n_refine;     /* uniform refinement steps (refinement_n_initial_steps) */
  int structured_n; /* cells per edge for PMCD_STRUCTURED_* */
  int order;        /* ANSATZ_ORDER: 1 or 2 */
  int supg;         /* space_discretization_type: 0 galerkin, 1 supg */
  int linear_quad;  /* INTERNAL_PROBLEM_LINEAR: quadrature degree 2p instead of 3p */
  int mg_levels;    /* multigrid_n_levels (1: finest level only) */
  int example;      /* PMCD_EX_* */
  double eps;       /* diffusion coefficient */
  double b[3];      /* convection */
  double c;         /* reaction (PMCD_EX_CONST_RHS) */
  double f0;        /* right-hand side (PMCD_EX_CONST_RHS) */
  const char* mesh_file; /* PMCD_MESH_FILE: path of a 3D medit .mesh file (geo_file ending in
                            .mesh, read as TDomain::InitFromMesh, src/Geometry/Domain.C:257-270,
                            878-992, 2912-3330); NULL otherwise */
  /* Adaptive refinement: after the n_refine regular refinements,
   * adapt_steps times TDomain::RefineByIndicator(indicator)
   * (src/Geometry/Domain.C:1094-1207, test/adaptive_hanging_test.c++:
   * 120-135): cells whose corner values of indicator(x, indicator_ctx)
   * (x = {x, y, z}) lie on both sides of zero are refined regularly, then the
   * mesh is closed to a 1-irregular one. With adapt_conforming = 0
   * (conforming_closure = false) the hanging vertices' DOFs are coupled to
   * their coarse edges as in Assemble2D.C:3350-3427. 2D, order 1 or 2,
   * mg_levels 1; steady problems only
   * (no AFC, no time stepping, no batches). The indicator runs on the host
   * during pmcd_system_create / pmcd_host_create (and on a server worker for
   * a served request, so the context must stay valid until it finishes).
   * adapt_steps 0: regular refinement only (indicator fields unused). */
  int adapt_steps;
  double (*indicator)(const double* x, void* ctx);
  void* indicator_ctx;
  /* conforming_closure = true (the reference's default; triangles only):
   * TDomain::MakeConfClosure (Domain.C:1219-1360) instead of hanging
   * vertices: before each step the last closure's bisections are removed,
   * after it triangles with two or three refined edges are refined
   * regularly and triangles with one are bisected on it. */
  int adapt_conforming;
} pmcd_problem_desc;

/* solver_type / iterative_solver_type: direct, cg, fgmres, richardson, and the
 * SOR / SSOR iterations (Iteration_sor, src/Solver/Iteration_sor.C:27-120) */
enum { PMCD_DIRECT = 0, PMCD_CG = 1, PMCD_FGMRES = 2, PMCD_RICHARDSON = 3, PMCD_SOR_ITER = 4,
       PMCD_SSOR_ITER = 5 };
/* preconditioner: no_preconditioner, jacobi, multigrid, sor, ssor
 * (Solver.C:214-241; SOR/SSOR apply one sweep from zero, Iteration_sor.C:125-140) */
enum { PMCD_NO_PRECOND = 0, PMCD_JACOBI = 1, PMCD_MULTIGRID = 2, PMCD_SOR = 3, PMCD_SSOR = 4 };
/* multigrid_smoother: jacobi, sor, ssor (Multigrid.C:19-30, MultigridLevel.C:40-66) */
enum { PMCD_SMOOTH_JACOBI = 0, PMCD_SMOOTH_SOR = 1, PMCD_SMOOTH_SSOR = 2 };
/* multigrid_cycle_type */
enum { PMCD_CYCLE_V = 0, PMCD_CYCLE_W = 1, PMCD_CYCLE_F = 2 };
/* multigrid_smoother_coarse: the direct solve, or the coarse smoother
 * repeated while the coarse defect exceeds coarse_residual (Multigrid.C:
 * 200-240) */
enum { PMCD_COARSE_DIRECT = 0, PMCD_COARSE_JACOBI = 1, PMCD_COARSE_SOR = 2, PMCD_COARSE_SSOR = 3 };

typedef struct {
  int method;          /* solver_type / iterative_solver_type */
  int precond;         /* preconditioner */
  int max_iter;        /* max_n_iterations */
  int min_iter;        /* min_n_iterations */
  double res_tol;      /* residual_tolerance */
  double res_red;      /* residual_reduction */
  double div_factor;   /* divergence_factor */
  int restart;         /* gmres_restart */
  int n_pre, n_post;   /* multigrid_n_pre_smooth, multigrid_n_post_smooth */
  double jacobi_damp;  /* jacobi_damp */
  double mg_damp;      /* multigrid_correction_damp_factor */
  int mixed;           /* 1: multigrid V-cycle in FP32 inside the FP64 outer solver */
  int smoother;        /* multigrid_smoother: PMCD_SMOOTH_* */
  double sor_omega;    /* sor_omega (reference default 1.5, Solver.C:154-156) */
  int cycle;           /* multigrid_cycle_type: PMCD_CYCLE_V, _W, _F (CycleControl.C:17-111) */
  int coarse_smoother; /* multigrid_smoother_coarse: PMCD_COARSE_* (0, the default, direct solve) */
  double coarse_residual; /* multigrid_coarse_residual (reference default 0.1, absolute) */
  int coarse_max_iter; /* multigrid_coarse_max_n_iterations (reference default 10) */
} pmcd_solver_desc;

typedef struct {
  int iterations;
  int converged;
  double initial_residual;
  double final_residual;
  double seconds;
} pmcd_solve_stats;

/* Message of the last failure on the calling thread. */
const char* pmcd_last_error(void);
/* Number of visible GPUs. */
int pmcd_device_count(int* n);
/* Free and total device memory in bytes (hipMemGetInfo); no reference
 * counterpart (admission by free-HBM budget in the request loop). */
int pmcd_mem_info(unsigned long long* free_bytes, unsigned long long* total_bytes);

/* Host-only setup (no GPU needed): mesh hierarchy, FE spaces and sparsity,
 * the part of ConvectionDiffusion<d>'s constructor that stays on the host:
 * TDomain::refine_and_get_hierarchy_of_collections (Domain.C:2796-2900),
 * TFESpace3D::ConstructSpace + compute_various_numbers (FESpace3D.C:186-804,
 * FESpace.C:1863-2099), TStructure(space) (Structure.C:226-618). Levels are
 * numbered 0 = coarsest, -1 = finest; info as in pmcd_system_info. */
typedef struct pmcd_host pmcd_host;
int pmcd_host_create(const pmcd_problem_desc* desc, pmcd_host** out);
int pmcd_host_destroy(pmcd_host* h);
int pmcd_host_info(pmcd_host* h, int level, long long* info);
int pmcd_host_dofmap(pmcd_host* h, int level, int* out);
int pmcd_host_pattern(pmcd_host* h, int level, int* row_ptr, int* col);
int pmcd_host_cell_coords(pmcd_host* h, int level, double* out);
/* Hanging DOFs of a level (meshes refined by indicator) in the sorted order
 * of TFESpace::get_sorted_hanging_nodes (src/FE/FESpace.C:2094-2160;
 * hanging DOFs follow the non-Dirichlet boundary DOFs, FESpace.h:71-92):
 * *n = their count; if out is not NULL, out[4t] = hanging DOF and
 * out[4t+1 .. 4t+3] = its coupling DOFs (THangingNode::GetDOF, -1 beyond
 * their count: 2 for order 1, 3 for order 2); if w is not NULL,
 * w[3t .. 3t+2] = the coupling weights (THangingNode::GetCoeff, 0 beyond). */
int pmcd_host_hanging(pmcd_host* h, int level, int* out, double* w, int* n);
/* Writes the level's mesh as a 3D medit .mesh file (vertices, boundary faces
 * with marker = boundary component + 1, cells), the format TCollection::
 * writeMesh writes (src/Geometry/Collection.C:806-); it reads back through
 * PMCD_MESH_FILE to the same cells, vertices and boundary joints. */
int pmcd_host_write_mesh(pmcd_host* h, int level, const char* path);
/* Row classes of the device SOR sweep for a level (host computation): rows
 * (n_dof ints) grouped by class, class_ptr (n_dof + 1 ints, may be NULL; the
 * first n_classes + 1 are set). exact = 1 when the classes are the level
 * schedule of the row-order dependencies (the device sweep then equals the
 * serial TMatrix::sor_sweep, Matrix.C:1370-1395, bit for bit), 0 for greedy
 * multicolour order (used when there are more than max_exact_classes levels).
 * No reference counterpart. */
int pmcd_host_sweep_order(pmcd_host* h, int level, int max_exact_classes, int* rows, int* class_ptr,
                          int* n_classes, int* exact);
/* Sliced ELL layout of the level's active rows as the device builds it:
 * perm and len (n_active + 63 ints each: slot -> row, -1 on padding, and the
 * row length), slice_off ((n_active + 63) / 64 + 1 entries: first entry of
 * each slice of 64 slots; entry k of the row in slot t sits at
 * slice_off[t / 64] + 64 k + t % 64). No reference counterpart. */
int pmcd_host_sell_layout(pmcd_host* h, int level, int* perm, int* len, long long* slice_off, int* n_slots);
/* Column footprints of that layout's blocks of 256 slots (4-slice windows
 * only), as the LDS-staged kernels (sliced ELL mode bit 4) use them: block
 * b's distinct columns ascending in fp[fp_off[b] .. fp_off[b+1]], and for
 * each layout entry the position of its column in its block's list (lcol, 0
 * on padding). sizes: fp_off, fp and lcol lengths and the widest footprint;
 * all 0 when a block has more than 4096 distinct columns (the kernels then
 * read x directly). Null arrays: sizes only. No reference counterpart. */
int pmcd_host_sell_footprint(pmcd_host* h, int level, int* fp_off, int* fp, unsigned short* lcol,
                             long long* sizes);

/* Mesh hierarchy, FE spaces, sparsity, device buffers.
 * Replaces ConvectionDiffusion<d>::ConvectionDiffusion
 * (src/System/ConvectionDiffusion.C:112-180): TDomain refinement,
 * TFESpace2D/3D (FESpace3D.C:186-804), TStructure (Structure.C:226-618). */
int pmcd_system_create(const pmcd_problem_desc* desc, pmcd_system** out);
int pmcd_system_destroy(pmcd_system* s);
/* Replace the PDE data (example, eps, b, c, f0) of a system, keeping its mesh
 * hierarchy, spaces, sparsity and device buffers; the other fields of desc
 * must equal the system's. Reassemble afterwards. Lets a request loop keep
 * meshes, matrices and hierarchies resident across requests (the reference
 * builds a new ConvectionDiffusion<d> per problem, ConvectionDiffusion.C:112). */
int pmcd_system_set_data(pmcd_system* s, const pmcd_problem_desc* desc);

/* info[0..7] = n_cells, n_dof, n_active, nnz, n_local, n_quad_points,
 * n_levels, n_dirichlet of level `level` (0 = coarsest, -1 = finest). */
int pmcd_system_info(pmcd_system* s, int level, long long* info);
/* Hanging DOFs of a level, as pmcd_host_hanging. */
int pmcd_get_hanging(pmcd_system* s, int level, int* out, double* w, int* n);
/* GlobalNumbers of the level's FE space (TFESpace::GetGlobalDOF),
 * n_cells * n_local ints. */
int pmcd_get_dofmap(pmcd_system* s, int level, int* out);
/* CSR pattern (TStructure::get_row_ptr / get_vector_columns). */
int pmcd_get_pattern(pmcd_system* s, int level, int* row_ptr, int* col);
/* vertex coordinates per cell, n_cells * n_vertices_per_cell * 3 doubles */
int pmcd_get_cell_coords(pmcd_system* s, int level, double* out);

/* Assembly of matrix and right-hand side on every level.
 * Replaces ConvectionDiffusion<d>::assemble (ConvectionDiffusion.C:238-254)
 * -> LocalAssembling<d> + Assemble2D/Assemble3D (Assemble3D.C:39-917,
 * Assemble2D.C:2798-3430). `ms` (may be NULL) receives the time. */
int pmcd_assemble(pmcd_system* s, double* ms);
int pmcd_get_matrix(pmcd_system* s, int level, double* val);
int pmcd_get_rhs(pmcd_system* s, int level, double* rhs);

/* Solve on the finest level. Replaces Solver<BlockFEMatrix,BlockVector>::solve
 * (src/Solver/Solver.C:601-696) with CG (Iteration_cg.C:27-147), flexible
 * GMRES (Iteration_gmres.C:430-620), Jacobi (Iteration_jacobi.C) and
 * multigrid (Multigrid.C:157-410) preconditioning. */
int pmcd_solve(pmcd_system* s, const pmcd_solver_desc* sd, pmcd_solve_stats* stats);
int pmcd_get_solution(pmcd_system* s, double* out);
int pmcd_set_solution(pmcd_system* s, const double* in);

/* L2, H1-semi, SD, DG-volume and L-infinity errors of the finest solution.
 * Replaces TFEFunction3D::GetErrors / TFEFunction2D::GetErrors with
 * conv_diff_l2_h1_linf_error<d> (ConvDiff.C:2124-2183). */
int pmcd_errors(pmcd_system* s, double* out5);

/* y = A x (mode 0, BlockFEMatrix::apply) or y += a A x (mode 1,
 * BlockFEMatrix::apply_scaled_add), BlockFEMatrix.C:969-1095, on host arrays
 * of the level's length. */
int pmcd_apply(pmcd_system* s, int level, const double* x, double* y, double a, int mode);
/* One multigrid cycle z = MG(v) on the finest level with the given solver
 * settings (Iteration_multigrid::apply, Iteration_multigrid.C:23-34; V, W or
 * F cycle by sd->cycle). */
int pmcd_vcycle(pmcd_system* s, const pmcd_solver_desc* sd, const double* v, double* z);
/* One SOR sweep x <- sweep(b, x) on level `level`, in place on host arrays
 * (TMatrix::sor_sweep, src/Matrix/Matrix.C:1308-1420): flag 0 forward, 1
 * backward, 2 forward then backward (SSOR). *exact (may be NULL) is set to 1
 * when the device visits rows in an order equivalent to the serial sweep
 * (bit-identical result), 0 when it uses multicolour order. */
int pmcd_sor_sweep(pmcd_system* s, int level, const double* b, double* x, double omega, int flag,
                   int* exact);
/* One damped Jacobi step x <- x + omega D^{-1} (b - A x) on level `level`, in
 * place on host arrays (JacobiSmoother::smooth, src/Multigrid/
 * JacobiSmoother.C:22-30; non-active rows with A = I), with the kernel the
 * smoother uses (CSR or sliced ELL, tuning key 3). */
int pmcd_jacobi_sweep(pmcd_system* s, int level, const double* b, double* x, double omega);

/* Average device time (HIP events on the system's stream) of `iters`
 * back-to-back launches of one kernel on level `level`:
 * 0 FP64 SpMV, 1 FP64 Jacobi sweep, 2 FP32 SpMV, 3 local-matrix assembly,
 * 4 row gather, 5 FP32 Jacobi sweep, 6 streaming copy of the level's matrix
 * values (HBM calibration), 7 the same copy with four loads in flight per
 * thread, 8 a read-only stream over the values. No reference counterpart
 * (measurement). */
int pmcd_time_kernel(pmcd_system* s, int which, int level, int iters, double* avg_ms);
/* The level's sliced ELL layout as built (for byte counts of its kernels):
 * out[0] slots (-1: not built), out[1] layout entries (values, with
 * padding), out[2] footprint blocks, out[3] footprint entries, out[4] the
 * widest footprint (out[2..4] 0 without footprints), out[5] the launch mode
 * its kernels use now. No reference counterpart (measurement). */
int pmcd_sell_stats(pmcd_system* s, int level, long long* out);
/* The V-cycle's grid transfer between `level` (-1 = finest) and the next
 * coarser level, with the kernels the V-cycle launches, on host arrays:
 * which 0 = prolongation y += a P x on the active fine rows (x: coarse
 * vector, y: fine vector), 1 = restriction y = R x on the active coarse rows
 * and 0 on the coarse Dirichlet tail (x fine, y coarse; a not read). The
 * counterparts of GridTransfer::Prolongate / DefectRestriction
 * (GridTransfer.C:603-1115) on the system's own hierarchy. */
int pmcd_apply_transfer(pmcd_system* s, int level, int which, const double* x, double* y, double a);

/* Launch-shape tuning (process-wide; no reference counterpart, knobs of the
 * autotuning sweep): key 0 / 1 = lanes per matrix row of the SpMV / Jacobi
 * kernels (4, 8, 16, 32; defaults 16 and 8); key 2 / 3 = their variant
 * (default 2; 0 plain, 1 streaming with
 * XCD-contiguous rows, 2 unrolled with all loads of a lane in flight at once,
 * taken where the level's longest row fits 16 entries per lane, 3 unrolled
 * with non-temporal matrix loads, 4 sliced ELL: one lane per row, each row
 * summed serially in column order as TMatrix::multiply does, so the row sums
 * are bitwise those of the serial loop; calls on raw CSR arrays take variant
 * 2 instead); key 9 = entries per lane loaded per step by the sliced ELL
 * kernels (4, 8, 16; default 16); key 10 = rows of a sliced ELL
 * length-sorting window, in slices of 64 (default 4; read when a level's
 * layout is first built); key 11 = levels with fewer active rows than this
 * keep the unrolled CSR kernels under variant 4 (default 0); key 5 = the
 * largest number of level-schedule classes for which SOR sweeps follow the
 * exact serial order (default 24; deeper schedules use multicolour order),
 * read when a level's sweep order is first built; key 6 = 1 to replay
 * multigrid V-cycles from captured HIP graphs, 0 to launch them kernel by
 * kernel (default); key 4 = assembly kernel (0 reference operation order on
 * vector ALUs, 1 FP64 MFMA for 27-DOF hexahedra and the vector kernels for the
 * other elements, 2 (default) reference-tensor GEMMs on FP64 MFMA for every
 * element but P1 simplices); key 7 = 1 (default) to run CG with no or Jacobi
 * preconditioning as two fused kernels per iteration with its scalars on the
 * device, 0 for the loop that reads every scalar back to the host; key 8 =
 * the largest grid of the device CG's SpMV kernel (1..4096 blocks); key 12 =
 * largest sweep class handled by the one-launch sweep kernel; key 13 = sliced
 * ELL launch mode bits (1 XCD-contiguous slices, 2 row-ordered output through
 * LDS, 4 non-temporal matrix loads, 8 matrix loads pipelined one step ahead,
 * 16 x staged in LDS per 256-row block through its column footprint, with 2;
 * default 23; same bits in every mode); key 18 = sliced ELL blocks resident
 * per CU at most (0 no cap, the default, or 3..16); key 19 = 1 (default) for
 * a sliced ELL copy of the prolongation on levels whose Jacobi sweeps use
 * the sliced ELL kernels (serial row sums), 0 for the CSR kernel;
 * key 14 = 1 (default) for XCD-contiguous rows in the unrolled CSR kernels;
 * key 15 = FGMRES look-ahead factor (the next iteration is queued before the
 * host test while the residual exceeds this many times the threshold;
 * default 4, 0 off); key 16 = 1 (default) to replay each batched FGMRES
 * iteration from a captured HIP graph; key 17 = 1 for one cooperative launch
 * per batched multicolour sweep (default 0: one launch per class; with 1 the
 * iterations are launched kernel by kernel, without graphs). */
int pmcd_set_tuning(int key, int value);
/* Directions of one multigrid cycle (CycleControl, src/Multigrid/
 * CycleControl.C:17-111): writes at most cap entries (0 down, 1 up, 2 end,
 * 3 the leading dummy) and returns the full count in *n. Host computation. */
int pmcd_cycle_directions(int type, int n_levels, int* out, int cap, int* n);
/* Current value of a launch-shape knob (keys as for pmcd_set_tuning). */
int pmcd_get_tuning(int key, int* value);

/* Lane-layout check of the 16x16x4 FP64 MFMA used by the assembly: writes the
 * 64 lanes x 4 result registers of one product of known operands (256
 * doubles). No reference counterpart (hardware check). */
int pmcd_probe_mfma(double* out256);

/* Device-pointer entry points (all pointers are device memory, `stream` a
 * hipStream_t or NULL). FEMatrix::multiplyActive + addScaledNonActive:
 * y[r] = beta*z[r] + alpha*(A x)[r] for r < n_active, beta*z[r] + alpha*x[r]
 * otherwise. */
int pmcd_spmv_device(int n, int n_active, const int* row_ptr, const int* col, const double* val,
                     const double* x, const double* z, double* y, double alpha, double beta,
                     void* stream);
/* Ddot (Blas1.C:21-38): deterministic two-pass reduction, result to host. */
int pmcd_dot_device(int n, const double* x, const double* y, double* result, void* stream);
/* pmcd_dot_device without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). The
 * reduction's partial sums and result live in a fixed per-device scratch
 * array of 2048 doubles (16 KiB, held until the process ends); every n >= 0
 * fits it, and a reduction that would need more is rejected with an error.
 * Same kernels and bits as pmcd_dot_device. Calls that overlap on one device
 * (other streams or threads) share the array and race. */
int pmcd_dot_device_noalloc(int n, const double* x, const double* y, double* result, void* stream);

/* ---- time-dependent problems ----------------------------------------------
 * TimeConvectionDiffusion<d> (src/System/TimeConvectionDiffusion.C:259-452)
 * with TimeDiscretization (src/TimeDiscRoutines/TimeDiscretizations.C):
 * pmcd_tcd_init = assemble_initial_time (stiffness, mass and rhs at
 * time_start, initial condition = the example's exact solution interpolated
 * at the nodes); pmcd_tcd_step = one step of length tau (assemble at the new
 * time, the scheme's rhs and system matrix on every level, solve); *time
 * receives the new time. Errors at the current time: pmcd_errors. */
enum { PMCD_BACKWARD_EULER = 0, PMCD_CRANK_NICOLSON = 1, PMCD_BDF_TWO = 2 };
typedef struct {
  int scheme;     /* time_discretization: PMCD_BACKWARD_EULER, PMCD_CRANK_NICOLSON, PMCD_BDF_TWO */
  int bdf_first;  /* bdf_two_first_step: PMCD_BACKWARD_EULER or PMCD_CRANK_NICOLSON */
  double tau;     /* time_step_length */
  double t0;      /* time_start */
} pmcd_tcd_desc;
int pmcd_tcd_init(pmcd_system* s, const pmcd_tcd_desc* t);
int pmcd_tcd_step(pmcd_system* s, const pmcd_solver_desc* sd, pmcd_solve_stats* stats, double* time);

/* ---- algebraic flux correction ---------------------------------------------
 * ConvectionDiffusion_AFC<d>::assemble(0); solve(0); solve(k) until it
 * returns true (src/System/ConvectionDiffusion_AFC.C:132-686, the loop of
 * test/cd2d_afc_test.c++:60-74) with afc_limiter = kuzmin or BJK17 (P1
 * triangles and tetrahedra), afc_iteration_scheme = fixed_point_rhs,
 * fixed_point_matrix or newton (AlgebraicFluxCorrection.C:2794-2981, the
 * Jacobians of :224-578 and :957-1140), dynamic damping, with or without
 * Anderson acceleration (anderson.f90), afc_initial_iterate = afc_zero.
 * Single level, Galerkin.
 * prm[19]: afc_nonlinloop_epsilon, afc_nonlinloop_maxit,
 * afc_nonlinloop_damping_factor, _max, _min, _max_global, _increase,
 * _decrease, _max_increase, _max_decrease, _min_tol, limiter (0 kuzmin,
 * 1 BJK17), scheme (0 fixed_point_rhs, 1 fixed_point_matrix, 2 newton),
 * afc_fixed_point_matrix_weight, afc_fixed_point_derivative_weight,
 * afc_change_method_threshold, afc_nonlinloop_anderson_acc (0 no, 1 yes),
 * afc_nonlinloop_anderson_acc_vec (1..32), afc_nonlinloop_anderson_acc_start.
 * out[4]: nonlinear iterations, final residual norm, rejected steps, initial
 * residual norm. stats: linear iterations summed over the loop. The solution
 * is left in the system (pmcd_get_solution, pmcd_errors). */
int pmcd_afc_solve(pmcd_system* s, const pmcd_solver_desc* sd, const double* prm, double* out,
                   pmcd_solve_stats* stats);

/* Files in the reference's formats (host arrays; no GPU use):
 * pmcd_vector_write / pmcd_vector_read = BlockVector::write_to_stream /
 * read_from_stream (src/Matrix/BlockVector.C:1064-1110: the length on its own
 * line, then the raw doubles; the restart format of CheckpointIO), reading
 * fails unless the file holds exactly n entries; pmcd_vector_write_market =
 * BlockVector::write (:598-615, MatrixMarket array); pmcd_matrix_write_market
 * = TMatrix::write (src/Matrix/Matrix.C:97-130, MatrixMarket coordinate,
 * 1-based, 16 significant digits). */
int pmcd_vector_write(const char* path, long long n, const double* v);
int pmcd_vector_read(const char* path, long long n, double* v);
int pmcd_vector_write_market(const char* path, long long n, const double* v);
int pmcd_matrix_write_market(const char* path, int n_rows, int n_cols, const int* row_ptr, const int* col,
                             const double* val);

/* ---- batched instances ------------------------------------------------------
 * Many independent problems (own mesh, coefficients and boundary data; shared
 * ansatz order, discretisation and quadrature) assembled and solved together:
 * one device launch per step covers every instance. The reference solves one
 * ConvectionDiffusion<d> at a time (ConvectionDiffusion.C:112-452); a batch
 * is the same path for a set of them. Instances are single-level; the solver
 * is flexible GMRES (Iteration_gmres.C:430-620) with no, Jacobi, SOR or SSOR
 * preconditioning (FP32 sweeps when sd->mixed), per-instance convergence. */
typedef struct pmcd_batch pmcd_batch;
int pmcd_batch_create(const pmcd_problem_desc* descs, int n_inst, pmcd_batch** out);
int pmcd_batch_destroy(pmcd_batch* b);
/* inst < 0: info[0..4] = n_inst, n_dof, n_active, nnz, n_cells of the whole
 * batch; inst >= 0: info[0..3] = 1, n_dof, n_active, nnz of that instance. */
int pmcd_batch_info(pmcd_batch* b, int inst, long long* info);
int pmcd_batch_assemble(pmcd_batch* b, double* ms);
/* stats: one entry per instance (may be NULL). */
int pmcd_batch_solve(pmcd_batch* b, const pmcd_solver_desc* sd, pmcd_solve_stats* stats);
/* instance-local numbering (the instance's own DOF order and CSR pattern) */
int pmcd_batch_get_solution(pmcd_batch* b, int inst, double* out);
int pmcd_batch_set_solution(pmcd_batch* b, int inst, const double* in);
int pmcd_batch_get_matrix(pmcd_batch* b, int inst, double* val);
int pmcd_batch_get_rhs(pmcd_batch* b, int inst, double* rhs);

/* ---- Drop-in entries on caller-owned spaces, structures and matrices ----
 *
 * These bind at the reference's own call sites: ConvectionDiffusion<d>::
 * call_assembling_routine (src/System/ConvectionDiffusion.C:410-452) calls
 * Assemble3D / Assemble2D with the caller's FE space, structure, matrix
 * entries and rhs; ConvectionDiffusion<d>::solve (:258-281) calls
 * Solver<BlockFEMatrix, BlockVector>::solve on the caller's matrix. */

/* Cell shapes of a pmcd_fe_space, as in include/Geometry/Enumerations_geometry.h */
enum { PMCD_SHAPE_TRIANGLE = 0, PMCD_SHAPE_QUADRANGLE = 1, PMCD_SHAPE_TETRAHEDRON = 2, PMCD_SHAPE_HEXAHEDRON = 3 };

/* A caller's FE space (TFESpace2D / TFESpace3D, include/FE/FESpace.h):
 * GetGlobalNumbers() and GetBeginIndex() (cell c owns global_numbers
 * [begin_index[c] .. begin_index[c+1])), get_n_dof() and get_n_active()
 * (inner and non-Dirichlet DOFs first, then the hanging DOFs, Dirichlet DOFs
 * last), and the cells' vertex coordinates in the element's local vertex
 * order (TBaseCell::GetVertex). Continuous Lagrange elements P1/P2 on
 * simplices, Q1/Q2 on quadrilaterals and hexahedra; affine cells.
 * Hanging DOFs (get_n_hanging(), FESpace.h:361-391, in 2D): the last
 * n_hanging of the n_active DOFs; hanging[4i] is one of them and
 * hanging[4i+1..4i+3] its coupling DOFs (THangingNode::GetDOF of
 * get_hanging_node(), HangingNode.h:39-47, -1 beyond GetN_Nodes()),
 * hanging_weights[3i..3i+2] their weights (GetCoeff). The
 * structure must then hold each hanging row's columns in the rows of its
 * active coupling DOFs, as the reference's structure does. n_hanging = 0
 * for a conforming space (the two pointers are not read). */
typedef struct {
  int shape;                  /* PMCD_SHAPE_* */
  int order;                  /* 1 or 2 */
  int n_cells;
  int n_dof;
  int n_active;               /* hanging DOFs included */
  const double* vertices;     /* [n_cells][vertices per cell][3] */
  const int* global_numbers;  /* sum over cells of the local DOF counts */
  const int* begin_index;     /* n_cells + 1 */
  int n_hanging;
  const int* hanging;              /* [n_hanging][4] */
  const double* hanging_weights;   /* [n_hanging][3] */
} pmcd_fe_space;

/* A caller's square structure (TStructure, include/Matrix/Structure.h):
 * get_row_ptr() and get_vector_columns() of its n_rows rows; it must hold
 * every coupling of two DOFs of one cell (columns in any order). */
typedef struct {
  int n_rows;
  const int* row_ptr;  /* n_rows + 1 */
  const int* col;      /* row_ptr[n_rows] */
} pmcd_structure;

/* Assemble3D / Assemble2D with the convection-diffusion LocalAssembling
 * (include/AssembleRoutines/Assemble3D.h:28-34, LocalAssembling.h:229-231):
 * the Galerkin (+ SUPG if la->supg) local matrices and rhs of every cell
 * added into the caller's `entries` (TMatrix::GetEntries, row_ptr[n_rows]
 * values) and `rhs` (n_dof values), as the reference's scatter does
 * (Assemble3D.C:258-300,346-360): active rows get the cell contributions
 * added in increasing cell order, each Dirichlet row gets 1 on its diagonal
 * (the rest of the row is left unchanged), the rhs gets the contributions
 * added on active DOFs and the Dirichlet values set on the others. With
 * hanging DOFs the sums are coupled first (Assemble2D.C:3350-3427: each
 * hanging row, weighted, added to the rows of its active coupling DOFs in
 * the order the couplings require), the coupled sums are added to the
 * non-hanging active rows, and each hanging row is set to its constraint
 * (1 on the diagonal, minus the weights at the coupling DOFs, 0 elsewhere)
 * with rhs 0.
 * dirichlet_values: the n_dof - n_active boundary values in DOF order (the
 * reference evaluates BoundaryValues through the nodal functionals; host
 * callbacks cannot run on the device), NULL for zero. The caller resets the
 * entries and rhs first when a fresh system is wanted, as
 * ConvectionDiffusion.C:429-430 does; the sums are then the reference's.
 * la: the PDE data and forms (example, eps, b, c, f0, supg, linear_quad);
 * its geometry fields are not read. Host arrays. Neumann boundary integrals
 * are not added. */
int pmcd_assemble_into(const pmcd_fe_space* space, const pmcd_structure* st, const pmcd_problem_desc* la,
                       const double* dirichlet_values, double* entries, double* rhs);
/* The same with the space and structure uploaded once: create, then assemble
 * as often as needed (host arrays, or device arrays on `stream`). */
typedef struct pmcd_assembler pmcd_assembler;
int pmcd_assembler_create(const pmcd_fe_space* space, const pmcd_structure* st, const pmcd_problem_desc* la,
                          pmcd_assembler** out);
int pmcd_assembler_destroy(pmcd_assembler* a);
int pmcd_assembler_assemble(pmcd_assembler* a, const double* dirichlet_values, double* entries, double* rhs);
int pmcd_assembler_assemble_device(pmcd_assembler* a, const double* dirichlet_values, double* entries, double* rhs,
                                   void* stream);

/* One level of a caller's matrix hierarchy, coarsest first (the
 * BlockFEMatrix of each SystemPerGrid, ConvectionDiffusion.h:44-78, and the
 * multigrid levels built from them, MultigridLevel.C:22-28): CSR with the
 * Dirichlet rows n_active.. last (they act as identity rows, their entries
 * are not read) and, on every level but the coarsest, the prolongation from
 * the next coarser level as an n x n_coarse CSR matrix (the cell-wise
 * interpolation that GridTransfer::Prolongate applies, GridTransfer.C:603-865;
 * restriction uses its transpose). */
typedef struct {
  int n;
  int n_active;
  const int* row_ptr;
  const int* col;
  const double* entries;
  int n_coarse;
  const int* p_row_ptr;
  const int* p_col;
  const double* p_entries;
} pmcd_csr_level;

/* Solver<BlockFEMatrix, BlockVector> (include/Solver/Solver.h:45-94) on a
 * caller-owned matrix: create = Solver(db) with the solver keys of
 * pmcd_solver_desc; update_matrix = update_matrix(A) with the multigrid
 * levels (the coarse factorization and derived copies are redone; values
 * only when the patterns are unchanged); solve(b, x) with x the initial
 * iterate on entry and the solution on return (host arrays, finest level);
 * is_using_multigrid. */
typedef struct pmcd_solver pmcd_solver;
int pmcd_solver_create(const pmcd_solver_desc* sd, pmcd_solver** out);
int pmcd_solver_destroy(pmcd_solver* s);
int pmcd_solver_update_matrix(pmcd_solver* s, int n_levels, const pmcd_csr_level* levels);
int pmcd_solver_solve(pmcd_solver* s, const double* b, double* x, pmcd_solve_stats* stats);
int pmcd_solver_is_using_multigrid(const pmcd_solver* s, int* out);

/* The prolongation of a system's level (level -1 = finest) from the next
 * coarser one, as pmcd_csr_level takes it: sizes first (null arrays), then
 * row_ptr (n_fine + 1), col and val (nnz). No reference counterpart
 * (GridTransfer applies it cell by cell without storing it). */
int pmcd_get_transfer(pmcd_system* s, int level, int* n_fine, int* n_coarse, int* nnz, int* row_ptr, int* col,
                      double* val);

/* Device-pointer building blocks (all pointers device memory, `stream` a
 * hipStream_t or NULL): Daxpy y += a x (src/General/Blas1.C), Dnorm
 * sqrt(x.x) (two-pass deterministic reduction, result to host); one damped
 * Jacobi step x_out = x + omega (b - A x) / diag, A acting as the identity on
 * the rows n_active.. (JacobiSmoother::smooth, src/Multigrid/
 * JacobiSmoother.C:22-30, with FEMatrix::multiplyActive);
 * prolongation fine += a P coarse on the active fine rows and restriction
 * coarse = R fine on the active coarse rows, 0 on the others
 * (GridTransfer::Prolongate / DefectRestriction with the non-active rules of
 * Multigrid.C:310-410). */
int pmcd_axpy_device(int n, double a, const double* x, double* y, void* stream);
int pmcd_norm_device(int n, const double* x, double* result, void* stream);
int pmcd_jacobi_device(int n, int n_active, const int* row_ptr, const int* col, const double* entries,
                       const double* diag, const double* b, const double* x, double* x_out, double omega,
                       void* stream);
int pmcd_prolongate_device(int n_fine, int n_active_fine, const int* p_row_ptr, const int* p_col,
                           const double* p_entries, const double* coarse, double* fine, double a, void* stream);
int pmcd_restrict_device(int n_coarse, int n_active_coarse, const int* r_row_ptr, const int* r_col,
                         const double* r_entries, const double* fine, double* coarse, void* stream);

/* Batched workload instances: std::mt19937_64 seeded `seed`, per instance in
 * this order N = uniform_int_distribution(n_lo, n_hi), eps = 10^u with u =
 * uniform_real_distribution(-6, 0), g = three normal_distribution(0, 1)
 * draws, b = g / |g| (n_out[n_inst], eps_out[n_inst], b_out[3 n_inst]). No
 * reference counterpart (the synthetic batch of independent problems). */
int pmcd_batch_instances(unsigned long long seed, int n_inst, int n_lo, int n_hi, int* n_out, double* eps_out,
                         double* b_out);

/* ---- Device memory of a system ----
 * Bytes the system's device buffers hold now and at most so far (each
 * allocation counted in whole 4 KiB pages). */
int pmcd_system_device_bytes(pmcd_system* s, long long* bytes, long long* peak);
/* Upper estimate of those bytes after a solve with `sd`, from the coarse
 * mesh's vertex, edge, face and cell counts and the regular-refinement
 * recurrences (no mesh is refined and nothing is allocated). */
int pmcd_estimate_device_bytes(const pmcd_problem_desc* pd, const pmcd_solver_desc* sd, long long* bytes);
/* The finest level's counts behind the estimate: out3 = cells, DoFs (both
 * exact), matrix-entry bound. */
int pmcd_estimate_counts(const pmcd_problem_desc* pd, long long* out3);

/* ---- Launch-shape autotuner ----
 * Times the SpMV and Jacobi kernels of `level` (-1: finest) in every CSR
 * variant (0 plain, 1 streaming, 2 unrolled, 3 unrolled non-temporal) at 8,
 * 16 and 32 lanes per row and as sliced ELL, with HIP events over `iters`
 * launches each, and sets the fastest (pmcd_set_tuning keys 0-3); when
 * sliced ELL wins, coarser levels on which it loses keep the CSR kernels
 * (key 11). `mixed`: the Jacobi sweep is timed in FP32 (the system must have
 * solved with the FP32 V-cycle once). ms_out[26]: SpMV candidates 0..12, then
 * Jacobi candidates 0..12, candidate 3 * variant + (0, 1, 2 for 8, 16, 32
 * lanes), candidate 12 sliced ELL. chosen_out[5]: SpMV variant, SpMV lanes,
 * Jacobi variant, Jacobi lanes, sliced-ELL minimum rows. No reference
 * counterpart (the reference's kernels have no launch shape). */
int pmcd_autotune(pmcd_system* s, int level, int iters, int mixed, double* ms_out, int* chosen_out);

/* ---- Request server ----
 * A request loop over device-resident systems for callers that run many
 * solves of mixed sizes. The reference runs one problem per process:
 * ConvectionDiffusion<d>(domain, db), assemble(), solve(), output()
 * (3DPrograms/CD3D_ParMooN.C:40-53, 2DPrograms/CD2D_ParMooN.C:19-22,
 * src/System/ConvectionDiffusion.C:112-452); a served request runs the same
 * assemble and solve on a system that stays resident for the next request on
 * the same mesh (same geometry, refinement, order, discretization,
 * quadrature mode, multigrid levels and mesh file), which only receives the
 * new PDE data (pmcd_system_set_data).
 *
 * Requests wait in FIFO order. A worker admits the head request when fewer
 * than max_slots run and either an idle system for its mesh exists or the
 * estimated bytes (pmcd_estimate_device_bytes) fit the budget once idle
 * systems, oldest first, are destroyed; a request larger than the budget
 * runs when nothing else runs. Each running request works on its own HIP
 * stream, so requests overlap on the device. Times are seconds on the
 * server's steady clock (pmcd_server_now). */
typedef struct pmcd_server pmcd_server;
typedef struct {
  long long id;
  int status;            /* 0 solved, 1 failed (message in error) */
  int reused;            /* 1: the request ran on a resident system */
  double arrival, admitted, finished;
  double assemble_ms;
  long long n_dof;
  long long reserved_bytes;  /* device bytes reserved at admission */
  pmcd_solve_stats stats;
  char error[256];
} pmcd_served_result;
/* budget_bytes <= 0: 90% of the free device memory; keep_solution: keep each
 * finest-level solution on the host for pmcd_server_solution */
int pmcd_server_create(int max_slots, long long budget_bytes, int keep_solution, pmcd_server** out);
/* runs every queued request to the end, then frees the resident systems */
int pmcd_server_destroy(pmcd_server* s);
int pmcd_server_now(pmcd_server* s, double* t);
/* arrival < 0: now */
int pmcd_server_submit(pmcd_server* s, const pmcd_problem_desc* pd, const pmcd_solver_desc* sd, double arrival,
                       long long* id);
/* creates and parks up to `count` idle systems for pd's mesh */
int pmcd_server_warm(pmcd_server* s, const pmcd_problem_desc* pd, const pmcd_solver_desc* sd, int count);
/* done = 1 once request `id` finished; waits at most timeout_s (< 0: no limit) */
int pmcd_server_wait(pmcd_server* s, long long id, double timeout_s, int* done);
int pmcd_server_result(pmcd_server* s, long long id, pmcd_served_result* out);
int pmcd_server_solution(pmcd_server* s, long long id, double* out, long long n);
/* forgets a finished request's record and solution */
int pmcd_server_release(pmcd_server* s, long long id);
/* out8: queued, running, reserved bytes, idle bytes, idle systems, budget,
 * finished requests, systems created */
int pmcd_server_info(pmcd_server* s, long long* out8);

#ifdef __cplusplus
}
#endif

#endif /* PMCD_H */
