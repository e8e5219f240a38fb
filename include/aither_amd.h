/*
 * aither_amd: MI355X-native implementation of aither's per-nonlinear-iteration
 * hot path (ghost-cell boundary conditions, face-flux residual, spectral-radius
 * time step, implicit operator, solution update), behind a plain C ABI.
 *
 * All arrays crossing this boundary are host arrays owned by the caller and
 * copied synchronously; the library keeps no pointer to them after a call
 * returns.  Device memory is owned by the context.  State arrays are
 * nondimensional primitive variables [rho, u, v, w, p] per physical cell,
 * AoS, i fastest then j then k -- the ordering of the reference's
 * blkMultiArray3d<primitive> (include/multiArray3d.hpp:104-113).
 *
 * Every call returns ATH_OK (0) or a negative ATH_ERR_* status; the message of
 * the last failure on the calling thread is available from ath_last_error().
 * A device-side error (ATH_ERR_NONPHYSICAL, ATH_ERR_DEVICE) is reported by
 * the call that reads the norms after it and then cleared: the context stays
 * usable, with the state as the failing iteration left it.
 * The library never calls exit().  One context is driven by one host thread
 * at a time and owns one HIP stream.
 */
#ifndef AITHER_AMD_H
#define AITHER_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define ATH_OK 0
#define ATH_ERR_ARG (-1)         /* bad argument or unsupported combination */
#define ATH_ERR_CONFIG (-2)      /* input file / grid / BC table problem */
#define ATH_ERR_IO (-3)          /* file could not be opened or read */
#define ATH_ERR_NOMEM (-4)       /* device or host allocation failed */
#define ATH_ERR_NONPHYSICAL (-5) /* rho <= 0, p <= 0 or NaN after an update */
#define ATH_ERR_HIP (-6)         /* HIP runtime error */
#define ATH_ERR_DEVICE (-7)      /* device-side check failed (e.g. spin limit) */

#define ATH_MAX_EQ 8

typedef struct ath_ctx ath_ctx;

/* case summary (input::ReadInput results plus derived quantities) */
typedef struct {
  int num_blocks;
  int num_equations;     /* input::NumEquations, src/input.cpp:655-671 */
  int num_ghosts;        /* input::NumberGhostLayers, src/input.cpp:1127-1143 */
  int iterations;        /* "iterations" */
  int nonlinear_iterations;
  int implicit;          /* input::IsImplicit, src/input.cpp:674-682 */
  double total_cells;
  double r_ref, t_ref, l_ref, a_ref;
  double gas_constant;   /* nondimensional R of the single species */
  double gamma;
  int block_matrix;      /* input::IsBlockMatrix, src/input.cpp:713-720 */
} ath_info;

/* per-iteration norms (what src/main.cpp:252-275 reduces and prints) */
typedef struct {
  int num_equations;
  double l2_sq[ATH_MAX_EQ];    /* sum over cells of R^2 per equation */
  double l2[ATH_MAX_EQ];       /* sqrt of l2_sq (residual::SquareRoot) */
  double normalized[ATH_MAX_EQ];/* PrintResiduals values, [0] = mass */
  double linf;                 /* resid::Linf (signed max, first wins) */
  int linf_block, linf_i, linf_j, linf_k, linf_eqn;
  double matrix_resid;         /* sqrt(CycleAtLevel/(cells*neq)) */
  double cfl;
} ath_norms;

/* precision of the implicit operator (D^-1, off-diagonals, du) */
#define ATH_IMPLICIT_FP64 0
#define ATH_IMPLICIT_FP32 1

/* ---- context ----------------------------------------------------------- */

/* Read a reference-format input file (.inp) and its binary Plot3D grid
 * (gridName.xyz next to it), build FP64 geometry on the host and upload the
 * case to device `device`.  Replaces, for the hot path, the sequence
 * input::ReadInput (src/input.cpp:167-643), ReadP3dGrid
 * (src/plot3d.cpp:363-445), the procBlock constructor and InitializeStates
 * (src/procBlock.cpp:59-174, 280-355), AssignGhostCellsGeom
 * (src/procBlock.cpp:2160-2268) and CalcCellWidths (src/procBlock.cpp:6397)
 * as called from src/main.cpp:101-164. */
int ath_open(const char *inp_path, int device, ath_ctx **out);
void ath_close(ath_ctx *ctx);
const char *ath_last_error(void);

int ath_get_info(const ath_ctx *ctx, ath_info *out);
int ath_block_dims(const ath_ctx *ctx, int block, int *ni, int *nj, int *nk);

/* implicit-operator precision (default ATH_IMPLICIT_FP64).  No reference
 * counterpart: the reference is FP64 throughout. */
int ath_set_implicit_precision(ath_ctx *ctx, int precision);
/* LU-SGS execution order.  ATH_SWEEP_PENCIL (default): 8x8 columns walking k
 * on a skewed wavefront, one launch per sweep pass.  ATH_SWEEP_TILED: 8x8x4
 * tiles in tile-hyperplane order, dataflow-synchronised, one launch per pass.
 * ATH_SWEEP_HYPERPLANE: one launch per i+j+k hyperplane in the order of
 * HyperplaneReorder (src/utility.cpp:377-398).  ATH_SWEEP_ROWS (block
 * solvers): 8x4x4 tiles in tile-hyperplane order with each cell's five rows
 * on their own lanes, one launch per pass.  All are topological orders of the
 * same Gauss-Seidel dependencies and give identical results. */
#define ATH_SWEEP_TILED 0
#define ATH_SWEEP_HYPERPLANE 1
#define ATH_SWEEP_PENCIL 2
#define ATH_SWEEP_ROWS 3
int ath_set_sweep_ordering(ath_ctx *ctx, int ordering);
/* the order the sweeps actually run in: block solvers (blusgs, bdplur) map
 * ATH_SWEEP_PENCIL to ATH_SWEEP_ROWS with the FP64 operator and to
 * ATH_SWEEP_TILED with the FP32 operator, scalar LU-SGS maps ATH_SWEEP_ROWS
 * to ATH_SWEEP_TILED, and the approximateRoe flux Jacobian always runs
 * ATH_SWEEP_HYPERPLANE, whatever was set */
int ath_get_sweep_ordering(const ath_ctx *ctx, int *ordering);
/* Autotune the sweep execution order for this case: time `iterations`
 * nonlinear iterations in each order the case admits (pencil, tiled and
 * hyperplane for scalar LU-SGS; the all-block tiled sweep with one or two
 * workgroups per CU, and the hyperplane order, for the block solvers) and
 * keep the fastest per sweep pass.  Results do not depend on the order; the
 * state and the norm normalisation are restored.  No reference counterpart. */
int ath_autotune(ath_ctx *ctx, int iterations, int *chosen, double *pass_ms);
/* compute the matrix residual each iteration (default 1), as
 * linearSolver::Residual does (src/linearSolver.cpp:95-109) */
int ath_set_matrix_residual(ath_ctx *ctx, int on);

/* state access: procBlock::States / InsertBlock on physical cells */
int ath_get_state(ath_ctx *ctx, int block, double *prim);
int ath_set_state(ath_ctx *ctx, int block, const double *prim);

/* ---- hot path ---------------------------------------------------------- */

/* input::CalcCFL (src/input.cpp:647-649) + mgSolution::StoreOldSolution
 * (src/mgSolution.cpp:103-114): called once per outer step nn. */
int ath_begin_step(ath_ctx *ctx, int nn);

/* One nonlinear iteration: mgSolution::Iterate (src/mgSolution.cpp:246-269)
 * followed by the norm reductions and PrintResiduals normalisation of
 * src/main.cpp:252-275 / src/output.cpp:1028-1063.  `norms` may be NULL. */
int ath_iterate(ath_ctx *ctx, int nn, int mm, ath_norms *norms);

/* The main loop of src/main.cpp:231-302 for outer steps [nn0, nn0+nsteps),
 * each with the input's nonlinear iterations, replayed from a captured HIP
 * graph.  Norms of the last iteration go to `last` (may be NULL).  With
 * nsteps == 0 it only captures and instantiates the graph (no iteration
 * runs), so a caller can keep that cost out of a timed loop. */
int ath_run(ath_ctx *ctx, int nn0, int nsteps, ath_norms *last);

/* ---- granular steps (gridLevel / procBlock mirror) ---------------------- */

/* gridLevel::GetBoundaryConditions (src/gridLevel.cpp:287-319) */
int ath_assign_ghost_cells(ath_ctx *ctx);
/* gridLevel::CalcResidual + CalcTimeStep + linearSolver::AddDiagonalTerms +
 * Invert (src/gridLevel.cpp:372-406): one fused launch per block */
int ath_calc_residual(ath_ctx *ctx);
/* linearSolver::InitializeMatrixUpdate + Relax (src/linearSolver.cpp:111-144,
 * 430-470, 509-535); returns the matrix residual (CycleAtLevel value).  Uses
 * the last ath_calc_residual's residual, time step and diagonal (for a
 * one-sweep FP64 block LU-SGS relax in the row-parallel order also the
 * right-hand side that call formed). */
int ath_relax(ath_ctx *ctx, double *cycle_value);
/* gridLevel::UpdateBlocks + ResetDiagonal (src/gridLevel.cpp:418-431) */
int ath_update(ath_ctx *ctx, int mm, ath_norms *norms);

/* field getters (physical cells, i fastest): procBlock::Residual, Dt,
 * SpectralRadius; linearSolver A (diagonal before inversion) and X (du) */
int ath_get_residual(ath_ctx *ctx, int block, double *r);     /* neq/cell */
int ath_get_dt(ath_ctx *ctx, int block, double *dt);
int ath_get_spectral_radius(ath_ctx *ctx, int block, double *sr);
int ath_get_diagonal(ath_ctx *ctx, int block, double *a);
int ath_get_update(ath_ctx *ctx, int block, double *du);      /* neq/cell */
/* block matrix solvers (blusgs, bdplur): linearSolver a_ after
 * AddDiagonalTerms (inverse = 0) or aInv_ (inverse = 1), 25 per cell,
 * row-major (src/linearSolver.cpp:146-184); ath_get_diagonal refuses such
 * cases */
int ath_get_block_diagonal(ath_ctx *ctx, int block, int inverse, double *a);
/* procBlock::CalcWallDistance over the physical cells (src/procBlock.cpp:
 * 6030-6041, kdtree::NearestNeighbor, src/kdtree.cpp:123-224): the distance
 * from each cell centre to the nearest viscousWall face centre of any block,
 * one per cell, i fastest; sqrt of the largest double where the case has no
 * viscous wall, as the reference's empty tree gives.  Computed on the host
 * from the case geometry (a setup step in the reference). */
int ath_get_wall_distance(ath_ctx *ctx, int block, double *d);

/* ---- files of the output layer ----------------------------------------- */
/* Restart file in the reference's binary format (WriteRestart,
 * src/output.cpp:591-755): dimensional state of every block (and U^{n-1}
 * with BDF2), the species name and the residual normalisation. */
int ath_write_restart(ath_ctx *ctx, const char *path, int iteration);
/* ReadRestart (src/output.cpp:756-899): state (and U^{n-1}) and the
 * residual normalisation from the file, which the norms keep from then on
 * (the reference's restart does not re-normalise); *iteration = the file's
 * iteration, where the caller's outer-step count resumes. */
int ath_read_restart(ath_ctx *ctx, const char *path, int *iteration);
/* WriteOutput's cell-centre files (src/output.cpp:573-578) into directory
 * `dir`: <gridName>_center.xyz (cell centres, WriteCellCenter :58-98),
 * <simName>_<iteration>_center.fun (the input's outputVariables, WriteFunFile
 * :206-439) and the ParaView meta file <simName>_center.p3d (:903-956).  A
 * requested variable this library does not hold (turbulence, gradients other
 * than the block solvers' velocity gradients) fails with ATH_ERR_CONFIG. */
int ath_write_output(ath_ctx *ctx, const char *dir, int iteration);
/* residual and timing logs <simName>.resid and <simName>.tme in `dir`
 * (logFileManager, src/logFileManager.cpp:34-113; PrintResiduals lines,
 * src/output.cpp:1007-1085): every ath_iterate / ath_run iteration appends a
 * line; ath_run's Iter-Time is the run's mean per iteration */
int ath_open_logs(ath_ctx *ctx, const char *dir, int append);
int ath_close_logs(ath_ctx *ctx);

/* ---- timing of the last ath_run (HIP events on the context's stream) ---- */
typedef struct {
  double run_ms;            /* whole ath_run */
  double flux_ms;           /* summed residual stages: every residual kernel
                               launched once over all blocks */
  int flux_launches;        /* residual stages timed (one per iteration) */
  double sweep_ms;          /* summed sweep passes: each forward or backward
                               LU-SGS pass (or DP-LUR sweep) over all blocks,
                               including its one-thread prologue */
  int sweep_launches;       /* sweep kernel launches in those passes */
  int sweep_passes;
} ath_timing;
/* Timed runs launch every kernel eagerly (untimed runs replay each whole
 * nonlinear iteration from a captured HIP graph) so that events can bracket
 * every residual stage and sweep pass. */
int ath_enable_timing(ath_ctx *ctx, int on);
int ath_get_timing(const ath_ctx *ctx, ath_timing *out);

/* ---- host-only inspection (no device needed) --------------------------- */
/* Parse the case and build its host geometry exactly as ath_open does, then
 * return one field of block `block` (physical cells unless noted, i fastest):
 * which = 0 volume, 1..3 cell width in i/j/k (padded, ghosts included),
 * 4..6 face vectors in i/j/k (padded face arrays, 4 doubles per face: unit
 * normal and magnitude), 7 initial primitive state (5 per cell).  `count` is
 * the capacity of `out` in doubles; the needed size is returned in *needed. */
int ath_inspect_info(const char *inp_path, ath_info *out);
int ath_inspect_field(const char *inp_path, int block, int which, double *out,
                      long count, long *needed);

/* ---- kernel-level entry points used by the parity tests ---------------- */
/* Evaluate the device inviscid face flux for `count` faces.  flux: 0 Roe
 * (include/inviscidFlux.hpp:259-382), 1 AUSMPW+ (:395-481).  WL, WR: 5 per
 * face, n: 3 per face (unit normal), F: 5 per face.  Gas: nondimensional R
 * and n = Cv/R. */
int ath_eval_face_flux(int device, int flux, double gas_r, double gas_n,
                       const double *WL, const double *WR, const double *n,
                       int count, double *F);
/* The SST k-omega 2003 cell model (turbKWSst::EddyViscAndBlending, CalcTurbSrc
 * and TurbSrcJac, src/turbulence.cpp:573-760): per cell in[27] = primitive
 * state with k and omega (7), velocity gradient (9, the reference's tensor
 * order), k gradient (3), omega gradient (3), laminar viscosity, wall
 * distance, volume, viscous scaling and its inverse; out[7] = F1, F2, eddy
 * viscosity, k source, omega source, source Jacobian (0,0) and (1,1).  The
 * model alone: the RANS equations are not part of the iteration. */
int ath_eval_sst_cells(int device, const double *in, int count, double *out);
/* ath_eval_face_flux without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The inputs and outputs are staged in a fixed device array of the library
 * (9 MiB per device, allocated with the code object), so `count` is at most
 * 65536 (larger counts return ATH_ERR_ARG).  Every call on one device uses
 * the same array: calls that overlap on one device (from several threads)
 * race and must be serialised by the caller. */
int ath_eval_face_flux_static(int device, int flux, double gas_r, double gas_n,
                              const double *WL, const double *WR, const double *n,
                              int count, double *F);

/* ---- serving: many steady cases resident on one device ------------------
 * No reference counterpart: the reference runs one case per process
 * (src/main.cpp:231-302).  A server keeps every admitted case's state,
 * geometry and operator on the device and advances all running cases
 * together: one launch of each kernel covers the blocks of every running
 * case (the per-block unit grids of the single-case launches, concatenated
 * across cases), each case at its own outer step and CFL, each reduced to
 * its own norms by a tree that depends only on that case.  A case's results
 * are the same alone and in any mix of cases.  Every case of a server has
 * the same numerics (schemes, solver, gas and the implicit-operator
 * precision); its grid, boundary states and initial state are its own.
 * Cases are steady solves: implicit local time stepping, one nonlinear
 * iteration per outer step. */
typedef struct ath_server ath_server;

#define ATH_SLOT_RUNNING 0
#define ATH_SLOT_CONVERGED 1  /* the residual drop was reached */
#define ATH_SLOT_FAILED 2     /* `error` holds the ATH_ERR_* status */
#define ATH_SLOT_LIMIT 3      /* max_iterations done without converging */

typedef struct {
  int state;            /* ATH_SLOT_* */
  int error;            /* ATH_OK, or why the case failed */
  int iterations;       /* nonlinear iterations done */
  int num_blocks;
  double cells;         /* physical cells */
  double device_bytes;  /* device memory the case holds */
  double drop;          /* residual drop reached: the largest normalised L2
                           (PrintResiduals, src/output.cpp:1028-1063) over the
                           equations whose first residual carries a signal
                           (first L2 >= 1e-10 of the largest) */
  ath_norms last;       /* norms of the last iteration (matrix_resid is 0:
                           served cases skip the matrix residual) */
} ath_slot_info;

/* `hbm_budget`: device bytes the resident cases may reserve together (0: no
 * limit beyond the free device memory less a 2 GiB reserve); `max_slots`:
 * cases that may run at once; `precision`: ATH_IMPLICIT_FP64/FP32 for every
 * case. */
int ath_server_open(int device, double hbm_budget, int max_slots, int precision, ath_server **out);
void ath_server_close(ath_server *srv);
/* device bytes the resident cases have reserved (the sum of their
 * estimates), and the device's free memory */
int ath_server_memory(const ath_server *srv, double *used, double *free_hbm);
/* the device bytes admission reserves for a case: an upper bound of what it
 * allocates (no device needed) */
int ath_estimate_case_bytes(const char *inp_path, int precision, double *bytes);
/* Read the case (as ath_open does) and upload it if its estimated device
 * memory fits the free memory and the budget and a slot is free; else
 * ATH_ERR_NOMEM (nothing changed: try again after a retire).  The case runs
 * until its residual drop is <= `residual_drop` (checked after every batch
 * step) or `max_iterations` are done.  *slot identifies it from then on. */
int ath_admit(ath_server *srv, const char *inp_path, double residual_drop, int max_iterations, int *slot);
/* `iterations` nonlinear iterations of every running case, launched
 * together and replayed from a captured HIP graph; then each case's norms
 * are read and its state updated (converged, failed, limit).  A case that
 * fails (nonphysical state, singular block diagonal) is stopped alone; the
 * others go on.  *running: cases still running. */
int ath_step_batch(ath_server *srv, int iterations, int *running);
int ath_get_slot_info(ath_server *srv, int slot, ath_slot_info *out);
int ath_slot_block_dims(ath_server *srv, int slot, int block, int *ni, int *nj, int *nk);
/* physical-cell primitive state of block `block` of a case (as ath_get_state) */
int ath_get_slot_state(ath_server *srv, int slot, int block, double *prim);
int ath_set_slot_state(ath_server *srv, int slot, int block, const double *prim);
/* free the case's device memory and its slot */
int ath_retire(ath_server *srv, int slot);

#ifdef __cplusplus
}
#endif
#endif
