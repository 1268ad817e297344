/*
 * sifv_gpu.h - C ABI of the MI355X (gfx950) SemiImplicitFV hot path.
 *
 * Plain C linkage, plain pointers and sizes: no C++ or torch types cross
 * this boundary.  Every entry point returns an int status (SIFV_OK == 0);
 * sifv_last_error() returns a message for the most recent failure on the
 * calling thread.
 *
 * Field arrays use the padded, x-fastest layout of the reference's
 * RectilinearMesh::index (include/RectilinearMesh.hpp:70-73):
 *     idx = (i + ngx) + nxT * ((j + ngy) + nyT * (k + ngz))
 * with ghosts only in active dimensions, so one upload/download is one
 * contiguous copy of totalCells doubles per field.
 *
 * One handle owns all device memory of one solver instance and one HIP
 * stream.  Calls on one handle are not thread safe; distinct handles may be
 * driven from distinct host threads.
 */
#ifndef SIFV_GPU_H
#define SIFV_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SIFV_ABI_VERSION 3
#define SIFV_MAX_PHASES 8

/* ---- status codes ---------------------------------------------------- */
enum {
    SIFV_OK = 0,
    SIFV_ERR_INVALID = 1,      /* std::invalid_argument in the reference */
    SIFV_ERR_HIP = 2,          /* HIP runtime failure */
    SIFV_ERR_NONFINITE = 3,    /* NaN/Inf detected in the state */
    SIFV_ERR_UNSUPPORTED = 4,  /* configuration outside the GPU path */
    SIFV_ERR_NOMEM = 5
};

/* ---- enums (numeric values follow the reference's declaration order) -- */
/* ReconstructionOrder, include/SimulationConfig.hpp:13-20 */
enum { SIFV_WENO1 = 0, SIFV_WENO3 = 1, SIFV_WENO5 = 2,
       SIFV_UPWIND1 = 3, SIFV_UPWIND3 = 4, SIFV_UPWIND5 = 5 };
/* RiemannSolverType, include/RiemannSolver.hpp:16 */
enum { SIFV_LF = 0, SIFV_RUSANOV = 1, SIFV_HLLC = 2 };
/* BoundaryCondition, include/RectilinearMesh.hpp:14-20 */
enum { SIFV_BC_SYMMETRY = 0, SIFV_BC_OUTFLOW = 1, SIFV_BC_PERIODIC = 2,
       SIFV_BC_SLIPWALL = 3, SIFV_BC_NOSLIPWALL = 4 };
/* VarSet, include/SolutionState.hpp:15-19 */
enum { SIFV_VARSET_CONS = 0, SIFV_VARSET_PRIM = 1, SIFV_VARSET_ALL = 2 };
/* Face ids, include/RectilinearMesh.hpp:110-115 */
enum { SIFV_XLOW = 0, SIFV_XHIGH = 1, SIFV_YLOW = 2, SIFV_YHIGH = 3,
       SIFV_ZLOW = 4, SIFV_ZHIGH = 5 };
/* EOS kinds: IdealGasEOS (src/IdealGasEOS.cpp), StiffenedGasEOS
 * (src/StiffenedGasEOS.cpp) */
enum { SIFV_EOS_IDEAL = 0, SIFV_EOS_STIFFENED = 1 };
/* Arithmetic of the bulk WENO/flux work.  SIFV_FP64 is the plain method;
 * SIFV_MIXED (explicit solver) stores the state in FP32 and runs WENO/flux
 * arithmetic in FP32; a stiffened gas or two-phase state keeps its
 * energy<->pressure conversion in FP64 and stores its energy without the
 * stiffened-gas constant sum_k alpha_k gamma_k pInf_k/(gamma_k-1), so the FP32
 * state keeps the pressure part; an ideal gas converts and updates in FP32.
 * With the semi-implicit solver SIFV_MIXED keeps the FP64 state and runs the
 * Jacobi sweeps in FP32 with FP64 residual refinement, or the multigrid
 * V-cycle of the MG-PCG pressure solve in FP32 (sifv_mgpcg_solve). */
enum { SIFV_FP64 = 0, SIFV_MIXED = 1 };

/* Field ids for upload/download (SolutionState members,
 * include/SolutionState.hpp:73-113). */
enum {
    SIFV_F_RHO = 0, SIFV_F_RHOU = 1, SIFV_F_RHOV = 2, SIFV_F_RHOW = 3,
    SIFV_F_RHOE = 4, SIFV_F_VELU = 5, SIFV_F_VELV = 6, SIFV_F_VELW = 7,
    SIFV_F_PRES = 8, SIFV_F_SIGMA = 9,
    SIFV_F_ALPHARHO = 16, /* + phase, phase < SIFV_MAX_PHASES */
    SIFV_F_ALPHA = 24     /* + phase */
};

/* ---- configuration mirror -------------------------------------------- */
typedef struct sifv_phase_eos { double gamma, pInf; } sifv_phase_eos;

/* Mirror of SimulationConfig and its *Params structs
 * (include/SimulationConfig.hpp:22-91).  The body force is applied in the
 * stage update of the explicit and semi-implicit solvers at the step's config
 * time; the viscous and capillary stresses in the stage update of both
 * solvers (surface tension needs a multiphase state); the ensemble and the
 * adjoint carry the flux terms only. */
typedef struct sifv_config {
    int dim, nGhost, RKOrder, useIGR, semiImplicit, reconOrder;
    double wenoEps;
    int step;
    double time;
    struct { double cfl, constDt, maxDt, minDt; } explicitParams;
    struct { double cfl, maxDt, minDt; int maxPressureIters; double pressureTol; } semiImplicitParams;
    struct { double alphaCoeff; int IGRIters, IGRWarmStartIters; } igrParams;
    struct { int nPhases; sifv_phase_eos phases[SIFV_MAX_PHASES]; double alphaMin; } multiPhaseParams;
    struct { double a[3], b[3], c[3], d[3]; } bodyForceParams;
    double viscousMu;
    double surfaceTensionSigma;
    /* ViscousParams::phaseMu (empty: viscousNPhaseMu = 0, the scalar mu) */
    int viscousNPhaseMu;
    double viscousPhaseMu[SIFV_MAX_PHASES];
    /* SurfaceTensionParams::epsGradAlpha (> 0; the reference's default 1e-8) */
    double surfaceTensionEps;
} sifv_config;

/* Mirror of the RectilinearMesh constructor arguments plus the boundary
 * setup done through Runtime::setBoundaryCondition / createUniformMesh
 * periods (src/Runtime.cpp:39-127, src/RectilinearMesh.cpp:14-56).
 * yNodes/zNodes may be NULL for inactive dimensions ({0,1} is used). */
typedef struct sifv_mesh {
    int dim, nx, ny, nz, nGhost;
    const double* xNodes;   /* nx+1 physical node coordinates */
    const double* yNodes;   /* ny+1 */
    const double* zNodes;   /* nz+1 */
    int bc[6];              /* SIFV_BC_* per face */
    int periods[3];         /* MPI-style periodicity flags (wrap on one device) */
} sifv_mesh;

typedef struct sifv_eos { int kind; double gamma, pInf; } sifv_eos;

typedef struct sifv_handle sifv_handle;

/* ---- life cycle ------------------------------------------------------- */
/* Replaces the ExplicitSolver / SemiImplicitSolver constructors
 * (include/ExplicitSolver.hpp:23-29, include/SemiImplicitSolver.hpp:24-31,
 * src/ExplicitSolver.cpp:16-66): riemann is the enum the reference derives
 * from RiemannSolver::name() (src/ExplicitSolver.cpp:31-34).  Validates the
 * config as SimulationConfig::validate (include/SimulationConfig.hpp:123-209). */
int sifv_create(const sifv_mesh* mesh, const sifv_config* cfg, int riemann,
                const sifv_eos* eos, int precision, sifv_handle** out);
int sifv_destroy(sifv_handle* h);
const char* sifv_last_error(void);
int sifv_abi_version(void);

/* Number of padded cells (RectilinearMesh::totalCells,
 * src/RectilinearMesh.cpp:83-85). */
int sifv_total_cells(const sifv_handle* h, size_t* n);

/* ---- state transfer (SolutionState member vectors) -------------------- */
int sifv_upload(sifv_handle* h, int field, const double* host, size_t n);
int sifv_download(sifv_handle* h, int field, double* host, size_t n);

/* ---- explicit SSP-RK step --------------------------------------------- */
/* Replaces ExplicitSolver::step (src/ExplicitSolver.cpp:68-203): acoustic
 * dt (or constDt), dt = min(dt, targetDt) when targetDt > 0, RKOrder stages
 * of cons->prim, ghost fill, [IGR], RHS and stage update, then the final
 * cons->prim and ghost fill.  Blocks until dt is known; *dt_out = dt. */
int sifv_explicit_step(sifv_handle* h, const sifv_config* cfg, double targetDt,
                       double* dt_out);

/* Runs up to nsteps explicit steps with device-resident dt and time, the
 * way runTimeLoop does (src/RKTimeStepping.cpp:259-274) with targetDt =
 * endTime - time and no output in between.  Stops early when time reaches
 * endTime.  *time_inout is advanced, *steps_done counts the steps taken.
 * One host synchronisation at the end. */
int sifv_explicit_advance(sifv_handle* h, const sifv_config* cfg, int nsteps,
                          double endTime, double* time_inout, int* steps_done);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Runs the same steps as sifv_explicit_advance with the same
 * kernels and results, but calls no hipMalloc, hipFree or allocating library
 * routine: the dt block partials live in one static device array of the
 * library (65536 doubles, held per GPU until the process ends) instead of the
 * handle's lazily allocated buffer.  Every buffer the steps need is allocated
 * by sifv_create.  Returns SIFV_ERR_INVALID when the mesh needs more than
 * 65536 partials per step, SIFV_ERR_UNSUPPORTED with kernel timing on.  All
 * calls on one GPU share the array: calls that overlap on one GPU (from
 * several host threads) race.  The primitive arrays are refreshed at the end
 * of the call, so the next step on the handle takes its dt from them. */
int sifv_explicit_advance_noalloc(sifv_handle* h, const sifv_config* cfg, int nsteps,
                                  double endTime, double* time_inout, int* steps_done);

/* ---- semi-implicit step ------------------------------------------------ */
/* Replaces SemiImplicitSolver::step (src/SemiImplicitSolver.cpp:106-236)
 * with the Jacobi PressureSolver, or the PCG one after
 * sifv_set_pressure_solver; *pressure_iters is lastPressureIterations() of
 * the last stage. */
int sifv_semi_implicit_step(sifv_handle* h, const sifv_config* cfg, double targetDt,
                            double* dt_out, int* pressure_iters);

/* ---- pressure solver --------------------------------------------------- */
/* Replaces JacobiPressureSolver::solve (src/JacobiPressureSolver.cpp:9-55):
 * host arrays of totalCells doubles in, pressure updated in place (interior
 * cells; ghosts as left by the last sweep's ghost fill).  Returns the
 * iteration count through *iters (maxIter when not converged, as the
 * reference does).  residual_history may be NULL or hold maxIter doubles. */
int sifv_jacobi_solve(sifv_handle* h, const double* rho, const double* rhoc2,
                      const double* rhs, double* pressure, double dt,
                      double tolerance, int maxIter, int* iters,
                      double* residual_history);

/* Pressure solver kinds of the semi-implicit step. */
#define SIFV_PS_JACOBI 0
#define SIFV_PS_PCG 1
#define SIFV_PS_MGPCG 2

/* Replaces HyprePressureSolver::solve (src/HyprePressureSolver.cpp:312-408):
 * the same system as sifv_jacobi_solve, solved by conjugate gradients on
 * its volume-weighted symmetric form with a diagonal preconditioner (the
 * reference's StructPCG role, with a diagonal instead of a PFMG V-cycle);
 * stops when ||rhs - A p||_2 <= tolerance ||rhs||_2 as the reference's
 * two-norm setting does.  Only interior cells of pressure are written;
 * *iters = iterations (0 when the initial guess passes, maxIter when not
 * converged).  SIFV_ERR_UNSUPPORTED on a periodic axis whose ghost widths
 * differ from the wrapped cells' (the system is not symmetric there). */
int sifv_pcg_solve(sifv_handle* h, const double* rho, const double* rhoc2,
                   const double* rhs, double* pressure, double dt,
                   double tolerance, int maxIter, int* iters);

/* Replaces HyprePressureSolver::solve (src/HyprePressureSolver.cpp:312-408)
 * in its StructPCG + PFMG form (src/HyprePressureSolver.cpp:123-138): as
 * sifv_pcg_solve, preconditioned by one geometric-multigrid V-cycle
 * (weighted-Jacobi relaxation, one pre- and one post-sweep, 2:1 coarsening of
 * every axis down to at most 4 cells per axis, piecewise-constant transfers,
 * rediscretised coarse operators) instead of the diagonal.  Same arguments,
 * stopping rule, iteration count and refusals as sifv_pcg_solve.  On a
 * SIFV_MIXED semi-implicit handle the V-cycle runs in FP32 (operator rounded
 * to FP32 once per solve) inside FP64 conjugate gradients, whose residual
 * and stopping test stay FP64; the iteration count may then differ from the
 * FP64 form's by a few. */
int sifv_mgpcg_solve(sifv_handle* h, const double* rho, const double* rhoc2,
                     const double* rhs, double* pressure, double dt,
                     double tolerance, int maxIter, int* iters);

/* Chooses the PressureSolver of sifv_semi_implicit_step (the reference's
 * SemiImplicitSolver constructor argument, include/SemiImplicitSolver.hpp:24-31):
 * SIFV_PS_JACOBI (default), SIFV_PS_PCG (FP64 handles only) or
 * SIFV_PS_MGPCG (on a SIFV_MIXED handle with the FP32 V-cycle of
 * sifv_mgpcg_solve). */
int sifv_set_pressure_solver(sifv_handle* h, int kind);

/* ---- ghost fill and time step ------------------------------------------ */
/* Replaces RectilinearMesh::applyBoundaryConditions(state, varSet, halo)
 * (src/RectilinearMesh.cpp:498-524) on the device-resident state. */
int sifv_fill_ghosts(sifv_handle* h, int varset);
/* Replaces RectilinearMesh::fillScalarGhosts(field, halo)
 * (src/RectilinearMesh.cpp:526-624) for one device-resident field. */
int sifv_fill_scalar_ghosts(sifv_handle* h, int field);
/* Replaces computeAcousticTimeStep(mesh, state, eos, config, cfl, maxDt, comm)
 * (src/RKTimeStepping.cpp:43-159) on the device-resident primitive fields. */
int sifv_acoustic_dt(sifv_handle* h, double cfl, double maxDt, double* dt);
/* Replaces computeAdvectiveTimeStep (src/RKTimeStepping.cpp:15-41,125-134). */
int sifv_advective_dt(sifv_handle* h, double cfl, double maxDt, double* dt);
/* Recomputes the primitive fields from the conservative ones on interior
 * cells and fills ghosts (SolutionState::convertConservativeToPrimitiveVariables,
 * src/SolutionState.cpp:187-205, or MixtureEOS::convertConservativeToPrimitive,
 * src/MixtureEOS.cpp:99-143, followed by the PRIM ghost fill). */
int sifv_cons_to_prim(sifv_handle* h);

/* ---- measurement ---------------------------------------------------------- */
/* When enabled, the library records HIP events around every launch of the
 * fused RHS+stage kernel on its own stream.  sifv_timing_read returns the
 * summed kernel time and the launch count since the last reset. */
int sifv_timing_enable(sifv_handle* h, int on);
int sifv_timing_read(sifv_handle* h, double* kernel_ms, long* launches, int reset);
int sifv_synchronize(sifv_handle* h);
/* Launch-shape override for the fused stage kernel (rows marched per wave);
 * 0 restores the built-in choice.  Used by the autotuner. */
int sifv_set_tile(sifv_handle* h, int rows_per_wave);
/* 3D stages: the number of slabs of z segments a stage is split into, the x+y
 * pass of slab k+1 running on the library's second stream beside the z pass of
 * slab k (1..8; 1 = one launch of each pass).  0 restores the built-in choice
 * (4 slabs when each slab's z launch keeps >= 8192 waves, else 1).  Results are
 * the same bits for every value.  No counterpart in the reference (a
 * launch-shape switch). */
int sifv_set_stage_slabs(sifv_handle* h, int slabs);
/* sifv_explicit_advance replays one captured RK step (dt finalize + stage
 * launches) as a HIP graph per step when the step's launch arguments repeat
 * (RK2/RK3 without IGR, timing off), and the 2D/3D Jacobi pressure solve
 * replays each chunk of sweeps between residual reads from a graph captured
 * once per chunk and density buffer; on = 0 enqueues every launch instead.
 * Default off (plain launches measured 4-5% faster per explicit step, and
 * within 1.4% either way for 1000-sweep Jacobi solves at 64^3-256^3, on
 * MI355X).  No counterpart in the reference (a launch-mechanism switch).
 * sifv_graph_replays: graph launches since the handle was created. */
int sifv_set_graphs(sifv_handle* h, int on);

/* ---- output of the resident state (VTK XML, ASCII) ------------------------- */
/* VTKWriter::writeVTR (src/VTKWriter.cpp:21-192) of the handle's current state:
 * downloads the fields it writes (Pressure, Sigma with IGR, TotalEnergy,
 * Velocity, Momentum, Density or AlphaRho_k / Alpha_k) and writes them with
 * the reference's layout and 15-digit formatting; rank >= 0 adds the Rank
 * array as VTKSession::write does on that rank (src/VTKSession.cpp:23-31),
 * rank < 0 is the serial form.  Parent directories are created. */
int sifv_write_vtr(sifv_handle* h, const char* filename, int rank);
/* VTKWriter::writePVTR (src/VTKWriter.cpp:194-243) for the one-piece
 * decomposition whose piece file is piece_file. */
int sifv_write_pvtr(const sifv_handle* h, const char* filename, const char* piece_file);
/* VTKWriter::writePVD (src/VTKWriter.cpp:245-291): mode "w" (new
 * collection), "a" (append a dataset at `time`), "close" (no-op). */
int sifv_write_pvd(const char* filename, const char* mode, double time, const char* data_file);
int sifv_graph_replays(const sifv_handle* h, long* replays);

/* ---- discrete adjoint ----------------------------------------------------- */
/* No counterpart in the reference (it has no derivatives).  The transpose of
 * the explicit solver's step map for doubly periodic 2D meshes with WENO5,
 * Rusanov, SSP-RK3, a fixed time step (explicitParams.constDt > 0), single-
 * phase ideal gas, FP64.  sifv_adjoint_forward runs nsteps steps from the
 * resident state (as sifv_explicit_step would) and keeps checkpoints every
 * checkpoint_every steps; sifv_adjoint_backward recomputes each segment from
 * its checkpoint and returns grad = (dU^N/dU^0)^T seed and
 * dgamma = seed . dU^N/dgamma.  seed / grad: 4 fields (rho, rhoU, rhoV, rhoE)
 * x totalCells, padded layout, device pointers (the _host form takes host
 * arrays); grad is 0 on ghost cells. */
int sifv_adjoint_forward(sifv_handle* h, const sifv_config* cfg, int nsteps, int checkpoint_every);
int sifv_adjoint_backward(sifv_handle* h, const double* seed, double* grad, double* dgamma);
int sifv_adjoint_backward_host(sifv_handle* h, const double* seed, double* grad, double* dgamma);
/* Device pointer of the resident conservative state (field q at q*totalCells). */
int sifv_state_device_ptr(sifv_handle* h, double** ptr);

/* ---- batched 1D runs ------------------------------------------------------ */
/* No counterpart in the reference, which runs one problem per process: many
 * independent 1D explicit runs (each what ExplicitSolver::step driven by
 * runTimeLoop without output does on its own 1D mesh with outflow ends,
 * src/ExplicitSolver.cpp:68-203, src/RKTimeStepping.cpp:241-356) advanced to
 * one end time in a single persistent launch.  Tubes have uneven cell counts;
 * fields are packed tube after tube (interior cells only, no ghosts).
 * nodes: ncells[t]+1 node coordinates per tube, packed.  cfg: dim 1, WENO5,
 * RKOrder 2 or 3, single phase, explicit, no IGR (cfl, maxDt, constDt,
 * wenoEps are read).  precision: SIFV_FP64 or SIFV_MIXED. */
typedef struct sifv_ensemble sifv_ensemble;
int sifv_ensemble_create(int ntubes, const int* ncells, const double* nodes, const sifv_config* cfg, int riemann,
                         const sifv_eos* eos, int precision, sifv_ensemble** out);
int sifv_ensemble_destroy(sifv_ensemble* e);
int sifv_ensemble_total_cells(const sifv_ensemble* e, size_t* n);
/* fields: SIFV_F_RHO, SIFV_F_RHOU, SIFV_F_RHOE (state), SIFV_F_VELU, SIFV_F_PRES
 * (initial primitives: the first time step reads them, as the reference's
 * first step reads SolutionState's primitive arrays).  Download of VELU/PRES
 * converts the current state. */
int sifv_ensemble_upload(sifv_ensemble* e, int field, const double* host, size_t n);
int sifv_ensemble_download(sifv_ensemble* e, int field, double* host, size_t n);
/* Tile shape: workgroup size (64..1024, one cell per thread per tile) and
 * resident workgroups per CU. */
int sifv_ensemble_set_shape(sifv_ensemble* e, int block, int wg_per_cu);
/* Advances every tube from time 0 to endTime (at most maxSteps steps each);
 * steps/times (ntubes entries, may be NULL) receive each tube's step count and
 * final time, kernel_ms the launch's duration. */
int sifv_ensemble_run(sifv_ensemble* e, double endTime, int maxSteps, int* steps, double* times, double* kernel_ms);
/* Runs every candidate shape on the current state (restored after each
 * trial, at most maxSteps steps per tube) and keeps the fastest. */
int sifv_ensemble_autotune(sifv_ensemble* e, double endTime, int maxSteps, int* block_out, int* wg_out,
                           double* best_ms);

#ifdef __cplusplus
}
#endif

#endif /* SIFV_GPU_H */
