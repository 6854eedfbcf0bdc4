/* C ABI of the MI355X CongFD right-hand side and SSP-RK3 time advance.
 *
 * The reference solver (wanghongfei23/CFD_inWSL, zone_FD) has no FFI: its
 * seam is SpDistributor::rhsSolve() (zone_FD/src/src/sp_distributor.cpp:3),
 * called from BlockSolver::RK3_SSP (zone_FD/src/src/blockSolver.cpp:43-87)
 * after Equation::consToPrim and Bnds::update.  Each entry point below names
 * the reference routine it stands in for.
 *
 * Conventions
 *   - All functions return 0 on success and a negative CFD_E_* code on error;
 *     cfd_last_error() returns a message for the calling thread.
 *   - Host arrays are in the reference's layout: array-of-structures
 *     [cell][nv] in double precision with x fastest, cell index
 *     i + j*nx + k*nx*ny (zone_FD/src/include/macro.hpp:93-96).
 *     nv = dim + 2: (rho, rho u[, rho v[, rho w]], rho E).
 *   - A handle owns one HIP stream and its device buffers (structure of
 *     arrays, one plane per variable).  Handles are not thread-safe; calls on
 *     one handle are ordered on its stream.
 *   - The right-hand side follows the reference's sign: rhs = +sum_d dF_d/dx_d,
 *     so dU/dt = -rhs.
 */
#ifndef CONGFD_H
#define CONGFD_H

#ifdef __cplusplus
extern "C" {
#endif

#define CFD_OK 0
#define CFD_E_ARG (-1)     /* invalid argument or configuration */
#define CFD_E_HIP (-2)     /* HIP runtime error */
#define CFD_E_UNSUPP (-3)  /* scheme / option not implemented */

/* Reference enum values (zone_FD/src/include/macro.hpp:33-90). */
#define CFD_FLUX_HLLC 0     /* FluxMethod::HLLC */
#define CFD_FLUX_ROE 1      /* FluxMethod::ROE */
#define CFD_DIFF_MND6 3     /* DiffMethod::MND6 */
#define CFD_INTER_WCNS5 2   /* InterMethod::WCNS5 -> weno5_JSchen */
#define CFD_INTER_WCNSZ5 3  /* InterMethod::WCNSZ5 -> weno5_Z */
#define CFD_INTER_WCNS5CONGZ 8 /* -> Teno5_CongZ (TENO-S) */
#define CFD_INTER_TCNS5 11  /* -> Teno5_Z (TENO5) */
#define CFD_INTER_TCNS5CT4 12
#define CFD_INTER_TCNS5CT7 13

/* Reconstruction variables.  The reference's Euler path is always
 * characteristic-wise (SolverType::initReconer, zone_FD/src/src/solverType.cpp:217-594);
 * component-wise applies Recon5OrderFaceCenter (reconstructor5order.hpp:128-151)
 * to the primitive variables. */
#define CFD_RECON_CHAR 0
#define CFD_RECON_COMPONENT 1

/* Boundary codes: reference BndType values (macro.hpp:33-34) plus extensions. */
#define CFD_BC_PERIODIC 1        /* true periodic halo */
#define CFD_BC_SYMMETRY1D 2      /* mirror, u negated */
#define CFD_BC_DIRICHLET 3       /* fixed primitive ghost state (bc_prim) */
#define CFD_BC_OUTFLOW 7         /* SUPERSONICOUTLET: edge cell copied */
#define CFD_BC_SYMMETRYX 10      /* mirror, u negated */
#define CFD_BC_SYMMETRYY 11      /* mirror, v negated */
#define CFD_BC_SYMMETRYZ 101     /* extension: mirror, w negated */
#define CFD_BC_MIRROR_COPY 102   /* the reference's 2D Euler PERIODIC1D wiring
                                    (initializer.cpp:806-822): mirror, no flip */
#define CFD_BC_DMR_TOP 8         /* DoubleMachUp (oneDBnd.cpp:87-108), 2D y-high face:
                                    bc_prim = post-shock state where y > sqrt(3)(x - g(t)),
                                    (1.4, 0, 0, 1) elsewhere; needs origin and the time */
#define CFD_BC_DMR_WALL 103      /* 2D y-low face of the double-Mach wiring
                                    (initializer.cpp:916-934): Dirichlet bc_prim for
                                    cell centres x < 1/6, SYMMETRYY beyond */

typedef struct cfd_config {
  int dim;              /* 1, 2 or 3 (Info::dim) */
  int n[3];             /* cells per direction (Info::icMax(); unused = 1) */
  double h[3];          /* spacing per direction (Info::geth; the reference uses
                           one spacing for every direction) */
  int inter_method;     /* Info::interMethod */
  int flux_method;      /* Info::fluxMethod */
  int diff_method;      /* Info::diffMethod (MND6 only) */
  int recon;            /* CFD_RECON_* */
  int bc[6];            /* per face: x-, x+, y-, y+, z-, z+ */
  double bc_prim[6][5]; /* Dirichlet ghost state per face, primitive variables */
  int device;           /* HIP device ordinal */
  int precision;        /* CFD_PRECISION_FP64 (reference arithmetic, bitwise),
                           CFD_PRECISION_FP32 (FP32 storage and arithmetic of the
                           sweeps; U^n and the RK3 stage combination stay FP64;
                           dim 1-3, single zones and batches) or
                           CFD_PRECISION_FP64_FMA (below; single zones) */
  double origin;        /* cell centres at origin + (i + 0.5) h[0] in every direction
                           (Initializer::initUniformBlock, initializer.cpp:672-677);
                           used by the double-Mach boundaries */
  double gravity[3];    /* gravity source (SourceTerm::calGravitySource,
                           sourceTerm.cpp:3-20): rhs(rho u_d) -= rho g_d,
                           rhs(rho E) -= rho (u . g); all zero: no source (the
                           reference's RTI case uses g = (0, 1) when enabled) */
} cfd_config;

#define CFD_PRECISION_FP64 0
#define CFD_PRECISION_FP32 1
/* FP64 storage and arithmetic with FMA contraction (a*b + c rounded once)
 * in the sweeps, and divisions / square roots with one Newton step fewer
 * (within an ulp).  Opt-in: results differ from the reference arithmetic by
 * rounding (one rhs within 1e-12 of S_v * max(|u| + c) / h per variable,
 * S_v the variable's magnitude scale; a branch predicate tied at rounding
 * level -- a TENO cut-off, an HLLC wave region -- may be decided the other
 * way).  Single zones, WENO-JS / WENO-Z / TENO5 (and its cut-off variants) /
 * TENO-S, any flux, dimension and recon; no adjoint.  No counterpart in the
 * reference, which has one arithmetic. */
#define CFD_PRECISION_FP64_FMA 2

typedef struct cfd_handle cfd_handle;

const char* cfd_last_error(void);
const char* cfd_version(void);

/* Builds the solver: stands in for BlockSolver(Info*) with
 * SolverType(info) scheme selection (zone_FD/src/src/blockSolver.cpp:24-41,
 * solverType.cpp:20-27) and Initializer::initBnds (initializer.cpp:743-952). */
int cfd_create(const cfd_config* cfg, cfd_handle** out);
int cfd_destroy(cfd_handle* h);

/* Conservative state in / out (Equation::cons, equation.hpp). */
int cfd_set_state(cfd_handle* h, const double* cons_aos);
int cfd_get_state(cfd_handle* h, double* cons_aos);

/* rhs of the current state: consToPrim + Bnds::update + SpDistributor::rhsSolve
 * (blockSolver.cpp:51-54).  rhs_aos may be NULL (result stays on the device). */
int cfd_eval_rhs(cfd_handle* h, double* rhs_aos);

/* One SSP-RK3 step with a given dt: BlockSolver::RK3_SSP (blockSolver.cpp:43-87). */
int cfd_rk3_step(cfd_handle* h, double dt);

/* nsteps RK3 steps, each with dt = cfl / sum_d max (c + |u_d|)/h_d computed on
 * the device (BlockSolver::getTimeIntervalExplicit, blockSolver.cpp:318-348,
 * without the outputDt clamp of stepsLoopCFL).  t_out (may be NULL) receives
 * the elapsed simulated time; the call is asynchronous when t_out is NULL. */
int cfd_rk3_steps_cfl(cfd_handle* h, double cfl, int nsteps, double* t_out);

/* BlockSolver::stepsLoopCFL (blockSolver.cpp:170-233) on the device up to the
 * next output time: each step takes the CFL dt, clamped as the reference
 * clamps it (if output_t + dt > output_dt: dt = output_dt - output_t,
 * output_t = 0, an output is due; else output_t += dt), advances RK3 and the
 * clock, and records (t, dt, rhoLinf) -- the per-step log line of the
 * reference.  Stops after the step that reaches the output time or after
 * max_steps steps.  output_t (in/out) is the reference's Info::outputT; log
 * (host, 3 * max_steps doubles) receives the records of *nsteps steps;
 * *reached is 1 when an output time was reached.  Steps are queued in groups
 * without host synchronisation; the steps of a group after the output step
 * do nothing. */
int cfd_steps_to_output(cfd_handle* h, double cfl, double output_dt, double* output_t, int max_steps,
                        double* log, int* nsteps, int* reached);

/* cfd_rk3_steps_cfl without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The same steps, dt
 * and clock as cfd_rk3_steps_cfl(h, cfl, nsteps, NULL), asynchronous; the
 * per-workgroup CFL maxima that cfd_rk3_steps_cfl keeps in a buffer it
 * allocates on demand go to a fixed device array of the library instead:
 * 32768 doubles (256 KiB) per GPU, held until the process ends.  A launch
 * plan whose last sweep needs more (D x workgroups above 32768) is refused
 * with CFD_E_ARG.  The array is shared by every call on one GPU: calls that
 * overlap on the same GPU (other handles, other streams) race. */
int cfd_rk3_steps_cfl_noalloc(cfd_handle* h, double cfl, int nsteps);

/* When keep != 0 the RK3 calls also store the rhs of each stage (the last
 * stage's rhs stays readable, as the reference's Equation::rhs does after
 * RK3_SSP); off by default, which saves one rhs write per step. */
int cfd_set_keep_rhs(cfd_handle* h, int keep);

/* max_i |rhs[i][var]| of the stored rhs: Data::getLinf
 * (zone_FD/src/src/data.cpp:230), the per-step "rhoLinf" log value of
 * BlockSolver::stepsLoopCFL (blockSolver.cpp:204-206) for var = 0. */
int cfd_rhs_linf(cfd_handle* h, int var, double* out);

/* CFL time step of the current state (getTimeIntervalExplicit). */
int cfd_cfl_dt(cfd_handle* h, double cfl, double* dt);

/* Kahan-compensated domain totals of the conservative variables (sum over
 * cells, no cell volume).  No counterpart in the reference (it has no
 * conserved-total diagnostic); used to check conservation. */
int cfd_conserved_totals(cfd_handle* h, double* out_nv);

/* Waits for all work queued on the handle's stream. */
int cfd_synchronize(cfd_handle* h);

/* Measures the direction sweeps of one RK stage with HIP events on the
 * handle's stream: runs stage 1 from the current state into a scratch buffer
 * (the state is left unchanged) `reps` times; ms_out[d] receives the mean
 * duration of the direction-d sweep launch.  Diagnostic; no reference
 * counterpart (the reference times whole steps with std::chrono,
 * zone_FD/src/src/blockSolver.cpp:194-200). */
int cfd_time_sweeps(cfd_handle* h, int reps, double* ms_out);

/* Plain device copy of the state planes (nv * cells doubles) into a scratch
 * buffer, `reps` times: a reference kernel of known traffic for calibrating
 * hardware byte counters.  Diagnostic; no reference counterpart. */
int cfd_copy_probe(cfd_handle* h, int reps);

/* Chooses the tile shape and segment length of each direction's sweep by
 * timing the candidates of the launch model on the current state (reps
 * launches each, each direction with the work it does in a step: the last
 * sweep with a stage update into a scratch state, which leaves the state
 * unchanged).  tiles_out / segs_out (3 ints each, may be NULL) receive the
 * choice.  No reference counterpart. */
int cfd_autotune(cfd_handle* h, int reps, int* tiles_out, int* segs_out);

/* Current launch plan: tile shape and segment length per direction. */
int cfd_launch_plan(cfd_handle* h, int* tiles_out, int* segs_out);

/* Sets the launch plan (3 ints each; entries beyond dim ignored).  Tile codes:
 * x sweeps 0 (32x8) or 1 (16x16); y/z sweeps 1 (16x16) or 2 (8x32); the
 * segment length is 1..n[d].  Results do not depend on the plan. */
int cfd_set_launch_plan(cfd_handle* h, const int* tiles, const int* segs);

/* ---- batches of independent zones ---------------------------------------
 * Many independent structured zones (separate cases or blocks of uneven
 * sizes) advanced together.  The reference solves one block per process
 * (BlockSolver); this has no reference counterpart beyond running
 * BlockSolver::stepsLoopCFL on each zone.  Every zone takes its own CFL time
 * step.  Scheme, flux, recon, dim and precision (CFD_PRECISION_FP64 or
 * CFD_PRECISION_FP32) come from `common`; n, h and boundaries from each
 * zone. */
typedef struct cfd_zone {
  int n[3];
  double h[3];
  int bc[6];
  double bc_prim[6][5];
} cfd_zone;

typedef struct cfd_batch cfd_batch;

#define CFD_BATCH_GRAPH 0        /* persistent work-queue launch per sweep, RK3 step replayed as a hipGraph */
#define CFD_BATCH_PERSISTENT 1   /* persistent work-queue launches, no graph */
#define CFD_BATCH_PER_ZONE 2     /* the single-zone kernels launched zone after zone */

int cfd_batch_create(const cfd_config* common, int nzones, const cfd_zone* zones, cfd_batch** out);
int cfd_batch_destroy(cfd_batch* b);
int cfd_batch_set_state(cfd_batch* b, int zone, const double* cons_aos);
int cfd_batch_get_state(cfd_batch* b, int zone, double* cons_aos);
/* nsteps RK3 steps of every zone with its own device CFL dt (see
 * cfd_rk3_steps_cfl); asynchronous. */
int cfd_batch_steps_cfl(cfd_batch* b, double cfl, int nsteps, int mode);
/* Simulated time reached by each zone (nzones doubles). */
int cfd_batch_times(cfd_batch* b, double* t_out);
/* Picks the tile shape per direction and the persistent grid size by timing. */
int cfd_batch_autotune(cfd_batch* b, int reps);
int cfd_batch_synchronize(cfd_batch* b);

/* ---- discrete adjoint -----------------------------------------------------
 * Reverse-mode derivative of the discrete RHS and of fixed-dt SSP-RK3 steps,
 * hand-written (the reference has no adjoint; the map differentiated is
 * SpDistributor::rhsSolve, zone_FD/src/src/sp_distributor.cpp:9-78, with
 * BlockSolver::RK3_SSP, zone_FD/src/src/blockSolver.cpp:43-87).  Covers 2D,
 * FP64, characteristic-wise WCNS5 (WENO-JS) or a TENO-family scheme (TENO5,
 * TENO-S, TENO-A, the whf_* variants; not musclIn5) + HLLC, any boundary
 * codes; other configurations return CFD_E_UNSUPP.  Branches (HLLC wave
 * pattern, pressure clipping, positivity fallback, TENO stencil flags) take
 * the derivative of the branch the forward evaluation takes.  Arrays are host
 * AoS [cell][4].  A direction periodic on both faces is differentiated as a
 * ring, marched in segments or (where the planner finds it cheaper) one whole
 * pencil per lane group; the environment variable CFD_ADJ_FULL = 0 or 1, read
 * when a handle's first adjoint call sets up its buffers, forces one march
 * (both give the same gradient up to rounding). */

/* grad = (d rhs / d U)^T cot at the current state. */
int cfd_rhs_vjp(cfd_handle* h, const double* cot, double* grad);

/* Advances the current state by nsteps RK3 steps of the fixed dt, keeping the
 * state every checkpoint_every steps on the device for cfd_adjoint_backward
 * (which then recomputes each segment's stage states).  checkpoint_every = 0
 * keeps every stage state instead (3 nsteps + 1 states of device memory, no
 * recomputation). */
int cfd_adjoint_forward(cfd_handle* h, double dt, int nsteps, int checkpoint_every);

/* Pulls the cotangent of the state after the last cfd_adjoint_forward back to
 * its initial state: grad_initial = (d U(t_end) / d U(0))^T cot_final.  Each
 * checkpoint segment's stage states are recomputed on the device. */
int cfd_adjoint_backward(cfd_handle* h, const double* cot_final, double* grad_initial);

/* Simulated time of the handle (Info::t): set it before time-dependent
 * boundaries are used; each RK3 step advances it as BlockSolver::RK3_SSP +
 * solve do (t + dt - dt/2 - dt/2 + dt, blockSolver.cpp:62-123). */
int cfd_set_time(cfd_handle* h, double t);
int cfd_get_time(cfd_handle* h, double* t);

/* Device pointer of the current state (structure of arrays, nv planes of
 * n[0]*n[1]*n[2] doubles) and of the rhs buffer.  FP64 handles only: an FP32
 * handle (float rhs, sweeps reading an FP32 shadow of the state) returns
 * CFD_E_UNSUPP. */
int cfd_device_buffers(cfd_handle* h, double** state, double** rhs);

#ifdef __cplusplus
}
#endif
#endif /* CONGFD_H */
