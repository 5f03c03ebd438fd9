/*
 * tfi.h -- C-ABI of libtfi_hip.so, the MI355X-native body-fitted IGA
 * Navier-Stokes hot path.
 *
 * The reference has no library or FFI for this path: its GPU backend is the
 * process `iga_cuda navier-stokes` (solvers/cuda/src/iga_cuda.cu:734-1153)
 * built from header-only CUDA/C++ classes.  Each entry point below names the
 * reference routine it replaces.  Conventions:
 *   - every function returns 0 on success, non-zero on failure; the message is
 *     available from tfi_last_error(ctx) (no exceptions cross the ABI);
 *   - host arrays are borrowed for the duration of the call only;
 *   - device memory is owned by the library objects; one HIP stream per context;
 *   - a context is not thread-safe (one host thread drives it, like the
 *     reference's single-process driver, CudaExecution.hpp:13-32);
 *   - vectors are node-major [u_x, u_y, u_z, p] per node (iga_cuda.cu:360-363);
 *   - block values are 4x4 row-major (row field * 4 + column field)
 *     (IgaCudaKernels.cuh:389-395).
 */
#ifndef TFI_H
#define TFI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tfi_ctx tfi_ctx;
typedef struct tfi_mesh tfi_mesh;
typedef struct tfi_flow tfi_flow;
typedef struct tfi_amg tfi_amg;

enum { TFI_OK = 0, TFI_ERROR = 1 };
enum { TFI_PC_BJACOBI = 0, TFI_PC_AMG = 1 };
enum { TFI_FP64 = 0, TFI_MIXED = 1 };
enum { TFI_WARM_PREVIOUS = 0, TFI_WARM_LEARNED = 1, TFI_WARM_EXTRAPOLATE = 2, TFI_WARM_EXTRAPOLATE2 = 3,
	TFI_WARM_EXTRAPOLATE3 = 4, TFI_WARM_EXTRAPOLATE4 = 5, TFI_WARM_MPE = 6 };
/* AMG smoothers; TFI_SMOOTH_AUTO: Chebyshev for transient flow (dt > 0),
 * multicolour Gauss-Seidel for steady flow (the default). */
enum { TFI_SMOOTH_JACOBI = 0, TFI_SMOOTH_GAUSS_SEIDEL = 1, TFI_SMOOTH_CHEBYSHEV = 2, TFI_SMOOTH_AUTO = 3 };
enum { TFI_TANGENT_LEGACY = 0, TFI_TANGENT_EXACT = 1 };

/* ------------------------------------------------------------------ context */
/* Replaces the CUDA runtime setup + PrintDevice (iga_cuda.cu:278-303). */
int tfi_ctx_create(int device, tfi_ctx** out);
void tfi_ctx_destroy(tfi_ctx* ctx);
const char* tfi_last_error(const tfi_ctx* ctx);
/* One-line device description ("name=... cus=... memory_total_gib=..."). */
int tfi_device_info(tfi_ctx* ctx, char* buffer, int length);
/* Project device allocations: DeviceAllocationCounter (DeviceAllocation.hpp:12-28). */
int tfi_allocation_stats(tfi_ctx* ctx, long long* peak_bytes, long long* live_bytes);

/* ------------------------------------------------------------------ mesh */
typedef struct {
	int nodes, elements, maximum_basis;
	long long basis_entries, extraction_nonzeros, block_nonzeros, element_pairs;
	int colors;               /* element colours of the deterministic scatter */
	double minimum_determinant; /* min over (element, q) of 0.125 det J */
	long long bad_samples;    /* quadrature points with det <= 0 or non-finite */
	long long basis_table_bytes; /* device bytes of the precomputed basis table (0: none) */
} tfi_mesh_stats;

/* Uploads a FlatMesh (DeviceMesh.hpp:16-68, 70-118), builds the node-block
 * pattern on the GPU (BlockPattern, BlockCsr.hpp:27-55), the element colouring
 * and pair->block map of the deterministic scatter, and the per-quadrature-point
 * geometry (BuildGeometryKernel, IgaCudaKernels.cuh:61-111). */
int tfi_mesh_create(tfi_ctx* ctx, int nodes, int elements, const int32_t* element_offsets,
	const int32_t* connectivity, const int32_t* extraction_offsets, const uint8_t* extraction_columns,
	const double* extraction_values, const double* bezier_points, tfi_mesh** out);
void tfi_mesh_destroy(tfi_mesh* mesh);
int tfi_mesh_stats_get(tfi_mesh* mesh, tfi_mesh_stats* out);
/* Downloads the block pattern: row_offsets[N+1], columns[nnzb], diagonal[N]. */
int tfi_mesh_pattern(tfi_mesh* mesh, int32_t* row_offsets, int32_t* columns, int32_t* diagonal);
/* Element colouring: colour of each element (int32[E]). */
int tfi_mesh_colors(tfi_mesh* mesh, int32_t* element_color);
/* Basis table: N, grad N and lap N of every (element, quadrature point, basis
 * function), 5 doubles each, evaluated once from the extraction operator and
 * geometry so that every later assembly reads them instead of re-evaluating
 * the Bernstein sum-factorisation (the reference evaluates them at every
 * point of every assembly: EvaluateBasisDevice, IgaCudaKernels.cuh:113, called
 * at :303).  enable = 1 builds it, 0 frees it,
 * -1 builds it when it takes at most half of the free device memory (what
 * tfi_mesh_create does).  Assembly results are the same either way. */
int tfi_mesh_basis_table(tfi_mesh* mesh, int enable);

/* Pressure-traction load of labelled faces (IntegratePressureTractionForces,
 * solvers/cpu/include/PressureTractionFlow.hpp:41-64, with the face integral of
 * BoundaryFlow.hpp:445-477): traction [4N] is overwritten with
 * sum_k -pressure[k] int_{face k} N_a n dA in the velocity rows, summed in the
 * order the faces are given (pass them element-major, faces 0..5, to match the
 * reference's order).  Computed on the GPU, gathered on the host. */
int tfi_pressure_traction(tfi_mesh* mesh, int count, const int32_t* element, const int32_t* face,
	const double* pressure, double* traction);

/* ------------------------------------------------------------------ kernels (host in/out) */
/* Negated residual and tangent with a deterministic colour-ordered scatter:
 * tangent TFI_TANGENT_LEGACY = the GPU reference's tau-frozen tangent
 * (AssembleNavierStokesKernel, IgaCudaKernels.cuh:265-409), TFI_TANGENT_EXACT =
 * the CPU reference's consistent tangent (NavierStokesElement.hpp:244-313).
 * jacobian may be NULL (residual only).  Arrays: state/previous [4N],
 * jacobian [nnzb*16], negative_residual [4N]. */
int tfi_ns_assemble(tfi_mesh* mesh, const double* state, const double* previous, double density,
	double viscosity, double dt, double* jacobian, double* negative_residual, int tangent);
/* y = A x on the mesh pattern (BlockSpmvKernel<4>, SparseKernels.cuh:25-40);
 * precision TFI_FP64 or TFI_MIXED (values rounded to fp32). */
int tfi_bsr_spmv(tfi_mesh* mesh, const double* values, const double* x, double* y, int precision);
/* Device-pointer forms of the kernels above, for a Newton loop that keeps its
 * arrays in device memory (the reference driver's own loop, iga_cuda.cu:956-1046,
 * keeps state, Jacobian and rhs on the GPU): the same arithmetic with no host
 * copies (tfi_ns_assemble round-trips the nnzb*16 Jacobian).  All arrays are
 * device pointers of the sizes above; synchronous on the context stream. */
int tfi_ns_assemble_device(tfi_mesh* mesh, const double* state, const double* previous, double density,
	double viscosity, double dt, double* jacobian, double* negative_residual, int tangent);
int tfi_bsr_spmv_device(tfi_mesh* mesh, const double* values, const double* x, double* y);
int tfi_apply_dirichlet_device(tfi_mesh* mesh, const int32_t* velocity_mask, const int32_t* pressure_mask,
	const double* velocity_target, const double* pressure_target, const double* state, double* jacobian,
	double* rhs);
/* 4x4 diagonal-block inverses (BuildBlockInverseKernel<4>, SparseKernels.cuh:42-90). */
int tfi_block_inverse(tfi_mesh* mesh, const double* values, double* inverse, int* singular);
/* Allocation-free variant of tfi_block_inverse.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * values [nnzb*16] and inverse [16N] are DEVICE pointers; *singular (host)
 * receives the number of blocks that took the diagonal fallback.  The kernels
 * work in a fixed library-wide device array (one count per 256-row
 * workgroup, 256 KiB per GPU, held until the process ends), so the mesh may
 * have at most 16,777,216 nodes; larger meshes are rejected with an error.
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * race.  Synchronous on the context stream; allocates nothing. */
int tfi_block_inverse_noalloc(tfi_mesh* mesh, const double* values, double* inverse, int* singular);
/* Dirichlet rows and rhs (ApplyNavierStokesBoundaryKernel /
 * SetNavierStokesBoundaryRhsKernel, IgaCudaKernels.cuh:436-463), in place. */
int tfi_apply_dirichlet(tfi_mesh* mesh, const int32_t* velocity_mask, const int32_t* pressure_mask,
	const double* velocity_target, const double* pressure_target, const double* state, double* jacobian,
	double* rhs);
/* Solve A x = b (x holds the initial guess) with FGMRES; options as in
 * tfi_flow_options.  Replaces SolveGmres<4> (Gmres.hpp:69-213). */
int tfi_linear_solve(tfi_mesh* mesh, const double* values, const double* b, double* x, int preconditioner,
	int precision, int restart, int maximum_iterations, double relative_tolerance, int* iterations,
	double* residual);
/* Device-resident SpMV timing on the context stream with HIP events:
 * average milliseconds per launch over `repetitions` launches on a synthetic
 * matrix of the mesh pattern. */
int tfi_bench_spmv(tfi_mesh* mesh, int precision, int repetitions, double* milliseconds);
/* Device-resident assembly timing on a synthetic state (HIP events, average
 * milliseconds per assembly, all colour launches included); with_jacobian = 0
 * times the residual-only assembly.  No reference counterpart. */
int tfi_bench_assembly(tfi_mesh* mesh, int tangent, int with_jacobian, int repetitions, double* milliseconds);
/* Calibration copy for HBM counters: y = x over `count` doubles (a plain
 * grid-stride copy kernel), average milliseconds per launch.  No reference
 * counterpart. */
int tfi_bench_copy(tfi_ctx* ctx, long long count, int repetitions, double* milliseconds);
/* Measured HBM stream rates on `count` doubles (even): a 16-byte-vector copy
 * (GB/s counting read + written bytes) and a read-only sum (GB/s read), each
 * averaged over `repetitions` launches of 32 workgroups per CU, one contiguous
 * chunk per workgroup (HIP events).
 * The measured bandwidth beside the published peak.  No reference
 * counterpart. */
int tfi_bench_stream(tfi_ctx* ctx, long long count, int repetitions, double* copy_gbs, double* read_gbs);
/* Measured FP64 matrix-core throughput: a probe kernel of independent
 * v_mfma_f64_16x16x4f64 chains on every SIMD, TFLOP/s over `repetitions`
 * launches (HIP events).  The measured peak beside the published one.  No
 * reference counterpart. */
int tfi_bench_mfma_fp64(tfi_ctx* ctx, int repetitions, double* tflops);

/* ------------------------------------------------------------------ flow solver */
typedef struct {
	double density, viscosity, dt;    /* dt == 0: steady */
	int maximum_newton;               /* 12 (iga_cuda.cu:402) */
	double nonlinear_rtol;            /* 1e-5 (iga_cuda.cu:403) */
	double nonlinear_atol;            /* 1e-10 (iga_cuda.cu:404) */
	double mass_rtol;                 /* 1e-3 (iga_cuda.cu:405) */
	int linear_maximum_iterations;    /* 5000 (iga_cuda.cu:1015) */
	double linear_rtol;               /* 1e-8 (iga_cuda.cu:1015) */
	int restart;                      /* Krylov restart length */
	int preconditioner;               /* TFI_PC_BJACOBI | TFI_PC_AMG */
	int precision;                    /* TFI_FP64 | TFI_MIXED */
	int warm_start;                   /* TFI_WARM_* */
	int verbose;                      /* > 0: one stderr line per Newton iteration as it ends (and one when a
	                                     Chebyshev-smoothed solve falls back to Gauss-Seidel) */
	int amg_smoother;                 /* TFI_SMOOTH_* on the large levels */
	int tangent;                      /* TFI_TANGENT_* */
	int amg_coarse_nodes;             /* block rows of the coarsest dense level (0: 256, at most 256) */
	int amg_gs_min_rows;              /* levels below this many rows use block Jacobi (0: 8192) */
	int amg_gs_sweeps;                /* 2 (0): symmetric Gauss-Seidel pre and post; 1: forward pre, backward post */
	int chord_polish;                 /* 1 (default): a residual check that misses the tolerance by at most
	                                     1000x is closed by chord steps (solves with the tangent and
	                                     preconditioner already set up, each solved as far as the observed
	                                     contraction makes useful; up to 6, each required to cut the
	                                     residual 4x) before a new tangent is built.  A learned guess
	                                     predicted within that window starts this way, with the tangent
	                                     of an earlier step. */
	int amg_chebyshev_degree;         /* Chebyshev steps before and after the coarse correction (0: 2) */
	int amg_half_smoother;            /* TFI_MIXED: 1 (default) = Chebyshev-smoothed levels read D^-1 A
	                                     (block-Jacobi scaled) in FP16, half the bytes of FP32 blocks; a level
	                                     whose D^-1 A has an entry above 60000 keeps FP32; 0 = FP32 A */
} tfi_flow_options;

void tfi_flow_default_options(tfi_flow_options* options);

#define TFI_MAX_NEWTON_RECORDS 64
typedef struct {
	double residual_l2, continuity_l2, continuity_sum, net_boundary_flow, relative_mass_imbalance;
	double update_l2, linear_residual, assembly_s, linear_s;
	int linear_iterations, singular_diagonal_blocks;
	double step;             /* fraction of the Newton update applied (1, or less under the velocity step limit) */
} tfi_newton_record;

typedef struct {
	int converged, newton_iterations, linear_iterations, records;
	double initial_residual, final_residual, tolerance, assembly_s, linear_s, total_s;
	double warm_start_initial_ratio;  /* ||r(x0)|| / ||b|| of the first linear solve */
	int warm_start_used;              /* 1 if the learned guess was accepted */
	int chord_steps;                  /* chord steps taken (see tfi_flow_options.chord_polish) */
	double warm_start_blend;          /* learned mode: line-search scale t of the network's correction */
	double warm_start_base_ratio;     /* learned mode: residual of its base prediction alone, relative to
	                                     the previous-step start (compare warm_start_initial_ratio) */
	tfi_newton_record record[TFI_MAX_NEWTON_RECORDS];
} tfi_newton_report;

int tfi_flow_create(tfi_mesh* mesh, const tfi_flow_options* options, tfi_flow** out);
void tfi_flow_destroy(tfi_flow* flow);
/* state/previous [4N] (previous may be NULL: copied from state). */
int tfi_flow_set_state(tfi_flow* flow, const double* state, const double* previous);
int tfi_flow_get_state(tfi_flow* flow, double* state);
/* Resolved Dirichlet data (ResolvedBoundaryConditions, CaseInput.hpp:35-47)
 * and the assembled traction load [4N] (IntegratePressureTractionForces,
 * PressureTractionFlow.hpp:41-64), may be NULL. */
int tfi_flow_set_boundary(tfi_flow* flow, const int32_t* velocity_mask, const int32_t* pressure_mask,
	const double* velocity_target, const double* pressure_target, const double* traction);
/* Labelled boundary faces for the mass-balance test (MeasureFlowConvergence,
 * iga_cuda.cu:520-557): count faces given as (element, local face, label). */
int tfi_flow_set_boundary_faces(tfi_flow* flow, int count, const int32_t* element, const int32_t* face,
	const int32_t* label);
/* Start of a physical step: previous <- state (iga_cuda.cu:918-920). phase =
 * inflow phase features for the learned predictor (may be NULL). */
int tfi_flow_begin_step(tfi_flow* flow, const double* phase, int phase_count);
/* Newton loop at the current boundary data (iga_cuda.cu:956-1046). The first
 * call after tfi_flow_begin_step starts from the warm-start prediction (when
 * the mode predicts); later calls in the same step (outlet fixed-point passes,
 * iga_cuda.cu:933-1084) start from the state the previous call left. */
int tfi_flow_newton(tfi_flow* flow, tfi_newton_report* report);
/* End of a physical step: history for the learned warm start. */
int tfi_flow_commit_step(tfi_flow* flow);
/* Per-face volumetric flows of the current state for labels (outlets):
 * flows[i] = sum over faces with label labels[i] (IntegrateBoundaryFlow). */
int tfi_flow_boundary_flows(tfi_flow* flow, int count, const int32_t* labels, double* flows);

/* AMG hierarchy of the flow solver after its first setup: per level the block
 * rows, block nonzeros and Gauss-Seidel colours (0 = Jacobi-smoothed or the
 * dense coarsest level); *levels receives the level count.  No reference
 * counterpart (the reference GPU solver has no multigrid). */
int tfi_flow_amg_levels(tfi_flow* flow, int capacity, int* levels, int* rows, long long* block_nonzeros,
	int* colors);

/* ------------------------------------------------------------------ AMG */
/* The coupled 4x4-block aggregation AMG as a preconditioner of its own, for a
 * caller with its own Krylov method (the reference's GMRES uses block-Jacobi
 * scaling only, Gmres.hpp:69-213; no reference counterpart).  smoother:
 * TFI_SMOOTH_* (AUTO = Gauss-Seidel); precision TFI_MIXED builds the levels
 * in FP32 (Chebyshev levels read D^-1 A in FP16).  tfi_amg_setup builds the
 * aggregation (first call) and the numeric hierarchy; tfi_amg_update rebuilds
 * the numeric hierarchy for new values on the same pattern.  values [nnzb*16]
 * is a DEVICE pointer the hierarchy keeps reading: it must stay valid and
 * unchanged until the next update or tfi_amg_destroy.  tfi_amg_apply: z =
 * M^-1 v, one V-cycle, v and z DEVICE pointers [4N].  Destroy the handle
 * before its mesh. */
int tfi_amg_create(tfi_mesh* mesh, int precision, int smoother, tfi_amg** out);
int tfi_amg_setup(tfi_amg* amg, const double* values);
int tfi_amg_update(tfi_amg* amg, const double* values);
int tfi_amg_apply(tfi_amg* amg, const double* v, double* z);
/* Per level: block rows, block nonzeros, Gauss-Seidel colours (0: other smoother). */
int tfi_amg_levels(tfi_amg* amg, int capacity, int* levels, int* rows, long long* block_nonzeros, int* colors);
void tfi_amg_destroy(tfi_amg* amg);

/* Learned warm start (MLP on 27-control-point stencils).  The network
 * predicts the error of a base prediction: TFI_WARM_MPE (minimal polynomial
 * extrapolation of the step increments, the default) or TFI_WARM_EXTRAPOLATE4
 * (quartic extrapolation in time). */
typedef struct {
	int hidden;            /* width (128) */
	int stencil;           /* control points per stencil (27) */
	int train_steps;       /* optimisation steps per snapshot */
	double learning_rate;
	int min_snapshots;     /* snapshots before the predictor is used */
	int base;              /* TFI_WARM_MPE (0: default) or TFI_WARM_EXTRAPOLATE4 */
} tfi_warm_options;
int tfi_flow_warm_configure(tfi_flow* flow, const tfi_warm_options* options);
/* Training statistics: snapshots seen, last training loss, accepted guesses. */
int tfi_flow_warm_stats(tfi_flow* flow, int* snapshots, double* last_loss, int* accepted, int* rejected);
/* Scale the learned correction is applied with: the least-squares fit of the
 * network's prediction to the latest snapshot before training on it, in
 * [0, 1] (0: the prediction is the extrapolation base alone). */
int tfi_flow_warm_gate(tfi_flow* flow, double* gate);
/* The state a predicting mode would start the next step from, given the
 * states recorded by tfi_flow_commit_step so far (host array [4N]).  For
 * TFI_WARM_MPE, *order receives the number of increments the extrapolation
 * combined (may be NULL).  No reference counterpart. */
int tfi_flow_warm_predict(tfi_flow* flow, int mode, double* state, int* order);
/* Append a state (host array [4N]) to the warm-start history, as
 * tfi_flow_commit_step appends the converged state (without training), e.g.
 * to resume a restarted run with the states before its checkpoint.
 * tfi_flow_set_state clears the history.  No reference counterpart. */
int tfi_flow_warm_record(tfi_flow* flow, const double* state);
/* One training step of the learned predictor outside a Newton solve: the
 * state (host array [4N]) is appended to the history as tfi_flow_commit_step
 * appends a converged state, and the network trains on the snapshot it forms
 * (the base prediction's error at this state from the history before it),
 * train_steps Adam steps of tfi_warm_options, once five states precede it.
 * phase: up to 8 features (may be NULL with phase_count 0).  *loss (may be
 * NULL) receives the last training loss (0 before the first snapshot).  No
 * reference counterpart. */
int tfi_flow_warm_train(tfi_flow* flow, const double* state, const double* phase, int phase_count, double* loss);

/* ------------------------------------------------------------------ transport */
/* Two-species neuron transport on the mesh pattern with 2x2 node blocks
 * [N0, Nplus] (`iga_cuda transport`, solvers/cuda/src/iga_cuda.cu:621-733).
 * Parameters as TransportParameters (solvers/cpu/include/CaseInput.hpp:19-33)
 * without the step count and boundary values. */
typedef struct tfi_transport tfi_transport;
typedef struct {
	double diffusion, vplus, vminus, kplus, kminus, detach_plus, detach_minus, dt;
	double artificial_diffusion;
} tfi_transport_params;
/* Assembles the backward-Euler matrices L and P for the prescribed nodal
 * velocity [3N] (AssembleTransportKernel, IgaCudaKernels.cuh:155-237) and
 * applies the constrained rows (ApplyTransportBoundaryKernel, :411-424):
 * constrained [N] = 1 for a node whose N0 and Nplus are prescribed.  The
 * assembly reads the mesh's basis table and builds it when the mesh has none
 * (tfi_mesh_basis_table). */
int tfi_transport_create(tfi_mesh* mesh, const tfi_transport_params* params, const double* velocity,
	const int32_t* constrained, tfi_transport** out);
void tfi_transport_destroy(tfi_transport* transport);
/* Preconditioner of the step's FGMRES on D^-1 L: TFI_PC_BJACOBI (default;
 * none beyond the block-Jacobi scaling itself, as the reference's GMRES<2>)
 * or TFI_PC_AMG (the coupled aggregation AMG with FP32 V-cycles, its 2x2
 * blocks embedded in the 4x4 hierarchy with an identity on the two padding
 * fields: about 5x fewer iterations, each about 7x dearer).  The solution is
 * the same either way. */
int tfi_transport_set_preconditioner(tfi_transport* transport, int preconditioner);
/* Prescribed N0 / Nplus values [N] of the constrained nodes
 * (SetTransportBoundaryVectorKernel, IgaCudaKernels.cuh:426-434). */
int tfi_transport_set_boundary(tfi_transport* transport, const double* n0, const double* nplus);
/* State [2N] (N0, Nplus per node). */
int tfi_transport_set_state(tfi_transport* transport, const double* state);
int tfi_transport_get_state(tfi_transport* transport, double* state);
/* state <- prescribed values at the constrained nodes (iga_cuda.cu:674-677). */
int tfi_transport_apply_boundary(tfi_transport* transport);
/* One backward-Euler step (iga_cuda.cu:685-705): b = P c, constrained entries
 * of b set to their values, L c' = b solved as D^-1 L c' = D^-1 b (D the 2x2
 * diagonal blocks, the reference's left block-Jacobi preconditioning) by
 * FGMRES(50) from c' = 0 (zero_guess) or c, to ||D^-1 (b - L c')|| <= rtol
 * ||D^-1 b||; c <- c'.  A solve that does not converge is an error.  Destroy
 * the transport before its context. */
int tfi_transport_step(tfi_transport* transport, int zero_guess, int maximum_iterations, double relative_tolerance,
	int* iterations, double* relative_residual);
/* Host copies of the assembled 2x2 blocks of L and P [nnzb*4] and the count of
 * diagonal blocks inverted by their diagonal (either pointer may be NULL). */
int tfi_transport_matrices(tfi_transport* transport, double* left, double* previous, int* singular_blocks);
/* Device-resident timing of the transport's 2x2 SpMV (y = L x on the current
 * state) with HIP events around `repetitions` back-to-back launches: average
 * milliseconds per launch.  No reference counterpart. */
int tfi_transport_bench_spmv(tfi_transport* transport, int repetitions, double* milliseconds);

/* ------------------------------------------------------------------ configured transport */
/* A linear_transport equation system of 1..8 coupled scalar fields with
 * F x F node blocks (`iga_cuda solve`, solvers/cuda/include/ConfiguredTransportHost.inc:142-471;
 * kernels GenericTransportKernels.cuh).  Terms as CompiledTerm
 * (include/SimulationConfig.hpp:159-165): kind TFI_TERM_*, equation and trial
 * field indices, coefficient. */
typedef struct tfi_solve tfi_solve;
enum {
	TFI_TERM_TIME_DERIVATIVE = 0, TFI_TERM_DIFFUSION = 1, TFI_TERM_ADVECTION = 2, TFI_TERM_LINEAR_COUPLING = 3,
	TFI_TERM_VOLUME_SOURCE = 4
};
/* SUPG setting per equation (ConfiguredTransportHost.inc:186-193): none, or
 * SUPG with the diffusion weighted "galerkin", "supg_test" or "consistent". */
enum { TFI_SUPG_NONE = 0, TFI_SUPG_GALERKIN = 1, TFI_SUPG_TEST = 2, TFI_SUPG_CONSISTENT = 3 };
typedef struct {
	int32_t kind, equation, trial;
	double coefficient;
} tfi_solve_term;
typedef struct {
	int32_t fields;              /* 1..8 */
	int32_t term_count;          /* 0..64 */
	const tfi_solve_term* terms;
	const int32_t* supg;         /* [fields] TFI_SUPG_* */
	const double* tau_diffusivity;  /* [fields] (SupgTauDiffusivity, SimulationConfig.hpp:728-741) */
	double dt;
	/* flux / Robin faces (ConfiguredTransportHost.inc:211-255): element, face
	 * (0..5) and per-field flux, Robin coefficient and coefficient x exterior
	 * value [surface_count * fields] */
	int32_t surface_count;
	const int32_t* surface_element;
	const int32_t* surface_face;
	const double* surface_flux;
	const double* surface_robin;
	const double* surface_robin_source;
} tfi_solve_system;
/* Assembles the backward-Euler left matrix, previous-state matrix and source
 * vector for the prescribed nodal velocity [3N] (AssembleGenericTransportKernel,
 * AssembleGenericSurfaceKernel) and applies the constrained rows
 * (ApplyGenericBoundaryKernel): constrained [N*fields] = 1 for a prescribed
 * (node, field).  Elements of at most 80 basis functions. */
int tfi_solve_create(tfi_mesh* mesh, const tfi_solve_system* system, const double* velocity,
	const int32_t* constrained, tfi_solve** out);
void tfi_solve_destroy(tfi_solve* solve);
/* Re-assembles L, P and the source for a new nodal velocity [3N] and rebuilds
 * the block-Jacobi scaling, as the reference does before every step of a
 * snapshot-series velocity source (ConfiguredTransportHost.inc:337-358, 376). */
int tfi_solve_set_velocity(tfi_solve* solve, const double* velocity);
/* Preconditioner of the step's FGMRES on D^-1 L: TFI_PC_BJACOBI (default;
 * none beyond the block-Jacobi scaling itself, as the reference's GMRES) or
 * TFI_PC_AMG (fields <= 4: the coupled aggregation AMG with FP32 V-cycles, the
 * F x F blocks embedded in its 4x4 hierarchy with an identity on the padding
 * fields; set up again after each re-assembly).  No reference counterpart;
 * the solution is the same either way. */
int tfi_solve_set_preconditioner(tfi_solve* solve, int preconditioner);
/* Prescribed values [N*fields] of the constrained entries
 * (SetGenericBoundaryVectorKernel). */
int tfi_solve_set_boundary(tfi_solve* solve, const double* values);
int tfi_solve_set_state(tfi_solve* solve, const double* state);
int tfi_solve_get_state(tfi_solve* solve, double* state);
/* One backward-Euler step (ConfiguredTransportHost.inc:378-400): b = P c + s,
 * constrained entries of b set to their values, L c' = b solved as
 * D^-1 L c' = D^-1 b (D the F x F diagonal blocks, the reference's left
 * block-Jacobi GMRES) by FGMRES(50) from c' = 0 (zero_guess) or c, to
 * ||D^-1 (b - L c')|| <= rtol ||D^-1 (b - L c'_0)||; c <- c'.  A solve that
 * does not converge is an error. */
int tfi_solve_step(tfi_solve* solve, int zero_guess, int maximum_iterations, double relative_tolerance,
	int* iterations, double* relative_residual);
/* Host copies of the assembled blocks of L and P [nnzb*fields^2] (row-major
 * blocks), the source vector [N*fields] and the count of diagonal blocks
 * inverted by their diagonal (any pointer may be NULL). */
int tfi_solve_matrices(tfi_solve* solve, double* left, double* previous, double* source, int* singular_blocks);
/* Device-resident timing of the F x F SpMV (y = L x on the current state)
 * with HIP events around `repetitions` back-to-back launches: average
 * milliseconds per launch.  No reference counterpart. */
int tfi_solve_bench_spmv(tfi_solve* solve, int repetitions, double* milliseconds);

/* Kernel timers accumulated on the context stream (HIP events). */
typedef struct {
	double spmv_ms, spmv_calls, spmv_bytes;           /* fine-level fp64 SpMV */
	double spmv32_ms, spmv32_calls, spmv32_bytes;     /* fine-level fp32 SpMV */
	double assembly_ms, assembly_calls, assembly_flops;  /* residual + tangent: AssembleNsKernel launches */
	double assembly_launches;                         /* (one launch per element colour) */
	double residual_ms, residual_calls, residual_flops;  /* residual-only assemblies */
	double spmv2_ms, spmv2_calls, spmv2_bytes;        /* transport 2x2 fp64 SpMV */
	double transport_assembly_ms;                     /* transport L and P assembly (all colours) */
} tfi_kernel_timers;
int tfi_timers_get(tfi_ctx* ctx, tfi_kernel_timers* out);
int tfi_timers_enable(tfi_ctx* ctx, int enabled);
int tfi_timers_reset(tfi_ctx* ctx);

/* ------------------------------------------------------------------ test hooks
 * No reference counterpart: kernel-level entry points of the learned warm
 * start, used by the parity tests against the CPU oracle.
 * tfi_mlp_reference: FP32 forward + backward (MSE loss) of the stencil MLP on a
 * host batch x [batch][640], target [batch][4]; weights w1 [640][H], b1 [H],
 * w2 [H][H], b2 [H], w3 [H][4], b3 [4] (row-major, x W + b, tanh hidden).
 * tfi_mlp_infer_bf16: the BF16-MFMA inference kernel on the same layout. */
int tfi_mlp_reference(tfi_ctx* ctx, int hidden, int batch, const float* x, const float* target, const float* w1,
	const float* b1, const float* w2, const float* b2, const float* w3, const float* b3, float* out, float* loss,
	float* g_w1, float* g_b1, float* g_w2, float* g_b2, float* g_w3, float* g_b3);
int tfi_mlp_infer_bf16(tfi_ctx* ctx, int hidden, int rows, const float* x, const float* w1, const float* b1,
	const float* w2, const float* b2, const float* w3, const float* b3, float* out);

#ifdef __cplusplus
}
#endif

#endif /* TFI_H */
