// Flow solver (Newton loop over device-resident state) and the host-in/out
// kernel entry points of the C-ABI.
//
// tfi_flow_newton restates the nonlinear loop of the reference driver
// (solvers/cuda/src/iga_cuda.cu:956-1046): assemble residual + tangent,
// add the traction load, measure continuity and boundary mass imbalance,
// apply Dirichlet rows, test ||r|| <= max(atol sqrt(4N), rtol ||r0||) together
// with the imbalance tolerance, solve, update.  Unlike the reference it copies
// no vector to the host per Newton iteration: the boundary flows are
// integrated on the GPU and only per-face scalars come back.
#include "tfi_internal.hpp"
#include "tfi_solver.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <limits>
#include <string>
#include <map>

namespace tfi {

namespace {

using Clock = std::chrono::steady_clock;

__global__ void ContinityPartialKernel(const double* __restrict__ rhs, int nodes, double* __restrict__ partial)
{
	__shared__ double l2[4], sm[4];
	double a = 0.0, s = 0.0;
	for (int n = blockIdx.x*blockDim.x + threadIdx.x; n < nodes; n += gridDim.x*blockDim.x) {
		const double c = rhs[4*(size_t)n + 3];
		a += c*c;
		s += c;
	}
	for (int o = 32; o > 0; o >>= 1) {
		a += __shfl_xor(a, o, 64);
		s += __shfl_xor(s, o, 64);
	}
	if ((threadIdx.x & 63) == 0) {
		l2[threadIdx.x >> 6] = a;
		sm[threadIdx.x >> 6] = s;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		partial[2*blockIdx.x] = l2[0] + l2[1] + l2[2] + l2[3];
		partial[2*blockIdx.x + 1] = sm[0] + sm[1] + sm[2] + sm[3];
	}
}

// Largest |x| over the entries i with i % stride < width (stride 4, width 3:
// the velocity components of a state vector), per block of 256 threads.
__global__ void MaxAbsPartialKernel(const double* __restrict__ x, long long n, int stride, int width,
	double* __restrict__ partial)
{
	__shared__ double w[4];
	double v = 0.0;
	for (long long i = blockIdx.x*(long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x*blockDim.x)
		if (i % stride < width) v = fmax(v, fabs(x[i]));
	for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
	if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = v;
	__syncthreads();
	if (threadIdx.x == 0) partial[blockIdx.x] = fmax(fmax(w[0], w[1]), fmax(w[2], w[3]));
}

}  // namespace

LinearSolver::LinearSolver(tfi_mesh* mesh, int preconditioner, int precision, int restart, const AmgOptions& amg)
	: mesh_(mesh), ctx_(mesh->ctx), pc_(preconditioner), precision_(precision),
	  krylov_(mesh->ctx, 4LL*mesh->nodes, restart)
{
	const long long n = 4LL*mesh->nodes;
	inverse_.Allocate(ctx_, 16LL*mesh->nodes);
	work_.Allocate(ctx_, n);
	if (precision_ == TFI_MIXED) {
		values32_.Allocate(ctx_, mesh->nnzb*16);
		correction_.Allocate(ctx_, n);
		residual_.Allocate(ctx_, n);
	}
	if (pc_ == TFI_PC_AMG) amg_ = std::make_unique<Amg>(mesh, precision_, amg);
}

LinearSolver::~LinearSolver() = default;

void LinearSolver::Setup(const double* values)
{
	values_ = values;
	const BsrView a = mesh_->view();
	if (precision_ == TFI_MIXED) ToFloat(ctx_, values, values32_.data(), mesh_->nnzb*16);
	if (pc_ == TFI_PC_BJACOBI) {
		singular_ = BlockInverse64(ctx_, a, values, inverse_.data());
	} else {
		amg_->Setup(values, precision_ == TFI_MIXED ? values32_.data() : nullptr);
		singular_ = amg_->singular_blocks();
	}
}

KrylovResult LinearSolver::Solve(const double* b, double* x, int maxit, double rtol)
{
	constexpr int kChebyshevCap = 400;
	if (!amg_ || amg_->smoother() != TFI_SMOOTH_CHEBYSHEV) return SolveOnce(b, x, maxit, rtol);
	const long long n = 4LL*mesh_->nodes;
	if (!guess_.size()) guess_.Allocate(ctx_, n);
	Copy(ctx_, x, guess_.data(), n);
	KrylovResult first = SolveOnce(b, x, std::min(maxit, kChebyshevCap), rtol);
	if (first.converged || maxit <= kChebyshevCap) return first;
	if (verbose_ > 0)
		std::fprintf(stderr, "tfi_flow linear solve: Chebyshev-smoothed AMG stalled at %d iterations "
			"(residual %.3e); Gauss-Seidel smoothing from here on\n", first.iterations, first.residual);
	amg_->UseGaussSeidel();
	Setup(values_);
	Copy(ctx_, guess_.data(), x, n);
	KrylovResult r = SolveOnce(b, x, maxit - first.iterations, rtol);
	r.iterations += first.iterations;
	return r;
}

KrylovResult LinearSolver::SolveOnce(const double* b, double* x, int maxit, double rtol)
{
	const BsrView a = mesh_->view();
	const long long n = 4LL*mesh_->nodes;
	tfi_ctx* ctx = ctx_;
	Operator A64 = [&](const double* in, double* out) {
		StreamTimer t(ctx, &ctx->timers.spmv_ms);
		Spmv64(ctx, a, values_, in, out);
		ctx->timers.spmv_calls += 1;
		ctx->timers.spmv_bytes += SpmvBytes(a, 8);
	};
	Precond M = [&](const double* v, double* z) {
		if (pc_ == TFI_PC_BJACOBI) BlockApply64(ctx, mesh_->nodes, inverse_.data(), v, z);
		else amg_->Apply(v, z);
	};
	if (precision_ == TFI_FP64) return krylov_.Solve(A64, M, b, x, maxit, rtol);
	// Mixed precision: iterative refinement.  Outer residual, Gram-Schmidt and
	// updates in FP64; the inner operator reads FP32 matrix values and the
	// preconditioner runs in FP32.
	Operator A32 = [&](const double* in, double* out) {
		StreamTimer t(ctx, &ctx->timers.spmv32_ms);
		Spmv32(ctx, a, values32_.data(), in, out);
		ctx->timers.spmv32_calls += 1;
		ctx->timers.spmv32_bytes += SpmvBytes(a, 4);
	};
	KrylovResult total;
	total.rhs_norm = Norm2(ctx, b, n);
	const double target = rtol*total.rhs_norm;
	double rn0 = -1.0;
	for (int refinement = 0; refinement < 30; ++refinement) {
		Residual64(ctx, a, values_, x, b, residual_.data());
		const double rn = Norm2(ctx, residual_.data(), n);
		total.residual = rn;
		if (rn <= target || total.rhs_norm == 0.0) {
			total.converged = true;
			break;
		}
		if (rn0 < 0.0) rn0 = rn;
		if (total.iterations >= maxit || !std::isfinite(rn) || rn > 10.0*rn0) break;
		Fill(ctx, correction_.data(), 0.0, n);
		// ask the inner solve for what FP32 storage can deliver, never more
		// than the remaining reduction
		const double inner = std::max(target/rn, 1e-5);
		const auto r = krylov_.Solve(A32, M, residual_.data(), correction_.data(), maxit - total.iterations, inner);
		total.iterations += r.iterations;
		Axpy(ctx, 1.0, correction_.data(), x, n);
		++refinements_;
	}
	return total;
}

}  // namespace tfi

using namespace tfi;

struct tfi_flow {
	tfi_mesh* mesh = nullptr;
	tfi_ctx* ctx = nullptr;
	tfi_flow_options opt{};
	DevBuf<double> state, previous, update, rhs, traction, jac, vtarget, ptarget, face_flow, cont_partial;
	DevBuf<double> guess_base, resid_base;  // learned mode: extrapolation base and its residual
	DevBuf<double> face_vec;                 // per labelled face: int N_a n dA, [face][kMaxBasis][3] (negated)
	DevBuf<double> last_update;              // the last Newton update, for backtracking after a failed solve
	DevBuf<int> vmask, pmask, face_elem, face_id;
	std::vector<int> face_label;
	bool has_traction = false, has_boundary = false;
	std::unique_ptr<LinearSolver> solver;
	std::unique_ptr<WarmStart> warm;
	double phase[8] = {};
	int phase_count = 0;
	double newton_contraction = 0.0;  // largest ||r_k|| / ||r_{k-1}||^2 of the last step
	bool tangent_ready = false;       // the linear solver holds a set-up tangent (possibly of an earlier state)
	// no Newton solve has run since tfi_flow_begin_step: only the first solve of
	// a step starts from a prediction; later ones (outlet fixed-point passes)
	// start where the previous one ended
	bool fresh_step = true;
};

namespace {

struct Metrics {
	double continuity_l2 = 0, continuity_sum = 0, net = 0, absolute = 0, imbalance = 0;
};

Metrics Measure(tfi_flow* f)
{
	tfi_ctx* ctx = f->ctx;
	Metrics m;
	const int blocks = 256;
	hipLaunchKernelGGL(ContinityPartialKernel, dim3(blocks), dim3(256), 0, ctx->stream, f->rhs.data(), f->mesh->nodes,
		f->cont_partial.data());
	TFI_LAUNCH_CHECK("ContinityPartialKernel");
	std::vector<double> part(2*blocks);
	f->cont_partial.Download(part.data(), part.size());
	for (int b = 0; b < blocks; ++b) {
		m.continuity_l2 += part[2*b];
		m.continuity_sum += part[2*b + 1];
	}
	m.continuity_l2 = std::sqrt(m.continuity_l2);
	if (!f->face_label.empty()) {
		const int count = static_cast<int>(f->face_label.size());
		FaceFlowsFromVectors(f->mesh, count, f->face_elem.data(), f->face_vec.data(), f->state.data(),
			f->face_flow.data());
		std::vector<double> flows(count);
		f->face_flow.Download(flows.data(), count);
		std::map<int, double> per_label;
		for (int i = 0; i < count; ++i) per_label[f->face_label[i]] += flows[i];
		for (const auto& kv : per_label) {
			m.net += kv.second;
			m.absolute += std::abs(kv.second);
		}
		if (m.absolute > 0.0) m.imbalance = 2.0*std::abs(m.net)/m.absolute;
	}
	return m;
}

template <class F>
int Guard(tfi_ctx* ctx, F&& fn)
{
	try {
		fn();
		return TFI_OK;
	} catch (const std::exception& e) {
		if (ctx) ctx->error = e.what();
		return TFI_ERROR;
	}
}

// max |x_i| over the components i % stride < width (device vector, host result)
double MaxAbs(tfi_flow* f, const double* x, long long n, int stride, int width)
{
	const int blocks = 256;
	hipLaunchKernelGGL(MaxAbsPartialKernel, dim3(blocks), dim3(256), 0, f->ctx->stream, x, n, stride, width,
		f->cont_partial.data());
	TFI_LAUNCH_CHECK("MaxAbsPartialKernel");
	std::vector<double> part(blocks);
	f->cont_partial.Download(part.data(), blocks);
	double v = 0.0;
	for (double p : part) v = std::max(v, p);
	return v;
}

}  // namespace

namespace {
// %.16e, the precision of the reference driver's failure messages
std::string Precise(double v)
{
	char buf[64];
	std::snprintf(buf, sizeof buf, "%.16e", v);
	return buf;
}
}  // namespace

// A standalone AMG hierarchy for callers with their own Krylov solver.
struct tfi_amg {
	tfi_mesh* mesh = nullptr;
	int precision = TFI_FP64;
	std::unique_ptr<Amg> amg;
	DevBuf<float> values32;
	bool ready = false;
};

namespace {
void AmgNumeric(tfi_amg* a, const double* values)
{
	tfi_mesh* m = a->mesh;
	if (!values) throw Error("the AMG set-up needs the matrix values");
	const float* v32 = nullptr;
	if (a->precision == TFI_MIXED) {
		if (!a->values32.size()) a->values32.Allocate(m->ctx, m->nnzb*16);
		ToFloat(m->ctx, values, a->values32.data(), m->nnzb*16);
		v32 = a->values32.data();
	}
	a->amg->Setup(values, v32);
	Sync(m->ctx);
	a->ready = true;
}
}  // namespace

extern "C" {

void tfi_flow_default_options(tfi_flow_options* o)
{
	if (!o) return;
	*o = tfi_flow_options{};
	o->density = 1.0;
	o->viscosity = 0.1;
	o->dt = 0.0;
	o->maximum_newton = 12;
	o->nonlinear_rtol = 1e-5;
	o->nonlinear_atol = 1e-10;
	o->mass_rtol = 1e-3;
	o->linear_maximum_iterations = 5000;
	o->linear_rtol = 1e-8;
	o->restart = 200;
	o->preconditioner = TFI_PC_AMG;
	o->precision = TFI_FP64;
	o->warm_start = TFI_WARM_PREVIOUS;
	o->verbose = 0;
	o->amg_smoother = TFI_SMOOTH_AUTO;
	o->tangent = TFI_TANGENT_EXACT;
	o->amg_coarse_nodes = 0;
	o->amg_gs_min_rows = 0;
	o->amg_gs_sweeps = 0;
	o->chord_polish = 1;
	o->amg_half_smoother = 1;
}

int tfi_flow_create(tfi_mesh* mesh, const tfi_flow_options* options, tfi_flow** out)
{
	if (!mesh || !options || !out) return TFI_ERROR;
	*out = nullptr;
	return Guard(mesh->ctx, [&] {
		if (!(options->density > 0.0) || !(options->viscosity > 0.0) || !(options->dt >= 0.0))
			throw Error("invalid Navier-Stokes density, dynamic viscosity, or time step");
		if (options->amg_smoother < TFI_SMOOTH_JACOBI || options->amg_smoother > TFI_SMOOTH_AUTO)
			throw Error("unknown AMG smoother " + std::to_string(options->amg_smoother));
		if (options->warm_start < TFI_WARM_PREVIOUS || options->warm_start > TFI_WARM_MPE)
			throw Error("unknown warm-start mode " + std::to_string(options->warm_start));
		auto f = std::make_unique<tfi_flow>();
		f->mesh = mesh;
		f->ctx = mesh->ctx;
		f->opt = *options;
		tfi_ctx* ctx = mesh->ctx;
		const long long n = 4LL*mesh->nodes;
		f->state.Allocate(ctx, n);
		f->previous.Allocate(ctx, n);
		f->update.Allocate(ctx, n);
		f->last_update.Allocate(ctx, n);
		if (f->opt.warm_start == TFI_WARM_LEARNED) {
			f->guess_base.Allocate(ctx, n);
			f->resid_base.Allocate(ctx, n);
		}
		f->rhs.Allocate(ctx, n);
		f->traction.Allocate(ctx, n);
		f->jac.Allocate(ctx, mesh->nnzb*16);
		f->vtarget.Allocate(ctx, 3LL*mesh->nodes);
		f->ptarget.Allocate(ctx, mesh->nodes);
		f->vmask.Allocate(ctx, mesh->nodes);
		f->pmask.Allocate(ctx, mesh->nodes);
		f->cont_partial.Allocate(ctx, 512);
		for (auto* b : {&f->state, &f->previous, &f->update, &f->rhs, &f->traction, &f->vtarget, &f->ptarget})
			Fill(ctx, b->data(), 0.0, b->size());
		FillInt(ctx, f->vmask.data(), 0, mesh->nodes);
		FillInt(ctx, f->pmask.data(), 0, mesh->nodes);
		AmgOptions amg;
		// auto: Chebyshev (one fused launch per degree) for a transient step, whose
		// mass term keeps D^-1 A well inside the smoothing interval; multicolour
		// Gauss-Seidel for the steady (saddle-point dominated) equations
		amg.smoother = options->amg_smoother == TFI_SMOOTH_AUTO
			? (options->dt > 0.0 ? TFI_SMOOTH_CHEBYSHEV : TFI_SMOOTH_GAUSS_SEIDEL) : options->amg_smoother;
		if (options->amg_coarse_nodes > 0) amg.coarse_nodes = options->amg_coarse_nodes;
		if (options->amg_gs_min_rows > 0) amg.gs_min_rows = options->amg_gs_min_rows;
		if (options->amg_gs_sweeps > 0) amg.gs_sweeps = options->amg_gs_sweeps;
		if (options->amg_chebyshev_degree < 0 || options->amg_chebyshev_degree > 8)
			throw Error("amg_chebyshev_degree must be between 0 (default: 2) and 8");
		if (options->amg_chebyshev_degree > 0) amg.chebyshev_degree = options->amg_chebyshev_degree;
		amg.half_smoother = options->amg_half_smoother != 0 && options->precision == TFI_MIXED;
		f->solver = std::make_unique<LinearSolver>(mesh, options->preconditioner, options->precision,
			options->restart, amg);
		f->solver->set_verbose(options->verbose);
		f->warm = std::make_unique<WarmStart>(mesh);
		Sync(ctx);
		*out = f.release();
	});
}

void tfi_flow_destroy(tfi_flow* f)
{
	if (f && f->ctx) (void)hipStreamSynchronize(f->ctx->stream);
	delete f;
}

int tfi_flow_set_state(tfi_flow* f, const double* state, const double* previous)
{
	return Guard(f->ctx, [&] {
		const size_t n = 4ULL*f->mesh->nodes;
		f->state.Upload(state, n);
		f->previous.Upload(previous ? previous : state, n);
		f->warm->Reset();
	});
}

int tfi_flow_get_state(tfi_flow* f, double* state)
{
	return Guard(f->ctx, [&] { f->state.Download(state, 4ULL*f->mesh->nodes); });
}

int tfi_flow_set_boundary(tfi_flow* f, const int32_t* vmask, const int32_t* pmask, const double* vtarget,
	const double* ptarget, const double* traction)
{
	return Guard(f->ctx, [&] {
		const int N = f->mesh->nodes;
		f->vmask.Upload(vmask, N);
		f->pmask.Upload(pmask, N);
		f->vtarget.Upload(vtarget, 3ULL*N);
		f->ptarget.Upload(ptarget, N);
		f->has_traction = traction != nullptr;
		if (traction) f->traction.Upload(traction, 4ULL*N);
		f->has_boundary = true;
	});
}

int tfi_flow_set_boundary_faces(tfi_flow* f, int count, const int32_t* element, const int32_t* face,
	const int32_t* label)
{
	return Guard(f->ctx, [&] {
		if (count < 0) throw Error("negative face count");
		for (int i = 0; i < count; ++i)
			if (element[i] < 0 || element[i] >= f->mesh->elements || face[i] < 0 || face[i] > 5)
				throw Error("invalid boundary face");
		f->face_label.assign(label, label + count);
		f->face_elem.Allocate(f->ctx, count);
		f->face_id.Allocate(f->ctx, count);
		f->face_flow.Allocate(f->ctx, count);
		f->face_vec.Allocate(f->ctx, (size_t)count*kMaxBasis*3);
		if (count) {
			f->face_elem.Upload(element, count);
			f->face_id.Upload(face, count);
			// the face flows are linear in the state: their area-weighted basis
			// vectors are integrated once here (unit-pressure traction vectors)
			std::vector<double> ones(count, 1.0);
			DevBuf<double> unit(f->ctx, count);
			unit.Upload(ones.data(), count);
			FaceTraction(f->mesh, count, f->face_elem.data(), f->face_id.data(), unit.data(), f->face_vec.data());
			Sync(f->ctx);
		}
	});
}

int tfi_flow_begin_step(tfi_flow* f, const double* phase, int phase_count)
{
	return Guard(f->ctx, [&] {
		Copy(f->ctx, f->state.data(), f->previous.data(), 4LL*f->mesh->nodes);
		f->fresh_step = true;
		f->phase_count = std::min(phase_count, 8);
		for (int i = 0; i < f->phase_count; ++i) f->phase[i] = phase[i];
	});
}

int tfi_flow_commit_step(tfi_flow* f)
{
	return Guard(f->ctx, [&] {
		// history is kept for the predicting modes only; the network trains in learned mode
		if (f->opt.warm_start != TFI_WARM_PREVIOUS)
			f->warm->Commit(f->state.data(), f->phase, f->phase_count, f->opt.warm_start == TFI_WARM_LEARNED);
	});
}

// A residual check that misses the Newton tolerance by at most this factor is
// closed by a chord step (a solve with the tangent and preconditioner already
// set up, to a tenth of the tolerance) instead of a new tangent.
constexpr double kChordWindow = 1000.0;
constexpr int kMaxChords = 6;
// A learned guess whose linearised residual is below this fraction of the
// previous-step state's starts with chord steps on the tangent of an earlier
// step, each solved only as far as the observed contraction makes useful.
constexpr double kStaleStart = 1e-3;
// A Newton update that raises the residual, or after which the next solve
// fails, is backtracked by halves (at most this many times in a step).
constexpr int kMaxBacktracks = 6;
// Velocity step limit: a Newton update whose largest velocity change exceeds
// this fraction of the largest prescribed velocity is shortened to it (an
// impulsive start otherwise overshoots into states whose tangents the
// Krylov solver cannot handle).  Ordinary steps change the velocity far less
// and take the full update.
constexpr double kVelocityStep = 0.5;
constexpr double kChordContraction = 4.0;
// A learned correction whose gate (its least-squares scale on the latest unseen
// snapshot) is below this is not evaluated: the guess is the base prediction.
constexpr double kGateOff = 0.05;

int tfi_flow_newton(tfi_flow* f, tfi_newton_report* report)
{
	return Guard(f->ctx, [&] {
		if (!f->has_boundary) throw Error("tfi_flow_set_boundary must be called before tfi_flow_newton");
		tfi_ctx* ctx = f->ctx;
		tfi_mesh* m = f->mesh;
		const long long n = 4LL*m->nodes;
		const NsParams params{f->opt.density, f->opt.viscosity, f->opt.dt, f->opt.tangent};
		tfi_newton_report rep{};
		const auto t_start = Clock::now();
		const double scale = std::sqrt(static_cast<double>(n));
		double initial = -1.0;
		// Keep the iteration's record; with verbose > 0 also report it on stderr
		// as it happens (a long solve is visibly alive).
		auto push = [&](const tfi_newton_record& r, int iteration) {
			if (rep.records < TFI_MAX_NEWTON_RECORDS) rep.record[rep.records++] = r;
			if (f->opt.verbose > 0)
				std::fprintf(stderr, "tfi_flow newton=%d residual=%.6e relative_mass_imbalance=%.3e "
					"linear_iterations=%d linear_residual=%.3e step=%.3g assembly_s=%.3f linear_s=%.3f\n",
					iteration, r.residual_l2, r.relative_mass_imbalance, r.linear_iterations, r.linear_residual,
					r.step, r.assembly_s, r.linear_s);
		};
		// Residual (and tangent when jac is given) at the current state: traction
		// load, continuity / boundary-flow measurement, Dirichlet rows.
		auto assemble = [&](bool with_jacobian, Metrics* mt) {
			AssembleNs(m, f->state.data(), f->previous.data(), params, with_jacobian ? f->jac.data() : nullptr,
				f->rhs.data());
			if (f->has_traction) Axpy(ctx, 1.0, f->traction.data(), f->rhs.data(), n);
			if (mt) *mt = Measure(f);
			ApplyDirichlet(m, f->vmask.data(), f->pmask.data(), f->vtarget.data(), f->ptarget.data(), f->state.data(),
				with_jacobian ? f->jac.data() : nullptr, f->rhs.data());
			return Norm2(ctx, f->rhs.data(), n);
		};
		// learned / extrapolated initial state, accepted only if it lowers the
		// residual; the tangent assembled at an accepted guess is reused by the
		// first Newton iteration
		bool preassembled = false, pre_jacobian = true;
		Metrics pre_mt;
		double pre_norm = 0.0, pre_asm = 0.0;
		const bool fresh = f->fresh_step;
		f->fresh_step = false;
		// until the network has trained on its first snapshots, the learned mode
		// starts from its base prediction alone
		int mode = f->opt.warm_start;
		if (mode == TFI_WARM_LEARNED && !f->warm->Ready(mode)) mode = f->warm->BaseMode();
		if (fresh && mode != TFI_WARM_PREVIOUS && f->warm->Ready(mode)) {
			const auto a0 = Clock::now();
			const double base = assemble(false, nullptr);
			Copy(ctx, f->state.data(), f->update.data(), n);  // keep the plain guess
			double blend = 1.0;
			double predicted = std::numeric_limits<double>::infinity();  // linearised ||r|| of the guess
			bool guess_assembled = false;  // rhs and pre_mt already hold the guess's residual
			if (mode == TFI_WARM_LEARNED) {
				// Residual line search along the learned correction: with u_b the
				// base prediction and u_L = u_b + g c the network's guess, take
				// u = u_b + t (u_L - u_b) with t minimising the linearised
				// ||r_b + t (r_L - r_b)||, t in [0, 2].  t = 0 is the base alone.
				// A correction the gate has switched off is not evaluated: the
				// guess is u_b, whose residual the base assembly leaves in rhs.
				const bool network = f->warm->Gate() >= kGateOff;
				f->warm->PredictBase(f->state.data());
				SetDirichletState(m, f->vmask.data(), f->pmask.data(), f->vtarget.data(), f->ptarget.data(),
					f->state.data());
				const double nb = assemble(false, network ? nullptr : &pre_mt);
				rep.warm_start_base_ratio = base > 0.0 ? nb/base : 0.0;
				if (!network) {
					blend = 0.0;
					predicted = nb;
					guess_assembled = true;
				} else {
					Copy(ctx, f->rhs.data(), f->resid_base.data(), n);
					Copy(ctx, f->state.data(), f->guess_base.data(), n);
					f->warm->Predict(TFI_WARM_LEARNED, f->state.data(), f->phase, f->phase_count);
					SetDirichletState(m, f->vmask.data(), f->pmask.data(), f->vtarget.data(), f->ptarget.data(),
						f->state.data());
					const double nl = assemble(false, nullptr);
					const double bl = Dot(ctx, f->resid_base.data(), f->rhs.data(), n);
					const double dd = nl*nl - 2.0*bl + nb*nb;
					blend = dd > 1e-24*nb*nb ? std::min(2.0, std::max(0.0, (nb*nb - bl)/dd)) : 1.0;
					predicted = std::sqrt(std::max(0.0, nb*nb + 2.0*blend*(bl - nb*nb) + blend*blend*dd));
					Axpby(ctx, 1.0 - blend, f->guess_base.data(), blend, f->state.data(), n);
				}
			} else {
				// a plain prediction: its residual decides between a chord start and a new tangent
				f->warm->Predict(mode, f->state.data(), f->phase, f->phase_count);
				SetDirichletState(m, f->vmask.data(), f->pmask.data(), f->vtarget.data(), f->ptarget.data(),
					f->state.data());
				predicted = assemble(false, &pre_mt);
				guess_assembled = true;
			}
			rep.warm_start_blend = blend;
			// A guess predicted within the chord window of the tolerance is checked
			// with a residual-only assembly and, if it is not converged already,
			// closed by a chord step with the tangent already set up (from an
			// earlier state); a new tangent is assembled only if that fails.
			const double tol_guess = std::max(f->opt.nonlinear_atol*scale, f->opt.nonlinear_rtol*base);
			pre_jacobian = !(f->opt.chord_polish && f->tangent_ready
				&& predicted <= std::max(kChordWindow*tol_guess, kStaleStart*base));
			const double guessed = guess_assembled && !pre_jacobian ? predicted : assemble(pre_jacobian, &pre_mt);
			pre_asm = std::chrono::duration<double>(Clock::now() - a0).count();
			rep.assembly_s += pre_asm;
			rep.warm_start_initial_ratio = base > 0.0 ? guessed/base : 0.0;
			if (guessed < base) {
				rep.warm_start_used = 1;
				f->warm->Accepted(true);
				initial = base;  // relative tolerance measured against the plain start
				preassembled = true;
				pre_norm = guessed;
			} else {
				Copy(ctx, f->update.data(), f->state.data(), n);
				f->warm->Accepted(false);
			}
		}
		bool converged = false;
		// largest prescribed velocity of the step: the scale of the velocity step limit
		const double velocity_scale = MaxAbs(f, f->vtarget.data(), 3LL*m->nodes, 3, 3);
		double prev_norm = -1.0, contraction = 0.0;  // observed ||r_k|| / ||r_{k-1}||^2 this step
		double applied = 0.0;                         // fraction of last_update in the state (0: none)
		double norm_before = 0.0;                     // residual of the state last_update was computed at
		int backtracks = 0;
		for (int it = 0; it < f->opt.maximum_newton; ++it) {
			tfi_newton_record rec{};
			const auto a0 = Clock::now();
			Metrics mt;
			double norm = 0.0;
			bool have_jacobian = true;
			if (preassembled) {
				mt = pre_mt;
				norm = pre_norm;
				have_jacobian = pre_jacobian;
				preassembled = false;
			} else {
				// When quadratic convergence predicts this iterate converged, check
				// it with a residual-only assembly first and build the tangent only
				// if it did not.
				// (With chord polish, "converged" includes a miss a chord step closes.)
				const double c = std::max(contraction, f->newton_contraction);
				const double reach = f->opt.chord_polish ? kChordWindow : 0.25;
				const bool predict = prev_norm > 0.0 && initial > 0.0 && c > 0.0
					&& c*prev_norm*prev_norm <= reach*std::max(f->opt.nonlinear_atol*scale, f->opt.nonlinear_rtol*initial);
				have_jacobian = !predict;
				norm = assemble(have_jacobian, &mt);
			}
			const double tol_now = std::max(f->opt.nonlinear_atol*scale, f->opt.nonlinear_rtol*(initial < 0.0 ? norm : initial));
			double chord_s = 0.0;  // linear-solve time of a chord step inside this assembly phase
			if (!have_jacobian && !(norm <= tol_now && mt.imbalance <= f->opt.mass_rtol)) {
				// The prediction missed.  A small miss is closed by a chord step: one
				// more solve with the tangent and preconditioner of the last iteration
				// (still set up), to a tenth of the tolerance, then a residual check.
				// Further chord steps follow while each cuts the residual at least
				// kChordContraction-fold (a tangent of an earlier state may need two).
				// A step that starts on an earlier tangent (it == 0) has a wider
				// window, and each of its chord solves stops at a quarter of the last
				// observed contraction (the stale tangent limits what a more accurate
				// solve could gain); after a fresh tangent they go to 0.1 x tolerance.
				const bool stale = it == 0;
				const double window = stale ? std::max(kChordWindow*tol_now, kStaleStart*initial)
					: kChordWindow*tol_now;
				double rho = 0.04;  // assumed contraction of the first chord on an earlier tangent
				for (int chord = 0; chord < kMaxChords && f->opt.chord_polish && norm <= window
					&& !(norm <= tol_now && mt.imbalance <= f->opt.mass_rtol); ++chord) {
					const auto c0 = Clock::now();
					const double floor_eta = std::max(f->opt.linear_rtol, 0.1*tol_now/norm);
					const double eta = std::min(0.1, std::max(floor_eta, stale ? 0.25*rho : 0.0));
					Fill(ctx, f->update.data(), 0.0, n);
					const auto r = f->solver->Solve(f->rhs.data(), f->update.data(), f->opt.linear_maximum_iterations,
						eta);
					Sync(ctx);
					chord_s += std::chrono::duration<double>(Clock::now() - c0).count();
					rep.linear_iterations += r.iterations;
					if (!r.converged) break;
					Axpy(ctx, 1.0, f->update.data(), f->state.data(), n);
					++rep.chord_steps;
					const double before = norm;
					norm = assemble(false, &mt);
					rho = norm/before;
					if (!(norm*kChordContraction <= before)) break;
					// give up early when the observed contraction cannot reach the
					// tolerance within the chord budget
					if (norm > tol_now && rho > 0.0
						&& chord + 1 + std::ceil(std::log(norm/tol_now)/std::log(1.0/rho)) > kMaxChords) break;
				}
				rep.linear_s += chord_s;
				if (!(norm <= tol_now && mt.imbalance <= f->opt.mass_rtol))
					norm = assemble(true, &mt);  // full assembly at the current state
			}
			Sync(ctx);
			const auto a1 = Clock::now();
			rec.assembly_s = std::chrono::duration<double>(a1 - a0).count() - chord_s + (it == 0 ? pre_asm : 0.0);
			rep.assembly_s += std::chrono::duration<double>(a1 - a0).count() - chord_s;
			rec.residual_l2 = norm;
			rec.continuity_l2 = mt.continuity_l2;
			rec.continuity_sum = mt.continuity_sum;
			rec.net_boundary_flow = mt.net;
			rec.relative_mass_imbalance = mt.imbalance;
			if (initial < 0.0) initial = norm;
			if (it == 0) rep.initial_residual = norm;
			// quadratic-convergence constant ||r_k|| / ||r_{k-1}||^2, from residuals
			// clear of the rounding floor (a floor-level ||r_k|| says nothing about it)
			if (prev_norm > 0.0 && norm > 1e-11*initial)
				contraction = std::max(contraction, norm/(prev_norm*prev_norm));
			const double tol = std::max(f->opt.nonlinear_atol*scale, f->opt.nonlinear_rtol*initial);
			rep.tolerance = tol;
			rep.final_residual = norm;
			if (norm <= tol && mt.imbalance <= f->opt.mass_rtol) {
				converged = true;
				push(rec, it);
				rep.newton_iterations = it;
				break;
			}
			// Damped Newton: an update that raised the residual is halved (back
			// toward the state it started from) and the residual assembled again.
			if (applied > 0.0 && norm > norm_before && backtracks < kMaxBacktracks) {
				Axpy(ctx, -0.5*applied, f->last_update.data(), f->state.data(), n);
				applied *= 0.5;
				++backtracks;
				push(rec, it);
				rep.newton_iterations = it + 1;
				prev_norm = -1.0;  // the quadratic model does not hold across a backtrack
				continue;
			}
			const auto l0 = Clock::now();
			Fill(ctx, f->update.data(), 0.0, n);
			f->solver->Setup(f->jac.data());
			f->tangent_ready = true;
			const auto r = f->solver->Solve(f->rhs.data(), f->update.data(), f->opt.linear_maximum_iterations,
				f->opt.linear_rtol);
			Sync(ctx);
			const auto l1 = Clock::now();
			rec.linear_s = std::chrono::duration<double>(l1 - l0).count();
			rep.linear_s += rec.linear_s;
			rec.linear_iterations = r.iterations;
			rec.linear_residual = r.residual;
			rec.singular_diagonal_blocks = f->solver->singular_blocks();
			rep.linear_iterations += r.iterations;
			if (!r.converged && applied > 0.0 && backtracks < kMaxBacktracks) {
				// the tangent at this iterate defeated the solver: step back to the
				// middle of the last update and assemble there
				Axpy(ctx, -0.5*applied, f->last_update.data(), f->state.data(), n);
				applied *= 0.5;
				++backtracks;
				push(rec, it);
				rep.newton_iterations = it + 1;
				prev_norm = -1.0;  // the quadratic model does not hold across a backtrack
				continue;
			}
			if (!r.converged) {
				push(rec, it);
				rep.total_s = std::chrono::duration<double>(Clock::now() - t_start).count();
				if (report) *report = rep;
				throw Error("Navier-Stokes linear solve failed at nonlinear iteration " + std::to_string(it)
					+ " (residual=" + Precise(r.residual) + ", nonlinear_residual=" + Precise(norm)
					+ ", continuity_l2=" + Precise(mt.continuity_l2) + ", net_boundary_flow=" + Precise(mt.net)
					+ ", relative_mass_imbalance=" + Precise(mt.imbalance) + ")");
			}
			double step = 1.0;
			if (velocity_scale > 0.0) {
				const double dv = MaxAbs(f, f->update.data(), n, 4, 3);
				if (dv > kVelocityStep*velocity_scale) step = kVelocityStep*velocity_scale/dv;
			}
			rec.update_l2 = step*Norm2(ctx, f->update.data(), n);
			rec.step = step;
			Axpy(ctx, step, f->update.data(), f->state.data(), n);
			Copy(ctx, f->update.data(), f->last_update.data(), n);
			applied = step;
			norm_before = norm;
			push(rec, it);
			rep.newton_iterations = it + 1;
			prev_norm = norm;
		}
		if (contraction > 0.0) f->newton_contraction = contraction;
		Sync(ctx);
		rep.converged = converged ? 1 : 0;
		rep.total_s = std::chrono::duration<double>(Clock::now() - t_start).count();
		if (report) *report = rep;
		if (!converged) {
			const tfi_newton_record& last = rep.record[rep.records > 0 ? rep.records - 1 : 0];
			throw Error("Navier-Stokes nonlinear solve reached MAX_NEWTON (nonlinear_residual="
				+ Precise(rep.final_residual) + ", continuity_l2=" + Precise(last.continuity_l2)
				+ ", net_boundary_flow=" + Precise(last.net_boundary_flow) + ", relative_mass_imbalance="
				+ Precise(last.relative_mass_imbalance) + ")");
		}
	});
}

int tfi_flow_boundary_flows(tfi_flow* f, int count, const int32_t* labels, double* flows)
{
	return Guard(f->ctx, [&] {
		const int faces = static_cast<int>(f->face_label.size());
		std::vector<double> per(faces, 0.0);
		if (faces) {
			FaceFlowsFromVectors(f->mesh, faces, f->face_elem.data(), f->face_vec.data(), f->state.data(),
				f->face_flow.data());
			f->face_flow.Download(per.data(), faces);
		}
		for (int i = 0; i < count; ++i) {
			double s = 0.0;
			for (int k = 0; k < faces; ++k)
				if (f->face_label[k] == labels[i]) s += per[k];
			flows[i] = s;
		}
	});
}

int tfi_flow_amg_levels(tfi_flow* f, int capacity, int* levels, int* rows, long long* block_nonzeros, int* colors)
{
	return Guard(f->ctx, [&] {
		std::vector<int> r, c;
		std::vector<long long> z;
		if (f->solver->amg()) f->solver->amg()->Describe(r, z, c);
		*levels = static_cast<int>(r.size());
		for (int l = 0; l < static_cast<int>(r.size()) && l < capacity; ++l) {
			rows[l] = r[l];
			block_nonzeros[l] = z[l];
			colors[l] = c[l];
		}
	});
}

int tfi_flow_warm_configure(tfi_flow* f, const tfi_warm_options* o)
{
	return Guard(f->ctx, [&] { f->warm->Configure(*o); });
}

int tfi_flow_warm_stats(tfi_flow* f, int* snapshots, double* last_loss, int* accepted, int* rejected)
{
	return Guard(f->ctx, [&] { f->warm->Stats(snapshots, last_loss, accepted, rejected); });
}

int tfi_flow_warm_gate(tfi_flow* f, double* gate)
{
	return Guard(f->ctx, [&] {
		if (!gate) throw Error("tfi_flow_warm_gate needs an output pointer");
		*gate = f->warm->Gate();
	});
}

int tfi_flow_warm_record(tfi_flow* f, const double* state)
{
	return Guard(f->ctx, [&] {
		if (!state) throw Error("tfi_flow_warm_record needs a state");
		const long long n = 4LL*f->mesh->nodes;
		f->update.Upload(state, n);
		const double none[1] = {0.0};
		f->warm->Commit(f->update.data(), none, 0, false);
		Sync(f->ctx);
	});
}

int tfi_flow_warm_predict(tfi_flow* f, int mode, double* state, int* order)
{
	return Guard(f->ctx, [&] {
		if (!state) throw Error("tfi_flow_warm_predict needs an output array");
		if (mode == TFI_WARM_PREVIOUS || !f->warm->Ready(mode))
			throw Error("warm-start mode " + std::to_string(mode) + " has no prediction for the recorded history");
		const long long n = 4LL*f->mesh->nodes;
		Copy(f->ctx, f->state.data(), f->update.data(), n);
		f->warm->Predict(mode, f->update.data(), f->phase, f->phase_count);
		f->update.Download(state, n);
		if (order) *order = mode == TFI_WARM_MPE ? f->warm->MpeOrder() : 0;
	});
}

// ---------------------------------------------------------------- host in/out kernels

int tfi_ns_assemble(tfi_mesh* m, const double* state, const double* previous, double density, double viscosity,
	double dt, double* jacobian, double* negative_residual, int tangent)
{
	return Guard(m->ctx, [&] {
		if (!(density > 0.0) || !(viscosity > 0.0) || !(dt >= 0.0))
			throw Error("invalid Navier-Stokes density, dynamic viscosity, or time step");
		tfi_ctx* ctx = m->ctx;
		const size_t n = 4ULL*m->nodes;
		DevBuf<double> s(ctx, n), p(ctx, n), r(ctx, n), j;
		s.Upload(state, n);
		p.Upload(previous ? previous : state, n);
		if (jacobian) j.Allocate(ctx, m->nnzb*16);
		AssembleNs(m, s.data(), p.data(), NsParams{density, viscosity, dt, tangent}, jacobian ? j.data() : nullptr,
			r.data());
		r.Download(negative_residual, n);
		if (jacobian) j.Download(jacobian, m->nnzb*16);
	});
}

int tfi_bsr_spmv(tfi_mesh* m, const double* values, const double* x, double* y, int precision)
{
	return Guard(m->ctx, [&] {
		tfi_ctx* ctx = m->ctx;
		const size_t n = 4ULL*m->nodes;
		DevBuf<double> v(ctx, m->nnzb*16), xd(ctx, n), yd(ctx, n);
		v.Upload(values, m->nnzb*16);
		xd.Upload(x, n);
		if (precision == TFI_FP64) {
			Spmv64(ctx, m->view(), v.data(), xd.data(), yd.data());
		} else {
			DevBuf<float> v32(ctx, m->nnzb*16);
			ToFloat(ctx, v.data(), v32.data(), m->nnzb*16);
			Spmv32(ctx, m->view(), v32.data(), xd.data(), yd.data());
			Sync(ctx);
		}
		yd.Download(y, n);
	});
}

int tfi_ns_assemble_device(tfi_mesh* m, const double* state, const double* previous, double density,
	double viscosity, double dt, double* jacobian, double* negative_residual, int tangent)
{
	return Guard(m->ctx, [&] {
		if (!(density > 0.0) || !(viscosity > 0.0) || !(dt >= 0.0))
			throw Error("invalid Navier-Stokes density, dynamic viscosity, or time step");
		if (!state || !negative_residual) throw Error("tfi_ns_assemble_device needs a state and a residual array");
		if (tangent != TFI_TANGENT_EXACT && tangent != TFI_TANGENT_LEGACY) throw Error("unknown tangent");
		AssembleNs(m, state, previous ? previous : state, NsParams{density, viscosity, dt, tangent}, jacobian,
			negative_residual);
		Sync(m->ctx);
	});
}

int tfi_bsr_spmv_device(tfi_mesh* m, const double* values, const double* x, double* y)
{
	return Guard(m->ctx, [&] {
		if (!values || !x || !y) throw Error("tfi_bsr_spmv_device needs values, x and y");
		Spmv64(m->ctx, m->view(), values, x, y);
		Sync(m->ctx);
	});
}

int tfi_apply_dirichlet_device(tfi_mesh* m, const int32_t* velocity_mask, const int32_t* pressure_mask,
	const double* velocity_target, const double* pressure_target, const double* state, double* jacobian, double* rhs)
{
	return Guard(m->ctx, [&] {
		if (!velocity_mask || !pressure_mask || !velocity_target || !pressure_target || !state)
			throw Error("tfi_apply_dirichlet_device needs masks, targets and the state");
		ApplyDirichlet(m, velocity_mask, pressure_mask, velocity_target, pressure_target, state, jacobian, rhs);
		Sync(m->ctx);
	});
}

int tfi_block_inverse(tfi_mesh* m, const double* values, double* inverse, int* singular)
{
	return Guard(m->ctx, [&] {
		tfi_ctx* ctx = m->ctx;
		DevBuf<double> v(ctx, m->nnzb*16), inv(ctx, 16ULL*m->nodes);
		v.Upload(values, m->nnzb*16);
		const int s = BlockInverse64(ctx, m->view(), v.data(), inv.data());
		inv.Download(inverse, 16ULL*m->nodes);
		if (singular) *singular = s;
	});
}

int tfi_apply_dirichlet(tfi_mesh* m, const int32_t* vmask, const int32_t* pmask, const double* vtarget,
	const double* ptarget, const double* state, double* jacobian, double* rhs)
{
	return Guard(m->ctx, [&] {
		tfi_ctx* ctx = m->ctx;
		const int N = m->nodes;
		DevBuf<int> vm(ctx, N), pm(ctx, N);
		DevBuf<double> vt(ctx, 3ULL*N), pt(ctx, N), s(ctx, 4ULL*N), r, j;
		vm.Upload(vmask, N);
		pm.Upload(pmask, N);
		vt.Upload(vtarget, 3ULL*N);
		pt.Upload(ptarget, N);
		s.Upload(state, 4ULL*N);
		if (jacobian) {
			j.Allocate(ctx, m->nnzb*16);
			j.Upload(jacobian, m->nnzb*16);
		}
		if (rhs) {
			r.Allocate(ctx, 4ULL*N);
			r.Upload(rhs, 4ULL*N);
		}
		ApplyDirichlet(m, vm.data(), pm.data(), vt.data(), pt.data(), s.data(), jacobian ? j.data() : nullptr,
			rhs ? r.data() : nullptr);
		if (jacobian) j.Download(jacobian, m->nnzb*16);
		if (rhs) r.Download(rhs, 4ULL*N);
	});
}

int tfi_pressure_traction(tfi_mesh* m, int count, const int32_t* element, const int32_t* face,
	const double* pressure, double* traction)
{
	return Guard(m->ctx, [&] {
		tfi_ctx* ctx = m->ctx;
		const size_t n = 4ULL*m->nodes;
		std::fill(traction, traction + n, 0.0);
		if (count <= 0) return;
		for (int k = 0; k < count; ++k)
			if (element[k] < 0 || element[k] >= m->elements || face[k] < 0 || face[k] >= 6)
				throw Error("pressure-traction face " + std::to_string(k) + " has an invalid element or face index");
		DevBuf<int> el(ctx, count), fa(ctx, count);
		DevBuf<double> pr(ctx, count), local(ctx, static_cast<size_t>(count)*kMaxBasis*3);
		el.Upload(element, count);
		fa.Upload(face, count);
		pr.Upload(pressure, count);
		FaceTraction(m, count, el.data(), fa.data(), pr.data(), local.data());
		std::vector<double> host(static_cast<size_t>(count)*kMaxBasis*3);
		local.Download(host.data(), host.size());
		// sum in face order (elements ascending, faces 0..5 within an element), as
		// IntegratePressureTractionForces does
		for (int k = 0; k < count; ++k) {
			const int e = element[k];
			const int b0 = m->h_eoff[e], nen = m->h_eoff[e + 1] - b0;
			for (int a = 0; a < nen; ++a) {
				const size_t node = static_cast<size_t>(m->h_conn[b0 + a]);
				for (int c = 0; c < 3; ++c) traction[4*node + c] += host[(static_cast<size_t>(k)*kMaxBasis + a)*3 + c];
			}
		}
	});
}

int tfi_linear_solve(tfi_mesh* m, const double* values, const double* b, double* x, int preconditioner,
	int precision, int restart, int maxit, double rtol, int* iterations, double* residual)
{
	return Guard(m->ctx, [&] {
		tfi_ctx* ctx = m->ctx;
		const size_t n = 4ULL*m->nodes;
		DevBuf<double> v(ctx, m->nnzb*16), bd(ctx, n), xd(ctx, n);
		v.Upload(values, m->nnzb*16);
		bd.Upload(b, n);
		xd.Upload(x, n);
		LinearSolver solver(m, preconditioner, precision, restart);
		solver.Setup(v.data());
		const auto r = solver.Solve(bd.data(), xd.data(), maxit, rtol);
		xd.Download(x, n);
		if (iterations) *iterations = r.iterations;
		if (residual) *residual = r.residual;
		if (!r.converged) throw Error("linear solve did not converge");
	});
}

int tfi_bench_assembly(tfi_mesh* m, int tangent, int with_jacobian, int repetitions, double* milliseconds)
{
	return Guard(m->ctx, [&] {
		if (repetitions <= 0) throw Error("assembly benchmark needs a positive repetition count");
		tfi_ctx* ctx = m->ctx;
		const size_t n = 4ULL*m->nodes;
		DevBuf<double> s(ctx, n), p(ctx, n), r(ctx, n), j;
		std::vector<double> h(n);
		for (size_t i = 0; i < n; ++i) h[i] = 0.01*std::sin(0.37*static_cast<double>(i)) + (i%4 == 2 ? 0.05 : 0.0);
		s.Upload(h.data(), n);
		p.Upload(h.data(), n);
		if (with_jacobian) j.Allocate(ctx, m->nnzb*16);
		const NsParams params{1060.0*4e-6, 3.5e-3, 1e-3, tangent};
		auto run = [&] {
			AssembleNs(m, s.data(), p.data(), params, with_jacobian ? j.data() : nullptr, r.data());
		};
		run();
		Sync(ctx);
		Check(hipEventRecord(ctx->ev_begin, ctx->stream), "event");
		for (int i = 0; i < repetitions; ++i) run();
		Check(hipEventRecord(ctx->ev_end, ctx->stream), "event");
		Check(hipEventSynchronize(ctx->ev_end), "event sync");
		float ms = 0.0f;
		Check(hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end), "elapsed");
		*milliseconds = ms/repetitions;
	});
}

int tfi_bench_copy(tfi_ctx* ctx, long long count, int repetitions, double* milliseconds)
{
	return Guard(ctx, [&] {
		if (count <= 0 || repetitions <= 0) throw Error("copy benchmark needs a positive size and repetition count");
		DevBuf<double> a(ctx, count), b(ctx, count);
		Fill(ctx, a.data(), 1.0, count);
		Copy(ctx, a.data(), b.data(), count);
		Sync(ctx);
		Check(hipEventRecord(ctx->ev_begin, ctx->stream), "event");
		for (int i = 0; i < repetitions; ++i) Copy(ctx, a.data(), b.data(), count);
		Check(hipEventRecord(ctx->ev_end, ctx->stream), "event");
		Check(hipEventSynchronize(ctx->ev_end), "event sync");
		float ms = 0.0f;
		Check(hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end), "elapsed");
		*milliseconds = ms/repetitions;
	});
}

int tfi_bench_stream(tfi_ctx* ctx, long long count, int repetitions, double* copy_gbs, double* read_gbs)
{
	return Guard(ctx, [&] {
		if (count <= 0 || (count & 1) || repetitions <= 0 || !copy_gbs || !read_gbs)
			throw Error("stream probe needs an even positive size, a positive repetition count and outputs");
		const int blocks = 32*ctx->cus;
		DevBuf<double> a(ctx, count), b(ctx, count), partial(ctx, blocks);
		Fill(ctx, a.data(), 1.0, count);
		Fill(ctx, b.data(), 0.0, count);
		StreamCopy(ctx, a.data(), b.data(), count, blocks);
		StreamRead(ctx, a.data(), count, partial.data(), blocks);
		Sync(ctx);
		auto time = [&](auto&& launch) {
			Check(hipEventRecord(ctx->ev_begin, ctx->stream), "event");
			for (int i = 0; i < repetitions; ++i) launch();
			Check(hipEventRecord(ctx->ev_end, ctx->stream), "event");
			Check(hipEventSynchronize(ctx->ev_end), "event sync");
			float ms = 0.0f;
			Check(hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end), "elapsed");
			return static_cast<double>(ms)/repetitions;
		};
		const double copy_ms = time([&] { StreamCopy(ctx, a.data(), b.data(), count, blocks); });
		const double read_ms = time([&] { StreamRead(ctx, a.data(), count, partial.data(), blocks); });
		*copy_gbs = 16.0*count/(copy_ms*1e-3)/1e9;
		*read_gbs = 8.0*count/(read_ms*1e-3)/1e9;
	});
}

int tfi_bench_mfma_fp64(tfi_ctx* ctx, int repetitions, double* tflops)
{
	return Guard(ctx, [&] {
		if (repetitions <= 0 || !tflops) throw Error("FP64 MFMA probe needs a positive repetition count and an output");
		const int blocks = 8*ctx->cus, threads = 256, iterations = 20000;
		DevBuf<double> out(ctx, (size_t)blocks*threads);
		MfmaPeak(ctx, blocks, threads, iterations, out.data());
		Sync(ctx);
		Check(hipEventRecord(ctx->ev_begin, ctx->stream), "event");
		for (int i = 0; i < repetitions; ++i) MfmaPeak(ctx, blocks, threads, iterations, out.data());
		Check(hipEventRecord(ctx->ev_end, ctx->stream), "event");
		Check(hipEventSynchronize(ctx->ev_end), "event sync");
		float ms = 0.0f;
		Check(hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end), "elapsed");
		const double flops = 2048.0*4.0*iterations*(blocks*(threads/64.0))*repetitions;
		*tflops = flops/(ms*1e-3)/1e12;
	});
}

int tfi_bench_spmv(tfi_mesh* m, int precision, int repetitions, double* milliseconds)
{
	return Guard(m->ctx, [&] {
		tfi_ctx* ctx = m->ctx;
		const size_t n = 4ULL*m->nodes;
		DevBuf<double> v(ctx, m->nnzb*16), xd(ctx, n), yd(ctx, n);
		Fill(ctx, v.data(), 1e-3, m->nnzb*16);
		Fill(ctx, xd.data(), 1.0, n);
		DevBuf<float> v32;
		if (precision != TFI_FP64) {
			v32.Allocate(ctx, m->nnzb*16);
			ToFloat(ctx, v.data(), v32.data(), m->nnzb*16);
		}
		auto run = [&] {
			if (precision == TFI_FP64) Spmv64(ctx, m->view(), v.data(), xd.data(), yd.data());
			else Spmv32(ctx, m->view(), v32.data(), xd.data(), yd.data());
		};
		for (int i = 0; i < 3; ++i) run();
		Sync(ctx);
		Check(hipEventRecord(ctx->ev_begin, ctx->stream), "event");
		for (int i = 0; i < repetitions; ++i) run();
		Check(hipEventRecord(ctx->ev_end, ctx->stream), "event");
		Check(hipEventSynchronize(ctx->ev_end), "event sync");
		float ms = 0.0f;
		Check(hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end), "elapsed");
		*milliseconds = ms/std::max(1, repetitions);
	});
}

// AMG handle and learned-predictor training (C-ABI)
int tfi_amg_create(tfi_mesh* m, int precision, int smoother, tfi_amg** out)
{
	if (!m) return TFI_ERROR;
	return Guard(m->ctx, [&] {
		if (!out) throw Error("tfi_amg_create needs an output handle");
		if (precision != TFI_FP64 && precision != TFI_MIXED) throw Error("unknown precision");
		if (smoother < TFI_SMOOTH_JACOBI || smoother > TFI_SMOOTH_AUTO)
			throw Error("unknown AMG smoother " + std::to_string(smoother));
		auto a = std::make_unique<tfi_amg>();
		a->mesh = m;
		a->precision = precision;
		AmgOptions o;
		// without a time step to decide by, AUTO takes Gauss-Seidel (any matrix)
		o.smoother = smoother == TFI_SMOOTH_AUTO ? TFI_SMOOTH_GAUSS_SEIDEL : smoother;
		o.half_smoother = precision == TFI_MIXED && o.smoother == TFI_SMOOTH_CHEBYSHEV;
		a->amg = std::make_unique<Amg>(m, precision, o);
		*out = a.release();
	});
}

void tfi_amg_destroy(tfi_amg* a)
{
	delete a;
}

int tfi_amg_setup(tfi_amg* a, const double* values)
{
	if (!a) return TFI_ERROR;
	return Guard(a->mesh->ctx, [&] { AmgNumeric(a, values); });
}

int tfi_amg_update(tfi_amg* a, const double* values)
{
	if (!a) return TFI_ERROR;
	return Guard(a->mesh->ctx, [&] {
		if (!a->ready) throw Error("tfi_amg_update needs a hierarchy from tfi_amg_setup");
		AmgNumeric(a, values);
	});
}

int tfi_amg_apply(tfi_amg* a, const double* v, double* z)
{
	if (!a) return TFI_ERROR;
	return Guard(a->mesh->ctx, [&] {
		if (!a->ready) throw Error("tfi_amg_apply needs a hierarchy from tfi_amg_setup");
		if (!v || !z) throw Error("tfi_amg_apply needs v and z");
		a->amg->Apply(v, z);
		Sync(a->mesh->ctx);
	});
}

int tfi_amg_levels(tfi_amg* a, int capacity, int* levels, int* rows, long long* block_nonzeros, int* colors)
{
	if (!a) return TFI_ERROR;
	return Guard(a->mesh->ctx, [&] {
		if (!a->ready) throw Error("tfi_amg_levels needs a hierarchy from tfi_amg_setup");
		std::vector<int> r, c;
		std::vector<long long> nz;
		a->amg->Describe(r, nz, c);
		if (levels) *levels = static_cast<int>(r.size());
		for (int i = 0; i < capacity && i < static_cast<int>(r.size()); ++i) {
			if (rows) rows[i] = r[i];
			if (block_nonzeros) block_nonzeros[i] = nz[i];
			if (colors) colors[i] = c[i];
		}
	});
}

int tfi_flow_warm_train(tfi_flow* f, const double* state, const double* phase, int phase_count, double* loss)
{
	return Guard(f->ctx, [&] {
		if (!state) throw Error("tfi_flow_warm_train needs a state");
		if (phase_count < 0 || phase_count > 8 || (phase_count > 0 && !phase))
			throw Error("tfi_flow_warm_train takes at most 8 phase features");
		const long long n = 4LL*f->mesh->nodes;
		f->update.Upload(state, n);
		const double none[1] = {0.0};
		f->warm->Commit(f->update.data(), phase_count > 0 ? phase : none, phase_count, true);
		Sync(f->ctx);
		if (loss) {
			int snapshots = 0, accepted = 0, rejected = 0;
			f->warm->Stats(&snapshots, loss, &accepted, &rejected);
		}
	});
}

}  // extern "C"
