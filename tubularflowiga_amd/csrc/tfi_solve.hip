// Configured linear transport of 1..8 coupled scalar fields (`iga_cuda solve`,
// solvers/cuda/include/ConfiguredTransportHost.inc:142-471) with F x F node
// blocks on the Navier-Stokes mesh pattern.
//
// Assembly (AssembleGenericTransportKernel,
// solvers/cuda/include/GenericTransportKernels.cuh:26-143) builds, once per
// run, the backward-Euler left matrix L, the previous-state matrix P and the
// source s from the compiled weak-form terms: with the SUPG test function
// w_e = N_a + tau_e u.grad N_a of a stabilised equation e (else N_a),
//   time derivative  L, P += c w_e N_b dV
//   diffusion        L += dt c g_e dV,  g_e = grad N_a . grad N_b ("galerkin", no SUPG),
//                    w_e grad N_a . grad N_b ("supg_test"),
//                    grad N_a . grad N_b - tau_e u.grad N_a lap N_b ("consistent")
//   advection        L += dt c w_e u.grad N_b dV
//   linear coupling  L += dt c w_e N_b dV
//   volume source    s += dt c w_e dV
//   tau_e = (1/tau_s^2 + 1/tau_t^2 + [D_e > 0] 1/tau_d^2)^-1/2, tau_s = 1/sum_a |u.grad N_a|,
//   tau_t = dt/2, tau_d = |u|^2 / (D_e (1/tau_s)^2).
// The per-pair Gauss-point sums are kept per equation (w_e N_b, w_e u.grad N_b,
// g_e, w_e) and each term adds its multiple of one of them.  One workgroup per
// element stages the element's basis records in LDS; one thread per (a, b)
// pair; one element colour per launch, so the sum order is fixed and no
// atomics are needed.  Flux and Robin faces (AssembleGenericSurfaceKernel,
// :146-215) add dt r int N_a N_b dA to the diagonal entries of L and
// dt (q + r c_ext) int N_a dA to s, from per-face masses, in launches of faces
// whose elements share no node.  Constrained (node, field) rows become
// identity rows of L and zero rows of P (ApplyGenericBoundaryKernel, :217-233).
#include "tfi_internal.hpp"
#include "tfi_solver.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <type_traits>

struct tfi_solve {
	tfi_ctx* ctx = nullptr;
	tfi_mesh* mesh = nullptr;
	int fields = 0;
	tfi::DevBuf<double> left, previous, source, inverse, scaled_left, state, rhs, next, scaled, values, work;
	tfi::DevBuf<int> constrained;
	std::unique_ptr<tfi::Fgmres> krylov;
	int restart = 50;
	int singular = 0;
	// the compiled system, kept for re-assembly with a new velocity
	std::vector<tfi_solve_term> terms;
	std::vector<int> supg;
	std::vector<double> tau_diffusivity;
	double dt = 0.0;
	tfi::DevBuf<double> velocity;
	// flux / Robin faces: masses, loads, coefficients and launch groups
	int surface_count = 0;
	tfi::DevBuf<double> face_mass, face_load, face_robin, face_total;
	tfi::DevBuf<int> face_element, face_order;
	std::vector<std::pair<int, int>> face_spans;
	// optional coupled AMG on D^-1 L, its F x F blocks embedded in the 4x4
	// hierarchy with an identity on the padding fields (F <= 4)
	int preconditioner = TFI_PC_BJACOBI;
	std::unique_ptr<tfi::Amg> amg;
	bool amg_stale = true;
	tfi::DevBuf<double> padded64, pad_in, pad_out;
	tfi::DevBuf<float> padded32;
};

namespace tfi {

namespace {

constexpr int kThreads = 256;
constexpr int kRec = 5;  // basis table record: N, dN/dx, dN/dy, dN/dz, lap N
constexpr int kMaxFields = 8;
constexpr int kMaxTerms = 64;

struct SolveTerm {
	int kind, equation, trial;
	double coefficient;
};

struct SolveArgs {
	const int* eoff;
	const int* conn;
	const double* geometry;
	const double* table;
	const long long* pair_offsets;
	const int* pair_block;
	const int* elements;
	int count;
	const double* velocity;  // [3N]
	double dt;
	int term_count;
	SolveTerm terms[kMaxTerms];
	int supg[kMaxFields];
	double tau_diffusivity[kMaxFields];
	double* left;
	double* previous;
	double* source;
};

template <class Fn>
void Dispatch(int fields, Fn&& fn)
{
	switch (fields) {
	case 1: fn(std::integral_constant<int, 1>{}); break;
	case 2: fn(std::integral_constant<int, 2>{}); break;
	case 3: fn(std::integral_constant<int, 3>{}); break;
	case 4: fn(std::integral_constant<int, 4>{}); break;
	case 5: fn(std::integral_constant<int, 5>{}); break;
	case 6: fn(std::integral_constant<int, 6>{}); break;
	case 7: fn(std::integral_constant<int, 7>{}); break;
	case 8: fn(std::integral_constant<int, 8>{}); break;
	default: throw Error("configured transport supports 1 through 8 coupled scalar fields");
	}
}

// kStaged: the element's basis records are copied to LDS as [q][a][5]
template <int F, bool kStaged>
__global__ void __launch_bounds__(kThreads) SolveAssembleKernel(SolveArgs args)
{
	extern __shared__ double lds[];
	constexpr int kPoint = 4 + F;       // u (3), dV, tau per field
	double* point = lds;                // [q][kPoint]
	double* staged = lds + kQ*kPoint;   // [q][a][kRec] when kStaged
	const int item = blockIdx.x;
	if (item >= args.count) return;
	const int e = args.elements[item];
	const int b0 = args.eoff[e];
	const int nen = args.eoff[e + 1] - b0;
	const double* tab = args.table + (size_t)b0*kQ*kRec;
	if (kStaged) {
		for (int t = threadIdx.x; t < kQ*nen*kRec; t += kThreads) staged[t] = tab[t];
		__syncthreads();
	}
	const double* basis = kStaged ? staged : tab;
	if (threadIdx.x < kQ) {
		const int q = threadIdx.x;
		const double* bq = basis + (size_t)kRec*q*nen;
		double u[3] = {0.0, 0.0, 0.0};
		for (int a = 0; a < nen; ++a) {
			const size_t node = args.conn[b0 + a];
			for (int d = 0; d < 3; ++d) u[d] += bq[kRec*a]*args.velocity[3*node + d];
		}
		double inverse_length = 0.0;
		for (int a = 0; a < nen; ++a)
			inverse_length += fabs(u[0]*bq[kRec*a + 1] + u[1]*bq[kRec*a + 2] + u[2]*bq[kRec*a + 3]);
		const double tau_space = inverse_length > 0.0 ? 1.0/inverse_length : 0.0;
		const double tau_time = args.dt/2.0;
		const double speed2 = u[0]*u[0] + u[1]*u[1] + u[2]*u[2];
		double* pt = point + kPoint*q;
		pt[0] = u[0];
		pt[1] = u[1];
		pt[2] = u[2];
		pt[3] = args.geometry[(size_t)e*kGeomFields*kQ + q];
		for (int f = 0; f < F; ++f) {
			double tau = 0.0;
			if (tau_space > 0.0) {
				double inverse_square = 1.0/(tau_space*tau_space) + 1.0/(tau_time*tau_time);
				const double d = args.tau_diffusivity[f];
				if (d > 0.0) {
					const double tau_diffusion = speed2/(d*inverse_length*inverse_length);
					inverse_square += 1.0/(tau_diffusion*tau_diffusion);
				}
				tau = 1.0/sqrt(inverse_square);
			}
			pt[4 + f] = tau;
		}
	}
	__syncthreads();
	const long long pbase = args.pair_offsets[e];
	for (int pair = threadIdx.x; pair < nen*nen; pair += kThreads) {
		const int a = pair/nen, b = pair - a*nen;
		const bool first = b == 0;
		double mass[F], adv[F], dif[F], src[F];
#pragma unroll
		for (int f = 0; f < F; ++f) mass[f] = adv[f] = dif[f] = src[f] = 0.0;
		for (int q = 0; q < kQ; ++q) {
			const double* ba = basis + (size_t)kRec*(q*nen + a);
			const double* bb = basis + (size_t)kRec*(q*nen + b);
			const double* pt = point + kPoint*q;
			const double adv_a = pt[0]*ba[1] + pt[1]*ba[2] + pt[2]*ba[3];
			const double adv_b = pt[0]*bb[1] + pt[1]*bb[2] + pt[2]*bb[3];
			const double gradient_dot = ba[1]*bb[1] + ba[2]*bb[2] + ba[3]*bb[3];
			const double dv = pt[3];
#pragma unroll
			for (int f = 0; f < F; ++f) {
				const int s = args.supg[f];
				const double test = ba[0] + (s ? pt[4 + f]*adv_a : 0.0);
				mass[f] += test*bb[0]*dv;
				adv[f] += test*adv_b*dv;
				double g = s == TFI_SUPG_TEST ? test*gradient_dot : gradient_dot;
				if (s == TFI_SUPG_CONSISTENT) g -= pt[4 + f]*adv_a*bb[4];
				dif[f] += g*dv;
				if (first) src[f] += test*dv;
			}
		}
		const size_t k = args.pair_block[pbase + pair];
		double* l = args.left + (size_t)F*F*k;
		double* p = args.previous + (size_t)F*F*k;
		const size_t row = args.conn[b0 + a];
		for (int t = 0; t < args.term_count; ++t) {
			const SolveTerm term = args.terms[t];
			double m = 0.0, v = 0.0, d = 0.0, w = 0.0;
#pragma unroll
			for (int f = 0; f < F; ++f)
				if (f == term.equation) {
					m = mass[f];
					v = adv[f];
					d = dif[f];
					w = src[f];
				}
			const int entry = term.equation*F + term.trial;
			const double c = term.coefficient;
			switch (term.kind) {
			case TFI_TERM_TIME_DERIVATIVE:
				l[entry] += c*m;
				p[entry] += c*m;
				break;
			case TFI_TERM_DIFFUSION: l[entry] += args.dt*c*d; break;
			case TFI_TERM_ADVECTION: l[entry] += args.dt*c*v; break;
			case TFI_TERM_LINEAR_COUPLING: l[entry] += args.dt*c*m; break;
			case TFI_TERM_VOLUME_SOURCE:
				if (first) args.source[F*row + term.equation] += args.dt*c*w;
				break;
			default: break;
			}
		}
	}
}

// Robin and flux faces: one workgroup per face of one group (elements of a
// group share no node), one thread per (a, b) pair of the face's element
template <int F>
__global__ void __launch_bounds__(kThreads) SolveSurfaceKernel(int count, const int* __restrict__ faces,
	const int* __restrict__ felem, const int* __restrict__ eoff, const int* __restrict__ conn,
	const long long* __restrict__ pair_offsets, const int* __restrict__ pair_block, const double* __restrict__ mass,
	const double* __restrict__ load, int stride, const double* __restrict__ robin,
	const double* __restrict__ flux_total, double dt, double* __restrict__ left, double* __restrict__ source)
{
	if ((int)blockIdx.x >= count) return;
	const int f = faces[blockIdx.x];
	const int e = felem[f];
	const int b0 = eoff[e];
	const int nen = eoff[e + 1] - b0;
	const long long pbase = pair_offsets[e];
	for (int pair = threadIdx.x; pair < nen*nen; pair += kThreads) {
		const int a = pair/nen, b = pair - a*nen;
		const double m = mass[((size_t)f*stride + a)*stride + b];
		double* l = left + (size_t)F*F*pair_block[pbase + pair];
		for (int field = 0; field < F; ++field) l[field*F + field] += dt*robin[(size_t)f*F + field]*m;
		if (b == 0) {
			const double s = load[(size_t)f*stride + a];
			const size_t row = conn[b0 + a];
			for (int field = 0; field < F; ++field) source[F*row + field] += dt*flux_total[(size_t)f*F + field]*s;
		}
	}
}

template <int F>
__global__ void SolveBoundaryKernel(int nodes, const int* __restrict__ ro, const int* __restrict__ diag,
	const int* __restrict__ constrained, double* __restrict__ left, double* __restrict__ previous)
{
	const int node = blockIdx.x*blockDim.x + threadIdx.x;
	if (node >= nodes) return;
	for (int field = 0; field < F; ++field) {
		if (!constrained[(size_t)F*node + field]) continue;
		for (int k = ro[node]; k < ro[node + 1]; ++k)
			for (int c = 0; c < F; ++c) {
				left[(size_t)F*F*k + F*field + c] = 0.0;
				previous[(size_t)F*F*k + F*field + c] = 0.0;
			}
		left[(size_t)F*F*diag[node] + F*field + field] = 1.0;
	}
}

// Inverse of each diagonal block by Gauss-Jordan elimination with partial
// pivoting.  A pivot below 1e-14 in magnitude (or not finite) marks the block
// singular, as BuildBlockInverseKernel (solvers/cuda/include/SparseKernels.cuh:42-90)
// does: it is then inverted by its diagonal (1 where |d_ii| <= 1e-14) and counted.
template <int F>
__global__ void SolveBlockInverseKernel(int rows, const int* __restrict__ diag, const double* __restrict__ vals,
	double* __restrict__ inv, int* __restrict__ singular)
{
	const int row = blockIdx.x*blockDim.x + threadIdx.x;
	if (row >= rows) return;
	const double* d = vals + (size_t)F*F*diag[row];
	double a[F][F], x[F][F];
	for (int i = 0; i < F; ++i)
		for (int j = 0; j < F; ++j) {
			a[i][j] = d[F*i + j];
			x[i][j] = i == j ? 1.0 : 0.0;
		}
	bool ok = true;
	for (int c = 0; c < F && ok; ++c) {
		int best = c;
		for (int r = c + 1; r < F; ++r)
			if (fabs(a[r][c]) > fabs(a[best][c])) best = r;
		const double pivot = a[best][c];
		if (!isfinite(pivot) || fabs(pivot) < 1e-14) {
			ok = false;
			break;
		}
		for (int j = 0; j < F; ++j) {
			const double ta = a[c][j], tx = x[c][j];
			a[c][j] = a[best][j] / pivot;
			x[c][j] = x[best][j] / pivot;
			if (best != c) {
				a[best][j] = ta;
				x[best][j] = tx;
			}
		}
		for (int r = 0; r < F; ++r) {
			if (r == c) continue;
			const double factor = a[r][c];
			for (int j = 0; j < F; ++j) {
				a[r][j] -= factor*a[c][j];
				x[r][j] -= factor*x[c][j];
			}
		}
	}
	double* o = inv + (size_t)F*F*row;
	if (ok) {
		for (int i = 0; i < F; ++i)
			for (int j = 0; j < F; ++j) o[F*i + j] = x[i][j];
	} else {
		for (int i = 0; i < F; ++i)
			for (int j = 0; j < F; ++j)
				o[F*i + j] = i != j ? 0.0 : (fabs(d[F*i + i]) > 1e-14 ? 1.0/d[F*i + i] : 1.0);
		atomicAdd(singular, 1);
	}
}

// S = D^-1 L row by row
template <int F>
__global__ void SolveScaleRowsKernel(int rows, const int* __restrict__ ro, const double* __restrict__ inv,
	const double* __restrict__ left, double* __restrict__ scaled)
{
	const int row = blockIdx.x*blockDim.x + threadIdx.x;
	if (row >= rows) return;
	const double* o = inv + (size_t)F*F*row;
	for (int k = ro[row]; k < ro[row + 1]; ++k) {
		const double* a = left + (size_t)F*F*k;
		double* s = scaled + (size_t)F*F*k;
		for (int i = 0; i < F; ++i)
			for (int j = 0; j < F; ++j) {
				double v = 0.0;
				for (int m = 0; m < F; ++m) v += o[F*i + m]*a[F*m + j];
				s[F*i + j] = v;
			}
	}
}

template <int F>
__global__ void SolveBlockApplyKernel(int rows, const double* __restrict__ inv, const double* __restrict__ v,
	double* __restrict__ z)
{
	const int row = blockIdx.x*blockDim.x + threadIdx.x;
	if (row >= rows) return;
	const double* o = inv + (size_t)F*F*row;
	double in[F];
	for (int j = 0; j < F; ++j) in[j] = v[(size_t)F*row + j];
	for (int i = 0; i < F; ++i) {
		double s = 0.0;
		for (int j = 0; j < F; ++j) s += o[F*i + j]*in[j];
		z[(size_t)F*row + i] = s;
	}
}

// y = A x with F x F blocks: one 64-lane wave per block row, each lane over
// every 64th block of its row, reduced across the wave
template <int F>
__global__ void __launch_bounds__(kThreads) SolveSpmvKernel(int rows, const int* __restrict__ ro,
	const int* __restrict__ cols, const double* __restrict__ vals, const double* __restrict__ x, double* __restrict__ y)
{
	const int sub = threadIdx.x & 63;
	const int row = (blockIdx.x*kThreads + threadIdx.x) >> 6;
	double acc[F];
#pragma unroll
	for (int i = 0; i < F; ++i) acc[i] = 0.0;
	if (row < rows) {
		for (int k = ro[row] + sub; k < ro[row + 1]; k += 64) {
			const double* v = vals + (size_t)F*F*k;
			const double* xc = x + (size_t)F*cols[k];
			double xv[F];
#pragma unroll
			for (int j = 0; j < F; ++j) xv[j] = xc[j];
#pragma unroll
			for (int i = 0; i < F; ++i)
#pragma unroll
				for (int j = 0; j < F; ++j) acc[i] += v[F*i + j]*xv[j];
		}
	}
#pragma unroll
	for (int i = 0; i < F; ++i)
		for (int o = 32; o > 0; o >>= 1) acc[i] += __shfl_xor(acc[i], o, 64);
	if (row < rows && sub < F) {
		double out = 0.0;
#pragma unroll
		for (int i = 0; i < F; ++i)
			if (i == sub) out = acc[i];
		y[(size_t)F*row + sub] = out;
	}
}

// 4x4 embedding of the F x F blocks (F <= 4): [[S, 0], [0, I if diagonal block else 0]]
template <int F>
__global__ void SolvePadBlocksKernel(int rows, const int* __restrict__ ro, const int* __restrict__ diag,
	const double* __restrict__ s, double* __restrict__ out)
{
	const int row = blockIdx.x*blockDim.x + threadIdx.x;
	if (row >= rows) return;
	for (int k = ro[row]; k < ro[row + 1]; ++k) {
		const double* a = s + (size_t)F*F*k;
		double* o = out + 16*(size_t)k;
		for (int i = 0; i < 4; ++i)
			for (int j = 0; j < 4; ++j)
				o[4*i + j] = (i < F && j < F) ? a[F*i + j] : (i == j && k == diag[row] ? 1.0 : 0.0);
	}
}

// [F per node] <-> [4 per node, padding zero]
template <int F>
__global__ void SolvePadVectorKernel(int rows, const double* __restrict__ v, double* __restrict__ p)
{
	const int row = blockIdx.x*blockDim.x + threadIdx.x;
	if (row >= rows) return;
	for (int i = 0; i < 4; ++i) p[4*(size_t)row + i] = i < F ? v[(size_t)F*row + i] : 0.0;
}

template <int F>
__global__ void SolveUnpadVectorKernel(int rows, const double* __restrict__ p, double* __restrict__ v)
{
	const int row = blockIdx.x*blockDim.x + threadIdx.x;
	if (row >= rows) return;
	for (int i = 0; i < F; ++i) v[(size_t)F*row + i] = p[4*(size_t)row + i];
}

__global__ void SolveValuesKernel(long long n, const int* __restrict__ constrained, const double* __restrict__ values,
	double* __restrict__ v)
{
	const long long i = (long long)blockIdx.x*blockDim.x + threadIdx.x;
	if (i < n && constrained[i]) v[i] = values[i];
}

int Grid(long long n) { return static_cast<int>((n + kThreads - 1)/kThreads); }

template <class Fn>
int Guard(tfi_ctx* ctx, Fn&& fn)
{
	try {
		fn();
		return TFI_OK;
	} catch (const std::exception& e) {
		if (ctx) ctx->error = e.what();
		return TFI_ERROR;
	}
}

void Spmv(tfi_solve* s, const double* vals, const double* x, double* y)
{
	const BsrView a = s->mesh->view();
	Dispatch(s->fields, [&](auto fc) {
		constexpr int F = decltype(fc)::value;
		hipLaunchKernelGGL(SolveSpmvKernel<F>, dim3((a.rows + kThreads/64 - 1)/(kThreads/64)), dim3(kThreads), 0,
			s->ctx->stream, a.rows, a.row_offsets, a.columns, vals, x, y);
	});
	TFI_LAUNCH_CHECK("SolveSpmvKernel");
}

void BlockApply(tfi_solve* s, const double* v, double* z)
{
	const int N = s->mesh->nodes;
	Dispatch(s->fields, [&](auto fc) {
		constexpr int F = decltype(fc)::value;
		hipLaunchKernelGGL(SolveBlockApplyKernel<F>, dim3(Grid(N)), dim3(kThreads), 0, s->ctx->stream, N,
			s->inverse.data(), v, z);
	});
	TFI_LAUNCH_CHECK("SolveBlockApplyKernel");
}

void Validate(const tfi_mesh* mesh, const tfi_solve_system* sys)
{
	if (!sys) throw Error("tfi_solve_create needs a system");
	if (sys->fields < 1 || sys->fields > kMaxFields)
		throw Error("configured transport supports 1 through 8 coupled scalar fields");
	if (sys->term_count < 0 || sys->term_count > kMaxTerms)
		throw Error("configured transport supports at most 64 weak-form terms");
	if (sys->term_count > 0 && !sys->terms) throw Error("configured transport terms are missing");
	if (!sys->supg || !sys->tau_diffusivity) throw Error("configured transport needs supg and tau_diffusivity");
	if (!(sys->dt > 0.0)) throw Error("configured transport dt must be positive");
	for (int t = 0; t < sys->term_count; ++t) {
		const tfi_solve_term& term = sys->terms[t];
		if (term.kind < TFI_TERM_TIME_DERIVATIVE || term.kind > TFI_TERM_VOLUME_SOURCE)
			throw Error("configured transport term kind out of range");
		if (term.equation < 0 || term.equation >= sys->fields || term.trial < 0 || term.trial >= sys->fields)
			throw Error("configured transport term field out of range");
	}
	for (int f = 0; f < sys->fields; ++f)
		if (sys->supg[f] < TFI_SUPG_NONE || sys->supg[f] > TFI_SUPG_CONSISTENT)
			throw Error("configured transport SUPG setting out of range");
	if (sys->surface_count < 0) throw Error("configured transport surface count is negative");
	if (sys->surface_count > 0) {
		if (!sys->surface_element || !sys->surface_face || !sys->surface_flux || !sys->surface_robin
			|| !sys->surface_robin_source)
			throw Error("configured transport surface arrays are missing");
		for (int i = 0; i < sys->surface_count; ++i) {
			if (sys->surface_element[i] < 0 || sys->surface_element[i] >= mesh->elements)
				throw Error("configured transport surface element out of range");
			if (sys->surface_face[i] < 0 || sys->surface_face[i] > 5)
				throw Error("configured transport surface face out of range");
		}
	}
	if (mesh->max_nen > kMaxBasis) throw Error("configured transport supports at most 80 basis functions");
}

// Face masses of the flux / Robin faces (geometry only, computed once) and
// their grouping by (colour of the element, occurrence of the element in the
// list): the elements of one group are distinct and share no node
void SetupSurfaces(tfi_solve* s, const tfi_solve_system* sys)
{
	tfi_ctx* ctx = s->ctx;
	tfi_mesh* mesh = s->mesh;
	const int count = sys->surface_count;
	const int F = s->fields;
	s->surface_count = count;
	if (count == 0) return;
	std::map<std::pair<int, int>, std::vector<int>> groups;
	std::vector<int> seen(mesh->elements, 0);
	for (int i = 0; i < count; ++i) {
		const int e = sys->surface_element[i];
		groups[{mesh->h_color[e], seen[e]++}].push_back(i);
	}
	std::vector<double> total((size_t)count*F);
	for (size_t i = 0; i < total.size(); ++i) total[i] = sys->surface_flux[i] + sys->surface_robin_source[i];
	DevBuf<int> face;
	face.Allocate(ctx, count);
	face.Upload(sys->surface_face, count);
	s->face_element.Allocate(ctx, count);
	s->face_element.Upload(sys->surface_element, count);
	s->face_robin.Allocate(ctx, (size_t)count*F);
	s->face_total.Allocate(ctx, (size_t)count*F);
	s->face_robin.Upload(sys->surface_robin, (size_t)count*F);
	s->face_total.Upload(total.data(), total.size());
	const int S = mesh->max_nen;
	s->face_mass.Allocate(ctx, (size_t)count*S*S);
	s->face_load.Allocate(ctx, (size_t)count*S);
	FaceMass(mesh, count, s->face_element.data(), face.data(), s->face_mass.data(), s->face_load.data());
	std::vector<int> sorted;
	for (const auto& g : groups) {
		s->face_spans.push_back({static_cast<int>(sorted.size()), static_cast<int>(g.second.size())});
		sorted.insert(sorted.end(), g.second.begin(), g.second.end());
	}
	s->face_order.Allocate(ctx, count);
	s->face_order.Upload(sorted.data(), sorted.size());
	Sync(ctx);
}

// L, P and the source for the nodal velocity [3N] on the device: element
// colours, then the faces, then the constrained rows, the diagonal-block
// inverses and D^-1 L
void Assemble(tfi_solve* s)
{
	tfi_ctx* ctx = s->ctx;
	tfi_mesh* mesh = s->mesh;
	const int F = s->fields;
	const int N = mesh->nodes;
	Fill(ctx, s->left.data(), 0.0, s->left.size());
	Fill(ctx, s->previous.data(), 0.0, s->previous.size());
	Fill(ctx, s->source.data(), 0.0, s->source.size());
	SolveArgs a{};
	a.eoff = mesh->eoff.data();
	a.conn = mesh->conn.data();
	a.geometry = mesh->geometry.data();
	a.table = mesh->basis_table.data();
	a.pair_offsets = mesh->pair_offsets.data();
	a.pair_block = mesh->pair_block.data();
	a.velocity = s->velocity.data();
	a.dt = s->dt;
	a.term_count = static_cast<int>(s->terms.size());
	for (size_t t = 0; t < s->terms.size(); ++t)
		a.terms[t] = {s->terms[t].kind, s->terms[t].equation, s->terms[t].trial, s->terms[t].coefficient};
	for (int f = 0; f < F; ++f) {
		a.supg[f] = s->supg[f];
		a.tau_diffusivity[f] = s->tau_diffusivity[f];
	}
	a.left = s->left.data();
	a.previous = s->previous.data();
	a.source = s->source.data();
	const size_t point_bytes = sizeof(double)*kQ*(4 + F);
	const size_t staged_bytes = sizeof(double)*(size_t)kQ*mesh->max_nen*kRec;
	const bool staged = point_bytes + staged_bytes <= 160*1024;
	const size_t lds = point_bytes + (staged ? staged_bytes : 0);
	{
		StreamTimer timer(ctx, &ctx->timers.transport_assembly_ms);
		Dispatch(F, [&](auto fc) {
			constexpr int FF = decltype(fc)::value;
			const void* kernel = staged ? reinterpret_cast<const void*>(&SolveAssembleKernel<FF, true>)
				: reinterpret_cast<const void*>(&SolveAssembleKernel<FF, false>);
			if (lds > 65536)
				Check(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)),
					"hipFuncSetAttribute");
			for (int c = 0; c < mesh->colors; ++c) {
				a.elements = mesh->color_elements.data() + mesh->color_offsets[c];
				a.count = mesh->color_offsets[c + 1] - mesh->color_offsets[c];
				if (a.count == 0) continue;
				if (staged)
					hipLaunchKernelGGL((SolveAssembleKernel<FF, true>), dim3(a.count), dim3(kThreads), lds, ctx->stream,
						a);
				else
					hipLaunchKernelGGL((SolveAssembleKernel<FF, false>), dim3(a.count), dim3(kThreads), lds,
						ctx->stream, a);
				TFI_LAUNCH_CHECK("SolveAssembleKernel");
			}
			for (const auto& span : s->face_spans) {
				hipLaunchKernelGGL(SolveSurfaceKernel<FF>, dim3(span.second), dim3(kThreads), 0, ctx->stream,
					span.second, s->face_order.data() + span.first, s->face_element.data(), mesh->eoff.data(),
					mesh->conn.data(), mesh->pair_offsets.data(), mesh->pair_block.data(), s->face_mass.data(),
					s->face_load.data(), mesh->max_nen, s->face_robin.data(), s->face_total.data(), s->dt,
					s->left.data(), s->source.data());
				TFI_LAUNCH_CHECK("SolveSurfaceKernel");
			}
		});
	}
	const BsrView v = mesh->view();
	DevBuf<int> singular;
	singular.Allocate(ctx, 1);
	FillInt(ctx, singular.data(), 0, 1);
	Dispatch(F, [&](auto fc) {
		constexpr int FF = decltype(fc)::value;
		hipLaunchKernelGGL(SolveBoundaryKernel<FF>, dim3(Grid(N)), dim3(kThreads), 0, ctx->stream, N,
			v.row_offsets, v.diagonal, s->constrained.data(), s->left.data(), s->previous.data());
		TFI_LAUNCH_CHECK("SolveBoundaryKernel");
		hipLaunchKernelGGL(SolveBlockInverseKernel<FF>, dim3(Grid(N)), dim3(kThreads), 0, ctx->stream, N,
			v.diagonal, s->left.data(), s->inverse.data(), singular.data());
		TFI_LAUNCH_CHECK("SolveBlockInverseKernel");
		hipLaunchKernelGGL(SolveScaleRowsKernel<FF>, dim3(Grid(N)), dim3(kThreads), 0, ctx->stream, N,
			v.row_offsets, s->inverse.data(), s->left.data(), s->scaled_left.data());
		TFI_LAUNCH_CHECK("SolveScaleRowsKernel");
	});
	int count = 0;
	singular.Download(&count, 1);
	s->singular = std::max(s->singular, count);
	s->amg_stale = true;
	Sync(ctx);
}

// The AMG hierarchy for the current D^-1 L (numeric set-up again after a
// re-assembly); FP32 V-cycles
void SetupAmg(tfi_solve* s)
{
	tfi_ctx* ctx = s->ctx;
	const BsrView a = s->mesh->view();
	const int N = s->mesh->nodes;
	if (!s->amg) {
		s->padded64.Allocate(ctx, (size_t)a.nnzb*16);
		s->padded32.Allocate(ctx, (size_t)a.nnzb*16);
		s->pad_in.Allocate(ctx, 4LL*N);
		s->pad_out.Allocate(ctx, 4LL*N);
		s->amg = std::make_unique<Amg>(s->mesh, TFI_MIXED, AmgOptions());
	}
	Dispatch(s->fields, [&](auto fc) {
		constexpr int FF = decltype(fc)::value;
		if constexpr (FF <= 4) {
			hipLaunchKernelGGL(SolvePadBlocksKernel<FF>, dim3(Grid(N)), dim3(kThreads), 0, ctx->stream, N,
				a.row_offsets, a.diagonal, s->scaled_left.data(), s->padded64.data());
			TFI_LAUNCH_CHECK("SolvePadBlocksKernel");
		} else {
			throw Error("the AMG preconditioner of the configured transport supports at most 4 fields");
		}
	});
	ToFloat(ctx, s->padded64.data(), s->padded32.data(), (long long)a.nnzb*16);
	s->amg->Setup(s->padded64.data(), s->padded32.data());
	s->amg_stale = false;
}

}  // namespace

}  // namespace tfi

using namespace tfi;

int tfi_solve_create(tfi_mesh* mesh, const tfi_solve_system* system, const double* velocity,
	const int32_t* constrained, tfi_solve** out)
{
	*out = nullptr;
	return Guard(mesh->ctx, [&] {
		Validate(mesh, system);
		if (!velocity || !constrained) throw Error("tfi_solve_create needs velocity and constrained mask");
		auto s = std::make_unique<tfi_solve>();
		tfi_ctx* ctx = mesh->ctx;
		s->ctx = ctx;
		s->mesh = mesh;
		const int F = system->fields;
		s->fields = F;
		s->terms.assign(system->terms, system->terms + system->term_count);
		s->supg.assign(system->supg, system->supg + F);
		s->tau_diffusivity.assign(system->tau_diffusivity, system->tau_diffusivity + F);
		s->dt = system->dt;
		const int N = mesh->nodes;
		const long long n = (long long)F*N;
		if (!mesh->basis_table.size()) SetBasisTable(mesh, 1);
		s->left.Allocate(ctx, (size_t)mesh->nnzb*F*F);
		s->previous.Allocate(ctx, (size_t)mesh->nnzb*F*F);
		s->scaled_left.Allocate(ctx, (size_t)mesh->nnzb*F*F);
		s->inverse.Allocate(ctx, (size_t)N*F*F);
		for (auto* b : {&s->source, &s->state, &s->rhs, &s->next, &s->scaled, &s->values, &s->work})
			b->Allocate(ctx, n);
		s->constrained.Allocate(ctx, n);
		s->constrained.Upload(constrained, n);
		for (auto* b : {&s->state, &s->values}) Fill(ctx, b->data(), 0.0, n);
		s->velocity.Allocate(ctx, 3LL*N);
		s->velocity.Upload(velocity, 3LL*N);
		SetupSurfaces(s.get(), system);
		Assemble(s.get());
		s->krylov = std::make_unique<Fgmres>(ctx, n, s->restart);
		Sync(ctx);
		*out = s.release();
	});
}

int tfi_solve_set_velocity(tfi_solve* s, const double* velocity)
{
	return Guard(s->ctx, [&] {
		if (!velocity) throw Error("tfi_solve_set_velocity needs a velocity");
		s->velocity.Upload(velocity, 3LL*s->mesh->nodes);
		Assemble(s);
	});
}

void tfi_solve_destroy(tfi_solve* s)
{
	if (s && s->ctx) (void)hipStreamSynchronize(s->ctx->stream);
	delete s;
}

int tfi_solve_set_boundary(tfi_solve* s, const double* values)
{
	return Guard(s->ctx, [&] {
		s->values.Upload(values, (long long)s->fields*s->mesh->nodes);
		Sync(s->ctx);
	});
}

int tfi_solve_set_state(tfi_solve* s, const double* state)
{
	return Guard(s->ctx, [&] {
		s->state.Upload(state, (long long)s->fields*s->mesh->nodes);
		Sync(s->ctx);
	});
}

int tfi_solve_get_state(tfi_solve* s, double* state)
{
	return Guard(s->ctx, [&] { s->state.Download(state, (long long)s->fields*s->mesh->nodes); });
}

int tfi_solve_step(tfi_solve* s, int zero_guess, int maximum_iterations, double relative_tolerance,
	int* iterations, double* residual)
{
	return Guard(s->ctx, [&] {
		tfi_ctx* ctx = s->ctx;
		const long long n = (long long)s->fields*s->mesh->nodes;
		Spmv(s, s->previous.data(), s->state.data(), s->rhs.data());
		Axpy(ctx, 1.0, s->source.data(), s->rhs.data(), n);
		hipLaunchKernelGGL(SolveValuesKernel, dim3(Grid(n)), dim3(kThreads), 0, ctx->stream, n, s->constrained.data(),
			s->values.data(), s->rhs.data());
		TFI_LAUNCH_CHECK("SolveValuesKernel");
		if (zero_guess) Fill(ctx, s->next.data(), 0.0, n);
		else Copy(ctx, s->state.data(), s->next.data(), n);
		BlockApply(s, s->rhs.data(), s->scaled.data());
		Operator A = [&](const double* x, double* y) { Spmv(s, s->scaled_left.data(), x, y); };
		Precond M = [&](const double* v, double* z) { Copy(ctx, v, z, n); };
		if (s->preconditioner == TFI_PC_AMG) {
			if (s->amg_stale) SetupAmg(s);
			const int N = s->mesh->nodes;
			M = [&, N](const double* v, double* z) {
				Dispatch(s->fields, [&](auto fc) {
					constexpr int FF = decltype(fc)::value;
					if constexpr (FF <= 4) {
						hipLaunchKernelGGL(SolvePadVectorKernel<FF>, dim3(Grid(N)), dim3(kThreads), 0, ctx->stream, N,
							v, s->pad_in.data());
						TFI_LAUNCH_CHECK("SolvePadVectorKernel");
						s->amg->Apply(s->pad_in.data(), s->pad_out.data());
						hipLaunchKernelGGL(SolveUnpadVectorKernel<FF>, dim3(Grid(N)), dim3(kThreads), 0, ctx->stream,
							N, s->pad_out.data(), z);
						TFI_LAUNCH_CHECK("SolveUnpadVectorKernel");
					}
				});
			};
		}
		// the reference's GMRES stops at rtol times the initial preconditioned
		// residual (Gmres.hpp:105-124)
		A(s->next.data(), s->work.data());
		Axpby(ctx, 1.0, s->scaled.data(), -1.0, s->work.data(), n);
		const double initial = Norm2(ctx, s->work.data(), n);
		int its = 0;
		double rel = 0.0;
		if (Norm2(ctx, s->scaled.data(), n) == 0.0) {
			Fill(ctx, s->next.data(), 0.0, n);
		} else if (initial > 0.0) {
			const KrylovResult r = s->krylov->Solve(A, M, s->scaled.data(), s->next.data(), maximum_iterations, 0.0,
				relative_tolerance*initial);
			its = r.iterations;
			rel = r.residual/initial;
			if (!r.converged)
				throw Error("configured transport FGMRES did not converge (relative residual " + std::to_string(rel)
					+ ")");
		}
		if (iterations) *iterations = its;
		if (residual) *residual = rel;
		std::swap(s->state, s->next);
		Sync(ctx);
	});
}

int tfi_solve_matrices(tfi_solve* s, double* left, double* previous, double* source, int* singular_blocks)
{
	return Guard(s->ctx, [&] {
		if (left) s->left.Download(left, s->left.size());
		if (previous) s->previous.Download(previous, s->previous.size());
		if (source) s->source.Download(source, s->source.size());
		if (singular_blocks) *singular_blocks = s->singular;
	});
}

int tfi_solve_bench_spmv(tfi_solve* s, int repetitions, double* milliseconds)
{
	return Guard(s->ctx, [&] {
		tfi_ctx* ctx = s->ctx;
		for (int i = 0; i < 3; ++i) Spmv(s, s->left.data(), s->state.data(), s->work.data());
		Sync(ctx);
		Check(hipEventRecord(ctx->ev_begin, ctx->stream), "event");
		for (int i = 0; i < repetitions; ++i) Spmv(s, s->left.data(), s->state.data(), s->work.data());
		Check(hipEventRecord(ctx->ev_end, ctx->stream), "event");
		Check(hipEventSynchronize(ctx->ev_end), "event sync");
		float ms = 0.0f;
		Check(hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end), "elapsed");
		*milliseconds = ms/std::max(1, repetitions);
	});
}

int tfi_solve_set_preconditioner(tfi_solve* s, int preconditioner)
{
	return Guard(s->ctx, [&] {
		if (preconditioner != TFI_PC_AMG && preconditioner != TFI_PC_BJACOBI)
			throw Error("configured transport preconditioner must be TFI_PC_AMG or TFI_PC_BJACOBI");
		if (preconditioner == TFI_PC_AMG && s->fields > 4)
			throw Error("the AMG preconditioner of the configured transport supports at most 4 fields");
		s->preconditioner = preconditioner;
	});
}
