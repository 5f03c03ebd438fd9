// IGA two-species neuron transport on the Navier-Stokes mesh pattern, with
// 2x2 node blocks: the bulk concentration N0 and the advected species Nplus
// (`iga_cuda transport`, solvers/cuda/src/iga_cuda.cu:621-733).
//
// Assembly (AssembleTransportKernel, solvers/cuda/include/IgaCudaKernels.cuh:155-237)
// builds, once per run, the left matrix L and the previous-state matrix P of
// backward Euler for a prescribed velocity u:
//   L00 = ((1 + dt(k+ + k-)) N_a N_b + dt D grad N_a . grad N_b) dV
//   L01 = -k'+ dt N_a N_b dV
//   L10 = -k+ dt w_a N_b dV
//   L11 = ((1 + dt k'+) w_a N_b + dt w_a u . grad N_b + dt v+ eps w_a grad N_a . grad N_b) dV
//   P00 = N_a N_b dV,  P11 = w_a N_b dV,  w_a = N_a + tau u . grad N_a (SUPG test function),
//   tau = (1/tau_s^2 + 1/tau_t^2)^-1/2, tau_s = 1/sum_a |u . grad N_a|, tau_t = dt/2.
// One workgroup per element stages the element's basis records (N, grad N) in
// LDS; one thread per (a, b) pair sums the 64 Gauss points and adds its 2x2
// blocks through the pair -> block map, one element colour per launch, so the
// sum order is fixed and no atomics are needed.  Constrained nodes get
// identity rows in L and zero rows in P (ApplyTransportBoundaryKernel,
// IgaCudaKernels.cuh:411-424).  Each step solves L c' = P c with the
// constrained values in the right-hand side (iga_cuda.cu:685-705) by FGMRES
// with a 2x2 block-Jacobi preconditioner.
#include "tfi_internal.hpp"
#include "tfi_solver.hpp"

#include <cmath>

struct tfi_transport {
	tfi_ctx* ctx = nullptr;
	tfi_mesh* mesh = nullptr;
	tfi_transport_params params{};
	tfi::DevBuf<double> left, previous, inverse, state, rhs, next, scaled, work, n0, nplus;
	tfi::DevBuf<int> constrained;
	std::unique_ptr<tfi::Fgmres> krylov;
	// D^-1 L (the solved operator) and, for the AMG preconditioner, its 2x2
	// blocks embedded in 4x4 blocks with an identity for the two padding fields
	tfi::DevBuf<double> scaled_left, padded64, pad_in, pad_out;
	tfi::DevBuf<float> padded32;
	std::unique_ptr<tfi::Amg> amg;
	int preconditioner = TFI_PC_BJACOBI;
	int restart = 50;
	int singular = 0;
};

namespace tfi {

namespace {

constexpr int kThreads = 256;
constexpr int kRec = 5;  // basis table record: N, dN/dx, dN/dy, dN/dz, lap N

struct TransportArgs {
	const int* eoff;
	const int* conn;
	const double* geometry;
	const double* table;
	const long long* pair_offsets;
	const int* pair_block;
	const int* elements;
	int count;
	const double* velocity;  // [3N]
	tfi_transport_params p;
	double* left;
	double* previous;
};

// kStaged: the element's basis records are copied to LDS as [q][a][N, grad N]
// (elements up to 64 basis functions); otherwise they are read from the table.
template <bool kStaged>
__global__ void __launch_bounds__(kThreads) TransportAssembleKernel(TransportArgs args)
{
	extern __shared__ double lds[];
	double* point = lds;                // [q][u (3), tau, dV]
	double* staged = lds + kQ*5;        // [q][a][4] when kStaged
	const int item = blockIdx.x;
	if (item >= args.count) return;
	const int e = args.elements[item];
	const int b0 = args.eoff[e];
	const int nen = args.eoff[e + 1] - b0;
	const double* tab = args.table + (size_t)b0*kQ*kRec;
	if (kStaged) {
		for (int t = threadIdx.x; t < kQ*nen; t += kThreads) {
			const double* r = tab + (size_t)t*kRec;
			for (int k = 0; k < 4; ++k) staged[4*t + k] = r[k];
		}
		__syncthreads();
	}
	const double* basis = kStaged ? staged : tab;
	constexpr int kStride = kStaged ? 4 : kRec;
	if (threadIdx.x < kQ) {
		const int q = threadIdx.x;
		const double* bq = basis + (size_t)kStride*q*nen;
		double u[3] = {0.0, 0.0, 0.0};
		for (int a = 0; a < nen; ++a) {
			const size_t node = args.conn[b0 + a];
			for (int d = 0; d < 3; ++d) u[d] += bq[kStride*a]*args.velocity[3*node + d];
		}
		double inverse_length = 0.0;
		for (int a = 0; a < nen; ++a)
			inverse_length += fabs(u[0]*bq[kStride*a + 1] + u[1]*bq[kStride*a + 2] + u[2]*bq[kStride*a + 3]);
		const double tau_space = inverse_length > 0.0 ? 1.0/inverse_length : 0.0;
		const double tau_time = args.p.dt/2.0;
		const double tau = tau_space > 0.0
			? 1.0/sqrt(1.0/(tau_space*tau_space) + 1.0/(tau_time*tau_time)) : 0.0;
		double* pt = point + 5*q;
		pt[0] = u[0];
		pt[1] = u[1];
		pt[2] = u[2];
		pt[3] = tau;
		pt[4] = args.geometry[(size_t)e*kGeomFields*kQ + q];
	}
	__syncthreads();
	const tfi_transport_params& P = args.p;
	const double dt = P.dt;
	const long long pbase = args.pair_offsets[e];
	for (int pair = threadIdx.x; pair < nen*nen; pair += kThreads) {
		const int a = pair/nen, b = pair - a*nen;
		double l00 = 0.0, l0p = 0.0, lp0 = 0.0, lpp = 0.0, p00 = 0.0, ppp = 0.0;
		for (int q = 0; q < kQ; ++q) {
			const double* ba = basis + (size_t)kStride*(q*nen + a);
			const double* bb = basis + (size_t)kStride*(q*nen + b);
			const double* pt = point + 5*q;
			const double adv_a = pt[0]*ba[1] + pt[1]*ba[2] + pt[2]*ba[3];
			const double adv_b = pt[0]*bb[1] + pt[1]*bb[2] + pt[2]*bb[3];
			const double test_plus = ba[0] + pt[3]*adv_a;
			const double gradient_dot = ba[1]*bb[1] + ba[2]*bb[2] + ba[3]*bb[3];
			const double measure = pt[4];
			const double mass = ba[0]*bb[0]*measure;
			l00 += ((1.0 + dt*(P.kplus + P.kminus))*ba[0]*bb[0] + dt*P.diffusion*gradient_dot)*measure;
			l0p += -P.detach_plus*dt*mass;
			lp0 += -P.kplus*dt*test_plus*bb[0]*measure;
			lpp += ((1.0 + dt*P.detach_plus)*test_plus*bb[0] + dt*test_plus*adv_b
				+ dt*P.vplus*P.artificial_diffusion*test_plus*gradient_dot)*measure;
			p00 += mass;
			ppp += test_plus*bb[0]*measure;
		}
		const size_t k = args.pair_block[pbase + pair];
		double* l = args.left + 4*k;
		double* p = args.previous + 4*k;
		l[0] += l00;
		l[1] += l0p;
		l[2] += lp0;
		l[3] += lpp;
		p[0] += p00;
		p[3] += ppp;
	}
}

__global__ void TransportBoundaryKernel(int nodes, const int* __restrict__ ro, const int* __restrict__ diag,
	const int* __restrict__ constrained, double* __restrict__ left, double* __restrict__ previous)
{
	const int node = blockIdx.x*blockDim.x + threadIdx.x;
	if (node >= nodes || !constrained[node]) return;
	for (int k = ro[node]; k < ro[node + 1]; ++k)
		for (int i = 0; i < 4; ++i) {
			left[4*(size_t)k + i] = 0.0;
			previous[4*(size_t)k + i] = 0.0;
		}
	double* d = left + 4*(size_t)diag[node];
	d[0] = 1.0;
	d[3] = 1.0;
}

__global__ void TransportValuesKernel(int nodes, const int* __restrict__ constrained, const double* __restrict__ n0,
	const double* __restrict__ nplus, double* __restrict__ v)
{
	const int node = blockIdx.x*blockDim.x + threadIdx.x;
	if (node < nodes && constrained[node]) {
		v[2*(size_t)node] = n0[node];
		v[2*(size_t)node + 1] = nplus[node];
	}
}

// y = A x with 2x2 blocks: kRowLanes lanes per block row (one row per
// wave; a row of the quadratic meshes has about 100 blocks), each lane over
// every kRowLanes-th block of its row, reduced within the lane group
constexpr int kRowLanes = 64;

__global__ void __launch_bounds__(kThreads) Bsr2SpmvKernel(int rows, const int* __restrict__ ro,
	const int* __restrict__ cols, const double* __restrict__ vals, const double* __restrict__ x, double* __restrict__ y)
{
	const int sub = threadIdx.x & (kRowLanes - 1);
	const int row = (blockIdx.x*kThreads + threadIdx.x)/kRowLanes;
	double y0 = 0.0, y1 = 0.0;
	if (row < rows) {
		for (int k = ro[row] + sub; k < ro[row + 1]; k += kRowLanes) {
			const double2* v = reinterpret_cast<const double2*>(vals + 4*(size_t)k);
			const double2 top = v[0], bottom = v[1];
			const double2 xc = reinterpret_cast<const double2*>(x)[cols[k]];
			y0 += top.x*xc.x + top.y*xc.y;
			y1 += bottom.x*xc.x + bottom.y*xc.y;
		}
	}
	for (int o = kRowLanes/2; o > 0; o >>= 1) {
		y0 += __shfl_xor(y0, o, 64);
		y1 += __shfl_xor(y1, o, 64);
	}
	if (row < rows && sub == 0)
		reinterpret_cast<double2*>(y)[row] = make_double2(y0, y1);
}

// 2x2 diagonal-block inverses; a block with |det| <= 1e-14 |a d| (or a
// non-finite determinant) falls back to the inverse of its diagonal and is counted
__global__ void Block2InverseKernel(int rows, const int* __restrict__ diag, const double* __restrict__ vals,
	double* __restrict__ inv, int* __restrict__ singular)
{
	const int row = blockIdx.x*blockDim.x + threadIdx.x;
	if (row >= rows) return;
	const double* d = vals + 4*(size_t)diag[row];
	double* o = inv + 4*(size_t)row;
	const double det = d[0]*d[3] - d[1]*d[2];
	if (isfinite(det) && fabs(det) > 1e-14*fabs(d[0]*d[3]) && det != 0.0) {
		o[0] = d[3]/det;
		o[1] = -d[1]/det;
		o[2] = -d[2]/det;
		o[3] = d[0]/det;
	} else {
		o[0] = d[0] != 0.0 ? 1.0/d[0] : 1.0;
		o[1] = 0.0;
		o[2] = 0.0;
		o[3] = d[3] != 0.0 ? 1.0/d[3] : 1.0;
		atomicAdd(singular, 1);
	}
}

// S = D^-1 L row by row (D_i^-1 times every block of row i)
__global__ void ScaleRowsKernel(int rows, const int* __restrict__ ro, const double* __restrict__ inv,
	const double* __restrict__ left, double* __restrict__ scaled)
{
	const int row = blockIdx.x*blockDim.x + threadIdx.x;
	if (row >= rows) return;
	const double* o = inv + 4*(size_t)row;
	for (int k = ro[row]; k < ro[row + 1]; ++k) {
		const double* a = left + 4*(size_t)k;
		double* s = scaled + 4*(size_t)k;
		s[0] = o[0]*a[0] + o[1]*a[2];
		s[1] = o[0]*a[1] + o[1]*a[3];
		s[2] = o[2]*a[0] + o[3]*a[2];
		s[3] = o[2]*a[1] + o[3]*a[3];
	}
}

// 4x4 embedding of the 2x2 blocks: [[S, 0], [0, I if diagonal block else 0]]
__global__ void PadBlocksKernel(int rows, const int* __restrict__ ro, const int* __restrict__ diag,
	const double* __restrict__ s, double* __restrict__ out)
{
	const int row = blockIdx.x*blockDim.x + threadIdx.x;
	if (row >= rows) return;
	for (int k = ro[row]; k < ro[row + 1]; ++k) {
		const double* a = s + 4*(size_t)k;
		double* o = out + 16*(size_t)k;
		const double one = k == diag[row] ? 1.0 : 0.0;
		for (int i = 0; i < 16; ++i) o[i] = 0.0;
		o[0] = a[0];
		o[1] = a[1];
		o[4] = a[2];
		o[5] = a[3];
		o[10] = one;
		o[15] = one;
	}
}

// [N0, Nplus] per node <-> [N0, Nplus, 0, 0] per node
__global__ void PadVectorKernel(int rows, const double* __restrict__ v, double* __restrict__ p)
{
	const int row = blockIdx.x*blockDim.x + threadIdx.x;
	if (row >= rows) return;
	p[4*(size_t)row] = v[2*(size_t)row];
	p[4*(size_t)row + 1] = v[2*(size_t)row + 1];
	p[4*(size_t)row + 2] = 0.0;
	p[4*(size_t)row + 3] = 0.0;
}

__global__ void UnpadVectorKernel(int rows, const double* __restrict__ p, double* __restrict__ v)
{
	const int row = blockIdx.x*blockDim.x + threadIdx.x;
	if (row >= rows) return;
	v[2*(size_t)row] = p[4*(size_t)row];
	v[2*(size_t)row + 1] = p[4*(size_t)row + 1];
}

__global__ void Block2ApplyKernel(int rows, const double* __restrict__ inv, const double* __restrict__ v,
	double* __restrict__ z)
{
	const int row = blockIdx.x*blockDim.x + threadIdx.x;
	if (row >= rows) return;
	const double* o = inv + 4*(size_t)row;
	const double v0 = v[2*(size_t)row], v1 = v[2*(size_t)row + 1];
	z[2*(size_t)row] = o[0]*v0 + o[1]*v1;
	z[2*(size_t)row + 1] = o[2]*v0 + o[3]*v1;
}

int Grid(long long n) { return static_cast<int>((n + kThreads - 1)/kThreads); }

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

// algorithmic bytes of one 2x2 SpMV: values and column indices once, row
// offsets, x and y once each
double Spmv2Bytes(const BsrView& a)
{
	return static_cast<double>(a.nnzb)*(4*8 + 4) + (a.rows + 1)*4.0 + 2.0*a.rows*2*8;
}

void Spmv2(tfi_transport* t, const double* vals, const double* x, double* y)
{
	tfi_ctx* ctx = t->ctx;
	const BsrView a = t->mesh->view();
	StreamTimer timer(ctx, &ctx->timers.spmv2_ms);
	hipLaunchKernelGGL(Bsr2SpmvKernel, dim3((a.rows + kThreads/kRowLanes - 1)/(kThreads/kRowLanes)), dim3(kThreads), 0,
		ctx->stream, a.rows,
		a.row_offsets, a.columns, vals, x, y);
	TFI_LAUNCH_CHECK("Bsr2SpmvKernel");
	ctx->timers.spmv2_calls += 1;
	ctx->timers.spmv2_bytes += Spmv2Bytes(a);
}

}  // namespace

}  // namespace tfi

using namespace tfi;

int tfi_transport_create(tfi_mesh* mesh, const tfi_transport_params* params, const double* velocity,
	const int32_t* constrained, tfi_transport** out)
{
	*out = nullptr;
	return Guard(mesh->ctx, [&] {
		if (!params || !velocity || !constrained) throw Error("tfi_transport_create needs parameters, velocity and mask");
		if (!(params->dt > 0.0)) throw Error("transport dt must be positive");
		auto t = std::make_unique<tfi_transport>();
		t->ctx = mesh->ctx;
		t->mesh = mesh;
		t->params = *params;
		tfi_ctx* ctx = mesh->ctx;
		const int N = mesh->nodes;
		if (mesh->max_nen > kMaxBasis) throw Error("transport assembly supports at most 80 basis functions");
		if (!mesh->basis_table.size()) SetBasisTable(mesh, 1);
		t->left.Allocate(ctx, (size_t)mesh->nnzb*4);
		t->previous.Allocate(ctx, (size_t)mesh->nnzb*4);
		t->inverse.Allocate(ctx, (size_t)N*4);
		for (auto* b : {&t->state, &t->rhs, &t->next, &t->scaled, &t->work}) b->Allocate(ctx, 2LL*N);
		t->n0.Allocate(ctx, N);
		t->nplus.Allocate(ctx, N);
		t->constrained.Allocate(ctx, N);
		t->constrained.Upload(constrained, N);
		Fill(ctx, t->left.data(), 0.0, t->left.size());
		Fill(ctx, t->previous.data(), 0.0, t->previous.size());
		Fill(ctx, t->state.data(), 0.0, t->state.size());
		Fill(ctx, t->n0.data(), 0.0, N);
		Fill(ctx, t->nplus.data(), 0.0, N);
		DevBuf<double> vel;
		vel.Allocate(ctx, 3LL*N);
		vel.Upload(velocity, 3LL*N);
		TransportArgs a{};
		a.eoff = mesh->eoff.data();
		a.conn = mesh->conn.data();
		a.geometry = mesh->geometry.data();
		a.table = mesh->basis_table.data();
		a.pair_offsets = mesh->pair_offsets.data();
		a.pair_block = mesh->pair_block.data();
		a.velocity = vel.data();
		a.p = *params;
		a.left = t->left.data();
		a.previous = t->previous.data();
		const bool staged = mesh->max_nen <= 64;
		const size_t lds = sizeof(double)*(kQ*5 + (staged ? (size_t)kQ*mesh->max_nen*4 : 0));
		if (lds > 65536)
			Check(hipFuncSetAttribute(reinterpret_cast<const void*>(&TransportAssembleKernel<true>),
				hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)), "hipFuncSetAttribute");
		{
		StreamTimer timer(ctx, &ctx->timers.transport_assembly_ms);
		for (int c = 0; c < mesh->colors; ++c) {
			a.elements = mesh->color_elements.data() + mesh->color_offsets[c];
			a.count = mesh->color_offsets[c + 1] - mesh->color_offsets[c];
			if (a.count == 0) continue;
			if (staged)
				hipLaunchKernelGGL(TransportAssembleKernel<true>, dim3(a.count), dim3(kThreads), lds, ctx->stream, a);
			else
				hipLaunchKernelGGL(TransportAssembleKernel<false>, dim3(a.count), dim3(kThreads), lds, ctx->stream, a);
			TFI_LAUNCH_CHECK("TransportAssembleKernel");
		}
		}
		const BsrView v = mesh->view();
		hipLaunchKernelGGL(TransportBoundaryKernel, dim3(Grid(N)), dim3(kThreads), 0, ctx->stream, N, v.row_offsets,
			v.diagonal, t->constrained.data(), t->left.data(), t->previous.data());
		TFI_LAUNCH_CHECK("TransportBoundaryKernel");
		DevBuf<int> singular;
		singular.Allocate(ctx, 1);
		FillInt(ctx, singular.data(), 0, 1);
		hipLaunchKernelGGL(Block2InverseKernel, dim3(Grid(N)), dim3(kThreads), 0, ctx->stream, N, v.diagonal,
			t->left.data(), t->inverse.data(), singular.data());
		TFI_LAUNCH_CHECK("Block2InverseKernel");
		singular.Download(&t->singular, 1);
		t->scaled_left.Allocate(ctx, (size_t)mesh->nnzb*4);
		hipLaunchKernelGGL(ScaleRowsKernel, dim3(Grid(N)), dim3(kThreads), 0, ctx->stream, N, v.row_offsets,
			t->inverse.data(), t->left.data(), t->scaled_left.data());
		TFI_LAUNCH_CHECK("ScaleRowsKernel");
		t->krylov = std::make_unique<Fgmres>(ctx, 2LL*N, t->restart);
		Sync(ctx);
		*out = t.release();
	});
}

void tfi_transport_destroy(tfi_transport* t)
{
	if (t && t->ctx) (void)hipStreamSynchronize(t->ctx->stream);
	delete t;
}

int tfi_transport_set_preconditioner(tfi_transport* t, int preconditioner)
{
	return Guard(t->ctx, [&] {
		if (preconditioner != TFI_PC_AMG && preconditioner != TFI_PC_BJACOBI)
			throw Error("transport preconditioner must be TFI_PC_AMG or TFI_PC_BJACOBI");
		t->preconditioner = preconditioner;
	});
}

int tfi_transport_set_boundary(tfi_transport* t, const double* n0, const double* nplus)
{
	return Guard(t->ctx, [&] {
		t->n0.Upload(n0, t->mesh->nodes);
		t->nplus.Upload(nplus, t->mesh->nodes);
		Sync(t->ctx);
	});
}

int tfi_transport_set_state(tfi_transport* t, const double* state)
{
	return Guard(t->ctx, [&] {
		t->state.Upload(state, 2LL*t->mesh->nodes);
		Sync(t->ctx);
	});
}

int tfi_transport_get_state(tfi_transport* t, double* state)
{
	return Guard(t->ctx, [&] { t->state.Download(state, 2LL*t->mesh->nodes); });
}

int tfi_transport_apply_boundary(tfi_transport* t)
{
	return Guard(t->ctx, [&] {
		const int N = t->mesh->nodes;
		hipLaunchKernelGGL(TransportValuesKernel, dim3(Grid(N)), dim3(kThreads), 0, t->ctx->stream, N,
			t->constrained.data(), t->n0.data(), t->nplus.data(), t->state.data());
		TFI_LAUNCH_CHECK("TransportValuesKernel");
		Sync(t->ctx);
	});
}

int tfi_transport_step(tfi_transport* t, int zero_guess, int maximum_iterations, double relative_tolerance,
	int* iterations, double* residual)
{
	return Guard(t->ctx, [&] {
		tfi_ctx* ctx = t->ctx;
		const int N = t->mesh->nodes;
		const long long n = 2LL*N;
		Spmv2(t, t->previous.data(), t->state.data(), t->rhs.data());
		hipLaunchKernelGGL(TransportValuesKernel, dim3(Grid(N)), dim3(kThreads), 0, ctx->stream, N,
			t->constrained.data(), t->n0.data(), t->nplus.data(), t->rhs.data());
		TFI_LAUNCH_CHECK("TransportValuesKernel");
		if (zero_guess) Fill(ctx, t->next.data(), 0.0, n);
		else Copy(ctx, t->state.data(), t->next.data(), n);
		// Block-Jacobi scaled system D^-1 L c' = D^-1 b, as the reference's
		// left-preconditioned GMRES measures its residual (Gmres.hpp:69-213):
		// the identity rows of the constrained nodes and the mass-scaled rows
		// of the others then weigh alike in the stopping test.
		auto block_apply = [&](const double* v, double* z) {
			hipLaunchKernelGGL(Block2ApplyKernel, dim3(Grid(N)), dim3(kThreads), 0, ctx->stream, N, t->inverse.data(),
				v, z);
			TFI_LAUNCH_CHECK("Block2ApplyKernel");
		};
		block_apply(t->rhs.data(), t->scaled.data());
		Operator A = [&](const double* x, double* y) { Spmv2(t, t->scaled_left.data(), x, y); };
		Precond M = [&](const double* v, double* z) { Copy(ctx, v, z, n); };
		if (t->preconditioner == TFI_PC_AMG) {
			if (!t->amg) {
				// the operator is fixed for the run: one AMG set-up (FP32 cycles)
				const BsrView a = t->mesh->view();
				t->padded64.Allocate(ctx, (size_t)a.nnzb*16);
				t->padded32.Allocate(ctx, (size_t)a.nnzb*16);
				t->pad_in.Allocate(ctx, 4LL*N);
				t->pad_out.Allocate(ctx, 4LL*N);
				hipLaunchKernelGGL(PadBlocksKernel, dim3(Grid(N)), dim3(kThreads), 0, ctx->stream, N, a.row_offsets,
					a.diagonal, t->scaled_left.data(), t->padded64.data());
				TFI_LAUNCH_CHECK("PadBlocksKernel");
				ToFloat(ctx, t->padded64.data(), t->padded32.data(), (long long)a.nnzb*16);
				// symmetric Gauss-Seidel (as the flow solver's); on the 0.6M-DOF Y
				// 22-26 iterations per step against 120-145 with block Jacobi alone
				// (damped-Jacobi smoothing: 51; one sweep pre and post stagnates),
				// but at about 2 ms per V-cycle a step is not faster than block
				// Jacobi's at 0.3 ms per iteration, hence not the default
				t->amg = std::make_unique<Amg>(t->mesh, TFI_MIXED, AmgOptions());
				t->amg->Setup(t->padded64.data(), t->padded32.data());
			}
			M = [&](const double* v, double* z) {
				hipLaunchKernelGGL(PadVectorKernel, dim3(Grid(N)), dim3(kThreads), 0, ctx->stream, N, v,
					t->pad_in.data());
				TFI_LAUNCH_CHECK("PadVectorKernel");
				t->amg->Apply(t->pad_in.data(), t->pad_out.data());
				hipLaunchKernelGGL(UnpadVectorKernel, dim3(Grid(N)), dim3(kThreads), 0, ctx->stream, N,
					t->pad_out.data(), z);
				TFI_LAUNCH_CHECK("UnpadVectorKernel");
			};
		}
		const KrylovResult r = t->krylov->Solve(A, M, t->scaled.data(), t->next.data(), maximum_iterations,
			relative_tolerance);
		if (iterations) *iterations = r.iterations;
		if (residual) *residual = r.rhs_norm > 0.0 ? r.residual/r.rhs_norm : r.residual;
		if (!r.converged)
			throw Error("transport FGMRES did not converge (relative residual "
				+ std::to_string(r.rhs_norm > 0.0 ? r.residual/r.rhs_norm : r.residual) + ")");
		std::swap(t->state, t->next);
		Sync(ctx);
	});
}

int tfi_transport_bench_spmv(tfi_transport* t, int repetitions, double* milliseconds)
{
	return Guard(t->ctx, [&] {
		tfi_ctx* ctx = t->ctx;
		const BsrView a = t->mesh->view();
		auto run = [&] {
			hipLaunchKernelGGL(Bsr2SpmvKernel, dim3((a.rows + kThreads/kRowLanes - 1)/(kThreads/kRowLanes)),
				dim3(kThreads), 0, ctx->stream, a.rows, a.row_offsets, a.columns, t->left.data(), t->state.data(),
				t->work.data());
			TFI_LAUNCH_CHECK("Bsr2SpmvKernel");
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

int tfi_transport_matrices(tfi_transport* t, double* left, double* previous, int* singular_blocks)
{
	return Guard(t->ctx, [&] {
		if (left) t->left.Download(left, t->left.size());
		if (previous) t->previous.Download(previous, t->previous.size());
		if (singular_blocks) *singular_blocks = t->singular;
	});
}
