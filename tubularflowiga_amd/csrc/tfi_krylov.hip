// Flexible GMRES(m), right-preconditioned, with classical Gram-Schmidt applied
// twice (CGS2) through fused multi-vector kernels.
//
// Replaces SolveGmres<4> (solvers/cuda/include/Gmres.hpp:69-213), whose
// modified Gram-Schmidt issues j+1 cuBLAS Ddot/Daxpy pairs per iteration, each
// Ddot a host round trip.  Here one Arnoldi step is: preconditioner, SpMV,
// two passes of {multi-dot over all basis vectors, multi-axpy} (each pass reads
// the basis once from HBM), a fused norm, and one small device->host copy of
// the Hessenberg column for the Givens update on the host.  All reductions
// have a fixed order, so a solve is bitwise reproducible.  The stopping test is
// the true relative residual ||b - A x|| <= rtol ||b|| (PETSc FGMRES with a zero
// guess, TransientFlowRuntime.hpp:176-190), re-checked at every restart.
#include "tfi_internal.hpp"

#include <algorithm>
#include <cmath>

namespace tfi {

namespace {

constexpr int kThreads = 256;
constexpr int kPerThread = 8;  // vector entries per thread: eight 8-byte loads per basis vector in flight
constexpr int kGroup = 8;      // basis vectors per multi-dot pass (one set of wave reductions per pass)

__device__ inline double WaveSum(double v)
{
	for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
	return v;
}

// partial[i*blocks + blk] = sum over this block's rows of V_i . w, i < count.
__global__ void __launch_bounds__(kThreads) MultiDotKernel(const double* __restrict__ V, long long ld, int count,
	const double* __restrict__ w, long long n, double* __restrict__ partial)
{
	__shared__ double lds[kGroup][4];
	const long long base = (long long)blockIdx.x*kThreads*kPerThread;
	double wr[kPerThread];
#pragma unroll
	for (int t = 0; t < kPerThread; ++t) {
		const long long i = base + (long long)t*kThreads + threadIdx.x;
		wr[t] = i < n ? w[i] : 0.0;
	}
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	for (int g = 0; g < count; g += kGroup) {
		double acc[kGroup] = {};
#pragma unroll
		for (int v = 0; v < kGroup; ++v) {
			if (g + v >= count) break;
			const double* Vi = V + (long long)(g + v)*ld;
#pragma unroll
			for (int t = 0; t < kPerThread; ++t) {
				const long long i = base + (long long)t*kThreads + threadIdx.x;
				if (i < n) acc[v] += Vi[i]*wr[t];
			}
		}
#pragma unroll
		for (int v = 0; v < kGroup; ++v) {
			const double s = WaveSum(acc[v]);
			if (lane == 0) lds[v][wave] = s;
		}
		__syncthreads();
		if (threadIdx.x < kGroup && g + (int)threadIdx.x < count) {
			const int v = threadIdx.x;
			partial[(long long)(g + v)*gridDim.x + blockIdx.x] = lds[v][0] + lds[v][1] + lds[v][2] + lds[v][3];
		}
		__syncthreads();
	}
}

// out[i] = (accumulate ? out[i] : 0) + sum_b partial[i*blocks + b]   (one workgroup per i)
__global__ void ReducePartialsKernel(const double* __restrict__ partial, int blocks, double* __restrict__ out,
	int accumulate)
{
	__shared__ double lds[4];
	const int i = blockIdx.x;
	double s = 0.0;
	for (int b = threadIdx.x; b < blocks; b += blockDim.x) s += partial[(long long)i*blocks + b];
	s = WaveSum(s);
	if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
	__syncthreads();
	if (threadIdx.x == 0) {
		const double t = lds[0] + lds[1] + lds[2] + lds[3];
		out[i] = accumulate ? out[i] + t : t;
	}
}

// w += sign * sum_i coef[i] V_i ; optional per-block partial of ||w||^2.
__global__ void __launch_bounds__(kThreads) MultiAxpyKernel(const double* __restrict__ V, long long ld, int count,
	const double* __restrict__ coef, double sign, double* __restrict__ w, long long n, double* __restrict__ norm_partial)
{
	__shared__ double c[512];
	__shared__ double lds[4];
	for (int i = threadIdx.x; i < count; i += kThreads) c[i] = sign*coef[i];
	__syncthreads();
	const long long base = (long long)blockIdx.x*kThreads*kPerThread + threadIdx.x;
	// each entry accumulates over the vectors in order; the thread's entries
	// are independent, so their loads of one vector are in flight together
	double acc[kPerThread];
#pragma unroll
	for (int t = 0; t < kPerThread; ++t) {
		const long long i = base + (long long)t*kThreads;
		acc[t] = i < n ? w[i] : 0.0;
	}
	for (int v = 0; v < count; ++v) {
		const double cv = c[v];
		const double* Vv = V + (long long)v*ld;
#pragma unroll
		for (int t = 0; t < kPerThread; ++t) {
			const long long i = base + (long long)t*kThreads;
			if (i < n) acc[t] += cv*Vv[i];
		}
	}
	double sq = 0.0;
#pragma unroll
	for (int t = 0; t < kPerThread; ++t) {
		const long long i = base + (long long)t*kThreads;
		if (i < n) {
			w[i] = acc[t];
			sq += acc[t]*acc[t];
		}
	}
	if (norm_partial) {
		sq = WaveSum(sq);
		if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = sq;
		__syncthreads();
		if (threadIdx.x == 0) norm_partial[blockIdx.x] = lds[0] + lds[1] + lds[2] + lds[3];
	}
}

// dst = src / norm where norm = sqrt(*sq) (device scalar); writes the norm to *out_norm.
__global__ void NormalizeKernel(const double* __restrict__ src, double* __restrict__ dst, long long n,
	const double* __restrict__ sq)
{
	const double nrm = sqrt(*sq);
	const double s = nrm > 1e-300 ? 1.0/nrm : 0.0;
	for (long long i = blockIdx.x*(long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x*blockDim.x)
		dst[i] = src[i]*s;
}

}  // namespace

Fgmres::Fgmres(tfi_ctx* ctx, long long n, int restart) : ctx_(ctx), n_(n), m_(restart)
{
	if (restart < 1 || restart > 500) throw Error("FGMRES restart must be in [1, 500]");
	V_.Allocate(ctx, (size_t)(m_ + 1)*n);
	Z_.Allocate(ctx, (size_t)m_*n);
	w_.Allocate(ctx, n);
	r_.Allocate(ctx, n);
	const long long blocks = (n + kThreads*kPerThread - 1)/(kThreads*kPerThread);
	partial_.Allocate(ctx, (size_t)(m_ + 2)*blocks + 8);
	h_.Allocate(ctx, 2*(m_ + 2) + 8);
	H_.assign((size_t)(m_ + 1)*m_, 0.0);
	cs_.assign(m_, 0.0);
	sn_.assign(m_, 0.0);
	g_.assign(m_ + 1, 0.0);
	y_.assign(m_, 0.0);
}

KrylovResult Fgmres::Solve(const Operator& A, const Precond& M, const double* b, double* x, int maxit, double rtol,
	double absolute_target)
{
	tfi_ctx* ctx = ctx_;
	const long long n = n_;
	const int blocks = static_cast<int>((n + kThreads*kPerThread - 1)/(kThreads*kPerThread));
	const int eblocks = static_cast<int>(std::min<long long>((n + 255)/256, 4096));
	KrylovResult res;
	res.rhs_norm = Norm2(ctx, b, n);
	const double tol = std::max(rtol*res.rhs_norm, absolute_target);
	double* V = V_.data();
	double* Z = Z_.data();
	double* w = w_.data();
	double* h = h_.data();            // h[0..m]: projection coefficients; h[m+1]: squared norm
	double* hsq = h + (m_ + 1);
	double* partial = partial_.data();
	auto true_residual = [&]() {
		A(x, w);
		// r = b - w
		Copy(ctx, b, r_.data(), n);
		Axpy(ctx, -1.0, w, r_.data(), n);
		return Norm2(ctx, r_.data(), n);
	};
	double beta = true_residual();
	res.residual = beta;
	if (res.rhs_norm == 0.0 || beta <= tol) {
		res.converged = true;
		return res;
	}
	double* hcol = ctx->host_pinned;  // pinned staging
	const double beta0 = beta;
	while (res.iterations < maxit) {
		const double beta_cycle = beta;
		Copy(ctx, r_.data(), V, n);
		Scale(ctx, 1.0/beta, V, n);
		std::fill(H_.begin(), H_.end(), 0.0);
		std::fill(g_.begin(), g_.end(), 0.0);
		g_[0] = beta;
		int used = 0;
		bool breakdown = false;
		for (int j = 0; j < m_ && res.iterations < maxit; ++j) {
			double* vj = V + (long long)j*n;
			double* zj = Z + (long long)j*n;
			M(vj, zj);
			A(zj, w);
			// CGS pass 1: h = V^T w ; w -= V h
			hipLaunchKernelGGL(MultiDotKernel, dim3(blocks), dim3(kThreads), 0, ctx->stream, V, n, j + 1, w, n, partial);
			TFI_LAUNCH_CHECK("MultiDotKernel");
			hipLaunchKernelGGL(ReducePartialsKernel, dim3(j + 1), dim3(256), 0, ctx->stream, partial, blocks, h, 0);
			TFI_LAUNCH_CHECK("ReducePartialsKernel");
			hipLaunchKernelGGL(MultiAxpyKernel, dim3(blocks), dim3(kThreads), 0, ctx->stream, V, n, j + 1, h, -1.0, w, n,
				nullptr);
			TFI_LAUNCH_CHECK("MultiAxpyKernel");
			// CGS pass 2 (re-orthogonalisation), fused with the norm of w
			hipLaunchKernelGGL(MultiDotKernel, dim3(blocks), dim3(kThreads), 0, ctx->stream, V, n, j + 1, w, n, partial);
			TFI_LAUNCH_CHECK("MultiDotKernel");
			double* h2 = h + (m_ + 2);  // second-pass coefficients
			hipLaunchKernelGGL(ReducePartialsKernel, dim3(j + 1), dim3(256), 0, ctx->stream, partial, blocks, h2, 0);
			TFI_LAUNCH_CHECK("ReducePartialsKernel");
			double* npart = partial + (long long)(m_ + 1)*blocks;
			hipLaunchKernelGGL(MultiAxpyKernel, dim3(blocks), dim3(kThreads), 0, ctx->stream, V, n, j + 1, h2, -1.0, w,
				n, npart);
			TFI_LAUNCH_CHECK("MultiAxpyKernel");
			hipLaunchKernelGGL(ReducePartialsKernel, dim3(1), dim3(256), 0, ctx->stream, npart, blocks, hsq, 0);
			TFI_LAUNCH_CHECK("ReducePartialsKernel");
			hipLaunchKernelGGL(NormalizeKernel, dim3(eblocks), dim3(256), 0, ctx->stream, w, V + (long long)(j + 1)*n, n,
				hsq);
			TFI_LAUNCH_CHECK("NormalizeKernel");
			// bring back h (pass 1 + pass 2) and the squared norm
			Check(hipMemcpyAsync(hcol, h, sizeof(double)*(2*(m_ + 2)), hipMemcpyDeviceToHost, ctx->stream), "h download");
			Check(hipStreamSynchronize(ctx->stream), "arnoldi sync");
			for (int i = 0; i <= j; ++i) H_[(size_t)i*m_ + j] = hcol[i] + hcol[m_ + 2 + i];
			const double hn = std::sqrt(hcol[m_ + 1]);
			H_[(size_t)(j + 1)*m_ + j] = hn;
			for (int i = 0; i < j; ++i) {
				const double h0 = H_[(size_t)i*m_ + j], h1 = H_[(size_t)(i + 1)*m_ + j];
				H_[(size_t)i*m_ + j] = cs_[i]*h0 + sn_[i]*h1;
				H_[(size_t)(i + 1)*m_ + j] = -sn_[i]*h0 + cs_[i]*h1;
			}
			const double h0 = H_[(size_t)j*m_ + j], h1 = H_[(size_t)(j + 1)*m_ + j];
			const double mag = std::hypot(h0, h1);
			cs_[j] = mag > 0.0 ? h0/mag : 1.0;
			sn_[j] = mag > 0.0 ? h1/mag : 0.0;
			H_[(size_t)j*m_ + j] = mag;
			H_[(size_t)(j + 1)*m_ + j] = 0.0;
			const double gj = g_[j];
			g_[j] = cs_[j]*gj;
			g_[j + 1] = -sn_[j]*gj;
			++res.iterations;
			used = j + 1;
			res.residual = std::abs(g_[j + 1]);
			if (res.residual <= tol) break;
			if (hn <= 1e-300) {
				breakdown = true;
				break;
			}
		}
		for (int row = used - 1; row >= 0; --row) {
			double v = g_[row];
			for (int c = row + 1; c < used; ++c) v -= H_[(size_t)row*m_ + c]*y_[c];
			const double d = H_[(size_t)row*m_ + row];
			if (std::abs(d) < 1e-300) throw Error("FGMRES Hessenberg matrix is singular");
			y_[row] = v/d;
		}
		Check(hipMemcpyAsync(h, y_.data(), sizeof(double)*used, hipMemcpyHostToDevice, ctx->stream), "y upload");
		hipLaunchKernelGGL(MultiAxpyKernel, dim3(blocks), dim3(kThreads), 0, ctx->stream, Z, n, used, h, 1.0, x, n,
			nullptr);
		TFI_LAUNCH_CHECK("MultiAxpyKernel update");
		beta = true_residual();
		res.residual = beta;
		if (beta <= tol) {
			res.converged = true;
			break;
		}
		if (breakdown) break;
		// a restart cycle that made no progress, or a residual far above the
		// initial one, means the preconditioned operator defeats the method:
		// stop early instead of spending the remaining iterations
		if (!std::isfinite(beta) || beta >= beta_cycle || beta > 10.0*beta0) break;
	}
	return res;
}

}  // namespace tfi
