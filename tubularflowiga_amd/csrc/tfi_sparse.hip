// Node-block (4x4) sparse kernels: SpMV in FP64 and FP32 storage, residual,
// diagonal-block inverse and apply.
//
// The reference's BlockSpmvKernel<4> (solvers/cuda/include/SparseKernels.cuh:25-40)
// runs one thread per block row, so neighbouring threads read blocks ~16 KB
// apart.  Here one 64-lane wave owns a block row: lane (b, r) = (lane/4,
// lane%4) reads row r of block b (32 contiguous bytes), so a wave streams 16
// whole blocks (2 KB) per step with fully coalesced loads, and the 16 partial
// sums of each row are combined by a fixed xor-butterfly (deterministic).
// Block inverse: Gauss-Jordan with partial pivoting and the 1e-14 threshold /
// diagonal fallback of SparseKernels.cuh:42-104.
#include "tfi_internal.hpp"

#include <cmath>

namespace tfi {

namespace {

constexpr int kRowsPerBlock = 4;  // waves per workgroup

template <class V, class X, class Y, bool kResidual>
__global__ void __launch_bounds__(256) BsrSpmvKernel(int rows, const int* __restrict__ ro,
	const int* __restrict__ cols, const V* __restrict__ vals, const X* __restrict__ x, const double* __restrict__ b,
	Y* __restrict__ y)
{
	const int lane = threadIdx.x & 63;
	const int row = blockIdx.x*kRowsPerBlock + (threadIdx.x >> 6);
	if (row >= rows) return;
	const int sub = lane >> 2, r = lane & 3;
	const int begin = ro[row], end = ro[row + 1];
	double acc = 0.0;
	for (int k = begin + sub; k < end; k += 16) {
		const V* a = vals + (size_t)k*16 + 4*r;
		const X* xv = x + 4*(size_t)cols[k];
		acc += (double)a[0]*(double)xv[0] + (double)a[1]*(double)xv[1] + (double)a[2]*(double)xv[2]
			+ (double)a[3]*(double)xv[3];
	}
	acc += __shfl_xor(acc, 4, 64);
	acc += __shfl_xor(acc, 8, 64);
	acc += __shfl_xor(acc, 16, 64);
	acc += __shfl_xor(acc, 32, 64);
	if (sub == 0) {
		if (kResidual) acc = b[4*(size_t)row + r] - acc;
		y[4*(size_t)row + r] = static_cast<Y>(acc);
	}
}

__global__ void BlockInverseKernel(int rows, const int* __restrict__ diag, const double* __restrict__ vals,
	double* __restrict__ inv, unsigned int* __restrict__ singular)
{
	const int node = blockIdx.x*blockDim.x + threadIdx.x;
	if (node >= rows) return;
	const double* s = vals + (size_t)diag[node]*16;
	double m[4][8];
	for (int i = 0; i < 4; ++i)
		for (int j = 0; j < 8; ++j) m[i][j] = j < 4 ? s[4*i + j] : (j - 4 == i ? 1.0 : 0.0);
	bool failed = false;
	for (int p = 0; p < 4; ++p) {
		int best = p;
		for (int i = p + 1; i < 4; ++i)
			if (fabs(m[i][p]) > fabs(m[best][p])) best = i;
		if (!isfinite(m[best][p]) || fabs(m[best][p]) < 1e-14) {
			failed = true;
			break;
		}
		if (best != p)
			for (int j = 0; j < 8; ++j) {
				const double t = m[p][j];
				m[p][j] = m[best][j];
				m[best][j] = t;
			}
		const double sc = 1.0/m[p][p];
		for (int j = 0; j < 8; ++j) m[p][j] *= sc;
		for (int i = 0; i < 4; ++i) {
			if (i == p) continue;
			const double f = m[i][p];
			for (int j = 0; j < 8; ++j) m[i][j] -= f*m[p][j];
		}
	}
	double* d = inv + (size_t)node*16;
	if (!failed) {
		for (int i = 0; i < 4; ++i)
			for (int j = 0; j < 4; ++j) d[4*i + j] = m[i][4 + j];
	} else {
		atomicAdd(singular, 1u);
		for (int i = 0; i < 4; ++i)
			for (int j = 0; j < 4; ++j)
				d[4*i + j] = i == j && fabs(s[5*i]) > 1e-14 ? 1.0/s[5*i] : (i == j ? 1.0 : 0.0);
	}
}

__global__ void BlockApplyKernel(int rows, const double* __restrict__ inv, const double* __restrict__ x,
	double* __restrict__ y)
{
	const int node = blockIdx.x*blockDim.x + threadIdx.x;
	if (node >= rows) return;
	const double* d = inv + (size_t)node*16;
	const double* xv = x + 4*(size_t)node;
	const double x0 = xv[0], x1 = xv[1], x2 = xv[2], x3 = xv[3];
	for (int i = 0; i < 4; ++i) y[4*(size_t)node + i] = d[4*i]*x0 + d[4*i + 1]*x1 + d[4*i + 2]*x2 + d[4*i + 3]*x3;
}

__global__ void ToFloatBlocksKernel(long long n, const double* __restrict__ x, float* __restrict__ y)
{
	for (long long i = blockIdx.x*(long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x*blockDim.x)
		y[i] = static_cast<float>(x[i]);
}

int SpmvBlocks(int rows)
{
	return (rows + kRowsPerBlock - 1)/kRowsPerBlock;
}

}  // namespace

void Spmv64(tfi_ctx* ctx, const BsrView& a, const double* values, const double* x, double* y)
{
	hipLaunchKernelGGL((BsrSpmvKernel<double, double, double, false>), dim3(SpmvBlocks(a.rows)), dim3(256), 0,
		ctx->stream, a.rows, a.row_offsets, a.columns, values, x, nullptr, y);
	TFI_LAUNCH_CHECK("BsrSpmvKernel fp64");
}

void Spmv32(tfi_ctx* ctx, const BsrView& a, const float* values, const double* x, double* y)
{
	hipLaunchKernelGGL((BsrSpmvKernel<float, double, double, false>), dim3(SpmvBlocks(a.rows)), dim3(256), 0,
		ctx->stream, a.rows, a.row_offsets, a.columns, values, x, nullptr, y);
	TFI_LAUNCH_CHECK("BsrSpmvKernel fp32");
}

void Spmv32f(tfi_ctx* ctx, const BsrView& a, const float* values, const float* x, float* y)
{
	hipLaunchKernelGGL((BsrSpmvKernel<float, float, float, false>), dim3(SpmvBlocks(a.rows)), dim3(256), 0,
		ctx->stream, a.rows, a.row_offsets, a.columns, values, x, nullptr, y);
	TFI_LAUNCH_CHECK("BsrSpmvKernel fp32/fp32");
}

void Residual64(tfi_ctx* ctx, const BsrView& a, const double* values, const double* x, const double* b, double* r)
{
	hipLaunchKernelGGL((BsrSpmvKernel<double, double, double, true>), dim3(SpmvBlocks(a.rows)), dim3(256), 0,
		ctx->stream, a.rows, a.row_offsets, a.columns, values, x, b, r);
	TFI_LAUNCH_CHECK("BsrSpmvKernel residual");
}

int BlockInverse64(tfi_ctx* ctx, const BsrView& a, const double* values, double* inverse)
{
	DevBuf<unsigned int> singular(ctx, 1);
	Check(hipMemsetAsync(singular.data(), 0, sizeof(unsigned int), ctx->stream), "memset");
	hipLaunchKernelGGL(BlockInverseKernel, dim3((a.rows + 255)/256), dim3(256), 0, ctx->stream, a.rows, a.diagonal,
		values, inverse, singular.data());
	TFI_LAUNCH_CHECK("BlockInverseKernel");
	unsigned int s = 0;
	singular.Download(&s, 1);
	return static_cast<int>(s);
}

void BlockApply64(tfi_ctx* ctx, int rows, const double* inverse, const double* x, double* y)
{
	hipLaunchKernelGGL(BlockApplyKernel, dim3((rows + 255)/256), dim3(256), 0, ctx->stream, rows, inverse, x, y);
	TFI_LAUNCH_CHECK("BlockApplyKernel");
}

void BlockInverseTo32(tfi_ctx* ctx, int rows, const double* inverse, float* inverse32)
{
	const long long n = 16LL*rows;
	hipLaunchKernelGGL(ToFloatBlocksKernel, dim3(1024), dim3(256), 0, ctx->stream, n, inverse, inverse32);
	TFI_LAUNCH_CHECK("ToFloatBlocksKernel");
}

}  // namespace tfi
