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
#include <mutex>

// Working memory of tfi_block_inverse_noalloc: one singular-block count per
// 256-row workgroup (at most kNoAllocGroups of them, so meshes of up to
// 16,777,216 nodes), then the total at kNoAllocGroups.  Every GPU the library
// runs on holds its own copy (256 KiB) until the process ends.  Host code
// must not pass this name to a kernel: it is a host-side stand-in, and the
// device address of each GPU's copy comes from hipGetSymbolAddress.
constexpr int kNoAllocGroups = 65536;
static __device__ __align__(256) unsigned int scratch[kNoAllocGroups + 64];

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

// FP32 block values: a lane holds two rows of its block (32 bytes of values,
// as a lane of the FP64 kernel holds one row), so that a wave keeps as many
// value bytes in flight per pass as in FP64: 32 blocks per pass, lanes (2s,
// 2s+1) rows (0,1) and (2,3) of block s.
template <class X, class Y>
__global__ void __launch_bounds__(256) BsrSpmvF32Kernel(int rows, const int* __restrict__ ro,
	const int* __restrict__ cols, const float* __restrict__ vals, const X* __restrict__ x, Y* __restrict__ y)
{
	typedef float F4 __attribute__((ext_vector_type(4)));
	const int lane = threadIdx.x & 63;
	const int row = blockIdx.x*kRowsPerBlock + (threadIdx.x >> 6);
	if (row >= rows) return;
	const int sub = lane >> 1, h = lane & 1;
	const int begin = ro[row], end = ro[row + 1];
	double acc0 = 0.0, acc1 = 0.0;
	for (int k = begin + sub; k < end; k += 32) {
		const F4* a = reinterpret_cast<const F4*>(vals + (size_t)k*16 + 8*h);
		const F4 r0 = a[0], r1 = a[1];
		const X* xv = x + 4*(size_t)cols[k];
		const double x0 = xv[0], x1 = xv[1], x2 = xv[2], x3 = xv[3];
		acc0 += (double)r0.x*x0 + (double)r0.y*x1 + (double)r0.z*x2 + (double)r0.w*x3;
		acc1 += (double)r1.x*x0 + (double)r1.y*x1 + (double)r1.z*x2 + (double)r1.w*x3;
	}
#pragma unroll
	for (int o = 2; o < 64; o <<= 1) {
		acc0 += __shfl_xor(acc0, o, 64);
		acc1 += __shfl_xor(acc1, o, 64);
	}
	if (sub == 0) {
		y[4*(size_t)row + 2*h] = static_cast<Y>(acc0);
		y[4*(size_t)row + 2*h + 1] = static_cast<Y>(acc1);
	}
}

// Inverse of the diagonal block of `node` into inv[16 node]; returns true when
// a pivot fell below 1e-14 and the diagonal fallback was taken.
__device__ inline bool InvertDiagonalBlock(int node, const int* __restrict__ diag, const double* __restrict__ vals,
	double* __restrict__ inv)
{
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
		for (int i = 0; i < 4; ++i)
			for (int j = 0; j < 4; ++j)
				d[4*i + j] = i == j && fabs(s[5*i]) > 1e-14 ? 1.0/s[5*i] : (i == j ? 1.0 : 0.0);
	}
	return failed;
}

__global__ void BlockInverseKernel(int rows, const int* __restrict__ diag, const double* __restrict__ vals,
	double* __restrict__ inv, unsigned int* __restrict__ singular)
{
	const int node = blockIdx.x*blockDim.x + threadIdx.x;
	if (node >= rows) return;
	if (InvertDiagonalBlock(node, diag, vals, inv)) atomicAdd(singular, 1u);
}

// Same inverse; the singular-block count of each 256-row workgroup goes to
// partial[blockIdx.x] (every slot below gridDim.x is written, so nothing
// needs clearing beforehand).
__global__ void __launch_bounds__(256) BlockInverseCountKernel(int rows, const int* __restrict__ diag,
	const double* __restrict__ vals, double* __restrict__ inv, unsigned int* __restrict__ partial)
{
	const int node = blockIdx.x*blockDim.x + threadIdx.x;
	const bool failed = node < rows && InvertDiagonalBlock(node, diag, vals, inv);
	const int count = __syncthreads_count(failed);
	if (threadIdx.x == 0) partial[blockIdx.x] = static_cast<unsigned int>(count);
}

// total[0] = sum of partial[0 .. count), one workgroup of 256 threads.
__global__ void __launch_bounds__(256) CountSumKernel(const unsigned int* __restrict__ partial, int count,
	unsigned int* __restrict__ total)
{
	__shared__ unsigned int lds[256];
	unsigned int s = 0;
	for (int i = threadIdx.x; i < count; i += 256) s += partial[i];
	lds[threadIdx.x] = s;
	__syncthreads();
	for (int half = 128; half > 0; half >>= 1) {
		if (threadIdx.x < half) lds[threadIdx.x] += lds[threadIdx.x + half];
		__syncthreads();
	}
	if (threadIdx.x == 0) total[0] = lds[0];
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
	hipLaunchKernelGGL((BsrSpmvF32Kernel<double, double>), dim3(SpmvBlocks(a.rows)), dim3(256), 0,
		ctx->stream, a.rows, a.row_offsets, a.columns, values, x, y);
	TFI_LAUNCH_CHECK("BsrSpmvF32Kernel");
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

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned int* g_scratch[kMaxDevices] = {};  // device address of `scratch` per GPU

}  // namespace

void InitNoAllocScratch(int device)
{
	if (device < 0 || device >= kMaxDevices) throw Error("device index beyond the allocation-free scratch table");
	std::lock_guard<std::mutex> lock(g_scratch_mutex);
	if (g_scratch[device]) return;
	void* ptr = nullptr;
	Check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
	if (!ptr) throw Error("hipGetSymbolAddress(scratch) returned a null address");
	g_scratch[device] = static_cast<unsigned int*>(ptr);
}

}  // namespace tfi

extern "C" int tfi_block_inverse_noalloc(tfi_mesh* m, const double* values, double* inverse, int* singular)
{
	using namespace tfi;
	if (!m) return TFI_ERROR;
	tfi_ctx* ctx = m->ctx;
	try {
		if (!values || !inverse) throw Error("tfi_block_inverse_noalloc needs device arrays of values and inverses");
		const int rows = m->nodes;
		const int groups = (rows + 255)/256;
		if (groups > kNoAllocGroups)
			throw Error("tfi_block_inverse_noalloc supports at most " + std::to_string(256LL*kNoAllocGroups)
				+ " nodes (its fixed working memory), the mesh has " + std::to_string(rows));
		unsigned int* ptr = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_scratch[ctx->device] : nullptr;
		if (!ptr) throw Error("allocation-free working memory was not set up for this device");
		if (rows > 0) {
			const BsrView a = m->view();
			hipLaunchKernelGGL(BlockInverseCountKernel, dim3(groups), dim3(256), 0, ctx->stream, rows, a.diagonal,
				values, inverse, ptr);
			TFI_LAUNCH_CHECK("BlockInverseCountKernel");
		}
		hipLaunchKernelGGL(CountSumKernel, dim3(1), dim3(256), 0, ctx->stream, ptr, groups, ptr + kNoAllocGroups);
		TFI_LAUNCH_CHECK("CountSumKernel");
		// the total lands in the context's pinned host buffer (no allocation)
		Check(hipMemcpyAsync(ctx->host_pinned, ptr + kNoAllocGroups, sizeof(unsigned int), hipMemcpyDeviceToHost,
			ctx->stream), "download singular count");
		Check(hipStreamSynchronize(ctx->stream), "stream sync");
		unsigned int count = 0;
		std::memcpy(&count, ctx->host_pinned, sizeof(count));
		if (singular) *singular = static_cast<int>(count);
		return TFI_OK;
	} catch (const std::exception& e) {
		ctx->error = e.what();
		return TFI_ERROR;
	}
}
