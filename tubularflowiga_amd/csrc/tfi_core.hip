// Context, device buffers, deterministic vector kernels and the context ABI.
// Replaces the reference's CUDA runtime helpers and cuBLAS level-1 calls
// (CudaRuntime.hpp:34-102; cublasDdot/Dnrm2/Daxpy/Dscal/Dcopy call sites in
// Gmres.hpp and iga_cuda.cu).  Reductions are two-pass with a fixed grid, so
// every dot product is bitwise reproducible run to run.
#include "tfi_internal.hpp"

#include <cmath>
#include <cstdio>

namespace tfi {

namespace {

constexpr int kThreads = 256;
constexpr int kReduceBlocks = 512;

__device__ inline double WaveSum(double v)
{
	for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_xor(v, offset, 64);
	return v;
}

// Block-wide sum in fixed order: wave sums then wave 0 adds the 4 wave totals.
__device__ inline double BlockSum(double v, double* lds)
{
	v = WaveSum(v);
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	if (lane == 0) lds[wave] = v;
	__syncthreads();
	double total = 0.0;
	if (threadIdx.x == 0)
		for (int w = 0; w < (int)(blockDim.x >> 6); ++w) total += lds[w];
	__syncthreads();
	return total;
}

__global__ void DotPartialKernel(const double* __restrict__ x, const double* __restrict__ y, long long n,
	double* __restrict__ partial)
{
	__shared__ double lds[16];
	double s = 0.0;
	for (long long i = blockIdx.x*(long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x*blockDim.x)
		s += x[i]*y[i];
	s = BlockSum(s, lds);
	if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void FinalSumKernel(const double* __restrict__ partial, int count, double* __restrict__ out)
{
	__shared__ double lds[16];
	double s = 0.0;
	for (int i = threadIdx.x; i < count; i += blockDim.x) s += partial[i];
	s = BlockSum(s, lds);
	if (threadIdx.x == 0) *out = s;
}

__global__ void AxpbyKernel(double a, const double* __restrict__ x, double b, double* __restrict__ y, long long n)
{
	for (long long i = blockIdx.x*(long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x*blockDim.x)
		y[i] = a*x[i] + b*y[i];
}

__global__ void AxpyKernel(double a, const double* __restrict__ x, double* __restrict__ y, long long n)
{
	for (long long i = blockIdx.x*(long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x*blockDim.x)
		y[i] += a*x[i];
}

__global__ void ScaleKernel(double a, double* __restrict__ x, long long n)
{
	for (long long i = blockIdx.x*(long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x*blockDim.x)
		x[i] *= a;
}

typedef double doublex4_t __attribute__((ext_vector_type(4)));

// FP64 matrix-core throughput probe: every wave runs `iterations` rounds of
// four independent v_mfma_f64_16x16x4f64 chains (2048 flops each) and writes
// one value per lane so that nothing is optimised away.
__global__ void __launch_bounds__(256) MfmaPeakKernel(int iterations, double seed, double* __restrict__ out)
{
	const int lane = threadIdx.x & 63;
	const double a = seed + lane*1e-3, b = seed - lane*1e-3;
	doublex4_t c0 = {0.0, 0.0, 0.0, 0.0}, c1 = c0, c2 = c0, c3 = c0;
	for (int i = 0; i < iterations; ++i) {
		c0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c0, 0, 0, 0);
		c1 = __builtin_amdgcn_mfma_f64_16x16x4f64(b, a, c1, 0, 0, 0);
		c2 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, a, c2, 0, 0, 0);
		c3 = __builtin_amdgcn_mfma_f64_16x16x4f64(b, b, c3, 0, 0, 0);
	}
	out[blockIdx.x*(size_t)blockDim.x + threadIdx.x] = c0[0] + c1[1] + c2[2] + c3[3];
}

__global__ void CopyKernel(const double* __restrict__ x, double* __restrict__ y, long long n)
{
	for (long long i = blockIdx.x*(long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x*blockDim.x)
		y[i] = x[i];
}

// Stream probes for the measured HBM peak: each workgroup streams one
// contiguous chunk with 16-byte accesses, eight in flight per lane before the
// first store (copy) or add (read); n2 counts double2.  (A grid-stride walk
// measured 5-10 % slower.)
constexpr int kStreamUnroll = 8;

__global__ void __launch_bounds__(256) StreamCopyKernel(const double2* __restrict__ x, double2* __restrict__ y,
	long long n2)
{
	const long long per = (n2 + gridDim.x - 1)/gridDim.x;
	const long long b = blockIdx.x*per, e = b + per < n2 ? b + per : n2;
	long long i = b + threadIdx.x;
	for (; i + (kStreamUnroll - 1)*(long long)blockDim.x < e; i += kStreamUnroll*(long long)blockDim.x) {
		double2 v[kStreamUnroll];
#pragma unroll
		for (int k = 0; k < kStreamUnroll; ++k) v[k] = x[i + k*(long long)blockDim.x];
#pragma unroll
		for (int k = 0; k < kStreamUnroll; ++k) y[i + k*(long long)blockDim.x] = v[k];
	}
	for (; i < e; i += blockDim.x) y[i] = x[i];
}

__global__ void __launch_bounds__(256) StreamReadKernel(const double2* __restrict__ x, long long n2,
	double* __restrict__ partial)
{
	const long long per = (n2 + gridDim.x - 1)/gridDim.x;
	const long long b = blockIdx.x*per, e = b + per < n2 ? b + per : n2;
	long long i = b + threadIdx.x;
	double s = 0.0;
	for (; i + (kStreamUnroll - 1)*(long long)blockDim.x < e; i += kStreamUnroll*(long long)blockDim.x) {
		double2 v[kStreamUnroll];
#pragma unroll
		for (int k = 0; k < kStreamUnroll; ++k) v[k] = x[i + k*(long long)blockDim.x];
#pragma unroll
		for (int k = 0; k < kStreamUnroll; ++k) s += v[k].x + v[k].y;
	}
	for (; i < e; i += blockDim.x) s += x[i].x + x[i].y;
	__shared__ double lane_sum[256];
	lane_sum[threadIdx.x] = s;
	__syncthreads();
	for (int w = blockDim.x/2; w > 0; w >>= 1) {
		if (threadIdx.x < w) lane_sum[threadIdx.x] += lane_sum[threadIdx.x + w];
		__syncthreads();
	}
	if (threadIdx.x == 0) partial[blockIdx.x] = lane_sum[0];
}

__global__ void FillKernel(double* __restrict__ x, double v, long long n)
{
	for (long long i = blockIdx.x*(long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x*blockDim.x)
		x[i] = v;
}

__global__ void FillIntKernel(int* __restrict__ x, int v, long long n)
{
	for (long long i = blockIdx.x*(long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x*blockDim.x)
		x[i] = v;
}

__global__ void ToFloatKernel(const double* __restrict__ x, float* __restrict__ y, long long n)
{
	for (long long i = blockIdx.x*(long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x*blockDim.x)
		y[i] = static_cast<float>(x[i]);
}

int Blocks(long long n)
{
	long long b = (n + kThreads - 1)/kThreads;
	if (b > 8192) b = 8192;
	return b < 1 ? 1 : static_cast<int>(b);
}

}  // namespace

void EnsureScratch(tfi_ctx* ctx, size_t count)
{
	if (ctx->scratch_count >= count) return;
	if (ctx->scratch) (void)hipFree(ctx->scratch);
	ctx->scratch = nullptr;
	Check(hipMalloc(&ctx->scratch, count*sizeof(double)), "hipMalloc scratch");
	ctx->scratch_count = count;
}

double Dot(tfi_ctx* ctx, const double* x, const double* y, long long n)
{
	EnsureScratch(ctx, kReduceBlocks + 8);
	hipLaunchKernelGGL(DotPartialKernel, dim3(kReduceBlocks), dim3(kThreads), 0, ctx->stream, x, y, n, ctx->scratch);
	TFI_LAUNCH_CHECK("DotPartialKernel");
	hipLaunchKernelGGL(FinalSumKernel, dim3(1), dim3(kThreads), 0, ctx->stream, ctx->scratch, kReduceBlocks,
		ctx->scratch + kReduceBlocks);
	TFI_LAUNCH_CHECK("FinalSumKernel");
	Check(hipMemcpyAsync(ctx->host_pinned, ctx->scratch + kReduceBlocks, sizeof(double), hipMemcpyDeviceToHost,
		ctx->stream), "dot download");
	Check(hipStreamSynchronize(ctx->stream), "dot sync");
	return ctx->host_pinned[0];
}

double Norm2(tfi_ctx* ctx, const double* x, long long n)
{
	return std::sqrt(Dot(ctx, x, x, n));
}

void Axpy(tfi_ctx* ctx, double a, const double* x, double* y, long long n)
{
	hipLaunchKernelGGL(AxpyKernel, dim3(Blocks(n)), dim3(kThreads), 0, ctx->stream, a, x, y, n);
	TFI_LAUNCH_CHECK("AxpyKernel");
}

void Axpby(tfi_ctx* ctx, double a, const double* x, double b, double* y, long long n)
{
	hipLaunchKernelGGL(AxpbyKernel, dim3(Blocks(n)), dim3(kThreads), 0, ctx->stream, a, x, b, y, n);
	TFI_LAUNCH_CHECK("AxpbyKernel");
}

void Scale(tfi_ctx* ctx, double a, double* x, long long n)
{
	hipLaunchKernelGGL(ScaleKernel, dim3(Blocks(n)), dim3(kThreads), 0, ctx->stream, a, x, n);
	TFI_LAUNCH_CHECK("ScaleKernel");
}

void MfmaPeak(tfi_ctx* ctx, int blocks, int threads, int iterations, double* out)
{
	hipLaunchKernelGGL(MfmaPeakKernel, dim3(blocks), dim3(threads), 0, ctx->stream, iterations, 0.5, out);
	TFI_LAUNCH_CHECK("MfmaPeakKernel");
}

void StreamCopy(tfi_ctx* ctx, const double* x, double* y, long long n, int blocks)
{
	hipLaunchKernelGGL(StreamCopyKernel, dim3(blocks), dim3(kThreads), 0, ctx->stream,
		reinterpret_cast<const double2*>(x), reinterpret_cast<double2*>(y), n/2);
	TFI_LAUNCH_CHECK("StreamCopyKernel");
}

void StreamRead(tfi_ctx* ctx, const double* x, long long n, double* partial, int blocks)
{
	hipLaunchKernelGGL(StreamReadKernel, dim3(blocks), dim3(kThreads), 0, ctx->stream,
		reinterpret_cast<const double2*>(x), n/2, partial);
	TFI_LAUNCH_CHECK("StreamReadKernel");
}

void Copy(tfi_ctx* ctx, const double* x, double* y, long long n)
{
	hipLaunchKernelGGL(CopyKernel, dim3(Blocks(n)), dim3(kThreads), 0, ctx->stream, x, y, n);
	TFI_LAUNCH_CHECK("CopyKernel");
}

void Fill(tfi_ctx* ctx, double* x, double value, long long n)
{
	hipLaunchKernelGGL(FillKernel, dim3(Blocks(n)), dim3(kThreads), 0, ctx->stream, x, value, n);
	TFI_LAUNCH_CHECK("FillKernel");
}

void FillInt(tfi_ctx* ctx, int* x, int value, long long n)
{
	hipLaunchKernelGGL(FillIntKernel, dim3(Blocks(n)), dim3(kThreads), 0, ctx->stream, x, value, n);
	TFI_LAUNCH_CHECK("FillIntKernel");
}

void ToFloat(tfi_ctx* ctx, const double* x, float* y, long long n)
{
	hipLaunchKernelGGL(ToFloatKernel, dim3(Blocks(n)), dim3(kThreads), 0, ctx->stream, x, y, n);
	TFI_LAUNCH_CHECK("ToFloatKernel");
}

void Sync(tfi_ctx* ctx)
{
	Check(hipStreamSynchronize(ctx->stream), "stream synchronize");
}

hipEvent_t TakeEvent(tfi_ctx* ctx)
{
	if (!ctx->event_pool.empty()) {
		hipEvent_t e = ctx->event_pool.back();
		ctx->event_pool.pop_back();
		return e;
	}
	hipEvent_t e = nullptr;
	Check(hipEventCreate(&e), "hipEventCreate");
	return e;
}

void ResolveTimers(tfi_ctx* ctx)
{
	for (auto& p : ctx->pending) {
		float ms = 0.0f;
		Check(hipEventSynchronize(p.end), "timer sync");
		Check(hipEventElapsedTime(&ms, p.begin, p.end), "timer elapsed");
		*p.accumulate += ms;
		ctx->event_pool.push_back(p.begin);
		ctx->event_pool.push_back(p.end);
	}
	ctx->pending.clear();
}

}  // namespace tfi

using namespace tfi;

extern "C" {

int tfi_ctx_create(int device, tfi_ctx** out)
{
	if (!out) return TFI_ERROR;
	*out = nullptr;
	auto* ctx = new tfi_ctx;
	try {
		int count = 0;
		Check(hipGetDeviceCount(&count), "hipGetDeviceCount");
		if (device < 0 || device >= count) throw Error("no such HIP device");
		Check(hipSetDevice(device), "hipSetDevice");
		ctx->device = device;
		InitNoAllocScratch(device);
		Check(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
		Check(hipEventCreate(&ctx->ev_begin), "hipEventCreate");
		Check(hipEventCreate(&ctx->ev_end), "hipEventCreate");
		Check(hipHostMalloc(reinterpret_cast<void**>(&ctx->host_pinned), 4096*sizeof(double), hipHostMallocDefault),
			"hipHostMalloc");
		hipDeviceProp_t prop{};
		Check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
		ctx->cus = prop.multiProcessorCount;
		*out = ctx;
		return TFI_OK;
	} catch (const std::exception& e) {
		std::fprintf(stderr, "tfi_ctx_create: %s\n", e.what());
		delete ctx;
		return TFI_ERROR;
	}
}

void tfi_ctx_destroy(tfi_ctx* ctx)
{
	if (!ctx) return;
	if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
	if (ctx->scratch) (void)hipFree(ctx->scratch);
	if (ctx->host_pinned) (void)hipHostFree(ctx->host_pinned);
	for (auto& p : ctx->pending) {
		(void)hipEventDestroy(p.begin);
		(void)hipEventDestroy(p.end);
	}
	for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
	if (ctx->ev_begin) (void)hipEventDestroy(ctx->ev_begin);
	if (ctx->ev_end) (void)hipEventDestroy(ctx->ev_end);
	if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
	delete ctx;
}

const char* tfi_last_error(const tfi_ctx* ctx)
{
	return ctx ? ctx->error.c_str() : "null context";
}

int tfi_device_info(tfi_ctx* ctx, char* buffer, int length)
{
	try {
		hipDeviceProp_t prop{};
		Check(hipGetDeviceProperties(&prop, ctx->device), "hipGetDeviceProperties");
		size_t free_b = 0, total_b = 0;
		Check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
		int runtime = 0, driver = 0;
		(void)hipRuntimeGetVersion(&runtime);
		(void)hipDriverGetVersion(&driver);
		std::snprintf(buffer, length,
			"device=%d name=\"%s\" arch=%s compute_units=%d fp64=native memory_free_gib=%.3f "
			"memory_total_gib=%.3f hip_runtime=%d hip_driver=%d",
			ctx->device, prop.name, prop.gcnArchName, prop.multiProcessorCount, free_b/1073741824.0,
			total_b/1073741824.0, runtime, driver);
		return TFI_OK;
	} catch (const std::exception& e) {
		ctx->error = e.what();
		return TFI_ERROR;
	}
}

int tfi_allocation_stats(tfi_ctx* ctx, long long* peak_bytes, long long* live_bytes)
{
	if (!ctx) return TFI_ERROR;
	if (peak_bytes) *peak_bytes = ctx->peak_bytes;
	if (live_bytes) *live_bytes = ctx->live_bytes;
	return TFI_OK;
}

int tfi_timers_get(tfi_ctx* ctx, tfi_kernel_timers* out)
{
	if (!ctx || !out) return TFI_ERROR;
	try {
		ResolveTimers(ctx);
	} catch (const std::exception& e) {
		ctx->error = e.what();
		return TFI_ERROR;
	}
	*out = ctx->timers;
	return TFI_OK;
}

int tfi_timers_enable(tfi_ctx* ctx, int enabled)
{
	if (!ctx) return TFI_ERROR;
	ctx->timers_enabled = enabled != 0;
	return TFI_OK;
}

int tfi_timers_reset(tfi_ctx* ctx)
{
	if (!ctx) return TFI_ERROR;
	try {
		ResolveTimers(ctx);
	} catch (const std::exception& e) {
		ctx->error = e.what();
		return TFI_ERROR;
	}
	ctx->timers = tfi_kernel_timers{};
	return TFI_OK;
}

}  // extern "C"
