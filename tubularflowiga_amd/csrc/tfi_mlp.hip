// Learned warm start: a 3-layer stencil MLP (kIn -> H -> H -> 4, tanh) trained
// online on the solver's own converged steps, with hand-written kernels for
// the feature gather, FP32 forward/backward GEMMs (fixed-order sums, so a
// training run is bitwise reproducible), Adam, and BF16-MFMA inference
// (v_mfma_f32_16x16x16_bf16, FP32 accumulation).  The reference has no
// counterpart; this is the north-star "learned part".
#include "tfi_mlp.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>

namespace tfi {

namespace {

constexpr int kT = 256;

// ------------------------------------------------------------------ generic FP32 GEMM
// C[m][n] = epi( sum_k A(m,k) B(k,n) ),  A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn]
// epi: mode 0 plain, 1 tanh(acc + bias[n]), 2 acc + bias[n], 3 acc * (1 - H[m][n]^2)
struct GemmArgs {
	int M, N, K;
	const float* A;
	long long sam, sak;
	const float* B;
	long long sbk, sbn;
	float* C;
	long long ldc;
	const float* bias;
	const float* H;
	long long ldh;
	int mode;
	// split-K (mode 0 only): blockIdx.z sums k in [z*kchunk, (z+1)*kchunk) into
	// partial[z][M][N]; SplitReduceKernel then adds the slices in order.
	int kchunk = 0;
	float* partial = nullptr;
};

__global__ void __launch_bounds__(kT) GemmKernel(GemmArgs g)
{
	__shared__ float As[16][65];
	__shared__ float Bs[16][65];
	const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
	const int m0 = blockIdx.y*64, n0 = blockIdx.x*64;
	const int kbeg = g.kchunk ? blockIdx.z*g.kchunk : 0;
	const int kend = g.kchunk ? min(g.K, kbeg + g.kchunk) : g.K;
	float acc[4][4] = {};
	for (int k0 = kbeg; k0 < kend; k0 += 16) {
		for (int i = threadIdx.x; i < 16*64; i += kT) {
			const int kk = i >> 6, mm = i & 63;
			const int m = m0 + mm, k = k0 + kk;
			As[kk][mm] = (m < g.M && k < kend) ? g.A[m*g.sam + k*g.sak] : 0.0f;
			const int n = n0 + mm;
			Bs[kk][mm] = (n < g.N && k < kend) ? g.B[k*g.sbk + n*g.sbn] : 0.0f;
		}
		__syncthreads();
#pragma unroll
		for (int kk = 0; kk < 16; ++kk) {
			float a[4], b[4];
#pragma unroll
			for (int i = 0; i < 4; ++i) a[i] = As[kk][ty*4 + i];
#pragma unroll
			for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx*4 + j];
#pragma unroll
			for (int i = 0; i < 4; ++i)
#pragma unroll
				for (int j = 0; j < 4; ++j) acc[i][j] += a[i]*b[j];
		}
		__syncthreads();
	}
	if (g.kchunk) {
		float* out = g.partial + (size_t)blockIdx.z*g.M*g.N;
		for (int i = 0; i < 4; ++i)
			for (int j = 0; j < 4; ++j) {
				const int m = m0 + ty*4 + i, n = n0 + tx*4 + j;
				if (m < g.M && n < g.N) out[(size_t)m*g.N + n] = acc[i][j];
			}
		return;
	}
	for (int i = 0; i < 4; ++i)
		for (int j = 0; j < 4; ++j) {
			const int m = m0 + ty*4 + i, n = n0 + tx*4 + j;
			if (m >= g.M || n >= g.N) continue;
			float v = acc[i][j];
			if (g.mode == 1) v = tanhf(v + g.bias[n]);
			else if (g.mode == 2) v = v + g.bias[n];
			else if (g.mode == 3) {
				const float h = g.H[m*g.ldh + n];
				v = v*(1.0f - h*h);
			}
			g.C[m*g.ldc + n] = v;
		}
}

// C[i] = sum_z partial[z][i] in z order (deterministic split-K reduction)
__global__ void SplitReduceKernel(long long count, int splits, const float* __restrict__ partial,
	float* __restrict__ C)
{
	const long long i = blockIdx.x*(long long)blockDim.x + threadIdx.x;
	if (i >= count) return;
	float s = 0.0f;
	for (int z = 0; z < splits; ++z) s += partial[(size_t)z*count + i];
	C[i] = s;
}

constexpr int kSplitMax = 64;

// Tall-K products (weight gradients, K = batch) split K over up to kSplitMax
// workgroup slices so that the launch fills the chip; `scratch` holds
// kSplitMax*M*N floats.  Other products run as one K pass.
void Gemm(tfi_ctx* ctx, GemmArgs g, float* scratch = nullptr)
{
	const int tiles = ((g.N + 63)/64)*((g.M + 63)/64);
	int splits = 1;
	if (scratch && g.mode == 0 && tiles < 64 && g.K >= 512)
		splits = std::min(kSplitMax, std::max(1, g.K/256));
	if (splits > 1) {
		g.kchunk = ((g.K + splits - 1)/splits + 15)/16*16;
		splits = (g.K + g.kchunk - 1)/g.kchunk;
		g.partial = scratch;
		dim3 grid((g.N + 63)/64, (g.M + 63)/64, splits);
		hipLaunchKernelGGL(GemmKernel, grid, dim3(kT), 0, ctx->stream, g);
		TFI_LAUNCH_CHECK("GemmKernel");
		const long long count = (long long)g.M*g.N;
		hipLaunchKernelGGL(SplitReduceKernel, dim3((count + kT - 1)/kT), dim3(kT), 0, ctx->stream, count, splits,
			scratch, g.C);
		TFI_LAUNCH_CHECK("SplitReduceKernel");
		return;
	}
	dim3 grid((g.N + 63)/64, (g.M + 63)/64);
	hipLaunchKernelGGL(GemmKernel, grid, dim3(kT), 0, ctx->stream, g);
	TFI_LAUNCH_CHECK("GemmKernel");
}

// partial[z][n] = sum of rows [z*chunk, (z+1)*chunk) of D[:, n] (fixed order)
__global__ void ColumnSumPartialKernel(int M, int N, int chunk, const float* __restrict__ D,
	float* __restrict__ partial)
{
	const int n = blockIdx.x*blockDim.x + threadIdx.x;
	if (n >= N) return;
	const int m0 = blockIdx.y*chunk, m1 = min(M, m0 + chunk);
	float s = 0.0f;
	for (int m = m0; m < m1; ++m) s += D[(long long)m*N + n];
	partial[(size_t)blockIdx.y*N + n] = s;
}

// out[n] = sum_m D[m][n] in two fixed-order passes
void ColumnSum(tfi_ctx* ctx, int M, int N, const float* D, float* out, float* scratch)
{
	const int splits = std::min(kSplitMax, std::max(1, M/64));
	const int chunk = (M + splits - 1)/splits;
	hipLaunchKernelGGL(ColumnSumPartialKernel, dim3((N + 63)/64, splits), dim3(64), 0, ctx->stream, M, N, chunk, D,
		scratch);
	TFI_LAUNCH_CHECK("ColumnSumPartialKernel");
	hipLaunchKernelGGL(SplitReduceKernel, dim3((N + kT - 1)/kT), dim3(kT), 0, ctx->stream, (long long)N, splits,
		scratch, out);
	TFI_LAUNCH_CHECK("SplitReduceKernel");
}

// batch rows of one optimisation step: rows[b] = hash(seed, b) mod N
__global__ void BatchRowsKernel(int batch, int nodes, unsigned long long seed, int* __restrict__ rows)
{
	const int b = blockIdx.x*blockDim.x + threadIdx.x;
	if (b >= batch) return;
	unsigned long long z = seed + (unsigned long long)b*0x9e3779b97f4a7c15ULL;
	z = (z ^ (z >> 30))*0xbf58476d1ce4e5b9ULL;
	z = (z ^ (z >> 27))*0x94d049bb133111ebULL;
	z ^= z >> 31;
	rows[b] = static_cast<int>(z%static_cast<unsigned long long>(nodes));
}

// delta3 = 2 (Y - T)/(B*4), partial loss per block
__global__ void LossKernel(int B, const float* __restrict__ Y, const float* __restrict__ T, float* __restrict__ d3,
	float* __restrict__ partial)
{
	__shared__ float lds[kT];
	float s = 0.0f;
	const int total = B*kOut;
	for (int i = blockIdx.x*blockDim.x + threadIdx.x; i < total; i += gridDim.x*blockDim.x) {
		const float e = Y[i] - T[i];
		s += e*e;
		d3[i] = 2.0f*e/static_cast<float>(total);
	}
	lds[threadIdx.x] = s;
	__syncthreads();
	for (int o = kT/2; o > 0; o >>= 1) {
		if (threadIdx.x < o) lds[threadIdx.x] += lds[threadIdx.x + o];
		__syncthreads();
	}
	if (threadIdx.x == 0) partial[blockIdx.x] = lds[0]/static_cast<float>(total);
}

// <Y, T> and <Y, Y> partial sums over B rows (one pair per block)
__global__ void GateKernel(int B, const float* __restrict__ Y, const float* __restrict__ T, double* __restrict__ partial)
{
	__shared__ double lds[2][kT];
	double yt = 0.0, yy = 0.0;
	for (int i = blockIdx.x*blockDim.x + threadIdx.x; i < B*kOut; i += gridDim.x*blockDim.x) {
		yt += static_cast<double>(Y[i])*T[i];
		yy += static_cast<double>(Y[i])*Y[i];
	}
	lds[0][threadIdx.x] = yt;
	lds[1][threadIdx.x] = yy;
	__syncthreads();
	for (int o = kT/2; o > 0; o >>= 1) {
		if (threadIdx.x < o) {
			lds[0][threadIdx.x] += lds[0][threadIdx.x + o];
			lds[1][threadIdx.x] += lds[1][threadIdx.x + o];
		}
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		partial[2*blockIdx.x] = lds[0][0];
		partial[2*blockIdx.x + 1] = lds[1][0];
	}
}

__global__ void AdamKernel(int n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
	float* __restrict__ v, float lr, float b1, float b2, float c1, float c2, float eps)
{
	const int i = blockIdx.x*blockDim.x + threadIdx.x;
	if (i >= n) return;
	const float gi = g[i];
	const float mi = b1*m[i] + (1.0f - b1)*gi;
	const float vi = b2*v[i] + (1.0f - b2)*gi*gi;
	m[i] = mi;
	v[i] = vi;
	w[i] -= lr*(mi/c1)/(sqrtf(vi/c2) + eps);
}

// features of target nodes rows[r] (or r itself when rows == nullptr)
// history u_{n-j}, j = 0..kOrder
struct Hist {
	const double* h[kOrder + 1];
};

// backward difference d_k at index t (d_0 = u_n)
__device__ inline double Diff(const Hist& H, int k, long long t)
{
	double c = 1.0, s = 0.0;
	for (int j = 0; j <= k; ++j) {
		s += c*H.h[j][t];
		c = -c*static_cast<double>(k - j)/static_cast<double>(j + 1);
	}
	return s;
}

// order-kOrder extrapolation in time: sum_{j=0..P} (-1)^j C(P+1, j+1) u_{n-j}
__device__ inline double Extrapolate(const Hist& H, long long t)
{
	double c = static_cast<double>(kOrder + 1), s = 0.0;
	for (int j = 0; j <= kOrder; ++j) {
		s += c*H.h[j][t];
		c = -c*static_cast<double>(kOrder - j)/static_cast<double>(j + 2);
	}
	return s;
}

__global__ void FeatureKernel(int count, const int* __restrict__ rows, const int* __restrict__ stencil,
	const float* __restrict__ pos, const float* __restrict__ hscale, Hist H, const double* __restrict__ scales,
	const double* __restrict__ phase, float* __restrict__ X)
{
	const long long t = blockIdx.x*(long long)blockDim.x + threadIdx.x;
	if (t >= (long long)count*kIn) return;
	const int r = static_cast<int>(t/kIn), c = static_cast<int>(t - (long long)r*kIn);
	const int i = rows ? rows[r] : r;
	float v = 0.0f;
	if (c < kStencil*kPerNode) {
		const int s = c/kPerNode, f = c - s*kPerNode;
		const int j = stencil[(long long)i*kStencil + s];
		if (f < 4*(kOrder + 1)) {
			const int k = f >> 2, field = f & 3;
			v = static_cast<float>(Diff(H, k, 4LL*j + field)/scales[f]);
		} else {
			const int d = f - 4*(kOrder + 1);
			v = (pos[3LL*j + d] - pos[3LL*i + d])/hscale[i];
		}
	} else if (c < kStencil*kPerNode + 8) {
		v = static_cast<float>(phase[c - kStencil*kPerNode]);
	}
	X[t] = v;
}

// error of the base prediction (base, or the extrapolation when base is null)
// of the next state, scaled by s_P
__global__ void TargetKernel(int count, const int* __restrict__ rows, const double* __restrict__ next, Hist H,
	const double* __restrict__ base, const double* __restrict__ scales, float* __restrict__ T)
{
	const int t = blockIdx.x*blockDim.x + threadIdx.x;
	if (t >= count*kOut) return;
	const int r = t >> 2, f = t & 3;
	const long long k = 4LL*rows[r] + f;
	const double b = base ? base[k] : Extrapolate(H, k);
	T[t] = static_cast<float>((next[k] - b)/scales[4*kOrder + f]);
}

__global__ void ApplyPredictionKernel(int nodes, Hist H, const double* __restrict__ base, const float* __restrict__ Y,
	const double* __restrict__ scales, double gate, double* __restrict__ state)
{
	const int t = blockIdx.x*blockDim.x + threadIdx.x;
	if (t >= nodes*4) return;
	const int f = t & 3;
	const double b = base ? base[t] : Extrapolate(H, t);
	state[t] = b + gate*scales[4*kOrder + f]*static_cast<double>(Y[t]);
}

constexpr int kScales = 4*(kOrder + 1);

// per-field sums of squares of u_n and d_1..d_P (partials per block)
__global__ void ScalePartialKernel(int nodes, Hist H, double* __restrict__ partial)
{
	__shared__ double lds[kScales][kT];
	double s[kScales] = {};
	for (int n = blockIdx.x*blockDim.x + threadIdx.x; n < nodes; n += gridDim.x*blockDim.x)
		for (int f = 0; f < 4; ++f)
			for (int k = 0; k <= kOrder; ++k) {
				const double d = Diff(H, k, 4LL*n + f);
				s[4*k + f] += d*d;
			}
	for (int f = 0; f < kScales; ++f) lds[f][threadIdx.x] = s[f];
	__syncthreads();
	for (int o = kT/2; o > 0; o >>= 1) {
		if (threadIdx.x < o)
			for (int f = 0; f < kScales; ++f) lds[f][threadIdx.x] += lds[f][threadIdx.x + o];
		__syncthreads();
	}
	if (threadIdx.x == 0)
		for (int f = 0; f < kScales; ++f) partial[kScales*blockIdx.x + f] = lds[f][0];
}

// ------------------------------------------------------------------ BF16 MFMA inference
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ inline short ToBf16(float f)
{
	// round to nearest even
	uint32_t u = __float_as_uint(f);
	const uint32_t rounding = 0x7fffu + ((u >> 16) & 1u);
	u += rounding;
	return static_cast<short>(u >> 16);
}

__global__ void ToBf16Kernel(long long n, const float* __restrict__ x, short* __restrict__ y)
{
	for (long long i = blockIdx.x*(long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x*blockDim.x)
		y[i] = ToBf16(x[i]);
}

// Transposed bf16 copy of W [K][N] -> Wt [N][K] (so B fragments read K-contiguous)
__global__ void TransposeBf16Kernel(int K, int N, const float* __restrict__ W, short* __restrict__ Wt)
{
	const int t = blockIdx.x*blockDim.x + threadIdx.x;
	if (t >= K*N) return;
	const int k = t/N, n = t - k*N;
	Wt[(long long)n*K + k] = ToBf16(W[t]);
}

// One wave computes 16 rows.  Fragment layouts of v_mfma_f32_16x16x16_bf16:
// A (16x16): lane l holds row l%16, k = 4*(l/16)..+3; B (16x16): lane l holds
// column l%16, k = 4*(l/16)..+3; D (16x16): lane l holds column l%16, rows
// 4*(l/16)..+3.
template <int H>
__global__ void __launch_bounds__(256) MlpInferKernel(int rows, const short* __restrict__ X,
	const short* __restrict__ W1t, const float* __restrict__ b1, const short* __restrict__ W2t,
	const float* __restrict__ b2, const short* __restrict__ W3t, const float* __restrict__ b3, float* __restrict__ Y)
{
	__shared__ short hbuf[4][16][H + 8];
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
	// every wave reaches every barrier: a wave past the last row computes on a
	// clamped row and stores nothing
	const int row0 = (blockIdx.x*4 + wave)*16;
	const int lr = lane & 15, lk = (lane >> 4)*4;
	const int arow = max(0, min(row0 + lr, rows - 1));
	short (*hb)[H + 8] = hbuf[wave];
	// layer 1: [16 x kIn] x [kIn x H]
	floatx4 acc[H/16];
	for (int t = 0; t < H/16; ++t) acc[t] = floatx4{0, 0, 0, 0};
	for (int k0 = 0; k0 < kIn; k0 += 16) {
		const short* ap = X + (long long)arow*kIn + k0 + lk;
		const bf16x4 a = {ap[0], ap[1], ap[2], ap[3]};
		for (int t = 0; t < H/16; ++t) {
			const short* bp = W1t + (long long)(t*16 + lr)*kIn + k0 + lk;
			const bf16x4 b = {bp[0], bp[1], bp[2], bp[3]};
			acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
		}
	}
	for (int t = 0; t < H/16; ++t)
		for (int r = 0; r < 4; ++r) {
			const int col = t*16 + lr;
			hb[lk + r][col] = ToBf16(tanhf(acc[t][r] + b1[col]));
		}
	__syncthreads();
	// layer 2: [16 x H] x [H x H]
	for (int t = 0; t < H/16; ++t) acc[t] = floatx4{0, 0, 0, 0};
	for (int k0 = 0; k0 < H; k0 += 16) {
		const bf16x4 a = {hb[lr][k0 + lk], hb[lr][k0 + lk + 1], hb[lr][k0 + lk + 2], hb[lr][k0 + lk + 3]};
		for (int t = 0; t < H/16; ++t) {
			const short* bp = W2t + (long long)(t*16 + lr)*H + k0 + lk;
			const bf16x4 b = {bp[0], bp[1], bp[2], bp[3]};
			acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
		}
	}
	__syncthreads();
	for (int t = 0; t < H/16; ++t)
		for (int r = 0; r < 4; ++r) {
			const int col = t*16 + lr;
			hb[lk + r][col] = ToBf16(tanhf(acc[t][r] + b2[col]));
		}
	__syncthreads();
	// layer 3: [16 x H] x [H x 4] (output columns padded to 16)
	floatx4 out = {0, 0, 0, 0};
	for (int k0 = 0; k0 < H; k0 += 16) {
		const bf16x4 a = {hb[lr][k0 + lk], hb[lr][k0 + lk + 1], hb[lr][k0 + lk + 2], hb[lr][k0 + lk + 3]};
		bf16x4 b = {0, 0, 0, 0};
		if (lr < kOut) {
			const short* bp = W3t + (long long)lr*H + k0 + lk;
			b = bf16x4{bp[0], bp[1], bp[2], bp[3]};
		}
		out = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, out, 0, 0, 0);
	}
	if (lr < kOut)
		for (int r = 0; r < 4; ++r) {
			const int row = row0 + lk + r;
			if (row < rows) Y[(long long)row*kOut + lr] = out[r] + b3[lr];
		}
}

uint64_t SplitMix(uint64_t& s)
{
	uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
	z = (z ^ (z >> 30))*0xbf58476d1ce4e5b9ULL;
	z = (z ^ (z >> 27))*0x94d049bb133111ebULL;
	return z ^ (z >> 31);
}

}  // namespace

// ------------------------------------------------------------------ set-up on the device
// Control-point position of each basis entry (e, a): sum_k C[a][k] X_k over its
// extraction row, and the row sum (the weight of the entry).
__global__ void EntryPositionKernel(int elements, const int* __restrict__ eoff, const int* __restrict__ xoff,
	const uint8_t* __restrict__ xcol, const double* __restrict__ xval, const double* __restrict__ bezier,
	double* __restrict__ entry)
{
	const int e = blockIdx.x;
	if (e >= elements) return;
	const double* X = bezier + (size_t)e*192;
	for (int a = eoff[e] + threadIdx.x; a < eoff[e + 1]; a += blockDim.x) {
		double s[4] = {0.0, 0.0, 0.0, 0.0};
		for (int k = xoff[a]; k < xoff[a + 1]; ++k) {
			const double c = xval[k];
			const double* x = X + 3*xcol[k];
			s[0] += c*x[0];
			s[1] += c*x[1];
			s[2] += c*x[2];
			s[3] += c;
		}
		for (int d = 0; d < 4; ++d) entry[4*(size_t)a + d] = s[d];
	}
}

// Node position = (sum of its entries' positions) / (sum of their weights),
// entries summed in ascending order (node_entries lists them per node).
__global__ void NodePositionKernel(int nodes, const long long* __restrict__ offsets,
	const long long* __restrict__ node_entries, const double* __restrict__ entry, float* __restrict__ pos)
{
	const int n = blockIdx.x*blockDim.x + threadIdx.x;
	if (n >= nodes) return;
	double s[4] = {0.0, 0.0, 0.0, 0.0};
	for (long long k = offsets[n]; k < offsets[n + 1]; ++k)
		for (int d = 0; d < 4; ++d) s[d] += entry[4*node_entries[k] + d];
	const double w = s[3] != 0.0 ? s[3] : 1.0;
	for (int d = 0; d < 3; ++d) pos[3*(size_t)n + d] = static_cast<float>(s[d]/w);
}

// 27-point stencil of each node: the node, then its 26 nearest pattern
// neighbours by control-point distance (squared distance of the float
// positions summed in double; ties by column index), and the RMS distance of
// the chosen neighbours (1 when there are none).  One 64-lane wave per node:
// every round takes the wave-wide smallest (d2, column) not yet chosen.
constexpr int kStencilLanes = 64;
constexpr int kStencilPerLane = 8;   // rows of up to 512 blocks
__global__ void __launch_bounds__(256) StencilKernel(int nodes, const int* __restrict__ ro,
	const int* __restrict__ cols, const float* __restrict__ pos, int* __restrict__ stencil, float* __restrict__ hscale,
	int* __restrict__ status)
{
	const int lane = threadIdx.x & 63;
	const int n = blockIdx.x*4 + (threadIdx.x >> 6);
	if (n >= nodes) return;
	const int begin = ro[n], end = ro[n + 1];
	if (end - begin > kStencilLanes*kStencilPerLane) {
		if (lane == 0) atomicOr(status, 1);
		return;
	}
	double d2[kStencilPerLane];
	int col[kStencilPerLane];
	const float px = pos[3*(size_t)n], py = pos[3*(size_t)n + 1], pz = pos[3*(size_t)n + 2];
#pragma unroll
	for (int i = 0; i < kStencilPerLane; ++i) {
		const int k = begin + lane + i*kStencilLanes;
		col[i] = k < end ? cols[k] : -1;
		d2[i] = __builtin_huge_val();
		if (col[i] >= 0 && col[i] != n) {
			const int j = col[i];
			const double tx = static_cast<double>(pos[3*(size_t)j] - px);
			const double ty = static_cast<double>(pos[3*(size_t)j + 1] - py);
			const double tz = static_cast<double>(pos[3*(size_t)j + 2] - pz);
			d2[i] = tx*tx + ty*ty + tz*tz;
		} else {
			col[i] = -1;
		}
	}
	int* out = stencil + (size_t)n*kStencil;
	if (lane == 0) out[0] = n;
	double h2 = 0.0;
	int used = 0;
	for (int s = 1; s < kStencil; ++s) {
		// lane-local smallest (d2, column)
		double bd = __builtin_huge_val();
		int bc = 0x7fffffff, bi = -1;
#pragma unroll
		for (int i = 0; i < kStencilPerLane; ++i)
			if (col[i] >= 0 && (d2[i] < bd || (d2[i] == bd && col[i] < bc))) {
				bd = d2[i];
				bc = col[i];
				bi = i;
			}
		// wave-wide smallest
		double wd = bd;
		int wc = bc;
		for (int off = 32; off > 0; off >>= 1) {
			const double od = __shfl_xor(wd, off, 64);
			const int oc = __shfl_xor(wc, off, 64);
			if (od < wd || (od == wd && oc < wc)) {
				wd = od;
				wc = oc;
			}
		}
		if (wc == 0x7fffffff) {
			if (lane == 0) out[s] = n;
			continue;
		}
		if (bi >= 0 && bc == wc) col[bi] = -1;  // columns are unique in a row: one lane owns it
		if (lane == 0) out[s] = wc;
		h2 += wd;
		++used;
	}
	if (lane == 0) {
		float h = static_cast<float>(used ? sqrt(h2/used) : 1.0);
		if (!(h > 0.0f)) h = 1.0f;
		hscale[n] = h;
	}
}

struct Mlp::Impl {
	tfi_mesh* mesh;
	tfi_ctx* ctx;
	tfi_warm_options opt;
	int H = 128;
	int batch = 4096;
	DevBuf<int> stencil, rows;
	DevBuf<float> pos, hscale;
	// parameters + Adam state: packed [w1 b1 w2 b2 w3 b3]
	DevBuf<float> params, grads, adam_m, adam_v;
	long long off[7] = {};
	DevBuf<float> X, H1, H2, Y, T, D3, D2, D1, loss_partial, split;
	DevBuf<double> scales, scale_partial, phase, gate_partial;
	DevBuf<short> Xb, W1t, W2t, W3t;
	DevBuf<float> Xall, Yall;  // inference features / outputs of every node
	double gate = 0.0;         // least-squares scale of the correction on the latest unseen snapshot
	int adam_t = 0;
	uint64_t rng = 20261202ULL;

	float* P(int i) { return params.data() + off[i]; }
	float* G(int i) { return grads.data() + off[i]; }

	void Scales(const Hist& H)
	{
		const int blocks = 128;
		hipLaunchKernelGGL(ScalePartialKernel, dim3(blocks), dim3(kT), 0, ctx->stream, mesh->nodes, H,
			scale_partial.data());
		TFI_LAUNCH_CHECK("ScalePartialKernel");
		std::vector<double> part(kScales*blocks), s(kScales, 0.0);
		scale_partial.Download(part.data(), part.size());
		for (int b = 0; b < blocks; ++b)
			for (int f = 0; f < kScales; ++f) s[f] += part[kScales*b + f];
		for (int f = 0; f < kScales; ++f) {
			s[f] = std::sqrt(s[f]/std::max(1, mesh->nodes));
			if (!(s[f] > 1e-300)) s[f] = 1.0;
		}
		scales.Upload(s.data(), kScales);
	}

	void Forward(int count, const float* x)
	{
		Gemm(ctx, {count, H, kIn, x, kIn, 1, P(0), H, 1, H1.data(), H, P(1), nullptr, 0, 1});
		Gemm(ctx, {count, H, H, H1.data(), H, 1, P(2), H, 1, H2.data(), H, P(3), nullptr, 0, 1});
		Gemm(ctx, {count, kOut, H, H2.data(), H, 1, P(4), kOut, 1, Y.data(), kOut, P(5), nullptr, 0, 2});
	}

	// gradients of the mean squared error for `count` rows (targets in T); the
	// loss is read back only when asked for (it costs a stream synchronisation)
	float Backward(int count, const float* x, bool want_loss = true)
	{
		const int lb = 64;
		hipLaunchKernelGGL(LossKernel, dim3(lb), dim3(kT), 0, ctx->stream, count, Y.data(), T.data(), D3.data(),
			loss_partial.data());
		TFI_LAUNCH_CHECK("LossKernel");
		float* sc = split.data();
		// dW3 = H2^T D3 ; db3
		Gemm(ctx, {H, kOut, count, H2.data(), 1, H, D3.data(), kOut, 1, G(4), kOut, nullptr, nullptr, 0, 0}, sc);
		ColumnSum(ctx, count, kOut, D3.data(), G(5), sc);
		// D2 = (D3 W3^T) * (1 - H2^2)
		Gemm(ctx, {count, H, kOut, D3.data(), kOut, 1, P(4), 1, kOut, D2.data(), H, nullptr, H2.data(), H, 3});
		Gemm(ctx, {H, H, count, H1.data(), 1, H, D2.data(), H, 1, G(2), H, nullptr, nullptr, 0, 0}, sc);
		ColumnSum(ctx, count, H, D2.data(), G(3), sc);
		// D1 = (D2 W2^T) * (1 - H1^2)
		Gemm(ctx, {count, H, H, D2.data(), H, 1, P(2), 1, H, D1.data(), H, nullptr, H1.data(), H, 3});
		Gemm(ctx, {kIn, H, count, x, 1, kIn, D1.data(), H, 1, G(0), H, nullptr, nullptr, 0, 0}, sc);
		ColumnSum(ctx, count, H, D1.data(), G(1), sc);
		if (!want_loss) return 0.0f;
		std::vector<float> part(lb);
		loss_partial.Download(part.data(), lb);
		float l = 0.0f;
		for (float v : part) l += v;
		return l;
	}

	void Allocate(int rows_max)
	{
		X.Allocate(ctx, (size_t)rows_max*kIn);
		H1.Allocate(ctx, (size_t)rows_max*H);
		H2.Allocate(ctx, (size_t)rows_max*H);
		Y.Allocate(ctx, (size_t)rows_max*kOut);
		T.Allocate(ctx, (size_t)rows_max*kOut);
		D3.Allocate(ctx, (size_t)rows_max*kOut);
		D2.Allocate(ctx, (size_t)rows_max*H);
		D1.Allocate(ctx, (size_t)rows_max*H);
		loss_partial.Allocate(ctx, 64);
		split.Allocate(ctx, (size_t)kSplitMax*std::max<long long>((long long)kIn*H, (long long)H*H));
	}

	void Layout()
	{
		const long long sizes[6] = {(long long)kIn*H, H, (long long)H*H, H, (long long)H*kOut, kOut};
		off[0] = 0;
		for (int i = 0; i < 6; ++i) off[i + 1] = off[i] + sizes[i];
		params.Allocate(ctx, off[6]);
		grads.Allocate(ctx, off[6]);
	}
};

Mlp::Mlp(tfi_mesh* mesh, const tfi_warm_options& options) : impl_(std::make_unique<Impl>())
{
	Impl& I = *impl_;
	I.mesh = mesh;
	I.ctx = mesh->ctx;
	I.opt = options;
	I.H = options.hidden > 0 ? options.hidden : 128;
	if (I.H%16 != 0 || I.H > 256) throw Error("MLP width must be a multiple of 16 and at most 256");
	if (options.stencil != kStencil) throw Error("the MLP stencil has 27 control points");
	const int N = mesh->nodes;
	// control-point positions: extraction-weighted average of the Bezier points,
	// per basis entry on the device, then per node over its entries in order
	const long long entries = mesh->basis_entries;
	std::vector<long long> noff(N + 1, 0);
	for (long long a = 0; a < entries; ++a) ++noff[mesh->h_conn[a] + 1];
	for (int n = 0; n < N; ++n) noff[n + 1] += noff[n];
	std::vector<long long> nent(entries);
	{
		std::vector<long long> fill(noff.begin(), noff.end() - 1);
		for (long long a = 0; a < entries; ++a) nent[fill[mesh->h_conn[a]]++] = a;
	}
	{
		DevBuf<double> entry(I.ctx, 4*(size_t)entries);
		DevBuf<long long> d_noff(I.ctx, N + 1), d_nent(I.ctx, entries);
		d_noff.Upload(noff.data(), N + 1);
		d_nent.Upload(nent.data(), entries);
		hipLaunchKernelGGL(EntryPositionKernel, dim3(mesh->elements), dim3(64), 0, I.ctx->stream, mesh->elements,
			mesh->eoff.data(), mesh->xoff.data(), mesh->xcol.data(), mesh->xval.data(), mesh->bezier.data(),
			entry.data());
		TFI_LAUNCH_CHECK("EntryPositionKernel");
		I.pos.Allocate(I.ctx, 3*(size_t)N);
		hipLaunchKernelGGL(NodePositionKernel, dim3((N + kT - 1)/kT), dim3(kT), 0, I.ctx->stream, N, d_noff.data(),
			d_nent.data(), entry.data(), I.pos.data());
		TFI_LAUNCH_CHECK("NodePositionKernel");
		Sync(I.ctx);
	}
	// 27-point stencils: the node and its 26 nearest pattern neighbours
	I.stencil.Allocate(I.ctx, (size_t)N*kStencil);
	I.hscale.Allocate(I.ctx, N);
	{
		DevBuf<int> status(I.ctx, 1);
		FillInt(I.ctx, status.data(), 0, 1);
		hipLaunchKernelGGL(StencilKernel, dim3((N + 3)/4), dim3(256), 0, I.ctx->stream, N, mesh->row_offsets.data(),
			mesh->columns.data(), I.pos.data(), I.stencil.data(), I.hscale.data(), status.data());
		TFI_LAUNCH_CHECK("StencilKernel");
		int st = 0;
		status.Download(&st, 1);
		if (st) throw Error("a pattern row has more blocks than the stencil search takes (512)");
	}
	I.batch = std::min(4096, N);
	I.rows.Allocate(I.ctx, I.batch);
	I.Layout();
	// deterministic Xavier-uniform initialisation; small last layer
	std::vector<float> p(I.off[6], 0.0f);
	uint64_t s = 1234567ULL;
	auto uni = [&]() { return static_cast<float>((SplitMix(s) >> 11)*(1.0/9007199254740992.0))*2.0f - 1.0f; };
	const float a1 = std::sqrt(6.0f/(kIn + I.H)), a2 = std::sqrt(6.0f/(2*I.H)), a3 = 0.01f*std::sqrt(6.0f/(I.H + kOut));
	for (long long i = I.off[0]; i < I.off[1]; ++i) p[i] = a1*uni();
	for (long long i = I.off[2]; i < I.off[3]; ++i) p[i] = a2*uni();
	for (long long i = I.off[4]; i < I.off[5]; ++i) p[i] = a3*uni();
	I.params.Upload(p.data(), p.size());
	I.adam_m.Allocate(I.ctx, I.off[6]);
	I.adam_v.Allocate(I.ctx, I.off[6]);
	std::vector<float> z(I.off[6], 0.0f);
	I.adam_m.Upload(z.data(), z.size());
	I.adam_v.Upload(z.data(), z.size());
	I.Allocate(I.batch);
	I.scales.Allocate(I.ctx, kScales);
	I.scale_partial.Allocate(I.ctx, kScales*128);
	I.phase.Allocate(I.ctx, 8);
}

Mlp::~Mlp() = default;

double Mlp::gate() const
{
	return impl_->gate;
}

void Mlp::Train(const double* next, const double* const* history, const double* base, const double* phase)
{
	Impl& I = *impl_;
	tfi_ctx* ctx = I.ctx;
	const int N = I.mesh->nodes;
	Hist H;
	for (int j = 0; j <= kOrder; ++j) H.h[j] = history[j];
	I.Scales(H);
	I.phase.Upload(phase, 8);
	// Gate: before it trains on this snapshot, the network predicts a batch of
	// it; the scale g minimising ||T - g Y|| (clamped to [0, 1]) is what the
	// next prediction applies, so a correction that does not help is dropped
	// and the prediction falls back to the extrapolation base.
	{
		hipLaunchKernelGGL(BatchRowsKernel, dim3((I.batch + kT - 1)/kT), dim3(kT), 0, ctx->stream, I.batch, N,
			static_cast<unsigned long long>(SplitMix(I.rng)), I.rows.data());
		TFI_LAUNCH_CHECK("BatchRowsKernel");
		const long long items = (long long)I.batch*kIn;
		hipLaunchKernelGGL(FeatureKernel, dim3((items + kT - 1)/kT), dim3(kT), 0, ctx->stream, I.batch, I.rows.data(),
			I.stencil.data(), I.pos.data(), I.hscale.data(), H, I.scales.data(), I.phase.data(), I.X.data());
		TFI_LAUNCH_CHECK("FeatureKernel");
		hipLaunchKernelGGL(TargetKernel, dim3((I.batch*kOut + kT - 1)/kT), dim3(kT), 0, ctx->stream, I.batch,
			I.rows.data(), next, H, base, I.scales.data(), I.T.data());
		TFI_LAUNCH_CHECK("TargetKernel");
		I.Forward(I.batch, I.X.data());
		constexpr int kGateBlocks = 32;
		if (I.gate_partial.size() < 2*kGateBlocks) I.gate_partial.Allocate(ctx, 2*kGateBlocks);
		hipLaunchKernelGGL(GateKernel, dim3(kGateBlocks), dim3(kT), 0, ctx->stream, I.batch, I.Y.data(), I.T.data(),
			I.gate_partial.data());
		TFI_LAUNCH_CHECK("GateKernel");
		double part[2*kGateBlocks], yt = 0.0, yy = 0.0;
		I.gate_partial.Download(part, 2*kGateBlocks);
		for (int b = 0; b < kGateBlocks; ++b) {
			yt += part[2*b];
			yy += part[2*b + 1];
		}
		I.gate = yy > 0.0 ? std::min(1.0, std::max(0.0, yt/yy)) : 0.0;
	}
	const int steps = std::max(1, I.opt.train_steps);
	float loss = 0.0f;
	for (int step = 0; step < steps; ++step) {
		hipLaunchKernelGGL(BatchRowsKernel, dim3((I.batch + kT - 1)/kT), dim3(kT), 0, ctx->stream, I.batch, N,
			static_cast<unsigned long long>(SplitMix(I.rng)), I.rows.data());
		TFI_LAUNCH_CHECK("BatchRowsKernel");
		const long long items = (long long)I.batch*kIn;
		hipLaunchKernelGGL(FeatureKernel, dim3((items + kT - 1)/kT), dim3(kT), 0, ctx->stream, I.batch, I.rows.data(),
			I.stencil.data(), I.pos.data(), I.hscale.data(), H, I.scales.data(), I.phase.data(), I.X.data());
		TFI_LAUNCH_CHECK("FeatureKernel");
		hipLaunchKernelGGL(TargetKernel, dim3((I.batch*kOut + kT - 1)/kT), dim3(kT), 0, ctx->stream, I.batch,
			I.rows.data(), next, H, base, I.scales.data(), I.T.data());
		TFI_LAUNCH_CHECK("TargetKernel");
		I.Forward(I.batch, I.X.data());
		loss = I.Backward(I.batch, I.X.data(), step + 1 == steps);
		++I.adam_t;
		const float b1 = 0.9f, b2 = 0.999f;
		const float c1 = 1.0f - std::pow(b1, static_cast<float>(I.adam_t));
		const float c2 = 1.0f - std::pow(b2, static_cast<float>(I.adam_t));
		const int n = static_cast<int>(I.off[6]);
		hipLaunchKernelGGL(AdamKernel, dim3((n + kT - 1)/kT), dim3(kT), 0, ctx->stream, n, I.params.data(),
			I.grads.data(), I.adam_m.data(), I.adam_v.data(), static_cast<float>(I.opt.learning_rate), b1, b2, c1, c2,
			1e-8f);
		TFI_LAUNCH_CHECK("AdamKernel");
	}
	last_loss_ = loss;
	++snapshots_;
}

void Mlp::Predict(const double* const* history, const double* base, const double* phase, double* state)
{
	Impl& I = *impl_;
	tfi_ctx* ctx = I.ctx;
	const int N = I.mesh->nodes;
	Hist H;
	for (int j = 0; j <= kOrder; ++j) H.h[j] = history[j];
	I.Scales(H);
	I.phase.Upload(phase, 8);
	if (I.Xb.size() < (size_t)N*kIn) {
		I.Xb.Allocate(ctx, (size_t)N*kIn);
		I.W1t.Allocate(ctx, (size_t)kIn*I.H);
		I.W2t.Allocate(ctx, (size_t)I.H*I.H);
		I.W3t.Allocate(ctx, (size_t)I.H*kOut);
	}
	if (I.Xall.size() < (size_t)N*kIn) {
		I.Xall.Allocate(ctx, (size_t)N*kIn);
		I.Yall.Allocate(ctx, (size_t)N*kOut);
	}
	float* X = I.Xall.data();
	float* Y = I.Yall.data();
	const long long items = (long long)N*kIn;
	hipLaunchKernelGGL(FeatureKernel, dim3((items + kT - 1)/kT), dim3(kT), 0, ctx->stream, N, nullptr,
		I.stencil.data(), I.pos.data(), I.hscale.data(), H, I.scales.data(), I.phase.data(), X);
	TFI_LAUNCH_CHECK("FeatureKernel");
	hipLaunchKernelGGL(ToBf16Kernel, dim3(4096), dim3(kT), 0, ctx->stream, items, X, I.Xb.data());
	hipLaunchKernelGGL(TransposeBf16Kernel, dim3((kIn*I.H + kT - 1)/kT), dim3(kT), 0, ctx->stream, kIn, I.H, I.P(0),
		I.W1t.data());
	hipLaunchKernelGGL(TransposeBf16Kernel, dim3((I.H*I.H + kT - 1)/kT), dim3(kT), 0, ctx->stream, I.H, I.H, I.P(2),
		I.W2t.data());
	hipLaunchKernelGGL(TransposeBf16Kernel, dim3((I.H*kOut + kT - 1)/kT), dim3(kT), 0, ctx->stream, I.H, kOut, I.P(4),
		I.W3t.data());
	TFI_LAUNCH_CHECK("TransposeBf16Kernel");
	const int blocks = (N + 63)/64;
	if (I.H == 128) {
		hipLaunchKernelGGL(MlpInferKernel<128>, dim3(blocks), dim3(256), 0, ctx->stream, N, I.Xb.data(), I.W1t.data(),
			I.P(1), I.W2t.data(), I.P(3), I.W3t.data(), I.P(5), Y);
	} else if (I.H == 64) {
		hipLaunchKernelGGL(MlpInferKernel<64>, dim3(blocks), dim3(256), 0, ctx->stream, N, I.Xb.data(), I.W1t.data(),
			I.P(1), I.W2t.data(), I.P(3), I.W3t.data(), I.P(5), Y);
	} else {
		throw Error("BF16 inference kernels are built for widths 64 and 128");
	}
	TFI_LAUNCH_CHECK("MlpInferKernel");
	hipLaunchKernelGGL(ApplyPredictionKernel, dim3((4*N + kT - 1)/kT), dim3(kT), 0, ctx->stream, N, H, base, Y,
		I.scales.data(), I.gate, state);
	TFI_LAUNCH_CHECK("ApplyPredictionKernel");
	Sync(ctx);
}

// ------------------------------------------------------------------ test hooks
void Mlp::Reference(tfi_ctx* ctx, int hidden, int batch, const float* x, const float* target, const Weights& w,
	float* out, float* loss, Weights& grad)
{
	Impl I;
	I.ctx = ctx;
	I.mesh = nullptr;
	I.H = hidden;
	I.Layout();
	std::vector<float> p(I.off[6]);
	const std::vector<float>* parts[6] = {&w.w1, &w.b1, &w.w2, &w.b2, &w.w3, &w.b3};
	for (int i = 0; i < 6; ++i) {
		if ((long long)parts[i]->size() != I.off[i + 1] - I.off[i]) throw Error("MLP weight size mismatch");
		std::copy(parts[i]->begin(), parts[i]->end(), p.begin() + I.off[i]);
	}
	I.params.Upload(p.data(), p.size());
	I.Allocate(batch);
	DevBuf<float> xd(ctx, (size_t)batch*kIn);
	xd.Upload(x, (size_t)batch*kIn);
	I.T.Upload(target, (size_t)batch*kOut);
	I.Forward(batch, xd.data());
	*loss = I.Backward(batch, xd.data());
	I.Y.Download(out, (size_t)batch*kOut);
	std::vector<float> g(I.off[6]);
	I.grads.Download(g.data(), g.size());
	std::vector<float>* gp[6] = {&grad.w1, &grad.b1, &grad.w2, &grad.b2, &grad.w3, &grad.b3};
	for (int i = 0; i < 6; ++i) gp[i]->assign(g.begin() + I.off[i], g.begin() + I.off[i + 1]);
}

void Mlp::InferBf16(tfi_ctx* ctx, int hidden, int rows, const float* x, const Weights& w, float* out)
{
	if (hidden != 64 && hidden != 128) throw Error("BF16 inference kernels are built for widths 64 and 128");
	DevBuf<float> xd(ctx, (size_t)rows*kIn), w1(ctx, w.w1.size()), b1(ctx, w.b1.size()), w2(ctx, w.w2.size()),
		b2(ctx, w.b2.size()), w3(ctx, w.w3.size()), b3(ctx, w.b3.size()), y(ctx, (size_t)rows*kOut);
	xd.Upload(x, (size_t)rows*kIn);
	w1.Upload(w.w1.data(), w.w1.size());
	b1.Upload(w.b1.data(), w.b1.size());
	w2.Upload(w.w2.data(), w.w2.size());
	b2.Upload(w.b2.data(), w.b2.size());
	w3.Upload(w.w3.data(), w.w3.size());
	b3.Upload(w.b3.data(), w.b3.size());
	DevBuf<short> xb(ctx, (size_t)rows*kIn), t1(ctx, (size_t)kIn*hidden), t2(ctx, (size_t)hidden*hidden),
		t3(ctx, (size_t)hidden*kOut);
	hipLaunchKernelGGL(ToBf16Kernel, dim3(1024), dim3(kT), 0, ctx->stream, (long long)rows*kIn, xd.data(), xb.data());
	hipLaunchKernelGGL(TransposeBf16Kernel, dim3((kIn*hidden + kT - 1)/kT), dim3(kT), 0, ctx->stream, kIn, hidden,
		w1.data(), t1.data());
	hipLaunchKernelGGL(TransposeBf16Kernel, dim3((hidden*hidden + kT - 1)/kT), dim3(kT), 0, ctx->stream, hidden, hidden,
		w2.data(), t2.data());
	hipLaunchKernelGGL(TransposeBf16Kernel, dim3((hidden*kOut + kT - 1)/kT), dim3(kT), 0, ctx->stream, hidden, kOut,
		w3.data(), t3.data());
	const int blocks = (rows + 63)/64;
	if (hidden == 128)
		hipLaunchKernelGGL(MlpInferKernel<128>, dim3(blocks), dim3(256), 0, ctx->stream, rows, xb.data(), t1.data(),
			b1.data(), t2.data(), b2.data(), t3.data(), b3.data(), y.data());
	else
		hipLaunchKernelGGL(MlpInferKernel<64>, dim3(blocks), dim3(256), 0, ctx->stream, rows, xb.data(), t1.data(),
			b1.data(), t2.data(), b2.data(), t3.data(), b3.data(), y.data());
	TFI_LAUNCH_CHECK("MlpInferKernel");
	y.Download(out, (size_t)rows*kOut);
}

}  // namespace tfi

using namespace tfi;

extern "C" {

// Test entry points (declared in include/tfi_test.h).
int tfi_mlp_reference(tfi_ctx* ctx, int hidden, int batch, const float* x, const float* target, const float* w1,
	const float* b1, const float* w2, const float* b2, const float* w3, const float* b3, float* out, float* loss,
	float* g_w1, float* g_b1, float* g_w2, float* g_b2, float* g_w3, float* g_b3)
{
	try {
		Mlp::Weights w, g;
		w.w1.assign(w1, w1 + (size_t)kIn*hidden);
		w.b1.assign(b1, b1 + hidden);
		w.w2.assign(w2, w2 + (size_t)hidden*hidden);
		w.b2.assign(b2, b2 + hidden);
		w.w3.assign(w3, w3 + (size_t)hidden*kOut);
		w.b3.assign(b3, b3 + kOut);
		Mlp::Reference(ctx, hidden, batch, x, target, w, out, loss, g);
		std::copy(g.w1.begin(), g.w1.end(), g_w1);
		std::copy(g.b1.begin(), g.b1.end(), g_b1);
		std::copy(g.w2.begin(), g.w2.end(), g_w2);
		std::copy(g.b2.begin(), g.b2.end(), g_b2);
		std::copy(g.w3.begin(), g.w3.end(), g_w3);
		std::copy(g.b3.begin(), g.b3.end(), g_b3);
		return TFI_OK;
	} catch (const std::exception& e) {
		ctx->error = e.what();
		return TFI_ERROR;
	}
}

int tfi_mlp_infer_bf16(tfi_ctx* ctx, int hidden, int rows, const float* x, const float* w1, const float* b1,
	const float* w2, const float* b2, const float* w3, const float* b3, float* out)
{
	try {
		Mlp::Weights w;
		w.w1.assign(w1, w1 + (size_t)kIn*hidden);
		w.b1.assign(b1, b1 + hidden);
		w.w2.assign(w2, w2 + (size_t)hidden*hidden);
		w.b2.assign(b2, b2 + hidden);
		w.w3.assign(w3, w3 + (size_t)hidden*kOut);
		w.b3.assign(b3, b3 + kOut);
		Mlp::InferBf16(ctx, hidden, rows, x, w, out);
		return TFI_OK;
	} catch (const std::exception& e) {
		ctx->error = e.what();
		return TFI_ERROR;
	}
}

}  // extern "C"
