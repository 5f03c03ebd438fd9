// Node-block (4x4) aggregation AMG preconditioner.
//
// The reference GPU path has no AMG (it preconditions GMRES with node-block
// Jacobi, solvers/cuda/include/Gmres.hpp:47-67); its CPU path uses PETSc
// FIELDSPLIT + GAMG (solvers/cpu/include/TransientFlowRuntime.hpp:176-215) and
// the GPU tet solver a scalar plain-aggregation AMG
// (solvers/cuda/include/AggregationAmg.hpp).  This is a coupled (all four
// fields per node) plain-aggregation hierarchy:
//   * symbolic phase (host, once per mesh, deterministic): strength graph from
//     the first matrix (field-diagonal couplings), greedy aggregation in node
//     order, coarse block pattern, and for every coarse block the ascending
//     list of fine blocks that sum into it;
//   * numeric phase (GPU, every Newton step): Galerkin coarse blocks as
//     fixed-order segmented sums (no atomics), 4x4 diagonal inverses, and a
//     dense Gauss-Jordan inverse of the coarsest operator;
//   * V-cycle: symmetric multicolour block Gauss-Seidel (greedy node colouring
//     per level, one launch per colour, in place -- rows of one colour are
//     uncoupled), damped block Jacobi with omega = 1.33 / lambda_max(D^-1 A),
//     or a Chebyshev polynomial in D^-1 A on [0.1, 1.1] lambda_max
//     (power-iteration estimate; one fused launch per degree, no colours);
//     restriction = aggregate sums, prolongation = injection; optionally
//     entirely in FP32.
// The prolongator is the 0/1 aggregate map, stored as integers, so it is exact
// in any precision.
#include "tfi_solver.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace tfi {

namespace {

constexpr double kStrength = 0.25;     // row-relative strength threshold
constexpr int kMaxLevels = 16;
constexpr double kOmegaScale = 1.33;   // block-Jacobi damping = kOmegaScale / lambda_max(D^-1 A)
constexpr int kPre = 2, kPost = 2;      // Jacobi sweeps / Chebyshev degree
constexpr double kChebLow = 0.1, kChebHigh = 1.1;  // Chebyshev interval, in units of lambda_max(D^-1 A)
constexpr int kPowerIterations = 12;

template <class T>
__device__ inline T WaveSumT(T v)
{
	for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
	return v;
}

// This lane's share of row `row` of A x for one block-row lane r: the wave
// splits the row's blocks over 16 groups of 4 lanes (lane r of a group takes
// block row r).  In FP32 each lane keeps four blocks in flight (column
// indices, then block rows and x segments as 16-byte vector loads) so that a
// row costs about one memory latency instead of one per block; in FP64 the
// plain loop measured faster (the four 32-byte loads in flight cost
// occupancy).
template <class T>
__device__ inline T RowPartial(const int* __restrict__ ro, const int* __restrict__ cols, const T* __restrict__ vals,
	const T* __restrict__ x, int row, int sub, int r)
{
	typedef T T4 __attribute__((ext_vector_type(4)));
	const int end = ro[row + 1];
	T acc = 0;
	if constexpr (sizeof(T) == 8) {
		for (int k = ro[row] + sub; k < end; k += 16) {
			const T* a = vals + (size_t)k*16 + 4*r;
			const T* xv = x + 4*(size_t)cols[k];
			acc += a[0]*xv[0] + a[1]*xv[1] + a[2]*xv[2] + a[3]*xv[3];
		}
		return acc;
	}
	constexpr int kD = 4;  // blocks in flight per lane (8 measured slower: register pressure)
	for (int k0 = ro[row] + sub; k0 < end; k0 += 16*kD) {
		int c[kD];
#pragma unroll
		for (int u = 0; u < kD; ++u) c[u] = k0 + 16*u < end ? cols[k0 + 16*u] : -1;
		T4 a[kD], xv[kD];
#pragma unroll
		for (int u = 0; u < kD; ++u)
			if (c[u] >= 0) {
				a[u] = *reinterpret_cast<const T4*>(vals + (size_t)(k0 + 16*u)*16 + 4*r);
				xv[u] = *reinterpret_cast<const T4*>(x + 4*(size_t)c[u]);
			}
#pragma unroll
		for (int u = 0; u < kD; ++u)
			if (c[u] >= 0) acc += a[u].x*xv[u].x + a[u].y*xv[u].y + a[u].z*xv[u].z + a[u].w*xv[u].w;
	}
	return acc;
}

// x_out = x_in + omega D^-1 (b - A x_in); x_in == nullptr means x_in = 0.
template <class T>
__global__ void __launch_bounds__(256) SmoothKernel(int rows, const int* __restrict__ ro, const int* __restrict__ cols,
	const T* __restrict__ vals, const T* __restrict__ inv, const T* __restrict__ b, const T* __restrict__ xin,
	T* __restrict__ xout, T omega)
{
	const int lane = threadIdx.x & 63;
	const int row = blockIdx.x*4 + (threadIdx.x >> 6);
	if (row >= rows) return;
	const int sub = lane >> 2, r = lane & 3;
	T acc = 0;
	if (xin) {
		acc = RowPartial(ro, cols, vals, xin, row, sub, r);
		acc += __shfl_xor(acc, 4, 64);
		acc += __shfl_xor(acc, 8, 64);
		acc += __shfl_xor(acc, 16, 64);
		acc += __shfl_xor(acc, 32, 64);
	}
	// lanes 0..3 hold row residual components after the butterfly
	const T res = b[4*(size_t)row + r] - acc;
	const T r0 = __shfl(res, 0, 64), r1 = __shfl(res, 1, 64), r2 = __shfl(res, 2, 64), r3 = __shfl(res, 3, 64);
	if (sub == 0) {
		const T* d = inv + (size_t)row*16 + 4*r;
		const T corr = d[0]*r0 + d[1]*r1 + d[2]*r2 + d[3]*r3;
		xout[4*(size_t)row + r] = (xin ? xin[4*(size_t)row + r] : T(0)) + omega*corr;
	}
}

// One Chebyshev step (three-term form): z = D^-1 (b - A x_in), d = cd d + cz z,
// x_out = x_in + d; x_in == nullptr means x_in = 0 (and A x_in = 0), cd == 0
// ignores the old d.  d is per row, so it is updated in place.
template <class T>
__global__ void __launch_bounds__(256) ChebyshevKernel(int rows, const int* __restrict__ ro,
	const int* __restrict__ cols, const T* __restrict__ vals, const T* __restrict__ inv, const T* __restrict__ b,
	const T* __restrict__ xin, T* __restrict__ xout, T* __restrict__ d, T cd, T cz)
{
	const int lane = threadIdx.x & 63;
	const int row = blockIdx.x*4 + (threadIdx.x >> 6);
	if (row >= rows) return;
	const int sub = lane >> 2, r = lane & 3;
	T acc = 0;
	if (xin) {
		acc = RowPartial(ro, cols, vals, xin, row, sub, r);
		acc += __shfl_xor(acc, 4, 64);
		acc += __shfl_xor(acc, 8, 64);
		acc += __shfl_xor(acc, 16, 64);
		acc += __shfl_xor(acc, 32, 64);
	}
	const T res = b[4*(size_t)row + r] - acc;
	const T r0 = __shfl(res, 0, 64), r1 = __shfl(res, 1, 64), r2 = __shfl(res, 2, 64), r3 = __shfl(res, 3, 64);
	if (sub == 0) {
		const size_t i = 4*(size_t)row + r;
		const T* dv = inv + (size_t)row*16 + 4*r;
		const T z = dv[0]*r0 + dv[1]*r1 + dv[2]*r2 + dv[3]*r3;
		const T dn = (cd != T(0) ? cd*d[i] : T(0)) + cz*z;
		d[i] = dn;
		xout[i] = (xin ? xin[i] : T(0)) + dn;
	}
}

// r = b - A x
template <class T>
__global__ void __launch_bounds__(256) ResidualKernel(int rows, const int* __restrict__ ro, const int* __restrict__ cols,
	const T* __restrict__ vals, const T* __restrict__ b, const T* __restrict__ x, T* __restrict__ r)
{
	const int lane = threadIdx.x & 63;
	const int row = blockIdx.x*4 + (threadIdx.x >> 6);
	if (row >= rows) return;
	const int sub = lane >> 2, rr = lane & 3;
	T acc = 0;
	acc = RowPartial(ro, cols, vals, x, row, sub, rr);
	acc += __shfl_xor(acc, 4, 64);
	acc += __shfl_xor(acc, 8, 64);
	acc += __shfl_xor(acc, 16, 64);
	acc += __shfl_xor(acc, 32, 64);
	if (sub == 0) r[4*(size_t)row + rr] = b[4*(size_t)row + rr] - acc;
}

// Mixed precision, Chebyshev levels: the smoother reads S = D^-1 A, the
// block-Jacobi-scaled matrix, in FP16.  Its diagonal blocks are the identity
// and its other entries of order one, so FP16's 11-bit significand and range
// hold it (a level whose S has an entry above 60000 keeps the FP32 matrix).
// Half the bytes of FP32 blocks per smoothing pass.
// (S x) of one block row, by one wave: a lane takes whole blocks (all four
// rows, 32 bytes of S as two 16-byte loads, the 64 lanes 64 blocks a pass),
// then a butterfly over the 64 lanes; lane l returns row l % 4.
__device__ inline float RowProductHalf(const int* __restrict__ ro, const int* __restrict__ cols,
	const _Float16* __restrict__ s, const float* __restrict__ x, int row, int lane)
{
	typedef float F4 __attribute__((ext_vector_type(4)));
	typedef _Float16 H8 __attribute__((ext_vector_type(8)));
	const int end = ro[row + 1];
	float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f, a3 = 0.0f;
	for (int k = ro[row] + lane; k < end; k += 64) {
		const H8* blk = reinterpret_cast<const H8*>(s + (size_t)k*16);
		const H8 lo = blk[0], hi = blk[1];  // rows 0, 1 | rows 2, 3
		const F4 xv = *reinterpret_cast<const F4*>(x + 4*(size_t)cols[k]);
		a0 += (float)lo[0]*xv.x + (float)lo[1]*xv.y + (float)lo[2]*xv.z + (float)lo[3]*xv.w;
		a1 += (float)lo[4]*xv.x + (float)lo[5]*xv.y + (float)lo[6]*xv.z + (float)lo[7]*xv.w;
		a2 += (float)hi[0]*xv.x + (float)hi[1]*xv.y + (float)hi[2]*xv.z + (float)hi[3]*xv.w;
		a3 += (float)hi[4]*xv.x + (float)hi[5]*xv.y + (float)hi[6]*xv.z + (float)hi[7]*xv.w;
	}
#pragma unroll
	for (int o = 1; o < 64; o <<= 1) {
		a0 += __shfl_xor(a0, o, 64);
		a1 += __shfl_xor(a1, o, 64);
		a2 += __shfl_xor(a2, o, 64);
		a3 += __shfl_xor(a3, o, 64);
	}
	const int r = lane & 3;
	return r == 0 ? a0 : r == 1 ? a1 : r == 2 ? a2 : a3;
}

// S = D^-1 A in FP16 (D^-1 from the FP64 block inverses) and D in FP32; flag
// set when an entry of S is not finite or above 60000 in magnitude.
__global__ void __launch_bounds__(256) ScaleBlocksKernel(int rows, const int* __restrict__ ro,
	const int* __restrict__ diag, const double* __restrict__ vals, const double* __restrict__ inv,
	_Float16* __restrict__ s16, float* __restrict__ dblock, int* __restrict__ flag)
{
	const int lane = threadIdx.x & 63;
	const int row = blockIdx.x*4 + (threadIdx.x >> 6);
	if (row >= rows) return;
	const int sub = lane >> 2, r = lane & 3;
	const double* iv = inv + (size_t)row*16 + 4*r;
	bool bad = false;
	for (int k = ro[row] + sub; k < ro[row + 1]; k += 16) {
		const double* a = vals + (size_t)k*16;
#pragma unroll
		for (int c = 0; c < 4; ++c) {
			const double v = iv[0]*a[c] + iv[1]*a[4 + c] + iv[2]*a[8 + c] + iv[3]*a[12 + c];
			if (!(fabs(v) <= 60000.0)) bad = true;
			s16[(size_t)k*16 + 4*r + c] = static_cast<_Float16>(v);
		}
	}
	if (sub == 0)
		for (int c = 0; c < 4; ++c)
			dblock[(size_t)row*16 + 4*r + c] = static_cast<float>(vals[(size_t)diag[row]*16 + 4*r + c]);
	if (bad) atomicOr(flag, 1);
}

// The Chebyshev step of ChebyshevKernel on S: z = D^-1 b - S x_in.
__global__ void __launch_bounds__(256) ChebyshevScaledKernel(int rows, const int* __restrict__ ro,
	const int* __restrict__ cols, const _Float16* __restrict__ s, const float* __restrict__ inv,
	const float* __restrict__ b, const float* __restrict__ xin, float* __restrict__ xout, float* __restrict__ d,
	float cd, float cz)
{
	const int lane = threadIdx.x & 63;
	const int row = blockIdx.x*4 + (threadIdx.x >> 6);
	if (row >= rows) return;
	const int sub = lane >> 2, r = lane & 3;
	float acc = 0.0f;
	if (xin) acc = RowProductHalf(ro, cols, s, xin, row, lane);
	if (sub == 0) {
		const size_t i = 4*(size_t)row + r;
		const float* iv = inv + (size_t)row*16 + 4*r;
		const float* bv = b + 4*(size_t)row;
		const float z = iv[0]*bv[0] + iv[1]*bv[1] + iv[2]*bv[2] + iv[3]*bv[3] - acc;
		const float dn = (cd != 0.0f ? cd*d[i] : 0.0f) + cz*z;
		d[i] = dn;
		xout[i] = (xin ? xin[i] : 0.0f) + dn;
	}
}

// r = b - A x = D (D^-1 b - S x).
__global__ void __launch_bounds__(256) ResidualScaledKernel(int rows, const int* __restrict__ ro,
	const int* __restrict__ cols, const _Float16* __restrict__ s, const float* __restrict__ inv,
	const float* __restrict__ dblock, const float* __restrict__ b, const float* __restrict__ x,
	float* __restrict__ rout)
{
	const int lane = threadIdx.x & 63;
	const int row = blockIdx.x*4 + (threadIdx.x >> 6);
	if (row >= rows) return;
	const int sub = lane >> 2, r = lane & 3;
	const float acc = RowProductHalf(ro, cols, s, x, row, lane);
	const float* iv = inv + (size_t)row*16 + 4*r;
	const float* bv = b + 4*(size_t)row;
	const float t = iv[0]*bv[0] + iv[1]*bv[1] + iv[2]*bv[2] + iv[3]*bv[3] - acc;
	const float t0 = __shfl(t, 0, 64), t1 = __shfl(t, 1, 64), t2 = __shfl(t, 2, 64), t3 = __shfl(t, 3, 64);
	if (sub == 0) {
		const float* dv = dblock + (size_t)row*16 + 4*r;
		rout[4*(size_t)row + r] = dv[0]*t0 + dv[1]*t1 + dv[2]*t2 + dv[3]*t3;
	}
}

// y = D^-1 (A x)  (power iteration for lambda_max of the Jacobi-preconditioned operator)
__global__ void __launch_bounds__(256) DinvAKernel(int rows, const int* __restrict__ ro, const int* __restrict__ cols,
	const double* __restrict__ vals, const double* __restrict__ inv, const double* __restrict__ x,
	double* __restrict__ y)
{
	const int lane = threadIdx.x & 63;
	const int row = blockIdx.x*4 + (threadIdx.x >> 6);
	if (row >= rows) return;
	const int sub = lane >> 2, r = lane & 3;
	double acc = 0.0;
	for (int k = ro[row] + sub; k < ro[row + 1]; k += 16) {
		const double* a = vals + (size_t)k*16 + 4*r;
		const double* xv = x + 4*(size_t)cols[k];
		acc += a[0]*xv[0] + a[1]*xv[1] + a[2]*xv[2] + a[3]*xv[3];
	}
	acc += __shfl_xor(acc, 4, 64);
	acc += __shfl_xor(acc, 8, 64);
	acc += __shfl_xor(acc, 16, 64);
	acc += __shfl_xor(acc, 32, 64);
	const double a0 = __shfl(acc, 0, 64), a1 = __shfl(acc, 1, 64), a2 = __shfl(acc, 2, 64), a3 = __shfl(acc, 3, 64);
	if (sub == 0) {
		const double* d = inv + (size_t)row*16 + 4*r;
		y[4*(size_t)row + r] = d[0]*a0 + d[1]*a1 + d[2]*a2 + d[3]*a3;
	}
}

// Deterministic pseudo-random start vector in [-1, 1).
__global__ void HashVectorKernel(long long n, double* __restrict__ v)
{
	for (long long i = blockIdx.x*(long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x*blockDim.x) {
		unsigned long long z = (unsigned long long)i*0x9e3779b97f4a7c15ULL + 0x632be59bd9b4e019ULL;
		z = (z ^ (z >> 30))*0xbf58476d1ce4e5b9ULL;
		z = (z ^ (z >> 27))*0x94d049bb133111ebULL;
		z ^= z >> 31;
		v[i] = (double)(z >> 11)*(2.0/9007199254740992.0) - 1.0;
	}
}

// One colour of a block Gauss-Seidel sweep, in place: for the rows of this
// colour (no two of them are coupled) x_i += D_i^-1 (b_i - sum_j A_ij x_j).
template <class T>
__global__ void __launch_bounds__(256) GsColorKernel(int count, const int* __restrict__ list, const int* __restrict__ ro,
	const int* __restrict__ cols, const T* __restrict__ vals, const T* __restrict__ inv, const T* __restrict__ b,
	T* __restrict__ x)
{
	const int lane = threadIdx.x & 63;
	const int item = blockIdx.x*4 + (threadIdx.x >> 6);
	if (item >= count) return;
	const int row = list[item];
	const int sub = lane >> 2, r = lane & 3;
	T acc = 0;
	acc = RowPartial(ro, cols, vals, static_cast<const T*>(x), row, sub, r);
	acc += __shfl_xor(acc, 4, 64);
	acc += __shfl_xor(acc, 8, 64);
	acc += __shfl_xor(acc, 16, 64);
	acc += __shfl_xor(acc, 32, 64);
	const T res = b[4*(size_t)row + r] - acc;
	const T r0 = __shfl(res, 0, 64), r1 = __shfl(res, 1, 64), r2 = __shfl(res, 2, 64), r3 = __shfl(res, 3, 64);
	if (sub == 0) {
		const T* d = inv + (size_t)row*16 + 4*r;
		x[4*(size_t)row + r] += d[0]*r0 + d[1]*r1 + d[2]*r2 + d[3]*r3;
	}
}

// bc[I] = sum_{i in aggregate I} r[i] (ascending fine order)
template <class T>
__global__ void RestrictKernel(int coarse, const int* __restrict__ aoff, const int* __restrict__ amem,
	const T* __restrict__ r, T* __restrict__ bc)
{
	const int t = blockIdx.x*blockDim.x + threadIdx.x;
	const int I = t >> 2, c = t & 3;
	if (I >= coarse) return;
	T s = 0;
	for (int k = aoff[I]; k < aoff[I + 1]; ++k) s += r[4*(size_t)amem[k] + c];
	bc[4*(size_t)I + c] = s;
}

// x[i] += xc[agg[i]]
template <class T>
__global__ void ProlongKernel(int fine, const int* __restrict__ agg, const T* __restrict__ xc, T* __restrict__ x)
{
	const int t = blockIdx.x*blockDim.x + threadIdx.x;
	const int i = t >> 2, c = t & 3;
	if (i >= fine) return;
	x[4*(size_t)i + c] += xc[4*(size_t)agg[i] + c];
}

// coarse[K] = sum of fine blocks list[off[K]..off[K+1]) (16 lanes per block)
__global__ void GalerkinKernel(long long cblocks, const long long* __restrict__ off, const int* __restrict__ list,
	const double* __restrict__ fine, double* __restrict__ coarse)
{
	const long long t = blockIdx.x*(long long)blockDim.x + threadIdx.x;
	const long long K = t >> 4;
	const int e = t & 15;
	if (K >= cblocks) return;
	double s = 0.0;
	for (long long k = off[K]; k < off[K + 1]; ++k) s += fine[(size_t)list[k]*16 + e];
	coarse[(size_t)K*16 + e] = s;
}

// Augmented [A | I] of the coarsest level (n = 4 rows, leading dimension 2n),
// on a zero-filled buffer: aug[(4i+a)*2n + 4j+b] = A_ij[a][b], identity on the right.
__global__ void DenseAugKernel(int rows, const int* __restrict__ ro, const int* __restrict__ cols,
	const double* __restrict__ vals, double* __restrict__ aug)
{
	const int i = blockIdx.x;
	const int n = 4*rows, w = 2*n;
	for (int k = ro[i]; k < ro[i + 1]; ++k)
		for (int e = threadIdx.x; e < 16; e += blockDim.x) {
			const int a = e >> 2, b = e & 3;
			aug[(size_t)(4*i + a)*w + 4*cols[k] + b] = vals[(size_t)k*16 + e];
		}
	if (threadIdx.x < 4) aug[(size_t)(4*i + threadIdx.x)*w + n + 4*i + threadIdx.x] = 1.0;
}

// Gauss-Jordan step p, part 1 (one workgroup): partial pivot search in column
// p over rows >= p (largest magnitude, lowest row on ties), row swap, scaling
// of the pivot row, and the elimination multipliers m[i] = aug[i][p] (i != p).
__global__ void __launch_bounds__(1024) PivotKernel(int n, int p, double* __restrict__ aug, double* __restrict__ mult,
	int* __restrict__ status)
{
	__shared__ double sval[1024];
	__shared__ int sidx[1024];
	const int w = 2*n;
	double best = -1.0;
	int bi = p;
	for (int i = p + threadIdx.x; i < n; i += blockDim.x) {
		const double v = fabs(aug[(size_t)i*w + p]);
		if (v > best) { best = v; bi = i; }
	}
	sval[threadIdx.x] = best;
	sidx[threadIdx.x] = bi;
	__syncthreads();
	for (int s = blockDim.x/2; s > 0; s >>= 1) {
		if (threadIdx.x < s) {
			const double o = sval[threadIdx.x + s];
			const int oi = sidx[threadIdx.x + s];
			if (o > sval[threadIdx.x] || (o == sval[threadIdx.x] && oi < sidx[threadIdx.x])) {
				sval[threadIdx.x] = o;
				sidx[threadIdx.x] = oi;
			}
		}
		__syncthreads();
	}
	const int q = sidx[0];
	if (threadIdx.x == 0 && (!(sval[0] > 1e-300) || !isfinite(sval[0]))) *status = 1;
	if (q != p)
		for (int j = threadIdx.x; j < w; j += blockDim.x) {
			const double t = aug[(size_t)p*w + j];
			aug[(size_t)p*w + j] = aug[(size_t)q*w + j];
			aug[(size_t)q*w + j] = t;
		}
	__syncthreads();
	const double piv = aug[(size_t)p*w + p];
	const double inv = piv != 0.0 ? 1.0/piv : 0.0;
	__syncthreads();
	for (int j = threadIdx.x; j < w; j += blockDim.x) aug[(size_t)p*w + j] *= inv;
	for (int i = threadIdx.x; i < n; i += blockDim.x) mult[i] = i == p ? 0.0 : aug[(size_t)i*w + p];
}

// Gauss-Jordan step p, part 2 (whole grid): aug[i][j] -= m[i] aug[p][j] for
// i != p; column p of the other rows becomes exactly zero.
__global__ void EliminateKernel(int n, int p, double* __restrict__ aug, const double* __restrict__ mult)
{
	const int w = 2*n;
	const long long t = blockIdx.x*(long long)blockDim.x + threadIdx.x;
	if (t >= (long long)n*w) return;
	const int i = static_cast<int>(t/w), j = static_cast<int>(t - (long long)i*w);
	if (i == p) return;
	aug[t] = j == p ? 0.0 : aug[t] - mult[i]*aug[(size_t)p*w + j];
}

// inverse (n x n) from the right half of the augmented matrix, in T
template <class T>
__global__ void ExtractInverseKernel(int n, const double* __restrict__ aug, T* __restrict__ inv)
{
	const long long t = blockIdx.x*(long long)blockDim.x + threadIdx.x;
	if (t >= (long long)n*n) return;
	const int i = static_cast<int>(t/n), j = static_cast<int>(t - (long long)i*n);
	inv[t] = static_cast<T>(aug[(size_t)i*2*n + n + j]);
}

// y = Dinv b (dense), one wave per row
template <class T>
__global__ void DenseApplyKernel(int n, const T* __restrict__ inv, const T* __restrict__ b, T* __restrict__ y)
{
	const int row = blockIdx.x*4 + (threadIdx.x >> 6);
	const int lane = threadIdx.x & 63;
	if (row >= n) return;
	T acc = 0;
	for (int j = lane; j < n; j += 64) acc += inv[(size_t)row*n + j]*b[j];
	acc = WaveSumT(acc);
	if (lane == 0) y[row] = acc;
}

__global__ void ToFloatKernel2(long long n, const double* __restrict__ x, float* __restrict__ y)
{
	for (long long i = blockIdx.x*(long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x*blockDim.x)
		y[i] = static_cast<float>(x[i]);
}

__global__ void ToDoubleKernel(long long n, const float* __restrict__ x, double* __restrict__ y)
{
	for (long long i = blockIdx.x*(long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x*blockDim.x)
		y[i] = static_cast<double>(x[i]);
}

__global__ void BlockInverseT32Kernel(long long n, const double* __restrict__ x, float* __restrict__ y)
{
	for (long long i = blockIdx.x*(long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x*blockDim.x)
		y[i] = static_cast<float>(x[i]);
}

int Grid(long long threads, int block = 256)
{
	long long b = (threads + block - 1)/block;
	return static_cast<int>(std::max<long long>(1, std::min<long long>(b, 1LL << 30)));
}

}  // namespace

struct Level {
	int rows = 0;
	long long nnzb = 0;
	// pattern (level 0 borrows the mesh pattern)
	DevBuf<int> ro_own, cols_own, diag_own;
	const int* ro = nullptr;
	const int* cols = nullptr;
	const int* diag = nullptr;
	std::vector<int> h_ro, h_cols, h_diag;
	DevBuf<double> vals;          // own values (levels >= 1)
	const double* vals64 = nullptr;
	DevBuf<double> inv64;
	DevBuf<float> vals32_own, inv32;
	const float* vals32 = nullptr;
	// mixed precision, Chebyshev levels: S = D^-1 A in FP16 and D in FP32 (scaled: in use)
	DevBuf<_Float16> s16;
	DevBuf<float> dblock32;
	bool scaled = false;
	// transfer to the next level
	DevBuf<int> agg, aoff, amem;
	DevBuf<long long> goff;       // Galerkin lists for the NEXT level's blocks
	DevBuf<int> glist;
	// multicolour Gauss-Seidel (gs) or damped block Jacobi; rows grouped by colour
	bool gs = false;
	std::vector<int> color_offsets;
	DevBuf<int> color_rows;
	double lambda_max = 1.0;
	// vectors
	DevBuf<double> x64, y64, b64, r64;
	DevBuf<float> x32, y32, b32, r32;
	BsrView view() const { return {rows, nnzb, ro, cols, diag}; }
};

struct Amg::Impl {
	tfi_mesh* mesh;
	tfi_ctx* ctx;
	int precision;
	AmgOptions opt;
	bool built = false;
	std::vector<std::unique_ptr<Level>> L;
	int nc = 0;  // coarsest dense size (4 * rows)
	DevBuf<double> aug, mult, dinv64, v64tmp;
	DevBuf<float> dinv32;
	DevBuf<int> status;
	DevBuf<int> scale_flags;  // per level: S did not fit FP16

	void Symbolic(const double* fine_values);
	void Numeric(const double* fine_values64, const float* fine_values32);
	void Colour(Level& v);
	template <class T>
	void Cycle(int l);
};

namespace {

// Strength of connection between block rows i and j: the largest coupling of
// one field with itself, scaled by the two field diagonals,
//   s_ij = max_f |A_ij[f][f]| / sqrt(|A_ii[f][f]| |A_jj[f][f]|),
// symmetrised (max of s_ij, s_ji) so that Dirichlet rows (identity velocity
// rows) stay connected through their pressure couplings.  j is strong for i
// when s_ij >= theta * max(rowmax_i, rowmax_j) -- a row-relative threshold.
// Greedy aggregation in node order (deterministic): roots whose strong
// neighbours are all free, then attach, then leftovers.
std::vector<int> Aggregate(int n, const std::vector<int>& ro, const std::vector<int>& cols,
	const std::vector<double>& vals, double theta, int& count)
{
	std::vector<double> dg(4LL*n, 0.0);
	for (int i = 0; i < n; ++i)
		for (int k = ro[i]; k < ro[i + 1]; ++k)
			if (cols[k] == i)
				for (int f = 0; f < 4; ++f) dg[4LL*i + f] = std::abs(vals[(size_t)k*16 + 5*f]);
	const long long nnz = ro[n];
	std::vector<double> s(nnz, 0.0);
	for (int i = 0; i < n; ++i)
		for (int k = ro[i]; k < ro[i + 1]; ++k) {
			const int j = cols[k];
			if (j == i) continue;
			double best = 0.0;
			for (int f = 0; f < 4; ++f) {
				const double d = dg[4LL*i + f]*dg[4LL*j + f];
				if (d > 0.0) best = std::max(best, std::abs(vals[(size_t)k*16 + 5*f])/std::sqrt(d));
			}
			s[k] = best;
		}
	// symmetrise through the (structurally symmetric) pattern
	std::vector<double> sym(s);
	for (int i = 0; i < n; ++i)
		for (int k = ro[i]; k < ro[i + 1]; ++k) {
			const int j = cols[k];
			if (j == i) continue;
			const int* b = cols.data() + ro[j];
			const int* e = cols.data() + ro[j + 1];
			const int* it = std::lower_bound(b, e, i);
			if (it != e && *it == i) sym[k] = std::max(s[k], s[it - cols.data()]);
		}
	std::vector<double> rowmax(n, 0.0);
	for (int i = 0; i < n; ++i)
		for (int k = ro[i]; k < ro[i + 1]; ++k) rowmax[i] = std::max(rowmax[i], sym[k]);
	std::vector<int> sro(n + 1, 0), scols;
	std::vector<double> sval;
	for (int i = 0; i < n; ++i) {
		for (int k = ro[i]; k < ro[i + 1]; ++k) {
			const int j = cols[k];
			if (j == i || sym[k] <= 0.0) continue;
			if (sym[k] >= theta*std::max(rowmax[i], rowmax[j])) {
				scols.push_back(j);
				sval.push_back(sym[k]);
			}
		}
		sro[i + 1] = static_cast<int>(scols.size());
	}
	std::vector<int> agg(n, -1);
	count = 0;
	for (int i = 0; i < n; ++i) {
		if (agg[i] >= 0 || sro[i + 1] == sro[i]) continue;
		bool free = true;
		for (int k = sro[i]; k < sro[i + 1] && free; ++k) free = agg[scols[k]] < 0;
		if (!free) continue;
		agg[i] = count;
		for (int k = sro[i]; k < sro[i + 1]; ++k) agg[scols[k]] = count;
		++count;
	}
	std::vector<int> pass2(n, -1);
	for (int i = 0; i < n; ++i) {
		if (agg[i] >= 0) continue;
		double best = -1.0;
		for (int k = sro[i]; k < sro[i + 1]; ++k)
			if (agg[scols[k]] >= 0 && sval[k] > best) {
				best = sval[k];
				pass2[i] = agg[scols[k]];
			}
	}
	for (int i = 0; i < n; ++i)
		if (agg[i] < 0 && pass2[i] >= 0) agg[i] = pass2[i];
	for (int i = 0; i < n; ++i) {
		if (agg[i] >= 0) continue;
		agg[i] = count;
		for (int k = sro[i]; k < sro[i + 1]; ++k)
			if (agg[scols[k]] < 0) agg[scols[k]] = count;
		++count;
	}
	return agg;
}

}  // namespace

// Greedy node colouring of a Gauss-Seidel level (rows of a colour are uncoupled).
void Amg::Impl::Colour(Level& v)
{
	std::vector<int> col(v.rows, -1), mark;
	int ncol = 0;
	for (int i = 0; i < v.rows; ++i) {
		if ((int)mark.size() < ncol + 1) mark.resize(ncol + 1, -1);
		for (int k = v.h_ro[i]; k < v.h_ro[i + 1]; ++k) {
			const int c = col[v.h_cols[k]];
			if (c >= 0) mark[c] = i;
		}
		int c = 0;
		while (c < ncol && mark[c] == i) ++c;
		if (c == ncol) ++ncol;
		col[i] = c;
	}
	v.color_offsets.assign(ncol + 1, 0);
	for (int i = 0; i < v.rows; ++i) ++v.color_offsets[col[i] + 1];
	for (int c = 0; c < ncol; ++c) v.color_offsets[c + 1] += v.color_offsets[c];
	std::vector<int> rows(v.rows);
	std::vector<int> fill(v.color_offsets.begin(), v.color_offsets.end() - 1);
	for (int i = 0; i < v.rows; ++i) rows[fill[col[i]]++] = i;
	v.color_rows.Allocate(ctx, v.rows);
	v.color_rows.Upload(rows.data(), v.rows);
}

void Amg::Impl::Symbolic(const double* fine_values)
{
	L.clear();
	auto l0 = std::make_unique<Level>();
	l0->rows = mesh->nodes;
	l0->nnzb = mesh->nnzb;
	l0->ro = mesh->row_offsets.data();
	l0->cols = mesh->columns.data();
	l0->diag = mesh->diagonal.data();
	l0->h_ro.resize(mesh->nodes + 1);
	l0->h_cols.resize(mesh->nnzb);
	l0->h_diag.resize(mesh->nodes);
	mesh->row_offsets.Download(l0->h_ro.data(), mesh->nodes + 1);
	mesh->columns.Download(l0->h_cols.data(), mesh->nnzb);
	mesh->diagonal.Download(l0->h_diag.data(), mesh->nodes);
	std::vector<double> hv(mesh->nnzb*16);
	Check(hipMemcpyAsync(hv.data(), fine_values, hv.size()*sizeof(double), hipMemcpyDeviceToHost, ctx->stream),
		"amg download");
	Sync(ctx);
	L.push_back(std::move(l0));
	while ((int)L.size() < kMaxLevels && L.back()->rows > opt.coarse_nodes) {
		Level& f = *L.back();
		int count = 0;
		auto agg = Aggregate(f.rows, f.h_ro, f.h_cols, hv, kStrength, count);
		if (count >= f.rows*0.9 || count < 1) break;  // coarsening stalled
		// aggregate membership lists
		std::vector<int> aoff(count + 1, 0), amem(f.rows);
		for (int i = 0; i < f.rows; ++i) ++aoff[agg[i] + 1];
		for (int c = 0; c < count; ++c) aoff[c + 1] += aoff[c];
		{
			std::vector<int> fill(aoff.begin(), aoff.end() - 1);
			for (int i = 0; i < f.rows; ++i) amem[fill[agg[i]]++] = i;
		}
		// coarse pattern
		auto c = std::make_unique<Level>();
		c->rows = count;
		std::vector<std::vector<int>> rows(count);
		for (int i = 0; i < f.rows; ++i) {
			auto& rw = rows[agg[i]];
			for (int k = f.h_ro[i]; k < f.h_ro[i + 1]; ++k) rw.push_back(agg[f.h_cols[k]]);
		}
		c->h_ro.assign(count + 1, 0);
		c->h_diag.assign(count, -1);
		for (int I = 0; I < count; ++I) {
			auto& rw = rows[I];
			std::sort(rw.begin(), rw.end());
			rw.erase(std::unique(rw.begin(), rw.end()), rw.end());
			c->h_ro[I + 1] = c->h_ro[I] + static_cast<int>(rw.size());
		}
		c->nnzb = c->h_ro[count];
		c->h_cols.resize(c->nnzb);
		for (int I = 0; I < count; ++I) {
			std::copy(rows[I].begin(), rows[I].end(), c->h_cols.begin() + c->h_ro[I]);
			const auto it = std::lower_bound(rows[I].begin(), rows[I].end(), I);
			c->h_diag[I] = c->h_ro[I] + static_cast<int>(it - rows[I].begin());
			std::vector<int>().swap(rows[I]);
		}
		// fine block -> coarse block lists (ascending fine block index)
		std::vector<long long> cnt(c->nnzb + 1, 0);
		std::vector<int> target(f.nnzb);
		for (int i = 0; i < f.rows; ++i) {
			const int I = agg[i];
			const int* b = c->h_cols.data() + c->h_ro[I];
			const int* e = c->h_cols.data() + c->h_ro[I + 1];
			for (int k = f.h_ro[i]; k < f.h_ro[i + 1]; ++k) {
				const int J = agg[f.h_cols[k]];
				const int pos = static_cast<int>(std::lower_bound(b, e, J) - c->h_cols.data());
				target[k] = pos;
				++cnt[pos + 1];
			}
		}
		for (long long K = 0; K < c->nnzb; ++K) cnt[K + 1] += cnt[K];
		std::vector<int> list(f.nnzb);
		{
			std::vector<long long> fill(cnt.begin(), cnt.end() - 1);
			for (long long k = 0; k < f.nnzb; ++k) list[fill[target[k]]++] = static_cast<int>(k);
		}
		f.agg.Allocate(ctx, f.rows);
		f.agg.Upload(agg.data(), f.rows);
		f.aoff.Allocate(ctx, count + 1);
		f.aoff.Upload(aoff.data(), count + 1);
		f.amem.Allocate(ctx, f.rows);
		f.amem.Upload(amem.data(), f.rows);
		f.goff.Allocate(ctx, c->nnzb + 1);
		f.goff.Upload(cnt.data(), c->nnzb + 1);
		f.glist.Allocate(ctx, f.nnzb);
		f.glist.Upload(list.data(), f.nnzb);
		c->ro_own.Allocate(ctx, count + 1);
		c->ro_own.Upload(c->h_ro.data(), count + 1);
		c->cols_own.Allocate(ctx, c->nnzb);
		c->cols_own.Upload(c->h_cols.data(), c->nnzb);
		c->diag_own.Allocate(ctx, count);
		c->diag_own.Upload(c->h_diag.data(), count);
		c->ro = c->ro_own.data();
		c->cols = c->cols_own.data();
		c->diag = c->diag_own.data();
		c->vals.Allocate(ctx, c->nnzb*16);
		// coarse values now (the next aggregation needs them)
		hipLaunchKernelGGL(GalerkinKernel, dim3(Grid(c->nnzb*16)), dim3(256), 0, ctx->stream, c->nnzb, f.goff.data(),
			f.glist.data(), L.size() == 1 ? fine_values : f.vals.data(), c->vals.data());
		TFI_LAUNCH_CHECK("GalerkinKernel");
		hv.assign(c->nnzb*16, 0.0);
		c->vals.Download(hv.data(), hv.size());
		L.push_back(std::move(c));
	}
	// per-level buffers; greedy node colouring of every Gauss-Seidel level
	for (size_t l = 0; l < L.size(); ++l) {
		Level& v = *L[l];
		v.gs = l + 1 < L.size() && opt.smoother == TFI_SMOOTH_GAUSS_SEIDEL && v.rows >= opt.gs_min_rows;
		if (v.gs) Colour(v);
		const long long n = 4LL*v.rows;
		v.inv64.Allocate(ctx, 16LL*v.rows);
		if (precision == TFI_MIXED) {
			v.inv32.Allocate(ctx, 16LL*v.rows);
			v.x32.Allocate(ctx, n);
			v.y32.Allocate(ctx, n);
			v.b32.Allocate(ctx, n);
			v.r32.Allocate(ctx, n);
			if (l > 0) v.vals32_own.Allocate(ctx, v.nnzb*16);
			if (opt.half_smoother && !v.gs && l + 1 < L.size() && opt.smoother == TFI_SMOOTH_CHEBYSHEV) {
				v.s16.Allocate(ctx, v.nnzb*16);
				v.dblock32.Allocate(ctx, 16LL*v.rows);
			}
		} else {
			v.x64.Allocate(ctx, n);
			v.y64.Allocate(ctx, n);
			v.b64.Allocate(ctx, n);
			v.r64.Allocate(ctx, n);
		}
	}
	nc = 4*L.back()->rows;
	if (nc > 1024) throw Error("AMG coarsest level too large for the dense solve");
	aug.Allocate(ctx, (size_t)nc*2*nc);
	mult.Allocate(ctx, nc);
	if (precision == TFI_MIXED) dinv32.Allocate(ctx, (size_t)nc*nc);
	else dinv64.Allocate(ctx, (size_t)nc*nc);
	status.Allocate(ctx, 1);
	scale_flags.Allocate(ctx, L.size());
	built = true;
}

void Amg::Impl::Numeric(const double* fine64, const float* fine32)
{
	L[0]->vals64 = fine64;
	L[0]->vals32 = fine32;
	for (size_t l = 0; l + 1 < L.size(); ++l) {
		Level& f = *L[l];
		Level& c = *L[l + 1];
		hipLaunchKernelGGL(GalerkinKernel, dim3(Grid(c.nnzb*16)), dim3(256), 0, ctx->stream, c.nnzb, f.goff.data(),
			f.glist.data(), l == 0 ? fine64 : f.vals.data(), c.vals.data());
		TFI_LAUNCH_CHECK("GalerkinKernel");
		c.vals64 = c.vals.data();
	}
	for (size_t l = 0; l < L.size(); ++l) {
		Level& v = *L[l];
		BlockInverse64(ctx, v.view(), v.vals64, v.inv64.data());
		if (precision == TFI_MIXED) {
			hipLaunchKernelGGL(BlockInverseT32Kernel, dim3(256), dim3(256), 0, ctx->stream, 16LL*v.rows,
				v.inv64.data(), v.inv32.data());
			TFI_LAUNCH_CHECK("BlockInverseT32Kernel");
			if (l > 0) {
				hipLaunchKernelGGL(ToFloatKernel2, dim3(256), dim3(256), 0, ctx->stream, v.nnzb*16, v.vals64,
					v.vals32_own.data());
				TFI_LAUNCH_CHECK("ToFloatKernel2");
				v.vals32 = v.vals32_own.data();
			}
		}
	}
	// FP16 copies of D^-1 A for the Chebyshev levels of the mixed path
	bool any_scaled = false;
	for (size_t l = 0; l < L.size(); ++l) any_scaled |= L[l]->s16.size() > 0 && !L[l]->gs;
	if (any_scaled) {
		FillInt(ctx, scale_flags.data(), 0, static_cast<long long>(L.size()));
		for (size_t l = 0; l < L.size(); ++l) {
			Level& v = *L[l];
			if (!(v.s16.size() > 0 && !v.gs)) continue;
			hipLaunchKernelGGL(ScaleBlocksKernel, dim3((v.rows + 3)/4), dim3(256), 0, ctx->stream, v.rows, v.ro,
				v.diag, v.vals64, v.inv64.data(), v.s16.data(), v.dblock32.data(), scale_flags.data() + l);
			TFI_LAUNCH_CHECK("ScaleBlocksKernel");
		}
		std::vector<int> flags(L.size(), 0);
		scale_flags.Download(flags.data(), flags.size());
		for (size_t l = 0; l < L.size(); ++l) L[l]->scaled = L[l]->s16.size() > 0 && !L[l]->gs && flags[l] == 0;
	}
	// lambda_max(D^-1 A) of every Jacobi-smoothed level by power iteration (FP64)
	for (size_t l = 0; l + 1 < L.size(); ++l) {
			Level& v = *L[l];
			if (v.gs) continue;
			const long long n = 4LL*v.rows;
			DevBuf<double> a(ctx, n), b(ctx, n);
			hipLaunchKernelGGL(HashVectorKernel, dim3(Grid(n)), dim3(256), 0, ctx->stream, n, a.data());
			double lam = 1.0;
			for (int it = 0; it < kPowerIterations; ++it) {
				const double nrm = Norm2(ctx, a.data(), n);
				Scale(ctx, 1.0/nrm, a.data(), n);
				hipLaunchKernelGGL(DinvAKernel, dim3((v.rows + 3)/4), dim3(256), 0, ctx->stream, v.rows, v.ro, v.cols,
					v.vals64, v.inv64.data(), a.data(), b.data());
				TFI_LAUNCH_CHECK("DinvAKernel");
				lam = Norm2(ctx, b.data(), n);
				std::swap(a, b);
			}
			v.lambda_max = lam > 0.0 ? lam : 1.0;
		}
	// coarsest dense inverse, built and inverted on the device
	Level& c = *L.back();
	Fill(ctx, aug.data(), 0.0, (long long)nc*2*nc);
	hipLaunchKernelGGL(DenseAugKernel, dim3(c.rows), dim3(64), 0, ctx->stream, c.rows, c.ro, c.cols, c.vals64,
		aug.data());
	TFI_LAUNCH_CHECK("DenseAugKernel");
	FillInt(ctx, status.data(), 0, 1);
	// n pivot steps, each a pivot/scale launch and a grid-wide elimination
	// launch (the single-workgroup form is bound by one CU's L2 bandwidth)
	const long long cells = (long long)nc*2*nc;
	for (int p = 0; p < nc; ++p) {
		hipLaunchKernelGGL(PivotKernel, dim3(1), dim3(1024), 0, ctx->stream, nc, p, aug.data(), mult.data(),
			status.data());
		hipLaunchKernelGGL(EliminateKernel, dim3(Grid(cells)), dim3(256), 0, ctx->stream, nc, p, aug.data(),
			mult.data());
	}
	TFI_LAUNCH_CHECK("EliminateKernel");
	int st = 0;
	status.Download(&st, 1);
	if (st) throw Error("AMG coarsest operator is singular");
	if (precision == TFI_MIXED) {
		hipLaunchKernelGGL(ExtractInverseKernel<float>, dim3(Grid((long long)nc*nc)), dim3(256), 0, ctx->stream, nc,
			aug.data(), dinv32.data());
	} else {
		hipLaunchKernelGGL(ExtractInverseKernel<double>, dim3(Grid((long long)nc*nc)), dim3(256), 0, ctx->stream, nc,
			aug.data(), dinv64.data());
	}
	TFI_LAUNCH_CHECK("ExtractInverseKernel");
}

template <class T>
void Amg::Impl::Cycle(int l)
{
	Level& v = *L[l];
	const bool f32 = std::is_same<T, float>::value;
	T* x = f32 ? reinterpret_cast<T*>(v.x32.data()) : reinterpret_cast<T*>(v.x64.data());
	T* y = f32 ? reinterpret_cast<T*>(v.y32.data()) : reinterpret_cast<T*>(v.y64.data());
	T* b = f32 ? reinterpret_cast<T*>(v.b32.data()) : reinterpret_cast<T*>(v.b64.data());
	T* r = f32 ? reinterpret_cast<T*>(v.r32.data()) : reinterpret_cast<T*>(v.r64.data());
	const T* vals = f32 ? reinterpret_cast<const T*>(v.vals32) : reinterpret_cast<const T*>(v.vals64);
	const T* inv = f32 ? reinterpret_cast<const T*>(v.inv32.data()) : reinterpret_cast<const T*>(v.inv64.data());
	const int grid = (v.rows + 3)/4;
	if (l + 1 == (int)L.size()) {
		const T* dinv = f32 ? reinterpret_cast<const T*>(dinv32.data()) : reinterpret_cast<const T*>(dinv64.data());
		hipLaunchKernelGGL(DenseApplyKernel<T>, dim3((nc + 3)/4), dim3(256), 0, ctx->stream, nc, dinv, b, x);
		TFI_LAUNCH_CHECK("DenseApplyKernel");
		return;
	}
	const T omega = static_cast<T>(kOmegaScale/v.lambda_max);
	const bool gs = v.gs;
	const bool cheb = !gs && opt.smoother == TFI_SMOOTH_CHEBYSHEV;
	// Chebyshev steps on [kChebLow, kChebHigh] lambda_max from x (x_in == nullptr: from 0);
	// d lives in the level's r buffer, which the smoothing phases do not otherwise use
	auto chebyshev = [&](int degree, bool from_zero) {
		const double hi = kChebHigh*v.lambda_max, lo = kChebLow*v.lambda_max;
		const double theta = 0.5*(hi + lo), delta = 0.5*(hi - lo), sigma = theta/delta;
		double rho = 1.0/sigma;
		for (int k = 0; k < degree; ++k) {
			double cd = 0.0, cz = 1.0/theta;
			if (k > 0) {
				const double rn = 1.0/(2.0*sigma - rho);
				cd = rn*rho;
				cz = 2.0*rn/delta;
				rho = rn;
			}
			bool done = false;
			if constexpr (std::is_same<T, float>::value) {
				if (v.scaled) {
					hipLaunchKernelGGL(ChebyshevScaledKernel, dim3(grid), dim3(256), 0, ctx->stream, v.rows, v.ro, v.cols,
						v.s16.data(), inv, b, (k == 0 && from_zero) ? static_cast<const float*>(nullptr) : x, y, r,
						static_cast<float>(cd), static_cast<float>(cz));
					done = true;
				}
			}
			if (!done)
				hipLaunchKernelGGL(ChebyshevKernel<T>, dim3(grid), dim3(256), 0, ctx->stream, v.rows, v.ro, v.cols, vals,
					inv, b, (k == 0 && from_zero) ? static_cast<const T*>(nullptr) : static_cast<const T*>(x), y, r,
					static_cast<T>(cd), static_cast<T>(cz));
			std::swap(x, y);
		}
		TFI_LAUNCH_CHECK("ChebyshevKernel");
	};
	const int ncol = static_cast<int>(v.color_offsets.size()) - 1;
	// after_forward: this backward sweep directly follows a forward one, whose
	// last colour it would repeat with nothing changed in between (a colour's
	// update depends only on the other colours), so that colour is skipped
	auto gs_sweep = [&](bool forward, bool after_forward = false) {
		bool skip = after_forward;
		for (int cc = 0; cc < ncol; ++cc) {
			const int c = forward ? cc : ncol - 1 - cc;
			const int count = v.color_offsets[c + 1] - v.color_offsets[c];
			if (count == 0) continue;
			if (skip) {
				skip = false;
				continue;
			}
			hipLaunchKernelGGL(GsColorKernel<T>, dim3((count + 3)/4), dim3(256), 0, ctx->stream, count,
				v.color_rows.data() + v.color_offsets[c], v.ro, v.cols, vals, inv, b, x);
		}
		TFI_LAUNCH_CHECK("GsColorKernel");
	};
	if (gs) {
		// symmetric multicolour Gauss-Seidel from x = 0
		const long long n = 4LL*v.rows;
		if (f32) Check(hipMemsetAsync(x, 0, n*sizeof(T), ctx->stream), "memset");
		else Fill(ctx, reinterpret_cast<double*>(x), 0.0, n);
		gs_sweep(true);
		if (opt.gs_sweeps > 1) gs_sweep(false, true);
	} else if (cheb) {
		chebyshev(opt.chebyshev_degree, true);
	} else {
		// damped block Jacobi from x = 0
		hipLaunchKernelGGL(SmoothKernel<T>, dim3(grid), dim3(256), 0, ctx->stream, v.rows, v.ro, v.cols, vals, inv, b,
			static_cast<const T*>(nullptr), x, omega);
		for (int s = 1; s < kPre; ++s) {
			hipLaunchKernelGGL(SmoothKernel<T>, dim3(grid), dim3(256), 0, ctx->stream, v.rows, v.ro, v.cols, vals, inv,
				b, x, y, omega);
			std::swap(x, y);
		}
	}
	bool residual_done = false;
	if constexpr (std::is_same<T, float>::value) {
		if (cheb && v.scaled) {
			hipLaunchKernelGGL(ResidualScaledKernel, dim3(grid), dim3(256), 0, ctx->stream, v.rows, v.ro, v.cols,
				v.s16.data(), inv, v.dblock32.data(), b, x, r);
			residual_done = true;
		}
	}
	if (!residual_done)
		hipLaunchKernelGGL(ResidualKernel<T>, dim3(grid), dim3(256), 0, ctx->stream, v.rows, v.ro, v.cols, vals, b, x, r);
	TFI_LAUNCH_CHECK("ResidualKernel");
	Level& c = *L[l + 1];
	T* bc = f32 ? reinterpret_cast<T*>(c.b32.data()) : reinterpret_cast<T*>(c.b64.data());
	T* xc = f32 ? reinterpret_cast<T*>(c.x32.data()) : reinterpret_cast<T*>(c.x64.data());
	hipLaunchKernelGGL(RestrictKernel<T>, dim3(Grid(4LL*c.rows)), dim3(256), 0, ctx->stream, c.rows, v.aoff.data(),
		v.amem.data(), r, bc);
	TFI_LAUNCH_CHECK("RestrictKernel");
	Cycle<T>(l + 1);
	// the coarse solution lives in c.x (cycle output); when the coarse cycle
	// ended with an odd number of swaps it is in c.y -- Cycle always leaves it in x
	hipLaunchKernelGGL(ProlongKernel<T>, dim3(Grid(4LL*v.rows)), dim3(256), 0, ctx->stream, v.rows, v.agg.data(), xc,
		x);
	TFI_LAUNCH_CHECK("ProlongKernel");
	if (gs) {
		if (opt.gs_sweeps > 1) gs_sweep(true);
		gs_sweep(false, opt.gs_sweeps > 1);
	} else if (cheb) {
		chebyshev(opt.chebyshev_degree, false);
	} else {
		for (int s = 0; s < kPost; ++s) {
			hipLaunchKernelGGL(SmoothKernel<T>, dim3(grid), dim3(256), 0, ctx->stream, v.rows, v.ro, v.cols, vals, inv,
				b, x, y, omega);
			std::swap(x, y);
		}
	}
	TFI_LAUNCH_CHECK("SmoothKernel");
	// leave the result in the level's x buffer
	T* xbuf = f32 ? reinterpret_cast<T*>(v.x32.data()) : reinterpret_cast<T*>(v.x64.data());
	if (x != xbuf) {
		const long long n = 4LL*v.rows;
		if (f32) {
			Check(hipMemcpyAsync(xbuf, x, n*sizeof(T), hipMemcpyDeviceToDevice, ctx->stream), "copy");
		} else {
			Copy(ctx, reinterpret_cast<const double*>(x), reinterpret_cast<double*>(xbuf), n);
		}
	}
}

Amg::Amg(tfi_mesh* mesh, int precision, const AmgOptions& options) : impl_(std::make_unique<Impl>())
{
	if (options.coarse_nodes < 1 || options.coarse_nodes > 256)
		throw Error("AMG coarsest level must hold between 1 and 256 block rows");
	impl_->mesh = mesh;
	impl_->ctx = mesh->ctx;
	impl_->precision = precision;
	impl_->opt = options;
}

void Amg::Describe(std::vector<int>& rows, std::vector<long long>& nnzb, std::vector<int>& colors) const
{
	rows.clear();
	nnzb.clear();
	colors.clear();
	for (const auto& l : impl_->L) {
		rows.push_back(l->rows);
		nnzb.push_back(l->nnzb);
		colors.push_back(l->gs ? static_cast<int>(l->color_offsets.size()) - 1 : 0);
	}
}

Amg::~Amg() = default;

int Amg::smoother() const { return impl_->opt.smoother; }

void Amg::UseGaussSeidel()
{
	Impl& I = *impl_;
	I.opt.smoother = TFI_SMOOTH_GAUSS_SEIDEL;
	if (!I.built) return;
	for (size_t l = 0; l + 1 < I.L.size(); ++l) {
		Level& v = *I.L[l];
		v.gs = v.rows >= I.opt.gs_min_rows;
		if (v.gs && v.color_offsets.empty()) I.Colour(v);
	}
}

int Amg::levels() const { return static_cast<int>(impl_->L.size()); }

void Amg::Setup(const double* values64, const float* values32)
{
	if (!impl_->built) impl_->Symbolic(values64);
	impl_->Numeric(values64, values32);
	singular_ = 0;
}

void Amg::Apply(const double* v, double* z)
{
	Impl& I = *impl_;
	Level& l0 = *I.L[0];
	const long long n = 4LL*l0.rows;
	if (I.precision == TFI_MIXED) {
		hipLaunchKernelGGL(ToFloatKernel2, dim3(Grid(n)), dim3(256), 0, I.ctx->stream, n, v, l0.b32.data());
		TFI_LAUNCH_CHECK("ToFloatKernel2");
		I.Cycle<float>(0);
		hipLaunchKernelGGL(ToDoubleKernel, dim3(Grid(n)), dim3(256), 0, I.ctx->stream, n, l0.x32.data(), z);
		TFI_LAUNCH_CHECK("ToDoubleKernel");
	} else {
		Copy(I.ctx, v, l0.b64.data(), n);
		I.Cycle<double>(0);
		Copy(I.ctx, l0.x64.data(), z, n);
	}
}

}  // namespace tfi
