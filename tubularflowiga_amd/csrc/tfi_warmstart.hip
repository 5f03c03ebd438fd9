// Initial-state prediction for a time step.  Modes:
//   TFI_WARM_PREVIOUS    the step starts from u_n (the reference behaviour,
//                        iga_cuda.cu:917-920);
//   TFI_WARM_EXTRAPOLATE u_n + (u_n - u_{n-1});
//   TFI_WARM_EXTRAPOLATE2 3 u_n - 3 u_{n-1} + u_{n-2} (quadratic in time);
//   TFI_WARM_EXTRAPOLATE3 4 u_n - 6 u_{n-1} + 4 u_{n-2} - u_{n-3} (cubic in time);
//   TFI_WARM_EXTRAPOLATE4 quartic in time;
//   TFI_WARM_MPE         minimal polynomial extrapolation of the increments
//                        d_k = u_k - u_{k-1}: the coefficients c of the
//                        least-squares fit d_n ~ sum_{i=1..m} c_i d_{n-i} (m up
//                        to kMpe) continue the recurrence one step,
//                        u_{n+1} = u_n + sum_i c_i d_{n+1-i}.  Near a steady or
//                        periodic state the increments are dominated by a few
//                        modes of the linearised step map, which such a
//                        recurrence reproduces exactly;
//   TFI_WARM_LEARNED     a base prediction (TFI_WARM_MPE or
//                        TFI_WARM_EXTRAPOLATE4) + MLP(stencil features), see
//                        tfi_mlp.hpp.
// A prediction is only a guess: the Newton loop accepts it when it lowers the
// nonlinear residual and always converges to the same tolerance.
#include "tfi_solver.hpp"
#include "tfi_mlp.hpp"

#include <cmath>

namespace tfi {

namespace {

constexpr int kMpe = 12;             // largest number of increments combined
constexpr double kMpeRank = 1e-10;   // an increment whose part outside the span of
                                     // the newer ones is below this fraction of
                                     // its norm ends the fit (numerical rank)
constexpr int kThreads = 256;
constexpr int kDotBlocks = 256;

// order-P polynomial extrapolation in time from u_n, .., u_{n-P}:
// sum_{j=0..P} (-1)^j C(P+1, j+1) u_{n-j}
struct History {
	const double* h[kMpe + 2];
};

__global__ void ExtrapolateKernel(long long n, int order, History H, double* __restrict__ out)
{
	for (long long i = blockIdx.x*(long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x*blockDim.x) {
		double c = static_cast<double>(order + 1), s = 0.0;
		for (int j = 0; j <= order; ++j) {
			s += c*H.h[j][i];
			c = -c*static_cast<double>(order - j)/static_cast<double>(j + 2);
		}
		out[i] = s;
	}
}

// D[j] = u_{n-j} - u_{n-j-1}, j = 0..count-1 (column j at D + j*ld)
__global__ void IncrementKernel(long long n, int count, History H, double* __restrict__ D, long long ld)
{
	for (long long i = blockIdx.x*(long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x*blockDim.x)
		for (int j = 0; j < count; ++j) D[j*ld + i] = H.h[j][i] - H.h[j + 1][i];
}

__device__ inline double WaveSum(double v)
{
	for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
	return v;
}

// partial[v*gridDim.x + block] = this block's share of V_v . w, v < count <= kMpe + 1
__global__ void __launch_bounds__(kThreads) DotsKernel(const double* __restrict__ V, long long ld, int count,
	const double* __restrict__ w, long long n, double* __restrict__ partial)
{
	__shared__ double lds[kMpe + 1][kThreads/64];
	double acc[kMpe + 1];
	for (int v = 0; v <= kMpe; ++v) acc[v] = 0.0;
	for (long long i = blockIdx.x*(long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x*blockDim.x) {
		const double wi = w[i];
#pragma unroll
		for (int v = 0; v <= kMpe; ++v)
			if (v < count) acc[v] += V[v*ld + i]*wi;
	}
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
	for (int v = 0; v <= kMpe; ++v) {
		if (v >= count) break;
		const double s = WaveSum(acc[v]);
		if (lane == 0) lds[v][wave] = s;
	}
	__syncthreads();
	if (threadIdx.x < count) {
		double s = 0.0;
		for (int k = 0; k < kThreads/64; ++k) s += lds[threadIdx.x][k];
		partial[threadIdx.x*gridDim.x + blockIdx.x] = s;
	}
}

struct Coefficients {
	double c[kMpe + 1];
};

// out = base + sum_{v<count} c_v V_v ; norm_partial[block] = this block's share of ||out||^2
// (out may be base: each entry is read before it is written)
__global__ void __launch_bounds__(kThreads) CombineKernel(const double* base, const double* __restrict__ V,
	long long ld, int count, Coefficients c, double* out, long long n, double* __restrict__ norm_partial)
{
	__shared__ double lds[kThreads/64];
	double sq = 0.0;
	for (long long i = blockIdx.x*(long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x*blockDim.x) {
		double s = base[i];
		for (int v = 0; v < count; ++v) s += c.c[v]*V[v*ld + i];
		out[i] = s;
		sq += s*s;
	}
	if (!norm_partial) return;
	sq = WaveSum(sq);
	if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = sq;
	__syncthreads();
	if (threadIdx.x == 0) {
		double s = 0.0;
		for (int k = 0; k < kThreads/64; ++k) s += lds[k];
		norm_partial[blockIdx.x] = s;
	}
}

}  // namespace

struct WarmStart::Impl {
	tfi_mesh* mesh;
	tfi_ctx* ctx;
	static constexpr int kRing = kMpe + 2;
	DevBuf<double> hist[kRing];   // u_{n}, u_{n-1}, ... (ring)
	int count = 0;
	int head = 0;                 // index of u_n
	double phase_hist[kRing][8] = {};
	int accepted = 0, rejected = 0;
	std::unique_ptr<Mlp> mlp;
	tfi_warm_options opt{128, 27, 20, 1e-3, 2, TFI_WARM_MPE};
	// minimal polynomial extrapolation work: increments D[0..kMpe] and the
	// orthonormal basis Q[0..kMpe-1] of D[1..m], contiguous columns
	DevBuf<double> D, Q, partial;
	int mpe_order = 0;
	// base prediction of the learned mode for the history of `base_count` states
	DevBuf<double> base;
	int base_count = -1;
	const double* at(int back) const { return hist[(head - back + kRing)%kRing].data(); }
	long long n() const { return 4LL*mesh->nodes; }

	// sums[v] = V_v . w over v < count, reduced on the host in a fixed order
	void Dots(const double* V, int count, const double* w, double* sums)
	{
		const long long len = n();
		hipLaunchKernelGGL(DotsKernel, dim3(kDotBlocks), dim3(kThreads), 0, ctx->stream, V, len, count, w, len,
			partial.data());
		TFI_LAUNCH_CHECK("DotsKernel");
		double* h = ctx->host_pinned;
		Check(hipMemcpyAsync(h, partial.data(), sizeof(double)*count*kDotBlocks, hipMemcpyDeviceToHost, ctx->stream),
			"dots download");
		Check(hipStreamSynchronize(ctx->stream), "dots sync");
		for (int v = 0; v < count; ++v) {
			double s = 0.0;
			for (int b = 0; b < kDotBlocks; ++b) s += h[v*kDotBlocks + b];
			sums[v] = s;
		}
	}

	// out = base + sum c_v V_v; returns ||out|| when want_norm
	double Combine(const double* base_vec, const double* V, int count, const double* c, double* out, bool want_norm)
	{
		const long long len = n();
		Coefficients k{};
		for (int v = 0; v < count; ++v) k.c[v] = c[v];
		hipLaunchKernelGGL(CombineKernel, dim3(kDotBlocks), dim3(kThreads), 0, ctx->stream, base_vec, V, len, count, k,
			out, len, want_norm ? partial.data() : nullptr);
		TFI_LAUNCH_CHECK("CombineKernel");
		if (!want_norm) return 0.0;
		double* h = ctx->host_pinned;
		Check(hipMemcpyAsync(h, partial.data(), sizeof(double)*kDotBlocks, hipMemcpyDeviceToHost, ctx->stream),
			"norm download");
		Check(hipStreamSynchronize(ctx->stream), "norm sync");
		double s = 0.0;
		for (int b = 0; b < kDotBlocks; ++b) s += h[b];
		return std::sqrt(s);
	}

	// u_{n+1} ~ u_n + sum_{i=1..m} c_i d_{n+1-i}, c = argmin ||d_n - sum_i c_i d_{n-i}||
	// through a QR factorisation of [d_{n-1} .. d_{n-m}] by classical
	// Gram-Schmidt applied twice; columns (newest first) are taken while they
	// add a direction above the numerical-rank threshold.
	void Mpe(double* out)
	{
		const long long len = n();
		const int m = std::min(kMpe, count - 2);
		if (!D.size()) {
			D.Allocate(ctx, (size_t)(kMpe + 1)*len);
			Q.Allocate(ctx, (size_t)kMpe*len);
			partial.Allocate(ctx, (size_t)(kMpe + 1)*kDotBlocks);
		}
		History H{};
		for (int j = 0; j <= m + 1; ++j) H.h[j] = at(j);
		hipLaunchKernelGGL(IncrementKernel, dim3(2048), dim3(kThreads), 0, ctx->stream, len, m + 1, H, D.data(), len);
		TFI_LAUNCH_CHECK("IncrementKernel");
		double R[kMpe][kMpe] = {};
		int k = 0;
		for (; k < m; ++k) {
			const double* dk = D.data() + (long long)(k + 1)*len;
			double* qk = Q.data() + (long long)k*len;
			double self[1];
			Dots(dk, 1, dk, self);
			const double norm0 = std::sqrt(self[0]);
			if (!(norm0 > 0.0)) break;
			double coef[kMpe + 1] = {}, h[kMpe + 1];
			// pass 1 from d itself, pass 2 on the projected vector
			Copy(ctx, dk, qk, len);
			double nrm = norm0;
			for (int pass = 0; pass < 2 && k > 0; ++pass) {
				Dots(Q.data(), k, qk, h);
				double neg[kMpe + 1];
				for (int i = 0; i < k; ++i) {
					coef[i] += h[i];
					neg[i] = -h[i];
				}
				nrm = Combine(qk, Q.data(), k, neg, qk, true);
			}
			if (!(nrm > kMpeRank*norm0)) break;
			for (int i = 0; i < k; ++i) R[i][k] = coef[i];
			R[k][k] = nrm;
			Scale(ctx, 1.0/nrm, qk, len);
		}
		mpe_order = k;
		if (k == 0) {
			Copy(ctx, at(0), out, len);
			return;
		}
		// c = R^-1 Q^T d_n
		double z[kMpe + 1], c[kMpe + 1] = {};
		Dots(Q.data(), k, D.data(), z);
		for (int i = k - 1; i >= 0; --i) {
			double s = z[i];
			for (int j = i + 1; j < k; ++j) s -= R[i][j]*c[j];
			c[i] = s/R[i][i];
		}
		// next increment sum_i c_i d_{n+1-i}: column i-1 of D for coefficient i
		Combine(at(0), D.data(), k, c, out, false);
	}

	void Polynomial(int order, double* out)
	{
		History H{};
		for (int j = 0; j <= order; ++j) H.h[j] = at(j);
		hipLaunchKernelGGL(ExtrapolateKernel, dim3(1024), dim3(kThreads), 0, ctx->stream, n(), order, H, out);
		TFI_LAUNCH_CHECK("ExtrapolateKernel");
	}

	int BaseMode() const { return opt.base == TFI_WARM_EXTRAPOLATE4 ? TFI_WARM_EXTRAPOLATE4 : TFI_WARM_MPE; }

	// base prediction of the learned mode from the current history, cached per history length
	const double* Base()
	{
		if (!base.size()) base.Allocate(ctx, n());
		if (base_count != count) {
			if (BaseMode() == TFI_WARM_MPE && count >= 3) Mpe(base.data());
			else Polynomial(kOrder, base.data());
			base_count = count;
		}
		return base.data();
	}
};

WarmStart::WarmStart(tfi_mesh* mesh) : impl_(std::make_unique<Impl>())
{
	impl_->mesh = mesh;
	impl_->ctx = mesh->ctx;
	for (auto& h : impl_->hist) h.Allocate(mesh->ctx, 4LL*mesh->nodes);
}

WarmStart::~WarmStart() = default;

void WarmStart::Configure(const tfi_warm_options& o)
{
	if (o.base != 0 && o.base != TFI_WARM_MPE && o.base != TFI_WARM_EXTRAPOLATE4)
		throw Error("warm-start base must be TFI_WARM_MPE or TFI_WARM_EXTRAPOLATE4");
	impl_->opt = o;
	impl_->mlp.reset();
	impl_->base_count = -1;
}

void WarmStart::Reset()
{
	impl_->count = 0;
	impl_->head = 0;
	impl_->base_count = -1;
}

void WarmStart::Commit(const double* state, const double* phase, int phase_count, bool train)
{
	Impl& I = *impl_;
	const long long n = I.n();
	// learned mode: the base prediction of this state from the history before it
	// and the history itself form a training snapshot
	const bool snapshot = train && I.count >= kOrder + 1;
	const double* base = snapshot ? I.Base() : nullptr;
	I.head = (I.head + 1)%Impl::kRing;
	Copy(I.ctx, state, I.hist[I.head].data(), n);
	for (int i = 0; i < 8; ++i) I.phase_hist[I.head][i] = i < phase_count ? phase[i] : 0.0;
	++I.count;
	if (snapshot) {
		if (!I.mlp) I.mlp = std::make_unique<Mlp>(I.mesh, I.opt);
		const double* h[kOrder + 1];
		for (int j = 0; j <= kOrder; ++j) h[j] = I.at(j + 1);
		I.mlp->Train(I.at(0), h, base, I.phase_hist[I.head]);
	}
}

bool WarmStart::Ready(int mode) const
{
	const Impl& I = *impl_;
	if (mode == TFI_WARM_EXTRAPOLATE) return I.count >= 2;
	if (mode == TFI_WARM_EXTRAPOLATE2) return I.count >= 3;
	if (mode == TFI_WARM_EXTRAPOLATE3) return I.count >= 4;
	if (mode == TFI_WARM_EXTRAPOLATE4) return I.count >= 5;
	if (mode == TFI_WARM_MPE) return I.count >= 3;
	if (mode == TFI_WARM_LEARNED) return I.count >= kOrder + 1 && I.mlp && I.mlp->snapshots() >= I.opt.min_snapshots;
	return false;
}

void WarmStart::Predict(int mode, double* state, const double* phase, int phase_count)
{
	Impl& I = *impl_;
	const int order = mode == TFI_WARM_EXTRAPOLATE ? 1 : mode == TFI_WARM_EXTRAPOLATE2 ? 2
		: mode == TFI_WARM_EXTRAPOLATE3 ? 3 : mode == TFI_WARM_EXTRAPOLATE4 ? 4 : 0;
	if (order > 0) {
		I.Polynomial(order, state);
	} else if (mode == TFI_WARM_MPE) {
		I.Mpe(state);
	} else if (mode == TFI_WARM_LEARNED) {
		double ph[8] = {};
		for (int i = 0; i < 8 && i < phase_count; ++i) ph[i] = phase[i];
		const double* h[kOrder + 1];
		for (int j = 0; j <= kOrder; ++j) h[j] = I.at(j);
		I.mlp->Predict(h, I.Base(), ph, state);
	}
}

void WarmStart::PredictBase(double* state)
{
	Copy(impl_->ctx, impl_->Base(), state, impl_->n());
}

int WarmStart::MpeOrder() const
{
	return impl_->mpe_order;
}

int WarmStart::BaseMode() const
{
	return impl_->BaseMode();
}

void WarmStart::Accepted(bool used)
{
	if (used) ++impl_->accepted;
	else ++impl_->rejected;
}

void WarmStart::Stats(int* snapshots, double* last_loss, int* accepted, int* rejected) const
{
	const Impl& I = *impl_;
	if (snapshots) *snapshots = I.mlp ? I.mlp->snapshots() : 0;
	if (last_loss) *last_loss = I.mlp ? I.mlp->last_loss() : 0.0;
	if (accepted) *accepted = I.accepted;
	if (rejected) *rejected = I.rejected;
}

double WarmStart::Gate() const
{
	return impl_->mlp ? impl_->mlp->gate() : 0.0;
}

}  // namespace tfi
