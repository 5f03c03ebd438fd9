// Initial-state prediction for a time step.  Modes:
//   TFI_WARM_PREVIOUS    the step starts from u_n (the reference behaviour,
//                        iga_cuda.cu:917-920);
//   TFI_WARM_EXTRAPOLATE u_n + (u_n - u_{n-1});
//   TFI_WARM_EXTRAPOLATE2 3 u_n - 3 u_{n-1} + u_{n-2} (quadratic in time);
//   TFI_WARM_EXTRAPOLATE3 4 u_n - 6 u_{n-1} + 4 u_{n-2} - u_{n-3} (cubic in time);
//   TFI_WARM_EXTRAPOLATE4 quartic in time (the learned mode's base);
//   TFI_WARM_LEARNED     polynomial extrapolation of order kOrder + MLP(stencil
//                        features), see tfi_mlp.hpp.
// A prediction is only a guess: the Newton loop accepts it when it lowers the
// nonlinear residual and always converges to the same tolerance.
#include "tfi_solver.hpp"
#include "tfi_mlp.hpp"

namespace tfi {

namespace {

// order-P polynomial extrapolation in time from u_n, .., u_{n-P}:
// sum_{j=0..P} (-1)^j C(P+1, j+1) u_{n-j}
struct History {
	const double* h[5];
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

}  // namespace

struct WarmStart::Impl {
	tfi_mesh* mesh;
	tfi_ctx* ctx;
	static constexpr int kRing = kOrder + 2;
	DevBuf<double> hist[kRing];   // u_{n}, u_{n-1}, ... (ring)
	int count = 0;
	int head = 0;                 // index of u_n
	double phase_hist[kRing][8] = {};
	int accepted = 0, rejected = 0;
	std::unique_ptr<Mlp> mlp;
	tfi_warm_options opt{128, 27, 20, 1e-3, 2};
	const double* at(int back) const { return hist[(head - back + kRing)%kRing].data(); }
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
	impl_->opt = o;
	impl_->mlp.reset();
}

void WarmStart::Reset()
{
	impl_->count = 0;
	impl_->head = 0;
}

void WarmStart::Commit(const double* state, const double* phase, int phase_count, bool train)
{
	Impl& I = *impl_;
	const long long n = 4LL*I.mesh->nodes;
	I.head = (I.head + 1)%Impl::kRing;
	Copy(I.ctx, state, I.hist[I.head].data(), n);
	for (int i = 0; i < 8; ++i) I.phase_hist[I.head][i] = i < phase_count ? phase[i] : 0.0;
	++I.count;
	// learned mode: a new (u_{n-1-kOrder}, .., u_{n-1}) -> u_n snapshot trains the network
	if (train && I.count >= kOrder + 2) {
		if (!I.mlp) I.mlp = std::make_unique<Mlp>(I.mesh, I.opt);
		const double* h[kOrder + 1];
		for (int j = 0; j <= kOrder; ++j) h[j] = I.at(j + 1);
		I.mlp->Train(I.at(0), h, I.phase_hist[I.head]);
	}
}

bool WarmStart::Ready(int mode) const
{
	const Impl& I = *impl_;
	if (mode == TFI_WARM_EXTRAPOLATE) return I.count >= 2;
	if (mode == TFI_WARM_EXTRAPOLATE2) return I.count >= 3;
	if (mode == TFI_WARM_EXTRAPOLATE3) return I.count >= 4;
	if (mode == TFI_WARM_EXTRAPOLATE4) return I.count >= 5;
	if (mode == TFI_WARM_LEARNED) return I.count >= kOrder + 1 && I.mlp && I.mlp->snapshots() >= I.opt.min_snapshots;
	return false;
}

void WarmStart::Predict(int mode, double* state, const double* phase, int phase_count)
{
	Impl& I = *impl_;
	const long long n = 4LL*I.mesh->nodes;
	const int order = mode == TFI_WARM_EXTRAPOLATE ? 1 : mode == TFI_WARM_EXTRAPOLATE2 ? 2
		: mode == TFI_WARM_EXTRAPOLATE3 ? 3 : mode == TFI_WARM_EXTRAPOLATE4 ? 4 : 0;
	if (order > 0) {
		History H{};
		for (int j = 0; j <= order; ++j) H.h[j] = I.at(j);
		hipLaunchKernelGGL(ExtrapolateKernel, dim3(1024), dim3(256), 0, I.ctx->stream, n, order, H, state);
		TFI_LAUNCH_CHECK("ExtrapolateKernel");
	} else if (mode == TFI_WARM_LEARNED) {
		double ph[8] = {};
		for (int i = 0; i < 8 && i < phase_count; ++i) ph[i] = phase[i];
		const double* h[kOrder + 1];
		for (int j = 0; j <= kOrder; ++j) h[j] = I.at(j);
		I.mlp->Predict(h, ph, state);
	}
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
