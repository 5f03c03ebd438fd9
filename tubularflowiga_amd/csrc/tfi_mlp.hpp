// Stencil MLP that predicts a time step's increment (learned warm start).
#pragma once

#include "tfi_internal.hpp"

namespace tfi {

// Feature layout per target node i (kIn = 416 columns, fp32):
//   for each of the 27 stencil nodes s (i itself, then its 26 nearest pattern
//   neighbours by control-point distance): u_n(s)/su (4), d1_n(s)/s1 (4),
//   d2_n(s)/s2 (4), (x_s - x_i)/h_i (3)  -> 405 columns;  then up to 8 phase
//   features;  zeros.  d1_n = u_n - u_{n-1}, d2_n = d1_n - d1_{n-1}.
// Output (4): d3_{n+1}(i)/s2 with d3_{n+1} = u_{n+1} - (3 u_n - 3 u_{n-1} + u_{n-2}),
// so the prediction is the quadratic-in-time extrapolation plus a learned
// correction, u_{n+1} ~ 3 u_n - 3 u_{n-1} + u_{n-2} + s2 * MLP(x),
// with su, s1, s2 the per-field RMS of u_n, d1_n, d2_n.
constexpr int kStencil = 27;
constexpr int kPerNode = 15;
constexpr int kIn = 416;
constexpr int kOut = 4;

class Mlp {
public:
	Mlp(tfi_mesh* mesh, const tfi_warm_options& options);
	~Mlp();
	// One snapshot: (u_{n-2}, u_{n-1}, u_n) -> u_{n+1}.
	void Train(const double* un1, const double* un, const double* um, const double* umm, const double* phase);
	// state <- 3 u_n - 3 u_{n-1} + u_{n-2} + s2 * MLP(features(u_n, u_{n-1}, u_{n-2}, phase)).
	void Predict(const double* un, const double* um, const double* umm, const double* phase, double* state);
	int snapshots() const { return snapshots_; }
	double last_loss() const { return last_loss_; }

	// Test hooks: forward + backward of a host batch (x [B][kIn], target [B][4])
	// with the given fp32 weights; returns loss and gradients.
	struct Weights {
		std::vector<float> w1, b1, w2, b2, w3, b3;
	};
	static void Reference(tfi_ctx* ctx, int hidden, int batch, const float* x, const float* target, const Weights& w,
		float* out, float* loss, Weights& grad);
	static void InferBf16(tfi_ctx* ctx, int hidden, int rows, const float* x, const Weights& w, float* out);

	struct Impl;

private:
	std::unique_ptr<Impl> impl_;
	int snapshots_ = 0;
	double last_loss_ = 0.0;
};

}  // namespace tfi
