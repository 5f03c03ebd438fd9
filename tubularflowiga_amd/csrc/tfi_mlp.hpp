// Stencil MLP that predicts a time step's increment (learned warm start).
#pragma once

#include "tfi_internal.hpp"

namespace tfi {

// Feature layout per target node i (kIn columns, fp32), for history order
// P = kOrder: for each of the 27 stencil nodes s (i itself, then its 26 nearest
// pattern neighbours by control-point distance): u_n(s)/s0 (4), then the
// backward differences d_k(s)/s_k for k = 1..P (4 each), then
// (x_s - x_i)/h_i (3); then up to 8 phase features; zeros to kIn.
//   d_k = sum_j (-1)^j C(k, j) u_{n-j}   (d1 = u_n - u_{n-1}, ...)
// Output (4): the error of a base prediction b of the next state (minimal
// polynomial extrapolation of the increments, or the order-P polynomial
// extrapolation in time sum_{j=0..P} (-1)^j C(P+1, j+1) u_{n-j}) scaled by
// s_P, so the prediction is
//   u_{n+1} ~ b + s_P * MLP(x),
// with s_0..s_P the per-field RMS of u_n, d_1, .., d_P.
constexpr int kOrder = 4;
constexpr int kStencil = 27;
constexpr int kPerNode = 4*(kOrder + 1) + 3;
constexpr int kIn = (kStencil*kPerNode + 8 + 15)/16*16;
constexpr int kOut = 4;

class Mlp {
public:
	Mlp(tfi_mesh* mesh, const tfi_warm_options& options);
	~Mlp();
	// One snapshot: history h[j] = u_{n-j} (j = 0..kOrder), base prediction of
	// u_{n+1} (nullptr: the order-kOrder extrapolation) -> next = u_{n+1}.
	void Train(const double* next, const double* const* history, const double* base, const double* phase);
	// state <- base (nullptr: order-kOrder extrapolation) + s_P * MLP(features(history, phase)).
	void Predict(const double* const* history, const double* base, const double* phase, double* state);
	int snapshots() const { return snapshots_; }
	double last_loss() const { return last_loss_; }
	double gate() const;

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
