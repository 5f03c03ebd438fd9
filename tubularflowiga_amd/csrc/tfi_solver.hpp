// Linear solver, algebraic multigrid and warm-start declarations.
#pragma once

#include "tfi_internal.hpp"

namespace tfi {

// Node-block plain-aggregation AMG (tfi_amg.hip).
struct AmgOptions {
	int smoother = TFI_SMOOTH_GAUSS_SEIDEL;  // GS on levels with at least gs_min_rows rows; CHEBYSHEV on every level
	int coarse_nodes = 256;                  // block rows on the coarsest (dense direct) level
	int gs_min_rows = 8192;                  // smaller levels use damped block Jacobi
	int gs_sweeps = 2;                       // 2: symmetric (forward+backward) pre and post; 1: forward pre, backward post
	int chebyshev_degree = 2;                // Chebyshev steps pre and post (the first pre step from zero reads no matrix)
	int half_smoother = 0;                   // mixed precision: Chebyshev levels read D^-1 A in FP16
};

class Amg {
public:
	Amg(tfi_mesh* mesh, int precision, const AmgOptions& options);
	~Amg();
	// Numeric setup for new fine values (symbolic hierarchy built on first call).
	void Setup(const double* values64, const float* values32);
	// z = M^-1 v (one V-cycle), FP64 vectors at the interface.
	void Apply(const double* v, double* z);
	int singular_blocks() const { return singular_; }
	int levels() const;
	// rows, block nonzeros and Gauss-Seidel colours (0: Jacobi / direct) per level
	void Describe(std::vector<int>& rows, std::vector<long long>& nnzb, std::vector<int>& colors) const;
	int smoother() const;
	// switch a Chebyshev-smoothed hierarchy to multicolour Gauss-Seidel (for good);
	// Setup must run again before the next Apply
	void UseGaussSeidel();
	struct Impl;

private:
	std::unique_ptr<Impl> impl_;
	int singular_ = 0;
};

class LinearSolver {
public:
	LinearSolver(tfi_mesh* mesh, int preconditioner, int precision, int restart,
		const AmgOptions& amg = AmgOptions());
	~LinearSolver();
	void Setup(const double* values);
	// A Chebyshev-smoothed AMG whose solve has not converged within
	// kChebyshevCap iterations is switched to Gauss-Seidel and the solve rerun
	// from the same initial guess.
	KrylovResult Solve(const double* b, double* x, int maxit, double rtol);
	int singular_blocks() const { return singular_; }
	int refinements() const { return refinements_; }
	const Amg* amg() const { return amg_.get(); }
	void set_verbose(int verbose) { verbose_ = verbose; }

private:
	tfi_mesh* mesh_;
	tfi_ctx* ctx_;
	int pc_, precision_;
	Fgmres krylov_;
	const double* values_ = nullptr;
	DevBuf<double> inverse_, work_, correction_, residual_;
	DevBuf<float> values32_;
	std::unique_ptr<Amg> amg_;
	int singular_ = 0;
	int refinements_ = 0;
	DevBuf<double> guess_;
	int verbose_ = 0;
	KrylovResult SolveOnce(const double* b, double* x, int maxit, double rtol);
};

// Initial-state predictor for a time step (tfi_warmstart.hip).
class WarmStart {
public:
	explicit WarmStart(tfi_mesh* mesh);
	~WarmStart();
	void Configure(const tfi_warm_options& options);
	void Reset();
	// Record the converged state of a finished step (and train when learned).
	void Commit(const double* state, const double* phase, int phase_count, bool train);
	bool Ready(int mode) const;
	// state <- predicted state of the next step (state holds u_n on entry).
	void Predict(int mode, double* state, const double* phase, int phase_count);
	// state <- the learned mode's base prediction (its correction is added to it)
	void PredictBase(double* state);
	// increments combined by the last TFI_WARM_MPE prediction
	int MpeOrder() const;
	// the mode of the learned prediction's base (TFI_WARM_MPE or TFI_WARM_EXTRAPOLATE4)
	int BaseMode() const;
	void Accepted(bool used);
	void Stats(int* snapshots, double* last_loss, int* accepted, int* rejected) const;
	double Gate() const;
	struct Impl;

private:
	std::unique_ptr<Impl> impl_;
};

}  // namespace tfi
