// Internal declarations shared by the translation units of libtfi_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tfi.h"

namespace tfi {

struct Error : std::runtime_error {
	using std::runtime_error::runtime_error;
};

inline void Check(hipError_t status, const char* what)
{
	if (status != hipSuccess)
		throw Error(std::string(what) + ": " + hipGetErrorString(status));
}

#define TFI_LAUNCH_CHECK(name) ::tfi::Check(hipGetLastError(), name)

}  // namespace tfi

struct tfi_ctx {
	int device = 0;
	hipStream_t stream = nullptr;
	std::string error;
	long long live_bytes = 0, peak_bytes = 0;
	bool timers_enabled = false;
	tfi_kernel_timers timers{};
	hipEvent_t ev_begin = nullptr, ev_end = nullptr;
	int cus = 0;
	double* scratch = nullptr;       // reduction partials
	size_t scratch_count = 0;
	double* host_pinned = nullptr;   // small pinned host buffer for reductions
	// kernel timers: event pairs recorded on the stream, resolved lazily
	struct PendingTimer {
		hipEvent_t begin, end;
		double* accumulate;
	};
	std::vector<PendingTimer> pending;
	std::vector<hipEvent_t> event_pool;
};

namespace tfi {

// RAII device buffer counted against the context (peak / live bytes like the
// reference's DeviceAllocationCounter, DeviceAllocation.hpp:12-28).
template <class T>
class DevBuf {
public:
	DevBuf() = default;
	DevBuf(tfi_ctx* ctx, size_t n) { Allocate(ctx, n); }
	DevBuf(const DevBuf&) = delete;
	DevBuf& operator=(const DevBuf&) = delete;
	DevBuf(DevBuf&& o) noexcept { Swap(o); }
	DevBuf& operator=(DevBuf&& o) noexcept { Release(); Swap(o); return *this; }
	~DevBuf() { Release(); }

	void Allocate(tfi_ctx* ctx, size_t n)
	{
		Release();
		ctx_ = ctx;
		n_ = n;
		if (n == 0) return;
		Check(hipMalloc(&p_, n*sizeof(T)), "hipMalloc");
		ctx_->live_bytes += static_cast<long long>(n*sizeof(T));
		if (ctx_->live_bytes > ctx_->peak_bytes) ctx_->peak_bytes = ctx_->live_bytes;
	}
	void Release()
	{
		if (p_) {
			(void)hipFree(p_);
			ctx_->live_bytes -= static_cast<long long>(n_*sizeof(T));
		}
		p_ = nullptr;
		n_ = 0;
	}
	void Upload(const T* host, size_t n, size_t offset = 0)
	{
		if (offset + n > n_) throw Error("upload exceeds buffer");
		if (n) Check(hipMemcpyAsync(p_ + offset, host, n*sizeof(T), hipMemcpyHostToDevice, ctx_->stream), "upload");
		Check(hipStreamSynchronize(ctx_->stream), "upload sync");
	}
	void Download(T* host, size_t n, size_t offset = 0) const
	{
		if (offset + n > n_) throw Error("download exceeds buffer");
		if (n) Check(hipMemcpyAsync(host, p_ + offset, n*sizeof(T), hipMemcpyDeviceToHost, ctx_->stream), "download");
		Check(hipStreamSynchronize(ctx_->stream), "download sync");
	}
	T* data() const { return p_; }
	size_t size() const { return n_; }
	size_t bytes() const { return n_*sizeof(T); }

private:
	void Swap(DevBuf& o)
	{
		std::swap(p_, o.p_);
		std::swap(n_, o.n_);
		std::swap(ctx_, o.ctx_);
	}
	T* p_ = nullptr;
	size_t n_ = 0;
	tfi_ctx* ctx_ = nullptr;
};

// Block-CSR pattern view (node blocks, 4x4 values).
struct BsrView {
	int rows = 0;
	long long nnzb = 0;
	const int* row_offsets = nullptr;
	const int* columns = nullptr;
	const int* diagonal = nullptr;
};

// Per-quadrature-point geometry record: measure, J^-1 (9, [param][phys]),
// L = J^-1 J^-T (6: 00 11 22 01 02 12), h_c = sum_i d2xi_c/dx_i^2 (3), pad.
constexpr int kGeomFields = 20;
constexpr int kQ = 64;
constexpr int kMaxBasis = 80;

}  // namespace tfi

struct tfi_mesh {
	tfi_ctx* ctx = nullptr;
	int nodes = 0, elements = 0, max_nen = 0;
	long long basis_entries = 0, extraction_nonzeros = 0, element_pairs = 0;
	tfi::DevBuf<int> eoff, conn, xoff;
	tfi::DevBuf<uint8_t> xcol;
	tfi::DevBuf<double> xval, bezier;
	// pattern
	tfi::DevBuf<int> row_offsets, columns, diagonal;
	long long nnzb = 0;
	// deterministic scatter: colours and pair -> block map
	int colors = 0;
	std::vector<int> color_offsets;          // host [colors+1]
	tfi::DevBuf<int> color_elements;         // elements sorted by colour
	std::vector<int> h_color;                // colour per element
	tfi::DevBuf<long long> pair_offsets;     // [E+1]
	tfi::DevBuf<int> pair_block;             // [sum nen^2]
	// geometry [E][kGeomFields][64]
	tfi::DevBuf<double> geometry;
	// optional precomputed basis records: element e at eoff[e]*64*5, [q][a][5]
	// (N, grad N, lap N); empty when assembly evaluates the basis itself
	tfi::DevBuf<double> basis_table;
	double min_det = 0.0;
	long long bad_samples = 0;
	// host copies
	std::vector<int> h_eoff, h_conn, h_xoff;
	std::vector<uint8_t> h_xcol;
	std::vector<double> h_xval, h_bezier;
	tfi::BsrView view() const
	{
		return {nodes, nnzb, row_offsets.data(), columns.data(), diagonal.data()};
	}
};

namespace tfi {

// ---- core (tfi_core.hip)
void EnsureScratch(tfi_ctx* ctx, size_t count);
double Dot(tfi_ctx* ctx, const double* x, const double* y, long long n);
double Norm2(tfi_ctx* ctx, const double* x, long long n);
void Axpy(tfi_ctx* ctx, double a, const double* x, double* y, long long n);
void Axpby(tfi_ctx* ctx, double a, const double* x, double b, double* y, long long n);
void Scale(tfi_ctx* ctx, double a, double* x, long long n);
void Copy(tfi_ctx* ctx, const double* x, double* y, long long n);
// FP64 matrix-core throughput probe (blocks x threads lanes, iterations of 4 MFMA chains)
void MfmaPeak(tfi_ctx* ctx, int blocks, int threads, int iterations, double* out);
// n doubles (even), `blocks` workgroups of 256; partial receives one sum per workgroup
void StreamCopy(tfi_ctx* ctx, const double* x, double* y, long long n, int blocks);
void StreamRead(tfi_ctx* ctx, const double* x, long long n, double* partial, int blocks);
void Fill(tfi_ctx* ctx, double* x, double value, long long n);
void FillInt(tfi_ctx* ctx, int* x, int value, long long n);
void ToFloat(tfi_ctx* ctx, const double* x, float* y, long long n);
void Sync(tfi_ctx* ctx);

// Event timer on the context stream (no-op unless timers are enabled).  The
// begin/end events are recorded around the timed launches without any host
// synchronisation; ResolveTimers adds the elapsed times later.
hipEvent_t TakeEvent(tfi_ctx* ctx);
void ResolveTimers(tfi_ctx* ctx);
struct StreamTimer {
	tfi_ctx* ctx;
	double* accumulate;
	bool on;
	hipEvent_t begin = nullptr;
	StreamTimer(tfi_ctx* c, double* acc) : ctx(c), accumulate(acc), on(c->timers_enabled)
	{
		if (!on) return;
		begin = TakeEvent(ctx);
		Check(hipEventRecord(begin, ctx->stream), "event record");
	}
	~StreamTimer()
	{
		if (!on) return;
		hipEvent_t end = TakeEvent(ctx);
		if (hipEventRecord(end, ctx->stream) == hipSuccess) ctx->pending.push_back({begin, end, accumulate});
		if (ctx->pending.size() > 8192) ResolveTimers(ctx);
	}
};

// ---- mesh (tfi_mesh.hip)
void BuildMesh(tfi_mesh* mesh);

// ---- assembly (tfi_assembly.hip)
struct NsParams {
	double density, viscosity, dt;
	int tangent = TFI_TANGENT_LEGACY;
};
void AssembleNs(tfi_mesh* mesh, const double* state, const double* previous, const NsParams& p, double* jacobian,
	double* negative_residual);
double AssemblyFlops(const tfi_mesh* mesh, bool with_jacobian, bool exact);
// Builds (enable) or frees the per-element basis table; enable < 0 builds it
// when it takes at most half of the free device memory.
void SetBasisTable(tfi_mesh* mesh, int enable);
void ApplyDirichlet(tfi_mesh* mesh, const int* vmask, const int* pmask, const double* vtarget, const double* ptarget,
	const double* state, double* jacobian, double* rhs);
// state[constrained] = target
void SetDirichletState(tfi_mesh* mesh, const int* vmask, const int* pmask, const double* vtarget,
	const double* ptarget, double* state);
// Per-face pressure-traction vectors local[count][kMaxBasis][3] (device).
// flows[f] = -sum_(a, c) state[4 conn(a) + c] vec[f][a][c] with vec the
// per-face vectors FaceTraction gives for unit pressure (= int_face N_a n dA).
void FaceFlowsFromVectors(tfi_mesh* mesh, int count, const int* element, const double* vec, const double* state,
	double* flows);
void FaceTraction(tfi_mesh* mesh, int count, const int* element, const int* face, const double* pressure,
	double* local);
// Per-face surface mass[count][S][S] = int_face N_a N_b dA and
// load[count][S] = int_face N_a dA (device), S = mesh->max_nen, 4x4 Gauss rule.
void FaceMass(tfi_mesh* mesh, int count, const int* element, const int* face, double* mass, double* load);

// ---- sparse (tfi_sparse.hip)
// Algorithmic bytes of one SpMV: values + column indices once, row offsets,
// x (FP64) read once and y (FP64) written once.
inline double SpmvBytes(const BsrView& a, int value_bytes)
{
	return (double)a.nnzb*(16.0*value_bytes + 4.0) + 4.0*(a.rows + 1) + 2.0*4.0*a.rows*8.0;
}
void Spmv64(tfi_ctx* ctx, const BsrView& a, const double* values, const double* x, double* y);
void Spmv32(tfi_ctx* ctx, const BsrView& a, const float* values, const double* x, double* y);
// r = b - A x
void Residual64(tfi_ctx* ctx, const BsrView& a, const double* values, const double* x, const double* b, double* r);
int BlockInverse64(tfi_ctx* ctx, const BsrView& a, const double* values, double* inverse);
void BlockApply64(tfi_ctx* ctx, int rows, const double* inverse, const double* x, double* y);
void BlockInverseTo32(tfi_ctx* ctx, int rows, const double* inverse, float* inverse32);
// Looks up, once per device (the device current), the device address of the
// fixed working memory of tfi_block_inverse_noalloc.
void InitNoAllocScratch(int device);

// ---- Krylov (tfi_krylov.hip)
// Preconditioner: z = M^-1 v (both length 4N, device).
using Precond = std::function<void(const double* v, double* z)>;
using Operator = std::function<void(const double* x, double* y)>;
struct KrylovResult {
	int iterations = 0;
	double residual = 0.0, rhs_norm = 0.0;
	bool converged = false;
};
class Fgmres {
public:
	Fgmres(tfi_ctx* ctx, long long n, int restart);
	// Solves A x = b to ||b - A x|| <= rtol ||b|| (true residual).
	KrylovResult Solve(const Operator& A, const Precond& M, const double* b, double* x, int maxit, double rtol,
		double absolute_target = 0.0);
private:
	tfi_ctx* ctx_;
	long long n_;
	int m_;
	DevBuf<double> V_, Z_, w_, r_, h_, partial_;
	std::vector<double> H_, cs_, sn_, g_, y_;
};

}  // namespace tfi
