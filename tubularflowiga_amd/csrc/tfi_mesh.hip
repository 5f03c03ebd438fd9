// Mesh upload, node-block pattern, deterministic-scatter colouring and the
// per-quadrature-point geometry.
//
// * Pattern: for each node, the sorted unique union of the connectivities of
//   the elements that contain it, plus the diagonal position -- the
//   algorithm of BlockPattern (solvers/cuda/include/BlockCsr.hpp:27-55),
//   built on the GPU: one workgroup per node gathers its candidate columns into
//   LDS and bitonic-sorts them; a host exclusive scan turns the counts into row
//   offsets.  Output is bit-identical to the reference's pattern.
// * Colouring: greedy, in element order, so that no two elements of one colour
//   share a node; assembly adds colours in order, which makes every block sum
//   deterministic without atomics (the reference scatters with FP64
//   atomicAdd, IgaCudaKernels.cuh:400-408).
// * Geometry: per (element, Gauss point) the measure, J^-1 and the two
//   contractions of d(J^-1) the Laplacian needs (BuildGeometryKernel,
//   IgaCudaKernels.cuh:61-111, including its |det| < 1e-14 zero-inverse rule).
#include "tfi_internal.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>

namespace tfi {

namespace {

__constant__ double kGaussPointD[4] = {0.06943184420297371, 0.33000947820757187, 0.6699905217924281,
	0.9305681557970262};
__constant__ double kGaussWeightD[4] = {0.3478548451374539, 0.6521451548625461, 0.6521451548625461,
	0.3478548451374539};

__device__ inline void Bern(double x, double b[4], double db[4], double d2b[4])
{
	const double y = 1.0 - x;
	b[0] = y*y*y; b[1] = 3.0*y*y*x; b[2] = 3.0*y*x*x; b[3] = x*x*x;
	db[0] = -3.0*y*y; db[1] = 3.0 - 12.0*x + 9.0*x*x; db[2] = 3.0*(2.0 - 3.0*x)*x; db[3] = 3.0*x*x;
	d2b[0] = 6.0*y; d2b[1] = -12.0 + 18.0*x; d2b[2] = 6.0 - 18.0*x; d2b[3] = 6.0*x;
}

// One workgroup per node: candidate columns = connectivity of every element
// containing the node; sort in LDS; count (pass 0) or write (pass 1) the
// unique values.
__global__ void PatternKernel(int nodes, const int* __restrict__ n2e_offsets, const int* __restrict__ n2e,
	const int* __restrict__ eoff, const int* __restrict__ conn, int capacity, int pass,
	int* __restrict__ row_length, const int* __restrict__ row_offsets, int* __restrict__ columns,
	int* __restrict__ diagonal, int* __restrict__ status)
{
	extern __shared__ int keys[];
	const int node = blockIdx.x;
	if (node >= nodes) return;
	__shared__ int count;
	if (threadIdx.x == 0) {
		int c = 0;
		for (int k = n2e_offsets[node]; k < n2e_offsets[node + 1]; ++k) {
			const int e = n2e[k];
			c += eoff[e + 1] - eoff[e];
		}
		count = c;
	}
	__syncthreads();
	const int total = count;
	if (total > capacity) {
		if (threadIdx.x == 0) atomicOr(status, 1);
		return;
	}
	// gather: element list walked by thread 0 computing offsets is serial;
	// threads copy each element's connectivity in parallel.
	int base = 0;
	for (int k = n2e_offsets[node]; k < n2e_offsets[node + 1]; ++k) {
		const int e = n2e[k];
		const int b = eoff[e], nen = eoff[e + 1] - b;
		for (int a = threadIdx.x; a < nen; a += blockDim.x) keys[base + a] = conn[b + a];
		base += nen;
	}
	for (int i = total + threadIdx.x; i < capacity; i += blockDim.x) keys[i] = 0x7fffffff;
	__syncthreads();
	// bitonic sort of `capacity` (power of two) keys
	for (int size = 2; size <= capacity; size <<= 1)
		for (int stride = size >> 1; stride > 0; stride >>= 1) {
			for (int i = threadIdx.x; i < capacity; i += blockDim.x) {
				const int j = i ^ stride;
				if (j > i) {
					const bool up = (i & size) == 0;
					const int ki = keys[i], kj = keys[j];
					if ((ki > kj) == up) {
						keys[i] = kj;
						keys[j] = ki;
					}
				}
			}
			__syncthreads();
		}
	if (threadIdx.x != 0) return;
	int unique = 0;
	int diag = -1;
	int out = pass ? row_offsets[node] : 0;
	for (int i = 0; i < total; ++i) {
		const int v = keys[i];
		if (i > 0 && v == keys[i - 1]) continue;
		if (v == node) diag = unique;
		if (pass) columns[out + unique] = v;
		++unique;
	}
	if (pass == 0) {
		row_length[node] = unique;
		if (diag < 0) atomicOr(status, 2);
	} else {
		diagonal[node] = out + diag;
	}
}

// pair_block[pair_offsets[e] + a*nen + b] = block index of (conn[a], conn[b]).
__global__ void PairMapKernel(int elements, const int* __restrict__ eoff, const int* __restrict__ conn,
	const long long* __restrict__ pair_offsets, const int* __restrict__ row_offsets,
	const int* __restrict__ columns, int* __restrict__ pair_block, int* __restrict__ status)
{
	const int e = blockIdx.x;
	if (e >= elements) return;
	const int b0 = eoff[e], nen = eoff[e + 1] - b0;
	const long long base = pair_offsets[e];
	for (int p = threadIdx.x; p < nen*nen; p += blockDim.x) {
		const int a = p/nen, b = p - a*nen;
		const int row = conn[b0 + a], col = conn[b0 + b];
		int lo = row_offsets[row], hi = row_offsets[row + 1];
		while (lo < hi) {
			const int mid = lo + (hi - lo)/2;
			if (columns[mid] < col) lo = mid + 1;
			else hi = mid;
		}
		if (lo >= row_offsets[row + 1] || columns[lo] != col) atomicOr(status, 4);
		pair_block[base + p] = lo;
	}
}

// Geometry per (element, q): writes geometry[e][field][q] and det samples.
__global__ void GeometryKernel(int elements, const double* __restrict__ bezier, double* __restrict__ geometry,
	double* __restrict__ block_min, unsigned long long* __restrict__ bad)
{
	const long long index = blockIdx.x*(long long)blockDim.x + threadIdx.x;
	__shared__ double lds_min[256];
	double local_min = INFINITY;
	if (index < (long long)elements*kQ) {
		const int e = static_cast<int>(index/kQ);
		const int q = static_cast<int>(index - (long long)e*kQ);
		const int qx = q & 3, qy = (q >> 2) & 3, qz = q >> 4;
		double bx[4], dbx[4], d2bx[4], by[4], dby[4], d2by[4], bz[4], dbz[4], d2bz[4];
		Bern(kGaussPointD[qx], bx, dbx, d2bx);
		Bern(kGaussPointD[qy], by, dby, d2by);
		Bern(kGaussPointD[qz], bz, dbz, d2bz);
		const double* X = bezier + (size_t)e*192;
		double J[3][3] = {}, H[3][6] = {};  // H[x][00 11 22 01 02 12]
		for (int k = 0; k < 4; ++k)
			for (int j = 0; j < 4; ++j)
				for (int i = 0; i < 4; ++i) {
					const int p = i + 4*j + 16*k;
					const double g0 = dbx[i]*by[j]*bz[k], g1 = bx[i]*dby[j]*bz[k], g2 = bx[i]*by[j]*dbz[k];
					const double h00 = d2bx[i]*by[j]*bz[k], h11 = bx[i]*d2by[j]*bz[k], h22 = bx[i]*by[j]*d2bz[k];
					const double h01 = dbx[i]*dby[j]*bz[k], h02 = dbx[i]*by[j]*dbz[k], h12 = bx[i]*dby[j]*dbz[k];
					for (int x = 0; x < 3; ++x) {
						const double c = X[3*p + x];
						J[x][0] += c*g0; J[x][1] += c*g1; J[x][2] += c*g2;
						H[x][0] += c*h00; H[x][1] += c*h11; H[x][2] += c*h22;
						H[x][3] += c*h01; H[x][4] += c*h02; H[x][5] += c*h12;
					}
				}
		const double det = J[0][0]*(J[1][1]*J[2][2] - J[1][2]*J[2][1]) - J[0][1]*(J[1][0]*J[2][2] - J[1][2]*J[2][0])
			+ J[0][2]*(J[1][0]*J[2][1] - J[1][1]*J[2][0]);
		double inv[3][3] = {};
		if (isfinite(det) && fabs(det) >= 1e-14) {
			const double s = 1.0/det;
			inv[0][0] = (J[1][1]*J[2][2] - J[1][2]*J[2][1])*s;
			inv[0][1] = (J[0][2]*J[2][1] - J[0][1]*J[2][2])*s;
			inv[0][2] = (J[0][1]*J[1][2] - J[0][2]*J[1][1])*s;
			inv[1][0] = (J[1][2]*J[2][0] - J[1][0]*J[2][2])*s;
			inv[1][1] = (J[0][0]*J[2][2] - J[0][2]*J[2][0])*s;
			inv[1][2] = (J[0][2]*J[1][0] - J[0][0]*J[1][2])*s;
			inv[2][0] = (J[1][0]*J[2][1] - J[1][1]*J[2][0])*s;
			inv[2][1] = (J[0][1]*J[2][0] - J[0][0]*J[2][1])*s;
			inv[2][2] = (J[0][0]*J[1][1] - J[0][1]*J[1][0])*s;
		}
		double L[6];
		L[0] = inv[0][0]*inv[0][0] + inv[0][1]*inv[0][1] + inv[0][2]*inv[0][2];
		L[1] = inv[1][0]*inv[1][0] + inv[1][1]*inv[1][1] + inv[1][2]*inv[1][2];
		L[2] = inv[2][0]*inv[2][0] + inv[2][1]*inv[2][1] + inv[2][2]*inv[2][2];
		L[3] = inv[0][0]*inv[1][0] + inv[0][1]*inv[1][1] + inv[0][2]*inv[1][2];
		L[4] = inv[0][0]*inv[2][0] + inv[0][1]*inv[2][1] + inv[0][2]*inv[2][2];
		L[5] = inv[1][0]*inv[2][0] + inv[1][1]*inv[2][1] + inv[1][2]*inv[2][2];
		double M[3];
		for (int x = 0; x < 3; ++x)
			M[x] = H[x][0]*L[0] + H[x][1]*L[1] + H[x][2]*L[2] + 2.0*(H[x][3]*L[3] + H[x][4]*L[4] + H[x][5]*L[5]);
		double* g = geometry + (size_t)e*kGeomFields*kQ + q;
		const double w = kGaussWeightD[qx]*kGaussWeightD[qy]*kGaussWeightD[qz];
		g[0] = w*(0.125*det);
		for (int a = 0; a < 3; ++a)
			for (int x = 0; x < 3; ++x) g[(1 + 3*a + x)*kQ] = inv[a][x];
		for (int k = 0; k < 6; ++k) g[(10 + k)*kQ] = L[k];
		for (int c = 0; c < 3; ++c) g[(16 + c)*kQ] = -(inv[c][0]*M[0] + inv[c][1]*M[1] + inv[c][2]*M[2]);
		g[19*kQ] = 0.0;
		const double d8 = 0.125*det;
		local_min = d8;
		if (!isfinite(d8) || d8 <= 0.0) atomicAdd(bad, 1ull);
	}
	lds_min[threadIdx.x] = local_min;
	__syncthreads();
	for (int s = blockDim.x/2; s > 0; s >>= 1) {
		if (threadIdx.x < s) lds_min[threadIdx.x] = fmin(lds_min[threadIdx.x], lds_min[threadIdx.x + s]);
		__syncthreads();
	}
	if (threadIdx.x == 0) block_min[blockIdx.x] = lds_min[0];
}

int NextPow2(int v)
{
	int p = 1;
	while (p < v) p <<= 1;
	return p;
}

}  // namespace

void BuildMesh(tfi_mesh* m)
{
	tfi_ctx* ctx = m->ctx;
	const int N = m->nodes, E = m->elements;
	// node -> element adjacency (elements ascending)
	std::vector<int> n2e_off(N + 1, 0);
	for (int e = 0; e < E; ++e)
		for (int a = m->h_eoff[e]; a < m->h_eoff[e + 1]; ++a) ++n2e_off[m->h_conn[a] + 1];
	for (int n = 0; n < N; ++n) n2e_off[n + 1] += n2e_off[n];
	std::vector<int> n2e(n2e_off[N]);
	{
		std::vector<int> fill(n2e_off.begin(), n2e_off.end() - 1);
		for (int e = 0; e < E; ++e)
			for (int a = m->h_eoff[e]; a < m->h_eoff[e + 1]; ++a) n2e[fill[m->h_conn[a]]++] = e;
	}
	int max_candidates = 0;
	for (int n = 0; n < N; ++n) {
		int c = 0;
		for (int k = n2e_off[n]; k < n2e_off[n + 1]; ++k) c += m->h_eoff[n2e[k] + 1] - m->h_eoff[n2e[k]];
		max_candidates = std::max(max_candidates, c);
	}
	const int capacity = NextPow2(std::max(max_candidates, 2));
	if (capacity > 16384) throw Error("node couples too many element bases for the LDS pattern builder");
	DevBuf<int> d_n2e_off(ctx, N + 1), d_n2e(ctx, n2e.size()), row_len(ctx, N), status(ctx, 1);
	d_n2e_off.Upload(n2e_off.data(), N + 1);
	d_n2e.Upload(n2e.data(), n2e.size());
	FillInt(ctx, status.data(), 0, 1);
	const size_t lds = capacity*sizeof(int);
	hipLaunchKernelGGL(PatternKernel, dim3(N), dim3(256), lds, ctx->stream, N, d_n2e_off.data(), d_n2e.data(),
		m->eoff.data(), m->conn.data(), capacity, 0, row_len.data(), nullptr, nullptr, nullptr, status.data());
	TFI_LAUNCH_CHECK("PatternKernel count");
	int st = 0;
	status.Download(&st, 1);
	if (st & 1) throw Error("pattern candidate capacity exceeded");
	if (st & 2) throw Error("block pattern is missing a diagonal entry");
	std::vector<int> lens(N), ro(N + 1, 0);
	row_len.Download(lens.data(), N);
	long long total = 0;
	for (int n = 0; n < N; ++n) {
		total += lens[n];
		if (total > std::numeric_limits<int>::max()) throw Error("block CSR has more than INT_MAX blocks");
		ro[n + 1] = static_cast<int>(total);
	}
	m->nnzb = total;
	m->row_offsets.Allocate(ctx, N + 1);
	m->row_offsets.Upload(ro.data(), N + 1);
	m->columns.Allocate(ctx, total);
	m->diagonal.Allocate(ctx, N);
	hipLaunchKernelGGL(PatternKernel, dim3(N), dim3(256), lds, ctx->stream, N, d_n2e_off.data(), d_n2e.data(),
		m->eoff.data(), m->conn.data(), capacity, 1, nullptr, m->row_offsets.data(), m->columns.data(),
		m->diagonal.data(), status.data());
	TFI_LAUNCH_CHECK("PatternKernel fill");

	// pair -> block map
	std::vector<long long> poff(E + 1, 0);
	for (int e = 0; e < E; ++e) {
		const long long nen = m->h_eoff[e + 1] - m->h_eoff[e];
		poff[e + 1] = poff[e] + nen*nen;
	}
	m->element_pairs = poff[E];
	m->pair_offsets.Allocate(ctx, E + 1);
	m->pair_offsets.Upload(poff.data(), E + 1);
	m->pair_block.Allocate(ctx, poff[E]);
	hipLaunchKernelGGL(PairMapKernel, dim3(E), dim3(256), 0, ctx->stream, E, m->eoff.data(), m->conn.data(),
		m->pair_offsets.data(), m->row_offsets.data(), m->columns.data(), m->pair_block.data(), status.data());
	TFI_LAUNCH_CHECK("PairMapKernel");
	status.Download(&st, 1);
	if (st & 4) throw Error("element pair missing from the block pattern");

	// greedy colouring in element order: no two elements of a colour share a node
	m->h_color.assign(E, -1);
	std::vector<int> mark;
	int colors = 0;
	for (int e = 0; e < E; ++e) {
		if ((int)mark.size() < colors + 1) mark.resize(colors + 1, -1);
		for (int a = m->h_eoff[e]; a < m->h_eoff[e + 1]; ++a) {
			const int n = m->h_conn[a];
			for (int k = n2e_off[n]; k < n2e_off[n + 1]; ++k) {
				const int c = m->h_color[n2e[k]];
				if (c >= 0) mark[c] = e;
			}
		}
		int c = 0;
		while (c < colors && mark[c] == e) ++c;
		if (c == colors) ++colors;
		m->h_color[e] = c;
	}
	m->colors = colors;
	m->color_offsets.assign(colors + 1, 0);
	for (int e = 0; e < E; ++e) ++m->color_offsets[m->h_color[e] + 1];
	for (int c = 0; c < colors; ++c) m->color_offsets[c + 1] += m->color_offsets[c];
	std::vector<int> order(E);
	{
		std::vector<int> fill(m->color_offsets.begin(), m->color_offsets.end() - 1);
		for (int e = 0; e < E; ++e) order[fill[m->h_color[e]]++] = e;
	}
	m->color_elements.Allocate(ctx, E);
	m->color_elements.Upload(order.data(), E);

	// geometry
	m->geometry.Allocate(ctx, (size_t)E*kGeomFields*kQ);
	const long long items = (long long)E*kQ;
	const int blocks = static_cast<int>((items + 255)/256);
	DevBuf<double> bmin(ctx, blocks);
	DevBuf<unsigned long long> bad(ctx, 1);
	Check(hipMemsetAsync(bad.data(), 0, sizeof(unsigned long long), ctx->stream), "memset");
	hipLaunchKernelGGL(GeometryKernel, dim3(blocks), dim3(256), 0, ctx->stream, E, m->bezier.data(),
		m->geometry.data(), bmin.data(), bad.data());
	TFI_LAUNCH_CHECK("GeometryKernel");
	std::vector<double> hmin(blocks);
	bmin.Download(hmin.data(), blocks);
	unsigned long long hb = 0;
	bad.Download(&hb, 1);
	m->min_det = *std::min_element(hmin.begin(), hmin.end());
	m->bad_samples = static_cast<long long>(hb);
	SetBasisTable(m, -1);
}

}  // namespace tfi

using namespace tfi;

extern "C" {

int tfi_mesh_create(tfi_ctx* ctx, int nodes, int elements, const int32_t* element_offsets,
	const int32_t* connectivity, const int32_t* extraction_offsets, const uint8_t* extraction_columns,
	const double* extraction_values, const double* bezier_points, tfi_mesh** out)
{
	if (!ctx || !out) return TFI_ERROR;
	*out = nullptr;
	auto m = std::make_unique<tfi_mesh>();
	try {
		Check(hipSetDevice(ctx->device), "hipSetDevice");
		if (nodes <= 0 || elements <= 0) throw Error("mesh needs positive node and element counts");
		m->ctx = ctx;
		m->nodes = nodes;
		m->elements = elements;
		m->h_eoff.assign(element_offsets, element_offsets + elements + 1);
		if (m->h_eoff[0] != 0) throw Error("element offsets must start at 0");
		const long long entries = m->h_eoff[elements];
		m->basis_entries = entries;
		m->h_conn.assign(connectivity, connectivity + entries);
		m->h_xoff.assign(extraction_offsets, extraction_offsets + entries + 1);
		const long long nnz = m->h_xoff[entries];
		m->extraction_nonzeros = nnz;
		m->h_xcol.assign(extraction_columns, extraction_columns + nnz);
		m->h_xval.assign(extraction_values, extraction_values + nnz);
		m->h_bezier.assign(bezier_points, bezier_points + (size_t)elements*192);
		for (int e = 0; e < elements; ++e) {
			const int nen = m->h_eoff[e + 1] - m->h_eoff[e];
			if (nen <= 0) throw Error("element with no basis functions");
			if (nen > kMaxBasis) throw Error("kernels support at most 80 basis functions per element");
			m->max_nen = std::max(m->max_nen, nen);
			std::vector<int> c(m->h_conn.begin() + m->h_eoff[e], m->h_conn.begin() + m->h_eoff[e + 1]);
			for (int v : c)
				if (v < 0 || v >= nodes) throw Error("connectivity exceeds node count");
			std::sort(c.begin(), c.end());
			if (std::adjacent_find(c.begin(), c.end()) != c.end())
				throw Error("element connectivity repeats a node");
		}
		for (long long a = 0; a < entries; ++a)
			if (m->h_xoff[a + 1] < m->h_xoff[a]) throw Error("extraction offsets decrease");
		for (uint8_t c : m->h_xcol)
			if (c >= 64) throw Error("invalid extraction column");
		m->eoff.Allocate(ctx, elements + 1);
		m->eoff.Upload(m->h_eoff.data(), elements + 1);
		m->conn.Allocate(ctx, entries);
		m->conn.Upload(m->h_conn.data(), entries);
		m->xoff.Allocate(ctx, entries + 1);
		m->xoff.Upload(m->h_xoff.data(), entries + 1);
		m->xcol.Allocate(ctx, nnz);
		m->xcol.Upload(m->h_xcol.data(), nnz);
		m->xval.Allocate(ctx, nnz);
		m->xval.Upload(m->h_xval.data(), nnz);
		m->bezier.Allocate(ctx, (size_t)elements*192);
		m->bezier.Upload(m->h_bezier.data(), (size_t)elements*192);
		BuildMesh(m.get());
		*out = m.release();
		return TFI_OK;
	} catch (const std::exception& e) {
		ctx->error = e.what();
		return TFI_ERROR;
	}
}

void tfi_mesh_destroy(tfi_mesh* mesh)
{
	if (mesh && mesh->ctx) (void)hipStreamSynchronize(mesh->ctx->stream);
	delete mesh;
}

int tfi_mesh_stats_get(tfi_mesh* m, tfi_mesh_stats* out)
{
	if (!m || !out) return TFI_ERROR;
	out->nodes = m->nodes;
	out->elements = m->elements;
	out->maximum_basis = m->max_nen;
	out->basis_entries = m->basis_entries;
	out->extraction_nonzeros = m->extraction_nonzeros;
	out->block_nonzeros = m->nnzb;
	out->element_pairs = m->element_pairs;
	out->colors = m->colors;
	out->minimum_determinant = m->min_det;
	out->bad_samples = m->bad_samples;
	out->basis_table_bytes = static_cast<long long>(m->basis_table.size()*sizeof(double));
	return TFI_OK;
}

int tfi_mesh_basis_table(tfi_mesh* m, int enable)
{
	if (!m) return TFI_ERROR;
	try {
		SetBasisTable(m, enable);
		return TFI_OK;
	} catch (const std::exception& e) {
		m->ctx->error = e.what();
		return TFI_ERROR;
	}
}

int tfi_mesh_pattern(tfi_mesh* m, int32_t* row_offsets, int32_t* columns, int32_t* diagonal)
{
	try {
		if (row_offsets) m->row_offsets.Download(row_offsets, m->nodes + 1);
		if (columns) m->columns.Download(columns, m->nnzb);
		if (diagonal) m->diagonal.Download(diagonal, m->nodes);
		return TFI_OK;
	} catch (const std::exception& e) {
		m->ctx->error = e.what();
		return TFI_ERROR;
	}
}

int tfi_mesh_colors(tfi_mesh* m, int32_t* element_color)
{
	if (!m || !element_color) return TFI_ERROR;
	std::copy(m->h_color.begin(), m->h_color.end(), element_color);
	return TFI_OK;
}

}  // extern "C"
