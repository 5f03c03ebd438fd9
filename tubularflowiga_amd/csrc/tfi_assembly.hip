// Element assembly of the stabilised (VMS) Navier-Stokes residual and tangent
// (legacy tau-frozen or exact consistent), with a deterministic scatter.
//
// Arithmetic restated from solvers/cuda/include/IgaCudaKernels.cuh:239-409
// (StabilizationDevice + AssembleNavierStokesKernel, legacy tangent) and the
// CPU element solvers/cpu/include/NavierStokesElement.hpp:121-313 (residual and
// exact tangent, which differentiates tau_M, tau_C and the strong viscous term).  The work is organised
// for CDNA4 instead of the reference's thread-per-pair tiles with a serial
// thread-0 section:
//   * one workgroup (4 waves) per (element, tile of test rows); the element's
//     dense extraction operator, nodal state and per-point data sit in LDS;
//   * the 64 Gauss points are processed in two slabs of 32;
//   * basis values / gradients / Laplacians are evaluated per (basis, point)
//     by sum factorisation of the cubic Bernstein tensor product
//     (328 FMA instead of 640 per point, no table reads);
//   * per-point fields (u, grad u, lap u, u_prev) are reduced by 8 lanes per
//     point in fixed order; tau_M, tau_C and the fine scales follow;
//   * the tangent accumulates in registers on 2x2 (test, trial) tiles
//     (36 FMA per pair per point, 17 accumulators per pair);
//   * the scatter adds each element's blocks into the pattern with plain
//     read-modify-write: elements of one colour share no node, colours run
//     in order, so every block sum has a fixed order (no atomics).
#include "tfi_internal.hpp"

#include <algorithm>
#include <cmath>

namespace tfi {

namespace {

constexpr int kThreads = 256;
constexpr size_t kLdsLimit = 160*1024;  // gfx950 LDS per workgroup
constexpr int kQd = 56;       // per-point record in LDS
constexpr int kBasisRec = 7;  // N, dN/dx, dN/dy, dN/dz, lap N, u.grad N, w.grad N (w = u + u')
static_assert(kMaxBasis % 2 == 0, "the connectivity ints fill whole doubles of the LDS carve-up");

__constant__ double kGp[4] = {0.06943184420297371, 0.33000947820757187, 0.6699905217924281, 0.9305681557970262};

__device__ inline void Bern3(double x, double b[4], double db[4], double d2b[4])
{
	const double y = 1.0 - x;
	b[0] = y*y*y; b[1] = 3.0*y*y*x; b[2] = 3.0*y*x*x; b[3] = x*x*x;
	db[0] = -3.0*y*y; db[1] = 3.0 - 12.0*x + 9.0*x*x; db[2] = 3.0*(2.0 - 3.0*x)*x; db[3] = 3.0*x*x;
	d2b[0] = 6.0*y; d2b[1] = -12.0 + 18.0*x; d2b[2] = 6.0 - 18.0*x; d2b[3] = 6.0*x;
}

constexpr int kTableRec = 5;  // N, dN/dx, dN/dy, dN/dz, lap N
constexpr int kPointGeom = 10;  // geometry fields the point record reads: measure, J^-1

// Value, physical gradient and Laplacian of the basis function with Bezier
// extraction row c[64] at element point q (Bernstein sum-factorisation in
// xi, eta, zeta; gq = geometry record of the point, field stride kQ).
__device__ inline void EvalBasis(const double* __restrict__ c, const double* __restrict__ gq, int q,
	double out[kTableRec])
{
	double bx[4], dbx[4], d2bx[4], by[4], dby[4], d2by[4], bz[4], dbz[4], d2bz[4];
	Bern3(kGp[q & 3], bx, dbx, d2bx);
	Bern3(kGp[(q >> 2) & 3], by, dby, d2by);
	Bern3(kGp[q >> 4], bz, dbz, d2bz);
	double v = 0, g0 = 0, g1 = 0, g2 = 0, h00 = 0, h11 = 0, h22 = 0, h01 = 0, h02 = 0, h12 = 0;
	for (int k = 0; k < 4; ++k) {
		double s00 = 0, s10 = 0, s01 = 0, s20 = 0, s11 = 0, s02 = 0;
		for (int j = 0; j < 4; ++j) {
			const double* cr = c + 16*k + 4*j;
			const double x0 = cr[0]*bx[0] + cr[1]*bx[1] + cr[2]*bx[2] + cr[3]*bx[3];
			const double x1 = cr[0]*dbx[0] + cr[1]*dbx[1] + cr[2]*dbx[2] + cr[3]*dbx[3];
			const double x2 = cr[0]*d2bx[0] + cr[1]*d2bx[1] + cr[2]*d2bx[2] + cr[3]*d2bx[3];
			s00 += x0*by[j]; s10 += x1*by[j]; s20 += x2*by[j];
			s01 += x0*dby[j]; s11 += x1*dby[j]; s02 += x0*d2by[j];
		}
		v += s00*bz[k]; g0 += s10*bz[k]; g1 += s01*bz[k]; g2 += s00*dbz[k];
		h00 += s20*bz[k]; h11 += s02*bz[k]; h22 += s00*d2bz[k];
		h01 += s11*bz[k]; h02 += s10*dbz[k]; h12 += s01*dbz[k];
	}
	double inv[9];
#pragma unroll
	for (int i = 0; i < 9; ++i) inv[i] = gq[(1 + i)*kQ];
	out[0] = v;
	out[1] = g0*inv[0] + g1*inv[3] + g2*inv[6];
	out[2] = g0*inv[1] + g1*inv[4] + g2*inv[7];
	out[3] = g0*inv[2] + g1*inv[5] + g2*inv[8];
	out[4] = h00*gq[10*kQ] + h11*gq[11*kQ] + h22*gq[12*kQ]
		+ 2.0*(h01*gq[13*kQ] + h02*gq[14*kQ] + h12*gq[15*kQ])
		+ g0*gq[16*kQ] + g1*gq[17*kQ] + g2*gq[18*kQ];
}

// Per-point record from the field sums s (u, p, grad u, grad p, lap u, u_prev
// at the point; 22 values) and the point's geometry record gq (fields 0..9:
// measure and J^-1, field stride kGs): r[0] measure, r[1..2] tau_M, tau_C, r[3..5] u, r[6..8] rho (du/dt +
// w.grad u), r[9..17] stress rows, r[18] div u, r[19..21] u', r[22..24] w =
// u + u', r[25..33] grad u and, with exact_tangent, r[34..54] the
// linearisation of u' (K, R, c1) the consistent tangent needs.
template <int kGs>
__device__ inline void PointRecord(const double* __restrict__ s, const double* __restrict__ gq, double rho, double mu,
	double nu, double dt, bool exact_tangent, double* __restrict__ r)
{
	double inv[3][3];
	for (int a = 0; a < 3; ++a)
		for (int x = 0; x < 3; ++x) inv[a][x] = gq[(1 + 3*a + x)*kGs];
	const double u[3] = {s[0], s[1], s[2]};
	const double p = s[3];
	double grad[4][3];
	for (int f = 0; f < 4; ++f)
		for (int d = 0; d < 3; ++d) grad[f][d] = s[4 + 3*f + d];
	double dudt[3];
	for (int c = 0; c < 3; ++c) dudt[c] = dt > 0.0 ? (u[c] - s[19 + c])/dt : 0.0;
	double G[3][3] = {}, gv[3] = {};
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j) {
			for (int k = 0; k < 3; ++k) G[i][j] += inv[k][i]*inv[k][j];
			gv[i] += inv[j][i];
		}
	double GG = 0.0, gg = 0.0, uGu = 0.0;
	for (int i = 0; i < 3; ++i) {
		gg += gv[i]*gv[i];
		for (int j = 0; j < 3; ++j) {
			GG += G[i][j]*G[i][j];
			uGu += u[i]*G[i][j]*u[j];
		}
	}
	const double temporal = dt > 0.0 ? 4.0/(dt*dt) : 0.0;
	const double tau_m = 1.0/sqrt(temporal + uGu + (1.0/12.0)*nu*nu*GG);
	const double tau_c = 1.0/(tau_m*gg);
	double fu[3];
	for (int c = 0; c < 3; ++c) {
		const double conv = u[0]*grad[c][0] + u[1]*grad[c][1] + u[2]*grad[c][2];
		fu[c] = -tau_m*(dudt[c] + conv + grad[3][c]/rho - nu*s[16 + c]);
	}
	const double div = grad[0][0] + grad[1][1] + grad[2][2];
	const double fp = -rho*tau_c*div;
	r[0] = gq[0];
	r[1] = tau_m;
	r[2] = tau_c;
	for (int c = 0; c < 3; ++c) r[3 + c] = u[c];
	const double w[3] = {u[0] + fu[0], u[1] + fu[1], u[2] + fu[2]};
	for (int c = 0; c < 3; ++c) {
		r[6 + c] = rho*(dudt[c] + w[0]*grad[c][0] + w[1]*grad[c][1] + w[2]*grad[c][2]);
		for (int d = 0; d < 3; ++d)
			r[9 + 3*c + d] = mu*(grad[c][d] + grad[d][c]) - rho*fu[c]*w[d] - (c == d ? p + fp : 0.0);
	}
	r[18] = div;
	for (int c = 0; c < 3; ++c) r[19 + c] = fu[c];
	for (int c = 0; c < 3; ++c) r[22 + c] = w[c];
	for (int c = 0; c < 3; ++c)
		for (int d = 0; d < 3; ++d) r[25 + 3*c + d] = grad[c][d];
	if (exact_tangent) {
		// d(u')/d(u_f) = N_b K[.][f] - delta tau_M s_b with
		// K_cf = tau_M^3 strong_c (G u)_f - tau_M du_c/dx_f, and
		// R = grad u + grad u K, c1_f = rho tau_C tau_M^2 div (G u)_f.
		double Gu[3];
		for (int i = 0; i < 3; ++i) Gu[i] = G[i][0]*u[0] + G[i][1]*u[1] + G[i][2]*u[2];
		const double tm3 = tau_m*tau_m*tau_m;
		double K[3][3];
		for (int c = 0; c < 3; ++c) {
			const double strong = -fu[c]/tau_m;
			for (int f = 0; f < 3; ++f) K[c][f] = tm3*strong*Gu[f] - tau_m*grad[c][f];
		}
		for (int c = 0; c < 3; ++c)
			for (int f = 0; f < 3; ++f) {
				r[34 + 3*c + f] = K[c][f];
				r[43 + 3*c + f] = grad[c][f] + grad[c][0]*K[0][f] + grad[c][1]*K[1][f] + grad[c][2]*K[2][f];
			}
		for (int f = 0; f < 3; ++f) r[52 + f] = rho*tau_c*tau_m*tau_m*div*Gu[f];
	}
}

struct AsmArgs {
	const int* eoff;
	const int* conn;
	const int* xoff;
	const uint8_t* xcol;
	const double* xval;
	const double* geometry;
	const long long* pair_offsets;
	const int* pair_block;
	const int* elements;   // element ids of this colour
	int count;
	int a_tile;
	int tiles;
	const double* state;
	const double* previous;
	double rho, mu, dt;
	double* jac;
	double* negres;
	const double* table;   // per-element basis records [q][a][kTableRec] at eoff[e]*kQ*kTableRec, or null
	int qstride;           // doubles between the basis records of consecutive points (odd)
	int max_nen;           // largest element of the mesh (LDS carve-up)
};

typedef double doublex4 __attribute__((ext_vector_type(4)));

// basis record of function a at slab point ql: [N, dN/dx, dN/dy, dN/dz, lap N, u.grad N, w.grad N]
__device__ inline double* Brec(double* basis, int qstride, int a, int ql)
{
	return basis + (size_t)ql*qstride + (size_t)a*kBasisRec;
}

// C[tj] += A B_tj for one test-function tile and the two trial tiles.
__device__ inline void Term2(doublex4 (&c)[2], double a, double b0, double b1)
{
	c[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b0, c[0], 0, 0, 0);
	c[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b1, c[1], 0, 0, 0);
}

// kMfma: elements with at most 32 basis functions, one workgroup per element.
// The tangent is a sum over quadrature points of (test factor) x (trial
// factor) products, so each of the 16 block components is a small GEMM,
// T_c[a][b] = sum_(q, term) A_term(a, q) B_term(q, b), run on FP64 MFMA with
// K = 4 quadrature points per instruction (v_mfma_f64_16x16x4f64: lane l
// holds A[l%16][l/16] and B[l/16][l%16]).  The workgroup has 8 waves, two per
// SIMD, so that one wave's operand building overlaps the other's MFMAs: wave
// w owns test field w%4 (row w%4 of the 4x4 block) for the 16 test functions
// of tile w/4, lanes build the operands of their (basis function, point) from
// the LDS records, and two 16 x 16 output tiles per wave cover the trial
// functions of the element's (padded) 32 x 32 basis pairs.
// kJac: residual and tangent; otherwise the residual alone.
// kSlab: quadrature points per slab (32, or 16 / 8 for elements whose basis
// records would not fit the LDS at 32 points: nen up to 80).
template <bool kExact, bool kMfma, bool kJac, int kSlab>
__global__ void __launch_bounds__(kMfma ? 512 : 256) AssembleNsKernel(AsmArgs args)
{
	static_assert(kSlab == 32 || kSlab == 16 || kSlab == 8, "slab of 8, 16 or 32 points");
	static_assert(!kMfma || kSlab == 32, "the MFMA tangent works on slabs of 32 points");
	constexpr int kT = kMfma ? 512 : 256;  // threads of the workgroup
	extern __shared__ double lds[];
	const int tid = threadIdx.x;
	const int item = blockIdx.x/args.tiles;
	const int tile = blockIdx.x - item*args.tiles;
	if (item >= args.count) return;
	const int e = args.elements[item];
	const int b0 = args.eoff[e];
	const int nen = args.eoff[e + 1] - b0;
	const int a0 = tile*args.a_tile;
	const int a1 = min(nen, a0 + args.a_tile);
	const double rho = args.rho, mu = args.mu, dt = args.dt;
	const double nu = mu/rho;
	const double idt = dt > 0.0 ? 1.0/dt : 0.0;

	// LDS carve-up (the MFMA scatter stages the element's blocks over basis and qd)
	const int qstride = args.qstride;
	const double* table = args.table;
	double* nodal = lds;                                   // [nen][8]
	double* resid = nodal + kMaxBasis*8;                    // [a_tile][4]
	int* lconn = reinterpret_cast<int*>(resid + 4*kMaxBasis);
	double* gslab = resid + 4*kMaxBasis + kMaxBasis/2;      // [kPointGeom][kSlab] measure, J^-1 of the slab
	double* C = gslab + (kMfma ? kPointGeom*kSlab : 0);     // [nen][64], only without the table
	double* basis = C + (table ? 0 : (size_t)args.max_nen*64);  // [kSlab][qstride] records
	double* qd = basis + (size_t)kSlab*qstride;             // [kSlab][kQd]

	const double* geo = args.geometry + (size_t)e*kGeomFields*kQ;
	// P1 of a slab: basis values, gradients and Laplacians per (a, point), read
	// from the precomputed table when there is one, else evaluated here
	auto stage = [&](int slab) {
		const int q0 = slab*kSlab;
		if (kMfma)  // the point records read the slab's measure and J^-1 from LDS
			for (int i = tid; i < kPointGeom*kSlab; i += kT) {
				const int f = i/kSlab, ql = i - f*kSlab;
				gslab[i] = geo[f*kQ + q0 + ql];
			}
		if (kMfma || table) {  // the MFMA variants run only with the table
			const int rec = nen*kTableRec;
			const double* src = table + ((size_t)b0*kQ + (size_t)q0*nen)*kTableRec;
			// entry i = (point ql, offset r) of this thread, advanced by kT entries
			// per pass without a division (dr < rec, so one carry suffices)
			const int dq = kT/rec, dr = kT - dq*rec;
			int ql = tid/rec, r = tid - ql*rec;
#pragma unroll 4
			for (int i = tid; i < kSlab*rec; i += kT) {
				const int a = r/kTableRec, f = r - a*kTableRec;
				Brec(basis, qstride, a, ql)[f] = src[i];
				ql += dq;
				r += dr;
				if (r >= rec) {
					r -= rec;
					++ql;
				}
			}
		} else {
			for (int it = tid; it < nen*kSlab; it += kT) {
				const int a = it/kSlab, ql = it - a*kSlab;
				EvalBasis(C + 64*a, geo + q0 + ql, q0 + ql, Brec(basis, qstride, a, ql));
			}
		}
	};

	for (int a = tid; a < nen; a += kT) {
		const int node = args.conn[b0 + a];
		lconn[a] = node;
		for (int f = 0; f < 4; ++f) nodal[8*a + f] = args.state[4*(size_t)node + f];
		for (int f = 0; f < 3; ++f) nodal[8*a + 4 + f] = args.previous[4*(size_t)node + f];
	}
	if (kMfma) stage(0);  // the first slab's loads in flight with the nodal gather
	if (!kMfma && !table)
		for (int i = tid; i < nen*64; i += kT) C[i] = 0.0;
	for (int i = tid; i < 4*kMaxBasis; i += kT) resid[i] = 0.0;
	if (kMfma)  // padding functions nen..31 stay zero in every slab
		for (int i = tid; i < kSlab*(32 - nen)*kBasisRec; i += kT) {
			const int ql = i/((32 - nen)*kBasisRec), rest = i - ql*(32 - nen)*kBasisRec;
			Brec(basis, qstride, nen, ql)[rest] = 0.0;
		}
	__syncthreads();
	if (!kMfma && !table) {
		for (int a = tid; a < nen; a += kT)
			for (int k = args.xoff[b0 + a]; k < args.xoff[b0 + a + 1]; ++k) C[64*a + args.xcol[k]] = args.xval[k];
		__syncthreads();
	}

	// pair tile of this thread: rows ra0, ra0+1 ; columns rb0, rb0+1
	const int tiles_b = (nen + 1)/2;
	const int tiles_a = (a1 - a0 + 1)/2;
	const bool pair_active = !kMfma && kJac && tid < tiles_a*tiles_b;
	const int ta = pair_active ? tid/tiles_b : 0;
	const int tb = pair_active ? tid - ta*tiles_b : 0;
	const int ra[2] = {a0 + 2*ta, min(a0 + 2*ta + 1, a1 - 1)};
	const int rb[2] = {2*tb, min(2*tb + 1, nen - 1)};
	double dg[4] = {}, uu[4][9] = {}, i3[4][3] = {}, t3[4][3] = {}, k33[4] = {};
	doublex4 acc[4][2];  // MFMA path: [component of this wave][trial tile tj]
	for (int c = 0; c < 4; ++c)
		for (int t = 0; t < 2; ++t) acc[c][t] = doublex4{0.0, 0.0, 0.0, 0.0};

	for (int slab = 0; slab < kQ/kSlab; ++slab) {
		const int q0 = slab*kSlab;
		if (!kMfma || slab > 0) {  // ---- P1 (the MFMA path staged slab 0 above)
			stage(slab);
			__syncthreads();
		}
		// ---- P2: fields at the points (kT/kSlab lanes per point), tau, fine scales
		{
			constexpr int kL = kT/kSlab;
			const int ql = tid/kL, part = tid%kL, q = q0 + ql;
			double s[22] = {};
			for (int a = part; a < nen; a += kL) {
				const double* bb = Brec(basis, qstride, a, ql);
				const double* un = nodal + 8*a;
				for (int f = 0; f < 4; ++f) {
					s[f] += un[f]*bb[0];
					s[4 + 3*f] += un[f]*bb[1];
					s[5 + 3*f] += un[f]*bb[2];
					s[6 + 3*f] += un[f]*bb[3];
				}
				for (int c = 0; c < 3; ++c) {
					s[16 + c] += un[c]*bb[4];
					s[19 + c] += un[4 + c]*bb[0];
				}
			}
#pragma unroll
			for (int i = 0; i < 22; ++i)
#pragma unroll
				for (int o = 1; o < kL; o <<= 1) s[i] += __shfl_xor(s[i], o, 64);
			if (part == 0) {
				if constexpr (kMfma) PointRecord<kSlab>(s, gslab + ql, rho, mu, nu, dt, kExact && kJac, qd + ql*kQd);
				else PointRecord<kQ>(s, geo + q, rho, mu, nu, dt, kExact && kJac, qd + ql*kQd);
			}
		}
		__syncthreads();
		// ---- P3: u.grad N per (a, point); residual rows of this tile
		for (int it = tid; it < ((nen + 1) & ~1)*kSlab; it += kT) {
			const int a = it/kSlab, ql = it - a*kSlab;
			const bool valid = a < nen;
			const double* r = qd + ql*kQd;
			double R[4] = {0, 0, 0, 0};
			if (valid) {
				double* bb = Brec(basis, qstride, a, ql);
				const double N = bb[0], gx = bb[1], gy = bb[2], gz = bb[3];
				bb[5] = r[3]*gx + r[4]*gy + r[5]*gz;
				bb[6] = r[22]*gx + r[23]*gy + r[24]*gz;
				if (a >= a0 && a < a1) {
					const double m = r[0];
					for (int c = 0; c < 3; ++c)
						R[c] = -(N*r[6 + c] + gx*r[9 + 3*c] + gy*r[10 + 3*c] + gz*r[11 + 3*c])*m;
					R[3] = -(N*r[18] - gx*r[19] - gy*r[20] - gz*r[21])*m;
				}
			}
			// kSlab consecutive lanes hold the slab's points of one a: fixed-order tree sum
#pragma unroll
			for (int c = 0; c < 4; ++c)
				for (int off = 1; off < kSlab; off <<= 1) R[c] += __shfl_xor(R[c], off, 64);
			if (valid && ql == 0 && a >= a0 && a < a1)
				for (int c = 0; c < 4; ++c) resid[4*(a - a0) + c] += R[c];
		}
		__syncthreads();
		// ---- P4 (MFMA): 4 points per instruction; wave w: test field w%4, test tile w/4
		if (kMfma && kJac) {
			const int wave = tid >> 6, lane = tid & 63;
			const int fl = lane & 15, kq = lane >> 4;
			const int x = wave & 3, ti = wave >> 2;
			const int y1 = (x + 1)%3, y2 = (x + 2)%3;
			for (int chunk = 0; chunk < kSlab/4; ++chunk) {
				const int ql = chunk*4 + kq;
				const double* r = qd + ql*kQd;
				const double m = r[0], tm = r[1], tc = r[2];
				// trial side: the lane's basis functions fl, fl + 16 at point ql
				double N[2], G[2][3], U[2], W[2], Lp[2];
#pragma unroll
				for (int i = 0; i < 2; ++i) {
					const double* bb = Brec(basis, qstride, fl + 16*i, ql);
					N[i] = bb[0];
					G[i][0] = bb[1]; G[i][1] = bb[2]; G[i][2] = bb[3];
					Lp[i] = bb[4];
					U[i] = bb[5];
					W[i] = bb[6];
				}
				// test side: function 16 ti + fl, one of the two
				const double Na = ti ? N[1] : N[0], Ua = ti ? U[1] : U[0], Wa = ti ? W[1] : W[0];
				double mG[3];
#pragma unroll
				for (int d = 0; d < 3; ++d) mG[d] = m*(ti ? G[1][d] : G[0][d]);
				const double mN = m*Na;
				if (kExact) {
					const double* up = r + 19;
					const double* gu = r + 25;
					const double* K = r + 34;
					const double* R = r + 43;
					const double* c1 = r + 52;
					const double mA = m*Wa;
					double sb[2];
#pragma unroll
					for (int j = 0; j < 2; ++j) sb[j] = N[j]*idt + U[j] - nu*Lp[j];
					if (x < 3) {
						double Zx[3], Xbx[2], Yb[2][3];
#pragma unroll
						for (int y = 0; y < 3; ++y) {
							const double mP = K[y]*mG[0] + K[3 + y]*mG[1] + K[6 + y]*mG[2];
							Zx[y] = rho*(mN*R[3*x + y] - mA*K[3*x + y] - up[x]*(mG[y] + mP));
						}
#pragma unroll
						for (int j = 0; j < 2; ++j) {
#pragma unroll
							for (int y = 0; y < 3; ++y) Yb[j][y] = rho*tc*G[j][y] + c1[y]*N[j];
							Xbx[j] = mu*G[j][x] + rho*tm*up[x]*sb[j];
						}
						// (x, x)
						Term2(acc[0], mG[x], Xbx[0] + Yb[0][x] + mu*G[0][x], Xbx[1] + Yb[1][x] + mu*G[1][x]);
						Term2(acc[0], mu*mG[y1], G[0][y1], G[1][y1]);
						Term2(acc[0], mu*mG[y2], G[0][y2], G[1][y2]);
						Term2(acc[0], Zx[x], N[0], N[1]);
						Term2(acc[0], rho*tm*(mA - gu[4*x]*mN), sb[0], sb[1]);
						Term2(acc[0], rho*mN, N[0]*idt + W[0], N[1]*idt + W[1]);
						// (x, y1), (x, y2)
						Term2(acc[1], mG[y1], Xbx[0], Xbx[1]);
						Term2(acc[1], mG[x], Yb[0][y1], Yb[1][y1]);
						Term2(acc[1], Zx[y1], N[0], N[1]);
						Term2(acc[1], -rho*tm*gu[3*x + y1]*mN, sb[0], sb[1]);
						Term2(acc[2], mG[y2], Xbx[0], Xbx[1]);
						Term2(acc[2], mG[x], Yb[0][y2], Yb[1][y2]);
						Term2(acc[2], Zx[y2], N[0], N[1]);
						Term2(acc[2], -rho*tm*gu[3*x + y2]*mN, sb[0], sb[1]);
						// (x, 3)
						Term2(acc[3], -mG[x], N[0], N[1]);
#pragma unroll
						for (int d = 0; d < 3; ++d)
							Term2(acc[3], tm*(up[x]*mG[d] - mN*gu[3*x + d] + (d == x ? mA : 0.0)), G[0][d], G[1][d]);
					} else {
						// (3, y) and (3, 3)
#pragma unroll
						for (int y = 0; y < 3; ++y) {
							const double mP = K[y]*mG[0] + K[3 + y]*mG[1] + K[6 + y]*mG[2];
							Term2(acc[y], mN, G[0][y], G[1][y]);
							Term2(acc[y], -mP, N[0], N[1]);
							Term2(acc[y], tm*mG[y], sb[0], sb[1]);
						}
#pragma unroll
						for (int d = 0; d < 3; ++d) Term2(acc[3], (tm/rho)*mG[d], G[0][d], G[1][d]);
					}
				} else {
					// legacy tau-frozen tangent
					double sb[2];
#pragma unroll
					for (int j = 0; j < 2; ++j) sb[j] = N[j]*idt + U[j];
					if (x < 3) {
						Term2(acc[0], m*rho*(Na + tm*Ua), sb[0], sb[1]);
						Term2(acc[0], (2.0*mu + rho*tc)*mG[x], G[0][x], G[1][x]);
						Term2(acc[0], mu*mG[y1], G[0][y1], G[1][y1]);
						Term2(acc[0], mu*mG[y2], G[0][y2], G[1][y2]);
						Term2(acc[1], mu*mG[y1], G[0][x], G[1][x]);
						Term2(acc[1], rho*tc*mG[x], G[0][y1], G[1][y1]);
						Term2(acc[2], mu*mG[y2], G[0][x], G[1][x]);
						Term2(acc[2], rho*tc*mG[x], G[0][y2], G[1][y2]);
						Term2(acc[3], m*tm*Ua, G[0][x], G[1][x]);
						Term2(acc[3], -mG[x], N[0], N[1]);
					} else {
#pragma unroll
						for (int y = 0; y < 3; ++y) {
							Term2(acc[y], mN, G[0][y], G[1][y]);
							Term2(acc[y], tm*mG[y], sb[0], sb[1]);
						}
#pragma unroll
						for (int d = 0; d < 3; ++d) Term2(acc[3], (tm/rho)*mG[d], G[0][d], G[1][d]);
					}
				}
			}
		}
		// ---- P4: tangent on 2x2 (test, trial) tiles
		if (kExact && pair_active) {
			// consistent tangent, NavierStokesElement.hpp:244-313, in outer-product form
			const double nu_ = nu;
			for (int ql = 0; ql < kSlab; ++ql) {
				const double* r = qd + ql*kQd;
				const double m = r[0], tm = r[1], tc = r[2];
				const double* up = r + 19;
				const double* gu = r + 25;
				const double* K = r + 34;
				const double* R = r + 43;
				const double* c1 = r + 52;
				double mN[2], mG[2][3], mA[2], mP[2][3], Z[2][9];
#pragma unroll
				for (int i = 0; i < 2; ++i) {
					const double* bb = Brec(basis, qstride, ra[i], ql);
					mN[i] = m*bb[0];
					for (int d = 0; d < 3; ++d) mG[i][d] = m*bb[1 + d];
					mA[i] = m*bb[6];
					for (int f = 0; f < 3; ++f) mP[i][f] = K[f]*mG[i][0] + K[3 + f]*mG[i][1] + K[6 + f]*mG[i][2];
					for (int x = 0; x < 3; ++x)
						for (int f = 0; f < 3; ++f)
							Z[i][3*x + f] = rho*(mN[i]*R[3*x + f] - mA[i]*K[3*x + f] - up[x]*(mG[i][f] + mP[i][f]));
				}
				double Nb[2], B[2][3], sb[2], wb[2], Qb[2][3], Xb[2][3], Yb[2][3];
#pragma unroll
				for (int j = 0; j < 2; ++j) {
					const double* bb = Brec(basis, qstride, rb[j], ql);
					Nb[j] = bb[0];
					B[j][0] = bb[1]; B[j][1] = bb[2]; B[j][2] = bb[3];
					sb[j] = Nb[j]*idt + bb[5] - nu_*bb[4];
					wb[j] = bb[6];
					for (int x = 0; x < 3; ++x) {
						Qb[j][x] = gu[3*x]*B[j][0] + gu[3*x + 1]*B[j][1] + gu[3*x + 2]*B[j][2];
						Xb[j][x] = mu*B[j][x] + rho*tm*up[x]*sb[j];
						Yb[j][x] = rho*tc*B[j][x] + c1[x]*Nb[j];
					}
				}
#pragma unroll
				for (int i = 0; i < 2; ++i)
#pragma unroll
					for (int j = 0; j < 2; ++j) {
						const int pr = 2*i + j;
						const double gd = mG[i][0]*B[j][0] + mG[i][1]*B[j][1] + mG[i][2]*B[j][2];
						dg[pr] += rho*mN[i]*(Nb[j]*idt + wb[j]) + mu*gd + rho*tm*mA[i]*sb[j];
						k33[pr] += (tm/rho)*gd;
						const double c4 = -rho*tm*mN[i]*sb[j];
#pragma unroll
						for (int x = 0; x < 3; ++x) {
#pragma unroll
							for (int y = 0; y < 3; ++y)
								uu[pr][3*x + y] += mG[i][y]*Xb[j][x] + mG[i][x]*Yb[j][y] + Nb[j]*Z[i][3*x + y] + c4*gu[3*x + y];
							i3[pr][x] += -mG[i][x]*Nb[j] - tm*mN[i]*Qb[j][x] + tm*mA[i]*B[j][x] + tm*up[x]*gd;
							t3[pr][x] += mN[i]*B[j][x] - Nb[j]*mP[i][x] + tm*sb[j]*mG[i][x];
						}
					}
			}
		}
		if (!kExact && pair_active) {
			for (int ql = 0; ql < kSlab; ++ql) {
				const double* r = qd + ql*kQd;
				const double m = r[0], tm = r[1], tc = r[2];
				const double rtc = rho*tc, tmr = tm/rho;
				double mA[2][3], al[2], be[2], ga[2], P[2][3], Qv[2][3], T[2][3];
#pragma unroll
				for (int i = 0; i < 2; ++i) {
					const double* bb = Brec(basis, qstride, ra[i], ql);
					const double N = bb[0], gu = bb[5];
					for (int d = 0; d < 3; ++d) {
						mA[i][d] = m*bb[1 + d];
						P[i][d] = mu*mA[i][d];
						Qv[i][d] = rtc*mA[i][d];
						T[i][d] = tm*mA[i][d];
					}
					al[i] = m*rho*(N + tm*gu);
					be[i] = m*tm*gu;
					ga[i] = m*N;
				}
				double Nb[2], B[2][3], sb[2];
#pragma unroll
				for (int j = 0; j < 2; ++j) {
					const double* bb = Brec(basis, qstride, rb[j], ql);
					Nb[j] = bb[0];
					B[j][0] = bb[1]; B[j][1] = bb[2]; B[j][2] = bb[3];
					sb[j] = Nb[j]*idt + bb[5];
				}
#pragma unroll
				for (int i = 0; i < 2; ++i)
#pragma unroll
					for (int j = 0; j < 2; ++j) {
						const int pr = 2*i + j;
						const double gd = mA[i][0]*B[j][0] + mA[i][1]*B[j][1] + mA[i][2]*B[j][2];
						dg[pr] += al[i]*sb[j] + mu*gd;
						k33[pr] += tmr*gd;
#pragma unroll
						for (int x = 0; x < 3; ++x) {
#pragma unroll
							for (int y = 0; y < 3; ++y) uu[pr][3*x + y] += P[i][y]*B[j][x] + Qv[i][x]*B[j][y];
							i3[pr][x] += be[i]*B[j][x] - mA[i][x]*Nb[j];
							t3[pr][x] += ga[i]*B[j][x] + T[i][x]*sb[j];
						}
					}
			}
		}
		__syncthreads();
	}

	// ---- scatter (exclusive by colour: plain read-modify-write).  The MFMA
	// path stages the element's blocks in LDS (the slab buffers are free after
	// the last slab's barrier; block stride kStageStride = 17 doubles, so that
	// lanes writing neighbouring blocks hit different banks), then eight lanes
	// per block read-modify-write its 128 bytes as 16-byte pieces: whole cache
	// lines per block instead of 16 scattered 8-byte updates from four waves.
	if (kMfma && kJac) {
		constexpr int kStageStride = 17;
		double* stage = basis;
		const int wave = tid >> 6, lane = tid & 63;
		const int x = wave & 3, ti = wave >> 2;
		const int y1 = (x + 1)%3, y2 = (x + 2)%3;
		const int comp[4] = {x < 3 ? 5*x : 12, x < 3 ? 4*x + y1 : 13, x < 3 ? 4*x + y2 : 14, x < 3 ? 4*x + 3 : 15};
#pragma unroll
		for (int c = 0; c < 4; ++c)
#pragma unroll
			for (int tj = 0; tj < 2; ++tj) {
				const int b = 16*tj + (lane & 15);
#pragma unroll
				for (int i = 0; i < 4; ++i) {
					const int a = 16*ti + 4*i + (lane >> 4);  // output row of element i of the lane's accumulator
					if (a < nen && b < nen) stage[(a*nen + b)*kStageStride + comp[c]] = acc[c][tj][i];
				}
			}
		__syncthreads();
		// All of a thread's block slots, then all of its Jacobian pieces, are
		// loaded before the first store (the blocks of one colour are disjoint),
		// so the read-modify-write costs two memory round trips, not two per
		// piece.  Indices past the element's pieces are clamped to its last piece
		// for the loads and skipped for the stores.
		constexpr int kPieces = (32*32*8 + kT - 1)/kT;  // per thread, nen <= 32
		const long long pbase = args.pair_offsets[e];
		const int total = nen*nen*8;
		// the element's residual rows (thread a < nen: row a) ride along
		const bool row = tid < nen;
		double* nr = args.negres + 4*(size_t)lconn[row ? tid : 0];
		double nrv[4];
#pragma unroll
		for (int c = 0; c < 4; ++c) nrv[c] = nr[c];
		int slot[kPieces];
#pragma unroll
		for (int k = 0; k < kPieces; ++k) slot[k] = args.pair_block[pbase + (min(tid + k*kT, total - 1) >> 3)];
		double2 v[kPieces];
#pragma unroll
		for (int k = 0; k < kPieces; ++k)
			v[k] = reinterpret_cast<const double2*>(args.jac + (size_t)slot[k]*16)[min(tid + k*kT, total - 1) & 7];
#pragma unroll
		for (int k = 0; k < kPieces; ++k) {
			const int it = tid + k*kT;
			if (it < total) {
				const double* src = stage + (it >> 3)*kStageStride + 2*(it & 7);
				v[k].x += src[0];
				v[k].y += src[1];
				reinterpret_cast<double2*>(args.jac + (size_t)slot[k]*16)[it & 7] = v[k];
			}
		}
		if (row)
			for (int c = 0; c < 4; ++c) nr[c] = nrv[c] + resid[4*tid + c];
	}
	if (pair_active) {
		const long long pbase = args.pair_offsets[e];
#pragma unroll
		for (int i = 0; i < 2; ++i)
#pragma unroll
			for (int j = 0; j < 2; ++j) {
				const int a = a0 + 2*ta + i, b = 2*tb + j;
				if (a >= a1 || b >= nen) continue;
				const int pr = 2*i + j;
				double* J = args.jac + (size_t)args.pair_block[pbase + (long long)a*nen + b]*16;
				for (int x = 0; x < 3; ++x) {
					for (int y = 0; y < 3; ++y) J[4*x + y] += uu[pr][3*x + y] + (x == y ? dg[pr] : 0.0);
					J[4*x + 3] += i3[pr][x];
					J[12 + x] += t3[pr][x];
				}
				J[15] += k33[pr];
			}
	}
	if (!(kMfma && kJac))
		for (int a = a0 + tid; a < a1; a += kT) {
			const size_t node = lconn[a];
			for (int c = 0; c < 4; ++c) args.negres[4*node + c] += resid[4*(a - a0) + c];
		}
}

__global__ void DirichletMatrixKernel(int nodes, const int* __restrict__ ro, const int* __restrict__ diag,
	const int* __restrict__ vmask, const int* __restrict__ pmask, double* __restrict__ jac)
{
	const int node = blockIdx.x*blockDim.x + threadIdx.x;
	if (node >= nodes) return;
	const bool v = vmask[node] != 0, p = pmask[node] != 0;
	if (!v && !p) return;
	for (int k = ro[node]; k < ro[node + 1]; ++k) {
		double* J = jac + (size_t)k*16;
		for (int f = 0; f < 4; ++f) {
			if (!(f < 3 ? v : p)) continue;
			for (int c = 0; c < 4; ++c) J[4*f + c] = 0.0;
		}
	}
	double* D = jac + (size_t)diag[node]*16;
	for (int f = 0; f < 4; ++f)
		if (f < 3 ? v : p) D[5*f] = 1.0;
}

__global__ void DirichletRhsKernel(int nodes, const int* __restrict__ vmask, const int* __restrict__ pmask,
	const double* __restrict__ vt, const double* __restrict__ pt, const double* __restrict__ state,
	double* __restrict__ rhs)
{
	const int node = blockIdx.x*blockDim.x + threadIdx.x;
	if (node >= nodes) return;
	if (vmask[node])
		for (int c = 0; c < 3; ++c) rhs[4*(size_t)node + c] = vt[3*(size_t)node + c] - state[4*(size_t)node + c];
	if (pmask[node]) rhs[4*(size_t)node + 3] = pt[node] - state[4*(size_t)node + 3];
}

__global__ void DirichletStateKernel(int nodes, const int* __restrict__ vmask, const int* __restrict__ pmask,
	const double* __restrict__ vt, const double* __restrict__ pt, double* __restrict__ state)
{
	const int node = blockIdx.x*blockDim.x + threadIdx.x;
	if (node >= nodes) return;
	if (vmask[node])
		for (int c = 0; c < 3; ++c) state[4*(size_t)node + c] = vt[3*(size_t)node + c];
	if (pmask[node]) state[4*(size_t)node + 3] = pt[node];
}

// Volumetric flow through labelled faces (IntegrateBoundaryFlow,
// solvers/cpu/include/BoundaryFlow.hpp:164-200).  The flow is linear in the
// state, so it is taken from the face's area-weighted basis vectors
// int_face N_a n dA (integrated once, with the same 4x4 Gauss rule, by
// FaceTractionKernel at unit pressure): one 64-lane block per face,
// fixed-order wave sum.
__global__ void FaceFlowVecKernel(int count, const int* __restrict__ felem, const int* __restrict__ eoff,
	const int* __restrict__ conn, const double* __restrict__ vec, const double* __restrict__ state,
	double* __restrict__ flows)
{
	const int f = blockIdx.x;
	if (f >= count) return;
	const int e = felem[f], b0 = eoff[e], nen = eoff[e + 1] - b0;
	double s = 0.0;
	for (int a = threadIdx.x; a < nen; a += 64) {
		const double* v = vec + ((size_t)f*kMaxBasis + a)*3;
		const size_t node = conn[b0 + a];
		s += state[4*node]*v[0] + state[4*node + 1]*v[1] + state[4*node + 2]*v[2];
	}
	for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
	if (threadIdx.x == 0) flows[f] = -s;
}

// Pressure-traction load of one labelled face (IntegrateBoundaryPressureTraction,
// BoundaryFlow.hpp:445-477): local[a][c] = -p sum_pt s w_i w_j / 4 * cof_c(pt) N_a(pt),
// with cof = det J * (row `fixed axis` of J^-1) the face's area vector.  One
// 64-lane block per face; the per-face vectors are summed on the host in face
// order, so the load is deterministic and ordered like the reference's.
__global__ void FaceTractionKernel(int count, const int* __restrict__ felem, const int* __restrict__ fface,
	const double* __restrict__ pressure, const int* __restrict__ eoff, const int* __restrict__ xoff,
	const uint8_t* __restrict__ xcol, const double* __restrict__ xval, const double* __restrict__ bezier,
	double* __restrict__ local)
{
	const int f = blockIdx.x;
	if (f >= count) return;
	__shared__ double bern[16][64];
	__shared__ double wcof[16][3];
	const int e = felem[f], face = fface[f];
	const int fixed_axis[6] = {2, 1, 0, 1, 0, 2};
	const int var0[6] = {0, 0, 1, 0, 1, 0}, var1[6] = {1, 2, 2, 2, 2, 1};
	const double fixed_value[6] = {0.0, 0.0, 1.0, 1.0, 0.0, 1.0};
	const double sign[6] = {-1.0, -1.0, 1.0, 1.0, -1.0, 1.0};
	const double gw[4] = {0.3478548451374539, 0.6521451548625461, 0.6521451548625461, 0.3478548451374539};
	const int lane = threadIdx.x;  // 64 threads
	for (int idx = lane; idx < 16*64; idx += 64) {
		const int pt = idx >> 6, p = idx & 63;
		const int qi = pt >> 2, qj = pt & 3;
		double c[3];
		c[fixed_axis[face]] = fixed_value[face];
		c[var0[face]] = kGp[qi];
		c[var1[face]] = kGp[qj];
		double b[3][4], db[3][4], d2b[3][4];
		for (int d = 0; d < 3; ++d) Bern3(c[d], b[d], db[d], d2b[d]);
		bern[pt][p] = b[0][p & 3]*b[1][(p >> 2) & 3]*b[2][p >> 4];
	}
	if (lane < 16) {
		const int pt = lane, qi = pt >> 2, qj = pt & 3;
		double c[3];
		c[fixed_axis[face]] = fixed_value[face];
		c[var0[face]] = kGp[qi];
		c[var1[face]] = kGp[qj];
		double b[3][4], db[3][4], d2b[3][4];
		for (int d = 0; d < 3; ++d) Bern3(c[d], b[d], db[d], d2b[d]);
		const double* X = bezier + (size_t)e*192;
		double J[3][3] = {};
		for (int k = 0; k < 4; ++k)
			for (int j = 0; j < 4; ++j)
				for (int i = 0; i < 4; ++i) {
					const int p = i + 4*j + 16*k;
					const double g[3] = {db[0][i]*b[1][j]*b[2][k], b[0][i]*db[1][j]*b[2][k], b[0][i]*b[1][j]*db[2][k]};
					for (int x = 0; x < 3; ++x)
						for (int d = 0; d < 3; ++d) J[x][d] += X[3*p + x]*g[d];
				}
		const int a = fixed_axis[face];
		const int a1 = (a + 1)%3, a2 = (a + 2)%3;
		const double w = sign[face]*gw[qi]*gw[qj]*0.25;
		for (int x = 0; x < 3; ++x) {
			const int x1 = (x + 1)%3, x2 = (x + 2)%3;
			wcof[pt][x] = w*(J[x1][a1]*J[x2][a2] - J[x1][a2]*J[x2][a1]);
		}
	}
	__syncthreads();
	const int b0 = eoff[e], nen = eoff[e + 1] - b0;
	const double pf = pressure[f];
	for (int aa = lane; aa < nen; aa += 64) {
		double acc[3] = {0.0, 0.0, 0.0};
		for (int pt = 0; pt < 16; ++pt) {
			double N = 0.0;
			for (int k = xoff[b0 + aa]; k < xoff[b0 + aa + 1]; ++k) N += xval[k]*bern[pt][xcol[k]];
			for (int c = 0; c < 3; ++c) acc[c] += N*wcof[pt][c];
		}
		double* out = local + ((size_t)f*kMaxBasis + aa)*3;
		for (int c = 0; c < 3; ++c) out[c] = -pf*acc[c];
	}
}

// Face mass of one labelled face for the scalar flux / Robin terms
// (AssembleGenericSurfaceKernel, solvers/cuda/include/GenericTransportKernels.cuh:146-215):
// with m = w_i w_j / 4 |t_0 x t_1| at the 4x4 Gauss points of the face,
// mass[f][a][b] = sum_pt N_a N_b m and load[f][a] = sum_pt N_a m.  |t_0 x t_1|
// is the length of the cofactor vector FaceTractionKernel integrates.  One
// 64-lane block per face; `stride` (>= the element's basis count) in both outputs.
__global__ void FaceMassKernel(int count, int stride, const int* __restrict__ felem, const int* __restrict__ fface,
	const int* __restrict__ eoff, const int* __restrict__ xoff, const uint8_t* __restrict__ xcol,
	const double* __restrict__ xval, const double* __restrict__ bezier, double* __restrict__ mass,
	double* __restrict__ load)
{
	const int f = blockIdx.x;
	if (f >= count) return;
	__shared__ double bern[16][64];
	__shared__ double measure[16];
	__shared__ double basis[16][kMaxBasis];
	const int e = felem[f], face = fface[f];
	const int fixed_axis[6] = {2, 1, 0, 1, 0, 2};
	const int var0[6] = {0, 0, 1, 0, 1, 0}, var1[6] = {1, 2, 2, 2, 2, 1};
	const double fixed_value[6] = {0.0, 0.0, 1.0, 1.0, 0.0, 1.0};
	const double gw[4] = {0.3478548451374539, 0.6521451548625461, 0.6521451548625461, 0.3478548451374539};
	const int lane = threadIdx.x;  // 64 threads
	for (int idx = lane; idx < 16*64; idx += 64) {
		const int pt = idx >> 6, p = idx & 63;
		double c[3];
		c[fixed_axis[face]] = fixed_value[face];
		c[var0[face]] = kGp[pt >> 2];
		c[var1[face]] = kGp[pt & 3];
		double b[3][4], db[3][4], d2b[3][4];
		for (int d = 0; d < 3; ++d) Bern3(c[d], b[d], db[d], d2b[d]);
		bern[pt][p] = b[0][p & 3]*b[1][(p >> 2) & 3]*b[2][p >> 4];
	}
	if (lane < 16) {
		const int pt = lane, qi = pt >> 2, qj = pt & 3;
		double c[3];
		c[fixed_axis[face]] = fixed_value[face];
		c[var0[face]] = kGp[qi];
		c[var1[face]] = kGp[qj];
		double b[3][4], db[3][4], d2b[3][4];
		for (int d = 0; d < 3; ++d) Bern3(c[d], b[d], db[d], d2b[d]);
		const double* X = bezier + (size_t)e*192;
		double t0[3] = {}, t1[3] = {};
		for (int k = 0; k < 4; ++k)
			for (int j = 0; j < 4; ++j)
				for (int i = 0; i < 4; ++i) {
					const int p = i + 4*j + 16*k;
					const double g[3] = {db[0][i]*b[1][j]*b[2][k], b[0][i]*db[1][j]*b[2][k], b[0][i]*b[1][j]*db[2][k]};
					for (int x = 0; x < 3; ++x) {
						t0[x] += X[3*p + x]*g[var0[face]];
						t1[x] += X[3*p + x]*g[var1[face]];
					}
				}
		const double n[3] = {t0[1]*t1[2] - t0[2]*t1[1], t0[2]*t1[0] - t0[0]*t1[2], t0[0]*t1[1] - t0[1]*t1[0]};
		measure[pt] = gw[qi]*gw[qj]*0.25*sqrt(n[0]*n[0] + n[1]*n[1] + n[2]*n[2]);
	}
	__syncthreads();
	const int b0 = eoff[e], nen = eoff[e + 1] - b0;
	for (int idx = lane; idx < 16*nen; idx += 64) {
		const int pt = idx/nen, a = idx - pt*nen;
		double N = 0.0;
		for (int k = xoff[b0 + a]; k < xoff[b0 + a + 1]; ++k) N += xval[k]*bern[pt][xcol[k]];
		basis[pt][a] = N;
	}
	__syncthreads();
	double* mf = mass + (size_t)f*stride*stride;
	for (int pair = lane; pair < nen*nen; pair += 64) {
		const int a = pair/nen, b = pair - a*nen;
		double s = 0.0;
		for (int pt = 0; pt < 16; ++pt) s += basis[pt][a]*basis[pt][b]*measure[pt];
		mf[a*stride + b] = s;
	}
	for (int a = lane; a < nen; a += 64) {
		double s = 0.0;
		for (int pt = 0; pt < 16; ++pt) s += basis[pt][a]*measure[pt];
		load[(size_t)f*stride + a] = s;
	}
}

// Residual-only assembly from the basis table, for elements of at most 32
// basis functions: one workgroup of 256 threads per element and no slabs.
// Phase A: four lanes per point sum the fields over the basis functions and
// the first of them stores the point record; phase B: eight lanes per basis
// function sum its residual rows over the 64 points in a fixed order, so the
// result is deterministic.  Without the LDS basis staging and tangent
// registers of AssembleNsKernel, several elements share a compute unit.
constexpr int kQr = 35;  // point record r[0..33] in LDS, odd stride

__global__ void __launch_bounds__(kThreads, 2) NsResidualKernel(AsmArgs args)
{
	__shared__ double nodal[32*8];
	__shared__ double qd[kQ*kQr];
	__shared__ int lconn[32];
	const int tid = threadIdx.x;
	const int item = blockIdx.x;
	if (item >= args.count) return;
	const int e = args.elements[item];
	const int b0 = args.eoff[e];
	const int nen = args.eoff[e + 1] - b0;
	const double rho = args.rho, mu = args.mu, dt = args.dt;
	const double nu = mu/rho;
	for (int a = tid; a < nen; a += kThreads) {
		const int node = args.conn[b0 + a];
		lconn[a] = node;
		for (int f = 0; f < 4; ++f) nodal[8*a + f] = args.state[4*(size_t)node + f];
		for (int f = 0; f < 3; ++f) nodal[8*a + 4 + f] = args.previous[4*(size_t)node + f];
	}
	__syncthreads();
	const double* tab = args.table + (size_t)b0*kQ*kTableRec;
	{
		const int q = tid >> 2, part = tid & 3;
		double s[22] = {};
		for (int a = part; a < nen; a += 4) {
			const double* bb = tab + ((size_t)q*nen + a)*kTableRec;
			const double N = bb[0], gx = bb[1], gy = bb[2], gz = bb[3], L = bb[4];
			const double* un = nodal + 8*a;
#pragma unroll
			for (int f = 0; f < 4; ++f) {
				s[f] += un[f]*N;
				s[4 + 3*f] += un[f]*gx;
				s[5 + 3*f] += un[f]*gy;
				s[6 + 3*f] += un[f]*gz;
			}
#pragma unroll
			for (int c = 0; c < 3; ++c) {
				s[16 + c] += un[c]*L;
				s[19 + c] += un[4 + c]*N;
			}
		}
#pragma unroll
		for (int i = 0; i < 22; ++i) {
			s[i] += __shfl_xor(s[i], 1, 64);
			s[i] += __shfl_xor(s[i], 2, 64);
		}
		if (part == 0)
			PointRecord<kQ>(s, args.geometry + (size_t)e*kGeomFields*kQ + q, rho, mu, nu, dt, false, qd + q*kQr);
	}
	__syncthreads();
	if (tid < nen*8) {
		const int a = tid >> 3, sub = tid & 7;
		double R[4] = {0.0, 0.0, 0.0, 0.0};
		for (int q = sub; q < kQ; q += 8) {
			const double* bb = tab + ((size_t)q*nen + a)*kTableRec;
			const double* r = qd + q*kQr;
			const double N = bb[0], gx = bb[1], gy = bb[2], gz = bb[3], m = r[0];
#pragma unroll
			for (int c = 0; c < 3; ++c) R[c] -= (N*r[6 + c] + gx*r[9 + 3*c] + gy*r[10 + 3*c] + gz*r[11 + 3*c])*m;
			R[3] -= (N*r[18] - gx*r[19] - gy*r[20] - gz*r[21])*m;
		}
#pragma unroll
		for (int c = 0; c < 4; ++c) {
			R[c] += __shfl_xor(R[c], 1, 64);
			R[c] += __shfl_xor(R[c], 2, 64);
			R[c] += __shfl_xor(R[c], 4, 64);
		}
		if (sub == 0) {
			const size_t node = lconn[a];
			for (int c = 0; c < 4; ++c) args.negres[4*node + c] += R[c];
		}
	}
}

// Basis records of every (point, function) of one element per workgroup, in
// the table layout [q][a][kTableRec] the assembly reads slab by slab.
__global__ void __launch_bounds__(256) BasisTableKernel(const int* __restrict__ eoff, const int* __restrict__ xoff,
	const uint8_t* __restrict__ xcol, const double* __restrict__ xval, const double* __restrict__ geometry,
	double* __restrict__ table)
{
	extern __shared__ double C[];  // [nen][64] extraction rows
	const int e = blockIdx.x;
	const int b0 = eoff[e], nen = eoff[e + 1] - b0;
	for (int i = threadIdx.x; i < nen*64; i += blockDim.x) C[i] = 0.0;
	__syncthreads();
	for (int a = threadIdx.x; a < nen; a += blockDim.x)
		for (int k = xoff[b0 + a]; k < xoff[b0 + a + 1]; ++k) C[64*a + xcol[k]] = xval[k];
	__syncthreads();
	const double* geo = geometry + (size_t)e*kGeomFields*kQ;
	double* dst = table + (size_t)b0*kQ*kTableRec;
	for (int it = threadIdx.x; it < kQ*nen; it += blockDim.x) {
		const int q = it/nen, a = it - q*nen;
		double rec[kTableRec];
		EvalBasis(C + 64*a, geo + q, q, rec);
#pragma unroll
		for (int f = 0; f < kTableRec; ++f) dst[(size_t)it*kTableRec + f] = rec[f];
	}
}

}  // namespace

void SetBasisTable(tfi_mesh* m, int enable)
{
	tfi_ctx* ctx = m->ctx;
	const size_t count = (size_t)m->basis_entries*kQ*kTableRec;
	if (enable < 0) {
		size_t free_bytes = 0, total_bytes = 0;
		Check(hipMemGetInfo(&free_bytes, &total_bytes), "hipMemGetInfo");
		enable = count*sizeof(double) <= free_bytes/2 ? 1 : 0;
	}
	if (!enable) {
		Sync(ctx);
		m->basis_table.Release();
		return;
	}
	if (m->basis_table.size() == count) return;
	m->basis_table.Allocate(ctx, count);
	hipLaunchKernelGGL(BasisTableKernel, dim3(m->elements), dim3(256), sizeof(double)*64*m->max_nen, ctx->stream,
		m->eoff.data(), m->xoff.data(), m->xcol.data(), m->xval.data(), m->geometry.data(), m->basis_table.data());
	TFI_LAUNCH_CHECK("BasisTableKernel");
	Sync(ctx);
}

namespace {

// LDS bytes of one AssembleNsKernel workgroup (nodal records, extraction rows
// when there is no basis table, slab basis records, point records, residual
// rows, connectivity).
size_t AssemblyLds(int nen_max, bool table, int slab, int qstride, bool mfma)
{
	// the MFMA scatter stages nen^2 blocks of 17 doubles over the slab buffers
	const size_t slab_part = std::max((size_t)slab*qstride + (size_t)slab*kQd,
		mfma ? (size_t)nen_max*nen_max*17 : (size_t)0);
	return sizeof(double)*(kMaxBasis*8 + 4*kMaxBasis + kMaxBasis/2 + (mfma ? (size_t)kPointGeom*slab : 0)
		+ (table ? 0 : (size_t)nen_max*64) + slab_part);
}

template <bool kExact, bool kMfma, bool kJac, int kSlab>
void LaunchAssembly(tfi_mesh* m, AsmArgs a, int tiles, size_t lds)
{
	const void* kernel = reinterpret_cast<const void*>(AssembleNsKernel<kExact, kMfma, kJac, kSlab>);
	static size_t raised = 0;  // per instantiation
	if (lds > 65536 && lds > raised) {
		Check(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)),
			"hipFuncSetAttribute");
		raised = lds;
	}
	tfi_ctx* ctx = m->ctx;
	for (int c = 0; c < m->colors; ++c) {
		a.elements = m->color_elements.data() + m->color_offsets[c];
		a.count = m->color_offsets[c + 1] - m->color_offsets[c];
		if (a.count == 0) continue;
		hipLaunchKernelGGL((AssembleNsKernel<kExact, kMfma, kJac, kSlab>), dim3(a.count*tiles),
			dim3(kMfma ? 512 : kThreads), lds, ctx->stream, a);
		TFI_LAUNCH_CHECK("AssembleNsKernel");
	}
}

template <bool kExact, bool kJac>
void LaunchVectorAssembly(tfi_mesh* m, const AsmArgs& a, int tiles, size_t lds, int slab)
{
	switch (slab) {
	case 32: LaunchAssembly<kExact, false, kJac, 32>(m, a, tiles, lds); break;
	case 16: LaunchAssembly<kExact, false, kJac, 16>(m, a, tiles, lds); break;
	default: LaunchAssembly<kExact, false, kJac, 8>(m, a, tiles, lds); break;
	}
}

}  // namespace

void AssembleNs(tfi_mesh* m, const double* state, const double* previous, const NsParams& p, double* jacobian,
	double* negative_residual)
{
	const bool exact = p.tangent == TFI_TANGENT_EXACT;
	const bool jac = jacobian != nullptr;
	const int nen_max = m->max_nen;
	if (nen_max > kMaxBasis) throw Error("element assembly supports at most 80 basis functions per element");
	const bool table = m->basis_table.size() > 0;
	// elements of at most 32 basis functions take the FP64-MFMA tangent, which
	// reads the basis table (evaluating the basis in the same kernel would
	// push it past its 256 registers)
	const bool mfma = jac && nen_max <= 32 && table;
	tfi_ctx* ctx = m->ctx;
	if (jacobian) Fill(ctx, jacobian, 0.0, m->nnzb*16);
	Fill(ctx, negative_residual, 0.0, 4LL*m->nodes);
	const int tiles_b = (nen_max + 1)/2;
	int a_tile = mfma ? nen_max : std::min(nen_max, 2*(kThreads/tiles_b));
	if (a_tile < 2) a_tile = 2;
	const int tiles = (nen_max + a_tile - 1)/a_tile;
	// per-point stride of the basis records, odd so that lanes reading the
	// same field of different points hit different LDS banks
	const int records = mfma ? 32 : nen_max;
	const int qstride = records*kBasisRec | 1;
	// slab of quadrature points: the largest of 32 / 16 / 8 whose LDS fits
	int slab = 32;
	if (!mfma)
		while (slab > 8 && AssemblyLds(nen_max, table, slab, qstride, false) > kLdsLimit) slab >>= 1;
	const size_t lds = AssemblyLds(nen_max, table, slab, qstride, mfma);
	if (lds > kLdsLimit)
		throw Error("element assembly needs " + std::to_string(lds) + " bytes of LDS for " + std::to_string(nen_max)
			+ " basis functions per element, above the " + std::to_string(kLdsLimit) + " bytes of a workgroup");
	AsmArgs a{};
	a.eoff = m->eoff.data();
	a.conn = m->conn.data();
	a.xoff = m->xoff.data();
	a.xcol = m->xcol.data();
	a.xval = m->xval.data();
	a.geometry = m->geometry.data();
	a.pair_offsets = m->pair_offsets.data();
	a.pair_block = m->pair_block.data();
	a.a_tile = a_tile;
	a.tiles = tiles;
	a.state = state;
	a.previous = previous;
	a.rho = p.density;
	a.mu = p.viscosity;
	a.dt = p.dt;
	a.jac = jacobian;
	a.negres = negative_residual;
	a.table = table ? m->basis_table.data() : nullptr;
	a.qstride = qstride;
	a.max_nen = nen_max;
	StreamTimer timer(ctx, jac ? &ctx->timers.assembly_ms : &ctx->timers.residual_ms);
	if (!jac && table && nen_max <= 32) {
		for (int c = 0; c < m->colors; ++c) {
			a.elements = m->color_elements.data() + m->color_offsets[c];
			a.count = m->color_offsets[c + 1] - m->color_offsets[c];
			if (a.count == 0) continue;
			hipLaunchKernelGGL(NsResidualKernel, dim3(a.count), dim3(kThreads), 0, ctx->stream, a);
			TFI_LAUNCH_CHECK("NsResidualKernel");
		}
	} else if (mfma) {
		if (exact) LaunchAssembly<true, true, true, 32>(m, a, tiles, lds);
		else LaunchAssembly<false, true, true, 32>(m, a, tiles, lds);
	} else if (!jac) {
		LaunchVectorAssembly<false, false>(m, a, tiles, lds, slab);
	} else if (exact) {
		LaunchVectorAssembly<true, true>(m, a, tiles, lds, slab);
	} else {
		LaunchVectorAssembly<false, true>(m, a, tiles, lds, slab);
	}
	if (jac) {
		ctx->timers.assembly_calls += 1;
		ctx->timers.assembly_launches += m->colors;
		ctx->timers.assembly_flops += AssemblyFlops(m, true, exact);
	} else {
		ctx->timers.residual_calls += 1;
		ctx->timers.residual_flops += AssemblyFlops(m, false, exact);
	}
}

// Counted arithmetic of AssembleNsKernel per element with nen basis functions
// and Q = 64 points: sum-factorised basis and push-forward (656 + 42 flops per
// (a, q), none when the basis table supplies them), field interpolation (44),
// u.grad N and residual row (38), per-point scalars (~300 per q) and the
// tangent (72 flops per (a, b, q) for the legacy tangent; 132 per (a, b, q)
// plus ~100 per (a, q) for the exact one).
double AssemblyFlops(const tfi_mesh* m, bool with_jacobian, bool exact)
{
	const double basis = m->basis_table.size() ? 0.0 : 656.0 + 42.0;
	double total = 0.0;
	for (int e = 0; e < m->elements; ++e) {
		const double nen = m->h_eoff[e + 1] - m->h_eoff[e];
		double pair = 0.0;
		if (with_jacobian) pair = exact ? 132.0*nen*nen + 100.0*nen : 72.0*nen*nen;
		total += kQ*(nen*(basis + 44.0 + 38.0) + 300.0 + pair);
	}
	return total;
}

void ApplyDirichlet(tfi_mesh* m, const int* vmask, const int* pmask, const double* vtarget, const double* ptarget,
	const double* state, double* jacobian, double* rhs)
{
	tfi_ctx* ctx = m->ctx;
	const int blocks = (m->nodes + 255)/256;
	if (jacobian) {
		hipLaunchKernelGGL(DirichletMatrixKernel, dim3(blocks), dim3(256), 0, ctx->stream, m->nodes,
			m->row_offsets.data(), m->diagonal.data(), vmask, pmask, jacobian);
		TFI_LAUNCH_CHECK("DirichletMatrixKernel");
	}
	if (rhs) {
		hipLaunchKernelGGL(DirichletRhsKernel, dim3(blocks), dim3(256), 0, ctx->stream, m->nodes, vmask, pmask,
			vtarget, ptarget, state, rhs);
		TFI_LAUNCH_CHECK("DirichletRhsKernel");
	}
}

void SetDirichletState(tfi_mesh* m, const int* vmask, const int* pmask, const double* vtarget,
	const double* ptarget, double* state)
{
	hipLaunchKernelGGL(DirichletStateKernel, dim3((m->nodes + 255)/256), dim3(256), 0, m->ctx->stream, m->nodes, vmask,
		pmask, vtarget, ptarget, state);
	TFI_LAUNCH_CHECK("DirichletStateKernel");
}

void FaceFlowsFromVectors(tfi_mesh* m, int count, const int* element, const double* vec, const double* state,
	double* flows)
{
	if (count <= 0) return;
	hipLaunchKernelGGL(FaceFlowVecKernel, dim3(count), dim3(64), 0, m->ctx->stream, count, element, m->eoff.data(),
		m->conn.data(), vec, state, flows);
	TFI_LAUNCH_CHECK("FaceFlowVecKernel");
}

void FaceMass(tfi_mesh* m, int count, const int* element, const int* face, double* mass, double* load)
{
	if (count <= 0) return;
	if (m->max_nen > kMaxBasis) throw Error("face integrals support at most 80 basis functions");
	hipLaunchKernelGGL(FaceMassKernel, dim3(count), dim3(64), 0, m->ctx->stream, count, m->max_nen, element, face,
		m->eoff.data(), m->xoff.data(), m->xcol.data(), m->xval.data(), m->bezier.data(), mass, load);
	TFI_LAUNCH_CHECK("FaceMassKernel");
}

void FaceTraction(tfi_mesh* m, int count, const int* element, const int* face, const double* pressure,
	double* local)
{
	if (count <= 0) return;
	hipLaunchKernelGGL(FaceTractionKernel, dim3(count), dim3(64), 0, m->ctx->stream, count, element, face, pressure,
		m->eoff.data(), m->xoff.data(), m->xcol.data(), m->xval.data(), m->bezier.data(), local);
	TFI_LAUNCH_CHECK("FaceTractionKernel");
}

}  // namespace tfi
