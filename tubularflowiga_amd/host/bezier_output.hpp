// Bezier visualisation of the flow solution: the element Bezier points merged
// into one point set, its geometry checks and report, and the temporal VTKHDF
// file that iga_cuda writes by default for transient runs
// (solvers/cuda/src/iga_cuda.cu:52-70, 366-386; include/VtkOutput.hpp:27-32).
//
// Point identity, tolerances, checks, report keys, the VTKHDF group and
// dataset layout and the restart behaviour follow include/BezierVisualization.hpp
// (BuildBezierVisualizationMesh :491-738, ExtractBezierPointArrays :766-795)
// and include/TemporalVtkHdf.hpp (TemporalVtkHdfWriter :735-998), so that a
// file written here reads like the reference's:
//   * a Bezier point of an element is identified by its extraction column:
//     the (node, coefficient) pairs with non-zero coefficient, coefficients
//     quantised to 1e-12; equal keys are one global point (coordinates must
//     agree within 2e-6 max(1, |x|max)); a new point that lies on an earlier
//     one without sharing its key counts as coincident-unmerged;
//   * per element: repeated global points (collapsed), the Jacobian of the
//     merged Bezier map at the 4x4x4 Gauss points (non-positive), and, for
//     element pairs whose boxes overlap and that share fewer than 16 points,
//     27 sampled points of each tested for strict containment in the other;
//   * the solution's point arrays are the extraction-weighted sums of the
//     control values (the value of the spline at the Bezier control point);
//   * VTK_BEZIER_HEXAHEDRON (79) cells of degree 3 in VTK's higher-order
//     hexahedron point order.
#pragma once

#include "case_input.hpp"

#include <hdf5.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <limits>
#include <numeric>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace tfi_host {

namespace bezier {

constexpr std::uint8_t kBezierHexCell = 79;  // VTK_BEZIER_HEXAHEDRON
constexpr std::uint64_t kNone = std::numeric_limits<std::uint64_t>::max();

struct GeometryChecks {
	std::uint64_t elements = 0, local_points = 0, unique_points = 0, shared_point_references = 0;
	std::uint64_t signature_coordinate_repairs = 0, coincident_unmerged_points = 0, collapsed_element_points = 0;
	std::uint64_t nonpositive_jacobian_elements = 0, overlap_candidates = 0, overlapping_element_pairs = 0;
	std::uint64_t first_bad_element = kNone;
	std::array<std::uint64_t, 2> first_overlap_pair{{kNone, kNone}};
	std::array<std::uint64_t, 2> first_coincident_element_pair{{kNone, kNone}};
	double minimum_jacobian = std::numeric_limits<double>::infinity();
	double maximum_signature_coordinate_difference = 0.0;

	bool valid() const
	{
		return collapsed_element_points == 0 && nonpositive_jacobian_elements == 0 && coincident_unmerged_points == 0
			&& overlapping_element_pairs == 0;
	}
};

struct VisualMesh {
	std::vector<std::array<double, 3>> points;
	std::vector<std::int32_t> point_labels;
	std::vector<std::int64_t> connectivity, offsets{0};
	std::vector<std::uint8_t> types;
	std::vector<std::uint64_t> element_ids;
	std::vector<std::int32_t> element_owners, element_labels;
	std::vector<std::array<std::int32_t, 6>> face_labels;
	// extraction column of every global point: (node, coefficient), nodes ascending
	std::vector<std::uint64_t> column_offsets{0};
	std::vector<std::int32_t> column_nodes;
	std::vector<double> column_values;
	GeometryChecks checks;
};

// Position of tensor point (i, j, k) of a degree-3 hexahedron in VTK's
// higher-order point order: 8 vertices, the 2 interior points of each of the
// 12 edges (i-edges at j, k = 0/3, j-edges at i = 3 and i = 0, then the four
// k-edges in vertex order), the 4 interior points of the -i, +i, -j, +j, -k,
// +k faces, then the 8 interior points, each group in lexicographic order of
// its free indices.
inline int VtkCubicHexPosition(int i, int j, int k)
{
	constexpr int p = 3, m = p - 1;
	const bool bi = i == 0 || i == p, bj = j == 0 || j == p, bk = k == 0 || k == p;
	const int on_boundary = bi + bj + bk;
	const int corner = i ? (j ? 2 : 1) : (j ? 3 : 0);
	if (on_boundary == 3) return corner + (k ? 4 : 0);
	int offset = 8;
	if (on_boundary == 2) {
		if (!bi) return offset + (i - 1) + (j ? 2*m : 0) + (k ? 4*m : 0);
		if (!bj) return offset + (j - 1) + (i ? m : 3*m) + (k ? 4*m : 0);
		return offset + 8*m + m*corner + (k - 1);
	}
	offset += 12*m;
	if (on_boundary == 1) {
		if (bi) return offset + (i ? m*m : 0) + (j - 1) + m*(k - 1);
		offset += 2*m*m;
		if (bj) return offset + (j ? m*m : 0) + (i - 1) + m*(k - 1);
		offset += 2*m*m;
		return offset + (k ? m*m : 0) + (i - 1) + m*(j - 1);
	}
	offset += 6*m*m;
	return offset + (i - 1) + m*((j - 1) + m*(k - 1));
}

inline std::array<int, 64> VtkCubicHexOrder()
{
	std::array<int, 64> order{};
	for (int k = 0; k < 4; ++k)
		for (int j = 0; j < 4; ++j)
			for (int i = 0; i < 4; ++i) order[VtkCubicHexPosition(i, j, k)] = i + 4*j + 16*k;
	return order;
}

inline bool PointOnFace(int point, int face)
{
	const int i = point % 4, j = (point/4) % 4, k = point/16;
	switch (face) {
	case 0: return k == 0;
	case 1: return j == 0;
	case 2: return i == 3;
	case 3: return j == 3;
	case 4: return i == 0;
	default: return k == 3;
	}
}

// A point on several labelled faces: the wall (0) wins at a cap rim, two
// distinct non-wall labels mark a mixed junction (-2).
inline std::int32_t MergeLabel(std::int32_t current, std::int32_t candidate)
{
	if (candidate < 0) return current;
	if (current < 0 || current == candidate) return candidate;
	return current == 0 || candidate == 0 ? 0 : -2;
}

inline std::int32_t CellLabel(const std::array<std::int32_t, 6>& labels)
{
	std::int32_t result = -1;
	for (const auto l : labels) {
		if (l < 0) continue;
		if (result < 0) result = l;
		else if (result != l) return -2;
	}
	return result;
}

struct Box {
	std::array<double, 3> lo{{std::numeric_limits<double>::infinity(), std::numeric_limits<double>::infinity(),
		std::numeric_limits<double>::infinity()}};
	std::array<double, 3> hi{{-std::numeric_limits<double>::infinity(), -std::numeric_limits<double>::infinity(),
		-std::numeric_limits<double>::infinity()}};
	void Add(const std::array<double, 3>& x)
	{
		for (int d = 0; d < 3; ++d) {
			lo[d] = std::min(lo[d], x[d]);
			hi[d] = std::max(hi[d], x[d]);
		}
	}
};

inline bool BoxesOverlap(const Box& a, const Box& b, double tol)
{
	for (int d = 0; d < 3; ++d)
		if (std::min(a.hi[d], b.hi[d]) - std::max(a.lo[d], b.lo[d]) <= tol) return false;
	return true;
}

using Cell = std::array<std::array<double, 3>, 64>;

inline void Bernstein3(double t, double b[4], double db[4])
{
	const double s = 1.0 - t;
	b[0] = s*s*s; b[1] = 3.0*s*s*t; b[2] = 3.0*s*t*t; b[3] = t*t*t;
	db[0] = -3.0*s*s; db[1] = 3.0 - 12.0*t + 9.0*t*t; db[2] = 3.0*(2.0 - 3.0*t)*t; db[3] = 3.0*t*t;
}

// x(u) of the cell's Bezier map and, optionally, dx/du (row: physical, column: parameter)
inline std::array<double, 3> Map(const Cell& cell, const std::array<double, 3>& u, double (*jac)[3] = nullptr)
{
	double b[3][4], db[3][4];
	for (int d = 0; d < 3; ++d) Bernstein3(u[d], b[d], db[d]);
	std::array<double, 3> x{{0.0, 0.0, 0.0}};
	if (jac)
		for (int r = 0; r < 3; ++r)
			for (int c = 0; c < 3; ++c) jac[r][c] = 0.0;
	for (int k = 0; k < 4; ++k)
		for (int j = 0; j < 4; ++j)
			for (int i = 0; i < 4; ++i) {
				const auto& p = cell[i + 4*j + 16*k];
				const double w = b[0][i]*b[1][j]*b[2][k];
				const double g[3] = {db[0][i]*b[1][j]*b[2][k], b[0][i]*db[1][j]*b[2][k], b[0][i]*b[1][j]*db[2][k]};
				for (int r = 0; r < 3; ++r) {
					x[r] += p[r]*w;
					if (jac)
						for (int c = 0; c < 3; ++c) jac[r][c] += p[r]*g[c];
				}
			}
	return x;
}

inline double Det3(const double (*a)[3])
{
	return a[0][0]*(a[1][1]*a[2][2] - a[1][2]*a[2][1]) - a[0][1]*(a[1][0]*a[2][2] - a[1][2]*a[2][0])
		+ a[0][2]*(a[1][0]*a[2][1] - a[1][1]*a[2][0]);
}

// Newton inversion of the cell's map at target, started from the box-relative
// position: true when it converges (within tol) to a parameter strictly inside
// (1e-7 from every face); false when it leaves [-0.25, 1.25]^3, meets a
// near-singular Jacobian (|det| <= 1e-18) or takes 24 steps.
inline bool StrictlyInside(const Cell& cell, const Box& box, const std::array<double, 3>& target, double tol)
{
	std::array<double, 3> u{};
	for (int d = 0; d < 3; ++d) {
		const double w = box.hi[d] - box.lo[d];
		u[d] = std::clamp(w > tol ? (target[d] - box.lo[d])/w : 0.5, 0.0, 1.0);
	}
	for (int it = 0; it < 24; ++it) {
		double J[3][3];
		const auto x = Map(cell, u, J);
		const double r[3] = {x[0] - target[0], x[1] - target[1], x[2] - target[2]};
		if (r[0]*r[0] + r[1]*r[1] + r[2]*r[2] <= tol*tol) {
			constexpr double kIn = 1e-7;
			return u[0] > kIn && u[0] < 1.0 - kIn && u[1] > kIn && u[1] < 1.0 - kIn && u[2] > kIn && u[2] < 1.0 - kIn;
		}
		const double det = Det3(J);
		if (std::abs(det) <= 1e-18) return false;
		// Cramer's rule for J du = r
		for (int c = 0; c < 3; ++c) {
			double M[3][3];
			for (int i = 0; i < 3; ++i)
				for (int j = 0; j < 3; ++j) M[i][j] = j == c ? r[i] : J[i][j];
			u[c] -= Det3(M)/det;
		}
		for (int d = 0; d < 3; ++d)
			if (u[d] < -0.25 || u[d] > 1.25) return false;
	}
	return false;
}

inline bool CellsOverlap(const Cell& a, const Box& abox, const Cell& b, const Box& bbox, double tol)
{
	constexpr double s[3] = {0.2, 0.5, 0.8};
	for (const double x : s)
		for (const double y : s)
			for (const double z : s) {
				const std::array<double, 3> u{{x, y, z}};
				if (StrictlyInside(b, bbox, Map(a, u), tol)) return true;
				if (StrictlyInside(a, abox, Map(b, u), tol)) return true;
			}
	return false;
}

inline std::int64_t Quantise(double v)
{
	constexpr double kScale = 1e12;
	if (!std::isfinite(v) || std::abs(v) > static_cast<double>(std::numeric_limits<std::int64_t>::max())/kScale)
		throw std::runtime_error("Bezier extraction coefficient cannot be quantized");
	return static_cast<std::int64_t>(std::llround(v*kScale));
}

inline std::uint64_t Fnv(std::uint64_t h, std::uint64_t v)
{
	h ^= v;
	return h*1099511628211ULL;
}

}  // namespace bezier

// The merged Bezier point set of the database (normalised coordinates, then
// mapped to the source frame by the header's origin and units per normalised
// unit), with its checks.
inline bezier::VisualMesh BuildBezierVisualMesh(const FlatDatabase& db)
{
	using namespace bezier;
	VisualMesh m;
	const int E = db.elements();
	m.checks.elements = static_cast<std::uint64_t>(E);
	m.checks.local_points = 64ULL*E;
	const auto cell_of = [&](int e) {
		Cell c;
		for (int p = 0; p < 64; ++p)
			for (int d = 0; d < 3; ++d) c[p][d] = db.bezier[(size_t)e*192 + 3*p + d];
		return c;
	};
	Box all;
	for (int e = 0; e < E; ++e) {
		const Cell c = cell_of(e);
		for (const auto& x : c) {
			for (const double v : x)
				if (!std::isfinite(v)) throw std::runtime_error("Bezier geometry contains a non-finite coordinate");
			all.Add(x);
		}
	}
	double diag2 = 0.0, xmax = 0.0;
	for (int d = 0; d < 3; ++d) {
		diag2 += (all.hi[d] - all.lo[d])*(all.hi[d] - all.lo[d]);
		xmax = std::max({xmax, std::abs(all.lo[d]), std::abs(all.hi[d])});
	}
	const double geo_tol = 1e-11*std::max(std::sqrt(diag2), 1e-12);
	const double match_tol = std::max(geo_tol, 2e-6*std::max(1.0, xmax));
	const double match_tol2 = match_tol*match_tol;

	struct Key {
		std::array<std::int64_t, 3> v;
		bool operator==(const Key& o) const { return v == o.v; }
	};
	struct KeyHash {  // splitmix64 finaliser over the three bucket indices
		static std::uint64_t Mix(std::uint64_t z)
		{
			z += 0x9e3779b97f4a7c15ULL;
			z = (z ^ (z >> 30))*0xbf58476d1ce4e5b9ULL;
			z = (z ^ (z >> 27))*0x94d049bb133111ebULL;
			return z ^ (z >> 31);
		}
		std::size_t operator()(const Key& k) const
		{
			return static_cast<std::size_t>(Mix(static_cast<std::uint64_t>(k.v[0])
				^ Mix(static_cast<std::uint64_t>(k.v[1]) ^ Mix(static_cast<std::uint64_t>(k.v[2])))));
		}
	};
	std::unordered_map<std::uint64_t, std::vector<std::int64_t>> by_column;
	std::unordered_map<Key, std::vector<std::int64_t>, KeyHash> by_position;
	std::vector<std::uint64_t> first_element;        // element that created each global point
	std::vector<std::vector<std::pair<std::int32_t, std::int64_t>>> keys;  // quantised column per global point
	std::vector<std::array<std::int64_t, 64>> local(E);
	std::vector<Box> boxes(E);
	const auto order = VtkCubicHexOrder();
	constexpr double kGauss[4] = {0.06943184420297371, 0.33000947820757187, 0.6699905217924281, 0.9305681557970262};
	auto dist2 = [](const std::array<double, 3>& a, const std::array<double, 3>& b) {
		double s = 0.0;
		for (int d = 0; d < 3; ++d) s += (a[d] - b[d])*(a[d] - b[d]);
		return s;
	};
	for (int e = 0; e < E; ++e) {
		const Cell c = cell_of(e);
		const int a0 = db.element_offsets[e], nen = db.element_offsets[e + 1] - a0;
		const std::uint64_t id = db.ids[e];
		// the element's extraction columns: (node, coefficient) over its rows, per Bezier point
		std::array<std::vector<std::pair<std::int32_t, double>>, 64> columns;
		for (int a = 0; a < nen; ++a)
			for (int k = db.extraction_offsets[a0 + a]; k < db.extraction_offsets[a0 + a + 1]; ++k)
				if (db.extraction_values[k] != 0.0)
					columns[db.extraction_columns[k]].emplace_back(db.connectivity[a0 + a], db.extraction_values[k]);
		for (int p = 0; p < 64; ++p) {
			auto& column = columns[p];
			if (column.empty()) throw std::runtime_error("Bezier extraction contains an empty column");
			std::vector<std::pair<std::int32_t, std::int64_t>> key;
			key.reserve(column.size());
			for (const auto& [n, v] : column) key.emplace_back(n, Quantise(v));
			std::sort(key.begin(), key.end());
			std::stable_sort(column.begin(), column.end(), [](const auto& l, const auto& r) { return l.first < r.first; });
			std::uint64_t h = 1469598103934665603ULL;
			for (const auto& [n, q] : key) h = Fnv(Fnv(h, static_cast<std::uint64_t>(n)), static_cast<std::uint64_t>(q));
			std::int64_t match = -1;
			if (const auto f = by_column.find(h); f != by_column.end())
				for (const auto cand : f->second)
					if (keys[static_cast<size_t>(cand)] == key) {
						match = cand;
						break;
					}
			if (match >= 0) {
				const double d2 = dist2(m.points[static_cast<size_t>(match)], c[p]);
				if (d2 > match_tol2) {
					std::ostringstream msg;
					msg << std::setprecision(17) << "matching Bezier extraction signatures have inconsistent coordinates: "
						<< "first_element=" << first_element[static_cast<size_t>(match)] << " current_element=" << id
						<< " current_local_point=" << p << " distance=" << std::sqrt(d2);
					throw std::runtime_error(msg.str());
				}
				if (d2 > geo_tol*geo_tol) {
					++m.checks.signature_coordinate_repairs;
					m.checks.maximum_signature_coordinate_difference =
						std::max(m.checks.maximum_signature_coordinate_difference, std::sqrt(d2));
				}
				local[e][p] = match;
				++m.checks.shared_point_references;
				continue;
			}
			const auto g = static_cast<std::int64_t>(m.points.size());
			by_column[h].push_back(g);
			m.points.push_back(c[p]);
			m.point_labels.push_back(-1);
			first_element.push_back(id);
			keys.push_back(std::move(key));
			for (const auto& [n, v] : column) {
				m.column_nodes.push_back(n);
				m.column_values.push_back(v);
			}
			m.column_offsets.push_back(m.column_nodes.size());
			local[e][p] = g;
		}
		std::array<std::int32_t, 6> faces{};
		for (int f = 0; f < 6; ++f) faces[f] = db.labels[(size_t)e*6 + f];
		for (int f = 0; f < 6; ++f) {
			if (faces[f] < 0) continue;
			for (int p = 0; p < 64; ++p)
				if (PointOnFace(p, f)) {
					auto& l = m.point_labels[static_cast<size_t>(local[e][p])];
					l = MergeLabel(l, faces[f]);
				}
		}
		auto sorted = local[e];
		std::sort(sorted.begin(), sorted.end());
		if (std::unique(sorted.begin(), sorted.end()) != sorted.end()) {
			++m.checks.collapsed_element_points;
			m.checks.first_bad_element = std::min(m.checks.first_bad_element, id);
		}
		for (int p = 0; p < 64; ++p) boxes[e].Add(m.points[static_cast<size_t>(local[e][p])]);
		for (const int t : order) m.connectivity.push_back(local[e][t]);
		m.offsets.push_back(static_cast<std::int64_t>(m.connectivity.size()));
		m.types.push_back(kBezierHexCell);
		m.element_ids.push_back(id);
		m.element_owners.push_back(db.owners[e]);
		m.element_labels.push_back(CellLabel(faces));
		m.face_labels.push_back(faces);
	}
	m.checks.unique_points = m.points.size();
	const std::int64_t G = static_cast<std::int64_t>(m.points.size());
	auto bucket_of = [&](const std::array<double, 3>& x) {
		Key k;
		for (int d = 0; d < 3; ++d) k.v[d] = static_cast<std::int64_t>(std::llround(x[d]/match_tol));
		return k;
	};
	const int workers = std::max(1, std::min({16, static_cast<int>(std::thread::hardware_concurrency()), E/256 + 1}));
	auto merged_cell = [&](int e) {
		Cell c;
		for (int p = 0; p < 64; ++p) c[p] = m.points[static_cast<size_t>(local[e][p])];
		return c;
	};
	// points that lie on an earlier point (within the match tolerance) without
	// sharing its extraction column: each point looks through the 27
	// neighbouring coordinate buckets for a lower-numbered point, the buckets
	// in (dz, dy, dx) order and each bucket's points in creation order, as a
	// creation-ordered scan would meet them
	{
		for (std::int64_t g = 0; g < G; ++g) by_position[bucket_of(m.points[static_cast<size_t>(g)])].push_back(g);
		struct Part {
			std::uint64_t count = 0, first_bad = kNone;
			std::int64_t first_point = -1, first_near = -1;
		};
		std::vector<Part> parts(workers);
		auto scan = [&](int w) {
			Part& part = parts[w];
			for (std::int64_t g = G*w/workers; g < G*(w + 1)/workers; ++g) {
				const auto& x = m.points[static_cast<size_t>(g)];
				const Key bucket = bucket_of(x);
				std::int64_t near = -1;
				for (int dz = -1; dz <= 1 && near < 0; ++dz)
					for (int dy = -1; dy <= 1 && near < 0; ++dy)
						for (int dx = -1; dx <= 1 && near < 0; ++dx) {
							Key nb = bucket;
							nb.v[0] += dx;
							nb.v[1] += dy;
							nb.v[2] += dz;
							const auto f = by_position.find(nb);
							if (f == by_position.end()) continue;
							for (const auto cand : f->second) {
								if (cand >= g) break;
								if (dist2(m.points[static_cast<size_t>(cand)], x) <= match_tol2) {
									near = cand;
									break;
								}
							}
						}
				if (near < 0) continue;
				++part.count;
				part.first_bad = std::min(part.first_bad, first_element[static_cast<size_t>(g)]);
				if (part.first_point < 0) {
					part.first_point = g;
					part.first_near = near;
				}
			}
		};
		std::vector<std::thread> pool;
		for (int w = 1; w < workers; ++w) pool.emplace_back(scan, w);
		scan(0);
		for (auto& t : pool) t.join();
		for (const auto& part : parts) {
			m.checks.coincident_unmerged_points += part.count;
			m.checks.first_bad_element = std::min(m.checks.first_bad_element, part.first_bad);
			if (m.checks.first_coincident_element_pair[0] == kNone && part.first_point >= 0)
				m.checks.first_coincident_element_pair = {{first_element[static_cast<size_t>(part.first_near)],
					first_element[static_cast<size_t>(part.first_point)]}};
		}
	}
	// Jacobian of every merged cell at the 4x4x4 Gauss points (order-free:
	// a count, a minimum and the smallest bad element id)
	{
		struct Part {
			std::uint64_t bad = 0, first = kNone;
			double minimum = std::numeric_limits<double>::infinity();
		};
		std::vector<Part> parts(workers);
		auto scan = [&](int w) {
			Part& part = parts[w];
			for (int e = static_cast<int>(static_cast<long long>(E)*w/workers);
					e < static_cast<int>(static_cast<long long>(E)*(w + 1)/workers); ++e) {
				const Cell cell = merged_cell(e);
				bool bad = false;
				for (const double w3 : kGauss)
					for (const double v : kGauss)
						for (const double u : kGauss) {
							double J[3][3];
							Map(cell, {{u, v, w3}}, J);
							const double det = Det3(J);
							part.minimum = std::min(part.minimum, det);
							if (!(det > 0.0)) bad = true;
						}
				if (bad) {
					++part.bad;
					part.first = std::min(part.first, db.ids[e]);
				}
			}
		};
		std::vector<std::thread> pool;
		for (int w = 1; w < workers; ++w) pool.emplace_back(scan, w);
		scan(0);
		for (auto& t : pool) t.join();
		for (const auto& part : parts) {
			m.checks.nonpositive_jacobian_elements += part.bad;
			m.checks.first_bad_element = std::min(m.checks.first_bad_element, part.first);
			m.checks.minimum_jacobian = std::min(m.checks.minimum_jacobian, part.minimum);
		}
	}
	// overlap test: sweep over the longest axis of the domain
	int axis = 0;
	for (int d = 1; d < 3; ++d)
		if (all.hi[d] - all.lo[d] > all.hi[axis] - all.lo[axis]) axis = d;
	std::vector<int> sweep(E);
	std::iota(sweep.begin(), sweep.end(), 0);
	std::stable_sort(sweep.begin(), sweep.end(), [&](int l, int r) { return boxes[l].lo[axis] < boxes[r].lo[axis]; });
	// pairs (s, t > s) of the sweep order; workers take contiguous ranges of s,
	// so the first overlapping pair in (s, t) order is the first one found by
	// the lowest worker that finds any
	struct Part {
		std::uint64_t candidates = 0, overlapping = 0;
		std::array<std::uint64_t, 2> first{{kNone, kNone}};
	};
	std::vector<Part> parts(workers);
	auto scan = [&](int w) {
		Part& part = parts[w];
		const int s0 = static_cast<int>(static_cast<long long>(E)*w/workers);
		const int s1 = static_cast<int>(static_cast<long long>(E)*(w + 1)/workers);
		for (int s = s0; s < s1; ++s) {
			const int l = sweep[s];
			for (int t = s + 1; t < E; ++t) {
				const int r = sweep[t];
				if (boxes[r].lo[axis] >= boxes[l].hi[axis] - geo_tol) break;
				if (!BoxesOverlap(boxes[l], boxes[r], geo_tol)) continue;
				int shared = 0;
				for (const auto x : local[l])
					if (std::find(local[r].begin(), local[r].end(), x) != local[r].end()) ++shared;
				if (shared >= 16) continue;
				++part.candidates;
				if (CellsOverlap(merged_cell(l), boxes[l], merged_cell(r), boxes[r], geo_tol)) {
					++part.overlapping;
					if (part.first[0] == kNone) part.first = {{db.ids[l], db.ids[r]}};
				}
			}
		}
	};
	std::vector<std::thread> pool;
	for (int w = 1; w < workers; ++w) pool.emplace_back(scan, w);
	scan(0);
	for (auto& t : pool) t.join();
	for (const auto& part : parts) {
		m.checks.overlap_candidates += part.candidates;
		m.checks.overlapping_element_pairs += part.overlapping;
		if (m.checks.first_overlap_pair[0] == kNone && part.first[0] != kNone) m.checks.first_overlap_pair = part.first;
	}
	// source coordinates
	const double scale = db.transform[3];
	if (!(scale > 0.0) || !std::isfinite(scale) || !std::isfinite(db.transform[0]) || !std::isfinite(db.transform[1])
		|| !std::isfinite(db.transform[2]))
		throw std::runtime_error("invalid Bezier visualization geometry transform");
	for (auto& x : m.points)
		for (int d = 0; d < 3; ++d) x[d] = db.transform[d] + scale*x[d];
	m.checks.maximum_signature_coordinate_difference *= scale;
	if (std::isfinite(m.checks.minimum_jacobian)) m.checks.minimum_jacobian *= scale*scale*scale;
	return m;
}

// OUTPUT with its extension replaced by .bezier_geometry.json
inline std::filesystem::path BezierReportPath(const std::filesystem::path& output)
{
	auto p = output;
	p.replace_extension(".bezier_geometry.json");
	return p;
}

inline void WriteBezierReport(const std::filesystem::path& path, const bezier::GeometryChecks& c)
{
	if (!path.parent_path().empty()) std::filesystem::create_directories(path.parent_path());
	std::ofstream out(path);
	if (!out) throw std::runtime_error("cannot create Bezier geometry report: " + path.string());
	auto id = [](std::uint64_t v) { return v == bezier::kNone ? std::string("null") : std::to_string(v); };
	auto pair = [&](const std::array<std::uint64_t, 2>& v) {
		return v[0] == bezier::kNone ? std::string("null") : "[" + std::to_string(v[0]) + ", " + std::to_string(v[1]) + "]";
	};
	out << std::setprecision(17) << "{\n"
		<< "  \"valid\": " << (c.valid() ? "true" : "false") << ",\n"
		<< "  \"elements\": " << c.elements << ",\n"
		<< "  \"local_points\": " << c.local_points << ",\n"
		<< "  \"unique_points\": " << c.unique_points << ",\n"
		<< "  \"shared_point_references\": " << c.shared_point_references << ",\n"
		<< "  \"signature_coordinate_repairs\": " << c.signature_coordinate_repairs << ",\n"
		<< "  \"maximum_signature_coordinate_difference\": " << c.maximum_signature_coordinate_difference << ",\n"
		<< "  \"coincident_unmerged_points\": " << c.coincident_unmerged_points << ",\n"
		<< "  \"collapsed_element_points\": " << c.collapsed_element_points << ",\n"
		<< "  \"nonpositive_jacobian_elements\": " << c.nonpositive_jacobian_elements << ",\n"
		<< "  \"minimum_sampled_jacobian\": " << c.minimum_jacobian << ",\n"
		<< "  \"overlap_candidates\": " << c.overlap_candidates << ",\n"
		<< "  \"overlapping_element_pairs\": " << c.overlapping_element_pairs << ",\n"
		<< "  \"first_bad_element\": " << id(c.first_bad_element) << ",\n"
		<< "  \"first_overlap_pair\": " << pair(c.first_overlap_pair) << ",\n"
		<< "  \"first_coincident_element_pair\": " << pair(c.first_coincident_element_pair) << "\n}\n";
	out.close();
	if (!out) throw std::runtime_error("cannot write Bezier geometry report: " + path.string());
}

inline void RequireValidBezierMesh(const bezier::GeometryChecks& c)
{
	if (c.valid()) return;
	throw std::runtime_error("invalid Bezier visualization geometry: collapsed_points="
		+ std::to_string(c.collapsed_element_points) + " nonpositive_jacobian_elements="
		+ std::to_string(c.nonpositive_jacobian_elements) + " coincident_unmerged_points="
		+ std::to_string(c.coincident_unmerged_points) + " overlapping_element_pairs="
		+ std::to_string(c.overlapping_element_pairs));
}

// Values of the control-point field (components per node) at the merged Bezier points.
inline std::vector<double> BezierPointValues(const bezier::VisualMesh& m, const std::vector<double>& control,
	int components, const std::string& name)
{
	std::vector<double> out(m.points.size()*components, 0.0);
	for (std::size_t p = 0; p < m.points.size(); ++p)
		for (auto k = m.column_offsets[p]; k < m.column_offsets[p + 1]; ++k) {
			const auto node = static_cast<std::size_t>(m.column_nodes[k]);
			if (control.size() < (node + 1)*components)
				throw std::runtime_error("control-point array is shorter than Bezier extraction connectivity: " + name);
			for (int c = 0; c < components; ++c) out[p*components + c] += m.column_values[k]*control[node*components + c];
		}
	return out;
}

// OUTPUT's stem with the .vtkhdf extension, beside OUTPUT
inline std::filesystem::path VtkHdfOutputPath(const std::filesystem::path& output)
{
	return output.parent_path()/(output.stem().string() + ".vtkhdf");
}

namespace hdf {

// An HDF5 identifier closed with its own closer.
class Id {
public:
	Id() = default;
	Id(hid_t v, herr_t (*close)(hid_t), const std::string& what) : v_(v), close_(close)
	{
		if (v_ < 0) throw std::runtime_error(what);
	}
	Id(const Id&) = delete;
	Id& operator=(const Id&) = delete;
	Id(Id&& o) noexcept : v_(o.v_), close_(o.close_) { o.v_ = -1; }
	Id& operator=(Id&& o) noexcept
	{
		if (this != &o) {
			Reset();
			v_ = o.v_;
			close_ = o.close_;
			o.v_ = -1;
		}
		return *this;
	}
	~Id() { Reset(); }
	hid_t get() const { return v_; }
	void Close(const std::string& what)
	{
		if (v_ < 0) return;
		const herr_t s = close_(v_);
		v_ = -1;
		if (s < 0) throw std::runtime_error(what);
	}

private:
	void Reset()
	{
		if (v_ >= 0) close_(v_);
		v_ = -1;
	}
	hid_t v_ = -1;
	herr_t (*close_)(hid_t) = nullptr;
};

inline void Ok(herr_t s, const std::string& what)
{
	if (s < 0) throw std::runtime_error(what);
}

template <class T> hid_t Type();
template <> inline hid_t Type<double>() { return H5T_NATIVE_DOUBLE; }
template <> inline hid_t Type<std::int32_t>() { return H5T_NATIVE_INT32; }
template <> inline hid_t Type<std::int64_t>() { return H5T_NATIVE_INT64; }
template <> inline hid_t Type<std::uint8_t>() { return H5T_NATIVE_UINT8; }
template <> inline hid_t Type<std::uint64_t>() { return H5T_NATIVE_UINT64; }

inline Id Group(hid_t parent, const std::string& name, bool create)
{
	return create ? Id(H5Gcreate2(parent, name.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose,
			"cannot create HDF5 group " + name)
		: Id(H5Gopen2(parent, name.c_str(), H5P_DEFAULT), H5Gclose, "cannot open HDF5 group " + name);
}

inline Id Dataset(hid_t parent, const std::string& name)
{
	return Id(H5Dopen2(parent, name.c_str(), H5P_DEFAULT), H5Dclose, "cannot open HDF5 dataset " + name);
}

// fixed-length, null-padded string attribute
inline void StringAttribute(hid_t object, const std::string& name, const std::string& value)
{
	const std::size_t n = std::max<std::size_t>(1, value.size());
	Id type(H5Tcopy(H5T_C_S1), H5Tclose, "cannot copy HDF5 string type");
	Ok(H5Tset_size(type.get(), n), "cannot size HDF5 string attribute");
	Ok(H5Tset_strpad(type.get(), H5T_STR_NULLPAD), "cannot configure HDF5 string padding");
	Id space(H5Screate(H5S_SCALAR), H5Sclose, "cannot create HDF5 scalar space");
	Id attr(H5Acreate2(object, name.c_str(), type.get(), space.get(), H5P_DEFAULT, H5P_DEFAULT), H5Aclose,
		"cannot create HDF5 attribute " + name);
	std::vector<char> buf(n, '\0');
	std::copy(value.begin(), value.end(), buf.begin());
	Ok(H5Awrite(attr.get(), type.get(), buf.data()), "cannot write HDF5 attribute " + name);
}

inline std::string ReadStringAttribute(hid_t object, const std::string& name)
{
	Id attr(H5Aopen(object, name.c_str(), H5P_DEFAULT), H5Aclose, "cannot open HDF5 attribute " + name);
	Id type(H5Aget_type(attr.get()), H5Tclose, "cannot inspect HDF5 attribute " + name);
	std::vector<char> buf(H5Tget_size(type.get()) + 1, '\0');
	Ok(H5Aread(attr.get(), type.get(), buf.data()), "cannot read HDF5 attribute " + name);
	return std::string(buf.data());
}

template <class T>
inline void ScalarAttribute(hid_t object, const std::string& name, T value, bool replace = false)
{
	if (replace) {
		Id attr(H5Aopen(object, name.c_str(), H5P_DEFAULT), H5Aclose, "cannot open HDF5 attribute " + name);
		Ok(H5Awrite(attr.get(), Type<T>(), &value), "cannot update HDF5 attribute " + name);
		return;
	}
	Id space(H5Screate(H5S_SCALAR), H5Sclose, "cannot create scalar attribute space");
	Id attr(H5Acreate2(object, name.c_str(), Type<T>(), space.get(), H5P_DEFAULT, H5P_DEFAULT), H5Aclose,
		"cannot create HDF5 attribute " + name);
	Ok(H5Awrite(attr.get(), Type<T>(), &value), "cannot write HDF5 attribute " + name);
}

template <class T>
inline T ReadScalarAttribute(hid_t object, const std::string& name)
{
	Id attr(H5Aopen(object, name.c_str(), H5P_DEFAULT), H5Aclose, "cannot open HDF5 attribute " + name);
	T v{};
	Ok(H5Aread(attr.get(), Type<T>(), &v), "cannot read HDF5 attribute " + name);
	return v;
}

// a written dataset, chunked by min(rows, 65536) rows, deflated when level > 0
template <class T>
inline Id Fixed(hid_t parent, const std::string& name, const T* data, std::vector<hsize_t> dims, int level)
{
	Id space(H5Screate_simple(static_cast<int>(dims.size()), dims.data(), nullptr), H5Sclose,
		"cannot create HDF5 dataset space");
	Id plist(H5Pcreate(H5P_DATASET_CREATE), H5Pclose, "cannot create HDF5 dataset properties");
	std::vector<hsize_t> chunk = dims;
	chunk[0] = std::max<hsize_t>(1, std::min<hsize_t>(chunk[0], 65536));
	for (auto& c : chunk) c = std::max<hsize_t>(1, c);
	Ok(H5Pset_chunk(plist.get(), static_cast<int>(chunk.size()), chunk.data()), "cannot configure HDF5 chunks");
	if (level > 0) Ok(H5Pset_deflate(plist.get(), static_cast<unsigned>(level)), "cannot configure HDF5 compression");
	Id set(H5Dcreate2(parent, name.c_str(), Type<T>(), space.get(), H5P_DEFAULT, plist.get(), H5P_DEFAULT), H5Dclose,
		"cannot create HDF5 dataset " + name);
	Ok(H5Dwrite(set.get(), Type<T>(), H5S_ALL, H5S_ALL, H5P_DEFAULT, data), "cannot write HDF5 dataset " + name);
	return set;
}

// an empty dataset growing along rows (chunks of 4096 rows); rank 2 when
// components > 1 or two_d
template <class T>
inline Id Growing(hid_t parent, const std::string& name, int components, int level, bool two_d = false)
{
	const bool rank2 = components > 1 || two_d;
	std::vector<hsize_t> dims(rank2 ? 2 : 1, 0), maxd(dims.size(), H5S_UNLIMITED), chunk(dims.size(), 1);
	chunk[0] = 4096;
	if (rank2) dims[1] = maxd[1] = chunk[1] = static_cast<hsize_t>(components);
	Id space(H5Screate_simple(static_cast<int>(dims.size()), dims.data(), maxd.data()), H5Sclose,
		"cannot create expandable HDF5 space");
	Id plist(H5Pcreate(H5P_DATASET_CREATE), H5Pclose, "cannot create HDF5 dataset properties");
	Ok(H5Pset_chunk(plist.get(), static_cast<int>(chunk.size()), chunk.data()), "cannot configure expandable HDF5 chunks");
	if (level > 0)
		Ok(H5Pset_deflate(plist.get(), static_cast<unsigned>(level)), "cannot configure expandable HDF5 compression");
	return Id(H5Dcreate2(parent, name.c_str(), Type<T>(), space.get(), H5P_DEFAULT, plist.get(), H5P_DEFAULT), H5Dclose,
		"cannot create expandable HDF5 dataset " + name);
}

inline std::vector<hsize_t> Dims(hid_t set)
{
	Id space(H5Dget_space(set), H5Sclose, "cannot inspect HDF5 dataset space");
	const int rank = H5Sget_simple_extent_ndims(space.get());
	if (rank < 1) throw std::runtime_error("invalid HDF5 dataset rank");
	std::vector<hsize_t> d(static_cast<std::size_t>(rank));
	Ok(H5Sget_simple_extent_dims(space.get(), d.data(), nullptr), "cannot inspect HDF5 dataset dimensions");
	return d;
}

inline void SetRows(hid_t set, hsize_t rows)
{
	auto d = Dims(set);
	d[0] = rows;
	Ok(H5Dset_extent(set, d.data()), "cannot resize HDF5 dataset");
}

template <class T>
inline void PutRows(hid_t set, hsize_t first, hsize_t rows, int components, const T* data)
{
	auto d = Dims(set);
	if (d[0] < first + rows) SetRows(set, first + rows);
	Id file(H5Dget_space(set), H5Sclose, "cannot open HDF5 file space");
	std::vector<hsize_t> start(d.size(), 0), count(d.size(), 1);
	start[0] = first;
	count[0] = rows;
	if (d.size() == 2) count[1] = static_cast<hsize_t>(components);
	Ok(H5Sselect_hyperslab(file.get(), H5S_SELECT_SET, start.data(), nullptr, count.data(), nullptr),
		"cannot select HDF5 output rows");
	Id mem(H5Screate_simple(static_cast<int>(count.size()), count.data(), nullptr), H5Sclose,
		"cannot create HDF5 memory space");
	Ok(H5Dwrite(set, Type<T>(), mem.get(), file.get(), H5P_DEFAULT, data), "cannot write HDF5 output rows");
}

template <class T>
inline T GetRow(hid_t set, hsize_t row)
{
	Id file(H5Dget_space(set), H5Sclose, "cannot open HDF5 file space");
	const hsize_t start[1] = {row}, count[1] = {1};
	Ok(H5Sselect_hyperslab(file.get(), H5S_SELECT_SET, start, nullptr, count, nullptr), "cannot select HDF5 row");
	Id mem(H5Screate_simple(1, count, nullptr), H5Sclose, "cannot create HDF5 row space");
	T v{};
	Ok(H5Dread(set, Type<T>(), mem.get(), file.get(), H5P_DEFAULT, &v), "cannot read HDF5 row");
	return v;
}

inline std::uint64_t HashBytes(std::uint64_t h, const void* data, std::size_t n)
{
	const auto* c = static_cast<const unsigned char*>(data);
	for (std::size_t i = 0; i < n; ++i) h = bezier::Fnv(h, c[i]);
	return h;
}

}  // namespace hdf

// FNV-1a over the points, connectivity, offsets, point labels and face labels:
// the identity a resumed file must match
inline std::uint64_t BezierGeometryHash(const bezier::VisualMesh& m)
{
	std::uint64_t h = 1469598103934665603ULL;
	for (const auto& x : m.points) h = hdf::HashBytes(h, x.data(), sizeof(x));
	h = hdf::HashBytes(h, m.connectivity.data(), m.connectivity.size()*sizeof(std::int64_t));
	h = hdf::HashBytes(h, m.offsets.data(), m.offsets.size()*sizeof(std::int64_t));
	h = hdf::HashBytes(h, m.point_labels.data(), m.point_labels.size()*sizeof(std::int32_t));
	h = hdf::HashBytes(h, m.face_labels.data(), m.face_labels.size()*sizeof(std::array<std::int32_t, 6>));
	return h;
}

struct PointField {
	std::string name;
	int components;
	std::vector<double> values;  // per control point
};

// The temporal VTKHDF UnstructuredGrid of Bezier cells: the geometry once,
// then one block of point data per time (a time equal to the last one, within
// 1e-12 relative, replaces it).  A resumed file must hold the same geometry;
// steps past its NSteps (an interrupted append) are cut off.
class BezierVtkHdfSeries {
public:
	BezierVtkHdfSeries(const std::filesystem::path& path, const bezier::VisualMesh& mesh, bool resume,
		int compression = 4)
		: path_(path), mesh_(mesh), level_(compression)
	{
		if (level_ < 0 || level_ > 9) throw std::runtime_error("VTKHDF compression must be between 0 and 9");
		H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
		if (!path_.parent_path().empty()) std::filesystem::create_directories(path_.parent_path());
		if (resume && std::filesystem::exists(path_)) {
			file_ = hdf::Id(H5Fopen(path_.string().c_str(), H5F_ACC_RDWR, H5P_DEFAULT), H5Fclose,
				"cannot open VTKHDF output " + path_.string());
			Reopen();
		} else {
			file_ = hdf::Id(H5Fcreate(path_.string().c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose,
				"cannot create VTKHDF output " + path_.string());
			Create();
		}
	}

	void Append(double time, const std::vector<PointField>& fields)
	{
		if (closing_ || file_.get() < 0) throw std::runtime_error("VTKHDF writer is closing or closed: " + path_.string());
		if (!std::isfinite(time)) throw std::runtime_error("VTKHDF physical time is not finite");
		std::vector<std::vector<double>> values;
		for (const auto& f : fields) {
			if (f.components < 1) throw std::runtime_error("Bezier point array has no components: " + f.name);
			values.push_back(BezierPointValues(mesh_, f.values, f.components, f.name));
		}
		Schema(fields);
		auto steps = hdf::Group(root_.get(), "Steps", false);
		auto times = hdf::Dataset(steps.get(), "Values");
		const auto count = static_cast<std::uint64_t>(hdf::ReadScalarAttribute<std::int64_t>(steps.get(), "NSteps"));
		bool replace = false;
		if (count > 0) {
			const double last = hdf::GetRow<double>(times.get(), count - 1);
			const double tol = 1e-12*std::max({1.0, std::abs(last), std::abs(time)});
			if (time < last - tol) throw std::runtime_error("VTKHDF time steps must be monotone");
			replace = std::abs(time - last) <= tol;
		}
		const std::uint64_t step = replace ? count - 1 : count;
		const hsize_t np = mesh_.points.size();
		auto point_data = hdf::Group(root_.get(), "PointData", false);
		for (std::size_t i = 0; i < fields.size(); ++i) {
			auto set = hdf::Dataset(point_data.get(), fields[i].name);
			hdf::PutRows(set.get(), step*np, np, fields[i].components, values[i].data());
		}
		if (!replace) {
			hdf::PutRows(times.get(), step, 1, 1, &time);
			const std::int64_t zero = 0, one = 1;
			for (const char* name : {"PartOffsets", "PointOffsets", "CellOffsets", "ConnectivityIdOffsets"}) {
				auto set = hdf::Dataset(steps.get(), name);
				hdf::PutRows(set.get(), step, 1, 1, &zero);
			}
			auto parts = hdf::Dataset(steps.get(), "NumberOfParts");
			hdf::PutRows(parts.get(), step, 1, 1, &one);
			auto offsets = hdf::Group(steps.get(), "PointDataOffsets", false);
			const auto at = static_cast<std::int64_t>(step*np);
			for (const auto& f : fields) {
				auto set = hdf::Dataset(offsets.get(), f.name);
				hdf::PutRows(set.get(), step, 1, 1, &at);
			}
			hdf::ScalarAttribute<std::int64_t>(steps.get(), "NSteps", static_cast<std::int64_t>(count + 1), true);
		}
		hdf::Ok(H5Fflush(file_.get(), H5F_SCOPE_GLOBAL), "cannot flush VTKHDF output " + path_.string());
	}

	void Close()
	{
		if (file_.get() < 0) return;
		closing_ = true;
		hdf::Ok(H5Fflush(file_.get(), H5F_SCOPE_GLOBAL), "cannot finalize VTKHDF output " + path_.string());
		root_.Close("cannot close VTKHDF root group " + path_.string());
		if (H5Fget_obj_count(file_.get(), H5F_OBJ_ALL | H5F_OBJ_LOCAL) != 1)
			throw std::runtime_error("VTKHDF finalization requires only the file handle to remain open: " + path_.string());
		file_.Close("cannot close VTKHDF output " + path_.string());
	}

	const std::filesystem::path& path() const { return path_; }

private:
	void Create()
	{
		const auto& m = mesh_;
		root_ = hdf::Group(file_.get(), "VTKHDF", true);
		{
			const std::int64_t version[2] = {2, 1};
			const hsize_t two = 2;
			hdf::Id space(H5Screate_simple(1, &two, nullptr), H5Sclose, "cannot create VTKHDF version space");
			hdf::Id attr(H5Acreate2(root_.get(), "Version", H5T_NATIVE_INT64, space.get(), H5P_DEFAULT, H5P_DEFAULT),
				H5Aclose, "cannot create VTKHDF version");
			hdf::Ok(H5Awrite(attr.get(), H5T_NATIVE_INT64, version), "cannot write VTKHDF version");
		}
		hdf::StringAttribute(root_.get(), "Type", "UnstructuredGrid");
		const std::int64_t points = static_cast<std::int64_t>(m.points.size());
		const std::int64_t cells = static_cast<std::int64_t>(m.types.size());
		const std::int64_t ids = static_cast<std::int64_t>(m.connectivity.size());
		hdf::Fixed(root_.get(), "NumberOfPoints", &points, {1}, 0);
		hdf::Fixed(root_.get(), "NumberOfCells", &cells, {1}, 0);
		hdf::Fixed(root_.get(), "NumberOfConnectivityIds", &ids, {1}, 0);
		hdf::Fixed(root_.get(), "Points", m.points.front().data(), {m.points.size(), 3}, level_);
		hdf::Fixed(root_.get(), "Connectivity", m.connectivity.data(), {m.connectivity.size()}, level_);
		hdf::Fixed(root_.get(), "Offsets", m.offsets.data(), {m.offsets.size()}, level_);
		hdf::Fixed(root_.get(), "Types", m.types.data(), {m.types.size()}, level_);
		{
			auto pd = hdf::Group(root_.get(), "PointData", true);
			hdf::Fixed(pd.get(), "boundary_label", m.point_labels.data(), {m.point_labels.size()}, level_);
		}
		{
			auto cd = hdf::Group(root_.get(), "CellData", true);
			std::vector<std::int32_t> degrees(3*m.types.size(), 3);
			auto deg = hdf::Fixed(cd.get(), "HigherOrderDegrees", degrees.data(), {m.types.size(), 3}, level_);
			hdf::StringAttribute(deg.get(), "Attribute", "HigherOrderDegrees");
			auto eid = hdf::Fixed(cd.get(), "element_id", m.element_ids.data(), {m.element_ids.size()}, level_);
			hdf::StringAttribute(eid.get(), "Attribute", "GlobalIds");
			hdf::Fixed(cd.get(), "metis_owner", m.element_owners.data(), {m.element_owners.size()}, level_);
			hdf::Fixed(cd.get(), "boundary_label", m.element_labels.data(), {m.element_labels.size()}, level_);
			hdf::Fixed(cd.get(), "boundary_face_labels", m.face_labels.front().data(), {m.face_labels.size(), 6}, level_);
		}
		{
			auto steps = hdf::Group(root_.get(), "Steps", true);
			hdf::ScalarAttribute<std::int64_t>(steps.get(), "NSteps", 0);
			hdf::Growing<double>(steps.get(), "Values", 1, 0);
			hdf::Growing<std::int64_t>(steps.get(), "PartOffsets", 1, 0);
			hdf::Growing<std::int64_t>(steps.get(), "NumberOfParts", 1, 0);
			hdf::Growing<std::int64_t>(steps.get(), "PointOffsets", 1, 0);
			hdf::Growing<std::int64_t>(steps.get(), "CellOffsets", 1, 0, true);
			hdf::Growing<std::int64_t>(steps.get(), "ConnectivityIdOffsets", 1, 0, true);
			hdf::Group(steps.get(), "PointDataOffsets", true);
		}
		auto meta = hdf::Group(file_.get(), "TubularFlowIGA", true);
		hdf::ScalarAttribute<std::uint64_t>(meta.get(), "GeometryHash", BezierGeometryHash(m));
		hdf::ScalarAttribute<std::uint64_t>(meta.get(), "UniquePoints", m.points.size());
		hdf::ScalarAttribute<std::uint64_t>(meta.get(), "Elements", m.types.size());
		hdf::ScalarAttribute<std::uint64_t>(meta.get(), "SharedPointReferences", m.checks.shared_point_references);
		hdf::ScalarAttribute<std::uint64_t>(meta.get(), "SignatureCoordinateRepairs", m.checks.signature_coordinate_repairs);
		hdf::ScalarAttribute<std::uint64_t>(meta.get(), "OverlapCandidates", m.checks.overlap_candidates);
		hdf::ScalarAttribute<std::uint64_t>(meta.get(), "OverlappingElementPairs", m.checks.overlapping_element_pairs);
	}

	void Reopen()
	{
		root_ = hdf::Group(file_.get(), "VTKHDF", false);
		if (hdf::ReadStringAttribute(root_.get(), "Type") != "UnstructuredGrid")
			throw std::runtime_error("existing VTKHDF output is not an UnstructuredGrid");
		auto meta = hdf::Group(file_.get(), "TubularFlowIGA", false);
		if (hdf::ReadScalarAttribute<std::uint64_t>(meta.get(), "GeometryHash") != BezierGeometryHash(mesh_))
			throw std::runtime_error("existing VTKHDF geometry does not match the current .ntiga database");
		auto steps = hdf::Group(root_.get(), "Steps", false);
		const auto count = static_cast<hsize_t>(hdf::ReadScalarAttribute<std::int64_t>(steps.get(), "NSteps"));
		for (const char* name : {"Values", "PartOffsets", "NumberOfParts", "PointOffsets", "CellOffsets",
				"ConnectivityIdOffsets"}) {
			auto set = hdf::Dataset(steps.get(), name);
			hdf::SetRows(set.get(), count);
		}
		if (H5Aexists(meta.get(), "PointArraySchema") <= 0) return;
		std::istringstream schema(hdf::ReadStringAttribute(meta.get(), "PointArraySchema"));
		auto point_data = hdf::Group(root_.get(), "PointData", false);
		auto offsets = hdf::Group(steps.get(), "PointDataOffsets", false);
		std::string line;
		while (std::getline(schema, line)) {
			const auto tab = line.find('\t');
			if (tab == std::string::npos) continue;
			const std::string name = line.substr(0, tab);
			auto values = hdf::Dataset(point_data.get(), name);
			hdf::SetRows(values.get(), count*mesh_.points.size());
			auto at = hdf::Dataset(offsets.get(), name);
			hdf::SetRows(at.get(), count);
		}
	}

	// "name<TAB>components<LF>" per field; fixed by the first append
	void Schema(const std::vector<PointField>& fields)
	{
		std::set<std::string> names;
		std::string schema;
		for (const auto& f : fields) {
			if (f.name.empty() || f.name.find_first_of("/\n\t") != std::string::npos)
				throw std::runtime_error("invalid VTKHDF point-array name: " + f.name);
			if (!names.insert(f.name).second) throw std::runtime_error("duplicate VTKHDF point-array name: " + f.name);
			schema += f.name + '\t' + std::to_string(f.components) + '\n';
		}
		auto meta = hdf::Group(file_.get(), "TubularFlowIGA", false);
		if (H5Aexists(meta.get(), "PointArraySchema") > 0) {
			if (hdf::ReadStringAttribute(meta.get(), "PointArraySchema") != schema)
				throw std::runtime_error("VTKHDF point-array schema changed while appending");
			return;
		}
		hdf::StringAttribute(meta.get(), "PointArraySchema", schema);
		auto point_data = hdf::Group(root_.get(), "PointData", false);
		auto steps = hdf::Group(root_.get(), "Steps", false);
		auto offsets = hdf::Group(steps.get(), "PointDataOffsets", false);
		for (const auto& f : fields) {
			auto set = hdf::Growing<double>(point_data.get(), f.name, f.components, level_);
			if (f.components == 3 && f.name == "velocity") hdf::StringAttribute(set.get(), "Attribute", "Vectors");
			hdf::Growing<std::int64_t>(offsets.get(), f.name, 1, 0);
		}
	}

	std::filesystem::path path_;
	const bezier::VisualMesh& mesh_;
	int level_ = 4;
	bool closing_ = false;
	hdf::Id file_, root_;
};

}  // namespace tfi_host
