// tfi_pack: validate a case directory's spline-stage output and write the
// indexed `.ntiga` v5 database that iga_hip (and the reference solvers) read.
// It is the counterpart of the reference packer (solvers/cpu/src/iga_pack.cpp:
// 219-418) with the same command line, inputs, byte layout, stdout records,
// exit codes and error messages:
//
//   tfi_pack CASE_DIR RANKS OUTPUT.ntiga [--legacy-text]
//
// Inputs in CASE_DIR:
//   controlmesh.vtk              legacy-VTK hexahedral control mesh with a point
//                                scalar `label` (-1 interior, 0 wall, >0 tips)
//   spline_cache.igacache        sparse binary extraction cache (preferred), or
//   cmat.txt + bzpt.txt          the legacy text form (forced by --legacy-text)
//   bzmeshinfo.txt.epart.RANKS   element owners (METIS output), one per element
//   simulation_config.json       optional; a schema-4 file supplies
//                                geometry.length_scale_to_m and its whole 3D
//                                mesh block is validated
//
// The database: header (magic NTIGADB2, version 5, ranks, elements, nodes, 64,
// 0, rank-index offset, geometry transform = control-mesh bounding-box origin,
// smallest extent, metres per source unit), element offsets [E+1], owners [E],
// element records (id, type, owner, nen, six face labels, connectivity,
// extraction rows as (count, (column, value)...) with zeros dropped, 64 Bezier
// points), then the rank index (offsets [R+1] and, per rank, the elements
// whose connectivity touches a node the rank owns in the block distribution
// of nodes).  Host-side tool: no GPU code.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <limits>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "json.hpp"

namespace fs = std::filesystem;

namespace {

using Error = std::runtime_error;

constexpr int kPoints = 64;  // Bezier points per element
constexpr char kDatabaseMagic[8] = {'N', 'T', 'I', 'G', 'A', 'D', 'B', '2'};
constexpr std::uint32_t kDatabaseVersion = 5;
constexpr char kCacheMagic[8] = {'I', 'G', 'A', 'C', 'A', 'C', 'H', '1'};
constexpr std::uint32_t kCacheVersion = 1;

// ------------------------------------------------------------ binary helpers
template <class T>
void Put(std::ostream& out, const T& value)
{
	out.write(reinterpret_cast<const char*>(&value), sizeof(T));
	if (!out) throw Error("binary database write failed");
}

template <class T>
void PutArray(std::ostream& out, const std::vector<T>& values)
{
	out.write(reinterpret_cast<const char*>(values.data()), static_cast<std::streamsize>(values.size()*sizeof(T)));
}

template <class T>
T Take(std::istream& in)
{
	T value{};
	in.read(reinterpret_cast<char*>(&value), sizeof(T));
	if (!in) throw Error("preprocessing cache is truncated");
	return value;
}

// One whitespace-separated value of a legacy text file; the message names the
// file, the element and the field that was missing or malformed.
template <class T>
T Field(std::istream& in, const char* file, std::uint64_t element, const char* what)
{
	T value{};
	if (!(in >> value)) {
		std::ostringstream msg;
		msg << file << " is truncated or malformed at element " << element << " (" << what << ')';
		throw Error(msg.str());
	}
	return value;
}

// 64-bit FNV-1a of a file's bytes (the cache's record of its control mesh).
std::uint64_t Fnv1a(const fs::path& path)
{
	std::ifstream in(path, std::ios::binary);
	if (!in) throw Error("cannot hash control mesh: " + path.string());
	std::uint64_t h = 14695981039346656037ull;
	std::vector<char> buf(1 << 16);
	while (in) {
		in.read(buf.data(), static_cast<std::streamsize>(buf.size()));
		const std::streamsize got = in.gcount();
		for (std::streamsize i = 0; i < got; ++i) {
			h ^= static_cast<unsigned char>(buf[static_cast<std::size_t>(i)]);
			h *= 1099511628211ull;
		}
	}
	if (!in.eof()) throw Error("failed while hashing control mesh: " + path.string());
	return h;
}

// -------------------------------------------------------------- control mesh
struct ControlMesh {
	std::uint64_t nodes = 0;
	std::vector<double> xyz;                    // [nodes][3]
	std::vector<std::array<std::int32_t, 8>> hexes;
	std::vector<std::int32_t> label;
};

ControlMesh LoadControlMesh(const fs::path& path)
{
	std::ifstream in(path);
	if (!in) throw Error("cannot open mesh file: " + path.string());
	ControlMesh m;
	std::string word;
	while (in >> word) {
		if (word == "POINTS") {
			std::string type;
			if (!(in >> m.nodes >> type) || m.nodes == 0) throw Error("invalid POINTS record in controlmesh.vtk");
			m.xyz.resize(3*static_cast<std::size_t>(m.nodes));
			for (double& c : m.xyz)
				if (!(in >> c)) throw Error("truncated POINTS data in controlmesh.vtk");
		} else if (word == "CELLS") {
			std::uint64_t count = 0, entries = 0;
			if (!(in >> count >> entries) || count == 0 || entries != 9*count)
				throw Error("invalid CELLS record in controlmesh.vtk");
			m.hexes.resize(static_cast<std::size_t>(count));
			for (auto& hex : m.hexes) {
				int corners = 0;
				if (!(in >> corners) || corners != 8) throw Error("controlmesh.vtk requires hexahedral cells");
				for (auto& v : hex) {
					std::int64_t id = -1;
					if (!(in >> id) || id < 0 || static_cast<std::uint64_t>(id) >= m.nodes)
						throw Error("invalid cell connectivity in controlmesh.vtk");
					v = static_cast<std::int32_t>(id);
				}
			}
		} else if (word == "POINT_DATA") {
			std::uint64_t count = 0;
			std::string scalars, name, type, lookup, table;
			int components = 0;
			if (!(in >> count >> scalars >> name >> type >> components >> lookup >> table) || count != m.nodes
				|| scalars != "SCALARS" || name != "label" || components != 1 || lookup != "LOOKUP_TABLE")
				throw Error("invalid point-label data in controlmesh.vtk");
			m.label.resize(static_cast<std::size_t>(m.nodes));
			for (auto& l : m.label)
				if (!(in >> l)) throw Error("truncated point labels in controlmesh.vtk");
			break;
		}
	}
	if (m.nodes == 0 || m.hexes.empty() || m.label.size() != m.nodes)
		throw Error("controlmesh.vtk is missing points, cells, or labels");
	return m;
}

// Boundary label of each of the six faces of every cell (-1: interior).  A face
// held by one cell is on the boundary: it takes the positive label of its
// corners (two different positive labels are an error), or 0 when all four
// corners are wall nodes.  The face numbering is that of the .ntiga records
// (VTK hexahedron corner order): bottom, front, right, back, left, top.
std::vector<std::array<std::int32_t, 6>> FaceLabels(const ControlMesh& m)
{
	static constexpr int kFace[6][4] = {{0, 3, 2, 1}, {0, 1, 5, 4}, {1, 2, 6, 5},
		{2, 3, 7, 6}, {0, 4, 7, 3}, {4, 5, 6, 7}};
	struct Entry {
		std::array<std::int32_t, 4> key;  // sorted corner ids
		std::uint64_t cell;
		int face;
	};
	std::vector<Entry> faces;
	faces.reserve(6*m.hexes.size());
	for (std::size_t c = 0; c < m.hexes.size(); ++c)
		for (int f = 0; f < 6; ++f) {
			Entry en{{}, c, f};
			for (int k = 0; k < 4; ++k) en.key[k] = m.hexes[c][kFace[f][k]];
			std::sort(en.key.begin(), en.key.end());
			faces.push_back(en);
		}
	// grouped by key in ascending key order (the order in which faces are checked),
	// each group in cell-then-face order
	std::stable_sort(faces.begin(), faces.end(), [](const Entry& a, const Entry& b) { return a.key < b.key; });
	std::vector<std::array<std::int32_t, 6>> out(m.hexes.size());
	for (auto& o : out) o.fill(-1);
	for (std::size_t i = 0; i < faces.size();) {
		std::size_t j = i + 1;
		while (j < faces.size() && faces[j].key == faces[i].key) ++j;
		const std::size_t holders = j - i;
		if (holders != 2) {
			if (holders != 1) throw Error("non-manifold control-mesh face");
			const Entry& en = faces[i];
			std::int32_t positive = -1;
			bool wall = true;
			for (int k = 0; k < 4; ++k) {
				const std::int32_t l = m.label[static_cast<std::size_t>(m.hexes[en.cell][kFace[en.face][k]])];
				wall = wall && l == 0;
				if (l > 0) {
					if (positive >= 0 && positive != l) throw Error("boundary face contains multiple positive labels");
					positive = l;
				}
			}
			if (positive >= 0) out[en.cell][en.face] = positive;
			else if (wall) out[en.cell][en.face] = 0;
			else throw Error("boundary face has no unambiguous mesh label");
		}
		i = j;
	}
	return out;
}

// Rank of a node in the block distribution of `nodes` over `ranks` (the first
// nodes % ranks ranks hold one node more).
std::uint32_t OwnerOfNode(std::uint64_t node, std::uint64_t nodes, std::uint32_t ranks)
{
	if (node >= nodes) throw Error("connectivity exceeds controlmesh point count");
	const std::uint64_t base = nodes/ranks, extra = nodes%ranks;
	const std::uint64_t wide = (base + 1)*extra;
	if (node < wide) return static_cast<std::uint32_t>(node/(base + 1));
	if (base == 0) throw Error("more ranks than mesh nodes");
	return static_cast<std::uint32_t>(extra + (node - wide)/base);
}

// ------------------------------------------------ geometry transform + config
const std::string kCase = "case_config.json: ";         // prefix of the typed-value errors
const std::string kSim = "simulation_config.json: ";   // prefix of the mesh-block errors

const tfi_host::Json& ObjectOf(const tfi_host::Json& v, const std::string& ctx)
{
	if (v.kind != tfi_host::Json::Kind::Object) throw Error(kCase + ctx + " must be an object");
	return v;
}
double NumberOf(const tfi_host::Json& v, const std::string& ctx)
{
	if (v.kind != tfi_host::Json::Kind::Number || !std::isfinite(v.number))
		throw Error(kCase + ctx + " must be a finite number");
	return v.number;
}
int IntegerOf(const tfi_host::Json& v, const std::string& ctx)
{
	const double x = NumberOf(v, ctx);
	if (std::floor(x) != x || x < 0.0 || x > 2147483647.0) throw Error(kCase + ctx + " must be a non-negative integer");
	return static_cast<int>(x);
}
std::string StringOf(const tfi_host::Json& v, const std::string& ctx)
{
	if (v.kind != tfi_host::Json::Kind::String) throw Error(kCase + ctx + " must be a string");
	return v.string;
}
void BooleanOf(const tfi_host::Json& v, const std::string& ctx)
{
	if (v.kind != tfi_host::Json::Kind::Bool) throw Error(kCase + ctx + " must be a boolean");
}
const tfi_host::Json& Member(const tfi_host::Json& obj, const std::string& key, const std::string& ctx)
{
	const tfi_host::Json* v = obj.Find(key);
	if (!v) throw Error(kSim + ctx + " requires '" + key + "'");
	return *v;
}
// unknown keys are reported in sorted key order
void OnlyKeys(const tfi_host::Json& obj, std::set<std::string> allowed, const std::string& ctx)
{
	std::vector<std::string> keys;
	for (const auto& kv : obj.object) keys.push_back(kv.first);
	std::sort(keys.begin(), keys.end());
	for (const auto& k : keys)
		if (!allowed.count(k)) throw Error(kCase + "unknown " + ctx + " key '" + k + "'");
}
double PositiveOf(const tfi_host::Json& v, const std::string& ctx)
{
	const double x = NumberOf(v, ctx);
	if (!(x > 0.0)) throw Error(kSim + ctx + " must be positive");
	return x;
}
void FractionOf(const tfi_host::Json& v, const std::string& ctx)
{
	const double x = NumberOf(v, ctx);
	if (x < 0.0 || x > 1.0) throw Error(kSim + ctx + " must be in [0,1]");
}

// The schema-4 3D mesh case (geometry + mesh blocks), validated in the order
// of the reference's MeshConfig.hpp:120-270; returns geometry.length_scale_to_m.
double MeshCaseLengthScale(const tfi_host::Json& root)
{
	ObjectOf(root, "root");
	if (IntegerOf(Member(root, "schema_version", "root"), "schema_version") != 4)
		throw Error(kSim + "3d mesh configuration requires schema_version 4");
	if (StringOf(Member(root, "dimension", "root"), "dimension") != "3d")
		throw Error(kSim + "3d mesh configuration requires dimension '3d'");
	// geometry
	const auto& geo = ObjectOf(Member(root, "geometry", "root"), "geometry");
	OnlyKeys(geo, {"kind", "file", "length_scale_to_m"}, "geometry");
	const std::string kind = StringOf(Member(geo, "kind", "geometry"), "geometry.kind");
	const std::string file = StringOf(Member(geo, "file", "geometry"), "geometry.file");
	const double scale = PositiveOf(Member(geo, "length_scale_to_m", "geometry"), "geometry.length_scale_to_m");
	if (kind != "swc_network" && kind != "obj_network")
		throw Error(kSim + "geometry.kind must be 'swc_network' or 'obj_network'");
	const fs::path gp(file);
	if (file.empty() || gp.is_absolute() || gp.has_parent_path())
		throw Error(kSim + "geometry.file must name a file in the case directory");
	const bool obj = gp.extension() == ".obj";
	if ((kind == "obj_network") != obj || (!obj && gp.extension() != ".swc"))
		throw Error(kSim + "geometry.kind and geometry.file extension do not match");
	// mesh
	const auto& mesh = ObjectOf(Member(root, "mesh", "root"), "mesh");
	OnlyKeys(mesh, {"smoothing", "centerline", "junction", "quality", "cross_section"}, "mesh");
	if (const tfi_host::Json* cs = mesh.Find("cross_section")) {
		const auto& sec = ObjectOf(*cs, "mesh.cross_section");
		OnlyKeys(sec, {"kind", "target_size"}, "mesh.cross_section");
		const std::string k = StringOf(Member(sec, "kind", "mesh.cross_section"), "mesh.cross_section.kind");
		if (k == "default") {
			OnlyKeys(sec, {"kind"}, "mesh.cross_section (default templates)");
		} else if (k == "circle") {
			const double t = PositiveOf(Member(sec, "target_size", "mesh.cross_section"),
				"mesh.cross_section.target_size");
			if (t < 1.0/128.0 || t > 1.0)
				throw Error(kSim + "mesh.cross_section.target_size must be in [1/128,1] on the radius-1 disk");
		} else {
			throw Error(kSim + "mesh.cross_section.kind must be 'default' or 'circle'");
		}
	}
	const auto& sm = ObjectOf(Member(mesh, "smoothing", "mesh"), "mesh.smoothing");
	OnlyKeys(sm, {"iterations", "bifurcation_ratio", "noise_ratio"}, "mesh.smoothing");
	IntegerOf(Member(sm, "iterations", "mesh.smoothing"), "mesh.smoothing.iterations");
	FractionOf(Member(sm, "bifurcation_ratio", "mesh.smoothing"), "mesh.smoothing.bifurcation_ratio");
	FractionOf(Member(sm, "noise_ratio", "mesh.smoothing"), "mesh.smoothing.noise_ratio");

	const std::string cl = "mesh.centerline";
	const auto& c = ObjectOf(Member(mesh, "centerline", "mesh"), cl);
	OnlyKeys(c, {"target_spacing", "max_spacing_over_diameter", "max_turn_degrees", "max_diameter_change_fraction",
		"maximum_curvature_radius_product"}, cl);
	PositiveOf(Member(c, "target_spacing", cl), cl + ".target_spacing");
	PositiveOf(Member(c, "max_spacing_over_diameter", cl), cl + ".max_spacing_over_diameter");
	if (PositiveOf(Member(c, "max_turn_degrees", cl), cl + ".max_turn_degrees") > 90.0)
		throw Error(kSim + cl + ".max_turn_degrees must not exceed 90");
	if (PositiveOf(Member(c, "max_diameter_change_fraction", cl), cl + ".max_diameter_change_fraction") > 1.0)
		throw Error(kSim + cl + ".max_diameter_change_fraction must not exceed 1");
	if (PositiveOf(Member(c, "maximum_curvature_radius_product", cl), cl + ".maximum_curvature_radius_product") >= 1.0)
		throw Error(kSim + cl + ".maximum_curvature_radius_product must be below 1");

	const std::string jn = "mesh.junction";
	const auto& j = ObjectOf(Member(mesh, "junction", "mesh"), jn);
	OnlyKeys(j, {"max_spacing_over_diameter", "upstream_clearance_over_diameter", "downstream_clearance_over_diameter",
		"minimum_angle_degrees", "maximum_radius_ratio", "optimization_iterations"}, jn);
	PositiveOf(Member(j, "max_spacing_over_diameter", jn), jn + ".max_spacing_over_diameter");
	PositiveOf(Member(j, "upstream_clearance_over_diameter", jn), jn + ".upstream_clearance_over_diameter");
	PositiveOf(Member(j, "downstream_clearance_over_diameter", jn), jn + ".downstream_clearance_over_diameter");
	if (PositiveOf(Member(j, "minimum_angle_degrees", jn), jn + ".minimum_angle_degrees") >= 90.0)
		throw Error(kSim + jn + ".minimum_angle_degrees must be below 90");
	if (PositiveOf(Member(j, "maximum_radius_ratio", jn), jn + ".maximum_radius_ratio") < 1.0)
		throw Error(kSim + jn + ".maximum_radius_ratio must be at least 1");
	if (IntegerOf(Member(j, "optimization_iterations", jn), jn + ".optimization_iterations") > 1000)
		throw Error(kSim + jn + ".optimization_iterations must not exceed 1000");

	const std::string ql = "mesh.quality";
	const auto& q = ObjectOf(Member(mesh, "quality", "mesh"), ql);
	OnlyKeys(q, {"minimum_scaled_jacobian", "check_self_intersection", "collision_safety_factor"}, ql);
	if (PositiveOf(Member(q, "minimum_scaled_jacobian", ql), ql + ".minimum_scaled_jacobian") > 1.0)
		throw Error(kSim + ql + ".minimum_scaled_jacobian must not exceed 1");
	BooleanOf(Member(q, "check_self_intersection", ql), ql + ".check_self_intersection");
	PositiveOf(Member(q, "collision_safety_factor", ql), ql + ".collision_safety_factor");
	return scale;
}

struct Transform {
	double origin[3];
	double units_per_normalized;
	double metres_per_unit;
};

Transform GeometryOf(const ControlMesh& m, const fs::path& dir)
{
	Transform t{};
	double lo[3], hi[3];
	for (int d = 0; d < 3; ++d) lo[d] = hi[d] = m.xyz[static_cast<std::size_t>(d)];
	for (std::uint64_t n = 0; n < m.nodes; ++n)
		for (int d = 0; d < 3; ++d) {
			lo[d] = std::min(lo[d], m.xyz[3*n + d]);
			hi[d] = std::max(hi[d], m.xyz[3*n + d]);
		}
	double extent = std::numeric_limits<double>::infinity();
	for (int d = 0; d < 3; ++d) extent = std::min(extent, hi[d] - lo[d]);
	if (!(extent > 0.0) || !std::isfinite(extent)) throw Error("control mesh has a degenerate normalization extent");
	for (int d = 0; d < 3; ++d) t.origin[d] = lo[d];
	t.units_per_normalized = extent;
	t.metres_per_unit = 1.0;
	const fs::path cfg = dir/"simulation_config.json";
	if (fs::exists(cfg)) {
		std::ifstream in(cfg);
		if (!in) throw Error("cannot open mesh configuration: " + cfg.string());
		std::ostringstream text;
		text << in.rdbuf();
		if (in.bad()) throw Error("cannot read complete text stream");
		const tfi_host::Json root = tfi_host::ParseJson(text.str());
		ObjectOf(root, "root");
		if (const tfi_host::Json* v = root.Find("schema_version"))
			if (IntegerOf(*v, "schema_version") == 4) t.metres_per_unit = MeshCaseLengthScale(root);
	}
	return t;
}

// --------------------------------------------------------------- the records
struct Record {
	std::uint64_t id = 0;
	std::int32_t type = 0;
	std::vector<std::int32_t> conn;
	std::vector<std::array<double, kPoints>> rows;   // dense extraction rows
	std::array<double, 3*kPoints> bezier{};
};

Record LegacyRecord(std::istream& cmat, std::istream& bzpt, std::uint64_t e)
{
	Record r;
	const auto id = Field<std::uint64_t>(cmat, "cmat.txt", e, "id");
	const auto nen = Field<std::uint64_t>(cmat, "cmat.txt", e, "basis count");
	const auto type = Field<std::int64_t>(cmat, "cmat.txt", e, "type");
	if (id != e) throw Error("cmat.txt element ids are not contiguous");
	if (nen == 0 || nen > 4096) throw Error("invalid basis count in cmat.txt");
	if (type < std::numeric_limits<std::int32_t>::min() || type > std::numeric_limits<std::int32_t>::max())
		throw Error("element type is out of range");
	r.id = id;
	r.type = static_cast<std::int32_t>(type);
	r.conn.resize(static_cast<std::size_t>(nen));
	r.rows.resize(static_cast<std::size_t>(nen));
	for (auto& node : r.conn) {
		const auto v = Field<std::int64_t>(cmat, "cmat.txt", e, "connectivity");
		if (v < 0 || v > std::numeric_limits<std::int32_t>::max()) throw Error("connectivity index is out of range");
		node = static_cast<std::int32_t>(v);
	}
	for (auto& row : r.rows)
		for (double& x : row) x = Field<double>(cmat, "cmat.txt", e, "extraction coefficient");
	for (double& x : r.bezier) x = Field<double>(bzpt, "bzpt.txt", e, "Bezier point");
	return r;
}

Record CacheRecord(std::istream& in, std::uint64_t e)
{
	Record r;
	r.id = Take<std::uint64_t>(in);
	r.type = Take<std::int32_t>(in);
	const auto nen = Take<std::uint32_t>(in);
	if (r.id != e) throw Error("preprocessing cache element ids are not contiguous");
	if (nen == 0 || nen > 4096) throw Error("invalid basis count in preprocessing cache");
	r.conn.resize(nen);
	r.rows.resize(nen);
	for (auto& node : r.conn) {
		node = Take<std::int32_t>(in);
		if (node < 0) throw Error("negative connectivity in preprocessing cache");
	}
	for (auto& row : r.rows) {
		row.fill(0.0);
		bool used[kPoints] = {};
		const auto count = Take<std::uint8_t>(in);
		if (count > kPoints) throw Error("too many extraction entries in preprocessing cache row");
		for (int k = 0; k < count; ++k) {
			const auto col = Take<std::uint8_t>(in);
			const auto val = Take<double>(in);
			if (col >= kPoints || used[col] || val == 0.0) throw Error("invalid sparse extraction entry in preprocessing cache");
			used[col] = true;
			row[col] = val;
		}
	}
	for (double& x : r.bezier) x = Take<double>(in);
	return r;
}

std::vector<std::int32_t> LoadOwners(const fs::path& path, std::uint64_t elements, std::uint32_t ranks)
{
	std::ifstream in(path);
	if (!in) throw Error("cannot open partition file: " + path.string());
	std::vector<std::int32_t> owner;
	owner.reserve(static_cast<std::size_t>(elements));
	std::int64_t v = 0;
	while (in >> v) {
		if (v < 0 || v >= ranks) throw Error("partition contains an invalid rank");
		owner.push_back(static_cast<std::int32_t>(v));
	}
	if (owner.size() != elements) throw Error("partition element count does not match preprocessing data");
	return owner;
}

void NothingAfter(std::istream& in, const char* file)
{
	std::string extra;
	if (in >> extra) throw Error(std::string(file) + " has extra data after the declared elements");
}

int Usage()
{
	std::cerr << "usage: tfi_pack CASE_DIR RANKS OUTPUT.ntiga [--legacy-text]\n";
	return 2;
}

int Pack(int argc, char** argv)
{
	if (argc < 4 || argc > 5) return Usage();
	const bool legacy_only = argc == 5;
	if (legacy_only && std::string(argv[4]) != "--legacy-text") return Usage();
	const fs::path dir(argv[1]);
	const unsigned long long ranks_in = std::stoull(argv[2]);
	if (ranks_in == 0 || ranks_in > std::numeric_limits<std::uint32_t>::max()) throw Error("invalid rank count");
	const auto ranks = static_cast<std::uint32_t>(ranks_in);
	const fs::path output(argv[3]);
	const ControlMesh mesh = LoadControlMesh(dir/"controlmesh.vtk");
	const std::uint64_t nodes = mesh.nodes;
	const Transform transform = GeometryOf(mesh, dir);

	const fs::path cache_path = dir/"spline_cache.igacache";
	const bool from_cache = !legacy_only && fs::exists(cache_path);
	std::ifstream cache, cmat, bzpt;
	std::uint64_t elements = 0;
	if (from_cache) {
		cache.open(cache_path, std::ios::binary);
		if (!cache) throw Error("cannot open " + cache_path.string());
		char magic[8];
		cache.read(magic, 8);
		const auto version = Take<std::uint32_t>(cache);
		const auto cache_elements = Take<std::uint64_t>(cache);
		const auto cache_nodes = Take<std::uint64_t>(cache);
		const auto mesh_hash = Take<std::uint64_t>(cache);
		const auto points = Take<std::uint32_t>(cache);
		Take<std::uint32_t>(cache);
		if (std::memcmp(magic, kCacheMagic, 8) != 0) throw Error("invalid preprocessing cache magic");
		if (version != kCacheVersion) throw Error("unsupported preprocessing cache version");
		if (cache_elements == 0 || cache_nodes == 0) throw Error("empty preprocessing cache");
		if (points != kPoints) throw Error("unsupported preprocessing cache Bezier point count");
		if (cache_nodes != nodes) throw Error("preprocessing cache node count does not match controlmesh.vtk");
		if (mesh_hash != Fnv1a(dir/"controlmesh.vtk"))
			throw Error("preprocessing cache does not match controlmesh.vtk content");
		elements = cache_elements;
		std::cout << "using sparse preprocessing cache " << cache_path << '\n';
	} else {
		cmat.open(dir/"cmat.txt");
		bzpt.open(dir/"bzpt.txt");
		if (!cmat) throw Error("cannot open " + (dir/"cmat.txt").string());
		if (!bzpt) throw Error("cannot open " + (dir/"bzpt.txt").string());
		elements = Field<std::uint64_t>(cmat, "cmat.txt", 0, "element count");
		const auto values = Field<std::uint64_t>(bzpt, "bzpt.txt", 0, "point count");
		if (elements == 0) throw Error("cmat.txt declares zero elements");
		if (values != elements*kPoints) throw Error("bzpt.txt point count is not 64 times the element count");
		std::cout << "using legacy cmat.txt and bzpt.txt preprocessing data\n";
	}
	if (mesh.hexes.size() != elements) throw Error("controlmesh cell count does not match preprocessing data");
	const auto faces = FaceLabels(mesh);
	const auto owner = LoadOwners(dir/("bzmeshinfo.txt.epart." + std::to_string(ranks)), elements, ranks);

	std::ofstream out(output, std::ios::binary | std::ios::trunc);
	if (!out) throw Error("cannot create " + output.string());
	out.write(kDatabaseMagic, 8);
	Put(out, kDatabaseVersion);
	Put(out, ranks);
	Put(out, elements);
	Put(out, nodes);
	Put(out, static_cast<std::uint32_t>(kPoints));
	Put(out, std::uint32_t{0});
	const auto rank_index_slot = out.tellp();
	Put(out, std::uint64_t{0});
	for (double o : transform.origin) Put(out, o);
	Put(out, transform.units_per_normalized);
	Put(out, transform.metres_per_unit);
	const auto index_slot = out.tellp();
	std::vector<std::uint64_t> offset(static_cast<std::size_t>(elements) + 1, 0);
	PutArray(out, offset);
	PutArray(out, owner);
	std::vector<std::vector<std::uint64_t>> touched(ranks);  // per rank, elements reading its nodes

	for (std::uint64_t e = 0; e < elements; ++e) {
		offset[static_cast<std::size_t>(e)] = static_cast<std::uint64_t>(out.tellp());
		const Record r = from_cache ? CacheRecord(cache, e) : LegacyRecord(cmat, bzpt, e);
		Put(out, r.id);
		Put(out, r.type);
		Put(out, owner[static_cast<std::size_t>(e)]);
		Put(out, static_cast<std::uint32_t>(r.conn.size()));
		for (std::int32_t l : faces[static_cast<std::size_t>(e)]) Put(out, l);
		std::set<std::uint32_t> readers;
		for (std::int32_t node : r.conn) {
			Put(out, node);
			readers.insert(OwnerOfNode(static_cast<std::uint64_t>(node), nodes, ranks));
		}
		for (std::uint32_t rank : readers) touched[rank].push_back(e);
		for (const auto& row : r.rows) {
			std::uint8_t count = 0;
			for (double x : row) count += x != 0.0;
			Put(out, count);
			for (int col = 0; col < kPoints; ++col)
				if (row[col] != 0.0) {
					Put(out, static_cast<std::uint8_t>(col));
					Put(out, row[col]);
				}
		}
		out.write(reinterpret_cast<const char*>(r.bezier.data()), sizeof(r.bezier));
		if (!out) throw Error("failed to write element record");
		if ((e + 1)%10000 == 0) std::cerr << "packed " << (e + 1) << '/' << elements << " elements\n";
	}
	offset.back() = static_cast<std::uint64_t>(out.tellp());
	if (from_cache) {
		char extra = 0;
		if (cache.read(&extra, 1)) throw Error("preprocessing cache has extra trailing data");
		if (!cache.eof()) throw Error("failed while finalizing preprocessing cache read");
	} else {
		NothingAfter(cmat, "cmat.txt");
		NothingAfter(bzpt, "bzpt.txt");
	}
	const auto rank_index = static_cast<std::uint64_t>(out.tellp());
	std::vector<std::uint64_t> rank_offset(static_cast<std::size_t>(ranks) + 1, 0);
	for (std::uint32_t k = 0; k < ranks; ++k) rank_offset[k + 1] = rank_offset[k] + touched[k].size();
	PutArray(out, rank_offset);
	for (const auto& list : touched) PutArray(out, list);
	out.seekp(index_slot);
	PutArray(out, offset);
	if (!out) throw Error("failed to finalize database index");
	out.seekp(rank_index_slot);
	Put(out, rank_index);
	out.close();
	std::cout << "wrote " << output << " with " << elements << " validated elements for " << ranks << " ranks\n";
	std::cout.flush();
	if (!std::cout) throw Error("cannot write complete text stream");
	return 0;
}

}  // namespace

int main(int argc, char** argv)
{
	try {
		return Pack(argc, argv);
	} catch (const std::exception& e) {
		std::cerr << "tfi_pack: " << e.what() << '\n';
		return 1;
	}
}
