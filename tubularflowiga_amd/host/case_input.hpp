// Host-side case input of the `iga_hip navier-stokes` driver: the `.ntiga`
// reader, the labelled control mesh, the inflow profile, the flow subset of
// simulation_config.json (temporal functions, Dirichlet / pressure-traction
// boundaries), boundary resolution with the wall trace, flow checkpoints and
// the text / VTK / VTU / PVD writers.  Each piece names the reference routine
// whose behaviour (accepted input, values, error conditions) it reproduces.
#pragma once

#include "json.hpp"

#include <algorithm>
#include <array>
#include <cctype>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iterator>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace tfi_host {

namespace fs = std::filesystem;

// ---------------------------------------------------------------- .ntiga
// Flattened database (solvers/cpu/include/IgaDatabase.hpp:64-161 read path,
// solvers/cuda/include/DeviceMesh.hpp:27-57 flattening): zero extraction values
// are dropped and columns kept ascending.
struct FlatDatabase {
	std::uint32_t version = 0, ranks = 0;
	std::uint64_t nodes = 0;
	std::vector<std::int32_t> element_offsets{0}, connectivity, extraction_offsets{0};
	std::vector<std::uint8_t> extraction_columns;
	std::vector<double> extraction_values, bezier;
	std::vector<std::int32_t> labels;       // 6 per element (-1 for v3)
	std::vector<std::uint64_t> ids;
	std::vector<std::int32_t> owners;       // METIS owner rank per element
	std::array<double, 5> transform{{0.0, 0.0, 0.0, 1.0, 1.0}};
	int maximum_basis = 0;

	int elements() const { return static_cast<int>(element_offsets.size()) - 1; }
};

inline FlatDatabase ReadDatabase(const std::string& path)
{
	std::ifstream in(path, std::ios::binary);
	if (!in) throw std::runtime_error("cannot open database: " + path);
	std::vector<char> data((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
	auto need = [&](std::size_t pos, std::size_t len) {
		if (pos + len > data.size()) throw std::runtime_error("truncated .ntiga database: " + path);
	};
	auto read = [&](std::size_t pos, auto& value) {
		need(pos, sizeof(value));
		std::memcpy(&value, data.data() + pos, sizeof(value));
	};
	need(0, 48);
	if (std::memcmp(data.data(), "NTIGADB2", 8) != 0) throw std::runtime_error("not a supported .ntiga database: " + path);
	FlatDatabase db;
	std::uint64_t elements = 0, rank_index = 0;
	std::uint32_t bezier_points = 0, reserved = 0;
	read(8, db.version);
	read(12, db.ranks);
	read(16, elements);
	read(24, db.nodes);
	read(32, bezier_points);
	read(36, reserved);
	read(40, rank_index);
	if (db.version < 3 || db.version > 5) throw std::runtime_error("unsupported .ntiga version " + std::to_string(db.version));
	if (bezier_points != 64) throw std::runtime_error("unsupported Bezier point count");
	if (db.nodes == 0 || rank_index == 0) throw std::runtime_error("invalid database ownership index");
	if (db.nodes > 0x7fffffffULL || elements > 0x7fffffffULL)
		throw std::runtime_error("database exceeds 32-bit node or element indices");
	std::size_t pos = 48;
	if (db.version >= 5) {
		for (int i = 0; i < 5; ++i) read(pos + 8*i, db.transform[i]);
		pos += 40;
	}
	std::vector<std::uint64_t> offsets(elements + 1);
	need(pos, 8*(elements + 1));
	std::memcpy(offsets.data(), data.data() + pos, 8*(elements + 1));
	db.element_offsets.reserve(elements + 1);
	db.bezier.reserve(elements*192);
	db.labels.reserve(elements*6);
	for (std::uint64_t e = 0; e < elements; ++e) {
		std::size_t p = offsets[e];
		std::uint64_t id = 0;
		std::int32_t type = 0, owner = 0;
		std::uint32_t nen = 0;
		read(p, id);
		read(p + 8, type);
		read(p + 12, owner);
		read(p + 16, nen);
		p += 20;
		std::array<std::int32_t, 6> lab{{-1, -1, -1, -1, -1, -1}};
		if (db.version >= 4) {
			for (int f = 0; f < 6; ++f) read(p + 4*f, lab[f]);
			p += 24;
		}
		if (nen == 0 || nen > 4096) throw std::runtime_error("invalid element basis count");
		for (std::uint32_t a = 0; a < nen; ++a) {
			std::int32_t c = 0;
			read(p + 4*a, c);
			if (c < 0 || static_cast<std::uint64_t>(c) >= db.nodes) throw std::runtime_error("element connectivity exceeds node count");
			db.connectivity.push_back(c);
		}
		p += 4ULL*nen;
		for (std::uint32_t a = 0; a < nen; ++a) {
			std::uint8_t k = 0;
			read(p, k);
			p += 1;
			double row[64] = {};
			for (int t = 0; t < k; ++t) {
				std::uint8_t col = 0;
				double v = 0.0;
				read(p, col);
				read(p + 1, v);
				p += 9;
				if (col >= 64) throw std::runtime_error("invalid extraction column");
				row[col] = v;
			}
			for (int c = 0; c < 64; ++c)
				if (row[c] != 0.0) {
					db.extraction_columns.push_back(static_cast<std::uint8_t>(c));
					db.extraction_values.push_back(row[c]);
				}
			db.extraction_offsets.push_back(static_cast<std::int32_t>(db.extraction_values.size()));
		}
		need(p, 1536);
		const std::size_t b0 = db.bezier.size();
		db.bezier.resize(b0 + 192);
		std::memcpy(db.bezier.data() + b0, data.data() + p, 1536);
		db.labels.insert(db.labels.end(), lab.begin(), lab.end());
		db.ids.push_back(id);
		db.owners.push_back(owner);
		db.maximum_basis = std::max(db.maximum_basis, static_cast<int>(nen));
		db.element_offsets.push_back(static_cast<std::int32_t>(db.connectivity.size()));
	}
	return db;
}

// ---------------------------------------------------------------- control mesh
// The labelled control mesh controlmesh.vtk: eight-node hexahedra and the
// integer point label array.  What is accepted and each error message follow
// ReadLabeledHexMesh (solvers/cpu/include/BoundarySupport.hpp:25-73); the
// reader is a cursor over the whole file held in memory.
struct LabeledHexMesh {
	std::vector<std::array<std::int32_t, 8>> cells;
	std::vector<int> labels;
};

class VtkTokenCursor {
public:
	explicit VtkTokenCursor(const std::string& path)
	{
		std::ifstream in(path, std::ios::binary);
		if (!in) throw std::runtime_error("cannot open mesh: " + path);
		text_.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
	}
	// next whitespace-separated token; false at the end of the file
	bool Token(std::string* out)
	{
		const std::size_t n = text_.size();
		while (pos_ < n && std::isspace(static_cast<unsigned char>(text_[pos_]))) ++pos_;
		if (pos_ >= n) return false;
		const std::size_t begin = pos_;
		while (pos_ < n && !std::isspace(static_cast<unsigned char>(text_[pos_]))) ++pos_;
		out->assign(text_, begin, pos_ - begin);
		return true;
	}
	// moves past the next token equal to `keyword`; false if there is none
	bool After(const char* keyword)
	{
		std::string t;
		while (Token(&t))
			if (t == keyword) return true;
		return false;
	}
	bool Real(double* v)
	{
		std::string t;
		if (!Token(&t)) return false;
		char* end = nullptr;
		*v = std::strtod(t.c_str(), &end);
		return end && *end == '\0';
	}
	bool Integer(long long* v)
	{
		std::string t;
		if (!Token(&t)) return false;
		char* end = nullptr;
		errno = 0;
		*v = std::strtoll(t.c_str(), &end, 10);
		return end && *end == '\0' && errno == 0;
	}
	bool Count(std::uint64_t* v)
	{
		long long x = 0;
		if (!Integer(&x) || x < 0) return false;
		*v = static_cast<std::uint64_t>(x);
		return true;
	}

private:
	std::string text_;
	std::size_t pos_ = 0;
};

inline LabeledHexMesh ReadLabeledHexMesh(const std::string& path, std::uint64_t nodes_expected,
	std::uint64_t elements_expected)
{
	VtkTokenCursor vtk(path);
	std::string word;
	std::uint64_t points = 0;
	if (vtk.After("POINTS") && !(vtk.Count(&points) && vtk.Token(&word)))
		throw std::runtime_error("invalid POINTS record");
	if (points != nodes_expected) throw std::runtime_error("VTK point count does not match database nodes");
	double coordinate = 0.0;
	for (std::uint64_t k = 3*points; k > 0; --k)
		if (!vtk.Real(&coordinate)) throw std::runtime_error("VTK POINTS record is truncated");
	std::uint64_t cell_count = 0, cell_values = 0;
	if (vtk.After("CELLS") && !(vtk.Count(&cell_count) && vtk.Count(&cell_values)))
		throw std::runtime_error("invalid CELLS record");
	if (cell_count != elements_expected) throw std::runtime_error("VTK cell count does not match database elements");
	LabeledHexMesh mesh;
	mesh.cells.assign(cell_count, {});
	const long long limit = static_cast<long long>(points);
	for (std::uint64_t e = 0; e < cell_count; ++e) {
		long long corners = 0;
		if (!vtk.Integer(&corners) || corners != 8) throw std::runtime_error("only eight-node VTK hexahedra are supported");
		for (int c = 0; c < 8; ++c) {
			long long id = -1;
			if (!vtk.Integer(&id) || id < 0 || id >= limit || id > INT32_MAX)
				throw std::runtime_error("invalid hexahedron connectivity");
			mesh.cells[e][c] = static_cast<std::int32_t>(id);
		}
	}
	std::uint64_t point_data = 0;
	if (vtk.After("POINT_DATA") && !vtk.Count(&point_data)) throw std::runtime_error("invalid POINT_DATA record");
	if (point_data != points) throw std::runtime_error("VTK POINT_DATA count does not match database nodes");
	if (!(vtk.After("LOOKUP_TABLE") && vtk.Token(&word))) throw std::runtime_error("VTK point labels were not found");
	mesh.labels.assign(points, 0);
	for (auto& label : mesh.labels) {
		double value = 0.0;
		if (!vtk.Real(&value)) throw std::runtime_error("VTK point-label array is truncated");
		if (!(std::floor(value) == value) || std::abs(value) > 2147483647.0)
			throw std::runtime_error("VTK point label is not integral");
		label = static_cast<int>(value);
	}
	return mesh;
}

// Wall trace: the basis functions that do not vanish on the wall.  A face of
// the control mesh is external when exactly one hexahedron uses it (a face
// shared by more than two is an error), and a wall face when its four corner
// points are labelled 0; every basis function of the face's element with a
// non-zero extraction coefficient on one of the 16 Bezier points of that face
// is in the trace.  Same rule and messages as WallTraceBasis
// (solvers/cpu/include/BoundarySupport.hpp:75-150).  Here the faces are found
// by sorting one record per (element, local face) on its sorted corner ids,
// and each local face's Bezier points are a 64-bit mask.
inline std::vector<std::int32_t> WallTraceBasis(const FlatDatabase& db, const LabeledHexMesh& mesh)
{
	// corner tuples of the six local faces, and the Bezier points (i + 4j + 16k) on each
	static const int kFaceCorners[6][4] = {{0, 3, 2, 1}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {0, 4, 7, 3},
		{4, 5, 6, 7}};
	std::uint64_t face_mask[6] = {0, 0, 0, 0, 0, 0};
	for (int pt = 0; pt < 64; ++pt) {
		const int i = pt & 3, j = (pt >> 2) & 3, k = pt >> 4;
		const bool on[6] = {k == 0, j == 0, i == 3, j == 3, i == 0, k == 3};
		for (int f = 0; f < 6; ++f)
			if (on[f]) face_mask[f] |= std::uint64_t{1} << pt;
	}
	struct Record {
		std::array<std::int32_t, 4> key;
		std::uint32_t element;
		std::uint8_t face;
	};
	const std::size_t E = mesh.cells.size();
	std::vector<Record> records(6*E);
	for (std::size_t e = 0; e < E; ++e)
		for (int f = 0; f < 6; ++f) {
			Record& r = records[6*e + f];
			for (int c = 0; c < 4; ++c) r.key[c] = mesh.cells[e][kFaceCorners[f][c]];
			std::sort(r.key.begin(), r.key.end());
			r.element = static_cast<std::uint32_t>(e);
			r.face = static_cast<std::uint8_t>(f);
		}
	std::sort(records.begin(), records.end(), [](const Record& a, const Record& b) {
		return a.key != b.key ? a.key < b.key : 6ULL*a.element + a.face < 6ULL*b.element + b.face;
	});
	std::vector<const Record*> external;
	for (std::size_t lo = 0, hi = 0; lo < records.size(); lo = hi) {
		hi = lo + 1;
		while (hi < records.size() && records[hi].key == records[lo].key) ++hi;
		if (hi - lo > 2) throw std::runtime_error("non-manifold control-mesh face");
		if (hi - lo == 1) external.push_back(&records[lo]);
	}
	std::vector<char> in_trace(db.nodes, 0);
	for (const Record* r : external) {
		const std::size_t e = r->element;
		bool wall = true;
		for (int c = 0; c < 4; ++c) wall = wall && mesh.labels.at(mesh.cells[e][kFaceCorners[r->face][c]]) == 0;
		if (!wall) continue;
		if (db.ids.at(e) != e) throw std::runtime_error("database element IDs do not match VTK cell order");
		const std::uint64_t mask = face_mask[r->face];
		for (int a = db.element_offsets[e]; a < db.element_offsets[e + 1]; ++a) {
			std::uint64_t row = 0;
			for (int k = db.extraction_offsets[a]; k < db.extraction_offsets[a + 1]; ++k)
				if (db.extraction_values[k] != 0.0) row |= std::uint64_t{1} << db.extraction_columns[k];
			if (row & mask) in_trace.at(static_cast<std::size_t>(db.connectivity[a])) = 1;
		}
	}
	std::vector<std::int32_t> trace;
	for (std::size_t n = 0; n < in_trace.size(); ++n)
		if (in_trace[n]) trace.push_back(static_cast<std::int32_t>(n));
	return trace;
}

// ReadVelocity (solvers/cpu/include/CaseInput.hpp:72-83)
inline std::vector<std::array<double, 3>> ReadVelocity(const std::string& path, std::uint64_t nodes)
{
	std::ifstream in(path);
	if (!in) throw std::runtime_error("cannot open velocity field: " + path);
	std::vector<std::array<double, 3>> v(nodes);
	for (auto& x : v)
		if (!(in >> x[0] >> x[1] >> x[2])) throw std::runtime_error("velocity field is truncated");
	std::string extra;
	if (in >> extra) throw std::runtime_error("velocity field contains extra records");
	return v;
}

// ---------------------------------------------------------------- configuration
// The Navier-Stokes subset of ParseSimulationConfiguration
// (include/SimulationConfig.hpp:262-600).
enum class TemporalKind { Constant, Sinusoid, Fourier, PeriodicTable };

struct TemporalFunction {
	std::string name, units, file, interpolation;
	TemporalKind kind = TemporalKind::Constant;
	double value = 0.0, mean = 0.0, amplitude = 0.0, period = 0.0, phase = 0.0;
	std::vector<double> cosine, sine;
};

struct Condition {
	std::string field, type, profile, waveform;
	std::vector<double> value;
	double scale = 1.0;
	bool pressure_gauge = false;
	// lumped outlet models (resistance, windkessel_rc, windkessel_rcr)
	double resistance = 0.0, proximal_resistance = 0.0, distal_resistance = 0.0, capacitance = 0.0;
	double reference_pressure = 0.0, initial_pressure = 0.0;
};

struct Boundary {
	int label = 0;
	std::string name;
	std::vector<Condition> conditions;
};

struct FlowSystem {
	std::string name, velocity, pressure, time_integration = "steady";
	double viscosity = 0.0, density = 1.0;
};

struct SimulationConfig {
	int schema_version = 0;
	double dt = 0.0;
	int steps = 0;
	std::vector<TemporalFunction> functions;
	std::vector<FlowSystem> systems;
	std::vector<Boundary> boundaries;
};

inline std::vector<double> ValueVector(const Json& v, const std::string& ctx)
{
	if (v.kind == Json::Kind::Number) return {v.number};
	std::vector<double> out;
	const auto& arr = RequireArray(v, ctx).array;
	for (std::size_t i = 0; i < arr.size(); ++i) out.push_back(RequireNumber(arr[i], ctx + "[" + std::to_string(i) + "]"));
	return out;
}

inline SimulationConfig ReadSimulationConfig(const std::string& path)
{
	std::ifstream in(path);
	if (!in) throw std::runtime_error("cannot open simulation configuration: " + path);
	std::stringstream ss;
	ss << in.rdbuf();
	const Json root_value = ParseJson(ss.str());
	const Json& root = RequireObject(root_value, "simulation_config.json");
	RequireKnownKeys(root, {"schema_version", "dimension", "geometry", "mesh", "fields", "equation_systems",
		"boundaries", "time", "temporal_functions", "velocity_sources", "physiology", "simulation_scope", "perfusate",
		"external_circuit", "coupling"}, "simulation_config.json");
	for (const char* key : {"coupling", "perfusate", "external_circuit"})
		if (root.Find(key))
			throw std::runtime_error(std::string("simulation_config.json: '") + key
				+ "' is not supported by the Navier-Stokes driver (flow-only configurations)");
	SimulationConfig c;
	c.schema_version = RequireInteger(Required(root, "schema_version", "root"), "schema_version");
	const Json* dimension = root.Find("dimension");
	if (c.schema_version == 2) {
		if (dimension || root.Find("geometry") || root.Find("mesh"))
			throw std::runtime_error("simulation_config.json: schema_version 2 does not accept dimension, geometry, or mesh");
	} else if (c.schema_version == 3 || c.schema_version == 4) {
		if (!dimension || RequireString(*dimension, "dimension") != "3d")
			throw std::runtime_error("simulation_config.json: schema_version " + std::to_string(c.schema_version)
				+ " requires dimension '3d'");
		if (c.schema_version == 3 && (root.Find("geometry") || root.Find("mesh")))
			throw std::runtime_error("simulation_config.json: schema_version 3 3d configuration does not accept geometry or mesh");
	} else {
		throw std::runtime_error("simulation_config.json: supported schema versions are 2, 3, and 4");
	}
	std::set<std::string> fields;
	const auto& farr = RequireArray(Required(root, "fields", "root"), "fields").array;
	for (std::size_t i = 0; i < farr.size(); ++i) {
		const std::string ctx = "fields[" + std::to_string(i) + "]";
		const Json& f = RequireObject(farr[i], ctx);
		RequireKnownKeys(f, {"name", "kind", "initial_value"}, ctx);
		const std::string name = RequireString(Required(f, "name", ctx), ctx + ".name");
		if (name.empty() || !fields.insert(name).second)
			throw std::runtime_error("simulation_config.json: field names must be non-empty and unique");
	}
	if (fields.empty()) throw std::runtime_error("simulation_config.json: fields cannot be empty");
	const Json& time = RequireObject(Required(root, "time", "root"), "time");
	RequireKnownKeys(time, {"dt", "steps"}, "time");
	c.dt = RequireNumber(Required(time, "dt", "time"), "time.dt");
	c.steps = RequireInteger(Required(time, "steps", "time"), "time.steps");
	if (!(c.dt > 0.0)) throw std::runtime_error("simulation_config.json: time.dt must be positive");
	std::set<std::string> fnames;
	if (const Json* tf = root.Find("temporal_functions")) {
		const auto& arr = RequireArray(*tf, "temporal_functions").array;
		for (std::size_t i = 0; i < arr.size(); ++i) {
			const std::string ctx = "temporal_functions[" + std::to_string(i) + "]";
			const Json& o = RequireObject(arr[i], ctx);
			RequireKnownKeys(o, {"name", "kind", "units", "value", "mean", "amplitude", "period", "phase", "file",
				"interpolation", "cosine", "sine"}, ctx);
			TemporalFunction fn;
			fn.name = RequireString(Required(o, "name", ctx), ctx + ".name");
			const std::string kind = RequireString(Required(o, "kind", ctx), ctx + ".kind");
			fn.units = RequireString(Required(o, "units", ctx), ctx + ".units");
			if (kind == "constant") fn.kind = TemporalKind::Constant;
			else if (kind == "sinusoid") fn.kind = TemporalKind::Sinusoid;
			else if (kind == "fourier") fn.kind = TemporalKind::Fourier;
			else if (kind == "periodic_table") fn.kind = TemporalKind::PeriodicTable;
			else throw std::runtime_error("simulation_config.json: unknown temporal function kind '" + kind + "'");
			if (fn.name.empty() || !fnames.insert(fn.name).second)
				throw std::runtime_error("simulation_config.json: temporal function names must be non-empty and unique");
			if (fn.units.empty()) throw std::runtime_error("simulation_config.json: temporal function units cannot be empty");
			if (const Json* v = o.Find("value")) fn.value = RequireNumber(*v, ctx + ".value");
			if (const Json* v = o.Find("mean")) fn.mean = RequireNumber(*v, ctx + ".mean");
			if (const Json* v = o.Find("amplitude")) fn.amplitude = RequireNumber(*v, ctx + ".amplitude");
			if (const Json* v = o.Find("period")) fn.period = RequireNumber(*v, ctx + ".period");
			if (const Json* v = o.Find("phase")) fn.phase = RequireNumber(*v, ctx + ".phase");
			if (const Json* v = o.Find("file")) fn.file = RequireString(*v, ctx + ".file");
			if (const Json* v = o.Find("interpolation")) fn.interpolation = RequireString(*v, ctx + ".interpolation");
			if (const Json* v = o.Find("cosine")) fn.cosine = ValueVector(*v, ctx + ".cosine");
			if (const Json* v = o.Find("sine")) fn.sine = ValueVector(*v, ctx + ".sine");
			if (fn.kind == TemporalKind::Constant && !o.Find("value"))
				throw std::runtime_error("simulation_config.json: constant temporal function requires value");
			if (fn.kind == TemporalKind::Sinusoid && (!o.Find("mean") || !o.Find("amplitude") || !(fn.period > 0.0)))
				throw std::runtime_error("simulation_config.json: sinusoid requires mean, amplitude, and positive period");
			if (fn.kind == TemporalKind::PeriodicTable && (fn.file.empty() || !(fn.period > 0.0) || fn.interpolation != "linear"))
				throw std::runtime_error("simulation_config.json: periodic_table requires file, positive period, and linear interpolation");
			if (fn.kind == TemporalKind::Fourier && (!(fn.period > 0.0) || fn.cosine.size() != fn.sine.size() || fn.cosine.empty()))
				throw std::runtime_error("simulation_config.json: fourier requires positive period and equal non-empty cosine/sine arrays");
			c.functions.push_back(fn);
		}
	}
	const auto& sarr = RequireArray(Required(root, "equation_systems", "root"), "equation_systems").array;
	std::set<std::string> snames;
	for (std::size_t i = 0; i < sarr.size(); ++i) {
		const std::string ctx = "equation_systems[" + std::to_string(i) + "]";
		const Json& o = RequireObject(sarr[i], ctx);
		RequireKnownKeys(o, {"name", "kind", "unknowns", "terms", "stabilization", "viscosity", "density",
			"time_integration"}, ctx);
		const std::string name = RequireString(Required(o, "name", ctx), ctx + ".name");
		if (name.empty() || !snames.insert(name).second)
			throw std::runtime_error("simulation_config.json: equation-system names must be non-empty and unique");
		const std::string kind = RequireString(Required(o, "kind", ctx), ctx + ".kind");
		if (kind != "navier_stokes") continue;
		FlowSystem s;
		s.name = name;
		const auto& unk = RequireArray(Required(o, "unknowns", ctx), ctx + ".unknowns").array;
		if (unk.size() != 2) throw std::runtime_error("navier_stokes requires [velocity, pressure] unknowns");
		s.velocity = RequireString(unk[0], ctx + ".unknowns[0]");
		s.pressure = RequireString(unk[1], ctx + ".unknowns[1]");
		for (const auto& f : {s.velocity, s.pressure})
			if (!fields.count(f)) throw std::runtime_error("simulation_config.json: unknown field '" + f + "'");
		if (const Json* v = o.Find("viscosity")) s.viscosity = RequireNumber(*v, ctx + ".viscosity");
		if (const Json* v = o.Find("density")) s.density = RequireNumber(*v, ctx + ".density");
		if (const Json* v = o.Find("time_integration")) s.time_integration = RequireString(*v, ctx + ".time_integration");
		if (!(s.viscosity > 0.0)) throw std::runtime_error("simulation_config.json: navier_stokes viscosity must be positive");
		if (!(s.density > 0.0)) throw std::runtime_error("simulation_config.json: navier_stokes density must be positive");
		if (s.time_integration != "steady" && s.time_integration != "backward_euler")
			throw std::runtime_error("simulation_config.json: navier_stokes time_integration must be 'steady' or 'backward_euler'");
		c.systems.push_back(s);
	}
	if (sarr.empty()) throw std::runtime_error("simulation_config.json: equation_systems cannot be empty");
	const auto& barr = RequireArray(Required(root, "boundaries", "root"), "boundaries").array;
	std::set<int> labels;
	for (std::size_t i = 0; i < barr.size(); ++i) {
		const std::string ctx = "boundaries[" + std::to_string(i) + "]";
		const Json& o = RequireObject(barr[i], ctx);
		RequireKnownKeys(o, {"label", "name", "conditions"}, ctx);
		Boundary b;
		b.label = RequireInteger(Required(o, "label", ctx), ctx + ".label");
		b.name = RequireString(Required(o, "name", ctx), ctx + ".name");
		if (!labels.insert(b.label).second) throw std::runtime_error("simulation_config.json: duplicate boundary label");
		const auto& carr = RequireArray(Required(o, "conditions", ctx), ctx + ".conditions").array;
		std::set<std::string> conditioned;
		for (std::size_t j = 0; j < carr.size(); ++j) {
			const std::string cctx = ctx + ".conditions[" + std::to_string(j) + "]";
			const Json& co = RequireObject(carr[j], cctx);
			RequireKnownKeys(co, {"field", "type", "value", "coefficient", "exterior_value", "profile", "scale", "waveform",
				"resistance", "proximal_resistance", "distal_resistance", "capacitance", "reference_pressure",
				"initial_pressure", "pressure_gauge", "dirichlet_dofs"}, cctx);
			Condition cond;
			cond.field = RequireString(Required(co, "field", cctx), cctx + ".field");
			cond.type = RequireString(Required(co, "type", cctx), cctx + ".type");
			static const std::set<std::string> kinds = {"dirichlet", "no_flux", "flux", "robin", "advective_outflow",
				"pressure_traction", "resistance", "windkessel_rc", "windkessel_rcr"};
			if (!kinds.count(cond.type)) throw std::runtime_error("simulation_config.json: unknown boundary type '" + cond.type + "'");
			if (!fields.count(cond.field) || !conditioned.insert(cond.field).second)
				throw std::runtime_error("simulation_config.json: boundary fields must exist and be unique per boundary");
			if (const Json* v = co.Find("value")) cond.value = ValueVector(*v, cctx + ".value");
			if (const Json* v = co.Find("profile")) cond.profile = RequireString(*v, cctx + ".profile");
			if (const Json* v = co.Find("scale")) cond.scale = RequireNumber(*v, cctx + ".scale");
			if (const Json* v = co.Find("waveform")) cond.waveform = RequireString(*v, cctx + ".waveform");
			const std::pair<const char*, double*> model_fields[] = {{"resistance", &cond.resistance},
				{"proximal_resistance", &cond.proximal_resistance}, {"distal_resistance", &cond.distal_resistance},
				{"capacitance", &cond.capacitance}, {"reference_pressure", &cond.reference_pressure},
				{"initial_pressure", &cond.initial_pressure}};
			for (const auto& mf : model_fields)
				if (const Json* v = co.Find(mf.first)) *mf.second = RequireNumber(*v, cctx + "." + mf.first);
			// SimulationConfig.hpp:576-584
			if (cond.type == "resistance" && !(cond.resistance > 0.0))
				throw std::runtime_error("simulation_config.json: resistance outlet requires positive resistance");
			if (cond.type == "windkessel_rc" && (!(cond.resistance > 0.0) || !(cond.capacitance > 0.0)))
				throw std::runtime_error("simulation_config.json: windkessel_rc requires positive resistance and capacitance");
			if (cond.type == "windkessel_rcr" && (cond.proximal_resistance < 0.0 || !(cond.distal_resistance > 0.0)
				|| !(cond.capacitance > 0.0)))
				throw std::runtime_error("simulation_config.json: windkessel_rcr requires nonnegative proximal resistance "
					"and positive distal resistance and capacitance");
			if (const Json* v = co.Find("pressure_gauge")) {
				if (v->kind != Json::Kind::Bool) throw std::runtime_error(cctx + ".pressure_gauge must be a boolean");
				cond.pressure_gauge = v->boolean;
			}
			if (cond.type == "dirichlet" && cond.value.empty() && cond.profile.empty())
				throw std::runtime_error("simulation_config.json: dirichlet requires value or profile");
			if (!cond.value.empty() && !cond.profile.empty())
				throw std::runtime_error("simulation_config.json: boundary condition cannot set both value and profile");
			if (cond.type == "pressure_traction" && cond.value.size() != 1)
				throw std::runtime_error("simulation_config.json: pressure_traction requires one value");
			if (!cond.waveform.empty() && ((cond.type != "dirichlet" && cond.type != "pressure_traction")
				|| !fnames.count(cond.waveform)))
				throw std::runtime_error("simulation_config.json: waveform must name a temporal function on a Dirichlet or pressure_traction condition");
			b.conditions.push_back(cond);
		}
		c.boundaries.push_back(b);
	}
	return c;
}

// FindNavierStokesSystem (solvers/cpu/include/GenericCaseInput.hpp:159-176)
inline const FlowSystem& FindFlowSystem(const SimulationConfig& c, const std::string& name)
{
	if (c.systems.empty()) throw std::runtime_error("simulation configuration has no navier_stokes equation system");
	if (name.empty()) return c.systems.front();
	for (const auto& s : c.systems)
		if (s.name == name) return s;
	throw std::runtime_error("simulation configuration has no navier_stokes system named '" + name + "'");
}

// EvaluateTemporalFunction / MaterializeBoundaryWaveforms (include/TemporalFunction.hpp:97-170)
inline std::vector<std::pair<double, double>> ReadTemporalCsv(const std::string& path, double period)
{
	std::ifstream in(path);
	if (!in) throw std::runtime_error("cannot open temporal CSV: " + path);
	auto trim = [](const std::string& s) {
		const auto a = s.find_first_not_of(" \t\r");
		if (a == std::string::npos) return std::string();
		return s.substr(a, s.find_last_not_of(" \t\r") - a + 1);
	};
	std::vector<std::pair<double, double>> samples;
	std::string line;
	int number = 0;
	while (std::getline(in, line)) {
		++number;
		line = trim(line);
		if (line.empty() || line[0] == '#') continue;
		const auto comma = line.find(',');
		if (comma == std::string::npos || line.find(',', comma + 1) != std::string::npos)
			throw std::runtime_error(path + " line " + std::to_string(number) + " must contain exactly time,value");
		const std::string a = trim(line.substr(0, comma)), b = trim(line.substr(comma + 1));
		if (samples.empty() && a == "time" && b == "value") continue;
		auto num = [&](const std::string& t) {
			std::size_t used = 0;
			double v = 0.0;
			try {
				v = std::stod(t, &used);
			} catch (const std::exception&) {
				throw std::runtime_error(path + " line " + std::to_string(number) + " contains a non-numeric value");
			}
			if (used != t.size() || !std::isfinite(v))
				throw std::runtime_error(path + " line " + std::to_string(number) + " contains a non-finite or malformed value");
			return v;
		};
		const double t = num(a), v = num(b);
		if (t < 0.0 || t >= period) throw std::runtime_error(path + " sample times must be in [0, period)");
		if (!samples.empty() && !(t > samples.back().first)) throw std::runtime_error(path + " sample times must be strictly increasing");
		samples.emplace_back(t, v);
	}
	if (samples.size() < 2 || samples.front().first != 0.0)
		throw std::runtime_error(path + " requires at least two samples beginning at time zero");
	return samples;
}

inline double EvaluateTemporal(const TemporalFunction& fn, double time, const fs::path& case_dir)
{
	const double pi = 3.141592653589793238462643383279502884;
	if (fn.kind == TemporalKind::Constant) return fn.value;
	double wrapped = std::fmod(time, fn.period);
	if (wrapped < 0.0) wrapped += fn.period;
	const double theta = 2.0*pi*wrapped/fn.period + fn.phase;
	if (fn.kind == TemporalKind::Sinusoid) return fn.mean + fn.amplitude*std::sin(theta);
	if (fn.kind == TemporalKind::Fourier) {
		double v = fn.mean;
		for (std::size_t h = 0; h < fn.cosine.size(); ++h) {
			const double ang = static_cast<double>(h + 1)*theta;
			v += fn.cosine[h]*std::cos(ang) + fn.sine[h]*std::sin(ang);
		}
		return v;
	}
	if (fs::path(fn.file).is_absolute()) throw std::runtime_error("temporal CSV path must be relative to the case directory");
	const auto samples = ReadTemporalCsv((case_dir/fn.file).string(), fn.period);
	const auto upper = std::upper_bound(samples.begin(), samples.end(), wrapped,
		[](double v, const std::pair<double, double>& s) { return v < s.first; });
	std::pair<double, double> lo, hi;
	if (upper == samples.end()) {
		lo = samples.back();
		hi = {fn.period, samples.front().second};
	} else {
		lo = *(upper - 1);
		hi = *upper;
	}
	return lo.second + (wrapped - lo.first)/(hi.first - lo.first)*(hi.second - lo.second);
}

inline SimulationConfig MaterializeWaveforms(const SimulationConfig& c, const fs::path& case_dir, double time)
{
	SimulationConfig r = c;
	for (auto& b : r.boundaries)
		for (auto& cond : b.conditions) {
			if (cond.waveform.empty()) continue;
			const auto it = std::find_if(c.functions.begin(), c.functions.end(),
				[&](const TemporalFunction& f) { return f.name == cond.waveform; });
			if (it == c.functions.end()) throw std::runtime_error("unknown temporal function '" + cond.waveform + "'");
			const double factor = EvaluateTemporal(*it, time, case_dir);
			if (cond.profile.empty())
				for (auto& v : cond.value) v *= factor;
			else
				cond.scale *= factor;
			cond.waveform.clear();
		}
	return r;
}

// ResolveFlowBoundaries (GenericCaseInput.hpp:178-258) + ExtractPressureTractions
// (include/PressureTraction.hpp:11-29); the wall trace is applied by the caller
// (solvers/cuda/src/iga_cuda.cu:750-759).
struct ResolvedFlowBoundaries {
	std::vector<std::int32_t> velocity_mask, pressure_mask;
	std::vector<double> velocity, pressure;   // [3N], [N]
	std::size_t velocity_nodes = 0, pressure_nodes = 0;
	std::map<int, double> tractions;
};

// outlet_pressures: the current pressures of the lumped outlet models by label
// (MaterializeOutletPressures, OutletModel.hpp:150-170), applied as tractions.
inline ResolvedFlowBoundaries ResolveFlowBoundaries(const SimulationConfig& c, const FlowSystem& s,
	const std::vector<int>& labels, const std::vector<std::array<double, 3>>& profile,
	const std::map<int, double>* outlet_pressures = nullptr)
{
	const std::size_t N = labels.size();
	if (profile.size() != N) throw std::runtime_error("velocity profile and boundary labels have different node counts");
	ResolvedFlowBoundaries r;
	r.velocity_mask.assign(N, 0);
	r.pressure_mask.assign(N, 0);
	r.velocity.assign(3*N, 0.0);
	r.pressure.assign(N, 0.0);
	std::set<int> mesh_labels;
	for (int l : labels)
		if (l >= 0) mesh_labels.insert(l);
	std::set<int> configured;
	bool traction = false;
	for (const auto& b : c.boundaries) {
		if (!mesh_labels.count(b.label))
			throw std::runtime_error("simulation_config.json boundary label " + std::to_string(b.label)
				+ " is not present in controlmesh.vtk");
		configured.insert(b.label);
		for (const auto& cond : b.conditions) {
			if (cond.field != s.velocity && cond.field != s.pressure) continue;
			if (!cond.waveform.empty())
				throw std::runtime_error("time-dependent Navier-Stokes boundary execution is not implemented");
			const bool outlet = cond.type == "resistance" || cond.type == "windkessel_rc" || cond.type == "windkessel_rcr";
			if (cond.type == "pressure_traction" || outlet) {
				if (outlet && cond.field != s.pressure)
					throw std::runtime_error("outlet models require a Navier-Stokes velocity-pressure system");
				if (!outlet && (cond.field != s.pressure || cond.value.size() != 1))
					throw std::runtime_error("pressure_traction requires the Navier-Stokes pressure field");
				double value = outlet ? 0.0 : cond.value[0];
				if (outlet) {
					const auto it = outlet_pressures ? outlet_pressures->find(b.label) : std::map<int, double>::const_iterator{};
					if (!outlet_pressures || it == outlet_pressures->end())
						throw std::runtime_error("outlet model on label " + std::to_string(b.label) + " has no pressure");
					value = it->second;
				}
				if (cond.pressure_gauge)
					throw std::runtime_error("pressure_traction already fixes the pressure reference; remove pressure_gauge "
						"to preserve every continuity equation");
				traction = true;
				if (!r.tractions.emplace(b.label, value).second)
					throw std::runtime_error("duplicate pressure traction boundary label");
				continue;
			}
			if (cond.type != "dirichlet")
				throw std::runtime_error("navier_stokes boundary field '" + cond.field
					+ "' currently supports only dirichlet conditions");
			if (cond.field == s.velocity) {
				if (cond.profile.empty() && cond.value.size() != 3)
					throw std::runtime_error("velocity Dirichlet condition requires three values or a profile");
				if (!cond.profile.empty() && cond.profile != "initial_velocityfield.txt")
					throw std::runtime_error("unsupported velocity profile '" + cond.profile + "'");
				for (std::size_t n = 0; n < N; ++n)
					if (labels[n] == b.label) {
						r.velocity_mask[n] = 1;
						for (int k = 0; k < 3; ++k)
							r.velocity[3*n + k] = cond.profile.empty() ? cond.value[k] : cond.scale*profile[n][k];
					}
			} else {
				if (cond.value.size() != 1) throw std::runtime_error("pressure Dirichlet condition requires one value");
				for (std::size_t n = 0; n < N; ++n)
					if (labels[n] == b.label) {
						r.pressure_mask[n] = 1;
						r.pressure[n] = cond.value[0];
					}
			}
		}
	}
	for (int l : mesh_labels)
		if (!configured.count(l))
			throw std::runtime_error("mesh boundary label " + std::to_string(l)
				+ " has no simulation_config.json boundary definition");
	r.velocity_nodes = static_cast<std::size_t>(std::count(r.velocity_mask.begin(), r.velocity_mask.end(), 1));
	r.pressure_nodes = static_cast<std::size_t>(std::count(r.pressure_mask.begin(), r.pressure_mask.end(), 1));
	if (r.velocity_nodes == 0 || (r.pressure_nodes == 0 && !traction))
		throw std::runtime_error("navier_stokes requires velocity Dirichlet and pressure gauge/traction boundaries");
	return r;
}

// ---------------------------------------------------------------- checkpoints
// FlowCheckpointMetadata (include/FlowCheckpoint.hpp:19-215), raw_float64 state.
// FlowCheckpointOutlet (FlowCheckpoint.hpp:33-41)
struct CheckpointOutlet {
	int label = -1;
	std::string kind;
	double flow = 0.0, pressure = 0.0, capacitor_pressure = 0.0;
};

struct CheckpointMeta {
	int schema_version = 1;
	std::uint64_t nodes = 0;
	int fields = 4, completed_step = 0;
	double physical_time = 0.0, dt = 0.0, density = 0.0, viscosity = 0.0;
	std::string state_file, state_format = "raw_float64";
	std::vector<CheckpointOutlet> outlets;
};

inline std::string JsonEscape(const std::string& v)
{
	std::string r;
	for (char c : v) {
		if (c == '\\' || c == '"') r.push_back('\\');
		if (c == '\n') r += "\\n";
		else if (c == '\r') r += "\\r";
		else if (c == '\t') r += "\\t";
		else r.push_back(c);
	}
	return r;
}

inline void WriteCheckpoint(const fs::path& prefix, CheckpointMeta meta, const std::vector<double>& state)
{
	const fs::path state_path = prefix.string() + ".state";
	{
		std::ofstream out(state_path, std::ios::binary);
		if (!out) throw std::runtime_error("cannot create flow checkpoint state: " + state_path.string());
		out.write(reinterpret_cast<const char*>(state.data()), static_cast<std::streamsize>(state.size()*sizeof(double)));
		out.close();
		if (!out) throw std::runtime_error("cannot write flow checkpoint state: " + state_path.string());
	}
	meta.state_file = state_path.filename().string();
	meta.state_format = "raw_float64";
	const fs::path json_path = prefix.string() + ".json";
	std::ofstream out(json_path);
	if (!out) throw std::runtime_error("cannot create flow checkpoint metadata: " + json_path.string());
	out << std::setprecision(17) << "{\n"
		<< "  \"schema_version\": " << meta.schema_version << ",\n"
		<< "  \"nodes\": " << meta.nodes << ",\n"
		<< "  \"fields\": " << meta.fields << ",\n"
		<< "  \"completed_step\": " << meta.completed_step << ",\n"
		<< "  \"physical_time\": " << meta.physical_time << ",\n"
		<< "  \"dt\": " << meta.dt << ",\n"
		<< "  \"density\": " << meta.density << ",\n"
		<< "  \"viscosity\": " << meta.viscosity << ",\n"
		<< "  \"state_file\": \"" << JsonEscape(meta.state_file) << "\",\n"
		<< "  \"state_format\": \"" << JsonEscape(meta.state_format) << "\",\n"
		<< "  \"outlets\": [";
	for (std::size_t i = 0; i < meta.outlets.size(); ++i) {
		const auto& o = meta.outlets[i];
		out << (i ? ",\n" : "\n") << "    {\"label\": " << o.label << ", \"kind\": \"" << JsonEscape(o.kind)
			<< "\", \"flow\": " << o.flow << ", \"pressure\": " << o.pressure
			<< ", \"capacitor_pressure\": " << o.capacitor_pressure << "}";
	}
	if (!meta.outlets.empty()) out << '\n';
	out << "  ]\n}\n";
	if (!out) throw std::runtime_error("cannot write flow checkpoint metadata: " + json_path.string());
}

inline CheckpointMeta ReadCheckpointMeta(const fs::path& prefix)
{
	const fs::path path = prefix.string() + ".json";
	std::ifstream in(path);
	if (!in) throw std::runtime_error("cannot open flow checkpoint metadata: " + path.string());
	std::stringstream ss;
	ss << in.rdbuf();
	const Json root_value = ParseJson(ss.str());
	const Json& root = RequireObject(root_value, "flow checkpoint");
	RequireKnownKeys(root, {"schema_version", "nodes", "fields", "completed_step", "physical_time", "dt", "density",
		"viscosity", "state_file", "state_format", "outlets"}, "flow checkpoint");
	CheckpointMeta m;
	m.schema_version = RequireInteger(Required(root, "schema_version", "flow checkpoint"), "flow checkpoint.schema_version");
	const double nodes = RequireNumber(Required(root, "nodes", "flow checkpoint"), "flow checkpoint.nodes");
	if (!(nodes > 0.0) || std::floor(nodes) != nodes) throw std::runtime_error("flow checkpoint.nodes must be a positive integer");
	m.nodes = static_cast<std::uint64_t>(nodes);
	m.fields = RequireInteger(Required(root, "fields", "flow checkpoint"), "flow checkpoint.fields");
	m.completed_step = RequireInteger(Required(root, "completed_step", "flow checkpoint"), "flow checkpoint.completed_step");
	m.physical_time = RequireNumber(Required(root, "physical_time", "flow checkpoint"), "flow checkpoint.physical_time");
	m.dt = RequireNumber(Required(root, "dt", "flow checkpoint"), "flow checkpoint.dt");
	m.density = RequireNumber(Required(root, "density", "flow checkpoint"), "flow checkpoint.density");
	m.viscosity = RequireNumber(Required(root, "viscosity", "flow checkpoint"), "flow checkpoint.viscosity");
	m.state_file = RequireString(Required(root, "state_file", "flow checkpoint"), "flow checkpoint.state_file");
	m.state_format = "petsc_binary";
	if (const Json* f = root.Find("state_format")) m.state_format = RequireString(*f, "flow checkpoint.state_format");
	if (const Json* o = root.Find("outlets")) {
		// FlowCheckpoint.hpp:137-164
		const auto& arr = RequireArray(*o, "flow checkpoint.outlets").array;
		std::set<int> seen;
		for (std::size_t i = 0; i < arr.size(); ++i) {
			const std::string ctx = "flow checkpoint.outlets[" + std::to_string(i) + "]";
			const Json& item = RequireObject(arr[i], ctx);
			RequireKnownKeys(item, {"label", "kind", "flow", "pressure", "capacitor_pressure"}, ctx);
			CheckpointOutlet out;
			out.label = RequireInteger(Required(item, "label", ctx), ctx + ".label");
			out.kind = RequireString(Required(item, "kind", ctx), ctx + ".kind");
			out.flow = RequireNumber(Required(item, "flow", ctx), ctx + ".flow");
			out.pressure = RequireNumber(Required(item, "pressure", ctx), ctx + ".pressure");
			out.capacitor_pressure = RequireNumber(Required(item, "capacitor_pressure", ctx), ctx + ".capacitor_pressure");
			if (!seen.insert(out.label).second
				|| (out.kind != "resistance" && out.kind != "windkessel_rc" && out.kind != "windkessel_rcr")
				|| !std::isfinite(out.flow) || !std::isfinite(out.pressure) || !std::isfinite(out.capacitor_pressure))
				throw std::runtime_error(ctx + " contains invalid outlet state");
			m.outlets.push_back(out);
		}
	}
	if (m.schema_version != 1) throw std::runtime_error("unsupported flow checkpoint schema_version");
	if (m.nodes == 0 || m.fields != 4 || m.completed_step < 0 || m.physical_time < 0.0 || !(m.dt > 0.0)
		|| !(m.density > 0.0) || !(m.viscosity > 0.0) || m.state_file.empty()
		|| (m.state_format != "petsc_binary" && m.state_format != "raw_float64"))
		throw std::runtime_error("flow checkpoint metadata contains invalid values");
	return m;
}

// ValidateFlowCheckpoint (FlowCheckpoint.hpp:197-212)
inline void ValidateCheckpoint(const CheckpointMeta& m, std::uint64_t nodes, int steps, double dt, double density,
	double viscosity)
{
	auto close = [](double a, double b) { return std::abs(a - b) <= 1e-12*std::max({1.0, std::abs(a), std::abs(b)}); };
	if (m.nodes != nodes) throw std::runtime_error("flow checkpoint node count does not match database");
	if (m.completed_step > steps) throw std::runtime_error("flow checkpoint is beyond configured final step");
	if (!close(m.physical_time, m.completed_step*dt))
		throw std::runtime_error("flow checkpoint physical time is inconsistent with completed step");
	if (!close(m.dt, dt)) throw std::runtime_error("flow checkpoint dt does not match configuration");
	if (!close(m.density, density)) throw std::runtime_error("flow checkpoint density does not match configuration");
	if (!close(m.viscosity, viscosity)) throw std::runtime_error("flow checkpoint viscosity does not match configuration");
}

// ReadCudaFlowCheckpoint (solvers/cuda/src/iga_cuda.cu:572-591)
inline std::vector<double> ReadCheckpointState(const fs::path& prefix, const CheckpointMeta& m)
{
	if (m.state_format != "raw_float64") throw std::runtime_error("flow restart requires raw_float64 checkpoint state");
	fs::path path(m.state_file);
	if (path.is_relative()) path = fs::path(prefix.string() + ".json").parent_path()/path;
	std::ifstream in(path, std::ios::binary);
	if (!in) throw std::runtime_error("cannot open flow checkpoint state: " + path.string());
	std::vector<double> v(static_cast<std::size_t>(m.nodes)*m.fields);
	in.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(v.size()*sizeof(double)));
	char extra = 0;
	if (!in || in.read(&extra, 1)) throw std::runtime_error("flow checkpoint state is truncated or has trailing data");
	return v;
}

// ---------------------------------------------------------------- outputs
// path helpers (FlowCheckpoint.hpp:50-60, VtkOutput.hpp:150-173, VelocitySeries.hpp:87-90)
inline fs::path TimeIndexedPath(const fs::path& base, int step)
{
	std::ostringstream s;
	s << ".step" << std::setw(6) << std::setfill('0') << step;
	const std::string ext = base.extension().string();
	if (ext.empty()) return base.string() + s.str();
	return base.parent_path()/(base.stem().string() + s.str() + ext);
}
inline fs::path VtuStepPath(const fs::path& base, int step)
{
	std::ostringstream s;
	s << base.stem().string() << ".step" << std::setw(6) << std::setfill('0') << step << ".vtu";
	return base.parent_path()/s.str();
}
inline fs::path VtuFinalPath(const fs::path& base) { return base.parent_path()/(base.stem().string() + ".vtu"); }
inline fs::path PvdPath(const fs::path& base) { return base.parent_path()/(base.stem().string() + ".pvd"); }
inline fs::path SeriesPath(const fs::path& base) { return base.parent_path()/(base.stem().string() + ".series.csv"); }

// WriteNavierStokes (iga_cuda.cu:386-399): "vx vy vz" per node and PATH.pressure
inline void WriteStateText(const fs::path& path, const std::vector<double>& s)
{
	std::ofstream v(path), p(path.string() + ".pressure");
	if (!v || !p) throw std::runtime_error("cannot create Navier-Stokes output");
	v << std::setprecision(17);
	p << std::setprecision(17);
	for (std::size_t n = 0; n < s.size()/4; ++n) {
		v << s[4*n] << ' ' << s[4*n + 1] << ' ' << s[4*n + 2] << '\n';
		p << s[4*n + 3] << '\n';
	}
}

// WriteNavierStokesVtk (iga_cuda.cu:351-363): the control mesh followed by the fields
inline void WriteStateVtk(const fs::path& mesh_path, const fs::path& path, const std::vector<double>& s)
{
	std::ifstream mesh(mesh_path);
	std::ofstream out(path);
	if (!mesh) throw std::runtime_error("cannot open VTK mesh " + mesh_path.string());
	if (!out) throw std::runtime_error("cannot create " + path.string());
	out << mesh.rdbuf() << "\nVECTORS Velocity double\n" << std::setprecision(17);
	for (std::size_t n = 0; n < s.size()/4; ++n) out << s[4*n] << ' ' << s[4*n + 1] << ' ' << s[4*n + 2] << '\n';
	out << "SCALARS Pressure double 1\nLOOKUP_TABLE default\n";
	for (std::size_t n = 0; n < s.size()/4; ++n) out << s[4*n + 3] << '\n';
}

// WriteVtu (include/VtkOutput.hpp:92-148) with point arrays velocity / pressure
// Named point arrays of a VTU file (VtkPointArray, VtkOutput.hpp)
struct PointArray {
	std::string name;
	int components = 1;
	std::vector<double> values;  // [points * components]
};

inline void WritePointArraysVtu(const fs::path& mesh_path, const fs::path& path, const std::vector<PointArray>& arrays,
	double time)
{
	std::ifstream in(mesh_path);
	if (!in) throw std::runtime_error("cannot open VTK mesh: " + mesh_path.string());
	std::vector<double> points;
	std::vector<std::int64_t> conn, offsets;
	std::vector<unsigned> types;
	std::string token;
	std::size_t npts = 0, ncells = 0;
	while (in >> token) {
		if (token == "POINTS") {
			std::string type;
			in >> npts >> type;
			points.resize(3*npts);
			for (auto& v : points)
				if (!(in >> v)) throw std::runtime_error("truncated VTK points: " + mesh_path.string());
		} else if (token == "CELLS") {
			std::size_t entries = 0;
			in >> ncells >> entries;
			for (std::size_t c = 0; c < ncells; ++c) {
				std::size_t k = 0;
				if (!(in >> k)) throw std::runtime_error("truncated VTK cells: " + mesh_path.string());
				for (std::size_t i = 0; i < k; ++i) {
					std::int64_t idx = -1;
					if (!(in >> idx) || idx < 0 || static_cast<std::size_t>(idx) >= npts)
						throw std::runtime_error("invalid VTK cell connectivity: " + mesh_path.string());
					conn.push_back(idx);
				}
				offsets.push_back(static_cast<std::int64_t>(conn.size()));
			}
		} else if (token == "CELL_TYPES") {
			std::size_t k = 0;
			in >> k;
			types.resize(k);
			for (auto& t : types)
				if (!(in >> t)) throw std::runtime_error("truncated VTK cell types: " + mesh_path.string());
			break;
		}
	}
	if (npts == 0 || ncells == 0 || offsets.size() != ncells || types.size() != ncells)
		throw std::runtime_error("VTK mesh is not a complete ASCII unstructured grid: " + mesh_path.string());
	for (const auto& a : arrays)
		if (a.values.size() != static_cast<std::size_t>(a.components)*npts)
			throw std::runtime_error("VTU point-array size mismatch");
	if (!path.parent_path().empty()) fs::create_directories(path.parent_path());
	std::ofstream out(path);
	if (!out) throw std::runtime_error("cannot create VTU output: " + path.string());
	out << std::setprecision(17) << "<?xml version=\"1.0\"?>\n"
		<< "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
		<< "  <UnstructuredGrid>\n"
		<< "    <FieldData><DataArray type=\"Float64\" Name=\"TimeValue\" NumberOfTuples=\"1\" format=\"ascii\">" << time
		<< "</DataArray></FieldData>\n"
		<< "    <Piece NumberOfPoints=\"" << npts << "\" NumberOfCells=\"" << ncells << "\">\n      <PointData>\n";
	for (const auto& a : arrays) {
		out << "        <DataArray type=\"Float64\" Name=\"" << a.name << "\" NumberOfComponents=\"" << a.components
			<< "\" format=\"ascii\">\n          ";
		for (double v : a.values) out << v << ' ';
		out << "\n        </DataArray>\n";
	}
	out << "      </PointData>\n"
		<< "      <Points><DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n        ";
	for (double v : points) out << v << ' ';
	out << "\n      </DataArray></Points>\n      <Cells>\n"
		<< "        <DataArray type=\"Int64\" Name=\"connectivity\" format=\"ascii\">\n          ";
	for (auto v : conn) out << v << ' ';
	out << "\n        </DataArray>\n        <DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n          ";
	for (auto v : offsets) out << v << ' ';
	out << "\n        </DataArray>\n        <DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n          ";
	for (auto v : types) out << v << ' ';
	out << "\n        </DataArray>\n      </Cells>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n";
	out.close();
	if (!out) throw std::runtime_error("cannot write VTU output: " + path.string());
}

inline void WriteStateVtu(const fs::path& mesh_path, const fs::path& path, const std::vector<double>& s, double time)
{
	const std::size_t n = s.size()/4;
	PointArray velocity{"velocity", 3, std::vector<double>(3*n)}, pressure{"pressure", 1, std::vector<double>(n)};
	for (std::size_t i = 0; i < n; ++i) {
		for (int d = 0; d < 3; ++d) velocity.values[3*i + d] = s[4*i + d];
		pressure.values[i] = s[4*i + 3];
	}
	WritePointArraysVtu(mesh_path, path, {velocity, pressure}, time);
}

// WritePvd (VtkOutput.hpp:175-190)
inline void WritePvd(const fs::path& path, const std::vector<std::pair<double, fs::path>>& snaps)
{
	std::ofstream out(path);
	if (!out) throw std::runtime_error("cannot create PVD output: " + path.string());
	out << "<?xml version=\"1.0\"?>\n<VTKFile type=\"Collection\" version=\"0.1\" byte_order=\"LittleEndian\">\n  <Collection>\n";
	for (const auto& s : snaps)
		out << "    <DataSet timestep=\"" << std::setprecision(17) << s.first << "\" group=\"\" part=\"0\" file=\""
			<< s.second.filename().string() << "\"/>\n";
	out << "  </Collection>\n</VTKFile>\n";
	out.close();
	if (!out) throw std::runtime_error("cannot write PVD output: " + path.string());
}

// WriteVelocityManifest (VelocitySeries.hpp:92-103)
inline void WriteSeries(const fs::path& path, const std::vector<std::pair<double, fs::path>>& snaps)
{
	std::ofstream out(path);
	if (!out) throw std::runtime_error("cannot create velocity manifest: " + path.string());
	out << "time,file\n";
	for (const auto& s : snaps) out << std::setprecision(17) << s.first << ',' << s.second.filename().string() << '\n';
	out.close();
	if (!out) throw std::runtime_error("cannot write velocity manifest: " + path.string());
}

}  // namespace tfi_host
