// Input and output of the two-species neuron transport command
// (`iga_cuda transport`, solvers/cuda/src/iga_cuda.cu:621-733): the legacy
// parameter file, the optional case_config.json boundary rules, the boundary
// resolution and the transport output files.
#pragma once

#include "case_input.hpp"

#include <algorithm>
#include <array>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

namespace tfi_host {

// TransportParameters / ReadTransportParameters (solvers/cpu/include/CaseInput.hpp:19-33, 85-119):
// whitespace-separated "key value" pairs; the twelve keys below are required.
struct TransportParameters {
	double diffusion = 0.0, vplus = 0.0, vminus = 0.0, kplus = 0.0, kminus = 0.0, detach_plus = 0.0,
		detach_minus = 0.0, dt = 0.0;
	int steps = 0;
	double n0_bc = 0.0, nplus_bc = 0.0, nminus_bc = 0.0, artificial_diffusion = 0.0;
};

inline TransportParameters ReadTransportParameters(const std::string& path)
{
	std::ifstream in(path);
	if (!in) throw std::runtime_error("cannot open simulation parameters: " + path);
	std::map<std::string, double> values;
	std::string key;
	double value = 0.0;
	while (in >> key >> value) values[key] = value;
	auto need = [&](const char* name) {
		const auto it = values.find(name);
		if (it == values.end()) throw std::runtime_error(std::string("missing transport parameter: ") + name);
		return it->second;
	};
	TransportParameters p;
	p.diffusion = need("D");
	p.vplus = need("vplus");
	p.vminus = need("vminus");
	p.kplus = need("kplus");
	p.kminus = need("kminus");
	p.detach_plus = need("k'plus");
	p.detach_minus = need("k'minus");
	p.dt = need("dt");
	p.steps = static_cast<int>(need("nstep"));
	p.n0_bc = need("N0bc");
	p.nplus_bc = need("Nplusbc");
	p.nminus_bc = need("Nminusbc");
	if (values.count("artificial_diffusion")) p.artificial_diffusion = values["artificial_diffusion"];
	else if (values.count("artificial_diffusion_weight")) p.artificial_diffusion = values["artificial_diffusion_weight"];
	if (p.dt <= 0.0 || p.steps < 0) throw std::runtime_error("dt and nstep must be positive");
	return p;
}

// case_config.json boundary rules (include/CaseConfig.hpp:22-50, 281-387).
enum class RuleType { Interior, Wall, Inlet, Outlet };

struct BoundaryRule {
	int label = -1;
	RuleType type = RuleType::Interior;
	bool has_velocity = false, has_velocity_scale = false, has_pressure = false, has_transport = false;
	std::array<double, 3> velocity{};
	double velocity_scale = 1.0, pressure = 0.0, n0 = 0.0, nplus = 0.0;
};

struct CaseRules {
	bool present = false;
	bool inherit_legacy = true;
	std::vector<BoundaryRule> rules;
};

inline CaseRules ReadCaseRules(const std::string& path)
{
	CaseRules c;
	if (!std::filesystem::exists(path)) return c;
	std::ifstream in(path);
	if (!in) throw std::runtime_error("cannot open case configuration: " + path);
	const std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
	const Json root = ParseJson(text);
	const std::string cfg = "case_config.json: ";
	RequireObject(root, cfg + "root");
	RequireKnownKeys(root, {"schema_version", "boundaries"}, cfg + "root");
	if (const Json* v = root.Find("schema_version"))
		if (RequireInteger(*v, cfg + "schema_version") != 1)
			throw std::runtime_error("case_config.json: only schema_version 1 is supported");
	const Json* boundaries = root.Find("boundaries");
	if (!boundaries) throw std::runtime_error("case_config.json: missing boundaries object");
	RequireObject(*boundaries, cfg + "boundaries");
	RequireKnownKeys(*boundaries, {"inherit_legacy", "conditions"}, cfg + "boundaries");
	c.present = true;
	if (const Json* v = boundaries->Find("inherit_legacy")) {
		if (v->kind != Json::Kind::Bool) throw std::runtime_error(cfg + "boundaries.inherit_legacy must be a boolean");
		c.inherit_legacy = v->boolean;
	}
	const Json* conditions = boundaries->Find("conditions");
	if (!conditions) throw std::runtime_error("case_config.json: boundaries.conditions is required");
	RequireArray(*conditions, cfg + "boundaries.conditions");
	std::set<int> labels;
	for (std::size_t i = 0; i < conditions->array.size(); ++i) {
		const Json& o = conditions->array[i];
		const std::string ctx = "boundaries.conditions[" + std::to_string(i) + "]";
		RequireObject(o, cfg + ctx);
		RequireKnownKeys(o, {"label", "name", "type", "velocity", "velocity_scale", "pressure", "transport"}, cfg + ctx);
		const Json* label = o.Find("label");
		const Json* type = o.Find("type");
		if (!label || !type) throw std::runtime_error("case_config.json: " + ctx + " requires label and type");
		BoundaryRule r;
		r.label = RequireInteger(*label, cfg + ctx + ".label");
		const std::string& t = RequireString(*type, cfg + ctx + ".type");
		if (t == "interior") r.type = RuleType::Interior;
		else if (t == "wall") r.type = RuleType::Wall;
		else if (t == "inlet") r.type = RuleType::Inlet;
		else if (t == "outlet") r.type = RuleType::Outlet;
		else throw std::runtime_error("case_config.json: unsupported boundary type '" + t + "'");
		if (const Json* v = o.Find("velocity_scale")) {
			r.has_velocity_scale = true;
			r.velocity_scale = RequireNumber(*v, cfg + ctx + ".velocity_scale");
		}
		if (const Json* v = o.Find("velocity")) {
			RequireArray(*v, cfg + ctx + ".velocity");
			if (v->array.size() != 3) throw std::runtime_error("case_config.json: " + ctx + ".velocity must have three entries");
			r.has_velocity = true;
			for (int d = 0; d < 3; ++d)
				r.velocity[d] = RequireNumber(v->array[d], cfg + ctx + ".velocity[" + std::to_string(d) + "]");
		}
		if (const Json* v = o.Find("pressure")) {
			r.has_pressure = true;
			r.pressure = RequireNumber(*v, cfg + ctx + ".pressure");
		}
		if (const Json* v = o.Find("transport")) {
			RequireObject(*v, cfg + ctx + ".transport");
			RequireKnownKeys(*v, {"N0", "Nplus"}, cfg + ctx + ".transport");
			const Json* n0 = v->Find("N0");
			const Json* np = v->Find("Nplus");
			if (!n0 || !np) throw std::runtime_error("case_config.json: " + ctx + ".transport requires N0 and Nplus");
			r.has_transport = true;
			r.n0 = RequireNumber(*n0, cfg + ctx + ".transport.N0");
			r.nplus = RequireNumber(*np, cfg + ctx + ".transport.Nplus");
		}
		if (r.has_velocity && r.has_velocity_scale)
			throw std::runtime_error("case_config.json: " + ctx + " cannot set both velocity and velocity_scale");
		if (r.type == RuleType::Interior && (r.has_velocity || r.has_velocity_scale || r.has_pressure || r.has_transport))
			throw std::runtime_error("case_config.json: interior boundary rules cannot set values");
		if (r.type == RuleType::Wall && (r.has_velocity_scale || r.has_pressure || r.has_transport))
			throw std::runtime_error("case_config.json: wall rules only support an optional velocity");
		if (r.type == RuleType::Inlet && r.has_pressure)
			throw std::runtime_error("case_config.json: inlet rules cannot set pressure");
		if (r.type == RuleType::Outlet && (r.has_velocity || r.has_velocity_scale || r.has_transport))
			throw std::runtime_error("case_config.json: outlet rules only support an optional pressure");
		if (!labels.insert(r.label).second)
			throw std::runtime_error("case_config.json: duplicate boundary label " + std::to_string(r.label));
		c.rules.push_back(r);
	}
	return c;
}

// ResolveBoundaryConditions (solvers/cpu/include/CaseInput.hpp:162-259): the
// transport constraints, and the velocity / pressure Dirichlet data of the
// legacy flow input (velocity: wall zero or the rule's velocity, inlet the
// reference profile times vplus / the rule's scale or velocity; pressure:
// outlets zero or the rule's pressure).
struct TransportBoundaries {
	std::vector<std::int32_t> constrained;
	std::vector<double> n0, nplus;
	std::vector<std::int32_t> velocity_mask, pressure_mask;
	std::vector<double> velocity, pressure;   // [3N], [N]
	std::size_t velocity_nodes = 0, pressure_nodes = 0, transport_nodes = 0;
};

inline TransportBoundaries ResolveTransportBoundaries(const CaseRules& c, const std::vector<int>& labels,
	const TransportParameters& p, const std::vector<std::array<double, 3>>* profile = nullptr)
{
	const std::size_t n = labels.size();
	if (profile && profile->size() != n)
		throw std::runtime_error("boundary labels and velocity field have different node counts");
	TransportBoundaries r;
	std::vector<std::int32_t>& vel = r.velocity_mask;
	std::vector<std::int32_t>& pres = r.pressure_mask;
	vel.assign(n, 0);
	pres.assign(n, 0);
	r.velocity.assign(3*n, 0.0);
	r.pressure.assign(n, 0.0);
	r.constrained.assign(n, 0);
	r.n0.assign(n, 0.0);
	r.nplus.assign(n, 0.0);
	auto set_profile = [&](std::size_t i, double scale) {
		if (profile)
			for (int d = 0; d < 3; ++d) r.velocity[3*i + d] = scale*(*profile)[i][d];
	};
	if (!c.present || c.inherit_legacy) {
		for (std::size_t i = 0; i < n; ++i) {
			if (labels[i] == 0) {
				vel[i] = 1;
			} else if (labels[i] == 1) {
				vel[i] = 1;
				set_profile(i, p.vplus);
				r.constrained[i] = 1;
				r.n0[i] = p.n0_bc;
				r.nplus[i] = p.nplus_bc;
			} else if (labels[i] >= 2) {
				pres[i] = 1;
			}
		}
	}
	std::set<int> configured;
	for (const auto& rule : c.rules) {
		configured.insert(rule.label);
		bool found = false;
		for (std::size_t i = 0; i < n; ++i) {
			if (labels[i] != rule.label) continue;
			found = true;
			vel[i] = pres[i] = r.constrained[i] = 0;
			r.n0[i] = r.nplus[i] = 0.0;
			for (int d = 0; d < 3; ++d) r.velocity[3*i + d] = 0.0;
			r.pressure[i] = 0.0;
			if (rule.type == RuleType::Wall) {
				vel[i] = 1;
				if (rule.has_velocity)
					for (int d = 0; d < 3; ++d) r.velocity[3*i + d] = rule.velocity[d];
			} else if (rule.type == RuleType::Inlet) {
				vel[i] = 1;
				if (rule.has_velocity) {
					for (int d = 0; d < 3; ++d) r.velocity[3*i + d] = rule.velocity[d];
				} else {
					set_profile(i, rule.has_velocity_scale ? rule.velocity_scale : p.vplus);
				}
				r.constrained[i] = 1;
				r.n0[i] = rule.has_transport ? rule.n0 : p.n0_bc;
				r.nplus[i] = rule.has_transport ? rule.nplus : p.nplus_bc;
			} else if (rule.type == RuleType::Outlet) {
				pres[i] = 1;
				r.pressure[i] = rule.has_pressure ? rule.pressure : 0.0;
			}
		}
		if (!found)
			throw std::runtime_error("case_config.json boundary label " + std::to_string(rule.label)
				+ " is not present in controlmesh.vtk");
	}
	if (c.present && !c.inherit_legacy)
		for (int label : labels)
			if (label >= 0 && !configured.count(label))
				throw std::runtime_error("mesh boundary label " + std::to_string(label)
					+ " has no case_config.json rule while inherit_legacy is false");
	r.velocity_nodes = static_cast<std::size_t>(std::count(vel.begin(), vel.end(), 1));
	r.pressure_nodes = static_cast<std::size_t>(std::count(pres.begin(), pres.end(), 1));
	r.transport_nodes = static_cast<std::size_t>(std::count(r.constrained.begin(), r.constrained.end(), 1));
	if (r.velocity_nodes == 0) throw std::runtime_error("case has no velocity boundary nodes");
	if (r.pressure_nodes == 0) throw std::runtime_error("case has no pressure boundary nodes");
	return r;
}

// WriteTransport / WriteTransportVtk (solvers/cuda/src/iga_cuda.cu:327-349):
// "node N0 Nplus" per line, and the control mesh with N0 and Nplus scalars.
inline void WriteTransportText(const std::filesystem::path& path, const std::vector<double>& values)
{
	std::ofstream out(path);
	if (!out) throw std::runtime_error("cannot create " + path.string());
	out << std::setprecision(17);
	for (std::size_t node = 0; node < values.size()/2; ++node)
		out << node << ' ' << values[2*node] << ' ' << values[2*node + 1] << '\n';
}

inline void WriteTransportVtk(const std::filesystem::path& mesh_path, const std::filesystem::path& path,
	const std::vector<double>& values)
{
	std::ifstream mesh(mesh_path);
	std::ofstream out(path);
	if (!mesh) throw std::runtime_error("cannot open VTK mesh " + mesh_path.string());
	if (!out) throw std::runtime_error("cannot create " + path.string());
	out << mesh.rdbuf() << "\nSCALARS N0 double 1\nLOOKUP_TABLE default\n";
	out << std::setprecision(17);
	for (std::size_t node = 0; node < values.size()/2; ++node) out << values[2*node] << '\n';
	out << "SCALARS Nplus double 1\nLOOKUP_TABLE default\n";
	for (std::size_t node = 0; node < values.size()/2; ++node) out << values[2*node + 1] << '\n';
}

}  // namespace tfi_host
