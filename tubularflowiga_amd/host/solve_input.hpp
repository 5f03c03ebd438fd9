// Inputs of the configured linear transport (`iga_cuda solve`,
// solvers/cuda/include/ConfiguredTransportHost.inc:474-503): the
// linear_transport systems of simulation_config.json compiled to weak-form
// terms (CompileLinearSystem, include/SimulationConfig.hpp:743-785), the SUPG
// settings, the scalar boundaries (ResolveScalarBoundaries,
// solvers/cpu/include/GenericCaseInput.hpp:63-141), the flux / Robin faces
// (ConfiguredTransportHost.inc:204-255), text / VTU output and the
// raw_float64 transport checkpoint (include/TransportCheckpoint.hpp).
#pragma once

#include "case_input.hpp"
#include "tfi.h"

#include <functional>
#include <algorithm>
#include <array>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>

namespace tfi_host {

struct ScalarFieldDef {
	std::string name, kind = "scalar";
	double initial_value = 0.0;
};

struct WeakTermDef {
	std::string op, equation, trial, velocity;
	double coefficient = 1.0;
};

struct StabilizationDef {
	std::string equation, method, velocity, diffusion_weighting = "consistent", tau = "advective_diffusive";
};

struct LinearSystemDef {
	std::string name;
	std::vector<std::string> unknowns;
	std::vector<WeakTermDef> terms;
	std::vector<StabilizationDef> stabilization;
	std::string time_integration = "backward_euler";
};

struct ScalarConditionDef {
	std::string field, type, dirichlet_dofs = "face_support", waveform, profile;
	std::vector<double> value;
	double coefficient = 0.0, exterior_value = 0.0;
};

struct ScalarBoundaryDef {
	int label = -1;
	std::string name;
	std::vector<ScalarConditionDef> conditions;
};

struct VelocitySourceDef {
	std::string name, kind, manifest, interpolation, out_of_range = "error";
};

// PhysiologyDefinition (include/SimulationConfig.hpp:124-141), the parts a
// linear_transport run uses: metabolism sources and derived output fields
struct PhysiologyDef {
	bool enabled = false, oxygen_capacity = false, total_oxygen = false;
	std::map<std::string, double> metabolism;  // ordered by field name
	double hematocrit_percent = 0.0, oxygen_solubility = 0.0031, hemoglobin_g_dl = 15.0, p50_mmhg = 26.8,
		hill_exponent = 2.7, gas_molar_volume_ml_mmol = 22.4;
	std::vector<std::string> derived_fields;
};

struct LinearConfig {
	PhysiologyDef physiology;
	std::vector<VelocitySourceDef> sources;
	std::vector<ScalarFieldDef> fields;
	std::vector<LinearSystemDef> systems;   // linear_transport systems in file order
	std::vector<ScalarBoundaryDef> boundaries;
	std::set<std::string> velocity_sources;
	std::map<std::string, std::string> velocity_source_kind;
	double dt = 0.0;
	int steps = 0;
};

// The physiology section (include/SimulationConfig.hpp:594-697) with its checks
// against the transported species of the linear_transport systems
inline void ParsePhysiology(const Json& value, const std::set<std::string>& fields, LinearConfig& c)
{
	const Json& o = RequireObject(value, "physiology");
	RequireKnownKeys(o, {"enabled", "metabolism", "oxygen_capacity", "vasodilation", "derived_fields"}, "physiology");
	auto boolean = [](const Json& v, const std::string& ctx) {
		if (v.kind != Json::Kind::Bool) throw std::runtime_error(ctx + " must be a boolean");
		return v.boolean;
	};
	PhysiologyDef& p = c.physiology;
	p.enabled = o.Find("enabled") ? boolean(*o.Find("enabled"), "physiology.enabled") : true;
	if (const Json* m = o.Find("metabolism")) {
		const Json& rates = RequireObject(*m, "physiology.metabolism");
		for (const auto& kv : rates.object) {
			if (!fields.count(kv.first))
				throw std::runtime_error("simulation_config.json: metabolism references missing field '" + kv.first + "'");
			const double v = RequireNumber(kv.second, "physiology.metabolism." + kv.first);
			if (!std::isfinite(v)) throw std::runtime_error("simulation_config.json: metabolism rates must be finite");
			p.metabolism.emplace(kv.first, v);
		}
	}
	if (const Json* ox = o.Find("oxygen_capacity")) {
		const Json& it = RequireObject(*ox, "physiology.oxygen_capacity");
		RequireKnownKeys(it, {"enabled", "oxygen_state", "hematocrit_percent", "oxygen_solubility", "hemoglobin_g_dl",
			"p50_mmhg", "hill_exponent", "gas_molar_volume_ml_mmol"}, "physiology.oxygen_capacity");
		auto number = [&](const char* name, double fallback) {
			const Json* v = it.Find(name);
			return v ? RequireNumber(*v, std::string("physiology.oxygen_capacity.") + name) : fallback;
		};
		p.oxygen_capacity = it.Find("enabled") ? boolean(*it.Find("enabled"), "physiology.oxygen_capacity.enabled") : true;
		if (const Json* st = it.Find("oxygen_state")) {
			const std::string state = RequireString(*st, "physiology.oxygen_capacity.oxygen_state");
			if (state != "dissolved_oxygen" && state != "total_oxygen")
				throw std::runtime_error("unsupported oxygen transport state '" + state + "'");
			p.total_oxygen = state == "total_oxygen";
		}
		p.hematocrit_percent = number("hematocrit_percent", 0.0);
		p.oxygen_solubility = number("oxygen_solubility", 0.0031);
		p.hemoglobin_g_dl = number("hemoglobin_g_dl", 15.0);
		p.p50_mmhg = number("p50_mmhg", 26.8);
		p.hill_exponent = number("hill_exponent", 2.7);
		p.gas_molar_volume_ml_mmol = number("gas_molar_volume_ml_mmol", 22.4);
		if (p.hematocrit_percent < 0.0 || p.hematocrit_percent > 100.0 || !(p.oxygen_solubility > 0.0)
			|| p.hemoglobin_g_dl < 0.0 || !(p.p50_mmhg > 0.0) || !(p.hill_exponent > 0.0)
			|| !(p.gas_molar_volume_ml_mmol > 0.0))
			throw std::runtime_error("simulation_config.json: oxygen-capacity parameters are invalid");
	}
	if (const Json* vd = o.Find("vasodilation")) {
		const Json& it = RequireObject(*vd, "physiology.vasodilation");
		RequireKnownKeys(it, {"enabled", "field", "emax_radius_fraction", "ec50", "relaxation_tau"}, "physiology.vasodilation");
		if (it.Find("field")) RequireString(*it.Find("field"), "physiology.vasodilation.field");
		if (it.Find("enabled") ? boolean(*it.Find("enabled"), "physiology.vasodilation.enabled") : true)
			throw std::runtime_error("simulation_config.json: 3d vasodilation requires FSI and is not supported; disable it "
				"or use the 1d solver");
	}
	if (const Json* df = o.Find("derived_fields")) {
		static const std::set<std::string> allowed{"pO2", "pCO2", "pH", "SaO2", "SvO2", "dissolved_oxygen",
			"bound_oxygen", "total_oxygen", "hematocrit"};
		std::set<std::string> unique;
		const auto& arr = RequireArray(*df, "physiology.derived_fields").array;
		for (std::size_t i = 0; i < arr.size(); ++i) {
			const std::string name = RequireString(arr[i], "physiology.derived_fields[" + std::to_string(i) + "]");
			if (!allowed.count(name) || !unique.insert(name).second)
				throw std::runtime_error("simulation_config.json: unsupported or duplicate physiology derived field '" + name + "'");
			p.derived_fields.push_back(name);
		}
	}
	std::set<std::string> transported;
	for (const auto& s : c.systems) transported.insert(s.unknowns.begin(), s.unknowns.end());
	for (const auto& r : p.metabolism)
		if (!transported.count(r.first))
			throw std::runtime_error("simulation_config.json: metabolism field '" + r.first + "' is not a transported species");
	const std::string oxygen = p.total_oxygen ? "total_oxygen" : "oxygen";
	if (p.total_oxygen && !p.oxygen_capacity)
		throw std::runtime_error("simulation_config.json: total_oxygen transport requires oxygen_capacity");
	if (transported.count("oxygen") && transported.count("total_oxygen"))
		throw std::runtime_error("simulation_config.json: oxygen and total_oxygen cannot both be transported states");
	for (const auto& name : p.derived_fields) {
		if ((name == "pO2" || name == "SaO2" || name == "SvO2" || name == "dissolved_oxygen" || name == "bound_oxygen"
			|| name == "total_oxygen") && !transported.count(oxygen))
			throw std::runtime_error("simulation_config.json: derived field '" + name + "' requires transported " + oxygen);
		if (name == "pCO2" && !transported.count("carbon_dioxide"))
			throw std::runtime_error("simulation_config.json: pCO2 requires transported carbon_dioxide");
		if (name == "pH" && (!transported.count("carbon_dioxide") || !transported.count("bicarbonate")))
			throw std::runtime_error("simulation_config.json: pH requires transported carbon_dioxide and bicarbonate");
		if ((name == "bound_oxygen" || name == "total_oxygen" || name == "hematocrit") && !p.oxygen_capacity)
			throw std::runtime_error("simulation_config.json: derived field '" + name + "' requires oxygen_capacity");
	}
}

// Oxygen equilibrium (include/OxygenCapacity.hpp:57-138): Hill saturation
// S = p^n / (p^n + p50^n), dissolved = alpha p 10 / Vm, bound = 1.34 Hb S 10 / Vm
// (mol/m^3), pO2 from the transported dissolved or total oxygen (the total
// inverted by bisection on [0, 5000 mmHg], the upper end doubled until it
// brackets, to 1e-10 relative, at most 100 halvings)
struct OxygenState {
	double po2 = 0.0, saturation = 0.0, dissolved = 0.0, bound = 0.0, total = 0.0;
};

inline OxygenState OxygenFromPo2(double po2, const PhysiologyDef& p)
{
	OxygenState r;
	r.po2 = std::max(po2, 0.0);
	const double num = std::pow(r.po2, p.hill_exponent);
	r.saturation = num > 0.0 ? num/(num + std::pow(p.p50_mmhg, p.hill_exponent)) : 0.0;
	r.dissolved = p.oxygen_solubility*r.po2*10.0/p.gas_molar_volume_ml_mmol;
	r.bound = 1.34*p.hemoglobin_g_dl*r.saturation*10.0/p.gas_molar_volume_ml_mmol;
	r.total = r.dissolved + r.bound;
	return r;
}

inline OxygenState OxygenFromTransported(double concentration, const PhysiologyDef& p)
{
	if (!std::isfinite(concentration) || concentration < 0.0)
		throw std::runtime_error(std::string(p.total_oxygen ? "total" : "dissolved")
			+ " oxygen must be finite and nonnegative");
	if (!p.total_oxygen) return OxygenFromPo2(concentration*p.gas_molar_volume_ml_mmol/(10.0*p.oxygen_solubility), p);
	if (concentration == 0.0) return OxygenFromPo2(0.0, p);
	double lo = 0.0, hi = 5000.0;
	while (OxygenFromPo2(hi, p).total < concentration && hi < 1.0e7) hi *= 2.0;
	if (OxygenFromPo2(hi, p).total < concentration)
		throw std::runtime_error("total oxygen exceeds the equilibrium inversion bound");
	for (int it = 0; it < 100; ++it) {
		const double mid = 0.5*(lo + hi);
		if (OxygenFromPo2(mid, p).total < concentration) lo = mid;
		else hi = mid;
		if (hi - lo <= 1.0e-10*std::max(1.0, hi)) break;
	}
	return OxygenFromPo2(0.5*(lo + hi), p);
}

// ComputePhysiologyPointArrays (include/PhysiologyOutput.hpp:29-91): one
// point array per derived field from the interleaved state of `fields`
inline std::vector<PointArray> PhysiologyArrays(const PhysiologyDef& p, const std::vector<std::string>& fields,
	const std::vector<double>& state)
{
	std::vector<PointArray> out;
	if (!p.enabled || fields.empty()) return out;
	const std::size_t F = fields.size(), nodes = state.size()/F;
	auto column = [&](const std::string& name) {
		const auto it = std::find(fields.begin(), fields.end(), name);
		if (it == fields.end()) throw std::runtime_error("physiology output needs field '" + name + "'");
		const std::size_t f = static_cast<std::size_t>(it - fields.begin());
		std::vector<double> v(nodes);
		for (std::size_t n = 0; n < nodes; ++n) v[n] = state[n*F + f];
		return v;
	};
	for (const auto& name : p.derived_fields) {
		PointArray a{name, 1, std::vector<double>(nodes, 0.0)};
		if (name == "hematocrit") {
			std::fill(a.values.begin(), a.values.end(), p.hematocrit_percent);
		} else if (name == "pO2" || name == "SaO2" || name == "SvO2" || name == "dissolved_oxygen"
			|| name == "bound_oxygen" || name == "total_oxygen") {
			const auto oxygen = column(p.total_oxygen ? "total_oxygen" : "oxygen");
			for (std::size_t n = 0; n < nodes; ++n) {
				const OxygenState e = OxygenFromTransported(oxygen[n], p);
				a.values[n] = name == "pO2" ? e.po2 : (name == "SaO2" || name == "SvO2") ? e.saturation
					: name == "dissolved_oxygen" ? e.dissolved : name == "bound_oxygen" ? e.bound : e.total;
			}
		} else if (name == "pCO2") {
			const auto co2 = column("carbon_dioxide");
			for (std::size_t n = 0; n < nodes; ++n) a.values[n] = co2[n]/0.0301;
		} else if (name == "pH") {
			const auto co2 = column("carbon_dioxide"), hco3 = column("bicarbonate");
			for (std::size_t n = 0; n < nodes; ++n) {
				const double pco2 = std::max(co2[n]/0.0301, 1.0e-30);
				a.values[n] = 6.1 + std::log10(std::max(hco3[n], 1.0e-30)/(0.0301*pco2));
			}
		}
		out.push_back(std::move(a));
	}
	return out;
}

// WritePhysiologyManifest (include/PhysiologyOutput.hpp:93-117)
inline void WritePhysiologyManifest(const fs::path& path, const PhysiologyDef& p, const std::vector<std::string>& fields)
{
	std::ofstream out(path);
	if (!out) throw std::runtime_error("cannot create physiology manifest: " + path.string());
	out << "{\n  \"enabled\": " << (p.enabled ? "true" : "false") << ",\n  \"fields\": {";
	bool first = true;
	for (const auto& f : fields) {
		out << (first ? "\n" : ",\n") << "    \"" << JsonEscape(f) << "\": {\"status\": \"solved\"}";
		first = false;
	}
	for (const auto& f : p.derived_fields) {
		out << (first ? "\n" : ",\n") << "    \"" << JsonEscape(f) << "\": {\"status\": \"derived\"}";
		first = false;
	}
	if (!first) out << '\n';
	out << "  }\n}\n";
	out.close();
	if (!out) throw std::runtime_error("cannot write physiology manifest: " + path.string());
}

// The parts of ParseSimulationConfiguration (include/SimulationConfig.hpp:229-717)
// the linear_transport path reads, with its checks on them.
inline LinearConfig ReadLinearConfig(const std::string& path)
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
	LinearConfig c;
	const int schema = RequireInteger(Required(root, "schema_version", "root"), "schema_version");
	if (schema < 2 || schema > 4)
		throw std::runtime_error("simulation_config.json: supported schema versions are 2, 3, and 4");
	std::set<std::string> names;
	const auto& farr = RequireArray(Required(root, "fields", "root"), "fields").array;
	for (std::size_t i = 0; i < farr.size(); ++i) {
		const std::string ctx = "fields[" + std::to_string(i) + "]";
		const Json& f = RequireObject(farr[i], ctx);
		RequireKnownKeys(f, {"name", "kind", "initial_value"}, ctx);
		ScalarFieldDef d;
		d.name = RequireString(Required(f, "name", ctx), ctx + ".name");
		d.kind = RequireString(Required(f, "kind", ctx), ctx + ".kind");
		if (d.kind != "scalar" && d.kind != "vector3" && d.kind != "pressure")
			throw std::runtime_error("simulation_config.json: unsupported field kind '" + d.kind + "'");
		if (const Json* v = f.Find("initial_value")) d.initial_value = RequireNumber(*v, ctx + ".initial_value");
		if (d.name.empty() || !names.insert(d.name).second)
			throw std::runtime_error("simulation_config.json: field names must be non-empty and unique");
		c.fields.push_back(d);
	}
	if (c.fields.empty()) throw std::runtime_error("simulation_config.json: fields cannot be empty");
	const Json& time = RequireObject(Required(root, "time", "root"), "time");
	RequireKnownKeys(time, {"dt", "steps"}, "time");
	c.dt = RequireNumber(Required(time, "dt", "time"), "time.dt");
	c.steps = RequireInteger(Required(time, "steps", "time"), "time.steps");
	if (!(c.dt > 0.0)) throw std::runtime_error("simulation_config.json: time.dt must be positive");
	if (const Json* vs = root.Find("velocity_sources")) {
		const auto& arr = RequireArray(*vs, "velocity_sources").array;
		for (std::size_t i = 0; i < arr.size(); ++i) {
			const std::string ctx = "velocity_sources[" + std::to_string(i) + "]";
			const Json& o = RequireObject(arr[i], ctx);
			RequireKnownKeys(o, {"name", "kind", "manifest", "interpolation", "out_of_range"}, ctx);
			VelocitySourceDef d;
			d.name = RequireString(Required(o, "name", ctx), ctx + ".name");
			d.kind = RequireString(Required(o, "kind", ctx), ctx + ".kind");
			d.manifest = RequireString(Required(o, "manifest", ctx), ctx + ".manifest");
			d.interpolation = RequireString(Required(o, "interpolation", ctx), ctx + ".interpolation");
			if (const Json* v = o.Find("out_of_range")) d.out_of_range = RequireString(*v, ctx + ".out_of_range");
			if (d.name.empty() || d.name == "prescribed" || !c.velocity_sources.insert(d.name).second)
				throw std::runtime_error("simulation_config.json: velocity source names must be non-empty, unique, and not 'prescribed'");
			if (d.kind != "snapshot_series" || d.manifest.empty() || d.interpolation != "linear"
				|| (d.out_of_range != "error" && d.out_of_range != "hold"))
				throw std::runtime_error("simulation_config.json: snapshot velocity source requires manifest, linear "
					"interpolation, and error or hold out_of_range");
			if (fs::path(d.manifest).is_absolute())
				throw std::runtime_error("simulation_config.json: velocity source manifest must be relative to the case directory");
			c.velocity_source_kind[d.name] = d.kind;
			c.sources.push_back(d);
		}
	}
	for (const char* key : {"perfusate", "coupling", "external_circuit"})
		if (root.Find(key))
			throw std::runtime_error(std::string("simulation_config.json: '") + key
				+ "' (multidomain coupling) is not supported by iga_hip solve");
	const Json* physiology = root.Find("physiology");
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
		if (kind == "navier_stokes") continue;
		if (kind != "linear_transport")
			throw std::runtime_error("simulation_config.json: unsupported equation-system kind '" + kind + "'");
		LinearSystemDef s;
		s.name = name;
		std::set<std::string> local;
		const auto& unk = RequireArray(Required(o, "unknowns", ctx), ctx + ".unknowns").array;
		for (std::size_t j = 0; j < unk.size(); ++j) {
			const std::string f = RequireString(unk[j], ctx + ".unknowns[" + std::to_string(j) + "]");
			if (!names.count(f)) throw std::runtime_error("simulation_config.json: unknown field '" + f + "'");
			if (!local.insert(f).second) throw std::runtime_error("simulation_config.json: duplicate system unknown");
			s.unknowns.push_back(f);
		}
		if (s.unknowns.empty()) throw std::runtime_error("simulation_config.json: system unknowns cannot be empty");
		if (o.Find("density")) throw std::runtime_error("simulation_config.json: density applies only to navier_stokes");
		if (const Json* v = o.Find("time_integration")) s.time_integration = RequireString(*v, ctx + ".time_integration");
		if (s.time_integration != "backward_euler" && s.time_integration != "bdf2")
			throw std::runtime_error("simulation_config.json: linear_transport time_integration must be 'backward_euler' or 'bdf2'");
		if (const Json* terms = o.Find("terms")) {
			const auto& arr = RequireArray(*terms, ctx + ".terms").array;
			for (std::size_t j = 0; j < arr.size(); ++j) {
				const std::string tctx = ctx + ".terms[" + std::to_string(j) + "]";
				const Json& t = RequireObject(arr[j], tctx);
				RequireKnownKeys(t, {"operator", "equation", "trial", "coefficient", "velocity"}, tctx);
				WeakTermDef w;
				w.op = RequireString(Required(t, "operator", tctx), tctx + ".operator");
				static const std::set<std::string> ops = {"time_derivative", "diffusion", "advection", "linear_coupling",
					"volume_source"};
				if (!ops.count(w.op)) throw std::runtime_error("simulation_config.json: unknown term operator '" + w.op + "'");
				w.equation = RequireString(Required(t, "equation", tctx), tctx + ".equation");
				w.trial = w.equation;
				if (const Json* v = t.Find("trial")) w.trial = RequireString(*v, tctx + ".trial");
				if (const Json* v = t.Find("coefficient")) w.coefficient = RequireNumber(*v, tctx + ".coefficient");
				if (const Json* v = t.Find("velocity")) w.velocity = RequireString(*v, tctx + ".velocity");
				if (!local.count(w.equation) || !local.count(w.trial))
					throw std::runtime_error("simulation_config.json: term equation and trial must be unknowns of their system");
				if (w.op == "advection" && w.velocity.empty())
					throw std::runtime_error("simulation_config.json: advection requires a velocity source");
				if (w.op == "advection" && w.velocity != "prescribed" && !c.velocity_sources.count(w.velocity))
					throw std::runtime_error("simulation_config.json: advection references an unknown velocity source");
				s.terms.push_back(w);
			}
		}
		if (s.terms.empty()) throw std::runtime_error("simulation_config.json: linear_transport requires terms");
		if (const Json* st = o.Find("stabilization")) {
			const auto& arr = RequireArray(*st, ctx + ".stabilization").array;
			for (std::size_t j = 0; j < arr.size(); ++j) {
				const std::string ictx = ctx + ".stabilization[" + std::to_string(j) + "]";
				const Json& it = RequireObject(arr[j], ictx);
				RequireKnownKeys(it, {"equation", "method", "velocity", "diffusion_weighting", "tau"}, ictx);
				StabilizationDef d;
				d.equation = RequireString(Required(it, "equation", ictx), ictx + ".equation");
				d.method = RequireString(Required(it, "method", ictx), ictx + ".method");
				d.velocity = RequireString(Required(it, "velocity", ictx), ictx + ".velocity");
				if (const Json* v = it.Find("diffusion_weighting"))
					d.diffusion_weighting = RequireString(*v, ictx + ".diffusion_weighting");
				if (d.diffusion_weighting != "consistent" && d.diffusion_weighting != "galerkin"
					&& d.diffusion_weighting != "supg_test")
					throw std::runtime_error("simulation_config.json: stabilization diffusion_weighting must be 'consistent', 'galerkin' or 'supg_test'");
				if (const Json* v = it.Find("tau")) d.tau = RequireString(*v, ictx + ".tau");
				if (d.tau != "advective_diffusive" && d.tau != "advective")
					throw std::runtime_error("simulation_config.json: stabilization tau must be 'advective_diffusive' or 'advective'");
				if (!local.count(d.equation) || d.method != "supg")
					throw std::runtime_error("simulation_config.json: stabilization requires a system equation and supported method 'supg'");
				if (d.velocity != "prescribed" && !c.velocity_sources.count(d.velocity))
					throw std::runtime_error("simulation_config.json: stabilization references an unknown velocity source");
				s.stabilization.push_back(d);
			}
		}
		c.systems.push_back(s);
	}
	if (sarr.empty()) throw std::runtime_error("simulation_config.json: equation_systems cannot be empty");
	if (physiology) ParsePhysiology(*physiology, names, c);
	const auto& barr = RequireArray(Required(root, "boundaries", "root"), "boundaries").array;
	std::set<int> labels;
	for (std::size_t i = 0; i < barr.size(); ++i) {
		const std::string ctx = "boundaries[" + std::to_string(i) + "]";
		const Json& o = RequireObject(barr[i], ctx);
		RequireKnownKeys(o, {"label", "name", "conditions"}, ctx);
		ScalarBoundaryDef b;
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
			ScalarConditionDef d;
			d.field = RequireString(Required(co, "field", cctx), cctx + ".field");
			d.type = RequireString(Required(co, "type", cctx), cctx + ".type");
			static const std::set<std::string> kinds = {"dirichlet", "no_flux", "flux", "robin", "advective_outflow",
				"pressure_traction", "resistance", "windkessel_rc", "windkessel_rcr"};
			if (!kinds.count(d.type)) throw std::runtime_error("simulation_config.json: unknown boundary type '" + d.type + "'");
			if (!names.count(d.field) || !conditioned.insert(d.field).second)
				throw std::runtime_error("simulation_config.json: boundary fields must exist and be unique per boundary");
			if (const Json* v = co.Find("value")) d.value = ValueVector(*v, cctx + ".value");
			if (const Json* v = co.Find("coefficient")) d.coefficient = RequireNumber(*v, cctx + ".coefficient");
			if (const Json* v = co.Find("exterior_value")) d.exterior_value = RequireNumber(*v, cctx + ".exterior_value");
			if (const Json* v = co.Find("dirichlet_dofs")) d.dirichlet_dofs = RequireString(*v, cctx + ".dirichlet_dofs");
			if (d.dirichlet_dofs != "face_support" && d.dirichlet_dofs != "node_label")
				throw std::runtime_error("simulation_config.json: dirichlet_dofs must be 'face_support' or 'node_label'");
			if (const Json* v = co.Find("profile")) d.profile = RequireString(*v, cctx + ".profile");
			if (const Json* v = co.Find("waveform")) d.waveform = RequireString(*v, cctx + ".waveform");
			if (d.type == "dirichlet" && d.value.empty() && d.profile.empty())
				throw std::runtime_error("simulation_config.json: dirichlet requires value or profile");
			b.conditions.push_back(d);
		}
		c.boundaries.push_back(b);
	}
	return c;
}

// CompiledLinearSystem (include/SimulationConfig.hpp:167-185, 743-785) in the
// form the C-ABI takes
struct CompiledSystem {
	std::string name, velocity_source = "prescribed", time_integration;
	std::vector<std::string> fields;
	std::map<std::string, int> index;
	std::vector<tfi_solve_term> terms;
	std::vector<std::int32_t> supg;
	std::vector<double> tau_diffusivity, initial;
	double dt = 0.0;
	int steps = 0;
};

inline CompiledSystem CompileSystem(const LinearConfig& c, const std::string& requested)
{
	if (c.systems.empty()) throw std::runtime_error("simulation configuration has no linear_transport equation system");
	const LinearSystemDef* s = requested.empty() ? &c.systems.front() : nullptr;
	for (const auto& it : c.systems)
		if (!s && it.name == requested) s = &it;
	if (!s) throw std::runtime_error("unknown equation system '" + requested + "'");
	CompiledSystem r;
	r.name = s->name;
	r.fields = s->unknowns;
	r.dt = c.dt;
	r.steps = c.steps;
	r.time_integration = s->time_integration;
	if (r.time_integration != "backward_euler")
		throw std::runtime_error("CUDA transport supports backward_euler only; use iga_solve for bdf2");
	std::set<std::string> velocities;
	for (const auto& t : s->terms)
		if (t.op == "advection") velocities.insert(t.velocity);
	for (const auto& d : s->stabilization) velocities.insert(d.velocity);
	if (velocities.size() > 1)
		throw std::runtime_error("linear_transport currently supports one velocity source per equation system");
	if (!velocities.empty()) r.velocity_source = *velocities.begin();
	for (std::size_t i = 0; i < r.fields.size(); ++i) r.index[r.fields[i]] = static_cast<int>(i);
	if (r.fields.size() > 8) throw std::runtime_error("CUDA supports 1 through 8 coupled scalar fields per system");
	static const std::map<std::string, int> kind = {{"time_derivative", TFI_TERM_TIME_DERIVATIVE},
		{"diffusion", TFI_TERM_DIFFUSION}, {"advection", TFI_TERM_ADVECTION},
		{"linear_coupling", TFI_TERM_LINEAR_COUPLING}, {"volume_source", TFI_TERM_VOLUME_SOURCE}};
	for (const auto& t : s->terms) {
		if (t.op == "advection" && t.velocity != r.velocity_source)
			throw std::runtime_error("CUDA transport received inconsistent velocity sources");
		r.terms.push_back({kind.at(t.op), r.index.at(t.equation), r.index.at(t.trial), t.coefficient});
	}
	// metabolism rates as volume sources of their fields (CompileLinearSystem, :778-784)
	if (c.physiology.enabled)
		for (const auto& rate : c.physiology.metabolism) {
			const auto f = r.index.find(rate.first);
			if (f != r.index.end()) r.terms.push_back({TFI_TERM_VOLUME_SOURCE, f->second, f->second, rate.second});
		}
	if (r.terms.size() > 64) throw std::runtime_error("configured transport supports at most 64 weak-form terms");
	const std::size_t F = r.fields.size();
	r.supg.assign(F, TFI_SUPG_NONE);
	r.tau_diffusivity.assign(F, 0.0);
	for (const auto& d : s->stabilization) {
		if (d.method != "supg" || d.velocity != r.velocity_source)
			throw std::runtime_error("CUDA transport requires SUPG and advection to use the same velocity source");
		r.supg[r.index.at(d.equation)] = d.diffusion_weighting == "supg_test" ? TFI_SUPG_TEST
			: d.diffusion_weighting == "consistent" ? TFI_SUPG_CONSISTENT : TFI_SUPG_GALERKIN;
		// SupgTauDiffusivity (SimulationConfig.hpp:728-741)
		if (d.tau != "advective_diffusive") continue;
		const int f = r.index.at(d.equation);
		for (const auto& t : r.terms)
			if (t.kind == TFI_TERM_DIFFUSION && t.equation == f && t.trial == f) r.tau_diffusivity[f] += t.coefficient;
	}
	r.initial.assign(F, 0.0);
	for (std::size_t i = 0; i < F; ++i)
		for (const auto& f : c.fields)
			if (f.name == r.fields[i]) r.initial[i] = f.initial_value;
	return r;
}

// Velocity snapshot series (include/VelocitySeries.hpp:19-117): "time,file"
// lines (an optional "time,file" header, '#' comments), times finite,
// nonnegative and strictly increasing, files relative to the case directory
struct VelocitySnapshot {
	double time = 0.0;
	fs::path file;
};

inline std::string TrimToken(const std::string& t)
{
	const auto a = t.find_first_not_of(" \t\r");
	if (a == std::string::npos) return {};
	return t.substr(a, t.find_last_not_of(" \t\r") - a + 1);
}

inline std::vector<VelocitySnapshot> ReadVelocityManifest(const fs::path& path)
{
	std::ifstream in(path);
	if (!in) throw std::runtime_error("cannot open velocity manifest: " + path.string());
	const std::string ctx = path.string();
	std::vector<VelocitySnapshot> out;
	std::string line;
	std::size_t number = 0;
	while (std::getline(in, line)) {
		++number;
		line = TrimToken(line);
		if (line.empty() || line.front() == '#') continue;
		const auto comma = line.find(',');
		if (comma == std::string::npos || line.find(',', comma + 1) != std::string::npos)
			throw std::runtime_error(ctx + " line " + std::to_string(number) + " must contain exactly time,file");
		const std::string tt = TrimToken(line.substr(0, comma)), ft = TrimToken(line.substr(comma + 1));
		if (out.empty() && tt == "time" && ft == "file") continue;
		std::size_t used = 0;
		double t = 0.0;
		try {
			t = std::stod(tt, &used);
		} catch (const std::exception&) {
			throw std::runtime_error(ctx + " line " + std::to_string(number) + " has an invalid time");
		}
		if (used != tt.size() || !std::isfinite(t) || t < 0.0)
			throw std::runtime_error(ctx + " times must be finite and nonnegative");
		if (!out.empty() && !(t > out.back().time)) throw std::runtime_error(ctx + " times must be strictly increasing");
		const fs::path file(ft);
		if (file.empty() || file.is_absolute())
			throw std::runtime_error(ctx + " files must be non-empty paths relative to the case directory");
		out.push_back({t, file});
	}
	if (out.empty()) throw std::runtime_error(ctx + " contains no snapshots");
	return out;
}

// ResolveVelocityInterpolation (VelocitySeries.hpp:105-125): the bracketing
// snapshots and the upper one's weight; "hold" clamps outside the series
struct VelocityInterpolation {
	std::size_t lower = 0, upper = 0;
	double upper_weight = 0.0;
};

inline VelocityInterpolation ResolveVelocityInterpolation(const std::vector<VelocitySnapshot>& s, double time,
	const std::string& out_of_range)
{
	if (s.empty() || !std::isfinite(time))
		throw std::runtime_error("cannot interpolate an empty velocity series or non-finite time");
	if (time < s.front().time) {
		if (out_of_range == "hold") return {0, 0, 0.0};
		throw std::runtime_error("transport time precedes the first velocity snapshot");
	}
	if (time > s.back().time) {
		if (out_of_range == "hold") return {s.size() - 1, s.size() - 1, 0.0};
		throw std::runtime_error("transport time exceeds the last velocity snapshot");
	}
	std::size_t upper = 0;
	while (upper < s.size() && s[upper].time < time) ++upper;
	if (upper == s.size()) return {s.size() - 1, s.size() - 1, 0.0};
	if (s[upper].time == time || upper == 0) return {upper, upper, 0.0};
	return {upper - 1, upper, (time - s[upper - 1].time)/(s[upper].time - s[upper - 1].time)};
}

// NodeFaceLabels (GenericCaseInput.hpp:38-61): labels of the element faces on
// which each node's basis has a nonzero trace
inline std::vector<std::vector<int>> NodeFaceLabels(const FlatDatabase& db)
{
	static const int fixed_axis[6] = {2, 1, 0, 1, 0, 2};
	static const int fixed_index[6] = {0, 0, 3, 3, 0, 3};
	std::vector<std::vector<int>> out(db.nodes);
	for (int e = 0; e < db.elements(); ++e)
		for (int f = 0; f < 6; ++f) {
			const int label = db.labels[6*static_cast<std::size_t>(e) + f];
			if (label < 0) continue;
			for (int a = db.element_offsets[e]; a < db.element_offsets[e + 1]; ++a)
				for (int k = db.extraction_offsets[a]; k < db.extraction_offsets[a + 1]; ++k) {
					const int p = db.extraction_columns[k];
					const int ijk[3] = {p % 4, (p/4) % 4, p/16};
					if (ijk[fixed_axis[f]] != fixed_index[f]) continue;
					auto& l = out.at(static_cast<std::size_t>(db.connectivity[a]));
					if (std::find(l.begin(), l.end(), label) == l.end()) l.push_back(label);
					break;
				}
		}
	return out;
}

struct ScalarBoundaries {
	std::vector<std::int32_t> constrained;
	std::vector<double> value;
	long long constrained_dofs = 0;
	// flux / Robin faces
	std::vector<std::int32_t> surface_element, surface_face;
	std::vector<double> surface_flux, surface_robin, surface_robin_source;
};

// ResolveScalarBoundaries (GenericCaseInput.hpp:63-141) with the face-support
// Dirichlet entries, then the flux / Robin faces (ConfiguredTransportHost.inc:204-255)
inline ScalarBoundaries ResolveScalarBoundaries(const LinearConfig& c, const CompiledSystem& s,
	const std::vector<int>& labels, const FlatDatabase& db)
{
	for (const auto& name : s.fields) {
		const auto it = std::find_if(c.fields.begin(), c.fields.end(), [&](const ScalarFieldDef& f) { return f.name == name; });
		if (it == c.fields.end() || it->kind != "scalar")
			throw std::runtime_error("linear_transport currently requires scalar fields: " + name);
	}
	const std::size_t F = s.fields.size();
	ScalarBoundaries r;
	r.constrained.assign(labels.size()*F, 0);
	r.value.assign(labels.size()*F, 0.0);
	std::set<int> mesh_labels, configured;
	for (int l : labels)
		if (l >= 0) mesh_labels.insert(l);
	for (const auto& b : c.boundaries) {
		if (!mesh_labels.count(b.label))
			throw std::runtime_error("simulation_config.json boundary label " + std::to_string(b.label)
				+ " is not present in controlmesh.vtk");
		configured.insert(b.label);
		for (const auto& cond : b.conditions) {
			const auto f = s.index.find(cond.field);
			if (f == s.index.end()) continue;
			if (!cond.waveform.empty()) throw std::runtime_error("time-dependent scalar boundary execution is not implemented");
			if (cond.type == "no_flux" || cond.type == "advective_outflow" || cond.type == "flux" || cond.type == "robin")
				continue;
			if (cond.type != "dirichlet")
				throw std::runtime_error("unsupported scalar boundary condition for field '" + cond.field + "'");
			if (cond.value.size() != 1)
				throw std::runtime_error("scalar Dirichlet field '" + cond.field + "' requires one value");
			for (std::size_t n = 0; n < labels.size(); ++n)
				if (labels[n] == b.label) {
					r.constrained[n*F + f->second] = 1;
					r.value[n*F + f->second] = cond.value[0];
				}
		}
	}
	for (int l : mesh_labels)
		if (!configured.count(l))
			throw std::runtime_error("mesh boundary label " + std::to_string(l) + " has no simulation_config.json boundary definition");
	const auto touched = NodeFaceLabels(db);
	std::set<std::pair<int, int>> labelled;
	for (const auto& b : c.boundaries)
		for (const auto& cond : b.conditions) {
			const auto f = s.index.find(cond.field);
			if (f != s.index.end() && cond.type == "dirichlet") labelled.insert({b.label, f->second});
		}
	for (const auto& b : c.boundaries)
		for (const auto& cond : b.conditions) {
			const auto f = s.index.find(cond.field);
			if (f == s.index.end() || cond.type != "dirichlet" || cond.dirichlet_dofs != "face_support") continue;
			for (std::size_t n = 0; n < labels.size(); ++n) {
				const std::size_t dof = n*F + f->second;
				const auto& t = touched[n];
				if (r.constrained[dof] || labelled.count({labels[n], f->second})
					|| std::find(t.begin(), t.end(), b.label) == t.end())
					continue;
				r.constrained[dof] = 1;
				r.value[dof] = cond.value[0];
			}
		}
	r.constrained_dofs = std::count(r.constrained.begin(), r.constrained.end(), 1);
	// flux / Robin faces
	std::vector<int> required, found;
	for (const auto& b : c.boundaries)
		for (const auto& cond : b.conditions)
			if (s.index.count(cond.field) && (cond.type == "flux" || cond.type == "robin")
				&& std::find(required.begin(), required.end(), b.label) == required.end())
				required.push_back(b.label);
	for (int e = 0; e < db.elements(); ++e)
		for (int face = 0; face < 6; ++face) {
			const int label = db.labels[6*static_cast<std::size_t>(e) + face];
			if (label < 0) continue;
			const auto b = std::find_if(c.boundaries.begin(), c.boundaries.end(),
				[&](const ScalarBoundaryDef& x) { return x.label == label; });
			if (b == c.boundaries.end())
				throw std::runtime_error("element boundary face has no simulation_config.json definition");
			std::vector<double> flux(F, 0.0), robin(F, 0.0), source(F, 0.0);
			bool active = false;
			for (const auto& cond : b->conditions) {
				const auto f = s.index.find(cond.field);
				if (f == s.index.end()) continue;
				if (cond.type == "flux") {
					if (cond.value.empty()) throw std::runtime_error("scalar flux field '" + cond.field + "' requires a value");
					flux[f->second] = cond.value[0];
					active = true;
				} else if (cond.type == "robin") {
					robin[f->second] = cond.coefficient;
					source[f->second] = cond.coefficient*cond.exterior_value;
					active = true;
				}
			}
			if (!active) continue;
			if (std::find(found.begin(), found.end(), label) == found.end()) found.push_back(label);
			r.surface_element.push_back(e);
			r.surface_face.push_back(face);
			r.surface_flux.insert(r.surface_flux.end(), flux.begin(), flux.end());
			r.surface_robin.insert(r.surface_robin.end(), robin.begin(), robin.end());
			r.surface_robin_source.insert(r.surface_robin_source.end(), source.begin(), source.end());
		}
	for (int l : required)
		if (std::find(found.begin(), found.end(), l) == found.end())
			throw std::runtime_error("configured scalar surface boundary label " + std::to_string(l)
				+ " has no boundary faces in the .ntiga database; repack with iga_pack");
	return r;
}

// WriteCudaConfiguredOutput (ConfiguredTransportHost.inc:96-140): "node v0 v1 ..."
// lines at 17 digits, the field names in PATH.fields, and a VTU with one
// point array per field
// The configured text output and the point arrays (fields, then the derived
// physiology arrays) as a VTU, or handed to `visual` when it is given (the
// Bezier VTKHDF series).
inline void WriteConfiguredOutput(const fs::path& path, const std::vector<std::string>& fields,
	const std::vector<double>& state, const fs::path& mesh_vtk, const fs::path& vtu, double time,
	const PhysiologyDef& physiology,
	const std::function<void(double, const std::vector<PointArray>&)>& visual = nullptr)
{
	const std::size_t F = fields.size(), nodes = state.size()/F;
	if (!path.parent_path().empty()) fs::create_directories(path.parent_path());
	std::ofstream values(path), meta(path.string() + ".fields");
	if (!values || !meta) throw std::runtime_error("cannot create configured CUDA output");
	values << std::setprecision(17);
	for (const auto& name : fields) meta << name << '\n';
	for (std::size_t n = 0; n < nodes; ++n) {
		values << n;
		for (std::size_t f = 0; f < F; ++f) values << ' ' << state[n*F + f];
		values << '\n';
	}
	if (!values || !meta) throw std::runtime_error("cannot write configured CUDA output");
	std::vector<PointArray> arrays;
	for (std::size_t f = 0; f < F; ++f) {
		PointArray a{fields[f], 1, std::vector<double>(nodes)};
		for (std::size_t n = 0; n < nodes; ++n) a.values[n] = state[n*F + f];
		arrays.push_back(std::move(a));
	}
	for (auto& a : PhysiologyArrays(physiology, fields, state)) arrays.push_back(std::move(a));
	if (visual) visual(time, arrays);
	else WritePointArraysVtu(mesh_vtk, vtu, arrays, time);
}

// raw_float64 transport checkpoint, schema 1 (include/TransportCheckpoint.hpp:18-219)
struct TransportCheckpointMeta {
	std::uint64_t nodes = 0;
	std::vector<std::string> fields;
	std::string system, velocity_source, state_file, state_format = "raw_float64";
	int schema_version = 1, completed_step = 0;
	double physical_time = 0.0, dt = 0.0;
};

inline void WriteTransportCheckpoint(const fs::path& prefix, TransportCheckpointMeta m, const std::vector<double>& state)
{
	const fs::path state_path = prefix.string() + ".state";
	{
		std::ofstream out(state_path, std::ios::binary);
		if (!out) throw std::runtime_error("cannot create CUDA transport checkpoint state: " + state_path.string());
		out.write(reinterpret_cast<const char*>(state.data()), static_cast<std::streamsize>(state.size()*sizeof(double)));
		out.close();
		if (!out) throw std::runtime_error("cannot write CUDA transport checkpoint state: " + state_path.string());
	}
	m.state_file = state_path.filename().string();
	const fs::path meta_path = prefix.string() + ".json";
	std::ofstream out(meta_path);
	if (!out) throw std::runtime_error("cannot create transport checkpoint metadata: " + meta_path.string());
	out << std::setprecision(17) << "{\n  \"schema_version\": 1,\n  \"nodes\": " << m.nodes << ",\n  \"fields\": [";
	for (std::size_t i = 0; i < m.fields.size(); ++i) out << (i ? ", " : "") << '"' << JsonEscape(m.fields[i]) << '"';
	out << "],\n  \"system\": \"" << JsonEscape(m.system) << "\",\n  \"velocity_source\": \"" << JsonEscape(m.velocity_source)
		<< "\",\n  \"completed_step\": " << m.completed_step << ",\n  \"physical_time\": " << m.physical_time
		<< ",\n  \"dt\": " << m.dt << ",\n  \"state_file\": \"" << JsonEscape(m.state_file)
		<< "\",\n  \"state_format\": \"raw_float64\"\n}\n";
	if (!out) throw std::runtime_error("cannot write transport checkpoint metadata: " + meta_path.string());
}

inline TransportCheckpointMeta ReadTransportCheckpoint(const fs::path& prefix, std::vector<double>* state)
{
	const fs::path meta_path = prefix.string() + ".json";
	std::ifstream in(meta_path);
	if (!in) throw std::runtime_error("cannot open transport checkpoint metadata: " + meta_path.string());
	std::stringstream ss;
	ss << in.rdbuf();
	const Json v = ParseJson(ss.str());
	const Json& root = RequireObject(v, "transport checkpoint");
	RequireKnownKeys(root, {"schema_version", "nodes", "fields", "system", "velocity_source", "completed_step",
		"physical_time", "dt", "state_file", "state_format", "time_integration", "history_file"}, "transport checkpoint");
	auto req = [&](const char* key) -> const Json& {
		const Json* x = root.Find(key);
		if (!x) throw std::runtime_error(std::string("transport checkpoint requires '") + key + "'");
		return *x;
	};
	TransportCheckpointMeta m;
	m.schema_version = RequireInteger(req("schema_version"), "transport checkpoint.schema_version");
	const double nodes = RequireNumber(req("nodes"), "transport checkpoint.nodes");
	if (!(nodes > 0.0) || std::floor(nodes) != nodes)
		throw std::runtime_error("transport checkpoint.nodes must be a positive integer");
	m.nodes = static_cast<std::uint64_t>(nodes);
	for (const auto& f : RequireArray(req("fields"), "transport checkpoint.fields").array)
		m.fields.push_back(RequireString(f, "transport checkpoint.fields[]"));
	m.system = RequireString(req("system"), "transport checkpoint.system");
	m.velocity_source = RequireString(req("velocity_source"), "transport checkpoint.velocity_source");
	m.completed_step = RequireInteger(req("completed_step"), "transport checkpoint.completed_step");
	m.physical_time = RequireNumber(req("physical_time"), "transport checkpoint.physical_time");
	m.dt = RequireNumber(req("dt"), "transport checkpoint.dt");
	m.state_file = RequireString(req("state_file"), "transport checkpoint.state_file");
	m.state_format = RequireString(req("state_format"), "transport checkpoint.state_format");
	if (m.schema_version == 2) throw std::runtime_error("CUDA transport restart supports backward_euler checkpoints only");
	if (m.schema_version != 1 || m.fields.empty() || m.system.empty() || m.velocity_source.empty()
		|| m.completed_step < 0 || m.physical_time < 0.0 || !(m.dt > 0.0) || m.state_file.empty()
		|| (m.state_format != "petsc_binary" && m.state_format != "raw_float64"))
		throw std::runtime_error("transport checkpoint metadata contains invalid values");
	if (m.state_format != "raw_float64") throw std::runtime_error("CUDA transport restart requires raw_float64 state");
	fs::path path(m.state_file);
	if (path.is_relative()) path = meta_path.parent_path()/path;
	std::ifstream bin(path, std::ios::binary);
	if (!bin) throw std::runtime_error("cannot open CUDA transport checkpoint state: " + path.string());
	state->assign(m.nodes*m.fields.size(), 0.0);
	bin.read(reinterpret_cast<char*>(state->data()), static_cast<std::streamsize>(state->size()*sizeof(double)));
	char extra = 0;
	if (!bin || bin.read(&extra, 1)) throw std::runtime_error("CUDA transport checkpoint state is truncated or has trailing data");
	return m;
}

// ValidateTransportCheckpoint (TransportCheckpoint.hpp:190-216)
inline void ValidateTransportCheckpoint(const TransportCheckpointMeta& m, std::uint64_t nodes, const CompiledSystem& s)
{
	auto close = [](double a, double b) { return std::abs(a - b) <= 1e-12*std::max({1.0, std::abs(a), std::abs(b)}); };
	if (m.nodes != nodes) throw std::runtime_error("transport checkpoint node count does not match database");
	if (m.fields != s.fields) throw std::runtime_error("transport checkpoint fields do not match equation system");
	if (m.system != s.name) throw std::runtime_error("transport checkpoint system does not match configuration");
	if (m.velocity_source != s.velocity_source)
		throw std::runtime_error("transport checkpoint velocity source does not match configuration");
	if (m.completed_step > s.steps) throw std::runtime_error("transport checkpoint is beyond configured final step");
	if (!close(m.physical_time, m.completed_step*s.dt))
		throw std::runtime_error("transport checkpoint physical time is inconsistent with completed step");
	if (!close(m.dt, s.dt)) throw std::runtime_error("transport checkpoint dt does not match configuration");
}

}  // namespace tfi_host
