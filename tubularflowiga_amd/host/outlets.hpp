// Lumped outlet models coupled to the Navier-Stokes step: resistance, RC and
// RCR Windkessel outlets (include/OutletModel.hpp, include/OutletCheckpoint.hpp), the
// driver-level fixed point the reference runs around each time step
// (iga_cuda.cu:927-1084).  Each model turns the volumetric flow through its
// outlet label into a pressure that the step applies as a pressure traction;
// the step is repeated with relaxed pressures until they settle.
#pragma once

#include "case_input.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace tfi_host {

struct OutletModel {
	int label = -1;
	std::string kind;  // "resistance", "windkessel_rc" or "windkessel_rcr"
	double resistance = 0.0, proximal_resistance = 0.0, distal_resistance = 0.0, capacitance = 0.0;
	double reference_pressure = 0.0;
	double capacitor_pressure = 0.0;  // committed state of the capacitor
	double flow = 0.0;
	double pressure = 0.0;            // the pressure the next solve applies
};

inline bool IsOutletKind(const std::string& type)
{
	return type == "resistance" || type == "windkessel_rc" || type == "windkessel_rcr";
}

// InitializeOutletModels (include/OutletModel.hpp:41-68): one model per outlet
// condition on the system's pressure field; the first applied pressure is the
// reference pressure (resistance) or the initial capacitor pressure (RC, RCR).
inline std::vector<OutletModel> InitOutletModels(const SimulationConfig& c, const FlowSystem& s)
{
	std::vector<OutletModel> models;
	for (const auto& b : c.boundaries)
		for (const auto& cond : b.conditions) {
			if (cond.field != s.pressure || !IsOutletKind(cond.type)) continue;
			OutletModel m;
			m.label = b.label;
			m.kind = cond.type;
			m.resistance = cond.resistance;
			m.proximal_resistance = cond.proximal_resistance;
			m.distal_resistance = cond.distal_resistance;
			m.capacitance = cond.capacitance;
			m.reference_pressure = cond.reference_pressure;
			m.capacitor_pressure = cond.initial_pressure;
			m.pressure = cond.type == "resistance" ? cond.reference_pressure : cond.initial_pressure;
			models.push_back(m);
		}
	return models;
}

struct OutletPressure {
	double pressure = 0.0, capacitor_pressure = 0.0;
};

// EvaluateOutletModel (include/OutletModel.hpp:70-90).  Resistance: p = p_ref + R Q.
// RC / RCR: backward Euler for C dPc/dt = Q - (Pc - p_ref)/R_d from the
// capacitor pressure at the start of the step, and p = Pc (+ R_p Q for RCR).
inline OutletPressure EvaluateOutlet(const OutletModel& m, double capacitor_start, double flow, double dt)
{
	if (!std::isfinite(flow)) throw std::runtime_error("outlet flow must be finite");
	if (m.kind == "resistance") return {m.reference_pressure + m.resistance*flow, capacitor_start};
	if (!(dt > 0.0)) throw std::runtime_error("RC/RCR outlets require positive transient dt");
	const double r = m.kind == "windkessel_rc" ? m.resistance : m.distal_resistance;
	const double pc = (m.capacitance*capacitor_start/dt + flow + m.reference_pressure/r)/(m.capacitance/dt + 1.0/r);
	return {pc + (m.kind == "windkessel_rcr" ? m.proximal_resistance*flow : 0.0), pc};
}

struct OutletCoupling {
	std::vector<double> flow, pressure, capacitor_pressure;
	double maximum_pressure_change = 0.0;
	double pressure_scale = 1.0;
	// OutletCouplingTolerance (include/OutletModel.hpp:118-121)
	double Tolerance() const { return 1e-8 + 1e-6*pressure_scale; }
};

// EvaluateOutletCoupling (include/OutletModel.hpp:92-116)
inline OutletCoupling EvaluateOutlets(const std::vector<OutletModel>& models,
	const std::vector<double>& capacitor_start, const std::vector<double>& flow, double dt)
{
	if (models.size() != capacitor_start.size() || models.size() != flow.size())
		throw std::runtime_error("outlet coupling vectors have inconsistent sizes");
	OutletCoupling r;
	r.flow = flow;
	for (std::size_t i = 0; i < models.size(); ++i) {
		const OutletPressure e = EvaluateOutlet(models[i], capacitor_start[i], flow[i], dt);
		r.pressure.push_back(e.pressure);
		r.capacitor_pressure.push_back(e.capacitor_pressure);
		r.maximum_pressure_change = std::max(r.maximum_pressure_change, std::abs(e.pressure - models[i].pressure));
		r.pressure_scale = std::max(r.pressure_scale, std::abs(e.pressure));
	}
	return r;
}

// RelaxOutletCoupling (include/OutletModel.hpp:123-135): under-relaxed pressure update.
inline void RelaxOutlets(std::vector<OutletModel>& models, const OutletCoupling& e, double relaxation = 0.5)
{
	if (models.size() != e.flow.size() || models.size() != e.pressure.size() || !(relaxation > 0.0 && relaxation <= 1.0))
		throw std::runtime_error("invalid outlet coupling relaxation");
	for (std::size_t i = 0; i < models.size(); ++i) {
		models[i].flow = e.flow[i];
		models[i].pressure = (1.0 - relaxation)*models[i].pressure + relaxation*e.pressure[i];
	}
}

// CommitOutletCoupling (include/OutletModel.hpp:137-149)
inline void CommitOutlets(std::vector<OutletModel>& models, const OutletCoupling& e)
{
	if (models.size() != e.flow.size() || models.size() != e.pressure.size() || models.size() != e.capacitor_pressure.size())
		throw std::runtime_error("invalid outlet coupling commit");
	for (std::size_t i = 0; i < models.size(); ++i) {
		models[i].flow = e.flow[i];
		models[i].pressure = e.pressure[i];
		models[i].capacitor_pressure = e.capacitor_pressure[i];
	}
}

// MaterializeOutletPressures (include/OutletModel.hpp:151-170): the pressures the
// boundary resolution applies as tractions on the outlet labels.
inline std::map<int, double> OutletPressures(const std::vector<OutletModel>& models)
{
	std::map<int, double> p;
	for (const auto& m : models) p.emplace(m.label, m.pressure);
	return p;
}

// AppendOutletCheckpoint / RestoreOutletCheckpoint (include/OutletCheckpoint.hpp:22-48)
inline std::vector<CheckpointOutlet> OutletCheckpoint(const std::vector<OutletModel>& models)
{
	std::vector<CheckpointOutlet> out;
	for (const auto& m : models) out.push_back({m.label, m.kind, m.flow, m.pressure, m.capacitor_pressure});
	return out;
}

inline void RestoreOutlets(const std::vector<CheckpointOutlet>& saved, std::vector<OutletModel>& models)
{
	if (saved.size() != models.size())
		throw std::runtime_error("flow checkpoint outlet count does not match configuration");
	for (auto& m : models) {
		const auto it = std::find_if(saved.begin(), saved.end(), [&](const CheckpointOutlet& o) { return o.label == m.label; });
		if (it == saved.end() || it->kind != m.kind)
			throw std::runtime_error("flow checkpoint outlet labels or kinds do not match configuration");
		m.flow = it->flow;
		m.pressure = it->pressure;
		m.capacitor_pressure = it->capacitor_pressure;
	}
}

}  // namespace tfi_host
