// Lumped outlet models coupled to the Navier-Stokes step: resistance, RC and
// RCR Windkessel outlets.  The models, their tolerance and the under-relaxed
// fixed point are the reference's contract (include/OutletModel.hpp:70-149,
// the driver loop of iga_cuda.cu:927-1084, checkpoint fields of
// include/OutletCheckpoint.hpp:22-48); the organisation here is different.
//
// All outlets of a case live in one OutletBank, as parallel arrays.  Over one
// time step every model is affine in its outlet flow Q, because the
// capacitor pressure it starts from is fixed for the step:
//   resistance       p = p_ref + R Q
//   RC  (R_d = R)    pc = (C pc0/dt + p_ref/R_d + Q) / (C/dt + 1/R_d),  p = pc
//   RCR              pc as RC with the distal R_d,                      p = pc + R_p Q
// (backward Euler for C dpc/dt = Q - (pc - p_ref)/R_d).  StartStep() fixes
// pc0; each outlet pass then evaluates the responses of all outlets to the
// measured flows, compares them with the pressures the pass applied, and
// either relaxes the applied pressures by 1/2 or commits the responses.
#pragma once

#include "case_input.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace tfi_host {

enum class OutletKind { kResistance, kRc, kRcr };

inline bool ParseOutletKind(const std::string& name, OutletKind* kind)
{
	static const std::pair<const char*, OutletKind> table[] = {
		{"resistance", OutletKind::kResistance}, {"windkessel_rc", OutletKind::kRc}, {"windkessel_rcr", OutletKind::kRcr}};
	for (const auto& entry : table)
		if (name == entry.first) {
			if (kind) *kind = entry.second;
			return true;
		}
	return false;
}

inline const char* OutletKindName(OutletKind kind)
{
	switch (kind) {
	case OutletKind::kResistance: return "resistance";
	case OutletKind::kRc: return "windkessel_rc";
	default: return "windkessel_rcr";
	}
}

class OutletBank {
public:
	// One outlet per condition of an outlet kind on the flow system's
	// pressure field, in boundary order.  The first applied pressure is the
	// reference pressure (resistance) or the initial capacitor pressure.
	static OutletBank FromConfig(const SimulationConfig& config, const FlowSystem& system)
	{
		OutletBank bank;
		for (const auto& boundary : config.boundaries)
			for (const auto& cond : boundary.conditions) {
				OutletKind kind;
				if (cond.field != system.pressure || !ParseOutletKind(cond.type, &kind)) continue;
				const double first = kind == OutletKind::kResistance ? cond.reference_pressure : cond.initial_pressure;
				bank.Add(boundary.label, kind, cond.resistance, cond.proximal_resistance, cond.distal_resistance,
					cond.capacitance, cond.reference_pressure, cond.initial_pressure, first);
			}
		return bank;
	}

	void Add(int label, OutletKind kind, double resistance, double proximal, double distal, double capacitance,
		double reference_pressure, double capacitor_pressure, double applied_pressure)
	{
		label_.push_back(label);
		kind_.push_back(kind);
		r_.push_back(resistance);
		rp_.push_back(proximal);
		rd_.push_back(distal);
		c_.push_back(capacitance);
		pref_.push_back(reference_pressure);
		capacitor_.push_back(capacitor_pressure);
		applied_.push_back(applied_pressure);
		flow_.push_back(0.0);
		capacitor_start_.push_back(capacitor_pressure);
		response_.push_back(0.0);
		response_capacitor_.push_back(capacitor_pressure);
	}

	bool empty() const { return label_.empty(); }
	std::size_t size() const { return label_.size(); }
	const std::vector<int>& labels() const { return label_; }
	double applied(std::size_t i) const { return applied_[i]; }
	double flow(std::size_t i) const { return flow_[i]; }
	double capacitor(std::size_t i) const { return capacitor_[i]; }
	const char* kind(std::size_t i) const { return OutletKindName(kind_[i]); }

	// Outlet pressures the next solve applies as tractions, by label.
	std::map<int, double> Tractions() const
	{
		std::map<int, double> out;
		for (std::size_t i = 0; i < size(); ++i) out.emplace(label_[i], applied_[i]);
		return out;
	}

	// Fixes the capacitor pressures the step's passes start from.
	void StartStep() { capacitor_start_ = capacitor_; }

	// Responses of every outlet to the measured flows; returns the largest
	// change against the applied pressures.  Tolerance() then holds
	// 1e-8 + 1e-6 max(1, |responses|).
	double Respond(const std::vector<double>& flows, double dt)
	{
		if (flows.size() != size()) throw std::runtime_error("outlet coupling vectors have inconsistent sizes");
		double change = 0.0;
		scale_ = 1.0;
		for (std::size_t i = 0; i < size(); ++i) {
			const double q = flows[i];
			if (!std::isfinite(q)) throw std::runtime_error("outlet flow must be finite");
			double p, pc = capacitor_start_[i];
			if (kind_[i] == OutletKind::kResistance) {
				p = pref_[i] + r_[i]*q;
			} else {
				if (!(dt > 0.0)) throw std::runtime_error("RC/RCR outlets require positive transient dt");
				const double rd = kind_[i] == OutletKind::kRc ? r_[i] : rd_[i];
				const double g = c_[i]/dt + 1.0/rd;
				pc = (c_[i]*capacitor_start_[i]/dt + pref_[i]/rd)/g + q/g;
				p = kind_[i] == OutletKind::kRcr ? pc + rp_[i]*q : pc;
			}
			response_[i] = p;
			response_capacitor_[i] = pc;
			change = std::max(change, std::abs(p - applied_[i]));
			scale_ = std::max(scale_, std::abs(p));
		}
		measured_ = flows;
		return change;
	}

	double Tolerance() const { return 1e-8 + 1e-6*scale_; }

	// Applied pressures move halfway (by `weight`) to the last responses.
	void Relax(double weight = 0.5)
	{
		if (measured_.size() != size() || !(weight > 0.0 && weight <= 1.0))
			throw std::runtime_error("invalid outlet coupling relaxation");
		for (std::size_t i = 0; i < size(); ++i) {
			flow_[i] = measured_[i];
			applied_[i] += weight*(response_[i] - applied_[i]);
		}
	}

	// The last responses become the committed outlet state.
	void Commit()
	{
		if (measured_.size() != size()) throw std::runtime_error("invalid outlet coupling commit");
		for (std::size_t i = 0; i < size(); ++i) {
			flow_[i] = measured_[i];
			applied_[i] = response_[i];
			capacitor_[i] = response_capacitor_[i];
		}
	}

	std::vector<CheckpointOutlet> Checkpoint() const
	{
		std::vector<CheckpointOutlet> out(size());
		for (std::size_t i = 0; i < size(); ++i)
			out[i] = CheckpointOutlet{label_[i], OutletKindName(kind_[i]), flow_[i], applied_[i], capacitor_[i]};
		return out;
	}

	void Restore(const std::vector<CheckpointOutlet>& saved)
	{
		if (saved.size() != size()) throw std::runtime_error("flow checkpoint outlet count does not match configuration");
		std::map<int, const CheckpointOutlet*> by_label;
		for (const auto& s : saved) by_label.emplace(s.label, &s);
		for (std::size_t i = 0; i < size(); ++i) {
			const auto it = by_label.find(label_[i]);
			if (it == by_label.end() || it->second->kind != OutletKindName(kind_[i]))
				throw std::runtime_error("flow checkpoint outlet labels or kinds do not match configuration");
			flow_[i] = it->second->flow;
			applied_[i] = it->second->pressure;
			capacitor_[i] = it->second->capacitor_pressure;
		}
	}

private:
	std::vector<int> label_;
	std::vector<OutletKind> kind_;
	std::vector<double> r_, rp_, rd_, c_, pref_;
	std::vector<double> applied_, capacitor_, flow_;       // committed / applied state
	std::vector<double> capacitor_start_;                  // capacitor pressures of the step start
	std::vector<double> response_, response_capacitor_;    // last responses
	std::vector<double> measured_;                         // flows of the last responses
	double scale_ = 1.0;
};

}  // namespace tfi_host
