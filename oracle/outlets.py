"""Lumped outlet models and the per-step outlet fixed point, restated in plain
Python.  TEST INFRASTRUCTURE ONLY: the checker for the drop-in driver's outlet
coupling, never part of the product path.

Follows include/OutletModel.hpp (model state, evaluation, coupling tolerance,
relaxation, commit) and the driver loop of solvers/cuda/src/iga_cuda.cu:927-1084
(up to 12 passes per physical step, each solving the step with the models'
current pressures as outlet tractions, then evaluating the models on the outlet
flows).  Pinned by tests/test_oracle_outlets.py against the reference's own
header outputs (tests/golden/outlet_vectors.json, written through
oracle/_ref/ref_probe), the values asserted in
solvers/cpu/tests/test_outlet_model.cpp and closed forms (the RC backward-Euler
recursion's steady state p_ref + R Q and its exact geometric decay).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, replace

KINDS = ("resistance", "windkessel_rc", "windkessel_rcr")


@dataclass
class Outlet:
    """OutletModelState (include/OutletModel.hpp:15-26)."""
    label: int
    kind: str
    resistance: float = 0.0
    proximal_resistance: float = 0.0
    distal_resistance: float = 0.0
    capacitance: float = 0.0
    reference_pressure: float = 0.0
    capacitor_pressure: float = 0.0
    flow: float = 0.0
    pressure: float = 0.0


def init_models(boundaries, pressure_field="pressure"):
    """InitializeOutletModels (include/OutletModel.hpp:41-68) over the
    `boundaries` list of a simulation_config.json."""
    models = []
    for b in boundaries:
        for c in b["conditions"]:
            if c["field"] != pressure_field or c["type"] not in KINDS:
                continue
            models.append(Outlet(
                label=int(b["label"]), kind=c["type"], resistance=float(c.get("resistance", 0.0)),
                proximal_resistance=float(c.get("proximal_resistance", 0.0)),
                distal_resistance=float(c.get("distal_resistance", 0.0)),
                capacitance=float(c.get("capacitance", 0.0)),
                reference_pressure=float(c.get("reference_pressure", 0.0)),
                capacitor_pressure=float(c.get("initial_pressure", 0.0)),
                pressure=float(c.get("reference_pressure", 0.0)) if c["type"] == "resistance"
                else float(c.get("initial_pressure", 0.0))))
    return models


def evaluate(model: Outlet, flow: float, dt: float):
    """EvaluateOutletModel (include/OutletModel.hpp:70-90) -> (pressure,
    capacitor pressure).  Backward Euler for C dPc/dt = Q - (Pc - p_ref)/R."""
    if not math.isfinite(flow):
        raise ValueError("outlet flow must be finite")
    if model.kind == "resistance":
        return model.reference_pressure + model.resistance * flow, model.capacitor_pressure
    if not dt > 0.0:
        raise ValueError("RC/RCR outlets require positive transient dt")
    r = model.resistance if model.kind == "windkessel_rc" else model.distal_resistance
    pc = (model.capacitance * model.capacitor_pressure / dt + flow + model.reference_pressure / r) \
        / (model.capacitance / dt + 1.0 / r)
    return pc + (model.proximal_resistance * flow if model.kind == "windkessel_rcr" else 0.0), pc


@dataclass
class Coupling:
    """OutletCouplingEvaluation (include/OutletModel.hpp:33-39)."""
    flow: list
    pressure: list
    capacitor_pressure: list
    maximum_pressure_change: float = 0.0
    pressure_scale: float = 1.0

    def tolerance(self):
        """OutletCouplingTolerance (include/OutletModel.hpp:118-121)."""
        return 1e-8 + 1e-6 * self.pressure_scale


def evaluate_coupling(models, capacitor_start, flows, dt):
    """EvaluateOutletCoupling (include/OutletModel.hpp:92-116): every model from
    the capacitor pressure the step started with."""
    if len(models) != len(capacitor_start) or len(models) != len(flows):
        raise ValueError("outlet coupling vectors have inconsistent sizes")
    out = Coupling(list(flows), [], [])
    for m, pc0, q in zip(models, capacitor_start, flows):
        p, pc = evaluate(replace(m, capacitor_pressure=pc0), q, dt)
        out.pressure.append(p)
        out.capacitor_pressure.append(pc)
        out.maximum_pressure_change = max(out.maximum_pressure_change, abs(p - m.pressure))
        out.pressure_scale = max(out.pressure_scale, abs(p))
    return out


def relax(models, ev: Coupling, relaxation=0.5):
    """RelaxOutletCoupling (include/OutletModel.hpp:123-135)."""
    if not 0.0 < relaxation <= 1.0:
        raise ValueError("invalid outlet coupling relaxation")
    for m, q, p in zip(models, ev.flow, ev.pressure):
        m.flow = q
        m.pressure = (1.0 - relaxation) * m.pressure + relaxation * p


def commit(models, ev: Coupling):
    """CommitOutletCoupling (include/OutletModel.hpp:137-149)."""
    for m, q, p, pc in zip(models, ev.flow, ev.pressure, ev.capacitor_pressure):
        m.flow, m.pressure, m.capacitor_pressure = q, p, pc


def coupled_step(models, dt, solve, flows, passes=12):
    """One physical step's outlet fixed point (iga_cuda.cu:927-1084).

    ``solve(pressures)`` runs the step's nonlinear solve with ``pressures``
    (label -> traction pressure) from wherever the previous pass left the
    state; ``flows(labels)`` returns the outlet flows of the current state.
    Returns the list of (maximum pressure change, tolerance) per pass; raises
    if the fixed point does not settle within ``passes``."""
    start = [m.capacitor_pressure for m in models]
    history = []
    for _ in range(passes if models else 1):
        solve({m.label: m.pressure for m in models})
        if not models:
            return history
        ev = evaluate_coupling(models, start, flows([m.label for m in models]), dt)
        history.append((ev.maximum_pressure_change, ev.tolerance()))
        if ev.maximum_pressure_change <= ev.tolerance():
            commit(models, ev)
            return history
        relax(models, ev)
    raise RuntimeError("outlet fixed-point iteration did not converge")
