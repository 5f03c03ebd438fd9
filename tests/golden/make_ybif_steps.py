"""Writes tests/golden/ybif_small_pulsatile_steps.npz: the CPU oracle's
backward-Euler states of the small Y-bifurcation (ybif_small) under the
pulsatile inflow Q(t)/Q0 = 1 + 0.6 sin 2 pi t + 0.2 sin 4 pi t at dt = 5 ms,
from the prescribed inflow at t = 0, every step solved by the oracle's
Newton (exact tangent, FGMRES + block Jacobi) to a relative residual of
1e-13.  States after steps 4 and 8 are stored.  Test infrastructure:
tests/test_gpu_flow.py checks the GPU's plain and mixed + predicted paths
against it (the oracle takes about ten seconds a step here, too slow for the
GPU test itself)."""
import math
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import oracle as O  # noqa: E402
from tubularflowiga_amd import boundary, casegen  # noqa: E402

RHO, MU = 1060.0 * (2e-3) ** 2, 3.5e-3
DT, STEPS, KEEP = 5e-3, 8, (4, 8)


def inflow_scale(t):
    return 1.0 + 0.6 * math.sin(2 * math.pi * t) + 0.2 * math.sin(4 * math.pi * t)


def main():
    case = casegen.workload("ybif_small")
    pattern = O.block_pattern(case.mesh)
    vmask, pmask, vt, pt = boundary.resolve_flow_boundaries(case)
    state = boundary.initial_state(vmask, pmask, vt * inflow_scale(0.0), pt)
    kept = {}
    for k in range(1, STEPS + 1):
        prev = state.copy()
        state, its, lin, _ = O.ns_newton_step(case.mesh, pattern, state, prev, RHO, MU, DT, vmask, pmask,
                                              vt * inflow_scale(k * DT), pt, max_newton=60, rtol=1e-13, atol=1e-16,
                                              linear_rtol=1e-12)
        print(f"step {k}: newton {its} linear {lin}", flush=True)
        if k in KEEP:
            kept[f"state_{k}"] = state.copy()
    np.savez_compressed(Path(__file__).with_name("ybif_small_pulsatile_steps.npz"), dt=DT, rho=RHO, mu=MU,
                        nodes=case.mesh.nodes, **kept)


if __name__ == "__main__":
    main()
