"""GPU parity on the reference's own element sizes: tricubic elements
(nen = 64, C2 cubic splines, the only kind the reference's iga_pack writes)
and larger elements as next to extraordinary edges (nen up to 80,
IgaCudaKernels.cuh:14, DeviceMesh.hpp:55-56), against the CPU oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from tubularflowiga_amd import boundary, casegen

from meshes import split_basis

pytestmark = pytest.mark.gpu

RHO, MU = 1060.0 * (1e-3) ** 2, 3.5e-3


@pytest.fixture(scope="module")
def ctx():
    from tubularflowiga_amd import device
    return device.Context(0)


@pytest.fixture(scope="module")
def cubic_tube():
    return casegen.workload("tube_c0", degree=3)


@pytest.fixture(scope="module", params=["nen64", "nen80"])
def mesh(request, cubic_tube):
    m = cubic_tube.mesh
    if request.param == "nen80":
        m = split_basis(m, every=3, extra=16)
        assert m.maximum_basis == 80
    else:
        assert m.maximum_basis == 64
    print(f"tricubic case {request.param}: elements={m.elements} nodes={m.nodes}", flush=True)
    return m


@pytest.fixture(scope="module")
def dmesh(ctx, mesh):
    from tubularflowiga_amd import device
    return device.DeviceMesh(ctx, mesh)


def _state(nodes, seed):
    rng = np.random.default_rng(seed)
    s = np.zeros((nodes, 4))
    s[:, :3] = 50.0 * rng.standard_normal((nodes, 3))
    s[:, 3] = 10.0 * rng.standard_normal(nodes)
    return s.ravel()


def test_pattern_bit_exact(dmesh, mesh):
    ro, cols, diag = dmesh.pattern()
    oro, ocols, odiag = O.block_pattern(mesh)
    assert np.array_equal(ro, oro) and np.array_equal(cols, ocols) and np.array_equal(diag, odiag)


@pytest.mark.parametrize("table", [1, 0])
@pytest.mark.parametrize("tangent", ["legacy", "exact"])
def test_assembly_matches_oracle(dmesh, mesh, tangent, table):
    state, prev = _state(mesh.nodes, 11), _state(mesh.nodes, 12)
    pattern = O.block_pattern(mesh)
    try:
        assert (dmesh.basis_table(table) > 0) == bool(table)
        jac, negres = dmesh.assemble(state, prev, RHO, MU, 1e-3, tangent=tangent)
        _, negres_only = dmesh.assemble(state, prev, RHO, MU, 1e-3, jacobian=False, tangent=tangent)
    finally:
        dmesh.basis_table(1)
    ojac, onegres = O.ns_assemble(mesh, pattern, state, prev, RHO, MU, 1e-3, tangent)
    assert np.max(np.abs(negres - onegres)) <= 1e-11 * np.max(np.abs(onegres))
    assert np.max(np.abs(jac - ojac)) <= 1e-11 * np.max(np.abs(ojac))
    assert np.max(np.abs(negres_only - negres)) <= 1e-13 * np.max(np.abs(onegres))


def test_backward_euler_step_matches_oracle(ctx, cubic_tube):
    """One Newton solve of a backward-Euler step on the tricubic tube (exact
    tangent, AMG-preconditioned FGMRES) within 1e-10 relative L2 of the
    oracle's Newton solution."""
    from tubularflowiga_amd import device
    case = cubic_tube
    rho, mu = casegen.flow_parameters(case)
    dm = device.DeviceMesh(ctx, case.mesh)
    vmask, pmask, vt, pt = boundary.resolve_flow_boundaries(case)
    s0 = boundary.initial_state(vmask, pmask, vt, pt)
    flow = device.FlowSolver(dm, density=rho, viscosity=mu, dt=1e-3, nonlinear_rtol=1e-12, nonlinear_atol=1e-16,
                             mass_rtol=1e30, linear_rtol=1e-12, maximum_newton=30)
    try:
        flow.set_state(s0)
        flow.set_boundary(vmask, pmask, vt, pt)
        flow.set_boundary_faces(*boundary.boundary_faces(case.mesh))
        flow.begin_step()
        flow.newton()
        gpu = flow.state().reshape(-1, 4)
    finally:
        flow.close()
        dm.close()
    pattern = O.block_pattern(case.mesh)
    ref, _, _, _ = O.ns_newton_step(case.mesh, pattern, s0, s0.copy(), rho, mu, 1e-3, vmask, pmask, vt, pt,
                                    max_newton=40, rtol=1e-13, atol=1e-16, linear_rtol=1e-12, tangent="exact")
    ref = ref.reshape(-1, 4)
    v = np.linalg.norm(gpu[:, :3] - ref[:, :3]) / np.linalg.norm(ref[:, :3])
    p = np.linalg.norm(gpu[:, 3] - ref[:, 3]) / np.linalg.norm(ref[:, 3])
    assert v <= 1e-10 and p <= 1e-10, (v, p)
