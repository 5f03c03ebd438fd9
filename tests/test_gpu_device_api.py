"""The device-pointer entry points, the stand-alone AMG handle and the
learned-predictor training call, through the C-ABI.

* tfi_ns_assemble_device / tfi_bsr_spmv_device / tfi_apply_dirichlet_device
  run the same kernels as their host-array forms (tfi_ns_assemble,
  tfi_bsr_spmv, tfi_apply_dirichlet, whose oracle parity test_gpu_kernels.py
  holds): the outputs must be bit-identical.
* tfi_amg_*: the V-cycle is a linear map fixed by the matrix, so for a matrix
  scaled by two every operation of the set-up and the cycle scales exactly
  (plain aggregation, Galerkin products, block LU, Gauss-Seidel and
  Chebyshev sweeps; FP32 and FP16 conversions of an exactly halved value):
  M(2A)^-1 v == M(A)^-1 v / 2 bit for bit, after tfi_amg_update on the same
  handle.  A right-preconditioned GMRES written here, on the host, with the
  handle as its preconditioner, must reach the library solver's solution.
* tfi_flow_warm_train: snapshots start once five states precede one, and the
  loss it returns is the predictor's own.
The device buffers are allocated here with the HIP runtime.
"""
import ctypes

import numpy as np
import pytest

from tubularflowiga_amd import boundary, casegen

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2
RHO, MU = 1060.0 * (2e-3) ** 2, 3.5e-3   # normalised density for L0 = 2 mm


def _hip():
    for path in ("/opt/rocm/lib/libamdhip64.so", "libamdhip64.so"):
        try:
            lib = ctypes.CDLL(path)
        except OSError:
            continue
        lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        lib.hipFree.argtypes = [ctypes.c_void_p]
        lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        return lib
    pytest.fail("HIP runtime library not found")


class DeviceArray:
    def __init__(self, hip, host: np.ndarray):
        self.hip, self.nbytes, self.dtype = hip, host.nbytes, host.dtype
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), self.nbytes) == 0
        self.upload(host)

    def upload(self, host: np.ndarray):
        host = np.ascontiguousarray(host, self.dtype)
        assert host.nbytes == self.nbytes
        assert self.hip.hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, H2D) == 0

    def download(self) -> np.ndarray:
        out = np.empty(self.nbytes // self.dtype.itemsize, self.dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


@pytest.fixture(scope="module")
def ctx():
    from tubularflowiga_amd import device
    return device.Context(0)


@pytest.fixture(scope="module")
def hip():
    return _hip()


@pytest.fixture(scope="module", params=["tube_c0", "ybif_small"])
def case(request):
    return casegen.workload(request.param)


@pytest.fixture(scope="module")
def dmesh(ctx, case):
    from tubularflowiga_amd import device
    dm = device.DeviceMesh(ctx, case.mesh)
    yield dm
    dm.close()


def _state(nodes, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    s = np.zeros((nodes, 4))
    s[:, :3] = 50.0 * scale * rng.standard_normal((nodes, 3))
    s[:, 3] = 10.0 * scale * rng.standard_normal(nodes)
    return s.ravel()


def test_device_assembly_matches_host_form(dmesh, case, hip):
    n4 = 4 * dmesh.nodes
    state, prev = _state(dmesh.nodes, 1), _state(dmesh.nodes, 2, 0.5)
    bufs = [DeviceArray(hip, a) for a in (state, prev, np.zeros(16 * dmesh.nnzb), np.zeros(n4))]
    ds, dp, dj, dr = bufs
    try:
        for tangent, code in (("exact", 1), ("legacy", 0)):
            print(f"device assembly: nodes={dmesh.nodes} tangent={tangent}", flush=True)
            jac, res = dmesh.assemble(state, prev, RHO, MU, 1e-3, tangent=tangent)
            dmesh.ctx.check(dmesh.lib.tfi_ns_assemble_device(dmesh.handle, ds.ptr, dp.ptr, RHO, MU, 1e-3, dj.ptr,
                                                             dr.ptr, code))
            assert np.array_equal(dj.download(), jac.ravel())
            assert np.array_equal(dr.download(), res)
        # residual only, and the steady form with no previous state
        _, res = dmesh.assemble(state, state, RHO, MU, 0.0, jacobian=False)
        dmesh.ctx.check(dmesh.lib.tfi_ns_assemble_device(dmesh.handle, ds.ptr, None, RHO, MU, 0.0, None, dr.ptr, 1))
        assert np.array_equal(dr.download(), res)
    finally:
        for b in bufs:
            b.free()


def test_device_spmv_and_dirichlet_match_host_forms(dmesh, case, hip):
    rng = np.random.default_rng(5)
    vals = rng.standard_normal(16 * dmesh.nnzb)
    x = rng.standard_normal(4 * dmesh.nodes)
    state = _state(dmesh.nodes, 9)
    vmask, pmask, vt, pt = boundary.resolve_flow_boundaries(case)
    pmask = pmask.copy()
    pmask[::17] = 1
    pt = np.linspace(-1.0, 1.0, len(pt))
    rhs = rng.standard_normal(4 * dmesh.nodes)
    arrays = [vals, x, np.zeros(4 * dmesh.nodes), np.ascontiguousarray(vmask, np.int32),
              np.ascontiguousarray(pmask, np.int32), np.ascontiguousarray(vt, np.float64).ravel(),
              np.ascontiguousarray(pt, np.float64), state, rhs]
    bufs = [DeviceArray(hip, a) for a in arrays]
    dv, dx, dy, dvm, dpm, dvt, dpt, ds, drhs = bufs
    try:
        print(f"device spmv / dirichlet: nodes={dmesh.nodes}", flush=True)
        dmesh.ctx.check(dmesh.lib.tfi_bsr_spmv_device(dmesh.handle, dv.ptr, dx.ptr, dy.ptr))
        assert np.array_equal(dy.download(), dmesh.spmv(vals, x))
        jac, r = dmesh.apply_dirichlet(vmask, pmask, vt, pt, state, vals.copy(), rhs.copy())
        dmesh.ctx.check(dmesh.lib.tfi_apply_dirichlet_device(dmesh.handle, dvm.ptr, dpm.ptr, dvt.ptr, dpt.ptr, ds.ptr,
                                                             dv.ptr, drhs.ptr))
        assert np.array_equal(dv.download(), jac.ravel())
        assert np.array_equal(drhs.download(), r)
    finally:
        for b in bufs:
            b.free()


def _newton_system(dmesh, case, dt=1e-3):
    vmask, pmask, vt, pt = boundary.resolve_flow_boundaries(case)
    s0 = boundary.initial_state(vmask, pmask, vt, pt)
    jac, rhs = dmesh.assemble(s0, s0, RHO, MU, dt)
    jac, rhs = dmesh.apply_dirichlet(vmask, pmask, vt, pt, s0, jac, rhs)
    return jac.ravel(), rhs


def _gmres(matvec, precondition, b, rtol, restart=60, maxit=600):
    """Right-preconditioned restarted GMRES (modified Gram-Schmidt)."""
    x = np.zeros_like(b)
    bnorm = np.linalg.norm(b)
    its = 0
    while its < maxit:
        r = b - matvec(x)
        beta = np.linalg.norm(r)
        if beta <= rtol * bnorm:
            break
        V, Z = [r / beta], []
        H = np.zeros((restart + 1, restart))
        for j in range(restart):
            Z.append(precondition(V[j]))
            w = matvec(Z[j])
            for i in range(j + 1):
                H[i, j] = w @ V[i]
                w = w - H[i, j] * V[i]
            H[j + 1, j] = np.linalg.norm(w)
            V.append(w / H[j + 1, j])
            its += 1
            e = np.zeros(j + 2)
            e[0] = beta
            y = np.linalg.lstsq(H[:j + 2, :j + 1], e, rcond=None)[0]
            if np.linalg.norm(H[:j + 2, :j + 1] @ y - e) <= rtol * bnorm or its >= maxit:
                break
        x = x + np.array(Z).T @ y
    return x, its


AMG_VARIANTS = [(0, 3), (0, 2), (0, 0), (1, 1), (1, 2)]   # (precision, smoother): FP64 AUTO / Chebyshev / Jacobi, mixed GS / Chebyshev


@pytest.mark.parametrize("precision,smoother", AMG_VARIANTS)
def test_amg_handle_update_scales_exactly(dmesh, case, hip, precision, smoother):
    from tubularflowiga_amd import device
    jac, rhs = _newton_system(dmesh, case)
    rng = np.random.default_rng(3)
    v = rng.standard_normal(4 * dmesh.nodes)
    bufs = [DeviceArray(hip, a) for a in (jac, 2.0 * jac, v, np.zeros_like(v))]
    da, d2a, dv, dz = bufs
    amg = device.AmgPreconditioner(dmesh, precision=precision, smoother=smoother)
    other = device.AmgPreconditioner(dmesh, precision=precision, smoother=smoother)
    try:
        print(f"amg handle: nodes={dmesh.nodes} precision={precision} smoother={smoother}", flush=True)
        with pytest.raises(device.TfiError):
            amg.apply(dv.ptr, dz.ptr)          # no hierarchy yet
        amg.setup(da.ptr)
        levels = amg.levels()
        assert len(levels) >= 1 and levels[0][0] == dmesh.nodes and levels[0][1] == dmesh.nnzb
        amg.apply(dv.ptr, dz.ptr)
        z1 = dz.download()
        assert np.all(np.isfinite(z1)) and np.linalg.norm(z1) > 0.0
        amg.update(d2a.ptr)
        assert amg.levels() == levels
        amg.apply(dv.ptr, dz.ptr)
        assert np.array_equal(dz.download(), 0.5 * z1)
        amg.update(da.ptr)
        amg.apply(dv.ptr, dz.ptr)
        assert np.array_equal(dz.download(), z1)
        other.setup(da.ptr)
        other.apply(dv.ptr, dz.ptr)
        assert np.array_equal(dz.download(), z1)
    finally:
        amg.close()
        other.close()
        for b in bufs:
            b.free()


@pytest.mark.parametrize("precision,smoother", AMG_VARIANTS)
def test_amg_handle_preconditions_a_host_gmres(dmesh, case, hip, precision, smoother):
    from tubularflowiga_amd import device
    jac, rhs = _newton_system(dmesh, case)
    bufs = [DeviceArray(hip, a) for a in (jac, rhs, np.zeros_like(rhs))]
    da, dv, dz = bufs
    amg = device.AmgPreconditioner(dmesh, precision=precision, smoother=smoother)
    try:
        amg.setup(da.ptr)

        def precondition(r):
            dv.upload(r)
            amg.apply(dv.ptr, dz.ptr)
            return dz.download()

        x, its = _gmres(lambda y: dmesh.spmv(jac, y), precondition, rhs, rtol=1e-8)
        ref, ref_its, _ = dmesh.solve(jac, rhs, rtol=1e-10)
        res = np.linalg.norm(rhs - dmesh.spmv(jac, x)) / np.linalg.norm(rhs)
        err = np.linalg.norm(x - ref) / np.linalg.norm(ref)
        print(f"host GMRES + AMG handle: nodes={dmesh.nodes} precision={precision} smoother={smoother} its={its} "
              f"relres={res:.2e} err={err:.2e} (library solve {ref_its} its)", flush=True)
        assert res <= 1.0000001e-8
        assert err <= 1e-6
        assert its <= 4 * max(ref_its, 10)
    finally:
        amg.close()
        for b in bufs:
            b.free()


def test_warm_train_snapshots_and_loss(ctx):
    from tubularflowiga_amd import device
    case = casegen.workload("ybif_small")
    dm = device.DeviceMesh(ctx, case.mesh)
    flow = device.FlowSolver(dm, density=RHO, viscosity=MU, dt=1e-3, warm_start=device.TFI_WARM_LEARNED)
    flow.configure_warm_start(hidden=64, train_steps=10, learning_rate=1e-3, min_snapshots=2)
    rng = np.random.default_rng(17)
    n = 4 * dm.nodes
    base, drift = rng.standard_normal(n), rng.standard_normal(n)
    flow.set_state(base)             # the current state; clears the history
    try:
        losses = []
        for k in range(9):
            u = base + 0.1 * k * drift + 0.01 * rng.standard_normal(n)
            losses.append(flow.warm_train(u, phase=[np.sin(0.1 * k), np.cos(0.1 * k)]))
            snapshots = flow.warm_stats()["snapshots"]
            print(f"warm train k={k} loss={losses[-1]:.3e} snapshots={snapshots}", flush=True)
            assert snapshots == max(0, k - 4)
        assert losses[:5] == [0.0] * 5
        assert all(np.isfinite(losses[5:])) and all(loss > 0.0 for loss in losses[5:])
        assert losses[-1] < losses[5]      # the same noise model each snapshot: the fit improves
        assert flow.warm_stats()["last_loss"] == losses[-1]
        pred, _ = flow.warm_predict(device.TFI_WARM_LEARNED)
        assert np.all(np.isfinite(pred))
        with pytest.raises(device.TfiError):
            flow.warm_train(base, phase=np.zeros(9))
    finally:
        flow.close()
        dm.close()
