"""pp_solve_device_noalloc: the FP64 block-Jacobi solve on caller device
buffers with every scratch buffer in the library's static device array.
Checked against pp_solve (same kernels, so bit-identical) on two calls in a
row, the second on a smaller plate, and against the FP64 oracle; oversized
plates and unsupported options are rejected.  Device buffers come from the
HIP runtime through ctypes."""
import ctypes as C

import numpy as np
import pytest

import oracle_api as O
from platepcg import Context, Plan, _lib
from platepcg import problem as P

pytestmark = pytest.mark.gpu


class Hip:
    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def upload(self, dst, a):
        assert self.lib.hipMemcpy(dst, a.ctypes.data, a.nbytes, 1) == 0      # host -> device

    def download(self, src, a):
        assert self.lib.hipMemcpy(a.ctypes.data, src, a.nbytes, 2) == 0      # device -> host

    def free(self, p):
        assert self.lib.hipFree(p) == 0


def noalloc_solve(hip, plan, F64, **kw):
    dF, dx = hip.alloc(F64.nbytes), hip.alloc(F64.nbytes)
    try:
        hip.upload(dF, F64)
        st = _lib.SolveStats()
        rc = plan.lib.pp_solve_device_noalloc(plan.h, dF, dx, C.byref(plan.options(**kw)),
                                              C.byref(st))
        x = np.empty_like(F64)
        if rc == 0:
            hip.download(dx, x)
        return rc, x, st
    finally:
        hip.free(dF)
        hip.free(dx)


def test_noalloc_matches_pp_solve_twice_second_smaller():
    hip = Hip()
    ctx = Context(0)
    for name in ("c2", "c1"):           # 61,347 DOF, then 3,267 DOF
        w = P.workload(name)
        plan = Plan(ctx, w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
        F32 = plan.pressure_load(w["q0"], w.get("patches"))
        x_ref, _, st_ref = plan.solve(F32, tolerance=1e-8, max_iterations=20000)
        rc, x, st = noalloc_solve(hip, plan, F32.astype(np.float64), tolerance=1e-8,
                                  max_iterations=20000)
        assert rc == 0, plan.lib.pp_last_error()
        print(f"{name}: noalloc {st.iterations} its, {st.device_ms:.2f} ms, "
              f"true residual {st.true_rel_residual:.2e}", flush=True)
        assert st.converged and st.iterations == st_ref["iterations"]
        assert st.true_rel_residual <= 1e-8
        np.testing.assert_array_equal(x, x_ref)
        if name == "c1":
            op = O.OracleProblem(w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
            xo, _ = op.pcg(F32.astype(np.float64), tol=1e-8)
            assert O.max_nodal_w_error(x, xo) < 1e-6
        plan.close()
    ctx.close()


def test_noalloc_rejects_what_it_cannot_run():
    hip = Hip()
    ctx = Context(0)
    w = P.workload("c1")
    plan = Plan(ctx, w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
    F = plan.pressure_load(w["q0"]).astype(np.float64)
    for kw in ({"precond": _lib.PRECOND_MULTIGRID}, {"precision": _lib.PRECISION_MIXED},
               {"flags": _lib.FLAG_CLASSIC_PCG}):
        rc, _, _ = noalloc_solve(hip, plan, F, **kw)
        assert rc == _lib.PP_ERR_UNSUPPORTED
    plan.close()
    # 500 x 500 elements: 754,509 DOFs, above the 16 MiB scratch limit
    mesh = P.rect_mesh(500)
    mask = P.constrained_mask(mesh, [{"type": "fixed", "location": "all_edges"}])
    big = Plan(ctx, mesh, 210e9, 0.3, 0.1, mask)
    rc, _, _ = noalloc_solve(hip, big, np.zeros(big.n_dof))
    assert rc == _lib.PP_ERR_INVALID and b"scratch limit" in big.lib.pp_last_error()
    big.close()
    ctx.close()
