"""Python host API over the C ABI, with the reference's solver call shape.

``solve_gpu(mesh, material, coloring, F, constrained_dofs, options)`` mirrors
``solveGPU`` (gpu/solver.ts:80-104) and returns the fields of
``GPUSolveResult`` (gpu/solver.ts:47-60) plus ``solution64`` and
``trueRelResidual``.  Unlike the reference there is no CPU fallback: a GPU or
library failure raises ``PlatePCGError`` (``forceCPU`` is rejected).
"""
import ctypes as C
import time

import numpy as np

from . import _lib
from ._lib import check


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


class Context:
    """One HIP device + stream (gpu/context.ts:31-106)."""

    def __init__(self, device=0):
        self.lib = _lib.load()
        h = C.c_void_p()
        check(self.lib.pp_ctx_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.pp_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def lognormal_thickness(ctx, nx, ny, width, height, corr_len, seed, t_median, sigma_log):
    """(g, t): the unit-variance Gaussian field (FP64) and the float32
    element thickness t_median exp(sigma_log g) of an nx x ny grid,
    synthesised on the device; both (ny, nx), x fastest."""
    g = np.empty((ny, nx), np.float64)
    t = np.empty((ny, nx), np.float32)
    check(ctx.lib.pp_lognormal_thickness(ctx.h, int(nx), int(ny), float(width), float(height),
                                         float(corr_len), int(seed), float(t_median),
                                         float(sigma_log), _ptr(g), _ptr(t)))
    return g, t


class Plan:
    """Set-up for one mesh/material/support set; solve many load cases."""

    def __init__(self, ctx, mesh, E, nu, t, mask, colouring=None, element_t=None):
        self.ctx, self.lib = ctx, ctx.lib
        self._keep = []
        nodes = np.ascontiguousarray(mesh["nodes"], np.float32)
        elems = np.ascontiguousarray(mesh["elements"], np.uint32)
        self.n_nodes = int(mesh["nodeCount"])
        self.n_dof = 3 * self.n_nodes
        m = _lib.Mesh(nodes.ctypes.data, elems.ctypes.data, self.n_nodes, int(mesh["elementCount"]),
                      int(mesh.get("nodesPerElement", 4)), int(mesh.get("gridNx", 0)),
                      int(mesh.get("gridNy", 0)))
        et = None if element_t is None else np.ascontiguousarray(element_t, np.float32)
        mat = _lib.Material(float(E), float(nu), float(t), et.ctypes.data if et is not None else None)
        mask = np.ascontiguousarray(mask, np.uint8)
        col_p = None
        if colouring is not None:
            order = np.ascontiguousarray(colouring["elements_by_colour"], np.uint32)
            offs = np.ascontiguousarray(colouring["colour_offsets"], np.uint32)
            col = _lib.Colouring(order.ctypes.data, offs.ctypes.data, int(colouring["colour_count"]))
            col_p = C.byref(col)
            self._keep += [order, offs]
        h = C.c_void_p()
        check(self.lib.pp_plan_create(ctx.h, C.byref(m), C.byref(mat), col_p, _ptr(mask), C.byref(h)))
        self.h = h
        self.structured = bool(self.lib.pp_plan_is_structured(h))

    def close(self):
        if self.h:
            self.lib.pp_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- unit entry points
    def apply_k(self, x, path=_lib.PATH_AUTO):
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty(self.n_dof, np.float64)
        check(self.lib.pp_apply_k(self.h, _ptr(x), _ptr(y), path))
        return y

    def apply_k_colour(self, colour, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty(self.n_dof, np.float64)
        check(self.lib.pp_apply_k_colour(self.h, colour, _ptr(x), _ptr(y)))
        return y

    def apply_precond(self, r):
        r = np.ascontiguousarray(r, np.float64)
        z = np.empty(self.n_dof, np.float64)
        check(self.lib.pp_apply_precond(self.h, _ptr(r), _ptr(z)))
        return z

    def block_inverse(self):
        b = np.empty(9 * self.n_nodes, np.float64)
        check(self.lib.pp_get_block_inverse(self.h, _ptr(b)))
        return b

    def pressure_load(self, q0, patches=None):
        pt = np.ascontiguousarray(np.zeros((0, 4)) if patches is None else patches, np.float64)
        F = np.empty(self.n_dof, np.float32)
        check(self.lib.pp_pressure_load(self.h, float(q0), _ptr(pt) if pt.size else None,
                                        int(pt.shape[0]), _ptr(F)))
        return F

    # ---- post-processing (postprocess.ts)
    def _post_x(self, x):
        if x is None:
            return None, None
        x = np.ascontiguousarray(x, np.float64)
        if x.size != self.n_dof:
            raise ValueError(f"x needs {self.n_dof} values")
        return x, _ptr(x)

    def moments(self, x=None):
        """Nodal Mx, My, Mxy (float32) of x, or of the plan's current solution."""
        keep, px = self._post_x(x)
        out = [np.empty(self.n_nodes, np.float32) for _ in range(3)]
        check(self.lib.pp_moments(self.h, px, *[_ptr(a) for a in out]))
        return tuple(out)

    def max_displacement(self, x=None):
        """{maxW, nodeIndex, x, y}: the largest |w| (signed), lowest node on ties."""
        keep, px = self._post_x(x)
        w, node = C.c_double(), C.c_int64()
        cx, cy = C.c_float(), C.c_float()
        check(self.lib.pp_max_displacement(self.h, px, C.byref(w), C.byref(node), C.byref(cx),
                                           C.byref(cy)))
        return {"maxW": w.value, "nodeIndex": node.value, "x": cx.value, "y": cy.value}

    # ---- solve
    @staticmethod
    def options(tolerance=1e-8, max_iterations=20000, precision=_lib.PRECISION_FP64,
                precond=_lib.PRECOND_BLOCK_JACOBI, path=_lib.PATH_AUTO, inner_tolerance=0.0,
                max_refinements=0, flags=0, poly_degree=0):
        return _lib.SolveOpts(tolerance, max_iterations, precision, precond, path,
                              inner_tolerance, max_refinements, flags, poly_degree)

    def set_rhs(self, F):
        F = np.ascontiguousarray(F, np.float32)
        check(self.lib.pp_set_rhs(self.h, _ptr(F)))

    def solve_resident(self, opts):
        st = _lib.SolveStats()
        check(self.lib.pp_solve_resident(self.h, C.byref(opts), C.byref(st)))
        return st

    def solution(self):
        x64 = np.empty(self.n_dof, np.float64)
        x32 = np.empty(self.n_dof, np.float32)
        check(self.lib.pp_get_solution(self.h, _ptr(x64), _ptr(x32)))
        return x64, x32

    def solve(self, F, **kw):
        self.set_rhs(F)
        st = self.solve_resident(self.options(**kw))
        x64, x32 = self.solution()
        return x64, x32, stats_dict(st)


class Net:
    """The learned preconditioner of one structured plan (pp_cnn_*)."""

    def __init__(self, plan, seed=1, channels=(16, 32, 64), bottleneck_dilation=1,
                 rotation_scale=0.3, smoother_weight=0.0, relu=False, fixed_step=False,
                 fcg_window=0, multigrid=False, mg_sweeps=0, mg_coarse_sweeps=0,
                 mg_jacobi=False, unfused_forward=False):
        """multigrid: the smoother around the network is a Galerkin V-cycle
        (mg_sweeps smoothing steps on the finest level before and after the
        coarse correction, 0 = 6 Chebyshev steps or 3 damped block-Jacobi
        sweeps with mg_jacobi; mg_coarse_sweeps on the coarser levels, 0 = 1)
        instead of w B.  unfused_forward: every convolution of the solver's
        forward pass as its own launch (the default runs the first two as
        one)."""
        self.plan, self.lib = plan, plan.lib
        cfg = _lib.CnnConfig((C.c_int32 * 3)(*channels), int(bottleneck_dilation),
                             float(rotation_scale), float(smoother_weight), int(bool(relu)),
                             int(bool(fixed_step)), int(fcg_window), int(bool(multigrid)),
                             int(mg_sweeps), int(mg_coarse_sweeps), int(bool(mg_jacobi)),
                             int(bool(unfused_forward)))
        h = C.c_void_p()
        check(self.lib.pp_cnn_create(plan.h, C.byref(cfg), int(seed), C.byref(h)))
        self.h = h
        self.n_params = int(self.lib.pp_cnn_param_count(h))
        self.omega = float(self.lib.pp_cnn_smoother_weight(h))

    def close(self):
        if self.h:
            self.lib.pp_cnn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.empty(self.n_params, np.float32)
        check(self.lib.pp_cnn_get_params(self.h, _ptr(out)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        check(self.lib.pp_cnn_set_params(self.h, _ptr(p)))

    def apply(self, r):
        r = np.ascontiguousarray(r, np.float64)
        z = np.empty_like(r)
        check(self.lib.pp_cnn_apply(self.h, _ptr(r), _ptr(z)))
        return z

    def forward(self, v):
        """The network alone: N(v) = B v + its learned correction."""
        v = np.ascontiguousarray(v, np.float64)
        u = np.empty_like(v)
        check(self.lib.pp_cnn_forward(self.h, _ptr(v), _ptr(u)))
        return u

    def mg_levels(self):
        """[(nx, ny, omega)] per multigrid level ([] without multigrid)."""
        nx, ny = (C.c_int32 * 16)(), (C.c_int32 * 16)()
        om = (C.c_double * 16)()
        n = self.lib.pp_cnn_mg_levels(self.h, nx, ny, om, 16)
        check(min(n, 0))
        return [(nx[l], ny[l], om[l]) for l in range(n)]

    def mg_apply(self, level, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty_like(x)
        check(self.lib.pp_cnn_mg_apply(self.h, int(level), _ptr(x), _ptr(y)))
        return y

    def mg_cycle(self, b):
        b = np.ascontiguousarray(b, np.float64)
        x = np.empty_like(b)
        check(self.lib.pp_cnn_mg_cycle(self.h, _ptr(b), _ptr(x)))
        return x

    def loss_grad(self, e):
        e = np.ascontiguousarray(e, np.float64)
        loss = C.c_double()
        g = np.empty(self.n_params, np.float32)
        check(self.lib.pp_cnn_loss_grad(self.h, _ptr(e), C.byref(loss), _ptr(g)))
        return loss.value, g

    def gen_error(self, seed):
        e = np.empty(self.plan.n_dof, np.float64)
        check(self.lib.pp_cnn_gen_error(self.h, int(seed), _ptr(e)))
        return e

    def train(self, pairs, epochs=1, seed0=1000, lr=1e-3, lr_final=0.0, max_grad_norm=0.0,
              source=_lib.TRAIN_SYNTHETIC, snapshots=8):
        o = _lib.TrainOpts(int(pairs), int(epochs), int(seed0), int(source), int(snapshots),
                           float(lr), float(lr_final), float(max_grad_norm))
        st = _lib.TrainStats()
        check(self.lib.pp_cnn_train(self.h, C.byref(o), C.byref(st)))
        return {"steps": st.steps, "loss_first": st.loss_first, "loss_last": st.loss_last,
                "seconds": st.seconds,
                "loss_curve": [st.loss_curve[i] for i in range(st.curve_points)]}

    def attach(self):
        check(self.lib.pp_plan_set_learned(self.plan.h, self.h))

    def detach(self):
        check(self.lib.pp_plan_set_learned(self.plan.h, None))


class Batch:
    """Many rectangular-grid plates (Q4 or split-quad DKT) solved by one
    persistent launch (pp_batch_*).  plates: list of dicts with mesh, E, nu,
    t, mask."""

    def __init__(self, ctx, plates):
        self.ctx, self.lib = ctx, ctx.lib
        n = len(plates)
        meshes = (_lib.Mesh * n)()
        mats = (_lib.Material * n)()
        masks = (C.c_void_p * n)()
        self._keep = []
        self.n_dof = []
        for i, p in enumerate(plates):
            m = p["mesh"]
            nodes = np.ascontiguousarray(m["nodes"], np.float32)
            elems = np.ascontiguousarray(m["elements"], np.uint32)
            mask = np.ascontiguousarray(p["mask"], np.uint8)
            self._keep += [nodes, elems, mask]
            meshes[i] = _lib.Mesh(nodes.ctypes.data, elems.ctypes.data, int(m["nodeCount"]),
                                  int(m["elementCount"]), int(m.get("nodesPerElement", 4)),
                                  int(m.get("gridNx", 0)), int(m.get("gridNy", 0)))
            mats[i] = _lib.Material(float(p["E"]), float(p["nu"]), float(p["t"]), None)
            masks[i] = mask.ctypes.data
            self.n_dof.append(3 * int(m["nodeCount"]))
        h = C.c_void_p()
        check(self.lib.pp_batch_create(ctx.h, n, meshes, mats, masks, C.byref(h)))
        self.h = h
        self.n = n

    def close(self):
        if self.h:
            self.lib.pp_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def pressure_load(self, plate, q0, patches=None):
        pt = np.ascontiguousarray(np.zeros((0, 4)) if patches is None else patches, np.float64)
        F = np.empty(self.n_dof[plate], np.float32)
        check(self.lib.pp_batch_pressure_load(self.h, plate, float(q0), _ptr(pt) if pt.size else None,
                                              int(pt.shape[0]), _ptr(F)))
        return F

    def pressure_loads(self, q0s, patches_list):
        """Every plate's load computed on the device straight into the
        batch's right-hand sides (plate p: uniform q0s[p] plus the (k, 4)
        array patches_list[p])."""
        q = np.ascontiguousarray(q0s, np.float64)
        if q.size != self.n or len(patches_list) != self.n:
            raise ValueError(f"need {self.n} loads")
        pts = [np.zeros((0, 4)) if p is None else np.asarray(p, np.float64).reshape(-1, 4)
               for p in patches_list]
        off = np.zeros(self.n + 1, np.int32)
        off[1:] = np.cumsum([len(p) for p in pts])
        flat = np.ascontiguousarray(np.concatenate(pts) if off[-1] else np.zeros(4), np.float64)
        check(self.lib.pp_batch_pressure_loads(self.h, _ptr(q), _ptr(flat), _ptr(off)))

    def set_rhs(self, plate, F):
        F = np.ascontiguousarray(F, np.float32)
        if 0 <= plate < self.n and F.size != self.n_dof[plate]:
            raise ValueError(f"plate {plate} needs {self.n_dof[plate]} load values")
        check(self.lib.pp_batch_set_rhs(self.h, int(plate), _ptr(F)))

    def schedule(self):
        s = _lib.BatchSchedule()
        wg = np.empty(self.n, np.int32)
        check(self.lib.pp_batch_get_schedule(self.h, C.byref(s), _ptr(wg)))
        return {"workgroups_per_cu": s.workgroups_per_cu, "tile_shape": s.tile_shape,
                "balance_iterations": s.balance_iterations,
                "resident_plates": s.resident_plates, "wg_per_plate": wg}

    def set_schedule(self, workgroups_per_cu=2, tile_shape=0, balance_iterations=0,
                     resident_plates=1):
        s = _lib.BatchSchedule(workgroups_per_cu, tile_shape, balance_iterations,
                               resident_plates)
        check(self.lib.pp_batch_set_schedule(self.h, C.byref(s)))

    @staticmethod
    def _stats(st):
        return {k: getattr(st, k) for k, _ in _lib.BatchStats._fields_}

    def autotune(self, tolerance=1e-8, max_iterations=200000):
        o = Plan.options(tolerance=tolerance, max_iterations=max_iterations)
        best = _lib.BatchSchedule()
        st = _lib.BatchStats()
        trials = C.c_int32()
        check(self.lib.pp_batch_autotune(self.h, C.byref(o), C.byref(best), C.byref(st),
                                         C.byref(trials)))
        return self.schedule(), self._stats(st), trials.value

    def solve(self, tolerance=1e-8, max_iterations=200000, flags=0):
        o = Plan.options(tolerance=tolerance, max_iterations=max_iterations, flags=flags)
        st = _lib.BatchStats()
        check(self.lib.pp_batch_solve(self.h, C.byref(o), C.byref(st)))
        return self._stats(st)

    def result(self, plate):
        x64 = np.empty(self.n_dof[plate], np.float64)
        st = _lib.SolveStats()
        check(self.lib.pp_batch_plate_result(self.h, plate, _ptr(x64), None, C.byref(st)))
        return x64, stats_dict(st)


def stats_dict(st):
    n = max(1, st.sweeps)
    return {
        "iterations": st.iterations, "converged": bool(st.converged),
        "breakdown": bool(st.breakdown), "sweeps": st.sweeps,
        "finalResidual": st.final_residual, "trueRelResidual": st.true_rel_residual,
        "deviceMs": st.device_ms,
        "sweepIterations": [st.sweep_iterations[i] for i in range(min(n, _lib.MAX_SWEEPS))],
        "sweepResidual": [st.sweep_residual[i] for i in range(min(n, _lib.MAX_SWEEPS))],
        "phaseUs": [st.phase_us[i] for i in range(6)],
    }


_ctx_cache = {}


def solve_gpu(mesh, material, coloring, F, constrained_dofs, options=None):
    """solveGPU (gpu/solver.ts:80-104) over the HIP library.

    mesh: dict as PlateMesh; material: {E, nu, t}; coloring: {colors: [..]}
    or the packed form; F: float32 load vector; constrained_dofs: iterable of
    DOF indices; options: {tolerance, maxIterations, forceCPU}.
    """
    from .problem import dofs_to_mask

    options = options or {}
    if options.get("forceCPU"):
        raise _lib.PlatePCGError(_lib.PP_ERR_UNSUPPORTED, "forceCPU: this library has no CPU path")
    t0 = time.perf_counter()
    ctx = _ctx_cache.get(0)
    if ctx is None:
        ctx = _ctx_cache[0] = Context(0)
    col = coloring
    if coloring is not None and "colors" in coloring:
        order = np.concatenate([np.asarray(c, np.uint32) for c in coloring["colors"]])
        offs = np.concatenate([[0], np.cumsum([len(c) for c in coloring["colors"]])])
        col = {"elements_by_colour": order, "colour_offsets": offs.astype(np.uint32),
               "colour_count": len(coloring["colors"])}
    mask = dofs_to_mask(int(mesh["nodeCount"]), constrained_dofs)
    plan = Plan(ctx, mesh, material["E"], material["nu"], material["t"], mask, col,
                material.get("element_t"))
    try:
        x64, x32, st = plan.solve(F, tolerance=options.get("tolerance", 1e-6),
                                  max_iterations=options.get("maxIterations", 1000))
    finally:
        plan.close()
    return {
        "solution": x32, "iterations": st["iterations"], "finalResidual": st["finalResidual"],
        "converged": st["converged"], "gpuTimeMs": (time.perf_counter() - t0) * 1e3,
        "usedGPU": True, "solution64": x64, "trueRelResidual": st["trueRelResidual"],
        "deviceMs": st["deviceMs"],
    }
