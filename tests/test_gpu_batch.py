"""Batched plates (one persistent launch over Q4 and DKT plates) against the
FP64 oracle and the library's single-plate solves."""
import numpy as np
import pytest

import oracle_api as O
from platepcg import Batch, Context, Plan, PlatePCGError
from platepcg import problem as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def load_batch(ctx, plates):
    b = Batch(ctx, plates)
    Fs = []
    for i, p in enumerate(plates):
        F = b.pressure_load(i, p["q0"], p["patches"])
        b.set_rhs(i, F)
        Fs.append(F)
    return b, Fs


def test_small_batch_matches_fp64_oracle(ctx):
    """12 plates (n <= 40, alternating Q4 / DKT, random materials, edges and
    loads): every plate within 1e-6 of the FP64 oracle (max nodal w error),
    true residual <= 1e-8, and the oracle's PCG iteration count to within
    3 + 1% (these plates have one workgroup per tile and run the
    single-reduction recurrences, whose rounding differs from pcg.ts's order
    and may add a residual-replacement restart)."""
    plates = P.batch_plates(count=12, seed=4, n_max=40)
    b, Fs = load_batch(ctx, plates)
    st = b.solve(tolerance=1e-8)
    assert st["plates"] == 12 and st["converged"] == 12, st
    assert st["worst_true_rel_residual"] <= 1e-8
    for i, p in enumerate(plates):
        op = O.OracleProblem(p["mesh"], p["E"], p["nu"], p["t"], p["mask"])
        F_ref = op.pressure_load(p["q0"], p["patches"])
        np.testing.assert_allclose(Fs[i], F_ref.astype(np.float32), rtol=2e-7, atol=1e-9 * np.abs(F_ref).max())
        x, s = b.result(i)
        xo, so = op.pcg(Fs[i].astype(np.float64), tol=1e-8)
        assert abs(s["iterations"] - so["iterations"]) <= 3 + so["iterations"] // 100, \
            (i, p["kind"], s["iterations"], so["iterations"])
        assert O.max_nodal_w_error(x, xo) < 1e-6, (i, p["kind"])
        assert op.true_rel_residual(x, Fs[i].astype(np.float64)) <= 1.5e-8
        assert s["trueRelResidual"] <= 1e-8
    b.close()


def test_batch_equals_single_plate_solves(ctx):
    """The batched kernel (cell-matrix tiles) and the single-plate paths
    (closed-form tiles for Q4, coloured stored-Ke passes for DKT) reach the
    same displacements."""
    plates = P.batch_plates(count=8, seed=11, n_max=64)
    b, Fs = load_batch(ctx, plates)
    b.solve(tolerance=1e-10)
    for i, p in enumerate(plates):
        mesh = p["mesh"]
        col = P.colour_split_quads(mesh) if p["kind"] == "dkt" else None
        plan = Plan(ctx, mesh, p["E"], p["nu"], p["t"], p["mask"], col)
        x1, _, s1 = plan.solve(Fs[i], tolerance=1e-10, max_iterations=100000)
        x, _ = b.result(i)
        assert O.max_nodal_w_error(x, x1) < 1e-6, (i, p["kind"])
        plan.close()
    b.close()


def test_resident_plates_equal_strided_plates(ctx):
    """Plates dealt one workgroup per tile run the register-resident
    single-reduction path (resident_plates=1) or the strided two-reduction
    path (0): same displacements, both at true residual <= 1e-8."""
    plates = P.batch_plates(count=10, seed=9, n_max=48)
    b, Fs = load_batch(ctx, plates)
    xs = []
    for res in (1, 0):
        b.set_schedule(workgroups_per_cu=2, tile_shape=0, balance_iterations=0, resident_plates=res)
        sc = b.schedule()
        assert sc["resident_plates"] == res
        tiles = np.array([((p["n"] + 1 + 14) // 15) ** 2 for p in plates])
        np.testing.assert_array_equal(sc["wg_per_plate"], tiles)  # one workgroup per tile
        st = b.solve(tolerance=1e-8)
        assert st["converged"] == len(plates) and st["worst_true_rel_residual"] <= 1e-8, st
        xs.append([b.result(i)[0] for i in range(len(plates))])
    for i in range(len(plates)):
        assert O.max_nodal_w_error(xs[0][i], xs[1][i]) < 1e-6, (i, plates[i]["kind"])
    with pytest.raises(PlatePCGError):
        b.set_schedule(resident_plates=2)
    b.close()


def test_autotune_keeps_the_answer_and_is_not_slower(ctx):
    plates = P.batch_plates(count=48, seed=4, n_max=96)
    b, Fs = load_batch(ctx, plates)
    st0 = b.solve(tolerance=1e-8)
    x0 = [b.result(i)[0] for i in range(len(plates))]
    sched, best, trials = b.autotune(tolerance=1e-8)
    print("autotune", trials, "trials:", {k: sched[k] for k in ("workgroups_per_cu", "tile_shape",
          "balance_iterations", "resident_plates")},
          "%.2f ms (default %.2f ms)" % (best["device_ms"], st0["device_ms"]))
    assert trials >= 2
    assert sched["wg_per_plate"].min() >= 1
    assert best["device_ms"] <= 1.05 * st0["device_ms"]
    st1 = b.solve(tolerance=1e-8)
    assert st1["converged"] == len(plates)
    for i in range(len(plates)):
        x, _ = b.result(i)
        assert O.max_nodal_w_error(x, x0[i]) < 1e-7
    b.close()


def test_batch_edge_cases(ctx):
    plates = P.batch_plates(count=3, seed=7, n_max=24)
    b = Batch(ctx, plates)
    # nothing loaded: every plate converges at iteration 0 with x = 0
    st = b.solve()
    assert st["converged"] == 3 and st["max_iterations"] == 0
    assert not b.result(0)[0].any()
    with pytest.raises(PlatePCGError):
        b.set_rhs(5, np.zeros(3, np.float32))
    b.close()
    # a mesh that is not a rectangular grid is refused
    bad = dict(plates[0])
    mesh = dict(bad["mesh"])
    nodes = mesh["nodes"].copy()
    nodes[2 * 5] += 0.01  # one interior column node moved
    mesh["nodes"] = nodes
    bad["mesh"] = mesh
    with pytest.raises(PlatePCGError):
        Batch(ctx, [bad])


def test_full_256_plate_batch(ctx):
    """The 256-plate workload (about 6.2M DOF): one autotuned persistent launch;
    every plate converges with true relative residual <= 1e-8 and agrees with
    the library's single-plate FP64 solve to 1e-4 (max nodal w error)."""
    plates = P.batch_plates()
    b, Fs = load_batch(ctx, plates)
    sched, best, trials = b.autotune(tolerance=1e-8)
    st = b.solve(tolerance=1e-8)
    print("256 plates: %d workgroups, %.1f ms, max iterations %d, worst residual %.2e" %
          (st["workgroups"], st["device_ms"], st["max_iterations"], st["worst_true_rel_residual"]))
    assert st["converged"] == 256 and st["worst_true_rel_residual"] <= 1e-8, st
    worst = 0.0
    for i, p in enumerate(plates):
        mesh = p["mesh"]
        col = P.colour_split_quads(mesh) if p["kind"] == "dkt" else None
        plan = Plan(ctx, mesh, p["E"], p["nu"], p["t"], p["mask"], col)
        x1, _, s1 = plan.solve(Fs[i], tolerance=1e-8, max_iterations=200000)
        x, _ = b.result(i)
        worst = max(worst, O.max_nodal_w_error(x, x1))
        plan.close()
    print("worst max-nodal w error vs single-plate FP64 solves: %.2e" % worst)
    assert worst < 1e-4
    b.close()


def test_iteration_budget_spans_restarts(ctx):
    """max_iterations bounds every plate's total over the first launch and the
    residual-replacement restarts, and a plate whose true residual is still
    above the tolerance is not reported converged."""
    plates = P.batch_plates(count=6, seed=4, n_max=40)
    b, _ = load_batch(ctx, plates)
    full = b.solve(tolerance=1e-8)
    cap = max(2, full["max_iterations"] // 2)
    st = b.solve(tolerance=1e-8, max_iterations=cap)
    assert st["max_iterations"] <= cap
    for i in range(len(plates)):
        _, s = b.result(i)
        assert s["iterations"] <= cap
        assert not s["converged"] or s["trueRelResidual"] <= 1e-8
    assert st["converged"] < len(plates)
    b.close()
