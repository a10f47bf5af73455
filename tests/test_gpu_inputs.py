"""Problem inputs built on the device, through the C ABI: the batched plates'
pressure loads (pp_batch_pressure_load / pp_batch_pressure_loads) against
the FP64 oracle's consistent load vector, and the log-normal thickness field
of the 1024 x 1024 workload (pp_lognormal_thickness) against the host's FFT
synthesis (platepcg.problem.gaussian_field, whose statistics the CPU tests
pin)."""
import ctypes as C

import numpy as np
import pytest

import oracle_api as O
from platepcg import Batch, Context, PlatePCGError, lognormal_thickness
from platepcg import problem as P
from platepcg._lib import check

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Context(0)
    yield c
    c.close()


def test_batch_pressure_loads_match_oracle(ctx):
    """Q4 and split-quad DKT plates with random edges and 1-3 Gaussian
    patches: the device's per-plate load equals the oracle's float32 load
    (2e-7), and loading every plate at once straight into the right-hand
    sides solves to the same displacements, bit for bit, as loading each
    plate through the host."""
    plates = P.batch_plates(count=12, seed=21, n_max=48)
    b = Batch(ctx, plates)
    Fs = []
    for i, p in enumerate(plates):
        F = b.pressure_load(i, p["q0"], p["patches"])
        op = O.OracleProblem(p["mesh"], p["E"], p["nu"], p["t"], p["mask"])
        Fo = op.pressure_load(p["q0"], p["patches"])
        np.testing.assert_allclose(F, Fo.astype(np.float32), rtol=2e-7,
                                   atol=1e-9 * np.abs(Fo).max(), err_msg=f"plate {i} {p['kind']}")
        assert not F[1::3].any() and not F[2::3].any()
        b.set_rhs(i, F)
        Fs.append(F)
    st = b.solve(tolerance=1e-8)
    assert st["converged"] == len(plates)
    xs = [b.result(i)[0].copy() for i in range(len(plates))]
    # a different load first, so the second solve cannot reuse the first's RHS
    b.pressure_loads([2e4] * len(plates), [None] * len(plates))
    st = b.solve(tolerance=1e-8)
    assert st["converged"] == len(plates)
    assert any(not np.array_equal(b.result(i)[0], xs[i]) for i in range(len(plates)))
    b.pressure_loads([p["q0"] for p in plates], [p["patches"] for p in plates])
    st = b.solve(tolerance=1e-8)
    assert st["converged"] == len(plates)
    for i in range(len(plates)):
        np.testing.assert_array_equal(b.result(i)[0], xs[i])
    b.close()


def test_batch_pressure_loads_validate_offsets(ctx):
    """Patch offsets that do not start at 0 or decrease are rejected before
    anything reaches the device."""
    plates = P.batch_plates(count=2, seed=3, n_max=20)
    b = Batch(ctx, plates)
    q = np.array([1e4, 1e4])
    pt = np.zeros(8)
    for bad in ([0, 2, 1], [1, 1, 2]):
        off = np.array(bad, np.int32)
        with pytest.raises(PlatePCGError):
            check(b.lib.pp_batch_pressure_loads(b.h, q.ctypes.data_as(C.c_void_p),
                                                pt.ctypes.data_as(C.c_void_p),
                                                off.ctypes.data_as(C.c_void_p)))
    b.close()


@pytest.mark.parametrize("nx,ny,corr,seed", [(8, 8, 5.0, 5), (48, 32, 0.5, 7), (160, 96, 0.5, 9)])
def test_lognormal_thickness_small(ctx, nx, ny, corr, seed):
    """Full-circle taps (correlation longer than the plate) and cut taps:
    the device field equals the FFT synthesis to 1e-12, the thickness to
    one float32 rounding."""
    g, t = lognormal_thickness(ctx, nx, ny, 10.0, 10.0, corr, seed, 0.1, 0.5)
    ref = P.gaussian_field(nx, ny, 10.0, 10.0, corr, seed)
    assert np.abs(g - ref).max() < 1e-12
    tr = 0.1 * np.exp(0.5 * ref)
    assert np.max(np.abs(t - tr) / tr) < 1.2e-7


def test_lognormal_thickness_full_size(ctx):
    """The 1024 x 1024 plate's field (seed 5, correlation 0.5 m, median
    0.1 m, sigma_log 0.5): equal to the host synthesis, zero mean and unit
    variance, and the thickness of the workload itself within one float32
    rounding."""
    g, t = lognormal_thickness(ctx, 1024, 1024, 10.0, 10.0, 0.5, 5, 0.1, 0.5)
    assert abs(g.mean()) < 1e-12 and abs(g.var() - 1) < 1e-12
    ref = P.gaussian_field(1024, 1024, 10.0, 10.0, 0.5, 5)
    assert np.abs(g - ref).max() < 1e-11
    w = P.workload("c5")
    et = w["element_t"].reshape(1024, 1024)
    assert np.max(np.abs(t - et) / et) < 1.2e-7
    assert np.mean(t != et) < 0.01
