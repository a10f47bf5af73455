"""The Node host (gpu-gauntlet_amd/js) over the N-API addon."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JS = os.path.join(ROOT, "gpu-gauntlet_amd", "js", "index.js")


def node(code):
    out = subprocess.run(["node", "-e", code], check=True, capture_output=True, text=True,
                         timeout=120)
    return out.stdout.strip()


def test_addon_loads_and_probes():
    out = node(f"const h = require({json.dumps(JS)}); console.log(JSON.stringify("
               "{avail: h.isGPUAvailable(), fns: Object.keys(h)}));")
    r = json.loads(out)
    assert "solveGPU" in r["fns"] and "PlateSolver" in r["fns"]
    assert isinstance(r["avail"], bool)


def test_host_helpers():
    out = node(f"const h = require({json.dumps(JS)});"
               "const m = h.constrainedMask(3, new Set([0, 4, 5, 6]));"
               "const c = h.packColouring({colors: [new Uint32Array([2, 0]), new Uint32Array([1])]});"
               "console.log(JSON.stringify({m: Array.from(m), o: Array.from(c.elementsByColour),"
               " f: Array.from(c.colourOffsets)}));")
    r = json.loads(out)
    assert r == {"m": [1, 6, 1], "o": [2, 0, 1], "f": [0, 2, 3]}


def test_solve_rejects_without_gpu():
    if json.loads(node(f"console.log(require({json.dumps(JS)}).isGPUAvailable())")):
        pytest.skip("a GPU is present")
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "solve_gpu_c1.js")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "platepcg error" in r.stderr


@pytest.mark.gpu
def test_solve_gpu_through_node():
    out = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "solve_gpu_c1.js")],
                         check=True, capture_output=True, text=True, timeout=120)
    r = json.loads(out.stdout.strip().splitlines()[-1])
    assert r["usedGPU"] and r["converged"] and r["forceCpuRejected"]
    assert r["trueRelResidual"] <= 1e-8
    assert r["maxWRelError"] < 1e-4
    # preconditioner: 'chebyshev' through the same host call
    assert r["chebConverged"] and r["chebTrueRelResidual"] <= 1e-8
    assert r["chebMaxWRelError"] < 1e-4 and r["chebIterations"] < r["iterations"] / 3
    # preconditioner: 'multigrid' (one V-cycle per PCG iteration)
    assert r["mgConverged"] and r["mgTrueRelResidual"] <= 1e-8
    assert r["mgMaxWRelError"] < 1e-4 and r["mgIterations"] < r["iterations"] / 10


@pytest.mark.gpu
def test_solve_plate_through_node():
    """solvePlateGPU: the PlateResult fields from the device (moments of the
    device solution) equal the oracle restatement of postprocess.ts on the
    same FP64 solution; w is the displacements' w; findMaxDisplacement is the
    first largest |w|."""
    import numpy as np
    import oracle_api as O
    from platepcg import problem as P
    out = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "solve_plate_c1.js")],
                         check=True, capture_output=True, text=True, timeout=120)
    r = json.loads(out.stdout.strip().splitlines()[-1])
    info = r["solverInfo"]
    assert info["converged"] and r["trueRelResidual"] <= 1e-8
    assert info["postprocessTimeMs"] >= 0 and info["totalTimeMs"] >= info["solveTimeMs"]
    x = np.array(r["solution64"])
    np.testing.assert_array_equal(np.array(r["w"], np.float32), np.array(r["displacements"], np.float32)[0::3])
    w = P.workload("c1")
    op = O.OracleProblem(w["mesh"], w["E"], w["nu"], w["t"], w["mask"])
    for got, want in zip((r["Mx"], r["My"], r["Mxy"]), op.moments(x)):
        np.testing.assert_allclose(got, want, rtol=2e-7, atol=2e-7 * np.abs(want).max())
    i = int(np.argmax(np.abs(x[0::3])))
    assert r["maxDisp"]["nodeIndex"] == i and r["maxDisp"]["maxW"] == x[3 * i]


@pytest.mark.gpu
def test_concurrent_solves_through_node():
    """Two solveGPU calls in flight together, two solve() calls on one
    PlateSolver together, and destroy() during a solve: every answer equals
    the same solve run alone, bit for bit, and the float32 restatement of
    solver.ts to the harness metric."""
    out = subprocess.run(["node", os.path.join(ROOT, "tests", "js", "concurrent_solves.js")],
                         check=True, capture_output=True, text=True, timeout=180)
    r = json.loads(out.stdout.strip().splitlines()[-1])
    assert r["converged"] and r["loadsDiffer"] and r["deviceMsPositive"]
    assert r["solveGpuSame"] and r["planSame"] and r["destroyedSame"]
    assert r["maxWRelError"] < 1e-4
