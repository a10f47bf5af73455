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
