"""The device element operators (matrix-free B^T D B on a rectangle, direct
and closed-form), run on the CPU, against the oracle's Q4 Ke (element.ts:258-483)."""
import os
import subprocess

import numpy as np
import pytest

import oracle_api as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "element_op_check.hip")


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    d = tmp_path_factory.mktemp("elop")
    out = str(d / "element_op_check")
    subprocess.check_call(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O2", "-std=c++17",
                           "-I", os.path.join(ROOT, "include"), "-o", out, SRC],
                          stderr=subprocess.DEVNULL)
    return out


def test_element_operators_match_oracle_ke(exe, tmp_path):
    rng = np.random.default_rng(7)
    E, n = 210e9, 40
    recs, kes = [], []
    for k in range(n):
        a, b = rng.uniform(0.01, 2.0, 2)
        t, nu = rng.uniform(0.005, 0.3), rng.uniform(0.0, 0.45)
        D = E * t ** 3 / (12 * (1 - nu ** 2))
        Ds = (5 / 6) * E / (2 * (1 + nu)) * t
        u = rng.standard_normal(12)
        x0, y0 = rng.uniform(-5, 5, 2)
        kes.append(O.ke_q4(np.array([x0, y0, x0 + a, y0, x0 + a, y0 + b, x0, y0 + b]), E, nu, t) @ u)
        recs.append(np.concatenate([[a, b, D, nu, Ds], u]))
    fi, fo = tmp_path / "in.bin", tmp_path / "out.bin"
    np.asarray(recs).tofile(fi)
    subprocess.check_call([exe, str(fi), str(fo)])
    ys = np.fromfile(fo).reshape(n, 2, 12)
    for k in range(n):
        ref = kes[k]
        scale = np.abs(ref).max()
        assert np.abs(ys[k, 0] - ref).max() < 1e-12 * scale, k
        assert np.abs(ys[k, 1] - ref).max() < 1e-12 * scale, k
