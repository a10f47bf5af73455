"""Host-side workload inputs: the SplitMix64 streams and the Gaussian random
thickness field of the 1024 x 1024 variable-thickness plate."""
import numpy as np

from platepcg import problem as P


def test_vectorised_normals_equal_the_scalar_stream():
    for seed in (0, 5, 2**63 + 11):
        r = P.SplitMix64(seed)
        ref = np.array([r.normal() for _ in range(64)])
        np.testing.assert_array_equal(P.splitmix_normals(seed, 64), ref)


def test_gaussian_field_statistics():
    """Unit variance, zero mean, and the squared-exponential correlation
    exp(-d^2 / (2 l^2)) at d = l and d = 2 l (within sampling error)."""
    n, L, l = 256, 10.0, 0.5
    fields = [P.gaussian_field(n, n, L, L, l, seed=s) for s in (5, 6, 7, 8)]
    g = fields[0]
    assert abs(g.mean()) < 1e-12 and abs(g.std() - 1.0) < 1e-12
    h = L / n
    # (the sample autocorrelation of a field 20 correlation lengths wide is
    # biased low at long lags by the mean subtraction, hence the wider 2 l bound)
    for d, tol in ((l / 2, 0.03), (l, 0.04), (2 * l, 0.06)):
        lag = int(round(d / h))
        c = np.mean([0.5 * (np.mean(f[:, :-lag] * f[:, lag:]) + np.mean(f[:-lag] * f[lag:]))
                     for f in fields])
        assert abs(c - np.exp(-0.5 * (lag * h / l) ** 2)) < tol, (d, c)
    # a different seed gives an independent field: the raw streams are
    # uncorrelated, and the sample cross-correlation of two fields is within
    # 3.5 sigma of 0 (sigma^2 = pi l^2 / area for this covariance)
    a, b = P.splitmix_normals(5, 100000), P.splitmix_normals(6, 100000)
    assert abs(np.corrcoef(a, b)[0, 1]) < 0.02
    sigma = np.sqrt(np.pi * l * l / (L * L))
    assert abs(np.mean(g * fields[1])) < 3.5 * sigma


def test_c5_workload_shape():
    w = P.workload("c5")
    t = w["element_t"]
    assert t.shape == (1024 * 1024,) and t.dtype == np.float32
    assert abs(np.median(t) - 0.1) < 0.005
    assert abs(np.std(np.log(t / 0.1)) - 0.5) < 1e-3


def test_split_quad_colouring_is_valid():
    mesh = P.split_quads(P.rect_mesh(9, 7))
    col = P.colour_split_quads(mesh)
    el = mesh["elements"].reshape(-1, 3)
    off = col["colour_offsets"]
    assert off[-1] == mesh["elementCount"]
    for c in range(col["colour_count"]):
        members = col["elements_by_colour"][off[c]:off[c + 1]]
        nodes = el[members].ravel()
        assert len(np.unique(nodes)) == nodes.size, c
