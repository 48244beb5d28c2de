"""Point-cloud generator: reference layout, orientation and stencil quality."""
import numpy as np

from mfcfd_amd.cloud import naca_ogrid, split_stencils


def test_layout_and_flags():
    c = naca_ogrid(64, 20, seed=2)
    W, R = 64, 20
    assert c.n == W * R
    assert (c.flag_1[:W] == 0).all() and (c.flag_1[-W:] == 2).all() and (c.flag_1[W:-W] == 1).all()
    assert (c.flag_2[:W] == 1).all() and (c.flag_2[W:] == 0).all()
    assert c.conn.shape[1] == 24 and (c.nbhs <= 24).all() and (c.nbhs >= 3).all()
    valid = c.conn >= 0
    assert (valid.sum(1) == c.nbhs).all()
    assert (c.conn[valid] != np.repeat(np.arange(c.n), 24).reshape(c.n, 24)[valid]).all()
    # min_dist = nearest neighbour distance
    k = np.where(valid, c.conn, 0)
    d = np.where(valid, np.hypot(c.x[k] - c.x[:, None], c.y[k] - c.y[:, None]), np.inf)
    assert np.allclose(c.min_dist, d.min(1))


def test_clockwise_orientation_normals_into_fluid():
    """rot90(right - left) points away from the airfoil at the wall and away from the
    domain at the far field (point_normals.F90:36-48 convention)."""
    c = naca_ogrid(64, 20, seed=2)
    W = 64
    wall = slice(0, W)
    out = slice(c.n - W, c.n)
    cx = 0.5
    for sl, sign in ((wall, 1.0), (out, 1.0)):
        rx, ry = c.x[sl] - cx, c.y[sl]
        assert (sign * (c.nx[sl] * rx + c.ny[sl] * ry) > 0).mean() > 0.95
    assert np.allclose(np.hypot(c.nx, c.ny), 1.0)


def test_half_stencils_filled():
    for W, R, kw in ((64, 20, {}), (48, 14, dict(first_spacing=0.01, far_radius=4.0))):
        c = naca_ogrid(W, R, seed=1, **kw)
        m = split_stencils(c, c.nx, c.ny)
        cnt = [((m >> b) & 1).sum(1) for b in range(4)]
        need = {0: (0, 1, 3), 1: (0, 1, 2, 3), 2: (0, 1, 2)}
        for cls, bits in need.items():
            sel = c.flag_1 == cls
            for b in bits:
                assert cnt[b][sel].min() >= 3


def test_seeded_and_deterministic():
    a = naca_ogrid(40, 12, seed=9)
    b = naca_ogrid(40, 12, seed=9)
    c = naca_ogrid(40, 12, seed=10)
    assert np.array_equal(a.x, b.x) and np.array_equal(a.conn, b.conn)
    assert not np.array_equal(a.x, c.x)
    # jitter only moves interior points
    assert np.array_equal(a.x[:40], c.x[:40]) and np.array_equal(a.x[-40:], c.x[-40:])


def test_c3_tangent_fixture_cloud():
    """The stored C3 tangent belongs to the cloud the generator makes today."""
    import json
    import os
    import sys
    from mfcfd_amd.cloud import naca_ogrid
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    sys.path.insert(0, here)
    import make_c3_tangent as M
    with open(os.path.join(here, "c3_tangent.json")) as f:
        ref = json.load(f)
    c = naca_ogrid(320, 125, seed=0)
    assert c.n == ref["points"]
    assert M.cloud_digest(c) == ref["cloud_sha256"]
