"""Synthetic NACA 4-digit O-type point clouds in the reference's point layout.

The reference reads its cloud from ``point/point.h5`` (point_preprocessor.F90:215-546):
per point ``original_id, x, y, nx, ny, min_dist, left, right, qtdepth, flag_1, flag_2,
nbhs, conn(1:nbhs)`` with 1-based indices.  No sample cloud ships with the reference,
so this module generates one of the same layout:

* ``n_wall`` points on the airfoil (ring 0), ``n_radial`` rings, the last ring a
  far-field circle of radius ``far_radius`` chords centred on mid-chord;
* geometric radial spacing starting at ``first_spacing`` chords;
* seeded uniform jitter of +-``jitter`` x local spacing (radial and tangential
  components) on interior points only;
* neighbours: the points within ``radius_factor`` x h_i, where h_i = max(radial,
  tangential) local spacing of point i, that see point i through the fluid, capped at
  ``max_nbhs``: picked round-robin over 8 angular sectors round the point, nearest
  first within each (a plain nearest-first cap leaves the half-stencil towards the
  next ring nearly empty where the cloud is strongly stretched), and stored nearest
  first.  A segment that crosses the airfoil, or joins two wall points on opposite
  sides of the thin trailing-edge wedge, is no stencil edge (a stencil reaching across
  the body couples the upper and lower surface flows);
* flags: wall ``flag_1=0, flag_2=1``; interior ``flag_1=1, flag_2=0``; outer ``flag_1=2``;
* ``left``/``right`` run clockwise round both the airfoil and the far-field circle, so
  the normal rot90(right-left) of point_normals.F90:36-48 points into the fluid at the
  wall (wall stencils take delta_n >= 0, generate_connectivity.F90:171-174) and out of
  the domain at the far field (outer stencils take delta_n <= 0, :241-244).

Arrays are returned in ring-major order (index = ring * n_wall + i), which is also the
``original_id`` order.
"""
from __future__ import annotations

import sys
from dataclasses import dataclass, field

import numpy as np


@dataclass
class PointCloud:
    """A point cloud in the reference's per-point layout (0-based arrays, 1-based ids).

    ``conn`` is ``(N, W)`` row-major with 0-based neighbour indices and -1 padding;
    ``left``/``right`` are 0-based (-1 for interior points).
    """

    x: np.ndarray
    y: np.ndarray
    nx: np.ndarray
    ny: np.ndarray
    min_dist: np.ndarray
    left: np.ndarray
    right: np.ndarray
    flag_1: np.ndarray
    flag_2: np.ndarray
    nbhs: np.ndarray
    conn: np.ndarray
    qtdepth: np.ndarray = None
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    @property
    def max_nbhs(self) -> int:
        return int(self.conn.shape[1])

    def __post_init__(self):
        if self.qtdepth is None:
            self.qtdepth = np.zeros(self.x.shape[0], dtype=np.int32)


def naca4_half_thickness(xc: np.ndarray, t: float) -> np.ndarray:
    """NACA 4-digit symmetric half-thickness with a closed trailing edge (-0.1036 x^4)."""
    xc = np.clip(xc, 0.0, 1.0)
    return 5.0 * t * (0.2969 * np.sqrt(xc) - 0.1260 * xc - 0.3516 * xc ** 2
                      + 0.2843 * xc ** 3 - 0.1036 * xc ** 4)


def _geometric_ratio(first: float, n_steps: int, total: float) -> float:
    """Ratio r with first*(r^n - 1)/(r - 1) = total (bisection)."""
    if first * n_steps >= total:
        return 1.0
    lo, hi = 1.0 + 1e-12, 2.0
    while first * (hi ** n_steps - 1.0) / (hi - 1.0) < total:
        hi *= 2.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        s = first * (mid ** n_steps - 1.0) / (mid - 1.0)
        if s < total:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def _wall_normals(ax: np.ndarray, ay: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Outward normals of a clockwise closed polyline (left of the l->r direction)."""
    dx = np.roll(ax, -1) - np.roll(ax, 1)
    dy = np.roll(ay, -1) - np.roll(ay, 1)
    nrm = np.hypot(dx, dy)
    return -dy / nrm, dx / nrm


class _Visibility:
    """Which point pairs see each other through the fluid round a convex wall polygon.

    A pair is hidden when its segment enters the airfoil, or when both are wall points
    whose shorter path along the wall is more than 1.5 x their distance (a chord across
    the body rather than along one surface).  The symmetric NACA 4-digit section is
    convex, so its wall polygon splits at the leading and trailing edges into an upper
    chain y_u(x) (concave) and a lower chain y_l(x) (convex); along a segment p + s (q - p)
    the depth min(y_u(x) - y, y - y_l(x)) is concave in s, and the segment enters the
    body exactly when its maximum (golden-section search) is positive.
    """

    def __init__(self, x: np.ndarray, y: np.ndarray, wall: np.ndarray):
        self.x, self.y = x, y
        ax, ay = x[wall], y[wall]
        elen = np.hypot(np.roll(ax, -1) - ax, np.roll(ay, -1) - ay)
        self.arc = np.concatenate([[0.0], np.cumsum(elen)])
        self.perim = float(self.arc[-1])
        self.pos = np.full(x.shape[0], -1, dtype=np.int64)
        self.pos[wall] = np.arange(wall.shape[0])
        nw = ax.shape[0]
        jl, jt = int(np.argmin(ax)), int(np.argmax(ax))
        c1 = np.arange(jt, jt + ((jl - jt) % nw) + 1) % nw
        c2 = np.arange(jl, jl + ((jt - jl) % nw) + 1) % nw
        chains = []
        for c in (c1, c2):
            o = np.argsort(ax[c], kind="stable")
            chains.append((ax[c][o], ay[c][o]))
        if chains[0][1].mean() < chains[1][1].mean():
            chains.reverse()
        (self.ux, self.uy), (self.lx, self.ly) = chains
        self.xmin, self.xmax = float(ax.min()), float(ax.max())
        self.ymin, self.ymax = float(ay.min()), float(ay.max())
        # lower bound on each point's distance to the wall polygon: nearest wall vertex
        # minus half the longest edge
        from scipy.spatial import cKDTree
        dv, _ = cKDTree(np.column_stack([ax, ay])).query(np.column_stack([x, y]), workers=-1)
        self.dlow = dv - 0.5 * float(elen.max())

    def _depth(self, xs, ys):
        return np.minimum(np.interp(xs, self.ux, self.uy) - ys, ys - np.interp(xs, self.lx, self.ly))

    def visible(self, i: np.ndarray, k: np.ndarray) -> np.ndarray:
        """Boolean mask over the pairs (i[j], k[j])."""
        i = np.asarray(i, dtype=np.int64).ravel()
        k = np.asarray(k, dtype=np.int64).ravel()
        out = np.ones(i.shape[0], dtype=bool)
        px, py, qx, qy = self.x[i], self.y[i], self.x[k], self.y[k]
        pi_, pk_ = self.pos[i], self.pos[k]
        ww = (pi_ >= 0) & (pk_ >= 0)
        if ww.any():  # wall-wall chords across the body
            L = np.hypot(qx[ww] - px[ww], qy[ww] - py[ww])
            fwd = np.mod(self.arc[pk_[ww]] - self.arc[pi_[ww]], self.perim)
            path = np.minimum(fwd, self.perim - fwd)
            out[np.where(ww)[0][path > 1.5 * L]] = False
        # other segments that can reach the body: every point of a segment lies within
        # half its length of an end, so one whose ends are both farther than that from
        # the wall stays outside; the rest also need a bounding box meeting the body's
        half = 0.5 * np.hypot(qx - px, qy - py)
        cand = np.where(~ww & (np.minimum(self.dlow[i], self.dlow[k]) <= half)
                        & (np.maximum(px, qx) >= self.xmin) & (np.minimum(px, qx) <= self.xmax)
                        & (np.maximum(py, qy) >= self.ymin) & (np.minimum(py, qy) <= self.ymax))[0]
        if cand.size == 0:
            return out
        px, py, qx, qy = px[cand], py[cand], qx[cand], qy[cand]
        dx, dy = qx - px, qy - py
        # the part of the segment with xmin <= x <= xmax
        with np.errstate(divide="ignore", invalid="ignore"):
            sa = np.where(dx != 0, (self.xmin - px) / dx, -np.inf)
            sb = np.where(dx != 0, (self.xmax - px) / dx, np.inf)
        s0 = np.clip(np.minimum(sa, sb), 0.0, 1.0)
        s1 = np.clip(np.maximum(sa, sb), 0.0, 1.0)
        vert_out = (dx == 0) & ((px < self.xmin) | (px > self.xmax))
        g = 0.5 * (np.sqrt(5.0) - 1.0)
        lo, hi = s0, s1
        m1, m2 = hi - g * (hi - lo), lo + g * (hi - lo)
        f1 = self._depth(px + m1 * dx, py + m1 * dy)
        f2 = self._depth(px + m2 * dx, py + m2 * dy)
        best = np.maximum(self._depth(px + s0 * dx, py + s0 * dy), self._depth(px + s1 * dx, py + s1 * dy))
        for _ in range(60):
            left = f1 > f2
            best = np.maximum(best, np.maximum(f1, f2))
            hi = np.where(left, m2, hi)
            lo = np.where(left, lo, m1)
            n1 = np.where(left, hi - g * (hi - lo), m2)
            n2 = np.where(left, m1, lo + g * (hi - lo))
            fn = self._depth(px + np.where(left, n1, n2) * dx, py + np.where(left, n1, n2) * dy)
            f1, f2 = np.where(left, fn, f2), np.where(left, f1, fn)
            m1, m2 = n1, n2
        best = np.maximum(best, np.maximum(f1, f2))
        hit = (best > 1e-12) & (s1 > s0) & ~vert_out
        out[cand[hit]] = False
        return out


_SECTORS = 8


def _select_stencils(x, y, nx, ny, hh, tree, vis, max_nbhs: int, radius_factor: float,
                     k_cand: int = 128, chunk: int = 65536, verbose: bool = False):
    """Neighbours of every point: the visible points within radius_factor x h_i, at most
    max_nbhs, picked round-robin over 8 angular sectors of the point's (t, n) frame
    (nearest of each sector first, then the second nearest, ...).  A plain nearest-first
    cap in a strongly stretched region fills the stencil from the point's own ring and
    leaves the half-stencil towards the next ring with a few nearly collinear points;
    balancing by direction keeps every half-plane populated.  The chosen neighbours
    are stored nearest first."""
    N = x.shape[0]
    K = min(k_cand + 1, N)
    conn = np.full((N, max_nbhs), -1, dtype=np.int32)
    nbhs = np.zeros(N, dtype=np.int32)
    for a in range(0, N, chunk):
        b = min(N, a + chunk)
        idx = np.arange(a, b)
        dd, ii = tree.query(np.column_stack([x[a:b], y[a:b]]), k=K, workers=-1)
        ok = (ii != idx[:, None]) & (dd <= radius_factor * hh[a:b, None]) & np.isfinite(dd)
        ii = np.where(ok, ii, 0)
        dx = x[ii] - x[idx, None]
        dy = y[ii] - y[idx, None]
        ds = dx * ny[idx, None] - dy * nx[idx, None]
        dn = dx * nx[idx, None] + dy * ny[idx, None]
        sec = np.floor((np.arctan2(dn, ds) + np.pi) * (_SECTORS / (2.0 * np.pi))).astype(np.int64)
        sec = np.minimum(sec, _SECTORS - 1)
        col = np.arange(K)[None, :]
        checked = np.zeros(ii.shape, dtype=bool)
        # pick, then test the visibility of what was picked; a hidden pick is dropped
        # and the point picks again (a few rounds: only points near the body change)
        while True:
            # rank of each candidate within its sector (candidates come nearest first)
            s_ok = np.where(ok, sec, _SECTORS)
            order = np.argsort(s_ok * K + col, axis=1, kind="stable")
            ss = np.take_along_axis(s_ok, order, axis=1)
            pos = np.broadcast_to(col, ss.shape)
            start = np.where(np.concatenate([np.ones((ss.shape[0], 1), dtype=bool), ss[:, 1:] != ss[:, :-1]],
                                            axis=1), pos, 0)
            start = np.maximum.accumulate(start, axis=1)
            rank = np.empty_like(order)
            np.put_along_axis(rank, order, pos - start, axis=1)
            key = np.where(ok, rank * K + col, np.iinfo(np.int64).max)
            pick = np.argsort(key, axis=1, kind="stable")[:, :max_nbhs]
            pick_ok = np.take_along_axis(ok, pick, axis=1)
            new = pick_ok & ~np.take_along_axis(checked, pick, axis=1)
            if not new.any():
                break
            rows = np.nonzero(new)[0]
            cols = pick[new]
            checked[rows, cols] = True
            hidden = ~vis.visible(idx[rows], ii[rows, cols])
            if not hidden.any():
                break
            ok[rows[hidden], cols[hidden]] = False
        pick = np.sort(np.where(pick_ok, pick, K), axis=1)  # nearest first, invalid last
        valid = pick < K
        chosen = np.take_along_axis(ii, np.minimum(pick, K - 1), axis=1)
        conn[a:b] = np.where(valid, chosen, -1)
        nbhs[a:b] = valid.sum(axis=1)
        if verbose:
            print(f"[cloud] stencils {b}/{N}", file=sys.stderr, flush=True)
    return conn, nbhs


def naca_ogrid(n_wall: int, n_radial: int, thickness: float = 0.12,
               first_spacing: float = 2e-3, far_radius: float = 15.0,
               jitter: float = 0.1, seed: int = 0, max_nbhs: int = 24,
               radius_factor: float = 2.5, smooth_passes: int | None = None,
               spacing: str = "max", verbose: bool = False) -> PointCloud:
    """Generate an O-type cloud round a NACA 00tt airfoil (see module docstring)."""
    from scipy.spatial import cKDTree

    W, R = int(n_wall), int(n_radial)
    if W < 8 or R < 3:
        raise ValueError("need n_wall >= 8 and n_radial >= 3")
    # clockwise parameter starting at the trailing edge: phi decreasing from 0
    phi = -2.0 * np.pi * np.arange(W) / W
    xc = 0.5 + 0.5 * np.cos(phi)
    ax = xc
    ay = np.sign(np.sin(phi)) * naca4_half_thickness(xc, thickness)
    anx, any_ = _wall_normals(ax, ay)
    # marching directions: wall normals smoothed round the airfoil so the trailing-edge
    # fan opens over several columns instead of one
    mx_, my_ = anx.copy(), any_.copy()
    if smooth_passes is None:
        # binomial smoothing width ~ sqrt(passes) columns: keep it a fixed arc fraction
        smooth_passes = int(40 * (W / 160.0) ** 2)
    for _ in range(smooth_passes):
        mx_ = np.roll(mx_, 1) + 2.0 * mx_ + np.roll(mx_, -1)
        my_ = np.roll(my_, 1) + 2.0 * my_ + np.roll(my_, -1)
        nrm = np.hypot(mx_, my_)
        mx_, my_ = mx_ / nrm, my_ / nrm

    cx, cy = 0.5, 0.0
    D = far_radius - 0.5
    r = _geometric_ratio(first_spacing, R - 1, D)
    d = np.zeros(R)
    if r == 1.0:
        d[1:] = np.linspace(first_spacing, D, R - 1)
    else:
        d[1:] = first_spacing * (r ** np.arange(1, R) - 1.0) / (r - 1.0)
    d[-1] = D
    b = (d / D) ** 2
    ux, uy = np.cos(phi), np.sin(phi)
    # P_ij = (1-b_j) (A_i + d_j n_i) + b_j (c + (0.5 + d_j) u_i)
    Px = (1.0 - b)[:, None] * (ax[None, :] + d[:, None] * mx_[None, :]) \
        + b[:, None] * (cx + (0.5 + d)[:, None] * ux[None, :])
    Py = (1.0 - b)[:, None] * (ay[None, :] + d[:, None] * my_[None, :]) \
        + b[:, None] * (cy + (0.5 + d)[:, None] * uy[None, :])
    Px[-1] = cx + far_radius * ux
    Py[-1] = cy + far_radius * uy
    Px[0], Py[0] = ax, ay

    # local spacings on the unjittered grid
    def dist(ax_, ay_, bx_, by_):
        return np.hypot(ax_ - bx_, ay_ - by_)

    h_tan = 0.5 * (dist(Px, Py, np.roll(Px, -1, 1), np.roll(Py, -1, 1))
                   + dist(Px, Py, np.roll(Px, 1, 1), np.roll(Py, 1, 1)))
    h_rad = np.empty_like(Px)
    h_rad[0] = dist(Px[0], Py[0], Px[1], Py[1])
    h_rad[-1] = dist(Px[-1], Py[-1], Px[-2], Py[-2])
    h_rad[1:-1] = 0.5 * (dist(Px[1:-1], Py[1:-1], Px[2:], Py[2:])
                         + dist(Px[1:-1], Py[1:-1], Px[:-2], Py[:-2]))
    h = {"max": np.maximum, "mean": lambda a, b: 0.5 * (a + b), "min": np.minimum}[spacing](h_tan, h_rad)

    if jitter > 0.0 and R > 2:
        rng = np.random.default_rng(seed)
        j1 = rng.uniform(-jitter, jitter, size=(R - 2, W))
        j2 = rng.uniform(-jitter, jitter, size=(R - 2, W))
        # radial unit: towards the next ring; tangential unit: towards the next point
        erx = Px[2:] - Px[:-2]
        ery = Py[2:] - Py[:-2]
        en = np.hypot(erx, ery)
        erx, ery = erx / en, ery / en
        etx = np.roll(Px[1:-1], -1, 1) - np.roll(Px[1:-1], 1, 1)
        ety = np.roll(Py[1:-1], -1, 1) - np.roll(Py[1:-1], 1, 1)
        en = np.hypot(etx, ety)
        etx, ety = etx / en, ety / en
        Px[1:-1] = Px[1:-1] + j1 * h_rad[1:-1] * erx + j2 * h_tan[1:-1] * etx
        Py[1:-1] = Py[1:-1] + j1 * h_rad[1:-1] * ery + j2 * h_tan[1:-1] * ety

    x = Px.reshape(-1).copy()
    y = Py.reshape(-1).copy()
    hh = h.reshape(-1)
    N = x.shape[0]

    flag_1 = np.ones(N, dtype=np.int32)
    flag_1[:W] = 0
    flag_1[(R - 1) * W:] = 2
    flag_2 = np.zeros(N, dtype=np.int32)
    flag_2[:W] = 1
    left = np.full(N, -1, dtype=np.int32)
    right = np.full(N, -1, dtype=np.int32)
    ring_i = np.arange(W)
    for base in (0, (R - 1) * W):
        left[base + ring_i] = base + (ring_i - 1) % W
        right[base + ring_i] = base + (ring_i + 1) % W

    nx = np.zeros(N)
    ny = np.ones(N)
    for base in (0, (R - 1) * W):
        sl = slice(base, base + W)
        lx, ly = x[left[sl]], y[left[sl]]
        rx, ry = x[right[sl]], y[right[sl]]
        vx, vy = 0.5 * ((y[sl] - ly) + (ry - y[sl])), 0.5 * ((x[sl] - lx) + (rx - x[sl]))
        det = np.sqrt(vx * vx + vy * vy)
        nx[sl], ny[sl] = -vx / det, vy / det

    tree = cKDTree(np.column_stack([x, y]))
    vis = _Visibility(x, y, np.arange(W))
    conn, nbhs = _select_stencils(x, y, nx, ny, hh, tree, vis, max_nbhs, radius_factor, verbose=verbose)

    cloud = PointCloud(x=x, y=y, nx=nx, ny=ny, min_dist=np.zeros(N), left=left,
                       right=right, flag_1=flag_1, flag_2=flag_2, nbhs=nbhs, conn=conn)
    repaired = _repair_stencils(cloud, tree, max_nbhs, vis)
    valid = cloud.conn >= 0
    kk = np.where(valid, cloud.conn, 0)
    dist_k = np.where(valid, np.hypot(x[kk] - x[:, None], y[kk] - y[:, None]), np.inf)
    cloud.min_dist = dist_k.min(axis=1)
    if np.any(cloud.nbhs < 2):
        raise RuntimeError("cloud has points with fewer than 2 neighbours")

    meta = dict(generator="naca_ogrid", n_wall=W, n_radial=R, thickness=thickness,
                first_spacing=first_spacing, far_radius=far_radius, jitter=jitter,
                seed=seed, max_nbhs=max_nbhs, radius_factor=radius_factor,
                radial_ratio=r, h_rule=f"{spacing}(radial, tangential) local spacing",
                repaired_points=int(repaired))
    cloud.meta = meta
    return cloud


# half-stencils each point class uses (flux_residual.F90:23-70): bit 0 xpos, 1 xneg,
# 2 ypos, 3 yneg
_REQUIRED_BITS = {0: (0, 1, 3), 1: (0, 1, 2, 3), 2: (0, 1, 2)}
_MIN_HALF = 3
# 1 - (correlation of delta_s, delta_n)^2 of a half-stencil's 2x2 normal matrix
_MIN_COND = 0.05


def _half_quality(ds: np.ndarray, dn: np.ndarray, member: np.ndarray) -> float:
    if member.sum() < _MIN_HALF:
        return 0.0
    a, b = ds[member], dn[member]
    sxx, syy, sxy = (a * a).sum(), (b * b).sum(), (a * b).sum()
    if sxx <= 0 or syy <= 0:
        return 0.0
    return float((sxx * syy - sxy * sxy) / (sxx * syy))


def _rotated(cloud: PointCloud, i: int, idx: np.ndarray):
    delx = cloud.x[idx] - cloud.x[i]
    dely = cloud.y[idx] - cloud.y[i]
    tx, ty = cloud.ny[i], -cloud.nx[i]
    return delx * tx + dely * ty, delx * cloud.nx[i] + dely * cloud.ny[i]


def _repair_stencils(cloud: PointCloud, tree, max_nbhs: int, vis: _Visibility | None = None) -> int:
    """Re-pick neighbours of points whose half-stencils are too small or degenerate.

    The nearest-first pick can leave a half-plane (e.g. behind a trailing-edge fan, or
    outward of a strongly stretched ring) with fewer than 3 points or with nearly
    collinear points, which makes that side's 2x2 least-squares system (nearly)
    singular.  For such points the stencil is re-chosen from a widening set of nearest
    candidates: each needed half-plane first gets its nearest members until it holds
    3 points and 1 - corr(delta_s, delta_n)^2 >= 0.05, then the rest is filled nearest
    first.
    """
    n_bad = 0
    # vectorised screen of every point's half-stencils
    N, W = cloud.conn.shape
    valid = cloud.conn >= 0
    kk = np.where(valid, cloud.conn, 0)
    delx = cloud.x[kk] - cloud.x[:, None]
    dely = cloud.y[kk] - cloud.y[:, None]
    dsa = delx * cloud.ny[:, None] - dely * cloud.nx[:, None]
    dna = delx * cloud.nx[:, None] + dely * cloud.ny[:, None]
    mask = split_stencils(cloud, cloud.nx, cloud.ny)
    bad = np.zeros(N, dtype=bool)
    for b in range(4):
        need = np.isin(cloud.flag_1, [c for c, bb in _REQUIRED_BITS.items() if b in bb])
        mem = ((mask >> b) & 1).astype(bool) & valid
        cnt = mem.sum(axis=1)
        sxx = np.where(mem, dsa * dsa, 0).sum(axis=1)
        syy = np.where(mem, dna * dna, 0).sum(axis=1)
        sxy = np.where(mem, dsa * dna, 0).sum(axis=1)
        with np.errstate(divide="ignore", invalid="ignore"):
            qual = np.where((sxx > 0) & (syy > 0), (sxx * syy - sxy * sxy) / (sxx * syy), 0.0)
        bad |= need & ((cnt < _MIN_HALF) | (qual < _MIN_COND))
    for i in np.where(bad)[0]:
        bits = _REQUIRED_BITS[int(cloud.flag_1[i])]
        kq = 4 * max_nbhs + 1
        while True:
            _, cand = tree.query([cloud.x[i], cloud.y[i]], k=min(kq, cloud.n))
            cand = np.array([int(c) for c in np.atleast_1d(cand) if c != i and c < cloud.n])
            if vis is not None:
                cand = cand[vis.visible(np.full(cand.shape[0], i), cand)]
            m = _mask_row(cloud, i, cand)
            ds, dn = _rotated(cloud, i, cand)
            chosen: list[int] = []
            # scarcest half first; each half may claim at most an equal share
            share = max(_MIN_HALF, max_nbhs // len(bits))
            order = sorted(bits, key=lambda b: int(((m >> b) & 1).sum()))
            for b in order:
                mem = [j for j in range(len(cand)) if (m[j] >> b) & 1]
                taken = [j for j in mem if cand[j] in chosen]
                for j in mem:
                    sel = np.zeros(len(cand), dtype=bool)
                    sel[taken] = True
                    if len(taken) >= _MIN_HALF and (
                            _half_quality(ds, dn, sel) >= _MIN_COND or len(taken) >= share):
                        break
                    if j not in taken:
                        taken.append(j)
                        if cand[j] not in chosen:
                            chosen.append(int(cand[j]))
            filled = all(sum(1 for c in chosen if (m[list(cand).index(c)] >> b) & 1) >= _MIN_HALF
                         for b in bits)
            if (filled and len(chosen) <= max_nbhs) or kq >= cloud.n:
                break
            kq *= 4
        for c in cand:  # fill the rest nearest first
            if len(chosen) >= max_nbhs:
                break
            if int(c) not in chosen:
                chosen.append(int(c))
        chosen = chosen[:max_nbhs]
        chosen.sort(key=lambda c: (cloud.x[c] - cloud.x[i]) ** 2 + (cloud.y[c] - cloud.y[i]) ** 2)
        cloud.conn[i, :] = -1
        cloud.conn[i, :len(chosen)] = chosen
        cloud.nbhs[i] = len(chosen)
        n_bad += 1
    return n_bad


def _mask_row(cloud: PointCloud, i: int, cand: np.ndarray) -> np.ndarray:
    delx = cloud.x[cand] - cloud.x[i]
    dely = cloud.y[cand] - cloud.y[i]
    tx, ty = cloud.ny[i], -cloud.nx[i]
    dels = delx * tx + dely * ty
    deln = delx * cloud.nx[i] + dely * cloud.ny[i]
    m = (dels <= 0).astype(np.uint8) | ((dels >= 0).astype(np.uint8) << 1)
    if cloud.flag_1[i] == 0:
        m |= 8
    elif cloud.flag_1[i] == 2:
        m |= 4
    else:
        m |= ((deln <= 0).astype(np.uint8) << 2) | ((deln >= 0).astype(np.uint8) << 3)
    return m


def split_stencils(cloud: PointCloud, nx: np.ndarray, ny: np.ndarray) -> np.ndarray:
    """Direction mask per connectivity slot, as generate_connectivity.F90:39-256 splits.

    bit 0 xpos (delta_s <= 0), bit 1 xneg (delta_s >= 0), bit 2 ypos (delta_n <= 0),
    bit 3 yneg (delta_n >= 0).  Wall points put every neighbour in yneg (:171-174) and
    outer points every neighbour in ypos (:241-244).  Returns uint8 ``(N, W)``.
    """
    N, W = cloud.conn.shape
    valid = cloud.conn >= 0
    k = np.where(valid, cloud.conn, 0)
    delx = cloud.x[k] - cloud.x[:, None]
    dely = cloud.y[k] - cloud.y[:, None]
    tx, ty = ny[:, None], -nx[:, None]
    dels = delx * tx + dely * ty
    deln = delx * nx[:, None] + dely * ny[:, None]
    m = (dels <= 0).astype(np.uint8) | ((dels >= 0).astype(np.uint8) << 1)
    wall = (cloud.flag_1 == 0)[:, None]
    outer = (cloud.flag_1 == 2)[:, None]
    interior = ~(wall | outer)
    ym = np.where(interior, (deln <= 0).astype(np.uint8) << 2 | (deln >= 0).astype(np.uint8) << 3,
                  np.where(wall, np.uint8(8), np.uint8(4)))
    m = (m | ym).astype(np.uint8)
    return np.where(valid, m, 0).astype(np.uint8)
