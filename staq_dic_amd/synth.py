"""Synthetic speckle image pairs with analytic displacement fields.

All arrays use the reference's layout (func_quadtree/ReadImageQuadtree.m:78,
the images are transposed on load): an image is an ``(H, W)`` array indexed
``img[x-1, y-1]`` with ``x`` the horizontal (dim-1) and ``y`` the vertical
(dim-2) coordinate, 1-based.  ``np.asfortranarray`` of such an array has the
exact memory of the MATLAB variable.

Deformed images follow the usual DIC convention: a material point ``X`` moves to
``x = X + u(X)``, so ``g(x) = f(X)``.  ``g`` is produced by inverting the map per
pixel (Newton) and sampling ``f`` with the Keys (a = -0.5) bicubic of
``ba_interp2.cpp`` at ``X``, then re-quantising.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

__all__ = [
    "speckle_image", "bicubic", "warp_image", "normalize", "AffineField", "KirschField",
    "SinusoidField", "StretchBumpField", "Frame", "make_frame", "plain_grid_nodes",
    "quadtree_nodes", "quadtree_radial", "IncrementalField", "config_plain", "config_hole",
    "config_adjoint", "config_large", "config_multi",
]


def speckle_image(H: int, W: int, n: int, seed: int, bits: int = 8) -> np.ndarray:
    """Sum of ``n`` Gaussian speckles (sigma U[1.5, 3] px, peak U[0.3, 1]) with
    centres uniform over [1, H] x [1, W], min-max normalised and quantised to
    ``bits`` bits.  Returns float64 ``(H, W)``."""
    rng = np.random.Generator(np.random.PCG64(seed))
    cx = rng.uniform(1, H, n)
    cy = rng.uniform(1, W, n)
    sig = rng.uniform(1.5, 3.0, n)
    pk = rng.uniform(0.3, 1.0, n)
    img = np.zeros((H, W))
    rad = 12
    offs = np.arange(-rad, rad + 1)
    for k in range(n):
        ix = int(round(cx[k]))
        iy = int(round(cy[k]))
        xs = ix + offs
        ys = iy + offs
        vx = (xs >= 1) & (xs <= H)
        vy = (ys >= 1) & (ys <= W)
        xs, ys = xs[vx], ys[vy]
        gx = np.exp(-((xs - cx[k]) ** 2) / (2 * sig[k] ** 2))
        gy = np.exp(-((ys - cy[k]) ** 2) / (2 * sig[k] ** 2))
        img[xs[0] - 1:xs[-1], ys[0] - 1:ys[-1]] += pk[k] * np.outer(gx, gy)
    img -= img.min()
    img /= img.max()
    q = (1 << bits) - 1
    return np.round(q * img)


def _keys(t):
    t2 = t * t
    t3 = t2 * t
    return (0.5 * (-t + 2 * t2 - t3), 0.5 * (2 - 5 * t2 + 3 * t3),
            0.5 * (t + 4 * t2 - 3 * t3), 0.5 * (-t2 + t3))


def bicubic(F: np.ndarray, xs: np.ndarray, ys: np.ndarray) -> np.ndarray:
    """Keys bicubic of ``F`` (H x W, ``F[x-1, y-1]``) at 1-based dim-1 coordinate
    ``xs`` and dim-2 coordinate ``ys``, taps clamped to the array
    (ba_interp2.cpp:60-130)."""
    H, W = F.shape
    xf = np.floor(xs)
    yf = np.floor(ys)
    wx = _keys(xs - xf)
    wy = _keys(ys - yf)
    ix = xf.astype(np.int64) - 1
    iy = yf.astype(np.int64) - 1
    out = np.zeros(np.broadcast(xs, ys).shape)
    for a in range(4):
        ra = np.clip(ix - 1 + a, 0, H - 1)
        s = np.zeros_like(out)
        for b in range(4):
            cb = np.clip(iy - 1 + b, 0, W - 1)
            s += wy[b] * F[ra, cb]
        out += wx[a] * s
    return out


# ---------------------------------------------------------------------------
# Displacement fields: u(X, Y) -> (u, v) with Jacobian, coordinates 1-based.
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class AffineField:
    """u = u0 + a11 x + a12 y, v = v0 + a21 x + a22 y."""
    u0: float
    v0: float
    a11: float
    a12: float
    a21: float
    a22: float

    def __call__(self, X, Y):
        return (self.u0 + self.a11 * X + self.a12 * Y, self.v0 + self.a21 * X + self.a22 * Y)

    def jac(self, X, Y):
        o = np.ones_like(np.asarray(X, dtype=float))
        return self.a11 * o, self.a12 * o, self.a21 * o, self.a22 * o


@dataclasses.dataclass
class KirschField:
    """Plane-stress displacement of an infinite plate with a circular hole of
    radius ``a`` centred at (xc, yc) under uniaxial x-tension (Kirsch), scaled
    by ``scale``.  With kappa = (3 - nu)/(1 + nu):
    u_r = r((k-1)/2 + cos2t) + a^2/r (1 + (1+k) cos2t) - a^4/r^3 cos2t,
    u_t = ((1-k) a^2/r - r - a^4/r^3) sin2t, both times sigma/(4 mu)."""
    xc: float
    yc: float
    a: float
    scale: float = 1.0
    nu: float = 0.3

    def _raw(self, X, Y):
        k = (3 - self.nu) / (1 + self.nu)
        dx = X - self.xc
        dy = Y - self.yc
        r = np.maximum(np.hypot(dx, dy), 0.5 * self.a)
        t = np.arctan2(dy, dx)
        c2, s2 = np.cos(2 * t), np.sin(2 * t)
        a2, a4 = self.a ** 2, self.a ** 4
        ur = r * ((k - 1) / 2 + c2) + a2 / r * (1 + (1 + k) * c2) - a4 / r ** 3 * c2
        ut = ((1 - k) * a2 / r - r - a4 / r ** 3) * s2
        ct, st = np.cos(t), np.sin(t)
        return ur * ct - ut * st, ur * st + ut * ct

    def __call__(self, X, Y):
        u, v = self._raw(X, Y)
        return self.scale * u, self.scale * v

    def jac(self, X, Y, h=1e-4):
        up, vp = self(X + h, Y)
        um, vm = self(X - h, Y)
        uq, vq = self(X, Y + h)
        un, vn = self(X, Y - h)
        return (up - um) / (2 * h), (uq - un) / (2 * h), (vp - vm) / (2 * h), (vq - vn) / (2 * h)


@dataclasses.dataclass
class SinusoidField:
    """u = A sin(2 pi x / Lx), v = B cos(2 pi y / Ly)."""
    A: float = 4.0
    Lx: float = 512.0
    B: float = 3.0
    Ly: float = 768.0

    def __call__(self, X, Y):
        return self.A * np.sin(2 * np.pi * X / self.Lx), self.B * np.cos(2 * np.pi * Y / self.Ly)

    def jac(self, X, Y):
        z = np.zeros_like(np.asarray(X, dtype=float))
        return (self.A * 2 * np.pi / self.Lx * np.cos(2 * np.pi * X / self.Lx), z, z,
                -self.B * 2 * np.pi / self.Ly * np.sin(2 * np.pi * Y / self.Ly))


@dataclasses.dataclass
class StretchBumpField:
    """Affine stretch ``s`` about the centre plus a Gaussian bump of amplitude
    ``amp`` px and width ``sig`` px in u."""
    xc: float
    yc: float
    s: float
    amp: float
    sig: float = 300.0

    def __call__(self, X, Y):
        g = np.exp(-((X - self.xc) ** 2 + (Y - self.yc) ** 2) / (2 * self.sig ** 2))
        return self.s * (X - self.xc) + self.amp * g, np.zeros_like(g)

    def jac(self, X, Y):
        g = np.exp(-((X - self.xc) ** 2 + (Y - self.yc) ** 2) / (2 * self.sig ** 2))
        z = np.zeros_like(g)
        return (self.s - self.amp * g * (X - self.xc) / self.sig ** 2,
                -self.amp * g * (Y - self.yc) / self.sig ** 2, z, z)


@dataclasses.dataclass
class IncrementalField:
    """Displacement between two frames of a sequence, in the earlier frame's
    coordinates: frame j shows the pattern moved by phi_j(X) = X + u_j(X), so a
    point x of frame ``a`` sits at X = phi_a^-1(x) in the pattern and at
    phi_b(X) in frame ``b``; u(x) = phi_b(phi_a^-1(x)) - x and
    grad u = (I + grad u_b)(I + grad u_a)^-1 - I at X."""
    a: object
    b: object

    def _pre(self, X, Y):
        X = np.asarray(X, dtype=float)
        Y = np.asarray(Y, dtype=float)
        return invert_map(self.a, X, Y)

    def __call__(self, X, Y):
        PX, PY = self._pre(X, Y)
        ub, vb = self.b(PX, PY)
        return PX + ub - X, PY + vb - Y

    def jac(self, X, Y):
        PX, PY = self._pre(X, Y)
        a11, a12, a21, a22 = self.a.jac(PX, PY)
        b11, b12, b21, b22 = self.b.jac(PX, PY)
        A11, A12, A21, A22 = 1 + a11, a12, a21, 1 + a22
        B11, B12, B21, B22 = 1 + b11, b12, b21, 1 + b22
        det = A11 * A22 - A12 * A21
        I11, I12, I21, I22 = A22 / det, -A12 / det, -A21 / det, A11 / det
        return (B11 * I11 + B12 * I21 - 1, B11 * I12 + B12 * I22,
                B21 * I11 + B22 * I21, B21 * I12 + B22 * I22 - 1)


def invert_map(field, x, y, iters: int = 30):
    """Solve X + u(X) = x per pixel by Newton's method."""
    X = np.array(x, dtype=float)
    Y = np.array(y, dtype=float)
    u, v = field(X, Y)
    X = x - u
    Y = y - v
    for _ in range(iters):
        u, v = field(X, Y)
        a11, a12, a21, a22 = field.jac(X, Y)
        rx = X + u - x
        ry = Y + v - y
        j11, j12, j21, j22 = 1 + a11, a12, a21, 1 + a22
        det = j11 * j22 - j12 * j21
        dX = (j22 * rx - j12 * ry) / det
        dY = (-j21 * rx + j11 * ry) / det
        X = X - dX
        Y = Y - dY
        if max(np.abs(dX).max(), np.abs(dY).max()) < 1e-10:
            break
    return X, Y


def warp_image(f: np.ndarray, field, bits: int = 8, hole=None):
    """Deformed image g(x) = f(X), x = X + u(X), by inverse mapping and Keys
    bicubic of ``f``; re-quantised to ``bits``.  ``hole`` = (xc, yc, a) masks
    points whose pre-image lies in the hole.  Returns (g, gmask)."""
    H, W = f.shape
    x, y = np.meshgrid(np.arange(1, H + 1, dtype=float), np.arange(1, W + 1, dtype=float),
                       indexing="ij")
    X, Y = invert_map(field, x, y)
    g = bicubic(f, X, Y)
    q = (1 << bits) - 1
    g = np.clip(np.round(g), 0, q)
    gmask = np.ones_like(g)
    if hole is not None:
        xc, yc, a = hole
        gmask = (np.hypot(X - xc, Y - yc) > a).astype(float)
    return g, gmask


def normalize(img: np.ndarray, roi_x=None, roi_y=None) -> np.ndarray:
    """(I - mean_ROI) / std_ROI with N-1 (func/funNormalizeImg.m:36-39)."""
    H, W = img.shape
    rx = roi_x or (1, H)
    ry = roi_y or (1, W)
    t = img[rx[0] - 1:rx[1], ry[0] - 1:ry[1]]
    return (img - t.mean()) / t.std(ddof=1)


@dataclasses.dataclass
class Frame:
    """One reference/deformed pair in the reference's layout.  ``f`` and ``g``
    are the normalised, masked images (main_ALDIC_Quadtree_v2.m:57-64),
    rounded to float32 so a float32 device copy is exact."""
    f: np.ndarray
    g: np.ndarray
    mask: np.ndarray
    field: object
    hole: tuple | None = None

    @property
    def shape(self):
        return self.f.shape


def _f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def make_frame(H, W, n_speckles, seed, field, bits=8, hole=None) -> Frame:
    f_raw = speckle_image(H, W, n_speckles, seed, bits)
    if hole is not None:
        xc, yc, a = hole
        xx, yy = np.meshgrid(np.arange(1, H + 1), np.arange(1, W + 1), indexing="ij")
        mask = (np.hypot(xx - xc, yy - yc) > a).astype(float)
    else:
        mask = np.ones((H, W))
    g_raw, gmask = warp_image(f_raw, field, bits, hole)
    f = _f32(normalize(f_raw) * mask)
    g = _f32(normalize(g_raw) * gmask)
    return Frame(f=f, g=g, mask=mask, field=field, hole=hole)


def plain_grid_nodes(H, W, winsize, step, margin=None):
    """Uniform node grid of main_ALDIC_Quadtree_v2.m:108-111:
    max(4 + w/2, ROI_lo) : step : min(size - w/2 - 3, ROI_hi)."""
    lo = 4 + winsize // 2 if margin is None else margin
    xs = np.arange(lo, H - winsize // 2 - 3 + 1, step)
    ys = np.arange(lo, W - winsize // 2 - 3 + 1, step)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    return np.stack([X.ravel(order="F"), Y.ravel(order="F")], axis=1).astype(float)


def quadtree_nodes(H, W, hole, leaf_max=32, leaf_min=8, band=(450, 150), margin=40):
    """Quadtree subset list refined at a circular hole edge.

    Cells of ``leaf_max`` px tile [margin, size - margin]; a cell within
    ``band[0]`` px of the hole edge is split to ``leaf_max/2`` and within
    ``band[1]`` px to ``leaf_min``.  Cells wholly inside the hole are dropped.
    Nodes are the corners of the remaining leaves outside the hole; each node
    gets window w = clamp(2 * smallest adjacent leaf, 16, 64).
    Returns (coords N x 2, winsize N)."""
    xc, yc, a = hole
    leaves = []

    def dist_to_edge(x0, y0, s):
        # distance range of the cell to the hole centre
        nx = min(max(xc, x0), x0 + s)
        ny = min(max(yc, y0), y0 + s)
        dmin = math.hypot(nx - xc, ny - yc)
        dmax = max(math.hypot(cx - xc, cy - yc) for cx in (x0, x0 + s) for cy in (y0, y0 + s))
        return dmin, dmax

    def split(x0, y0, s):
        dmin, dmax = dist_to_edge(x0, y0, s)
        if dmax <= a:
            return  # inside the hole
        edge = min(abs(dmin - a), abs(dmax - a)) if not (dmin <= a <= dmax) else 0.0
        target = leaf_max
        if edge <= band[1]:
            target = leaf_min
        elif edge <= band[0]:
            target = max(leaf_min, leaf_max // 2)
        if s > target:
            h = s // 2
            for dx in (0, h):
                for dy in (0, h):
                    split(x0 + dx, y0 + dy, h)
        else:
            leaves.append((x0, y0, s))

    for x0 in range(margin, H - margin - leaf_max + 1, leaf_max):
        for y0 in range(margin, W - margin - leaf_max + 1, leaf_max):
            split(x0, y0, leaf_max)
    best = {}
    for (x0, y0, s) in leaves:
        for cx in (x0, x0 + s):
            for cy in (y0, y0 + s):
                if math.hypot(cx - xc, cy - yc) <= a:
                    continue
                k = (cx, cy)
                best[k] = min(best.get(k, 1 << 30), s)
    keys = sorted(best.keys(), key=lambda k: (k[1], k[0]))
    coords = np.array(keys, dtype=float)
    win = np.array([min(64, max(16, 2 * best[k])) for k in keys], dtype=np.int64)
    return coords, win


def quadtree_radial(H, W, centre, radii, leaves, margin=64, wmin=8, wmax=64):
    """Quadtree subset list refined towards a point.

    Cells of ``leaves[-1]`` px tile [margin, size - margin]; a cell whose
    nearest point lies within ``radii[i]`` px of ``centre`` is split down to
    ``leaves[i]`` (``radii`` increasing, ``leaves`` increasing, one more leaf
    size than radii).  Nodes are the leaf corners; each node gets window
    w = clamp(2 * smallest adjacent leaf, wmin, wmax).
    Returns (coords N x 2, winsize N)."""
    xc, yc = centre
    top = leaves[-1]
    best = {}

    def target(x0, y0, s):
        nx = min(max(xc, x0), x0 + s)
        ny = min(max(yc, y0), y0 + s)
        d = math.hypot(nx - xc, ny - yc)
        for r, leaf in zip(radii, leaves):
            if d <= r:
                return leaf
        return top

    stack = [(x0, y0, top) for x0 in range(margin, H - margin - top + 1, top)
             for y0 in range(margin, W - margin - top + 1, top)]
    while stack:
        x0, y0, s = stack.pop()
        if s > target(x0, y0, s):
            h = s // 2
            stack.extend([(x0, y0, h), (x0 + h, y0, h), (x0, y0 + h, h), (x0 + h, y0 + h, h)])
            continue
        for cx in (x0, x0 + s):
            for cy in (y0, y0 + s):
                best[(cx, cy)] = min(best.get((cx, cy), 1 << 30), s)
    keys = sorted(best.keys(), key=lambda k: (k[1], k[0]))
    coords = np.array(keys, dtype=float)
    win = np.array([min(wmax, max(wmin, 2 * best[k])) for k in keys], dtype=np.int64)
    return coords, win


def truth_at(field, coords):
    """Analytic (u, v) and displacement gradient at nodes: P = [u_x v_x u_y v_y u v]."""
    X = coords[:, 0]
    Y = coords[:, 1]
    u, v = field(X, Y)
    a11, a12, a21, a22 = field.jac(X, Y)
    return np.stack([a11, a21, a12, a22, u, v], axis=1)


def config_plain(seed=1001, H=512, W=512, n_speckles=3000, winsize=32, step=16):
    """Plain method, 512^2 affine pair, uniform grid w=32 step 16 (900 nodes)."""
    field = AffineField(3.4, -1.7, 0.02, 0.01, 0.005, -0.01)
    fr = make_frame(H, W, n_speckles, seed, field)
    coords = plain_grid_nodes(H, W, winsize, step, margin=20)
    coords = coords[(coords[:, 0] <= H - 28) & (coords[:, 1] <= W - 28)]
    return fr, coords, np.full(len(coords), winsize, dtype=np.int64)


def config_hole(seed=1002, H=1024, W=1024, n_speckles=12000, max_disp=8.0, radius=150.0,
                window_rule="leaf", winsize=32, **qt):
    """Quadtree with a centred hole of radius ``radius`` px, Kirsch field
    scaled to ``max_disp`` px, leaves 8-32 px refined at the hole edge.

    ``window_rule`` "leaf": each node's window is tied to its smallest
    adjacent leaf, clamp(2 leaf, 16, 64).  "reference": every node gets the
    one window ``winsize`` as the reference's quadtree drivers do
    (main_ALDIC_Quadtree_v2.m:361, LocalICGNQuadtree.m:85 passes
    DICpara.winsize); windows then differ only where the IC-GN step grows
    them at the mask (funICGNQuadtree.m:66-88,139-145)."""
    hole = (H / 2 + 0.5, W / 2 + 0.5, float(radius))
    k = KirschField(hole[0], hole[1], hole[2])
    xx, yy = np.meshgrid(np.arange(1, H + 1, 4.0), np.arange(1, W + 1, 4.0), indexing="ij")
    out = np.hypot(xx - hole[0], yy - hole[1]) > hole[2]
    u, v = k(xx[out], yy[out])
    k.scale = max_disp / float(np.hypot(u, v).max())
    fr = make_frame(H, W, n_speckles, seed, k, hole=hole)
    coords, win = quadtree_nodes(H, W, hole, **qt)
    if window_rule == "reference":
        win = np.full(len(coords), winsize, dtype=np.int64)
    elif window_rule != "leaf":
        raise ValueError(f"unknown window rule {window_rule!r}")
    return fr, coords, win


def config_adjoint(seed=1003, H=2048, W=2048, n_speckles=48000, step=16, winsize=16):
    """Sinusoidal field u = 4 sin(2 pi x/512), v = 3 cos(2 pi y/768) on a
    ``step``-px node grid."""
    field = SinusoidField()
    fr = make_frame(H, W, n_speckles, seed, field)
    m = max(winsize // 2 + 16, 32)
    xs = np.arange(m, H - m + 1, step)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    coords = np.stack([X.ravel(order="F"), Y.ravel(order="F")], axis=1).astype(float)
    return fr, coords, np.full(len(coords), winsize, dtype=np.int64)


def stretch_bump_frame_field(k, H, W, n_frames=40):
    """Cumulative field of frame k of the large-deformation sequence: affine
    x-stretch 0.2 k/n_frames about the centre plus a centred Gaussian bump of
    40 k/n_frames px (sigma 300 px) in u."""
    return StretchBumpField(H / 2 + 0.5, W / 2 + 0.5, 0.2 * k / n_frames, 40.0 * k / n_frames, 300.0)


def config_large(k=40, seed=1004, H=4096, W=4096, n_speckles=192000, bits=16,
                 radii=(460, 900, 1500, 2100), leaves=(4, 8, 16, 32, 64), margin=64, wmin=16):
    """Large deformation: 16-bit speckle pattern (``n_speckles`` per 4096^2,
    seed 1004) shown in frames 0..40 under the cumulative stretch-and-bump
    field; returns the incremental pair (frame k-1, frame k) as a Frame whose
    ``field`` is the incremental displacement, and a quadtree refined towards
    the bump (leaves 4-64 px, windows clamp(2 leaf, wmin, 64))."""
    f0 = speckle_image(H, W, n_speckles, seed, bits)
    fa = stretch_bump_frame_field(k - 1, H, W)
    fb = stretch_bump_frame_field(k, H, W)
    ga, _ = warp_image(f0, fa, bits)
    gb, _ = warp_image(f0, fb, bits)
    del f0
    mask = np.ones((H, W))
    fr = Frame(f=_f32(normalize(ga)), g=_f32(normalize(gb)), mask=mask,
               field=IncrementalField(fa, fb))
    coords, win = quadtree_radial(H, W, (H / 2 + 0.5, W / 2 + 0.5), radii, leaves, margin=margin,
                                  wmin=wmin)
    return fr, coords, win


MULTI_SIZES = (512, 1024, 1536, 2048)


def config_multi(n_pairs=16, seed0=5000, sizes=MULTI_SIZES, density=3000 / 512 ** 2):
    """Multi-instance batch: pair j (seed seed0 + j) has a square size drawn
    from ``sizes``, a random affine field about its centre (translation
    |t| <= 10 px, each gradient component in [-0.05, 0.05]) and its own
    quadtree refined towards a random point (leaves 8-32 px, windows
    16-64 px).  Returns a list of (Frame, coords, winsize)."""
    out = []
    for j in range(n_pairs):
        rng = np.random.Generator(np.random.PCG64(seed0 + j))
        n = int(rng.choice(np.asarray(sizes)))
        ang = rng.uniform(0, 2 * np.pi)
        tr = 10.0 * math.sqrt(rng.uniform())
        a11, a12, a21, a22 = rng.uniform(-0.05, 0.05, 4)
        c = n / 2 + 0.5
        t = (tr * math.cos(ang), tr * math.sin(ang))
        field = AffineField(t[0] - a11 * c - a12 * c, t[1] - a21 * c - a22 * c, a11, a12, a21, a22)
        fr = make_frame(n, n, int(round(density * n * n)), seed0 + j, field)
        # the pattern's largest displacement decides the margin (window 64 + bicubic support)
        dmax = 10.0 + 0.1 * n / 2 * math.sqrt(2)
        margin = int(math.ceil(dmax)) + 40
        p = rng.uniform(0.25 * n, 0.75 * n, 2)
        coords, win = quadtree_radial(n, n, (p[0], p[1]), (0.2 * n, 0.45 * n), (8, 16, 32),
                                      margin=margin, wmin=16)
        out.append((fr, coords, win))
    return out
