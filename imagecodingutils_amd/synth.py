"""Seeded synthetic luma frames for motion-estimation workloads.

Texture: ``clamp(128 + sum_k A_k sin(2*pi*<f_k, (x,y)> + phi_k) + N(0, sigma))``
with six sinusoids, each with a period drawn from U[8, 64] px at a random
orientation and an amplitude from U[10, 30].  The texture is analytic, so a
shifted copy is evaluated exactly rather than resampled.

Motion: a "current" frame is the reference texture displaced by either one
global integer vector or a piecewise-constant field of rectangles painted in
order, then noised again.  A block whose true displacement is (dx, dy) has
cur(x, y) = tex(x + dx, y + dy), so the vector the estimator should find is
(dx, dy) in full pels.

Everything comes from ``numpy.random.default_rng(seed)`` so CPU and GPU runs
see identical frames.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

# Standard luma sizes (multiples of 16 where the codec needs them).
SIZES = {
    "QCIF": (176, 144),
    "CIF": (352, 288),
    "720p": (1280, 720),
    "1080p": (1920, 1088),
    "4K": (3840, 2160),
}


@dataclass
class Texture:
    freqs: np.ndarray  # (6, 2) cycles per pixel along x, y
    phases: np.ndarray  # (6,)
    amps: np.ndarray  # (6,)

    @staticmethod
    def make(rng: np.random.Generator, n: int = 6) -> "Texture":
        period = rng.uniform(8.0, 64.0, n)
        theta = rng.uniform(0.0, np.pi, n)
        freqs = np.stack([np.cos(theta), np.sin(theta)], axis=1) / period[:, None]
        return Texture(freqs, rng.uniform(0.0, 2 * np.pi, n), rng.uniform(10.0, 30.0, n))

    def eval(self, xs: np.ndarray, ys: np.ndarray) -> np.ndarray:
        """Noise-free texture at integer coordinates (broadcast xs, ys)."""
        v = np.full(np.broadcast(xs, ys).shape, 128.0)
        for (fx, fy), ph, a in zip(self.freqs, self.phases, self.amps):
            v += a * np.sin(2 * np.pi * (fx * xs + fy * ys) + ph)
        return v


def _to_u8(v: np.ndarray) -> np.ndarray:
    return np.clip(np.rint(v), 0, 255).astype(np.uint8)


@dataclass
class FramePair:
    cur: np.ndarray  # (H, W) uint8
    ref: np.ndarray  # (H, W) uint8
    true_mv: np.ndarray  # (H/16, W/16, 2) int, full pel, MB-centre displacement


def global_shift_pair(w: int, h: int, seed: int, max_mv: int = 8, sigma_ref: float = 4.0,
                      sigma_cur: float = 2.0, mv: tuple[int, int] | None = None) -> FramePair:
    rng = np.random.default_rng(seed)
    tex = Texture.make(rng)
    if mv is None:
        mv = (int(rng.integers(-max_mv, max_mv + 1)), int(rng.integers(-max_mv, max_mv + 1)))
    ys, xs = np.mgrid[0:h, 0:w]
    # Shared texture noise so that the displaced copy matches the reference
    # sample for sample; evaluate it on a canvas covering both.
    m = max_mv + 1
    noise = rng.normal(0.0, sigma_ref, (h + 2 * m, w + 2 * m))
    ref = tex.eval(xs, ys) + noise[m:m + h, m:m + w]
    dx, dy = mv
    cur = tex.eval(xs + dx, ys + dy) + noise[m + dy:m + dy + h, m + dx:m + dx + w]
    cur = cur + rng.normal(0.0, sigma_cur, (h, w))
    tm = np.zeros((h // 16, w // 16, 2), dtype=np.int32)
    tm[..., 0], tm[..., 1] = dx, dy
    return FramePair(_to_u8(cur), _to_u8(ref), tm)


def rect_motion_pair(w: int, h: int, seed: int, n_rects: int = 24, max_mv: int = 32,
                     sigma_ref: float = 4.0, sigma_cur: float = 3.0) -> FramePair:
    """Piecewise-constant motion: rectangles painted in order over a static
    background, each displaced by its own integer vector in [-max_mv, max_mv]^2."""
    rng = np.random.default_rng(seed)
    tex = Texture.make(rng)
    m = max_mv + 1
    ys, xs = np.mgrid[0:h, 0:w]
    noise = rng.normal(0.0, sigma_ref, (h + 2 * m, w + 2 * m))
    ref = tex.eval(xs, ys) + noise[m:m + h, m:m + w]
    dxm = np.zeros((h, w), dtype=np.int64)
    dym = np.zeros((h, w), dtype=np.int64)
    for _ in range(n_rects):
        rw = int(rng.integers(w // 8, w // 2 + 1))
        rh = int(rng.integers(h // 8, h // 2 + 1))
        x0 = int(rng.integers(0, w - rw + 1))
        y0 = int(rng.integers(0, h - rh + 1))
        dx = int(rng.integers(-max_mv, max_mv + 1))
        dy = int(rng.integers(-max_mv, max_mv + 1))
        dxm[y0:y0 + rh, x0:x0 + rw] = dx
        dym[y0:y0 + rh, x0:x0 + rw] = dy
    sx = xs + dxm
    sy = ys + dym
    cur = tex.eval(sx, sy) + noise[np.clip(sy + m, 0, h + 2 * m - 1), np.clip(sx + m, 0, w + 2 * m - 1)]
    cur = cur + rng.normal(0.0, sigma_cur, (h, w))
    tm = np.stack([dxm[8::16, 8::16], dym[8::16, 8::16]], axis=-1).astype(np.int32)
    return FramePair(_to_u8(cur), _to_u8(ref), tm)


def shifted_copies(w: int, h: int, seed: int, n: int, max_mv: int = 16,
                   sigmas: tuple[float, float] = (2.0, 6.0)) -> tuple[np.ndarray, list[np.ndarray], list[tuple[int, int]]]:
    """One current frame and ``n`` references, each a differently shifted and
    noised copy of one seeded texture (multi-reference workloads)."""
    rng = np.random.default_rng(seed)
    tex = Texture.make(rng)
    ys, xs = np.mgrid[0:h, 0:w]
    cur = _to_u8(tex.eval(xs, ys) + rng.normal(0.0, 2.0, (h, w)))
    refs, mvs = [], []
    for _ in range(n):
        dx = int(rng.integers(-max_mv, max_mv + 1))
        dy = int(rng.integers(-max_mv, max_mv + 1))
        s = float(rng.uniform(*sigmas))
        refs.append(_to_u8(tex.eval(xs - dx, ys - dy) + rng.normal(0.0, s, (h, w))))
        mvs.append((dx, dy))
    return cur, refs, mvs


class StreamSource:
    """Frames of one video stream: a seeded texture under a drifting
    piecewise-constant motion field.  The background pans by ``pan`` pels per
    frame; each of ``n_rects`` rectangles moves by its own integer velocity in
    [-max_mv, max_mv]^2 per frame, carrying its content with it (so the true
    motion of a block inside rectangle k between frames t-1 and t is minus
    that velocity).  Frame t adds fresh sensor noise N(0, sigma) seeded by
    (seed, t), so ``frame(t)`` is the same on every call."""

    def __init__(self, w: int, h: int, seed: int, n_rects: int = 12, max_mv: int = 8,
                 pan: tuple[int, int] | None = None, sigma: float = 3.0):
        rng = np.random.default_rng(seed)
        self.w, self.h, self.seed, self.sigma = w, h, seed, sigma
        self.tex = Texture.make(rng)
        self.pan = pan if pan is not None else (int(rng.integers(-2, 3)), int(rng.integers(-2, 3)))
        self.rects = []
        for _ in range(n_rects):
            rw = int(rng.integers(w // 8, w // 3 + 1))
            rh = int(rng.integers(h // 8, h // 3 + 1))
            self.rects.append((rw, rh, int(rng.integers(0, w - rw + 1)), int(rng.integers(0, h - rh + 1)),
                               int(rng.integers(-max_mv, max_mv + 1)), int(rng.integers(-max_mv, max_mv + 1))))

    def frame(self, t: int) -> np.ndarray:
        w, h = self.w, self.h
        ys, xs = np.mgrid[0:h, 0:w]
        dx = np.full((h, w), t * self.pan[0], dtype=np.int64)
        dy = np.full((h, w), t * self.pan[1], dtype=np.int64)
        for rw, rh, x0, y0, vx, vy in self.rects:
            # the rectangle's position wraps inside the picture
            px = (x0 + t * vx) % (w - rw + 1)
            py = (y0 + t * vy) % (h - rh + 1)
            dx[py:py + rh, px:px + rw] = t * vx
            dy[py:py + rh, px:px + rw] = t * vy
        noise = np.random.default_rng([self.seed, t]).normal(0.0, self.sigma, (h, w))
        return _to_u8(self.tex.eval(xs - dx, ys - dy) + noise)


# The mixed batched workload: 24 QCIF + 20 CIF + 12 720p + 8 1080p frame
# pairs, each stream with its own texture and 8..24 moving rectangles, and a
# search range drawn from {8, 16, 32, 64} (motion bounded by min(R, 32)).
BATCH_PLAN = (("QCIF", 24), ("CIF", 20), ("720p", 12), ("1080p", 8))


def mixed_batch(seed: int = 2026) -> list[tuple[int, int, int, FramePair]]:
    """(width, height, R, frame pair) per stream, in the plan's order."""
    rng = np.random.default_rng(seed)
    out = []
    for name, count in BATCH_PLAN:
        w, h = SIZES[name]
        h -= h % 16
        for _ in range(count):
            R = int(rng.choice([8, 16, 32, 64]))
            sd = int(rng.integers(1 << 30))
            fp = rect_motion_pair(w, h, sd, n_rects=int(rng.integers(8, 25)), max_mv=min(R, 32))
            out.append((w, h, R, fp))
    return out

