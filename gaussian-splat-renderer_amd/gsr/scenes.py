"""Deterministic synthetic scenes of the benchmark workloads.

Each attribute comes from its own SplitMix64 stream (counter based: value i of stream s is
mix64(seed' + (i + 1) * 0x9E3779B97F4A7C15)) and Box-Muller in double precision, then one
cast to float32, so the same seed gives the same scene on any host. Layout and activations
are those of the reference's GaussianScene (scene/gaussian_scene.h:1-36) as produced by
its PLY loader (scene/ply_loader.cpp:132-150):

    pos_alpha[i] = (mean, sigmoid(opacity logit))      float32 (N, 4)
    scales[i]    = (exp(log sigma), 0)                  float32 (N, 4)
    rotations[i] = normalize(q), q ~ N(0, I4), w first  float32 (N, 4)  (uniform on S^3)
    sh[i]        = K RGB triplets, k-major              float32 (N, K*3)
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix64(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def _stream_seed(seed: int, tag: int) -> np.uint64:
    with np.errstate(over="ignore"):
        return _mix64(np.array([(seed * 0x100000001B3 + tag * 0x9E3779B1) & 0xFFFFFFFFFFFFFFFF], np.uint64))[0]


def uniform01(seed: int, tag: int, n: int) -> np.ndarray:
    """n doubles in (0, 1]: ((u64 >> 11) + 1) * 2^-53."""
    s = _stream_seed(seed, tag)
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = _mix64(s + idx * _GOLDEN)
    return ((z >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)


def normal(seed: int, tag: int, n: int) -> np.ndarray:
    """n standard normals (double), Box-Muller on pairs of uniforms of one stream."""
    m = (n + 1) // 2
    u = uniform01(seed, tag, 2 * m)
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    a = 2.0 * np.pi * u2
    out = np.empty(2 * m)
    out[0::2] = r * np.cos(a)
    out[1::2] = r * np.sin(a)
    return out[:n]


@dataclass
class SyntheticScene:
    pos_alpha: np.ndarray
    scales: np.ndarray
    rotations: np.ndarray
    sh: np.ndarray
    sh_degree: int
    seed: int

    @property
    def count(self) -> int:
        return int(self.pos_alpha.shape[0])

    @property
    def K(self) -> int:
        return (self.sh_degree + 1) ** 2


def make_scene(count: int, seed: int = 0, mean_sigma: float = 2.0, sh_degree: int = 3, log_scale_mean: float = -4.5,
               log_scale_sigma: float = 0.8, logit_sigma: float = 1.5, dc_sigma: float = 0.5,
               rest_sigma: float = 0.1) -> SyntheticScene:
    """The benchmark distributions: means ~ N(0, mean_sigma^2 I), log-scales ~ N(-4.5, 0.8^2),
    uniform unit quaternions, opacity logits ~ N(0, 1.5^2), SH DC ~ N(0, 0.5^2) and the
    higher bands ~ N(0, 0.1^2)."""
    n = int(count)
    K = (sh_degree + 1) ** 2
    means = (normal(seed, 1, 3 * n) * mean_sigma).astype(np.float32).reshape(n, 3)
    log_s = (log_scale_mean + normal(seed, 2, 3 * n) * log_scale_sigma).astype(np.float32).reshape(n, 3)
    q = (4.0 * normal(seed, 3, 4 * n)).astype(np.float32).reshape(n, 4)
    logits = (normal(seed, 4, n) * logit_sigma).astype(np.float32)
    sh = np.empty((n, K, 3), np.float32)
    sh[:, 0, :] = (normal(seed, 5, 3 * n) * dc_sigma).astype(np.float32).reshape(n, 3)
    if K > 1:
        sh[:, 1:, :] = (normal(seed, 6, (K - 1) * 3 * n) * rest_sigma).astype(np.float32).reshape(n, K - 1, 3)
    # Activations as the PLY loader: sigmoid, exp, normalize (float32).
    alpha = (np.float32(1.0) / (np.float32(1.0) + np.exp(-logits))).astype(np.float32)
    sigma = np.exp(log_s).astype(np.float32)
    norm = np.sqrt((q.astype(np.float32) ** 2).sum(axis=1, dtype=np.float32)).astype(np.float32)
    norm[norm == 0] = 1.0
    qn = (q / norm[:, None]).astype(np.float32)
    pos_alpha = np.concatenate([means, alpha[:, None]], axis=1).astype(np.float32)
    scales = np.concatenate([sigma, np.zeros((n, 1), np.float32)], axis=1)
    return SyntheticScene(np.ascontiguousarray(pos_alpha), np.ascontiguousarray(scales), np.ascontiguousarray(qn),
                          np.ascontiguousarray(sh.reshape(n, K * 3)), sh_degree, seed)


def make_lowrank_scene(count: int, seed: int = 0, rank: int = 2, mean_sigma: float = 3.0, sh_degree: int = 3,
                       rest_sigma: float = 0.1) -> SyntheticScene:
    """make_scene's distributions except the view-dependent SH bands, which have rank `rank`:
    sh[i, k, c] = sum_r a[i, r] P[r, k, c] for k >= 1, with `rank` patterns P[r] ((K - 1) x 3,
    entries ~ N(0, rest_sigma^2)) shared by every gaussian and amplitudes a[i, r] ~ N(0, 1). Each
    gaussian's colour is then its own base colour plus a mix of `rank` view-dependent functions
    common to the scene (a material shared across the scene), the case a small per-gaussian latent
    and one shared network can represent; the DC terms stay independent per gaussian."""
    sc = make_scene(count, seed, mean_sigma, sh_degree)
    n, K = sc.count, sc.K
    sh = sc.sh.reshape(n, K, 3).copy()
    if K > 1:
        P = normal(seed, 11, rank * (K - 1) * 3).reshape(rank, K - 1, 3) * rest_sigma
        a = normal(seed, 12, n * rank).reshape(n, rank)
        sh[:, 1:, :] = np.einsum("nr,rkc->nkc", a, P).astype(np.float32)
    return SyntheticScene(sc.pos_alpha, sc.scales, sc.rotations, np.ascontiguousarray(sh.reshape(n, K * 3)),
                          sh_degree, seed)


def make_test_scene(count: int, seed: int = 99, sh_degree: int = 1) -> SyntheticScene:
    """The renderer test scene of test_tile_raster.cpp:246-275, restated with this module's
    generator: a 16^3 box around a camera at z = +6, some gaussians behind it, inside the
    near plane, off screen, large or almost transparent, every 7th alpha scaled by 0.01,
    every 3rd nearly opaque (early termination), and exact depth ties (every 4099th at
    z = -1)."""
    n = int(count)
    K = (sh_degree + 1) ** 2
    u = uniform01(seed, 7, n * 12).reshape(n, 12).astype(np.float32)
    p = u[:, 0:3] * np.float32(16.0) - np.float32(8.0)
    alpha = u[:, 3].copy()
    i = np.arange(n)
    alpha[i % 7 == 0] *= np.float32(0.01)
    opaque = i % 3 == 0
    alpha[opaque] = np.float32(0.95) + np.float32(0.04) * u[opaque, 4]
    p[i % 4099 == 0, 2] = np.float32(-1.0)
    s = np.exp(u[:, 5] * np.float32(4.0) - np.float32(5.5)).astype(np.float32)
    scales = np.stack([s, s * (np.float32(0.2) + u[:, 6]), s * (np.float32(0.2) + u[:, 7]), np.zeros(n, np.float32)],
                      axis=1).astype(np.float32)
    q = (uniform01(seed, 8, n * 4).reshape(n, 4) - 0.5).astype(np.float32)
    q /= np.sqrt((q * q).sum(axis=1, dtype=np.float32))[:, None]
    shv = ((uniform01(seed, 9, n * K * 3).reshape(n, K * 3) - 0.5)).astype(np.float32)
    shv[:, :3] *= np.float32(2.0)
    shv[:, 3:] *= np.float32(0.5)
    pos_alpha = np.concatenate([p, alpha[:, None]], axis=1).astype(np.float32)
    return SyntheticScene(np.ascontiguousarray(pos_alpha), np.ascontiguousarray(scales),
                          np.ascontiguousarray(q.astype(np.float32)), np.ascontiguousarray(shv), sh_degree, seed)


def make_needle_scene(count: int, seed: int = 77, sh_degree: int = 1) -> SyntheticScene:
    """The tight-bounds test scene of test_tight_bounds.cpp:181-211 (150,000 gaussians there),
    restated with this module's generator: the renderer test scene's 16^3 box, alpha uniform with
    every 7th scaled by 0.01 (many below 1/255) and every 3rd in [0.95, 1), exact depth ties (every
    4099th at z = -1), sigma = exp(U * 4 - 5.5), every 5th gaussian a needle with axes
    (3 s, 0.01 s, 0.02 s) (the others (s, s (0.2 + U), s (0.2 + U))), arbitrary rotations
    normalize(U4 - 0.5), SH DC in [-1, 1) and the other coefficients in [-0.25, 0.25)."""
    n = int(count)
    K = (sh_degree + 1) ** 2
    u = uniform01(seed, 17, n * 8).reshape(n, 8).astype(np.float32)
    p = u[:, 0:3] * np.float32(16.0) - np.float32(8.0)
    i = np.arange(n)
    alpha = u[:, 3].copy()
    alpha[i % 7 == 0] *= np.float32(0.01)
    opaque = i % 3 == 0
    alpha[opaque] = np.float32(0.95) + np.float32(0.05) * u[opaque, 4]
    p[i % 4099 == 0, 2] = np.float32(-1.0)
    s = np.exp(u[:, 5] * np.float32(4.0) - np.float32(5.5)).astype(np.float32)
    needle = i % 5 == 0
    sx = s.copy()
    sy = s * (np.float32(0.2) + u[:, 6])
    sz = s * (np.float32(0.2) + u[:, 7])
    sx[needle] = s[needle] * np.float32(3.0)
    sy[needle] = s[needle] * np.float32(0.01)
    sz[needle] = s[needle] * np.float32(0.02)
    scales = np.stack([sx, sy, sz, np.zeros(n, np.float32)], axis=1).astype(np.float32)
    q = (uniform01(seed, 18, n * 4).reshape(n, 4) - 0.5).astype(np.float32)
    q /= np.sqrt((q * q).sum(axis=1, dtype=np.float32))[:, None]
    shv = (uniform01(seed, 19, n * K * 3).reshape(n, K * 3) - 0.5).astype(np.float32)
    shv[:, :3] *= np.float32(2.0)
    shv[:, 3:] *= np.float32(0.5)
    pos_alpha = np.concatenate([p, alpha[:, None]], axis=1).astype(np.float32)
    return SyntheticScene(np.ascontiguousarray(pos_alpha), np.ascontiguousarray(scales),
                          np.ascontiguousarray(q.astype(np.float32)), np.ascontiguousarray(shv), sh_degree, seed)


# The benchmark workloads (names carry the sizes).
WORKLOADS = {
    "100k-720p": dict(count=100_000, mean_sigma=2.0, width=1280, height=720),
    "1.2M-1080p-orbit": dict(count=1_200_000, mean_sigma=3.0, width=1920, height=1080),
}
