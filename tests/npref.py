"""Independent numpy (float64) references used to pin the oracle on the CPU.

Nothing here shares code with the oracle or the HIP library.  The renderer
below solves the volume-rendering integral along each pixel ray by
Gauss-Legendre quadrature of

    C = sum_i T_i * integral_0^L_i sigma_i exp(-sigma_i s) c_i(s) ds,
    T_i = exp(-sum_{j<i} sigma_j L_j),

with c_i linear along the ray (base colour + g . (p - v0)), tets taken in the
reference's depth order (ascending float32 bit pattern of |centre - cam|^2 - r^2,
crates/rmesh-tile/src/wgsl/tile_gen_compute.wgsl), and the ray-tet chord from
the half-spaces of the four faces chosen by combination (not the reference's
face table).  It applies none of the reference's tile binning, scanline
spans, closed-form integral weights or early-termination thresholds, so an
error in any of those shows up as a difference.
"""

from __future__ import annotations

import itertools

import numpy as np


def pixel_rays(cam, w, h):
    """(origin[3], unit directions[h*w, 3]) of pixel centres, float64."""
    cp, _, ci = cam
    ci = np.asarray(ci, np.float64)
    c2w = np.stack([ci[0:3], ci[4:7], ci[8:11]], axis=1)  # columns
    fx, fy, cx, cy = ci[12:16]
    px, py = np.meshgrid(np.arange(w, dtype=np.float64), np.arange(h, dtype=np.float64))
    dc = np.stack([(px + 0.5 - cx) / fx, (py + 0.5 - cy) / fy, np.ones_like(px)], -1).reshape(-1, 3)
    d = dc @ c2w.T
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.asarray(cp, np.float64), d


def tet_chords(tv, o, d):
    """Entry/exit ray parameters of rays o + t d through tets tv[M, 4, 3].

    Returns (t0[P, M], t1[P, M], hit[P, M]).  Each face plane is oriented so
    that the opposite vertex is inside; the chord is the intersection of the
    four half-line intervals.
    """
    P, M = len(d), len(tv)
    t0 = np.full((P, M), -np.inf)
    t1 = np.full((P, M), np.inf)
    ok = np.ones((P, M), bool)
    for tri in itertools.combinations(range(4), 3):
        opp = ({0, 1, 2, 3} - set(tri)).pop()
        a, b, c = tv[:, tri[0]], tv[:, tri[1]], tv[:, tri[2]]
        n = np.cross(b - a, c - a)
        s = np.sign(np.einsum("mk,mk->m", n, tv[:, opp] - a))
        n = n * s[:, None]  # inward: n . (x - a) >= 0 inside
        num = np.einsum("mk,mk->m", n, a - o)[None, :]  # inside iff n.(o-a) + t n.d >= 0  <=> t n.d >= num
        den = d @ n.T
        with np.errstate(divide="ignore", invalid="ignore"):
            t = num / den
        enter = den > 0
        leave = den < 0
        t0 = np.where(enter, np.maximum(t0, t), t0)
        t1 = np.where(leave, np.minimum(t1, t), t1)
        ok &= ~((den == 0) & (num > 0))
    hit = ok & (t0 < t1)
    return t0, t1, hit


def depth_order(scene, cam_pos):
    """Reference compositing order: ascending uint32 bits of the float32
    |centre - cam|^2 - r^2 (keys are ~bits, sorted ascending, read back to front)."""
    cd = np.asarray(scene["circumdata"], np.float32).reshape(-1, 4)
    cam = np.asarray(cam_pos, np.float32)
    dx, dy, dz = cd[:, 0] - cam[0], cd[:, 1] - cam[1], cd[:, 2] - cam[2]
    dist = ((dx * dx + dy * dy) + dz * dz) - cd[:, 3]
    return np.argsort(dist.astype(np.float32).view(np.uint32), kind="stable")


_GL_X, _GL_W = np.polynomial.legendre.leggauss(12)


def render(scene, cam, w, h, colors=None, min_t=0.0):
    """float64 RGBA image [h, w, 4] by quadrature; `colors` defaults to base_colors."""
    o, d = pixel_rays(cam, w, h)
    V = np.asarray(scene["vertices"], np.float64).reshape(-1, 3)
    idx = np.asarray(scene["indices"], np.int64).reshape(-1, 4)
    tv = V[idx]
    sig = np.asarray(scene["densities"], np.float64)
    g = np.asarray(scene["color_grads"], np.float64).reshape(-1, 3)
    col = np.asarray(scene["base_colors"] if colors is None else colors, np.float64).reshape(-1, 3)
    P = len(d)
    origins = o[None, :] + min_t * d  # [P, 3]; the reference offsets the ray origin by min_t
    rgb = np.zeros((P, 3))
    logT = np.zeros(P)
    order = depth_order(scene, cam[0])
    for m in order:
        t0, t1, hit = _chord_per_origin(tv[m], origins, d)
        L = np.where(hit, t1 - t0, 0.0)
        if not hit.any():
            continue
        s_len = sig[m] * L
        K = int(np.ceil(max(s_len.max(), 1e-12))) + 1  # sub-intervals with sigma*ds <= 1
        seg = np.zeros((P, 3))
        for k in range(K):
            a = L * k / K
            b = L * (k + 1) / K
            for x, wq in zip(_GL_X, _GL_W):
                s = 0.5 * (b - a) * x + 0.5 * (a + b)
                p = origins + d * (t0 + s)[:, None]
                c = col[m][None, :] + (p - tv[m, 0][None, :]) @ g[m][:, None]
                seg += (0.5 * (b - a) * wq * sig[m] * np.exp(-sig[m] * s))[:, None] * c
        T = np.exp(logT)
        rgb += np.where(hit[:, None], T[:, None] * seg, 0.0)
        logT -= np.where(hit, s_len, 0.0)
    out = np.concatenate([rgb, (1.0 - np.exp(logT))[:, None]], axis=1)
    return out.reshape(h, w, 4)


def _chord_per_origin(tv, origins, d):
    """tet_chords for one tet with a different origin per ray."""
    t0 = np.full(len(d), -np.inf)
    t1 = np.full(len(d), np.inf)
    ok = np.ones(len(d), bool)
    for tri in itertools.combinations(range(4), 3):
        opp = ({0, 1, 2, 3} - set(tri)).pop()
        a, b, c = tv[tri[0]], tv[tri[1]], tv[tri[2]]
        n = np.cross(b - a, c - a)
        if np.dot(n, tv[opp] - a) < 0:
            n = -n
        num = (a - origins) @ n
        den = d @ n
        with np.errstate(divide="ignore", invalid="ignore"):
            t = num / den
        t0 = np.where(den > 0, np.maximum(t0, t), t0)
        t1 = np.where(den < 0, np.minimum(t1, t), t1)
        ok &= ~((den == 0) & (num > 0))
    return t0, t1, ok & (t0 < t1)


def endpoint_colours_positive(scene, cam, w, h, colors=None):
    """True when every chord's colour stays >= 0 (the reference clamps endpoint
    colours at 0, which the quadrature does not model)."""
    o, d = pixel_rays(cam, w, h)
    V = np.asarray(scene["vertices"], np.float64).reshape(-1, 3)
    idx = np.asarray(scene["indices"], np.int64).reshape(-1, 4)
    tv = V[idx]
    g = np.asarray(scene["color_grads"], np.float64).reshape(-1, 3)
    col = np.asarray(scene["base_colors"] if colors is None else colors, np.float64).reshape(-1, 3)
    for m in range(len(tv)):
        t0, t1, hit = _chord_per_origin(tv[m], o[None, :], d)
        for t in (t0, t1):
            p = o[None, :] + d * np.where(hit, t, 0.0)[:, None]
            c = col[m][None, :] + (p - tv[m, 0][None, :]) @ g[m][:, None]
            if np.any(hit[:, None] & (c < 0)):
                return False
    return True
