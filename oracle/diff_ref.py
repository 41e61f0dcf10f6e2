"""FP64 CPU reference of the differentiable interpolation (test
infrastructure only: the checker of cr_interpolate / cr_interpolate_backward,
never the product).

Given clip positions pos [V, 4], attributes attr [V, C], triangles tris
[T, 3] and a fixed triangle-id image [H, W] (row 0 at the bottom, -1 where
empty), each covered pixel (x, y) has NDC centre q = ((2x+1)/W - 1,
(2y+1)/H - 1) and perspective-correct barycentrics b_k = c_k / sum(c),
c_k = (v_{k+1} x v_{k+2}) . (qx, qy, 1) with v_k = (x_k, y_k, w_k) -- the
point of the triangle's plane that projects to q.  out = sum_k b_k attr_k.
The gradients follow by the chain rule (visibility fixed).  The reference
has no counterpart (its shaders interpolate with fixed-point planes,
FineRaster.inl:21-119, and are not differentiated).
"""
from __future__ import annotations

import numpy as np


def _pixel_geometry(pos, tris, tri_image):
    H, W = tri_image.shape
    ys, xs = np.nonzero(tri_image >= 0)
    t = tri_image[ys, xs].astype(np.int64)
    idx = tris[t].astype(np.int64)                       # [P, 3]
    qx = (2.0 * xs + 1.0) / W - 1.0
    qy = (2.0 * ys + 1.0) / H - 1.0
    p = pos.astype(np.float64)[idx]                      # [P, 3, 4]
    xsh = p[..., 0] - qx[:, None] * p[..., 3]
    ysh = p[..., 1] - qy[:, None] * p[..., 3]
    c = np.stack([xsh[:, (k + 1) % 3] * ysh[:, (k + 2) % 3] - ysh[:, (k + 1) % 3] * xsh[:, (k + 2) % 3]
                  for k in range(3)], axis=1)
    s = c.sum(axis=1)
    ok = s != 0
    b = np.where(ok[:, None], c / np.where(ok, s, 1.0)[:, None], 0.0)
    return ys, xs, idx, qx, qy, xsh, ysh, s, b, ok


def interpolate(pos, attr, tris, tri_image):
    """FP64 forward: out [H, W, C]."""
    H, W = tri_image.shape
    a = attr.astype(np.float64)
    out = np.zeros((H, W, a.shape[1]), dtype=np.float64)
    ys, xs, idx, qx, qy, xsh, ysh, s, b, ok = _pixel_geometry(pos, tris, tri_image)
    out[ys, xs] = np.einsum("pk,pkc->pc", b, a[idx])
    return out


def backward(pos, attr, tris, tri_image, grad_out):
    """FP64 dL/dpos [V, 4] (z column 0) and dL/dattr [V, C] given dL/dout."""
    a = attr.astype(np.float64)
    g = grad_out.astype(np.float64)
    V = pos.shape[0]
    gp = np.zeros((V, 4), dtype=np.float64)
    ga = np.zeros((V, a.shape[1]), dtype=np.float64)
    ys, xs, idx, qx, qy, xsh, ysh, s, b, ok = _pixel_geometry(pos, tris, tri_image)
    ys, xs, idx, qx, qy, xsh, ysh, s, b = (v[ok] for v in (ys, xs, idx, qx, qy, xsh, ysh, s, b))
    gpix = g[ys, xs]                                       # [P, C]
    for k in range(3):
        np.add.at(ga, idx[:, k], b[:, k:k + 1] * gpix)
    dB = np.einsum("pc,pkc->pk", gpix, a[idx])              # dL/db_k
    mean = (b * dB).sum(axis=1, keepdims=True)
    dc = (dB - mean) / s[:, None]                           # dL/dc_k
    dxs = np.zeros_like(xsh)
    dys = np.zeros_like(ysh)
    for k in range(3):
        k1, k2 = (k + 1) % 3, (k + 2) % 3
        dxs[:, k1] += dc[:, k] * ysh[:, k2]
        dys[:, k1] -= dc[:, k] * xsh[:, k2]
        dxs[:, k2] -= dc[:, k] * ysh[:, k1]
        dys[:, k2] += dc[:, k] * xsh[:, k1]
    for k in range(3):
        np.add.at(gp[:, 0], idx[:, k], dxs[:, k])
        np.add.at(gp[:, 1], idx[:, k], dys[:, k])
        np.add.at(gp[:, 3], idx[:, k], -qx * dxs[:, k] - qy * dys[:, k])
    return gp, ga


def l2_loss(out, target):
    d = out - target
    return float((d * d).sum())


def finite_difference(pos, attr, tris, tri_image, target, params, h=1e-4):
    """FP64 central differences of L = sum((out - target)^2) for the listed
    parameters, visibility fixed.  params: list of ("pos", v, comp) or
    ("attr", v, c).  Only the pixels of triangles that use vertex v change,
    so each difference is evaluated on those pixels alone."""
    pos64 = pos.astype(np.float64)
    attr64 = attr.astype(np.float64)
    H, W = tri_image.shape
    ys_all, xs_all = np.nonzero(tri_image >= 0)
    t_all = tri_image[ys_all, xs_all]
    uses = tris.astype(np.int64)[t_all]                     # [P, 3]
    res = []
    for kind, v, comp in params:
        sel = np.nonzero((uses == v).any(axis=1))[0]
        img = np.full((H, W), -1, dtype=np.int64)
        img[ys_all[sel], xs_all[sel]] = t_all[sel]
        vals = []
        for sgn in (1.0, -1.0):
            P = pos64.copy()
            A = attr64.copy()
            if kind == "pos":
                P[v, comp] += sgn * h
            else:
                A[v, comp] += sgn * h
            out = interpolate(P, A, tris, img)
            d = out[ys_all[sel], xs_all[sel]] - target[ys_all[sel], xs_all[sel]]
            vals.append((d * d).sum())
        res.append((vals[0] - vals[1]) / (2.0 * h))
    return np.array(res)
