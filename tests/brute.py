"""Brute-force restatement of the pipeline for tiny inputs (test infrastructure).

An independent numpy version of the per-pixel semantics: vertex transform,
28.4 snapping, det cull, the pixel-centre edge test with the top-left bias
(EDXRaster/Core/RasterTriangle.h:35-63,296-337), the sequential LEQUAL depth
test in submission order (Core/FrameBuffer.cpp:54-68) and the pixel shaders
(Core/Shader.h:142-170,209-282).  No tiles, quadrants, quads, fragment lists
or clipping: scenes passed here must lie inside the view frustum.
"""
import numpy as np

F = np.float32


def wrap32(a):
    """int64 -> int32 with two's-complement wrap (the reference's SSE integer ops)."""
    a = np.asarray(a, np.int64) & 0xFFFFFFFF
    return np.where(a >= 1 << 31, a - (1 << 32), a).astype(np.int64)


def xform(m, x, y, z, w):
    m = np.asarray(m, np.float32).reshape(4, 4)
    out = []
    for r in range(4):
        out.append(((m[r, 0] * x + m[r, 1] * y) + m[r, 2] * z) + m[r, 3] * w)
    return out


def setup(mesh, mvp, raster):
    """Per triangle: snapped x, y (T,3), det, invDet, z (T,3), inv_w (T,3), valid."""
    x, y, z = mesh.px.astype(F), mesh.py.astype(F), mesh.pz.astype(F)
    cx, cy, cz, cw = xform(mvp, x, y, z, F(1))
    idx = mesh.indices.reshape(-1, 3).astype(np.int64)
    hx, hy, hz = cx / cw, cy / cw, cz / cw
    rx, ry, rz, rw = xform(raster, hx, hy, hz, F(1))
    sx, sy = rx / rw, ry / rw
    X = np.trunc(sx * F(16)).astype(np.int64)[idx]
    Y = np.trunc(sy * F(16)).astype(np.int64)[idx]
    B = [wrap32(Y[:, k] - Y[:, (k + 1) % 3]) for k in range(3)]
    C = [wrap32(X[:, (k + 1) % 3] - X[:, k]) for k in range(3)]
    det = wrap32(C[2] * B[1] - C[1] * B[2])
    inv_w = (F(1) / cw)
    zdiv = (cz * inv_w)[idx]
    inv_det = np.where(det > 0, F(1) / np.maximum(det, 1).astype(F), F(0)).astype(F)
    return {"X": X, "Y": Y, "B": B, "C": C, "det": det, "inv_det": inv_det, "z": zdiv, "inv_w": inv_w[idx],
            "idx": idx, "valid": det > 0}


def rasterize(mesh, width, height, mvp, raster):
    """Returns (ids, depth, lambda0, lambda1) in raster (y-down) orientation."""
    s = setup(mesh, mvp, raster)
    ys, xs = np.mgrid[0:height, 0:width]
    px = xs.astype(np.int64) * 16 + 8
    py = ys.astype(np.int64) * 16 + 8
    D = np.full((height, width), F(1), np.float32)
    ids = np.full((height, width), 0xFFFFFFFF, np.uint32)
    L0 = np.zeros((height, width), np.float32)
    L1 = np.zeros((height, width), np.float32)
    X, Y, B, C = s["X"], s["Y"], s["B"], s["C"]
    for t in range(X.shape[0]):
        if not s["valid"][t]:
            continue
        cov = np.ones((height, width), bool)
        for k in range(3):
            a, b = k, (k + 1) % 3
            bias = -1 if (Y[t, b] > Y[t, a]) or (Y[t, a] == Y[t, b] and X[t, a] > X[t, b]) else 0
            e = wrap32(B[k][t] * (px - X[t, a]) + C[k][t] * (py - Y[t, a]) + bias)
            cov &= e >= 0
        if not cov.any():
            continue
        n0 = wrap32(B[1][t] * (px - X[t, 2]) + C[1][t] * (py - Y[t, 2]))
        n1 = wrap32(B[2][t] * (px - X[t, 2]) + C[2][t] * (py - Y[t, 2]))
        l0 = n0.astype(F) * s["inv_det"][t]
        l1 = n1.astype(F) * s["inv_det"][t]
        l2 = (F(1) - l0) - l1
        z = s["z"][t]
        d = (l0 * z[0] + l1 * z[1]) + l2 * z[2]
        win = cov & (d <= D)
        D = np.where(win, d, D)
        ids = np.where(win, np.uint32(t), ids)
        L0 = np.where(win, l0, L0)
        L1 = np.where(win, l1, L1)
    return ids, D, L0, L1, s


def normalize(v):
    v = [F(c) for c in v]
    s = F(1) / np.sqrt(F((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]))
    return [c * s for c in v]


def pow_decl(x, s):
    if 0 <= s <= 4096 and float(s).is_integer():
        n = int(s)
        result = np.ones_like(x)
        base = x.copy()
        while n:
            if n & 1:
                result = result * base
            n >>= 1
            if n:
                base = base * base
        return result
    return np.power(x, F(s)).astype(F)


def to_byte(c):
    c = np.minimum(np.maximum(np.nan_to_num(c, nan=0.0), F(0)), F(1)).astype(F)
    return np.trunc(c * F(255) + F(0.5)).astype(np.uint8)


def shade(mesh, ids, L0, L1, s, eye, light, shader, shininess, ambient=0.2, kd=3.0, ks=3.0, albedo=(0.9, 0.9, 0.9)):
    """float32 shading of the winners; returns RGBA8 in raster orientation."""
    H, W = ids.shape
    out = np.zeros((H, W, 4), np.uint8)
    m = ids != 0xFFFFFFFF
    if not m.any():
        return out
    t = ids[m].astype(np.int64)
    b0, b1 = L0[m], L1[m]
    b2 = (F(1) - b0) - b1
    iw = s["inv_w"][t]
    b0 = b0 * iw[:, 0]
    b1 = b1 * iw[:, 1]
    b2 = b2 * iw[:, 2]
    inv_b = F(1) / ((b0 + b1) + b2)
    b0 = b0 * inv_b
    b1 = b1 * inv_b
    b2 = (F(1) - b0) - b1
    vi = s["idx"][t]

    def interp(a):
        a = a.astype(F)
        return (b0 * a[vi[:, 0]] + b1 * a[vi[:, 1]]) + b2 * a[vi[:, 2]]

    P = [interp(mesh.px), interp(mesh.py), interp(mesh.pz)]
    N = [interp(mesh.nx), interp(mesh.ny), interp(mesh.nz)]
    w = F(1) / np.sqrt((N[0] * N[0] + N[1] * N[1]) + N[2] * N[2])
    n = [c * w for c in N]
    L = normalize(light)
    da = (L[0] * n[0] + L[1] * n[1]) + L[2] * n[2]
    da = np.where(da < F(0), F(0), da)
    diffuse = ((da + F(ambient)) * F(kd)) * F(1.0 / np.pi)
    if shader == 0:
        rgb = [diffuse * F(a) for a in albedo]
    else:
        e = [F(eye[k]) - P[k] for k in range(3)]
        w = F(1) / np.sqrt((e[0] * e[0] + e[1] * e[1]) + e[2] * e[2])
        e = [c * w for c in e]
        h = [L[k] + e[k] for k in range(3)]
        w = F(1) / np.sqrt((h[0] * h[0] + h[1] * h[1]) + h[2] * h[2])
        h = [c * w for c in h]
        sa = (n[0] * h[0] + n[1] * h[1]) + n[2] * h[2]
        spec = pow_decl(sa, shininess) * F(ks)
        if shader == 1:
            rgb = [diffuse + spec] * 3
        else:
            col = [interp(mesh.cr), interp(mesh.cg), interp(mesh.cb)]
            rgb = [diffuse * col[k] + spec for k in range(3)]
    for k in range(3):
        out[..., k][m] = to_byte(rgb[k])
    out[..., 3][m] = 255
    return out


def to_bottom_up(a):
    """Raster rows (y down) -> GetBackBuffer rows (row H-1-y)."""
    return a[::-1].copy()


# Sample offsets in 28.4 units per MSAA level (x, y pairs): the reference's
# FrameBuffer::MultiSampleOffsets table (Core/FrameBuffer.cpp:107-191).
SAMPLE_OFFSETS = [
    [(0, 0)],
    [(4, 4), (-4, -4)],
    [(-2, -6), (6, -2), (-6, 2), (2, 6)],
    [(1, -3), (-1, 3), (5, 1), (-3, -5), (-5, 5), (-7, -1), (3, 7), (7, -7)],
    [(1, 1), (-1, -3), (-3, 2), (4, -1), (-5, -2), (2, 5), (5, 3), (3, -5), (-2, 6), (0, -7), (-4, -6), (-6, 4),
     (-8, 0), (7, -4), (6, 7), (-7, -8)],
]
SAMPLE_OFFSETS.append(SAMPLE_OFFSETS[4] + [
    (1, 3), (-3, -3), (-3, 0), (6, -2), (-7, -1), (3, 4), (7, 3), (3, -6), (-2, 7), (0, -4), (-2, -5), (-7, 6),
    (-8, 3), (4, -1), (2, 7), (4, -8)])


def rasterize_ms(mesh, width, height, mvp, raster, level):
    """Multisampling (Core/Rasterizer.h:202-300,355-415): coverage, depth and the
    sequential LEQUAL test per sample at centre + offset; a winner is shaded with
    its barycentrics at the pixel centre.  Returns per-sample (ids, depth,
    centre lambda0, centre lambda1), shape (S, H, W), in raster orientation."""
    s = setup(mesh, mvp, raster)
    offs = SAMPLE_OFFSETS[level]
    S = len(offs)
    ys, xs = np.mgrid[0:height, 0:width]
    cx = xs.astype(np.int64) * 16 + 8
    cy = ys.astype(np.int64) * 16 + 8
    D = np.full((S, height, width), F(1), np.float32)
    ids = np.full((S, height, width), 0xFFFFFFFF, np.uint32)
    L0 = np.zeros((S, height, width), np.float32)
    L1 = np.zeros((S, height, width), np.float32)
    X, Y, B, C = s["X"], s["Y"], s["B"], s["C"]
    for t in range(X.shape[0]):
        if not s["valid"][t]:
            continue
        bias = [-1 if (Y[t, (k + 1) % 3] > Y[t, k]) or (Y[t, k] == Y[t, (k + 1) % 3] and X[t, k] > X[t, (k + 1) % 3])
                else 0 for k in range(3)]
        c0 = (wrap32(B[1][t] * (cx - X[t, 2]) + C[1][t] * (cy - Y[t, 2])).astype(F) * s["inv_det"][t])
        c1 = (wrap32(B[2][t] * (cx - X[t, 2]) + C[2][t] * (cy - Y[t, 2])).astype(F) * s["inv_det"][t])
        z = s["z"][t]
        for si, (ox, oy) in enumerate(offs):
            px, py = cx + ox, cy + oy
            cov = np.ones((height, width), bool)
            for k in range(3):
                a = k
                e = wrap32(B[k][t] * (px - X[t, a]) + C[k][t] * (py - Y[t, a]) + bias[k])
                cov &= e >= 0
            if not cov.any():
                continue
            l0 = wrap32(B[1][t] * (px - X[t, 2]) + C[1][t] * (py - Y[t, 2])).astype(F) * s["inv_det"][t]
            l1 = wrap32(B[2][t] * (px - X[t, 2]) + C[2][t] * (py - Y[t, 2])).astype(F) * s["inv_det"][t]
            l2 = (F(1) - l0) - l1
            d = (l0 * z[0] + l1 * z[1]) + l2 * z[2]
            win = cov & (d <= D[si])
            D[si] = np.where(win, d, D[si])
            ids[si] = np.where(win, np.uint32(t), ids[si])
            L0[si] = np.where(win, c0, L0[si])
            L1[si] = np.where(win, c1, L1[si])
    return ids, D, L0, L1, s


def resolve(samples_rgba):
    """FrameBuffer::Resolve with the declared conversions: the mean of
    b / 255.0f over the samples (summed in sample order), times 1 / S, then
    FromFloats."""
    S = samples_rgba.shape[0]
    acc = np.zeros(samples_rgba.shape[1:], np.float32)
    for si in range(S):
        acc = (acc + samples_rgba[si].astype(F) / F(255)).astype(F)
    return to_byte(acc * (F(1) / F(S)))
