"""FP64 closed forms of the pipeline's per-pixel values, clipping included
(test infrastructure).

An independent float64 restatement used to pin the oracle:

- ``clip_triangle``: homogeneous Sutherland-Hodgman clipping in plane order
  L R B T F N with the clipped coordinate placed on the plane, clip weights
  carried along and polygons with any w <= 0 dropped, then a fan (0, k-1, k)
  (EDXRaster/Core/Clipper.h:71-189,192-288).  Written from the textbook
  form (t = d0 / (d0 - d1) for the signed plane distance d), not from the
  reference's per-plane expressions.
- ``frame_reference``: for every covered pixel of an oracle frame, the FP64
  depth (screen-linear z/w, Core/RasterTriangle.h:324-337) and the
  perspective-correct attributes (Core/Shader.h:142-170) of the winning
  triangle.  Clip-created vertices take their attributes from the FP64 clip
  weights (Core/Clipper.h:136-151), so a wrong weight in the oracle's vertex
  combination shows up as a colour error.
- ``colour``: the pixel shaders of Core/Shader.h:209-282 in FP64 with
  FromFloats' clamp-and-round.

Visibility (which triangle won a pixel) is taken from the oracle; the 28.4
snapped vertex positions of clipped pieces are taken from the oracle's setup
records after ``check_records`` has pinned them to the FP64 clip within one
sub-pixel step.
"""
import numpy as np

from brute import wrap32

LEFT, RIGHT, BOTTOM, TOP, NEAR, FAR = 1, 2, 4, 8, 16, 32
# ClipPolygon's plane order (Core/Clipper.h:238-278): L R B T F N.
PLANE_ORDER = (LEFT, RIGHT, BOTTOM, TOP, FAR, NEAR)


def outcode(v):
    x, y, z, w = v
    c = 0
    c |= LEFT if x < -w else 0
    c |= RIGHT if x > w else 0
    c |= BOTTOM if y < -w else 0
    c |= TOP if y > w else 0
    c |= FAR if z > w else 0
    c |= NEAR if z < 0 else 0
    return c


def _distance(v, plane):
    """Signed distance to a plane, >= 0 on the kept side."""
    x, y, z, w = v
    return {LEFT: x + w, RIGHT: w - x, BOTTOM: y + w, TOP: w - y, FAR: w - z, NEAR: z}[plane]


def _onto_plane(v, plane):
    v = v.copy()
    if plane == LEFT:
        v[0] = -v[3]
    elif plane == RIGHT:
        v[0] = v[3]
    elif plane == BOTTOM:
        v[1] = -v[3]
    elif plane == TOP:
        v[1] = v[3]
    elif plane == FAR:
        v[2] = v[3]
    else:
        v[2] = 0.0
    return v


def clip_triangle(c):
    """c: (3, 4) clip-space vertices (float64).  Returns the clipped polygon
    as a list of (position (4,), weights (3,)) or [] when it is dropped."""
    codes = [outcode(v) for v in c]
    planes = (codes[0] ^ codes[1]) | (codes[1] ^ codes[2]) | (codes[2] ^ codes[0])
    poly = [(np.asarray(c[k], np.float64), np.eye(3)[k]) for k in range(3)]
    for plane in PLANE_ORDER:
        if not planes & plane:
            continue
        out = []
        n = len(poly)
        for i in range(n):
            (a, wa), (b, wb) = poly[i], poly[(i + 1) % n]
            da, db = _distance(a, plane), _distance(b, plane)
            if (da >= 0) != (db >= 0):
                t = da / (da - db)
                out.append((_onto_plane(a + t * (b - a), plane), wa + t * (wb - wa)))
            if db >= 0:
                out.append((b, wb))
        poly = out
    if any(v[3] <= 0 for v, _ in poly):
        return []
    return poly


def clip_space(mesh, mvp):
    P = np.stack([mesh.px, mesh.py, mesh.pz, np.ones_like(mesh.px)], 1).astype(np.float64)
    return P @ np.asarray(mvp, np.float64).T


def snapped(v, raster):
    """28.4 position of a clip-space vertex: the homogeneous projection, the
    raster matrix and truncation of x * 16 (Core/RasterTriangle.h:35-40)."""
    ndc = np.array([v[0] / v[3], v[1] / v[3], v[2] / v[3], 1.0])
    s = np.asarray(raster, np.float64) @ ndc
    return int(np.trunc(s[0] / s[3] * 16.0)), int(np.trunc(s[1] / s[3] * 16.0))


def clipped_pieces(mesh, mvp, raster):
    """{triangle: [piece, ...]} for every input triangle that goes through
    clipping (an outcode set, not all three sharing a bit).  A piece is a dict
    of its fan index, FP64 clip positions (3, 4), weights (3, 3) and the FP64
    snapped 28.4 coordinates (3, 2)."""
    cs = clip_space(mesh, mvp)
    idx = mesh.indices.reshape(-1, 3).astype(np.int64)
    out = {}
    for t in range(idx.shape[0]):
        c = cs[idx[t]]
        codes = [outcode(v) for v in c]
        if not (codes[0] | codes[1] | codes[2]) or (codes[0] & codes[1] & codes[2]):
            continue
        poly = clip_triangle(c)
        pieces = []
        for k in range(2, len(poly)):
            vs = [poly[0], poly[k - 1], poly[k]]
            pieces.append({"sub": k - 2, "pos": np.stack([v for v, _ in vs]),
                           "weights": np.stack([w for _, w in vs]),
                           "xy": np.array([snapped(v, raster) for v, _ in vs], np.int64)})
        out[t] = pieces
    return out


def _det(xy):
    B1, C1 = xy[1, 1] - xy[2, 1], xy[2, 0] - xy[1, 0]
    B2, C2 = xy[2, 1] - xy[0, 1], xy[0, 0] - xy[2, 0]
    return C2 * B1 - C1 * B2


def check_records(records, pieces, max_step=1, z_tol=2e-6):
    """Pins the oracle's setup records of clipped triangles to the FP64 clip:
    the same surviving fan pieces (det > 0 after snapping), vertices within
    max_step sub-pixel steps, z/w within z_tol and 1/w within 1e-4 relative
    (fp32 clipping of clip coordinates near 40 loses about 3e-5 of z/w and
    1.5e-5 of w, so scenes of that size need z_tol = 1e-5).
    Returns the number of pieces compared."""
    got = {}
    for r in records:
        if r.tri_id in pieces:
            got[(r.tri_id, r.sub)] = r
    want = {}
    for t, ps in pieces.items():
        for p in ps:
            if _det(p["xy"]) > 0:
                want[(t, p["sub"])] = p
    assert set(got) == set(want), "clipped pieces differ: oracle-only %s, FP64-only %s" % (
        sorted(set(got) - set(want))[:8], sorted(set(want) - set(got))[:8])
    for key, p in want.items():
        r = got[key]
        xy = np.array([[r.x[k], r.y[k]] for k in range(3)], np.int64)
        assert np.abs(xy - p["xy"]).max() <= max_step, "piece %s: 28.4 vertices %s vs FP64 %s" % (
            key, xy.tolist(), p["xy"].tolist())
        z64 = p["pos"][:, 2] / p["pos"][:, 3]
        np.testing.assert_allclose(np.array(r.z[:], np.float64), z64, rtol=0, atol=z_tol)
        np.testing.assert_allclose(np.array(r.inv_w[:], np.float64), 1.0 / p["pos"][:, 3], rtol=1e-4)
    return len(want)


def _covering_piece(recs, px, py):
    """The record whose pixel-centre edge test (with the top-left bias,
    Core/RasterTriangle.h:296-312) covers (px, py) in 28.4 units."""
    for r in recs:
        X, Y = list(r.x), list(r.y)
        ok = True
        for k in range(3):
            a, b = k, (k + 1) % 3
            B, C = Y[a] - Y[b], X[b] - X[a]
            bias = -1 if (Y[b] > Y[a]) or (Y[a] == Y[b] and X[a] > X[b]) else 0
            if wrap32(B * (px - X[a]) + C * (py - Y[a]) + bias) < 0:
                ok = False
                break
        if ok:
            return r
    return None


def frame_reference(ids_raster, mesh, mvp, raster, records=None, pieces=None):
    """FP64 depth and perspective-correct attributes per covered pixel.

    ids_raster: (H, W) winning input triangles in raster orientation.
    Unclipped winners use their input vertices; clipped winners use the fan
    piece that covers the pixel (snapped positions from the oracle records,
    everything else from the FP64 clip).  Returns a dict of pixel coordinates,
    depth, and interpolated position / normal / colour lists."""
    H, W = ids_raster.shape
    ys, xs = np.nonzero(ids_raster != 0xFFFFFFFF)
    tri = ids_raster[ys, xs].astype(np.int64)
    cs = clip_space(mesh, mvp)
    idx = mesh.indices.reshape(-1, 3).astype(np.int64)
    attrs = np.stack([mesh.px, mesh.py, mesh.pz, mesh.nx, mesh.ny, mesh.nz, mesh.cr, mesh.cg, mesh.cb],
                     1).astype(np.float64)
    n = len(ys)
    depth = np.zeros(n)
    val = np.zeros((n, 9))
    by_tri = {}
    if records is not None:
        for r in records:
            by_tri.setdefault(r.tri_id, []).append(r)
    pieces = pieces or {}
    for i in range(n):
        t = int(tri[i])
        px, py = int(xs[i]) * 16 + 8, int(ys[i]) * 16 + 8
        if t in pieces:
            r = _covering_piece(by_tri.get(t, []), px, py)
            assert r is not None, "pixel (%d, %d) of clipped triangle %d lies in no piece" % (xs[i], ys[i], t)
            p = [q for q in pieces[t] if q["sub"] == r.sub][0]
            X, Y = np.array(r.x[:], np.float64), np.array(r.y[:], np.float64)
            pos, wts = p["pos"], p["weights"]
            A = wts @ attrs[idx[t]]          # attributes of the piece's vertices
        else:
            pos = cs[idx[t]]
            ndc = pos[:, :3] / pos[:, 3:4]
            sc = np.concatenate([ndc, np.ones((3, 1))], 1) @ np.asarray(raster, np.float64).T
            X = np.trunc(sc[:, 0].astype(np.float32) * 16).astype(np.float64)
            Y = np.trunc(sc[:, 1].astype(np.float32) * 16).astype(np.float64)
            A = attrs[idx[t]]
        B1, C1 = Y[1] - Y[2], X[2] - X[1]
        B2, C2 = Y[2] - Y[0], X[0] - X[2]
        det = C2 * B1 - C1 * B2
        l0 = (B1 * (px - X[2]) + C1 * (py - Y[2])) / det
        l1 = (B2 * (px - X[2]) + C2 * (py - Y[2])) / det
        lam = np.array([l0, l1, 1.0 - l0 - l1])
        depth[i] = lam @ (pos[:, 2] / pos[:, 3])
        b = lam / pos[:, 3]
        b /= b.sum()
        val[i] = b @ A
    return {"ys": ys, "xs": xs, "tri": tri, "depth": depth, "P": val[:, 0:3], "N": val[:, 3:6], "C": val[:, 6:9]}


def eye_position(mv):
    """TransformPoint(0, MV^-1) in FP64."""
    e = np.linalg.inv(np.asarray(mv, np.float64)) @ np.array([0.0, 0.0, 0.0, 1.0])
    return e[:3] / e[3]


def colour(ref, eye, light, shader, shininess, ambient=0.2, kd=3.0, ks=3.0, albedo=(0.9, 0.9, 0.9)):
    """RGB bytes of the pixel shaders (Core/Shader.h:209-282) in FP64:
    0 Lambertian with a constant albedo, 1 Blinn-Phong (albedo 1), 2 Blinn-Phong
    with the interpolated vertex colour as albedo; then clamp to [0, 1],
    * 255 + 0.5, truncated (Color4b::FromFloats as declared)."""
    N = ref["N"] / np.linalg.norm(ref["N"], axis=1, keepdims=True)
    L = np.asarray(light, np.float64) / np.linalg.norm(light)
    diffuse = (np.maximum(N @ L, 0.0) + ambient) * kd / np.pi
    if shader == 0:
        c = diffuse[:, None] * np.asarray(albedo, np.float64)[None]
    else:
        V = np.asarray(eye, np.float64)[None] - ref["P"]
        V /= np.linalg.norm(V, axis=1, keepdims=True)
        Hh = L[None] + V
        Hh /= np.linalg.norm(Hh, axis=1, keepdims=True)
        spec = (N * Hh).sum(1) ** shininess * ks
        alb = np.ones_like(ref["C"]) if shader == 1 else ref["C"]
        c = diffuse[:, None] * alb + spec[:, None]
    return np.trunc(np.clip(c, 0.0, 1.0) * 255.0 + 0.5)
