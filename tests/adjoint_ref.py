"""FP64 CPU reference of the differentiable path (torch float64 autograd).

image_fp64 evaluates, for the pixels a forward pass covered (its triangle-ID
buffer, visibility held fixed), the clamped colour of the winning input
triangle as a smooth function of the per-vertex positions, normals, colours,
the light direction and (ambient, kd, ks, shininess):

  * perspective-correct barycentrics of the pixel centre from the continuous
    projection: with (X, Y, W) = rows 0, 1, 3 of MVP (p, 1) per vertex and the
    pixel centre (x, y) in NDC, A = X - x W, B = Y - y W, q = A x B,
    b = q / sum(q) (the weights that put sum b_i clip_i on the pixel's ray);
  * P, N, C interpolated with b; the pixel shader of EDXRaster
    (Core/Shader.h:209-282): n = N/|N|, L = l/|l|, diffuse =
    (max(n.L, 0) + ambient) kd / pi, V = normalize(eye - P), H =
    normalize(L + V), spec = (n.H)^s ks, colour = diffuse * C + spec
    (vertex colours), clamped to [0, 1].

The gradient is torch's float64 reverse mode of this function, an
independent derivation of what the HIP adjoint computes by hand; central
finite differences of the same function pin it.
"""
import numpy as np
import torch

BG = 0xFFFFFFFF


def eye_of(mv):
    inv = np.linalg.inv(np.asarray(mv, np.float64).reshape(4, 4))
    return inv[:3, 3] / inv[3, 3]


def covered_pixels(ids):
    """(rows, xs, y, tri) of covered pixels of a bottom-up ID image."""
    H, W = ids.shape
    rows, xs = np.nonzero(ids != BG)
    return rows, xs, H - 1 - rows, ids[rows, xs].astype(np.int64)


def image_fp64(theta, indices, view, ids, shininess_int=None):
    """Clamped colours (N, 3) of the covered pixels of one view.  theta: dict
    of float64 tensors px..cb (V), light (3), material (4)."""
    rows, xs, ys, tri = covered_pixels(ids)
    idx = torch.from_numpy(indices.reshape(-1, 3)[tri].astype(np.int64))
    # The library multiplies proj * mv in float (k = 0..3 left to right); use
    # those float32 values.
    mvp32 = np.zeros(16, np.float32)
    a = np.asarray(view.proj, np.float32).reshape(4, 4)
    b = np.asarray(view.mv, np.float32).reshape(4, 4)
    for i in range(4):
        for j in range(4):
            mvp32[4 * i + j] = np.float32(np.float32(np.float32(a[i, 0] * b[0, j]) + np.float32(a[i, 1] * b[1, j]))
                                          + np.float32(a[i, 2] * b[2, j])) + np.float32(a[i, 3] * b[3, j])
    mvp = torch.from_numpy(mvp32.astype(np.float64).reshape(4, 4))
    R = np.asarray(view.raster, np.float32).astype(np.float64).reshape(4, 4)
    xr = xs + 0.5 - R[0, 3]
    yr = ys + 0.5 - R[1, 3]
    det = R[0, 0] * R[1, 1] - R[0, 1] * R[1, 0]
    xn = torch.from_numpy((R[1, 1] * xr - R[0, 1] * yr) / det)[:, None]
    yn = torch.from_numpy((R[0, 0] * yr - R[1, 0] * xr) / det)[:, None]

    def gather(names):
        return torch.stack([theta[n][idx] for n in names], dim=2)  # (N, 3 vertices, 3)

    p, n, c = gather(("px", "py", "pz")), gather(("nx", "ny", "nz")), gather(("cr", "cg", "cb"))
    X = p @ mvp[0, :3] + mvp[0, 3]
    Y = p @ mvp[1, :3] + mvp[1, 3]
    Wc = p @ mvp[3, :3] + mvp[3, 3]
    A = X - xn * Wc
    B = Y - yn * Wc
    q = torch.stack([A[:, 1] * B[:, 2] - A[:, 2] * B[:, 1], A[:, 2] * B[:, 0] - A[:, 0] * B[:, 2],
                     A[:, 0] * B[:, 1] - A[:, 1] * B[:, 0]], dim=1)
    bary = q / q.sum(dim=1, keepdim=True)
    Pw = (bary[:, :, None] * p).sum(1)
    Nw = (bary[:, :, None] * n).sum(1)
    Cw = (bary[:, :, None] * c).sum(1)
    nh = Nw / Nw.norm(dim=1, keepdim=True)
    l = theta["light"]
    Lh = l / l.norm()
    amb, kd, ks, s = theta["material"][0], theta["material"][1], theta["material"][2], theta["material"][3]
    nl = (nh * Lh).sum(1)
    diffuse = (torch.clamp(nl, min=0.0) + amb) * kd / np.pi
    eye = torch.from_numpy(eye_of(view.mv))
    v = eye[None, :] - Pw
    vh = v / v.norm(dim=1, keepdim=True)
    h = Lh[None, :] + vh
    hh = h / h.norm(dim=1, keepdim=True)
    sa = (nh * hh).sum(1)
    k = int(round(float(s.detach()))) if shininess_int is None else shininess_int
    # sa^s with the integer exponent of the forward pass; its s-derivative is
    # sa^s ln(sa) for sa > 0 and 0 otherwise (first-order term at s = k).
    logsa = torch.where(sa > 0, torch.log(torch.clamp(sa, min=1e-300)), torch.zeros_like(sa))
    powk = sa ** k
    spec = powk * (1.0 + (s - k) * logsa) * ks
    col = diffuse[:, None] * Cw + spec[:, None]
    return torch.clamp(col, 0.0, 1.0), (rows, xs)


def theta_from(mesh, light, material, requires_grad=True):
    t = {}
    for name in ("px", "py", "pz", "nx", "ny", "nz", "cr", "cg", "cb"):
        t[name] = torch.tensor(getattr(mesh, name).astype(np.float64), requires_grad=requires_grad)
    t["light"] = torch.tensor(np.asarray(light, np.float64), requires_grad=requires_grad)
    t["material"] = torch.tensor(np.asarray(material, np.float64), requires_grad=requires_grad)
    return t


def loss_fp64(theta, indices, views, ids_list, grads_list):
    """sum over views and covered pixels of d_rgb . colour."""
    total = torch.zeros((), dtype=torch.float64)
    for view, ids, g in zip(views, ids_list, grads_list):
        col, (rows, xs) = image_fp64(theta, indices, view, ids)
        total = total + (col * torch.from_numpy(g[rows, xs].astype(np.float64))).sum()
    return total
