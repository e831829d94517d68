"""Host-side file formats: Mesh::LoadMesh's OBJ input (Utils/Mesh.cpp:11-34)
and WriteFrameToFile's BMP output (Core/Renderer.cpp:352-358).

EDXUtil's ObjMesh and Bitmap are absent, so the semantics are the ones the
header declares; these tests pin them by closed forms (transforms, fan
triangulation, plane normals), by round trips (mesh -> OBJ text -> mesh,
image -> BMP -> decoded image) and by the BMP file layout itself.
"""
import os
import struct

import numpy as np
import pytest

from edxraster_amd import Mesh, configs
from edxraster_amd import _lib as L


def write(path, text):
    with open(path, "w") as f:
        f.write(text)
    return str(path)


def decode_bmp(path):
    """(H, W, 3) RGB rows bottom-up, from a 24-bit BI_RGB file."""
    data = open(path, "rb").read()
    assert data[:2] == b"BM"
    size, off = struct.unpack_from("<I", data, 2)[0], struct.unpack_from("<I", data, 10)[0]
    hdr, w, h, planes, bpp, comp, img = struct.unpack_from("<IiiHHII", data, 14)
    assert (hdr, planes, bpp, comp) == (40, 1, 24, 0)
    assert size == len(data) and off == 54
    row = (3 * w + 3) & ~3
    assert img == row * h and len(data) == 54 + row * h
    out = np.zeros((h, w, 3), np.uint8)
    for y in range(h):
        line = np.frombuffer(data, np.uint8, 3 * w, 54 + y * row).reshape(w, 3)
        out[y] = line[:, ::-1]
    return out


def test_obj_quad_fan_transform_and_normals(built, tmp_path):
    obj = write(tmp_path / "quad.obj", """# a unit square in z = 0, one polygon
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
f 1 2 3 4
""")
    m = Mesh.load_obj(obj)
    assert m.num_vertices == 4 and m.num_triangles == 2
    np.testing.assert_array_equal(m.indices, [0, 1, 2, 0, 2, 3])
    np.testing.assert_array_equal(m.positions(), [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]])
    # counter-clockwise seen from +z: the area-weighted normal is +z everywhere
    np.testing.assert_array_equal(np.stack([m.nx, m.ny, m.nz], 1), np.tile([0, 0, 1], (4, 1)))
    np.testing.assert_array_equal(np.stack([m.cr, m.cg, m.cb], 1), np.full((4, 3), np.float32(0.9)))
    # p' = T R S p with R = Rz(90): (1,0,0) -> S (2,0,0) -> R (0,2,0) -> T (1,4,3)
    t = Mesh.load_obj(obj, pos=(1, 2, 3), scl=(2, 1, 1), rot=(0, 0, 90))
    np.testing.assert_allclose(t.positions()[1], [1, 4, 3], atol=1e-6)
    np.testing.assert_allclose(t.positions()[3], [0, 2, 3], atol=1e-6)  # (0,1,0) -> (-1,0,0) + T
    np.testing.assert_allclose(np.stack([t.nx, t.ny, t.nz], 1)[0], [0, 0, 1], atol=1e-7)
    # Rx(90) turns the +z normal to -y; Ry(90) turns +z to +x
    np.testing.assert_allclose(np.stack([*[getattr(Mesh.load_obj(obj, rot=(90, 0, 0)), c)[0] for c in ("nx", "ny", "nz")]]),
                               [0, -1, 0], atol=1e-7)
    np.testing.assert_allclose(np.stack([*[getattr(Mesh.load_obj(obj, rot=(0, 90, 0)), c)[0] for c in ("nx", "ny", "nz")]]),
                               [1, 0, 0], atol=1e-7)


def test_obj_materials_indices_and_normals(built, tmp_path):
    write(tmp_path / "two.mtl", """newmtl red
Kd 1 0 0
newmtl blue
Kd 0 0 0.5
""")
    obj = write(tmp_path / "two.obj", """mtllib two.mtl
v 0 0 0
v 1 0 0
v 0 1 0
v 1 1 0
vn 0 0 -1
vt 0.5 0.5
usemtl red
f 1//1 2//1 3//1
usemtl blue
f -2/1/1 -1/1/1 -3/1/1
usemtl missing
f 1 2 4
""")
    m = Mesh.load_obj(obj)
    # vertex 2 (index 1) and 3 (index 2) appear under both materials: distinct output vertices
    assert m.num_triangles == 3
    assert m.num_vertices == 3 + 3 + 3
    cols = np.stack([m.cr, m.cg, m.cb], 1)
    np.testing.assert_array_equal(cols[m.indices[0:3]], np.tile([1, 0, 0], (3, 1)))
    np.testing.assert_array_equal(cols[m.indices[3:6]], np.tile(np.float32([0, 0, 0.5]), (3, 1)))
    np.testing.assert_array_equal(cols[m.indices[6:9]], np.full((3, 3), np.float32(0.9)))
    # negative indices: -2 -> v3, -1 -> v4, -3 -> v2
    np.testing.assert_array_equal(m.positions()[m.indices[3:6]], [[0, 1, 0], [1, 1, 0], [1, 0, 0]])
    # given normals are used as written
    np.testing.assert_array_equal(np.stack([m.nx, m.ny, m.nz], 1)[m.indices[0:6]], np.tile([0, 0, -1], (6, 1)))


def test_obj_round_trip_of_a_sphere(built, tmp_path):
    s = Mesh.sphere(24, 12, 1.5, (0.25, -0.5, 2.0), color_seed=5)
    lines = ["v %.9g %.9g %.9g" % tuple(p) for p in s.positions()]
    lines += ["vn %.9g %.9g %.9g" % (a, b, c) for a, b, c in zip(s.nx, s.ny, s.nz)]
    lines += ["f %d//%d %d//%d %d//%d" % (a + 1, a + 1, b + 1, b + 1, c + 1, c + 1) for a, b, c in s.indices.reshape(-1, 3)]
    m = Mesh.load_obj(write(tmp_path / "s.obj", "\n".join(lines) + "\n"))
    assert (m.num_vertices, m.num_triangles) <= (s.num_vertices, s.num_triangles)
    # vertices renumbered by first use; map back through the index buffers
    remap = np.full(m.num_vertices, -1)
    remap[m.indices] = s.indices
    np.testing.assert_array_equal(m.positions(), s.positions()[remap])
    n_in = np.stack([s.nx, s.ny, s.nz], 1)[remap]
    np.testing.assert_allclose(np.stack([m.nx, m.ny, m.nz], 1), n_in / np.linalg.norm(n_in, axis=1)[:, None],
                               atol=1e-7)
    np.testing.assert_array_equal(remap[m.indices], s.indices)


def test_obj_errors(built, tmp_path):
    with pytest.raises(RuntimeError):
        Mesh.load_obj(tmp_path / "absent.obj")
    with pytest.raises(RuntimeError):
        Mesh.load_obj(write(tmp_path / "bad.obj", "v 0 0 0\nv 1 0 0\nf 1 2\n"))
    with pytest.raises(RuntimeError):
        Mesh.load_obj(write(tmp_path / "range.obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n"))
    hm = L.HostMesh()
    z = np.zeros(3, np.float32)
    assert L.lib().edxr_load_obj(os.fsencode(tmp_path / "absent.obj"), L.fptr(z), L.fptr(z + 1), L.fptr(z),
                                 hm) == L.EDXR_E_IO


@pytest.mark.parametrize("w,h", [(3, 2), (4, 1), (17, 5)])
def test_bmp_layout_and_round_trip(built, tmp_path, w, h):
    rng = np.random.default_rng(w * 31 + h)
    rgba = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    path = tmp_path / "f.bmp"
    assert L.lib().edxr_write_bmp(os.fsencode(path), rgba.ctypes.data, w, h) == 0
    np.testing.assert_array_equal(decode_bmp(path), rgba[:, :, :3])
    data = open(path, "rb").read()
    # first stored pixel is row 0 of the bottom-up buffer, as B, G, R
    assert tuple(data[54:57]) == (rgba[0, 0, 2], rgba[0, 0, 1], rgba[0, 0, 0])


@pytest.mark.gpu
def test_obj_mesh_renders_like_the_oracle_and_frames_are_written(built, tmp_path):
    """An OBJ scene (two materials, polygons, generated normals) through the
    GPU pipeline against the oracle, with SetWriteFrames on: each frame's BMP
    equals the back buffer, numbered by the frame counter."""
    from oracle.oracle import Oracle
    from edxraster_amd import Renderer
    from tests.test_gpu_parity import assert_frame_parity
    g = Mesh.load_heightfield(40, 0.3, 7, 9)
    write(tmp_path / "hf.mtl", "newmtl a\nKd 0.8 0.3 0.2\nnewmtl b\nKd 0.2 0.6 0.9\n")
    lines = ["mtllib hf.mtl"] + ["v %.9g %.9g %.9g" % tuple(p) for p in g.positions()]
    tris = g.indices.reshape(-1, 3)
    for k, (a, b, c) in enumerate(tris):
        if k % 2 == 0:
            lines.append("usemtl %s" % ("a" if (k // 64) % 2 else "b"))
        lines.append("f %d %d %d" % (a + 1, b + 1, c + 1))
    m = Mesh.load_obj(write(tmp_path / "hf.obj", "\n".join(lines) + "\n"), rot=(-30, 0, 0))
    w = configs.Workload("obj_heightfield", 320, 240, m, *configs.camera(320, 240, (0.0, 1.2, -2.2), (0, 0, 0)))
    r = Renderer(w.width, w.height)
    w.apply(r)
    r.set_write_frames(True, str(tmp_path))
    for _ in range(2):
        r.render_mesh(w.mesh)
    o = Oracle()
    o.render_workload(w)
    info = assert_frame_parity(r, o, "obj")
    assert info["covered"] > 1000
    assert r.frame_count == 2
    frames = sorted(os.listdir(tmp_path / "Frames"))
    assert frames == ["Frame00000.bmp", "Frame00001.bmp"]
    rgb = r.get_back_buffer()[:, :, :3]
    for f in frames:
        np.testing.assert_array_equal(decode_bmp(tmp_path / "Frames" / f), rgb)
    r.write_frame_to_file()  # explicit write: the counter has advanced to 2
    assert os.path.exists(tmp_path / "Frames" / "Frame00002.bmp")
    r.close()


def test_obj_texture_coordinates_slots_and_materials(built, tmp_path):
    write(tmp_path / "m.mtl", "newmtl img\nKd 0.5 0.5 0.5\nmap_Kd tex/checker.bmp\nnewmtl flat\nKd 0.1 0.2 0.3\n")
    obj = write(tmp_path / "t.obj", """mtllib m.mtl
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
vt 0 0
vt 1 0
vt 1 1
vt 0 1
vt 0.5 0.5
usemtl img
f 1/1 2/2 3/3
usemtl flat
f 1/5 3/3 4/4
f 2/2 3/3 4
""")
    m = Mesh.load_obj(obj)
    assert m.num_triangles == 3
    np.testing.assert_array_equal(m.tex_ids, [0, 1, 1])
    uv = np.stack([m.tu, m.tv], 1)
    np.testing.assert_array_equal(uv[m.indices[0:3]], [[0, 0], [1, 0], [1, 1]])
    np.testing.assert_array_equal(uv[m.indices[3:6]], [[0.5, 0.5], [1, 1], [0, 1]])
    np.testing.assert_array_equal(uv[m.indices[8]], [0, 0])  # no texture coordinate given
    # position 1 appears with two coordinates: two vertices
    assert len({tuple(uv[i]) for i in m.indices if tuple(m.positions()[i]) == (0.0, 0.0, 0.0)}) == 2
    mats = Mesh.obj_materials(obj)
    assert [x[0] for x in mats] == ["img", "flat"]
    np.testing.assert_array_equal(mats[1][1], np.float32([0.1, 0.2, 0.3]))
    assert mats[0][2] == str(tmp_path / "tex" / "checker.bmp") and mats[1][2] == ""
    # faces without a material use the slot after the materials
    m2 = Mesh.load_obj(write(tmp_path / "n.obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n"))
    np.testing.assert_array_equal(m2.tex_ids, [0])


def test_bmp_reader_round_trip_and_top_down(built, tmp_path):
    from edxraster_amd import load_bmp
    rng = np.random.default_rng(5)
    rgba = rng.integers(0, 256, (7, 9, 4), dtype=np.uint8)
    assert L.lib().edxr_write_bmp(os.fsencode(tmp_path / "a.bmp"), rgba.ctypes.data, 9, 7) == 0
    back = load_bmp(tmp_path / "a.bmp")
    np.testing.assert_array_equal(back[:, :, :3], rgba[:, :, :3])
    assert (back[:, :, 3] == 255).all()
    # a 32-bit top-down file (negative height) comes back bottom-up with alpha
    w, h = 3, 2
    px = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)  # file rows top-down, BGRA
    hdr = b"BM" + struct.pack("<IHHI", 54 + 4 * w * h, 0, 0, 54) + \
        struct.pack("<IiiHHIIiiII", 40, w, -h, 1, 32, 0, 4 * w * h, 2835, 2835, 0, 0)
    (tmp_path / "td.bmp").write_bytes(hdr + px.tobytes())
    got = load_bmp(tmp_path / "td.bmp")
    np.testing.assert_array_equal(got, px[::-1][:, :, [2, 1, 0, 3]])
    (tmp_path / "bad.bmp").write_bytes(b"BM" + b"\0" * 20)
    with pytest.raises(RuntimeError):
        load_bmp(tmp_path / "bad.bmp")


def write_textured_obj(tmp_path, place=None):
    """An OBJ heightfield with texture coordinates and a .bmp diffuse map
    (ImageTexture) beside a Kd-only material (ConstantTexture2D); `place`
    maps the (V, 3) positions before they are written."""
    (tmp_path / "tex").mkdir()
    img = configs.checker_texture(128, 8, 3)
    assert L.lib().edxr_write_bmp(os.fsencode(tmp_path / "tex" / "c.bmp"), img.ctypes.data, 128, 128) == 0
    write(tmp_path / "m.mtl", "newmtl a\nKd 0.4 0.4 0.4\nmap_Kd tex/c.bmp\nnewmtl b\nKd 0.9 0.5 0.2\n")
    g = Mesh.load_heightfield(48, 0.3, 7, 9)
    P = g.positions() if place is None else place(g.positions().astype(np.float64))
    lines = ["mtllib m.mtl"] + ["v %.9g %.9g %.9g" % tuple(p) for p in P]
    lines += ["vt %.9g %.9g" % ((x + 1) * 2, (z + 1) * 2) for x, z in zip(g.px, g.pz)]
    for k, (a, b, c) in enumerate(g.indices.reshape(-1, 3)):
        if k % 200 == 0:
            lines.append("usemtl %s" % ("a" if (k // 200) % 3 else "b"))
        lines.append("f %d/%d %d/%d %d/%d" % (a + 1, a + 1, b + 1, b + 1, c + 1, c + 1))
    return write(tmp_path / "hf.obj", "\n".join(lines) + "\n")


TEXTURED_OBJ_CAMERA = ((0.0, 1.2, -2.2), (0.0, 0.0, 0.0))


@pytest.mark.gpu
def test_obj_with_bmp_diffuse_map_renders_like_the_oracle(built, tmp_path):
    """The textured OBJ through the textured shader, its slots taken from the
    OBJ's materials, against the oracle."""
    from oracle.oracle import Oracle
    from edxraster_amd import Renderer
    from tests.test_gpu_parity import assert_frame_parity
    path = write_textured_obj(tmp_path)
    m = Mesh.load_obj(path)
    slots = Mesh.obj_texture_slots(path)
    assert isinstance(slots[0], np.ndarray) and len(slots) == 3
    w = configs.Workload("obj_textured", 320, 240, m, *configs.camera(320, 240, *TEXTURED_OBJ_CAMERA),
                        shader=L.SHADER_TEXTURED_LAMBERT, shininess=0.0, textures=slots)
    r = Renderer(w.width, w.height)
    w.apply(r)
    r.render_mesh(w.mesh)
    o = Oracle()
    o.render_workload(w)
    info = assert_frame_parity(r, o, "obj textured")
    assert info["covered"] > 1000
    r.close()


def tga_bytes(img, rle=False, top_down=False, bits=32):
    """A TGA file of an (H, W, 4) RGBA image whose rows are bottom-up (v = 0
    first), written independently of the library: type 2 or type 10 (with
    runs of repeated pixels where the image repeats), BGR(A) bytes."""
    H, W, _ = img.shape
    px = bits // 8
    rows = img[::-1] if top_down else img
    pix = np.concatenate([rows[..., 2::-1], rows[..., 3:4]], axis=-1)[..., :px].reshape(-1, px)
    hdr = bytes([0, 0, 10 if rle else 2, 0, 0, 0, 0, 0, 0, 0, 0, 0, W & 255, W >> 8, H & 255, H >> 8, bits,
                 (0x20 if top_down else 0) | (8 if bits == 32 else 0)])
    if not rle:
        return hdr + pix.tobytes()
    out = bytearray(hdr)
    i, n = 0, pix.shape[0]
    while i < n:
        j = i + 1
        while j < n and j - i < 128 and np.array_equal(pix[j], pix[i]):
            j += 1
        if j - i >= 2:  # a run
            out += bytes([0x80 | (j - i - 1)]) + pix[i].tobytes()
            i = j
        else:           # literals up to the next repeat
            j = i + 1
            while j < n and j - i < 128 and not (j + 1 < n and np.array_equal(pix[j], pix[j + 1])):
                j += 1
            out += bytes([j - i - 1]) + pix[i:j].tobytes()
            i = j
    return bytes(out)


@pytest.mark.parametrize("rle", [False, True], ids=["uncompressed", "rle"])
@pytest.mark.parametrize("top_down", [False, True], ids=["bottom_up", "top_down"])
@pytest.mark.parametrize("bits", [24, 32])
def test_tga_reader_round_trip(built, tmp_path, rle, top_down, bits):
    from edxraster_amd import load_image
    rng = np.random.default_rng(bits + 2 * rle + top_down)
    img = rng.integers(0, 256, (13, 21, 4), dtype=np.uint8)
    img[3:7, :, :] = img[3, 0]  # rows of one colour: runs for the RLE writer
    if bits == 24:
        img[..., 3] = 255
    p = tmp_path / "t.tga"
    p.write_bytes(tga_bytes(img, rle, top_down, bits))
    np.testing.assert_array_equal(load_image(p), img)


def test_ppm_reader_round_trip_and_errors(built, tmp_path):
    from edxraster_amd import load_image
    rng = np.random.default_rng(7)
    img = rng.integers(0, 256, (9, 14, 4), dtype=np.uint8)
    img[..., 3] = 255
    p = tmp_path / "t.ppm"
    # rows top-down in the file, a comment in the header
    p.write_bytes(b"P6\n# diffuse map\n14 9\n255\n" + img[::-1, :, :3].tobytes())
    np.testing.assert_array_equal(load_image(p), img)
    bad = tmp_path / "bad.ppm"
    bad.write_bytes(b"P6\n14 9\n65535\n" + bytes(9 * 14 * 6))
    with pytest.raises(L.EdxrError):
        load_image(bad)
    short = tmp_path / "short.tga"
    short.write_bytes(tga_bytes(img)[:-5])
    with pytest.raises(L.EdxrError):
        load_image(short)


def test_obj_with_tga_and_ppm_diffuse_maps_binds_images(built, tmp_path):
    """Mesh::LoadMesh's ImageTexture for .tga and .ppm diffuse maps: the slots
    hold the decoded images, not the Kd fallback."""
    img = configs.checker_texture(16, 4, 5)
    (tmp_path / "a.tga").write_bytes(tga_bytes(img, rle=True))
    (tmp_path / "b.ppm").write_bytes(b"P6\n16 16\n255\n" + img[::-1, :, :3].tobytes())
    write(tmp_path / "m.mtl", "newmtl a\nKd 0.1 0.2 0.3\nmap_Kd a.tga\nnewmtl b\nKd 0.4 0.5 0.6\nmap_Kd b.ppm\n")
    write(tmp_path / "q.obj", "mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl a\nf 1 2 3\nusemtl b\nf 1 3 2\n")
    slots = Mesh.obj_texture_slots(str(tmp_path / "q.obj"))
    assert len(slots) == 3
    np.testing.assert_array_equal(slots[0], img)
    np.testing.assert_array_equal(slots[1][..., :3], img[..., :3])
    assert (slots[1][..., 3] == 255).all()
