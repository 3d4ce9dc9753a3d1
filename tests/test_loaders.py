"""Asset readers: Radiance .hdr against an independent RGBE encoder, Wavefront .obj/.mtl
against the meshes the scene generators build and against closed forms."""
import math

import numpy as np
import pytest

from axomae_amd import loaders, scenes


def _encode_rgbe(rgb):
    """Float RGB to Radiance RGBE bytes (Ward's encoding: the largest component's frexp)."""
    rgb = np.asarray(rgb, dtype=np.float64).reshape(-1, 3)
    out = np.zeros((len(rgb), 4), dtype=np.uint8)
    for i, px in enumerate(rgb):
        v = px.max()
        if v < 1e-32:
            continue
        m, e = math.frexp(v)
        scale = m * 256.0 / v
        out[i, :3] = np.floor(px * scale).astype(np.uint8)
        out[i, 3] = e + 128
    return out


def _decode_rgbe(b):
    b = b.astype(np.float64)
    return np.where(b[:, 3:4] > 0, b[:, :3] * np.ldexp(1.0, (b[:, 3:4] - 136).astype(np.int64)), 0.0)


def _rle_channel(row):
    """One channel of a scanline as new-style Radiance runs: runs of >= 3 equal bytes, else
    literal spans (at most 128 and 127 long)."""
    out, x, n = bytearray(), 0, len(row)
    while x < n:
        run = 1
        while x + run < n and run < 127 and row[x + run] == row[x]:
            run += 1
        if run >= 3:
            out += bytes([128 + run, row[x]])
            x += run
            continue
        lit = 1
        while x + lit < n and lit < 128 and not (x + lit + 2 < n and row[x + lit] == row[x + lit + 1] == row[x + lit + 2]):
            lit += 1
        out += bytes([lit]) + bytes(row[x:x + lit])
        x += lit
    return bytes(out)


def _write_hdr(path, w, h, rgbe, rle):
    with open(path, "wb") as f:
        f.write(b"#?RADIANCE\n# test map\nFORMAT=32-bit_rle_rgbe\nEXPOSURE=1.0\n\n")
        f.write(f"-Y {h} +X {w}\n".encode())
        px = rgbe.reshape(h, w, 4)
        for y in range(h):
            if rle:
                f.write(bytes([2, 2, w >> 8, w & 255]))
                for c in range(4):
                    f.write(_rle_channel(px[y, :, c]))
            else:
                f.write(px[y].tobytes())


@pytest.mark.parametrize("rle", [False, True])
def test_hdr_decodes_rgbe_exactly(tmp_path, rle):
    w, h = 37, 5
    rng = np.random.default_rng(3)
    rgb = rng.random((h * w, 3)) * np.exp(rng.normal(0, 3, (h * w, 1)))
    rgb[:w] = [0.25, 0.5, 4.0]  # a constant first row: long runs
    rgb[w + 3] = 0.0
    rgbe = _encode_rgbe(rgb)
    p = tmp_path / "map.hdr"
    _write_hdr(p, w, h, rgbe, rle)
    tex = loaders.load_hdr(str(p))
    assert (tex.width, tex.height, tex.channels) == (w, h, 4) and tex.is_f32
    got = tex.data.reshape(-1, 4)
    np.testing.assert_array_equal(got[:, :3], _decode_rgbe(rgbe).astype(np.float32))
    np.testing.assert_array_equal(got[:, 3], 1.0)
    nz = rgb.max(axis=1) > 0
    err = np.abs(got[nz, :3] - rgb[nz]).max(axis=1) / rgb[nz].max(axis=1)
    assert err.max() < 1.0 / 128  # RGBE keeps 8 bits of the largest component


def test_hdr_rejects_other_orientations(tmp_path):
    p = tmp_path / "flip.hdr"
    p.write_bytes(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n+Y 1 +X 1\n\x80\x80\x80\x81")
    with pytest.raises(ValueError):
        loaders.load_hdr(str(p))


def _write(path, text):
    path.write_text(text)
    return str(path)


def test_obj_quad_matches_the_generator_quad(tmp_path):
    """A quad polygon with normals and UVs written so that FlipUVs gives quad_mesh's UVs
    reads back as exactly quad_mesh's arrays: fan triangulation (0 1 2, 0 2 3), one shared
    vertex per corner, the normal, and a tangent frame along +u / +v."""
    _write(tmp_path / "q.mtl", "newmtl wall\nKd 0.6 0.6 0.6\nNs 10\nKe 0.25 0.5 0.75\n")
    obj = _write(tmp_path / "q.obj", """mtllib q.mtl
v -1 -1 0
v 1 -1 0
v 1 1 0
v -1 1 0
vt 0 1
vt 1 1
vt 1 0
vt 0 0
vn 0 0 1
usemtl wall
f 1/1/1 2/2/1 3/3/1 4/4/1
""")
    sc = scenes.Scene()
    (mi,) = loaders.load_obj(obj, sc)
    got = sc.meshes[mi]
    ref = scenes.quad_mesh([(-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0)], got.material)
    for field in ("vertices", "normals", "tangents", "bitangents", "uv", "indices"):
        np.testing.assert_array_equal(getattr(got, field), getattr(ref, field), err_msg=field)
    m = sc.materials[got.material]
    assert m.emissive >= 0 and m.eta == (0.0, 0.0, 0.0)
    rough = (int(sc.textures[m.metallic].data[0]) >> 8) & 255
    assert rough == round(math.sqrt(2.0 / 12.0) * 255)


def test_obj_flat_normals_fans_and_negative_indices(tmp_path):
    """No normals: every face keeps its own corners with the face normal. A pentagon is a
    fan of three triangles; negative indices count back from the last vertex."""
    obj = _write(tmp_path / "p.obj", "\n".join(["v %f %f 0" % (math.cos(2 * math.pi * k / 5), math.sin(2 * math.pi * k / 5)) for k in range(5)]
                                              + ["v 0 0 1", "f -6 -5 -4 -3 -2", "f 1 2 6"]) + "\n")
    sc = scenes.Scene()
    (mi,) = loaders.load_obj(obj, sc)
    m = sc.meshes[mi]
    idx = m.indices.reshape(-1, 3)
    assert len(idx) == 4 and m.vertices.size // 3 == 8
    np.testing.assert_array_equal(idx[:3], [[0, 1, 2], [0, 2, 3], [0, 3, 4]])
    n = m.normals.reshape(-1, 3)
    np.testing.assert_allclose(n[:5], np.tile([0, 0, 1], (5, 1)), atol=1e-7)
    p = m.vertices.reshape(-1, 3).astype(np.float64)
    fn = np.cross(p[6] - p[5], p[7] - p[5])
    np.testing.assert_allclose(n[5:], np.tile(fn / np.linalg.norm(fn), (3, 1)), atol=1e-6)
    t, b = m.tangents.reshape(-1, 3), m.bitangents.reshape(-1, 3)
    np.testing.assert_allclose(np.sum(t * n, axis=1), 0, atol=1e-6)
    np.testing.assert_allclose(np.linalg.norm(t, axis=1), 1, atol=1e-6)
    np.testing.assert_allclose(np.linalg.norm(b, axis=1), 1, atol=1e-6)


def test_obj_tangents_follow_the_uv_map(tmp_path):
    """A triangle whose UVs grow along a rotated pair of axes: the tangent is the unit dp/du
    and the bitangent points along dp/dv (the plane is z = 0)."""
    c, s = math.cos(0.4), math.sin(0.4)
    obj = _write(tmp_path / "t.obj", f"""v 0 0 0
v {c} {s} 0
v {-s} {c} 0
vt 0.2 0.3
vt 1.2 0.3
vt 0.2 1.3
vn 0 0 1
f 1/1/1 2/2/1 3/3/1
""")
    sc = scenes.Scene()
    (mi,) = loaders.load_obj(obj, sc)
    m = sc.meshes[mi]
    t, b = m.tangents.reshape(-1, 3), m.bitangents.reshape(-1, 3)
    np.testing.assert_allclose(t, np.tile([c, s, 0], (3, 1)), atol=1e-6)
    np.testing.assert_allclose(b, np.tile([-(-s), -c, 0], (3, 1)), atol=1e-6)  # v flipped: dp/dv' = -(-s, c)
    np.testing.assert_allclose(m.uv.reshape(-1, 2), [[0.2, 0.7], [1.2, 0.7], [0.2, -0.3]], atol=1e-6)


def test_mtl_diffuse_map_becomes_a_packed_texture(tmp_path):
    from PIL import Image

    px = np.array([[[255, 0, 0, 255], [0, 255, 0, 128]], [[0, 0, 255, 255], [10, 20, 30, 40]]], dtype=np.uint8)
    Image.fromarray(px, "RGBA").save(tmp_path / "albedo.png")
    _write(tmp_path / "m.mtl", "newmtl tex\nKd 1 1 1\nmap_Kd albedo.png\nd 0.5\nNi 1.45\n")
    obj = _write(tmp_path / "m.obj", "mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 0 1\nusemtl tex\nf 1/1 2/2 3/3\n")
    sc = scenes.Scene()
    (mi,) = loaders.load_obj(obj, sc)
    mat = sc.materials[sc.meshes[mi].material]
    tex = sc.textures[mat.albedo]
    assert (tex.width, tex.height) == (2, 2) and tex.data.dtype == np.uint32
    np.testing.assert_array_equal(tex.data, scenes.pack_rgba8(px[..., 0], px[..., 1], px[..., 2], px[..., 3]).reshape(-1))
    assert mat.eta == pytest.approx((1.45,) * 3)


def test_loaded_scene_renders_like_the_generated_one(oracle, tmp_path):
    """The wall quad loaded from .obj/.mtl and from the generator give the same image."""
    _write(tmp_path / "w.mtl", "newmtl wall\nKd 0.6 0.6 0.6\nPr 0.5\nKe 0.25 0.5 0.75\n")
    obj = _write(tmp_path / "w.obj", "mtllib w.mtl\nv -1 -1 0\nv 1 -1 0\nv 1 1 0\nv -1 1 0\nvt 0 1\nvt 1 1\nvt 1 0\nvt 0 0\nvn 0 0 1\n"
                                     "usemtl wall\nf 1/1/1 2/2/1 3/3/1 4/4/1\n")
    imgs = []
    for load in (True, False):
        sc = scenes.Scene(name="wall")
        if load:
            loaders.load_obj(obj, sc)
        else:
            m = sc.principled((0.6, 0.6, 0.6), roughness=0.5, emission=(0.25, 0.5, 0.75))
            sc.meshes.append(scenes.quad_mesh([(-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0)], m))
        sc.envmap = sc.add_texture(scenes.Texture(2, 2, 4, np.full(16, 0.5, dtype=np.float32)))
        scenes.set_camera(sc, (0.1, 0.05, 4.0), (0.1, 0.05, 0.0), 40.0, 1.0)
        _, img = oracle.OracleScene(sc).render(24, 24, 2, 3)
        imgs.append(img)
    np.testing.assert_array_equal(imgs[0], imgs[1])
