"""Asset readers that feed :class:`axomae_amd.scenes.Scene`: Radiance ``.hdr`` environment
maps and Wavefront ``.obj`` meshes with their ``.mtl`` materials.

The reference imports assets through third-party libraries that are not part of this
build. Environment maps go through stb_image's ``stbi_loadf``
(``Loader_image.cpp:145-175``); three-channel data gets an alpha of 1. Meshes go through
assimp with ``aiProcess_CalcTangentSpace | aiProcess_Triangulate |
aiProcess_JoinIdenticalVertices | aiProcess_FlipUVs`` (``Loader_scene.cpp:84-85``). The
readers here follow the published file formats and apply the same post-processing steps:
- polygons are split into triangle fans;
- corners with identical (position, uv, normal) indices share one vertex;
- v becomes 1 - v;
- per-vertex tangents and bitangents come from the UV derivatives of the incident faces,
  made orthogonal to the normal.
Two differences from assimp are deliberate. A mesh without normals gets flat face normals,
where assimp would leave none. assimp's 45-degree tangent smoothing limit is not
reproduced: incident faces are averaged by area.
"""
import math
import os

import numpy as np

from .scenes import Mesh, Texture, pack_rgba8

# ----------------------------------------------------------------- Radiance HDR


def _rgbe_to_float(rgbe):
    """Radiance RGBE to linear float, as stb_image converts it: a mantissa byte times
    2^(e - 136); a zero exponent is black."""
    rgbe = np.asarray(rgbe, dtype=np.uint8).reshape(-1, 4)
    e = rgbe[:, 3].astype(np.int32)
    scale = np.where(e > 0, np.ldexp(np.float32(1.0), e - 136), 0.0).astype(np.float32)
    return rgbe[:, :3].astype(np.float32) * scale[:, None]


def _read_scanline(buf, pos, width):
    """One scanline of RGBE pixels starting at buf[pos]; returns (pixels (width, 4), new pos).
    New-style run-length scanlines start with 2, 2 and the width as a 16-bit big-endian
    number, then hold each of the four channels as runs (count > 128: count - 128 copies of
    one byte) or literal spans (count bytes). Anything else is flat 4-byte pixels."""
    if 8 <= width < 32768 and pos + 4 <= len(buf) and buf[pos] == 2 and buf[pos + 1] == 2 and (buf[pos + 2] & 0x80) == 0:
        if (buf[pos + 2] << 8 | buf[pos + 3]) != width:
            raise ValueError("Radiance scanline width does not match the header")
        pos += 4
        out = np.empty((4, width), dtype=np.uint8)
        for c in range(4):
            x = 0
            while x < width:
                count = buf[pos]
                pos += 1
                if count > 128:
                    count -= 128
                    if x + count > width:
                        raise ValueError("Radiance run overflows the scanline")
                    out[c, x:x + count] = buf[pos]
                    pos += 1
                else:
                    if count == 0 or x + count > width:
                        raise ValueError("bad Radiance literal span")
                    out[c, x:x + count] = np.frombuffer(buf, dtype=np.uint8, count=count, offset=pos)
                    pos += count
                x += count
        return out.T, pos
    if pos + 4 * width > len(buf):
        raise ValueError("truncated Radiance data")
    return np.frombuffer(buf, dtype=np.uint8, count=4 * width, offset=pos).reshape(width, 4), pos + 4 * width


def load_hdr(path):
    """A Radiance ``.hdr`` file as an RGBA32F :class:`Texture` (alpha 1), rows top to bottom
    as stored (stb_image's default orientation)."""
    with open(path, "rb") as f:
        buf = f.read()
    pos = 0

    def line():
        nonlocal pos
        end = buf.index(b"\n", pos)
        s = buf[pos:end].decode("ascii", "replace").strip()
        pos = end + 1
        return s

    magic = line()
    if magic not in ("#?RADIANCE", "#?RGBE"):
        raise ValueError(f"{path}: not a Radiance file")
    while True:
        s = line()
        if not s:
            break
        if s.startswith("FORMAT=") and s != "FORMAT=32-bit_rle_rgbe":
            raise ValueError(f"{path}: unsupported {s}")
    res = line().split()
    if len(res) != 4 or res[0] != "-Y" or res[2] != "+X":
        raise ValueError(f"{path}: unsupported orientation {' '.join(res)}")
    height, width = int(res[1]), int(res[3])
    rows = []
    for _ in range(height):
        px, pos = _read_scanline(buf, pos, width)
        rows.append(px)
    rgb = _rgbe_to_float(np.concatenate(rows))
    rgba = np.ones((width * height, 4), dtype=np.float32)
    rgba[:, :3] = rgb
    return Texture(width, height, 4, rgba.reshape(-1))


# ----------------------------------------------------------------- Wavefront OBJ


def _load_image_u32(path):
    """An LDR image as the reference's packed u32 texels ((a << 24) | (r << 16) | (g << 8) | b)."""
    from PIL import Image

    with Image.open(path) as im:
        a = np.asarray(im.convert("RGBA"), dtype=np.uint32)
    h, w = a.shape[:2]
    texels = pack_rgba8(a[..., 0], a[..., 1], a[..., 2], a[..., 3]).reshape(-1)
    return Texture(w, h, 4, np.ascontiguousarray(texels, dtype=np.uint32))


def _parse_mtl(path):
    mats, cur = {}, None
    with open(path) as f:
        for raw in f:
            tok = raw.split("#", 1)[0].split()
            if not tok:
                continue
            key, args = tok[0], tok[1:]
            if key == "newmtl":
                cur = mats.setdefault(" ".join(args), {})
            elif cur is not None and key in ("Kd", "Ke"):
                cur[key] = tuple(float(a) for a in args[:3])
            elif cur is not None and key in ("Ns", "Ni", "d", "Tr", "Pr", "Pm"):
                cur[key] = float(args[0])
            elif cur is not None and key == "map_Kd":
                cur[key] = os.path.join(os.path.dirname(path), args[-1])
    return mats


def _material(scene, spec):
    """Principled material of one .mtl entry. Roughness is Pr when given, else
    sqrt(2 / (Ns + 2)) (the Phong exponent's usual GGX equivalent); metal is Pm; a dissolve
    below 1 (d, or 1 - Tr) with an index of refraction Ni makes it a dielectric."""
    kd = spec.get("Kd", (0.8, 0.8, 0.8))
    rough = spec.get("Pr", math.sqrt(2.0 / (spec["Ns"] + 2.0)) if "Ns" in spec else 0.5)
    d = spec.get("d", 1.0 - spec.get("Tr", 0.0))
    eta = spec.get("Ni") if d < 1.0 and spec.get("Ni", 1.0) > 0.0 else None
    ke = spec.get("Ke")
    emission = ke if ke is not None and max(ke) > 0.0 else None
    scale = 1.0
    if emission is not None and max(emission) > 1.0:  # u32 texels hold [0, 1]: the excess goes to the scale
        scale = max(emission)
        emission = tuple(c / scale for c in emission)
    m = scene.principled(kd, roughness=min(max(rough, 0.0), 1.0), metal=spec.get("Pm", 0.0), emission=emission, emission_scale=scale, eta=eta)
    if "map_Kd" in spec:
        scene.materials[m].albedo = scene.add_texture(_load_image_u32(spec["map_Kd"]))
    return m


def _tangent_space(pos, nrm, uv, tris):
    """Per-vertex tangent and bitangent: each face's UV derivatives (dp/du, dp/dv), summed
    over the incident faces with the face area as weight, then Gram-Schmidt against the
    normal; the bitangent keeps the side of the summed dp/dv. Faces with degenerate UVs add
    nothing; a vertex with no UV direction takes an arbitrary frame about its normal."""
    p0, p1, p2 = pos[tris[:, 0]], pos[tris[:, 1]], pos[tris[:, 2]]
    t0, t1, t2 = uv[tris[:, 0]], uv[tris[:, 1]], uv[tris[:, 2]]
    e1, e2 = p1 - p0, p2 - p0
    d1, d2 = t1 - t0, t2 - t0
    det = d1[:, 0] * d2[:, 1] - d2[:, 0] * d1[:, 1]
    ok = np.abs(det) > 1e-20
    inv = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0)[:, None]
    dpdu = (e1 * d2[:, 1:2] - e2 * d1[:, 1:2]) * inv
    dpdv = (e2 * d1[:, 0:1] - e1 * d2[:, 0:1]) * inv
    area = 0.5 * np.linalg.norm(np.cross(e1, e2), axis=1)[:, None]
    tan = np.zeros_like(pos)
    bit = np.zeros_like(pos)
    for k in range(3):
        np.add.at(tan, tris[:, k], dpdu * area)
        np.add.at(bit, tris[:, k], dpdv * area)
    t = tan - nrm * np.sum(tan * nrm, axis=1, keepdims=True)
    tl = np.linalg.norm(t, axis=1, keepdims=True)
    up = np.where(np.abs(nrm[:, 1:2]) < 0.999, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    fallback = np.cross(up, nrm)
    fallback /= np.linalg.norm(fallback, axis=1, keepdims=True)
    t = np.where(tl > 1e-12, t / np.maximum(tl, 1e-30), fallback)
    b = np.cross(nrm, t)
    b = np.where(np.sum(b * bit, axis=1, keepdims=True) < 0.0, -b, b)
    return t, b


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1), dtype=np.float32)


def load_obj(path, scene, default_material=None):
    """Append the meshes of a Wavefront ``.obj`` file (one mesh per material used) and the
    materials of its ``.mtl`` libraries to `scene`; returns the new mesh indices. Faces that
    name no material use `default_material`, or a grey diffuse one when that is None."""
    base = os.path.dirname(path)
    v, vt, vn = [], [], []
    groups = {}  # material name -> list of faces, a face = list of (vi, ti, ni)
    libs = {}
    cur = None
    with open(path) as f:
        for raw in f:
            tok = raw.split("#", 1)[0].split()
            if not tok:
                continue
            key, args = tok[0], tok[1:]
            if key == "v":
                v.append([float(a) for a in args[:3]])
            elif key == "vt":
                vt.append([float(a) for a in (args + ["0"])[:2]])
            elif key == "vn":
                vn.append([float(a) for a in args[:3]])
            elif key == "f":
                face = []
                for a in args:
                    parts = (a.split("/") + ["", ""])[:3]
                    idx = []
                    for p, n in zip(parts, (len(v), len(vt), len(vn))):
                        if p == "":
                            idx.append(-1)
                        else:
                            i = int(p)
                            idx.append(i - 1 if i > 0 else n + i)
                    face.append(tuple(idx))
                if len(face) >= 3:
                    groups.setdefault(cur, []).append(face)
            elif key == "usemtl":
                cur = " ".join(args)
            elif key == "mtllib":
                for name in args:
                    p = os.path.join(base, name)
                    if os.path.exists(p):
                        libs.update(_parse_mtl(p))
    v, vt, vn = (np.asarray(a, dtype=np.float64).reshape(-1, w) for a, w in ((v, 3), (vt, 2), (vn, 3)))
    mat_ids = {}
    out = []
    for name, faces in groups.items():
        if name in libs:
            if name not in mat_ids:
                mat_ids[name] = _material(scene, libs[name])
            material = mat_ids[name]
        else:
            if default_material is None:
                default_material = scene.principled((0.8, 0.8, 0.8), roughness=0.5)
            material = default_material
        flat = any(c[2] < 0 for face in faces for c in face)
        corners, tris = {}, []
        keys = []
        for fi, face in enumerate(faces):
            ids = []
            for c in face:
                key = c + (fi,) if flat else c  # flat-shaded faces keep their own corners
                if key not in corners:
                    corners[key] = len(keys)
                    keys.append(key)
                ids.append(corners[key])
            tris += [(ids[0], ids[k], ids[k + 1]) for k in range(1, len(ids) - 1)]  # fan
        keys = np.asarray([k[:3] for k in keys], dtype=np.int64)
        tris = np.asarray(tris, dtype=np.int64)
        pos = v[keys[:, 0]]
        uv = np.where(keys[:, 1:2] >= 0, vt[np.maximum(keys[:, 1], 0)], 0.0) if len(vt) else np.zeros((len(keys), 2))
        uv = np.stack([uv[:, 0], 1.0 - uv[:, 1]], axis=1)  # aiProcess_FlipUVs
        if flat:
            nrm = np.zeros_like(pos)
            fn = np.cross(pos[tris[:, 1]] - pos[tris[:, 0]], pos[tris[:, 2]] - pos[tris[:, 0]])
            for k in range(3):
                np.add.at(nrm, tris[:, k], fn)
        else:
            nrm = vn[keys[:, 2]]
        nrm = nrm / np.maximum(np.linalg.norm(nrm, axis=1, keepdims=True), 1e-30)
        tan, bit = _tangent_space(pos, nrm, uv, tris)
        scene.meshes.append(Mesh(_f32(pos), _f32(nrm), _f32(tan), _f32(bit), _f32(uv), np.ascontiguousarray(tris.reshape(-1), dtype=np.uint32),
                                 material))
        out.append(len(scene.meshes) - 1)
    return out


def load_envmap(path, scene):
    """Load a Radiance map and make it the scene's environment; returns its texture index."""
    scene.envmap = scene.add_texture(load_hdr(path))
    return scene.envmap
