"""Deterministic synthetic workloads for the Nova hot path.

Each generator returns a :class:`Scene`: meshes with full vertex attributes
(positions, normals, tangents, bitangents, UVs), Principled materials built from
u32 texture packs (albedo, metallic = (_, roughness, metal), flat normal map,
emission), an FP32 RGBA environment map and a camera given as inverse projection
/ inverse view matrices (column-major, as the reference's camera holder stores
them, engine.cpp:380-429).

All arithmetic that defines geometry runs in float64 and is rounded once to
float32, so the same seed gives the same scene on every machine. Transforms are
identity (geometry generated in world space) except in `transformed_scene`.
"""
from dataclasses import dataclass, field
import math
import numpy as np

FLAT_NORMAL = 0xFF8080FF  # (a, r, g, b) = (255, 128, 128, 255)


def pack_rgba8(r, g, b, a=255):
    """Pack channel bytes the way the reference's u32 textures are read:
    (a << 24) | (r << 16) | (g << 8) | b (NovaTextureInterface.h:78-88, is_rgba = false)."""
    r, g, b, a = (np.asarray(x, dtype=np.uint32) for x in (r, g, b, a))
    return (a << 24) | (r << 16) | (g << 8) | b


@dataclass
class Texture:
    width: int
    height: int
    channels: int
    data: np.ndarray  # uint32 (width*height) or float32 (width*height*channels)
    invert_u: bool = False
    invert_v: bool = False

    @property
    def is_f32(self):
        return self.data.dtype == np.float32


@dataclass
class Material:
    albedo: int = -1
    metallic: int = -1
    normal: int = -1
    emissive: int = -1
    albedo_f32: int = -1
    roughness_f32: int = -1
    eta: tuple = (0.0, 0.0, 0.0)
    k: tuple = (0.0, 0.0, 0.0)
    anisotropy: float = 0.0
    emission_scale: float = 1.0


@dataclass
class Mesh:
    vertices: np.ndarray
    normals: np.ndarray
    tangents: np.ndarray
    bitangents: np.ndarray
    uv: np.ndarray
    indices: np.ndarray
    material: int
    transform: np.ndarray = field(default_factory=lambda: np.eye(4, dtype=np.float32).T.reshape(-1).copy())

    @property
    def triangle_count(self):
        return len(self.indices) // 3


@dataclass
class Scene:
    meshes: list = field(default_factory=list)
    materials: list = field(default_factory=list)
    textures: list = field(default_factory=list)
    envmap: int = -1
    inv_P: np.ndarray = None
    inv_V: np.ndarray = None
    name: str = ""

    @property
    def triangle_count(self):
        return sum(m.triangle_count for m in self.meshes)

    def add_texture(self, tex):
        self.textures.append(tex)
        return len(self.textures) - 1

    def add_material(self, mat):
        self.materials.append(mat)
        return len(self.materials) - 1

    def constant_u32(self, value):
        return self.add_texture(Texture(2, 2, 4, np.full(4, value, dtype=np.uint32)))

    def principled(self, albedo_rgb, roughness=0.0, metal=0.0, emission=None, emission_scale=1.0, eta=None):
        a = np.clip(np.round(np.asarray(albedo_rgb) * 255.0), 0, 255).astype(np.uint32)
        alb = self.constant_u32(int(pack_rgba8(a[0], a[1], a[2])))
        met = self.constant_u32(int(pack_rgba8(0, int(round(roughness * 255.0)), int(round(metal * 255.0)))))
        nrm = self.constant_u32(FLAT_NORMAL)
        mat = Material(albedo=alb, metallic=met, normal=nrm)
        if emission is not None:
            e = np.clip(np.round(np.asarray(emission) * 255.0), 0, 255).astype(np.uint32)
            mat.emissive = self.constant_u32(int(pack_rgba8(e[0], e[1], e[2])))
            mat.emission_scale = float(emission_scale)
        if eta is not None:
            mat.eta = (float(eta),) * 3
        return self.add_material(mat)


# ----------------------------------------------------------------- cameras

def perspective(fovy_deg, aspect, near, far):
    f = 1.0 / math.tan(math.radians(fovy_deg) / 2.0)
    P = np.zeros((4, 4))
    P[0, 0] = f / aspect
    P[1, 1] = f
    P[2, 2] = -(far + near) / (far - near)
    P[2, 3] = -(2.0 * far * near) / (far - near)
    P[3, 2] = -1.0
    return P


def look_at(eye, center, up):
    eye, center, up = (np.asarray(v, dtype=np.float64) for v in (eye, center, up))
    f = center - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    V = np.eye(4)
    V[0, :3], V[1, :3], V[2, :3] = s, u, -f
    V[0, 3], V[1, 3], V[2, 3] = -s @ eye, -u @ eye, f @ eye
    return V


def colmajor(M):
    """Row-major numpy 4x4 -> 16 floats in glm value_ptr (column-major) order."""
    return np.ascontiguousarray(np.asarray(M, dtype=np.float64).T.reshape(-1).astype(np.float32))


def set_camera(scene, eye, center, fovy, aspect, near=0.1, far=1000.0):
    P = perspective(fovy, aspect, near, far)
    V = look_at(eye, center, (0.0, 1.0, 0.0))
    scene.inv_P = colmajor(np.linalg.inv(P))
    scene.inv_V = colmajor(np.linalg.inv(V))


# --------------------------------------------------------------- geometry

def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).astype(np.float32))


def _normalize(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def quad_mesh(corners, material):
    """Quad p0,p1,p2,p3 (counter-clockwise seen from the normal side), UV (0,0),(1,0),(1,1),(0,1)."""
    p = np.asarray(corners, dtype=np.float64)
    t = _normalize(p[1] - p[0])
    b = _normalize(p[3] - p[0])
    n = _normalize(np.cross(t, b))
    uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], dtype=np.float64)
    return Mesh(_f32(p.reshape(-1)), _f32(np.tile(n, 4)), _f32(np.tile(t, 4)), _f32(np.tile(b, 4)), _f32(uv.reshape(-1)),
                np.array([0, 1, 2, 0, 2, 3], dtype=np.uint32), material)


def box_mesh(center, half, angle, material):
    """Axis-aligned box rotated about +y by `angle`; 6 faces x 2 triangles, per-face vertices."""
    c, s = math.cos(angle), math.sin(angle)
    R = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    faces = []
    for axis in range(3):
        for sign in (-1.0, 1.0):
            n = np.zeros(3)
            n[axis] = sign
            u = np.zeros(3)
            u[(axis + 1) % 3] = 1.0
            v = np.cross(n, u)
            quad = [n * half[axis] + (-u * half[(axis + 1) % 3]) + (-v * half[(axis + 2) % 3]),
                    n * half[axis] + (u * half[(axis + 1) % 3]) + (-v * half[(axis + 2) % 3]),
                    n * half[axis] + (u * half[(axis + 1) % 3]) + (v * half[(axis + 2) % 3]),
                    n * half[axis] + (-u * half[(axis + 1) % 3]) + (v * half[(axis + 2) % 3])]
            faces.append(([R @ q + np.asarray(center) for q in quad], R @ n, R @ u, R @ v))
    verts, nrm, tan, bit, uv, idx = [], [], [], [], [], []
    for fi, (quad, n, u, v) in enumerate(faces):
        verts += quad
        nrm += [n] * 4
        tan += [u] * 4
        bit += [v] * 4
        uv += [[0, 0], [1, 0], [1, 1], [0, 1]]
        base = 4 * fi
        idx += [base, base + 1, base + 2, base, base + 2, base + 3]
    return Mesh(_f32(np.reshape(verts, -1)), _f32(np.reshape(nrm, -1)), _f32(np.reshape(tan, -1)), _f32(np.reshape(bit, -1)),
                _f32(np.reshape(uv, -1)), np.asarray(idx, dtype=np.uint32), material)


def _tangent_frame(n):
    up = np.where(np.abs(n[:, 1:2]) < 0.999, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    t = _normalize(np.cross(up, n))
    b = np.cross(n, t)
    return t, b


def octasphere_mesh(center, radius, level, material):
    """Octahedron subdivided `level` times (8 * 4**level triangles), projected to a sphere.
    Vertices are per triangle (no sharing) so every triangle carries its own UV seam fix."""
    v = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], dtype=np.float64)
    faces = np.array([[4, 0, 2], [0, 5, 2], [5, 1, 2], [1, 4, 2], [0, 4, 3], [5, 0, 3], [1, 5, 3], [4, 1, 3]])
    tris = v[faces]  # (T, 3, 3)
    for _ in range(level):
        a, b, c = tris[:, 0], tris[:, 1], tris[:, 2]
        ab, bc, ca = _normalize(a + b), _normalize(b + c), _normalize(c + a)
        tris = np.stack([np.stack([a, ab, ca], 1), np.stack([ab, b, bc], 1), np.stack([ca, bc, c], 1), np.stack([ab, bc, ca], 1)], 1)
        tris = tris.reshape(-1, 3, 3)
    tris = _normalize(tris)
    n = tris.reshape(-1, 3)
    u = (np.arctan2(n[:, 2], n[:, 0]) / (2 * math.pi)) % 1.0
    vv = np.arccos(np.clip(n[:, 1], -1, 1)) / math.pi
    u = u.reshape(-1, 3)
    span = u.max(1) - u.min(1)
    u = np.where((span > 0.5)[:, None] & (u < 0.5), u + 1.0, u).reshape(-1)
    t, b = _tangent_frame(n)
    pos = np.asarray(center)[None, :] + radius * n
    idx = np.arange(len(pos), dtype=np.uint32)
    return Mesh(_f32(pos.reshape(-1)), _f32(n.reshape(-1)), _f32(t.reshape(-1)), _f32(b.reshape(-1)),
                _f32(np.stack([u, vv], 1).reshape(-1)), idx, material)


def _fbm(rng, n, octaves=6):
    """Value-noise fBm on an (n+1)^2 grid in [0,1)^2, float64."""
    h = np.zeros((n + 1, n + 1))
    amp, freq = 1.0, 4
    xs = np.linspace(0.0, 1.0, n + 1)
    for _ in range(octaves):
        g = rng.standard_normal((freq + 1, freq + 1))
        gx = xs * freq
        i0 = np.minimum(gx.astype(int), freq - 1)
        f = gx - i0
        f = f * f * (3 - 2 * f)
        a = g[i0][:, i0] * (1 - f)[None, :] + g[i0][:, i0 + 1] * f[None, :]
        b = g[i0 + 1][:, i0] * (1 - f)[None, :] + g[i0 + 1][:, i0 + 1] * f[None, :]
        h += amp * (a * (1 - f)[:, None] + b * f[:, None])
        amp *= 0.5
        freq *= 2
    return h


def terrain_mesh(rng, cells, extent, base_y, amplitude, material):
    """cells x cells grid over [-extent, extent]^2 in x/z, fBm heights; 2*cells^2 triangles."""
    h = base_y + amplitude * _fbm(rng, cells)
    xs = np.linspace(-extent, extent, cells + 1)
    X, Z = np.meshgrid(xs, xs)  # [row=z, col=x]
    P = np.stack([X, h, Z], -1)
    dx = np.gradient(h, xs, axis=1)
    dz = np.gradient(h, xs, axis=0)
    n = _normalize(np.stack([-dx, np.ones_like(h), -dz], -1)).reshape(-1, 3)
    t = _normalize(np.stack([np.ones_like(h), dx, np.zeros_like(h)], -1)).reshape(-1, 3)
    t = _normalize(t - (t * n).sum(1, keepdims=True) * n)
    b = np.cross(n, t)
    uvs = np.stack(np.meshgrid(np.linspace(0, 1, cells + 1), np.linspace(0, 1, cells + 1)), -1).reshape(-1, 2)
    r = np.arange(cells)
    i00 = (r[:, None] * (cells + 1) + r[None, :]).reshape(-1)
    i10, i01, i11 = i00 + 1, i00 + (cells + 1), i00 + (cells + 2)
    idx = np.stack([i00, i01, i11, i00, i11, i10], 1).reshape(-1).astype(np.uint32)
    return Mesh(_f32(P.reshape(-1)), _f32(n.reshape(-1)), _f32(t.reshape(-1)), _f32(b.reshape(-1)), _f32(uvs.reshape(-1)), idx, material)


def sky_envmap(width, height, sun_dir, sun_intensity=50.0, sun_size=0.02):
    """FP32 RGBA equirect sky: zenith/horizon gradient plus an analytic sun disc.
    Texel (i, j) covers the direction with phi = 2*pi*u, theta = pi*v, z-up,
    matching the envmap lookup (math_spherical.h:16-52)."""
    u = (np.arange(width) + 0.5) / width
    v = (np.arange(height) + 0.5) / height
    phi, theta = np.meshgrid(2 * math.pi * u, math.pi * v)
    d = np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)], -1)
    up = d[..., 1]  # scenes use +y as up
    t = np.clip(up * 0.5 + 0.5, 0, 1)[..., None]
    horizon = np.array([0.9, 0.85, 0.8])
    zenith = np.array([0.25, 0.45, 0.9])
    ground = np.array([0.2, 0.18, 0.16])
    col = np.where(up[..., None] >= 0, horizon * (1 - t) + zenith * t, ground)
    s = np.asarray(sun_dir, dtype=np.float64)
    s /= np.linalg.norm(s)
    cosang = np.clip((d * s).sum(-1), -1, 1)
    col = col + sun_intensity * np.exp(-(1 - cosang) / sun_size)[..., None] * np.array([1.0, 0.95, 0.85])
    rgba = np.concatenate([col, np.ones_like(col[..., :1])], -1)
    return Texture(width, height, 4, _f32(rgba.reshape(-1)))


# -------------------------------------------------------------- workloads

def box_room(seed=0xA70AE001, width=256, height=256):
    """Box room: 5 walls (open front), an emissive ceiling quad (radiance 15 via the
    emission-scale extension), two boxes rotated by seeded angles, one of them a
    rough dielectric (IOR 1.5, roughness 0.1). 36 triangles (5*2 + 2 + 2*12)."""
    rng = np.random.default_rng(seed)
    sc = Scene(name="box_room")
    alb = rng.uniform(0.2, 0.8, size=(8, 3))
    # "Lambertian" surfaces are Principled materials with metal 0 and a broad (roughness 0.8)
    # specular lobe; roughness 0 would make that always-present lobe a near-delta.
    wall = [sc.principled(alb[i], roughness=0.8) for i in range(5)]
    light = sc.principled((1, 1, 1), roughness=0.8, emission=(1.0, 1.0, 1.0), emission_scale=15.0)
    box_a = sc.principled(alb[5], roughness=0.8)
    glass = sc.principled(alb[6] * 0 + 1.0, roughness=0.1, eta=1.5)
    L = 1.0
    sc.meshes += [
        quad_mesh([(-L, -L, L), (L, -L, L), (L, -L, -L), (-L, -L, -L)], wall[0]),   # floor (+y)
        quad_mesh([(-L, L, -L), (L, L, -L), (L, L, L), (-L, L, L)], wall[1]),       # ceiling (-y)
        quad_mesh([(-L, -L, -L), (L, -L, -L), (L, L, -L), (-L, L, -L)], wall[2]),   # back (+z)
        quad_mesh([(-L, -L, L), (-L, -L, -L), (-L, L, -L), (-L, L, L)], wall[3]),   # left (+x)
        quad_mesh([(L, -L, -L), (L, -L, L), (L, L, L), (L, L, -L)], wall[4]),       # right (-x)
        quad_mesh([(-0.3, 0.99, -0.3), (0.3, 0.99, -0.3), (0.3, 0.99, 0.3), (-0.3, 0.99, 0.3)], light),
        box_mesh((-0.4, -0.55, -0.3), (0.3, 0.45, 0.3), rng.uniform(0, 2 * math.pi), box_a),
        box_mesh((0.45, -0.7, 0.25), (0.25, 0.3, 0.25), rng.uniform(0, 2 * math.pi), glass),
    ]
    sc.envmap = sc.add_texture(Texture(2, 2, 4, np.zeros(16, dtype=np.float32)))
    set_camera(sc, (0.0, 0.0, 3.6), (0.0, 0.0, 0.0), 40.0, width / height)
    return sc


def million_triangles(seed=0xA70AE002, width=1920, height=1080, terrain_cells=512, spheres=64, sphere_level=5,
                      sky=(2048, 1024)):
    """1,048,576 triangles: a 512x512-cell fBm terrain (524,288 triangles, Disney diffuse)
    plus 64 octahedral spheres at subdivision 5 (8,192 triangles each, GGX conductors with
    roughness U[0.05, 0.8]), centred in a 40^3 box around the origin; procedural sky."""
    rng = np.random.default_rng(seed)
    sc = Scene(name="million_triangles")
    terrain_mat = sc.principled((0.55, 0.5, 0.42), roughness=0.7)
    sc.meshes.append(terrain_mesh(rng, terrain_cells, 20.0, -6.0, 3.0, terrain_mat))
    for _ in range(spheres):
        c = rng.uniform(-20, 20, size=3)
        c[1] = rng.uniform(-6.0, 14.0)
        r = rng.uniform(0.1, 2.0)
        m = sc.principled(rng.uniform(0.3, 0.95, size=3), roughness=rng.uniform(0.05, 0.8), metal=1.0)
        sc.meshes.append(octasphere_mesh(c, r, sphere_level, m))
    sc.envmap = sc.add_texture(sky_envmap(sky[0], sky[1], (0.4, 0.8, 0.3)))
    set_camera(sc, (0.0, 10.0, 34.0), (0.0, -2.0, 0.0), 50.0, width / height)
    return sc


def _affine(rot_axis, angle, scale, translate):
    """Column-major 4x4 (glm value_ptr order) of translate * rotate(axis, angle) * scale."""
    a = np.asarray(rot_axis, dtype=np.float64)
    a /= np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    R = np.eye(3) + math.sin(angle) * K + (1 - math.cos(angle)) * (K @ K)
    M = np.eye(4)
    M[:3, :3] = R @ np.diag(np.asarray(scale, dtype=np.float64))
    M[:3, 3] = translate
    return colmajor(M)


def transformed_scene(seed=0xA70AE006, width=256, height=256, terrain_cells=64, spheres=12, sphere_level=3):
    """Meshes placed by non-identity model matrices (rotation, non-uniform scale,
    translation), as Object3D transforms place them in the reference: a terrain built around
    the origin and turned about a tilted axis, unit octaspheres and boxes scaled per axis,
    rotated about random axes and translated. Diffuse and conductor materials under a sky.
    The reference intersects such meshes in object space (Triangle.h:146-175)."""
    rng = np.random.default_rng(seed)
    sc = Scene(name="transformed")
    terrain = terrain_mesh(rng, terrain_cells, 1.0, 0.0, 0.15, sc.principled((0.55, 0.5, 0.42), roughness=0.7))
    terrain.transform = _affine((0.1, 1.0, 0.05), rng.uniform(0, 2 * math.pi), (9.0, 6.0, 9.0), (0.3, -3.0, -0.5))
    sc.meshes.append(terrain)
    for i in range(spheres):
        m = sc.principled(rng.uniform(0.3, 0.95, 3), roughness=rng.uniform(0.05, 0.8), metal=float(i % 2))
        mesh = octasphere_mesh((0.0, 0.0, 0.0), 1.0, sphere_level, m) if i % 3 else \
            box_mesh((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 0.0, m)
        mesh.transform = _affine(rng.standard_normal(3), rng.uniform(0, 2 * math.pi), rng.uniform(0.4, 1.6, 3),
                                 (rng.uniform(-6, 6), rng.uniform(-1.0, 3.0), rng.uniform(-6, 4)))
        sc.meshes.append(mesh)
    sc.envmap = sc.add_texture(sky_envmap(256, 128, (0.4, 0.8, 0.3), sun_intensity=20.0, sun_size=0.01))
    set_camera(sc, (0.0, 5.0, 14.0), (0.0, -0.5, 0.0), 55.0, width / height)
    return sc


def scaled_million(scale_cells=64, spheres=8, sphere_level=3, width=128, height=72, seed=0xA70AE002):
    """The million-triangle workload's generator at reduced size, for parity tests."""
    return million_triangles(seed=seed, width=width, height=height, terrain_cells=scale_cells, spheres=spheres,
                             sphere_level=sphere_level, sky=(256, 128))


def bake_envmaps(seed=0xA70AE005, count=32):
    """Equirect FP32 RGBA environment maps of uneven sizes with a seeded sun."""
    rng = np.random.default_rng(seed)
    sizes = [(512, 256), (1024, 512), (2048, 1024), (4096, 2048)]
    maps = []
    for _ in range(count):
        w, h = sizes[rng.integers(0, len(sizes))]
        d = rng.standard_normal(3)
        inten = float(np.exp(rng.uniform(np.log(1e2), np.log(1e5))))
        maps.append(sky_envmap(w, h, d, sun_intensity=inten, sun_size=0.002))
    return maps


def differentiable_scene(seed=0xA70AE004, width=1280, height=720, tex=512, terrain_cells=200, spheres=10, sphere_level=4, sky=(1024, 512)):
    """About 100k triangles (a 200x200-cell terrain, 80,000 triangles, plus 10 octahedral
    spheres at subdivision 4, 2,048 each) whose materials all read one UV-mapped FP32
    tex x tex RGB albedo texture (seeded noise in [0.1, 0.9]) and one FP32 roughness
    texture (in [0.05, 0.9]). The terrain is Disney diffuse + specular (metal 0); every
    other sphere is a conductor (metal 1), so albedo also reaches the specular Fresnel.
    Returns (scene, albedo texture id, roughness texture id)."""
    rng = np.random.default_rng(seed)
    sc = Scene(name="differentiable")
    alb = sc.add_texture(Texture(tex, tex, 3, _f32(_smooth_noise(rng, tex, 3, 0.1, 0.9))))
    rough = sc.add_texture(Texture(tex, tex, 1, _f32(_smooth_noise(rng, tex, 1, 0.05, 0.9))))

    def textured(metal):
        m = sc.materials[sc.principled((0.5, 0.5, 0.5), roughness=0.5, metal=metal)]
        m.albedo_f32, m.roughness_f32 = alb, rough
        return len(sc.materials) - 1

    terrain_mat = textured(0.0)
    sc.meshes.append(terrain_mesh(rng, terrain_cells, 12.0, -3.0, 2.0, terrain_mat))
    mats = [textured(0.0), textured(1.0)]
    for i in range(spheres):
        c = rng.uniform(-8, 8, size=3)
        c[1] = rng.uniform(-1.0, 4.0)
        sc.meshes.append(octasphere_mesh(c, rng.uniform(0.6, 1.8), sphere_level, mats[i % 2]))
    sc.envmap = sc.add_texture(sky_envmap(sky[0], sky[1], (0.3, 0.8, 0.5)))
    set_camera(sc, (0.0, 6.0, 18.0), (0.0, -1.0, 0.0), 50.0, width / height)
    return sc, alb, rough


def _smooth_noise(rng, n, channels, lo, hi):
    """n x n x channels value noise (a few octaves, float64) mapped to [lo, hi]."""
    out = np.stack([_fbm(rng, n - 1, octaves=5) for _ in range(channels)], -1)
    out -= out.min()
    out /= max(out.max(), 1e-12)
    return (lo + (hi - lo) * out).reshape(-1)


def sky_envmap_device(width, height, sun_dir, sun_intensity=50.0, sun_size=0.02):
    """sky_envmap evaluated with torch on the GPU (float64, cast once to float32), for
    workloads whose maps are too large to build quickly on the host. Returns a flat cuda
    tensor of width*height*4 floats."""
    import torch
    dev = "cuda"
    u = (torch.arange(width, device=dev, dtype=torch.float64) + 0.5) / width
    v = (torch.arange(height, device=dev, dtype=torch.float64) + 0.5) / height
    theta, phi = torch.meshgrid(math.pi * v, 2 * math.pi * u, indexing="ij")
    d = torch.stack([torch.sin(theta) * torch.cos(phi), torch.sin(theta) * torch.sin(phi), torch.cos(theta)], -1)
    up = d[..., 1]
    t = torch.clamp(up * 0.5 + 0.5, 0, 1)[..., None]
    horizon = torch.tensor([0.9, 0.85, 0.8], device=dev, dtype=torch.float64)
    zenith = torch.tensor([0.25, 0.45, 0.9], device=dev, dtype=torch.float64)
    ground = torch.tensor([0.2, 0.18, 0.16], device=dev, dtype=torch.float64)
    col = torch.where(up[..., None] >= 0, horizon * (1 - t) + zenith * t, ground)
    s = torch.tensor(np.asarray(sun_dir, dtype=np.float64) / np.linalg.norm(sun_dir), device=dev)
    cosang = torch.clamp((d * s).sum(-1), -1, 1)
    col = col + sun_intensity * torch.exp(-(1 - cosang) / sun_size)[..., None] * torch.tensor([1.0, 0.95, 0.85], device=dev, dtype=torch.float64)
    rgba = torch.cat([col, torch.ones_like(col[..., :1])], -1)
    return rgba.to(torch.float32).reshape(-1).contiguous()


def bake_envmap_specs(seed=0xA70AE005, count=32):
    """(width, height, sun_dir, sun_intensity, sun_size) of the IBL bake batch's maps: the same
    draws as bake_envmaps, without building the maps."""
    rng = np.random.default_rng(seed)
    sizes = [(512, 256), (1024, 512), (2048, 1024), (4096, 2048)]
    specs = []
    for _ in range(count):
        w, h = sizes[rng.integers(0, len(sizes))]
        d = rng.standard_normal(3)
        inten = float(np.exp(rng.uniform(np.log(1e2), np.log(1e5))))
        specs.append((w, h, d, inten, 0.002))
    return specs


def batch_scene_specs(seed=0xA70AE003, count=64):
    """Draws of the batched workload: per scene (seed, triangle target log-uniform in
    [1e3, 5e5], square resolution in {128, 256, 512, 1024}, spp in {4, 16, 64})."""
    rng = np.random.default_rng(seed)
    specs = []
    for k in range(count):
        tris = int(np.exp(rng.uniform(np.log(1e3), np.log(5e5))))
        res = int(rng.choice([128, 256, 512, 1024]))
        spp = int(rng.choice([4, 16, 64]))
        specs.append((int(rng.integers(0, 2**31)), tris, res, spp))
    return specs


def batch_scene(seed, tris, res):
    """A terrain + sphere mix of about `tris` triangles (about 60% terrain) under a small sky."""
    rng = np.random.default_rng(seed)
    sc = Scene(name=f"batch{seed}")
    cells = max(8, int(math.sqrt(0.6 * tris / 2)))
    sc.meshes.append(terrain_mesh(rng, cells, 15.0, -4.0, 2.5, sc.principled(rng.uniform(0.3, 0.8, 3), roughness=rng.uniform(0.3, 0.9))))
    rest = tris - 2 * cells * cells
    level = 3 if rest < 20000 else 4 if rest < 200000 else 5
    n = max(1, rest // (8 * 4 ** level))
    for _ in range(n):
        c = rng.uniform(-12, 12, 3)
        c[1] = rng.uniform(-2.0, 8.0)
        m = sc.principled(rng.uniform(0.3, 0.95, 3), roughness=rng.uniform(0.05, 0.8), metal=float(rng.integers(0, 2)))
        sc.meshes.append(octasphere_mesh(c, rng.uniform(0.3, 1.8), level, m))
    sc.envmap = sc.add_texture(sky_envmap(128, 64, rng.standard_normal(3) * 0.3 + np.array([0.3, 0.8, 0.3])))
    set_camera(sc, (0.0, 7.0, 26.0), (0.0, -1.0, 0.0), 50.0, 1.0)
    return sc
