"""Scenes and ray sets shared by the BVH and closest-hit tests (CPU and GPU)."""
import numpy as np

from axomae_amd import scenes


def soup_scene(n, seed, lo=-1.0, hi=1.0, size=0.05, duplicate=0):
    """n random triangles (one mesh) with full attributes; `duplicate` copies of the first triangle."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(lo, hi, size=(n, 1, 3))
    v = c + rng.uniform(-size, size, size=(n, 3, 3))
    if duplicate:
        v[1:1 + duplicate] = v[0]
    v = v.reshape(-1, 3)
    nrm = np.tile([0.0, 0.0, 1.0], (len(v), 1))
    tan = np.tile([1.0, 0.0, 0.0], (len(v), 1))
    bit = np.tile([0.0, 1.0, 0.0], (len(v), 1))
    uv = rng.uniform(0, 1, size=(len(v), 2))
    sc = scenes.Scene(name=f"soup{n}")
    m = sc.principled((0.5, 0.5, 0.5), roughness=0.5)
    f = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.float64).astype(np.float32).reshape(-1))
    sc.meshes.append(scenes.Mesh(f(v), f(nrm), f(tan), f(bit), f(uv), np.arange(len(v), dtype=np.uint32), m))
    sc.envmap = sc.add_texture(scenes.sky_envmap(64, 32, (0.3, 0.9, 0.2)))
    scenes.set_camera(sc, (0.0, 0.0, 3.0), (0.0, 0.0, 0.0), 50.0, 1.0)
    return sc


BVH_SCENES = {
    "box_room": lambda: scenes.box_room(),
    "terrain_spheres": lambda: scenes.scaled_million(scale_cells=48, spheres=6, sphere_level=3),
    "soup_5000": lambda: soup_scene(5000, 1),
    "soup_duplicates": lambda: soup_scene(600, 2, duplicate=200),
    "soup_int_max_fallback": lambda: soup_scene(3000, 3, lo=-3000.0, hi=3000.0, size=40.0),
    "soup_int_max_single_leaf": lambda: soup_scene(2000, 4, lo=100.0, hi=3000.0, size=40.0),
    "single_triangle": lambda: soup_scene(1, 5),
    "transformed": lambda: scenes.transformed_scene(),
}


def camera_rays(sc, w, h, frame=0):
    """Primary rays of the reference camera for pixel centres (no jitter), n x 6."""
    P = np.asarray(sc.inv_P, dtype=np.float64).reshape(4, 4).T
    V = np.asarray(sc.inv_V, dtype=np.float64).reshape(4, 4).T
    ys, xs = np.mgrid[0:h, 0:w]
    nx = (2 * xs - w) / w
    ny = (h - 2 * (h - ys)) / h
    pts = np.stack([nx, ny, -np.ones_like(nx), np.ones_like(nx)], -1) @ P.T
    o = (pts / pts[..., 3:]) @ V.T
    d = np.stack([nx, ny, -np.ones_like(nx), np.zeros_like(nx)], -1) @ P.T
    d = np.concatenate([d[..., :2], -np.ones_like(nx)[..., None], np.zeros_like(nx)[..., None]], -1) @ V.T
    d = d[..., :3] / np.linalg.norm(d[..., :3], axis=-1, keepdims=True)
    return np.concatenate([o[..., :3], d], -1).reshape(-1, 6).astype(np.float32)


def random_rays(n, seed, origin_box=1.5):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-origin_box, origin_box, size=(n, 3))
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], 1).astype(np.float32)


def axis_rays(n, seed, box):
    """Rays with zero direction components (one or two axes), half of them from origins on a
    quarter grid so that they run inside box face planes and through shared edges."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(-box, box, (n, 3))
    o[: n // 2] = np.round(o[: n // 2] * 4) / 4
    d = np.zeros((n, 3))
    ax = rng.integers(0, 3, n)
    d[np.arange(n), ax] = rng.choice([-1.0, 1.0], n)
    m = np.where(rng.random(n) < 0.5)[0]
    d[m, (ax[m] + 1) % 3] = rng.choice([-1.0, 1.0], len(m)) * rng.choice([1.0, 0.5], len(m))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], 1).astype(np.float32)


def near_axis_rays(n, seed, box):
    """axis_rays with each zero direction component replaced by a tiny one, +-10^-k for k from
    3 to 37: large finite reciprocals, up to near the largest float, instead of the infinite
    ones that make an axis a containment test."""
    r = axis_rays(n, seed, box).astype(np.float64)
    rng = np.random.default_rng(seed + 1)
    d = r[:, 3:]
    zero = d == 0.0
    tiny = rng.choice([1e-3, 1e-6, 1e-9, 1e-12, 1e-17, 1e-19, 1e-21, 1e-25, 1e-30, 1e-37], size=d.shape) * rng.choice([-1.0, 1.0], size=d.shape)
    d[zero] = tiny[zero]
    r[:, 3:] = d
    out = r.astype(np.float32)
    assert np.all(out[:, 3:] != 0.0)
    return out


def hit_test_rays(sc, name, n_random):
    """Pixel-centre camera rays, random rays and axis-parallel rays for one test scene."""
    scale = 1.0 if "int_max" not in name else 3000.0
    return np.concatenate([camera_rays(sc, 96, 96), random_rays(n_random, 7, 1.5 * scale), axis_rays(n_random // 4, 5, 1.2 * scale)])


def near_axis_test_rays(name, n):
    """Nearly axis-parallel rays for one test scene, origins scaled like hit_test_rays'."""
    scale = 1.0 if "int_max" not in name else 3000.0
    return near_axis_rays(n, 11, 1.2 * scale)
