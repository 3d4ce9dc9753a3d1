"""Packs a scenes.Scene into the plain arrays of the C++ facade driver
(tests/facade/facade_driver.cpp)."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "facade", "libfacade_driver.so")


def load():
    # torch's HIP runtime and the one the driver links share a soname: the first loaded serves
    # both, and torch needs its own, so torch initialises first
    import torch
    torch.cuda.is_available()
    L = C.CDLL(LIB)
    P = C.c_void_p
    common = [C.c_int, P, P, P, P, P, P, P, P, P, P, C.c_int, P, P, C.c_int, P, P, C.c_int, C.c_int]
    L.facade_flatten.argtypes = common + [P, P, P, P, C.c_char_p, C.c_int]
    L.facade_engine_loop.argtypes = common + [P, P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, P, P, C.c_char_p, C.c_int]
    return L


class Packed:
    """Keeps the arrays alive and exposes the argument list shared by both entry points."""

    def __init__(self, sc, omit_bitangents=False):
        self.sc = sc
        f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
        self.meshes = [(f32(m.vertices), f32(m.normals), f32(m.tangents), f32(m.bitangents), f32(m.uv), np.ascontiguousarray(m.indices, dtype=np.uint32))
                       for m in sc.meshes]
        ptrs = lambda k: (C.c_void_p * len(self.meshes))(*[a[k].ctypes.data for a in self.meshes])
        self.vp, self.np_, self.tp, self.bp, self.uvp, self.ip = (ptrs(k) for k in range(6))
        self.vcount = np.array([len(a[0]) // 3 for a in self.meshes], dtype=np.uint64)
        self.icount = np.array([len(a[5]) for a in self.meshes], dtype=np.uint64)
        self.xf = np.concatenate([f32(m.transform) for m in sc.meshes])
        self.mesh_mat = np.array([m.material for m in sc.meshes], dtype=np.int32)
        self.mat_tex = np.array([[m.albedo, m.metallic, m.normal, m.emissive, m.albedo_f32, m.roughness_f32] for m in sc.materials], dtype=np.int32)
        self.mat_f = np.array([list(m.eta) + list(m.k) + [m.anisotropy, m.emission_scale] for m in sc.materials], dtype=np.float32)
        self.tex_info = np.array([[t.width, t.height, t.channels, int(t.invert_u), int(t.invert_v), 1 if t.is_f32 else 0] for t in sc.textures],
                                 dtype=np.int32)
        self.tex_data = [np.ascontiguousarray(t.data) for t in sc.textures]
        self.tex_ptrs = (C.c_void_p * len(self.tex_data))(*[a.ctypes.data for a in self.tex_data])
        self.omit = int(omit_bitangents)

    def args(self):
        sc = self.sc
        return [len(self.meshes), self.vp, self.np_, self.tp, self.bp, self.uvp, self.ip, self.vcount.ctypes.data, self.icount.ctypes.data,
                self.xf.ctypes.data, self.mesh_mat.ctypes.data, len(sc.materials), self.mat_tex.ctypes.data, self.mat_f.ctypes.data,
                len(sc.textures), self.tex_info.ctypes.data, self.tex_ptrs, sc.envmap, self.omit]
