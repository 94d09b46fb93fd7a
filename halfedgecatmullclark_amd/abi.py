"""ctypes mirrors of the C ABI structs ``cc_Mesh`` and ``cc_Subd``.

Field order and types follow CatmullClark.h:58-71 (``cc_Mesh``) and :136-142
(``cc_Subd``); both structs are plain C with natural alignment, so one
definition serves every library that exports the ``ccs_*`` / ``ccm_*`` ABI.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .cage import CREASE_DTYPE, Cage

HALFEDGE_DTYPE = np.dtype([("twinID", "<i4"), ("edgeID", "<i4"), ("vertexID", "<i4"), ("uvID", "<i4")])
# cc_Halfedge_SemiRegular of a CC_DISABLE_UV build (no uvID): 12 bytes
HALFEDGE12_DTYPE = np.dtype([("twinID", "<i4"), ("edgeID", "<i4"), ("vertexID", "<i4")])


class cc_Mesh(ctypes.Structure):
    _fields_ = [
        ("vertexCount", ctypes.c_int32),
        ("uvCount", ctypes.c_int32),
        ("halfedgeCount", ctypes.c_int32),
        ("edgeCount", ctypes.c_int32),
        ("faceCount", ctypes.c_int32),
        ("vertexToHalfedgeIDs", ctypes.c_void_p),
        ("edgeToHalfedgeIDs", ctypes.c_void_p),
        ("faceToHalfedgeIDs", ctypes.c_void_p),
        ("vertexPoints", ctypes.c_void_p),
        ("uvs", ctypes.c_void_p),
        ("halfedges", ctypes.c_void_p),
        ("creases", ctypes.c_void_p),
    ]


class cc_Subd(ctypes.Structure):
    _fields_ = [
        ("cage", ctypes.POINTER(cc_Mesh)),
        ("vertexPoints", ctypes.c_void_p),
        ("halfedges", ctypes.c_void_p),
        ("creases", ctypes.c_void_p),
        ("maxDepth", ctypes.c_int32),
    ]


def crease_dtype(real=np.float32) -> np.dtype:
    """Crease record with the sharpness in ``real`` (natural C alignment)."""
    if np.dtype(real).itemsize == 4:
        return CREASE_DTYPE
    return np.dtype({"names": ["nextID", "prevID", "sharpness"], "formats": ["<i4", "<i4", "<f8"],
                     "offsets": [0, 4, 8], "itemsize": 16})


class MeshView:
    """A ``cc_Mesh`` struct pointing at numpy arrays it keeps alive.

    ``real`` selects the floating-point type of points, uvs and sharpness
    (float32 for the ABI; float64 only for the double build of the
    reference used by the tests)."""

    def __init__(self, cage: Cage, real=np.float32):
        c = cage.contiguous()
        real = np.dtype(real)
        self.cage = c
        self.points = np.ascontiguousarray(np.asarray(cage.vertex_points).reshape(-1, 3), dtype=real)
        self.uvs = np.ascontiguousarray(np.asarray(cage.uvs).reshape(-1, 2), dtype=real)
        cr = np.zeros(c.E, dtype=crease_dtype(real))
        cr["nextID"] = c.creases["nextID"]
        cr["prevID"] = c.creases["prevID"]
        cr["sharpness"] = c.creases["sharpness"]
        self.creases = cr
        self.halfedges = c.halfedges
        self.v2h, self.e2h, self.f2h = c.v2h, c.e2h, c.f2h
        self.struct = cc_Mesh(c.V, c.U, c.H, c.E, c.F,
                              _ptr(self.v2h), _ptr(self.e2h), _ptr(self.f2h),
                              _ptr(self.points), _ptr(self.uvs), _ptr(self.halfedges), _ptr(self.creases))

    @property
    def pointer(self):
        return ctypes.pointer(self.struct)


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data) if a.size else ctypes.c_void_p(0)


def level_counts(V: int, H: int, E: int, F: int, depth: int) -> dict:
    """Exact per-level counts (CatmullClark.h:455-585): F_{d+1} = H_d,
    E_{d+1} = 2E_d + H_d, H_{d+1} = 4H_d, C_{d+1} = 2C_d, V_{d+1} = V_d + E_d + F_d."""
    Hs, Es, Fs, Vs, Cs = [H], [E], [F], [V], [E]
    for _ in range(depth):
        Hs.append(4 * Hs[-1])
        Fs.append(Hs[-2])
        Es.append(2 * Es[-1] + Hs[-2])
        Vs.append(Vs[-1] + Es[-2] + Fs[-2])
        Cs.append(2 * Cs[-1])
    cum = lambda xs: [sum(xs[1:k + 1]) for k in range(len(xs))]
    return {"H": Hs, "E": Es, "F": Fs, "V": Vs, "C": Cs,
            "Hcum": cum(Hs), "Vcum": cum(Vs), "Ccum": cum(Cs)}
