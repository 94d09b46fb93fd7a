"""ctypes front end of libcchip.so, the GPU drop-in for CatmullClark.h's
refinement path.  Importing this module loads the library and fails loudly if
it has not been built; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .abi import HALFEDGE12_DTYPE, HALFEDGE_DTYPE, MeshView, cc_Mesh, cc_Subd, level_counts
from .cage import CREASE_DTYPE, Cage

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcchip.so")
# the library variants: 16-byte level records with uvIDs, and the 12-byte
# records of a CC_DISABLE_UV build (libcchip_nouv.so)
LIB_PATHS = {"uv": LIB_PATH, "nouv": os.path.join(HERE, "libcchip_nouv.so")}
RECORD_DTYPES = {"uv": HALFEDGE_DTYPE, "nouv": HALFEDGE12_DTYPE}

HALFEDGES, CREASES, UVS, POINTS, GATHER, NOCREASES, FP64 = 1, 2, 4, 8, 16, 32, 64
ALL = HALFEDGES | CREASES | UVS | POINTS
PTR_HALFEDGES, PTR_CREASES, PTR_POINTS, PTR_CAGE_POINTS = 0, 1, 2, 3
OPT_FUSED, OPT_STORE, OPT_XCD_REMAP, OPT_FAN, OPT_PATCH_TILE, OPT_PATCH_HPT, OPT_SCAN, OPT_ADJ_VERTEX = 0, 1, 2, 3, 4, 5, 6, 7
OPT_FUSED_BLOCK, OPT_FUSED_TILES, OPT_VERTEX_WAVES = 8, 9, 10
OPT_GRAPH = 100
(COUNT_HALFEDGES, COUNT_CREASES, COUNT_POINTS, COUNT_CAGE_POINTS, COUNT_FINEST_POINTS, COUNT_DEPTH,
 COUNT_CAGES) = range(7)

HOST_DRIVERS = ("ccs_Refine_Scatter", "ccs_Refine_Gather", "ccs_Refine_NoCreases_Scatter",
                "ccs_Refine_NoCreases_Gather", "ccs_RefineVertexPoints_Scatter",
                "ccs_RefineVertexPoints_Gather", "ccs_RefineVertexPoints_NoCreases_Scatter",
                "ccs_RefineVertexPoints_NoCreases_Gather", "ccs_RefineHalfedges", "ccs_RefineCreases",
                "ccs_RefineVertexUvs", "ccs_RefineCageHalfedges", "ccs_RefineCageCreases",
                "ccs_RefineCageVertexUvs", "ccs_RefineCageVertexPoints_Scatter",
                "ccs_RefineCageVertexPoints_Gather")


class CCError(RuntimeError):
    pass


_libs: dict = {}


def lib(variant: str = "uv") -> ctypes.CDLL:
    if variant in _libs:
        return _libs[variant]
    path = LIB_PATHS[variant]
    if not os.path.exists(path):
        raise CCError(f"{path} is missing: build it with `python -m halfedgecatmullclark_amd.build`")
    # One HIP runtime per process: when torch is installed, load it first so
    # that the library binds to the libamdhip64.so.7 torch brings instead of
    # a second copy from /opt/rocm (two runtimes in one process leave torch
    # without devices).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(path)
    P = ctypes.POINTER
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    sigs = {
        "ccs_Create": (P(cc_Subd), [P(cc_Mesh), i32]),
        "ccs_Release": (None, [P(cc_Subd)]),
        "ccs_LastError": (i32, []),
        "ccs_LastErrorString": (ctypes.c_char_p, []),
        "ccs_RefineVertexPoints_Adjoint": (None, [P(cc_Subd), vp, vp]),
        "ccs_RefineVertexPoints_Adjoint_f64": (None, [P(cc_Subd), vp, vp]),
        "ccd_Create": (vp, [P(cc_Mesh), i32]),
        "ccd_Release": (None, [vp]),
        "ccd_MaxDepth": (i32, [vp]),
        "ccd_Count": (i64, [vp, i32, i32]),
        "ccd_DevicePointer": (vp, [vp, i32]),
        "ccd_UploadCagePoints": (i32, [vp, vp]),
        "ccd_Refine": (i32, [vp, i32, vp]),
        "ccd_RefinePoints": (i32, [vp, vp, vp, i32, vp]),
        "ccd_RefinePointsAdjoint": (i32, [vp, vp, vp, i32, vp]),
        "ccd_Download": (i32, [vp, P(cc_Subd), i32]),
        "ccd_Synchronize": (i32, [vp]),
        "ccd_SetTiming": (i32, [vp, i32]),
        "ccd_SetOption": (i32, [vp, i32, i32]),
        "ccd_TimingCount": (i32, [vp]),
        "ccd_TimingName": (ctypes.c_char_p, [vp, i32]),
        "ccd_TimingMs": (ctypes.c_float, [vp, i32]),
        "ccd_Autotune": (i32, [vp, i32, vp]),
        "ccd_TileOption": (i32, [vp, i32, i32]),
        "ccb_Create": (vp, [P(P(cc_Mesh)), P(i32), i32]),
        "ccb_Release": (None, [vp]),
        "ccb_Count": (i64, [vp, i32, i32]),
        "ccb_Offset": (i64, [vp, i32, i32]),
        "ccb_DevicePointer": (vp, [vp, i32]),
        "ccb_Refine": (i32, [vp, i32, vp]),
        "ccb_RefinePoints": (i32, [vp, vp, i32, vp]),
        "ccb_FinestPoints": (i32, [vp, vp, vp]),
        "ccb_RefinePointsAdjoint": (i32, [vp, vp, vp, i32, vp]),
        "ccb_Download": (i32, [vp, i32, P(cc_Subd), i32]),
        "ccb_Synchronize": (i32, [vp]),
        "ccb_SetOption": (i32, [vp, i32, i32]),
        "ccb_Autotune": (i32, [vp, i32, vp]),
        "ccb_TileOption": (i32, [vp, i32, i32]),
        "ccb_SetTiming": (i32, [vp, i32]),
        "ccb_TimingCount": (i32, [vp]),
        "ccb_TimingName": (ctypes.c_char_p, [vp, i32]),
        "ccb_TimingLevel": (i32, [vp, i32]),
        "ccb_TimingMs": (ctypes.c_float, [vp, i32]),
        "ccm_Load": (P(cc_Mesh), [ctypes.c_char_p]),
        "ccm_Save": (ctypes.c_bool, [P(cc_Mesh), ctypes.c_char_p]),
        "ccm_Release": (None, [P(cc_Mesh)]),
        "ccd_RefineNoAlloc": (i32, [P(cc_Mesh), P(cc_Subd), i32, vp]),
        "ccd_HalfedgeNormals": (i32, [vp, vp, i32, vp]),
        "ccm_CreateFromFaces": (P(cc_Mesh), [i32, vp, i32, vp, vp, i32, vp, vp, i32, vp, vp, i32]),
        "ccm_LoadObj": (P(cc_Mesh), [ctypes.c_char_p]),
        "ccs_HalfedgeNormals": (None, [P(cc_Subd), vp]),
        "ccs_HalfedgeNormals_Fast": (None, [P(cc_Subd), vp]),
        "ccs_ExportToObj": (ctypes.c_bool, [P(cc_Subd), i32, ctypes.c_char_p]),
    }
    for name in HOST_DRIVERS:
        sigs[name] = (None, [P(cc_Subd)])
    for name in ("ccs_CumulativeHalfedgeCountAtDepth", "ccs_CumulativeVertexCountAtDepth",
                 "ccs_CumulativeEdgeCountAtDepth", "ccs_CumulativeFaceCountAtDepth",
                 "ccs_CumulativeCreaseCountAtDepth", "ccm_HalfedgeCountAtDepth", "ccm_CreaseCountAtDepth",
                 "ccm_FaceCountAtDepth", "ccm_EdgeCountAtDepth", "ccm_VertexCountAtDepth"):
        sigs[name] = (i32, [P(cc_Mesh), i32])
    for name in ("ccs_EdgeToHalfedgeID", "ccs_VertexPointToHalfedgeID", "ccs_VertexToHalfedgeID"):
        sigs[name] = (i32, [P(cc_Subd), i32, i32])
    for name, (res, args) in sigs.items():
        try:
            f = getattr(L, name)
        except AttributeError:
            continue  # calling it raises AttributeError
        f.restype = res
        f.argtypes = args
    _libs[variant] = L
    return L


def last_error(variant: str = "uv") -> tuple[int, str]:
    L = lib(variant)
    return int(L.ccs_LastError()), (L.ccs_LastErrorString() or b"").decode()


def _check(rc: int, what: str, variant: str = "uv"):
    if rc != 0:
        code, msg = last_error(variant)
        raise CCError(f"{what} failed ({rc}): {msg}")


class HostSubd:
    """A ``cc_Subd`` from ``ccs_Create`` (host arrays, GPU-refined)."""

    def __init__(self, cage: Cage, depth: int, variant: str = "uv"):
        self.variant = variant
        self.L = lib(variant)
        self.rec = RECORD_DTYPES[variant]
        self.mesh = MeshView(cage)
        self.depth = depth
        self.n = level_counts(cage.V, cage.H, cage.E, cage.F, depth)
        self.subd = self.L.ccs_Create(self.mesh.pointer, depth)
        if not self.subd:
            raise CCError(f"ccs_Create failed: {last_error(variant)[1]}")

    def run(self, *drivers: str) -> "HostSubd":
        for name in drivers:
            getattr(self.L, name)(self.subd)
            code, msg = last_error(self.variant)
            if code:
                raise CCError(f"{name}: {msg}")
        return self

    def adjoint(self, grad_finest: np.ndarray, real=np.float32) -> np.ndarray:
        """dL/d(cage points) from dL/d(finest points), on the GPU."""
        real = np.dtype(real)
        g = np.ascontiguousarray(np.asarray(grad_finest).reshape(-1, 3), dtype=real)
        out = np.zeros((self.n["V"][0], 3), dtype=real)
        name = "ccs_RefineVertexPoints_Adjoint" if real.itemsize == 4 else "ccs_RefineVertexPoints_Adjoint_f64"
        getattr(self.L, name)(self.subd, g.ctypes.data, out.ctypes.data)
        code, msg = last_error(self.variant)
        if code:
            raise CCError(f"{name}: {msg}")
        return out

    def normals(self, smooth: bool = True) -> np.ndarray:
        """Halfedge normals of level maxDepth from the subd's host arrays
        (ccs_HalfedgeNormals / ccs_HalfedgeNormals_Fast)."""
        out = np.zeros((self.n["H"][self.depth], 3), np.float32)
        name = "ccs_HalfedgeNormals" if smooth else "ccs_HalfedgeNormals_Fast"
        getattr(self.L, name)(self.subd, out.ctypes.data)
        code, msg = last_error(self.variant)
        if code:
            raise CCError(f"{name}: {msg}")
        return out

    def export_obj(self, depth: int, path: str):
        """Level ``depth`` as an OBJ file (ccs_ExportToObj)."""
        if not self.L.ccs_ExportToObj(self.subd, depth, os.fsencode(path)):
            raise CCError(f"ccs_ExportToObj: {last_error(self.variant)[1]}")

    def arrays(self, keys=("halfedges", "creases", "points")) -> dict:
        """Copies of the subd's host arrays named in ``keys``."""
        s = self.subd.contents
        nh, nc, nv = self.n["Hcum"][self.depth], self.n["Ccum"][self.depth], self.n["Vcum"][self.depth]
        out = {"counts": self.n}
        if "halfedges" in keys:
            words = self.rec.itemsize // 4
            he = np.ctypeslib.as_array(ctypes.cast(s.halfedges, ctypes.POINTER(ctypes.c_int32)), (nh * words,))
            out["halfedges"] = he.view(self.rec).copy()
        if "creases" in keys:
            cr = np.ctypeslib.as_array(ctypes.cast(s.creases, ctypes.POINTER(ctypes.c_int32)), (nc * 3,))
            out["creases"] = cr.view(CREASE_DTYPE).copy()
        if "points" in keys:
            pt = np.ctypeslib.as_array(ctypes.cast(s.vertexPoints, ctypes.POINTER(ctypes.c_float)), (nv * 3,))
            out["points"] = pt.reshape(nv, 3).copy()
        return out

    def close(self):
        if getattr(self, "subd", None):
            self.L.ccs_Release(self.subd)
            self.subd = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def refine_host(cage: Cage, depth: int, *drivers: str, variant: str = "uv") -> dict:
    """Run host-semantics drivers (default ``ccs_Refine_Scatter``) and return copies of the arrays."""
    s = HostSubd(cage, depth, variant)
    try:
        return s.run(*(drivers or ("ccs_Refine_Scatter",))).arrays()
    finally:
        s.close()


def refine_file(path: str, depth: int, *drivers: str, variant: str = "uv") -> dict:
    """Load a ``.ccm`` cage with the library's ``ccm_Load``, refine it with
    host-semantics drivers (default ``ccs_Refine_Scatter``) on a subd from
    ``ccs_Create``, and return copies of the subd's arrays."""
    L = lib(variant)
    rec = RECORD_DTYPES[variant]
    mesh = L.ccm_Load(os.fsencode(path))
    if not mesh:
        raise CCError(f"ccm_Load({path}) failed: {last_error(variant)[1]}")
    try:
        m = mesh.contents
        n = level_counts(m.vertexCount, m.halfedgeCount, m.edgeCount, m.faceCount, depth)
        subd = L.ccs_Create(mesh, depth)
        if not subd:
            raise CCError(f"ccs_Create failed: {last_error(variant)[1]}")
        try:
            for name in drivers or ("ccs_Refine_Scatter",):
                getattr(L, name)(subd)
                code, msg = last_error(variant)
                if code:
                    raise CCError(f"{name}: {msg}")
            s = subd.contents
            nh, nc, nv = n["Hcum"][depth], n["Ccum"][depth], n["Vcum"][depth]
            he = np.ctypeslib.as_array(ctypes.cast(s.halfedges, ctypes.POINTER(ctypes.c_int32)),
                                       (nh * (rec.itemsize // 4),))
            cr = np.ctypeslib.as_array(ctypes.cast(s.creases, ctypes.POINTER(ctypes.c_int32)), (nc * 3,))
            pt = np.ctypeslib.as_array(ctypes.cast(s.vertexPoints, ctypes.POINTER(ctypes.c_float)), (nv * 3,))
            return {"halfedges": he.view(rec).copy(), "creases": cr.view(CREASE_DTYPE).copy(),
                    "points": pt.reshape(nv, 3).copy(), "counts": n}
        finally:
            L.ccs_Release(subd)
    finally:
        L.ccm_Release(mesh)


def cage_from_mesh(mesh_ptr) -> Cage:
    """Copy a library-allocated cc_Mesh into a Cage (numpy arrays)."""
    m = mesh_ptr.contents
    V, U, H, E, F = m.vertexCount, m.uvCount, m.halfedgeCount, m.edgeCount, m.faceCount

    def arr(ptr, ctype, n, dtype):
        if n == 0:
            return np.zeros(0, dtype)
        return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctype)), (n,)).copy().view(dtype)

    return Cage(arr(m.vertexPoints, ctypes.c_float, 3 * V, np.float32).reshape(V, 3),
                arr(m.uvs, ctypes.c_float, 2 * U, np.float32).reshape(U, 2),
                arr(m.halfedges, ctypes.c_int32, 7 * H, np.int32).reshape(H, 7),
                arr(m.creases, ctypes.c_int32, 3 * E, np.int32).view(CREASE_DTYPE),
                arr(m.vertexToHalfedgeIDs, ctypes.c_int32, V, np.int32),
                arr(m.edgeToHalfedgeIDs, ctypes.c_int32, E, np.int32),
                arr(m.faceToHalfedgeIDs, ctypes.c_int32, F, np.int32))


def face_arrays(points, faces, uvs=None, face_uvs=None, creases=()) -> dict:
    """The flat arrays ccm_CreateFromFaces takes, from the arguments of
    cage.from_faces."""
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    sizes = np.array([len(f) for f in faces], np.int64)
    off = np.zeros(len(faces) + 1, np.int32)
    np.cumsum(sizes, out=off[1:])
    fv = np.ascontiguousarray(np.concatenate([np.asarray(f, np.int32) for f in faces]))
    uv = np.ascontiguousarray(uvs if uvs is not None else np.zeros((0, 2)), np.float32).reshape(-1, 2)
    fuv = (np.ascontiguousarray(np.concatenate([np.asarray(f, np.int32) for f in face_uvs]))
           if face_uvs is not None else None)
    tags = list(creases)
    tv = np.ascontiguousarray(np.array([[a, b] for a, b, _ in tags], np.int32).reshape(-1))
    ts = np.ascontiguousarray(np.array([s for _, _, s in tags], np.float32))
    return {"pts": pts, "off": off, "fv": fv, "uv": uv, "fuv": fuv, "tv": tv, "ts": ts}


def create_mesh(arrs: dict, sharpen_boundaries=True):
    """ccm_CreateFromFaces on face_arrays' arrays: the library-allocated
    cc_Mesh (free it with release_mesh)."""
    L = lib()
    p = lambda a: a.ctypes.data if a is not None and a.size else None
    m = L.ccm_CreateFromFaces(len(arrs["pts"]), p(arrs["pts"]), len(arrs["off"]) - 1, p(arrs["off"]), p(arrs["fv"]),
                              len(arrs["uv"]), p(arrs["uv"]), p(arrs["fuv"]), len(arrs["ts"]), p(arrs["tv"]),
                              p(arrs["ts"]), int(sharpen_boundaries))
    if not m:
        raise CCError(f"ccm_CreateFromFaces failed: {last_error()[1]}")
    return m


def release_mesh(m):
    lib().ccm_Release(m)


def build_cage(points, faces, uvs=None, face_uvs=None, creases=(), sharpen_boundaries=True) -> Cage:
    """Polygons to a cage with the GPU cage builder (ccm_CreateFromFaces);
    the arguments are those of cage.from_faces."""
    m = create_mesh(face_arrays(points, faces, uvs, face_uvs, creases), sharpen_boundaries)
    try:
        return cage_from_mesh(m)
    finally:
        release_mesh(m)


def load_obj_mesh(path: str):
    """ccm_LoadObj: the library-allocated cc_Mesh (free it with release_mesh)."""
    m = lib().ccm_LoadObj(os.fsencode(path))
    if not m:
        raise CCError(f"ccm_LoadObj({path}) failed: {last_error()[1]}")
    return m


def load_obj(path: str) -> Cage:
    """An OBJ file through ccm_LoadObj (parsed as the reference's converter
    parses it, built on the GPU)."""
    L = lib()
    m = L.ccm_LoadObj(os.fsencode(path))
    if not m:
        raise CCError(f"ccm_LoadObj({path}) failed: {last_error()[1]}")
    try:
        return cage_from_mesh(m)
    finally:
        L.ccm_Release(m)


class DeviceSubd:
    """A ``ccd_Subd``: the cage and levels 1..depth resident in HBM."""

    def __init__(self, cage: Cage, depth: int, variant: str = "uv"):
        self.variant = variant
        self.L = lib(variant)
        self.rec = RECORD_DTYPES[variant]
        self.cage = cage.contiguous()
        self.mesh = MeshView(self.cage)
        self.depth = depth
        self.n = level_counts(cage.V, cage.H, cage.E, cage.F, depth)
        self.handle = self.L.ccd_Create(self.mesh.pointer, depth)
        if not self.handle:
            raise CCError(f"ccd_Create failed: {last_error(variant)[1]}")

    def _ok(self, rc: int, what: str):
        _check(rc, what, self.variant)

    def refine(self, steps: int = ALL, stream: int | None = None):
        self._ok(self.L.ccd_Refine(self.handle, steps, stream), "ccd_Refine")

    def refine_points(self, cage_points_ptr: int | None, out_ptr: int | None, steps: int = 0,
                      stream: int | None = None):
        self._ok(self.L.ccd_RefinePoints(self.handle, cage_points_ptr, out_ptr, steps, stream), "ccd_RefinePoints")

    def refine_points_adjoint(self, grad_finest_ptr: int, grad_cage_ptr: int, steps: int = 0,
                              stream: int | None = None):
        self._ok(self.L.ccd_RefinePointsAdjoint(self.handle, grad_finest_ptr, grad_cage_ptr, steps, stream),
               "ccd_RefinePointsAdjoint")

    def synchronize(self):
        self._ok(self.L.ccd_Synchronize(self.handle), "ccd_Synchronize")

    def halfedge_normals(self, out_ptr: int, smooth: bool = True, stream: int | None = None):
        self._ok(self.L.ccd_HalfedgeNormals(self.handle, out_ptr, int(smooth), stream), "ccd_HalfedgeNormals")

    def pointer(self, which: int) -> int:
        return int(self.L.ccd_DevicePointer(self.handle, which) or 0)

    def upload_cage_points(self, points: np.ndarray):
        p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        self._ok(self.L.ccd_UploadCagePoints(self.handle, p.ctypes.data), "ccd_UploadCagePoints")

    def download(self, steps: int = ALL) -> dict:
        """Copy levels 1..depth to host numpy arrays (through a host cc_Subd)."""
        nh, nc, nv = self.n["Hcum"][self.depth], self.n["Ccum"][self.depth], self.n["Vcum"][self.depth]
        he = np.zeros(nh, self.rec)
        cr = np.zeros(nc, CREASE_DTYPE)
        pt = np.zeros((nv, 3), np.float32)
        sub = cc_Subd(self.mesh.pointer, pt.ctypes.data, he.ctypes.data, cr.ctypes.data, self.depth)
        self._ok(self.L.ccd_Download(self.handle, ctypes.byref(sub), steps), "ccd_Download")
        return {"halfedges": he, "creases": cr, "points": pt, "counts": self.n}

    def set_option(self, key: int, value: int):
        self._ok(self.L.ccd_SetOption(self.handle, key, value), "ccd_SetOption")

    def autotune(self, steps: int = ALL, stream: int | None = None):
        self._ok(self.L.ccd_Autotune(self.handle, steps, stream), "ccd_Autotune")

    def level_options(self) -> list[dict]:
        keys = {"kind": OPT_FUSED, "scan": OPT_SCAN, "tile": OPT_PATCH_TILE, "hpt": OPT_PATCH_HPT,
                "block": OPT_FUSED_BLOCK, "tiles": OPT_FUSED_TILES, "waves": OPT_VERTEX_WAVES}
        return [{k: int(self.L.ccd_TileOption(self.handle, v, d)) for k, v in keys.items()}
                for d in range(1, self.depth)]

    def set_timing(self, on: bool):
        self.L.ccd_SetTiming(self.handle, int(on))

    def timings(self) -> list[tuple[str, float]]:
        n = self.L.ccd_TimingCount(self.handle)
        return [(self.L.ccd_TimingName(self.handle, i).decode(), float(self.L.ccd_TimingMs(self.handle, i)))
                for i in range(n)]

    def close(self):
        if getattr(self, "handle", None):
            self.L.ccd_Release(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """A ``ccb_Batch``: many cages, each with its own depth, refined together
    (one launch per kernel kind and level, replayed as a hipGraph)."""

    def __init__(self, cages: list[Cage], depths: list[int], variant: str = "uv"):
        self.variant = variant
        self.L = lib(variant)
        self.rec = RECORD_DTYPES[variant]
        self.cages = [c.contiguous() for c in cages]
        self.meshes = [MeshView(c) for c in self.cages]
        self.depths = [int(d) for d in depths]
        arr = (ctypes.POINTER(cc_Mesh) * len(self.meshes))(*[m.pointer for m in self.meshes])
        dep = (ctypes.c_int32 * len(self.depths))(*self.depths)
        self.handle = self.L.ccb_Create(arr, dep, len(self.meshes))
        if not self.handle:
            raise CCError(f"ccb_Create failed: {last_error(variant)[1]}")

    def _ok(self, rc: int, what: str):
        _check(rc, what, self.variant)

    def __len__(self) -> int:
        return len(self.cages)

    def count(self, what: int, cage: int = -1) -> int:
        return int(self.L.ccb_Count(self.handle, what, cage))

    def offset(self, what: int, cage: int) -> int:
        return int(self.L.ccb_Offset(self.handle, what, cage))

    def pointer(self, which: int) -> int:
        return int(self.L.ccb_DevicePointer(self.handle, which) or 0)

    def refine(self, steps: int = ALL, stream: int | None = None):
        self._ok(self.L.ccb_Refine(self.handle, steps, stream), "ccb_Refine")

    def refine_points(self, cage_points_ptr: int | None = None, steps: int = 0, stream: int | None = None):
        self._ok(self.L.ccb_RefinePoints(self.handle, cage_points_ptr, steps, stream), "ccb_RefinePoints")

    def finest_points(self, out_ptr: int, stream: int | None = None):
        self._ok(self.L.ccb_FinestPoints(self.handle, out_ptr, stream), "ccb_FinestPoints")

    def refine_points_adjoint(self, grad_finest_ptr: int, grad_cage_ptr: int, steps: int = 0,
                              stream: int | None = None):
        self._ok(self.L.ccb_RefinePointsAdjoint(self.handle, grad_finest_ptr, grad_cage_ptr, steps, stream),
               "ccb_RefinePointsAdjoint")

    def synchronize(self):
        self._ok(self.L.ccb_Synchronize(self.handle), "ccb_Synchronize")

    def set_option(self, key: int, value: int):
        self._ok(self.L.ccb_SetOption(self.handle, key, value), "ccb_SetOption")

    def autotune(self, steps: int = ALL, stream: int | None = None):
        self._ok(self.L.ccb_Autotune(self.handle, steps, stream), "ccb_Autotune")

    def level_options(self) -> list[dict]:
        keys = {"kind": OPT_FUSED, "scan": OPT_SCAN, "tile": OPT_PATCH_TILE, "hpt": OPT_PATCH_HPT,
                "block": OPT_FUSED_BLOCK, "tiles": OPT_FUSED_TILES, "waves": OPT_VERTEX_WAVES}
        dmax = self.count(COUNT_DEPTH)
        return [{k: int(self.L.ccb_TileOption(self.handle, v, d)) for k, v in keys.items()} for d in range(1, dmax)]

    def set_timing(self, on: bool):
        self.L.ccb_SetTiming(self.handle, int(on))

    def timings(self) -> list[tuple[str, int, float]]:
        n = self.L.ccb_TimingCount(self.handle)
        return [(self.L.ccb_TimingName(self.handle, i).decode(), int(self.L.ccb_TimingLevel(self.handle, i)),
                 float(self.L.ccb_TimingMs(self.handle, i))) for i in range(n)]

    def download(self, i: int, steps: int = ALL) -> dict:
        """Copy cage i's levels 1..D_i to host numpy arrays."""
        c, d = self.cages[i], self.depths[i]
        n = level_counts(c.V, c.H, c.E, c.F, d)
        he = np.zeros(n["Hcum"][d], self.rec)
        cr = np.zeros(n["Ccum"][d], CREASE_DTYPE)
        pt = np.zeros((n["Vcum"][d], 3), np.float32)
        sub = cc_Subd(self.meshes[i].pointer, pt.ctypes.data, he.ctypes.data, cr.ctypes.data, d)
        self._ok(self.L.ccb_Download(self.handle, i, ctypes.byref(sub), steps), "ccb_Download")
        return {"halfedges": he, "creases": cr, "points": pt, "counts": n}

    def close(self):
        if getattr(self, "handle", None):
            self.L.ccb_Release(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
