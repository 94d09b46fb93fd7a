"""Shared test helpers: workload cages and array comparisons."""
from __future__ import annotations

import numpy as np

from halfedgecatmullclark_amd import cage as C


def small_cages():
    """(name, cage, depth) cases small enough for the oracle to finish in seconds."""
    return [
        ("cube", C.cube(0), 5),
        ("cube_s7", C.cube(7), 4),
        ("torus16x8", C.torus(16, 8, seed=1), 4),
        ("torus_r1.3", C.torus(6, 5, seed=4, ring_sharpness=((0, 1.3), (2, 0.4))), 4),
        ("irregular_l2", C.irregular(2, seed=3), 3),
        ("irregular_l3", C.irregular(3, seed=11), 3),
        ("irregular_soft", C.from_faces(*C.irregular_mesh(2, seed=5)[:4], creases=C.irregular_mesh(2, seed=5)[4],
                                        sharpen_boundaries=False), 3),
    ]


def assert_bytes_equal(a: np.ndarray, b: np.ndarray, what: str):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    if a.tobytes() != b.tobytes():
        av = a.view(np.uint8).reshape(len(a), -1)
        bv = b.view(np.uint8).reshape(len(b), -1)
        bad = np.nonzero(np.any(av != bv, axis=1))[0]
        raise AssertionError(f"{what}: {bad.size} of {len(a)} records differ; first {bad[:5].tolist()}: "
                             f"{a[bad[:3]]} vs {b[bad[:3]]}")


def point_error(p: np.ndarray, ref: np.ndarray) -> tuple[float, float]:
    """(max abs error, RMS error) of p against ref, both relative to the
    largest coordinate magnitude of ref."""
    ref = np.asarray(ref, np.float64)
    scale = max(float(np.abs(ref).max()), 1e-30)
    d = np.asarray(p, np.float64) - ref
    return float(np.abs(d).max()) / scale, float(np.sqrt(np.mean(d * d))) / scale


def assert_bits_equal(a: np.ndarray, b: np.ndarray, what: str, chunk: int = 1 << 26):
    """Bitwise equality of two large arrays, compared as 32-bit words in
    chunks (no full-size temporaries)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    assert a.shape == b.shape and a.dtype == b.dtype, f"{what}: {a.shape}/{a.dtype} vs {b.shape}/{b.dtype}"
    aw = a.reshape(-1).view(np.uint32)
    bw = b.reshape(-1).view(np.uint32)
    per = aw.size // max(len(a), 1)
    for lo in range(0, aw.size, chunk):
        x, y = aw[lo:lo + chunk], bw[lo:lo + chunk]
        if not np.array_equal(x, y):
            bad = lo + np.nonzero(x != y)[0]
            rec = np.unique(bad // per)
            raise AssertionError(f"{what}: records differ from {rec[:5].tolist()} "
                                 f"({rec.size} in words {lo}..{lo + x.size}); first {a[rec[:3]]} vs {b[rec[:3]]}")
