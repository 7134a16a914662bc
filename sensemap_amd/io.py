"""The reference's on-disk map formats (pm_mat_*, pm_normal_*).

Mat<T>::Write / Read (src/util/mat.h:178-215): text header "W&H&D&", then
the values little-endian, slice-major.  DepthMap files are Mat<float> with
depth 1; NormalMap files (src/mvs/normal_map.cc:121-178) hold one uint32 code
per pixel."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi

_DTYPES = {np.dtype(np.uint8): 1, np.dtype(np.uint16): 2, np.dtype(np.float32): 4, np.dtype(np.uint32): 4,
           np.dtype(np.int32): 4, np.dtype(np.float64): 8}


def _lib():
    return abi.load_engine()


def write_mat(path: str, data: np.ndarray) -> None:
    """data: (depth, height, width) or (height, width)."""
    a = np.ascontiguousarray(data)
    if a.ndim == 2:
        a = a[None]
    if a.dtype not in _DTYPES:
        raise ValueError(f"unsupported dtype {a.dtype}")
    d, h, w = a.shape
    abi.check(_lib().pm_mat_write(os.fsencode(path), w, h, d, _DTYPES[a.dtype], a.ctypes.data_as(C.c_void_p)))


def read_mat(path: str, dtype) -> np.ndarray:
    """Returns (depth, height, width)."""
    w, h, d = C.c_int32(), C.c_int32(), C.c_int32()
    abi.check(_lib().pm_mat_read_header(os.fsencode(path), C.byref(w), C.byref(h), C.byref(d)))
    a = np.empty((d.value, h.value, w.value), dtype=dtype)
    abi.check(_lib().pm_mat_read(os.fsencode(path), w.value, h.value, d.value, _DTYPES[a.dtype],
                                 a.ctypes.data_as(C.c_void_p)))
    return a


def encode_normals(normal: np.ndarray) -> np.ndarray:
    """(3, H, W) float32 -> (H, W) uint32 codes (NormalMap::Encode)."""
    n = np.ascontiguousarray(normal, dtype=np.float32)
    codes = np.empty(n.shape[1:], np.uint32)
    abi.check(_lib().pm_normal_encode(n[0].size, n.ctypes.data_as(C.POINTER(C.c_float)),
                                      codes.ctypes.data_as(C.POINTER(C.c_uint32))))
    return codes


def decode_normals(codes: np.ndarray) -> np.ndarray:
    """(H, W) uint32 codes -> (3, H, W) float32 (NormalMap::Decode)."""
    c = np.ascontiguousarray(codes, dtype=np.uint32)
    n = np.empty((3,) + c.shape, np.float32)
    abi.check(_lib().pm_normal_decode(c.size, c.ctypes.data_as(C.POINTER(C.c_uint32)),
                                      n.ctypes.data_as(C.POINTER(C.c_float))))
    return n


def write_depth_map(path: str, depth: np.ndarray) -> None:
    write_mat(path, np.asarray(depth, np.float32))


def read_depth_map(path: str) -> np.ndarray:
    return read_mat(path, np.float32)[0]


def write_normal_map(path: str, normal: np.ndarray) -> None:
    write_mat(path, encode_normals(normal))


def read_normal_map(path: str) -> np.ndarray:
    return decode_normals(read_mat(path, np.uint32)[0])
