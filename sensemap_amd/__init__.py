"""sensemap_amd: gfx950 (MI355X) PatchMatch multi-view-stereo engine.

Drop-in for the PatchMatch stage of SenseMap's src/mvs: the C-ABI in
include/sensemap_pm.h (libsensemap_pm.so, HIP kernels + C++ host engine),
with a Python mirror of PatchMatchOptions / Problem / PatchMatchACMMCuda in
sensemap_amd.patch_match and seeded synthetic scenes in sensemap_amd.scene.
"""
from .abi import load_engine, PMError  # noqa: F401

__all__ = ["load_engine", "PMError"]
