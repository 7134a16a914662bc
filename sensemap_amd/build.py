"""Build recipe for the native parts of sensemap_amd.

  libsensemap_pm.so  gfx950 HIP kernels + host engine + C-ABI (hipcc)
  libpm_scene.so     synthetic scene renderer (gcc, host only)

Both are built in-tree under sensemap_amd/lib/ so they travel with the
repository snapshot.  Every HIP translation unit is compiled with
-ffp-contract=off (the canonical FP32 evaluation order) and with packed FP32
instructions disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PM_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fvisibility=hidden", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}", "-Wall", "-Wno-unused-function",
]

HIP_SOURCES = ["pm_kernels.hip", "pm_engine.hip", "pm_server.hip", "pm_io.hip", "pm_cross.hip", "pm_model.hip",
               "pm_fusion.hip"]
# extra compile flags (tuning knobs such as -DPM_SWEEP_WAVES_PER_EU=6)
EXTRA = os.environ.get("PM_EXTRA_HIPFLAGS", "").split()


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")


def build_engine(out_path: str | None = None) -> str:
    import shutil
    import tempfile
    os.makedirs(LIB, exist_ok=True)
    objdir = tempfile.mkdtemp(prefix="pm_build_")
    objs = []
    jobs = []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        jobs.append([HIPCC, *HIP_FLAGS, *EXTRA, "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=len(jobs)) as ex:
        for f in [ex.submit(_run, j) for j in jobs]:
            f.result()
    out = out_path or os.path.join(LIB, "libsensemap_pm.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs])
    os.replace(tmp, out)
    shutil.rmtree(objdir, ignore_errors=True)
    return out


def build_scene() -> str:
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libpm_scene.so")
    _run(["gcc", "-O2", "-fopenmp", "-fPIC", "-shared", "-Wall", "-o", out,
          os.path.join(CSRC, "pm_scene.c"), "-lm"])
    return out


def build_all() -> None:
    build_scene()
    build_engine()


if __name__ == "__main__":
    build_all()
    print("built", os.listdir(LIB), file=sys.stderr)
