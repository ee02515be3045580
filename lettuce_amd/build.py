"""Build lettuce_amd/liblettuce_hip.so with hipcc for gfx950.

Each translation unit is compiled to an object in lettuce_amd/_build/ (in
parallel), then linked into the shared library in-tree, so the library travels
with the source tree.  Usage: ``python -m lettuce_amd.build`` or ``build()``.
"""
import concurrent.futures
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblettuce_hip.so")
ARCH = os.environ.get("LT_OFFLOAD_ARCH", "gfx950")

SOURCES = ["lt_api.cpp", "lt_hmc_api.cpp", "lt_batch_api.cpp", "lt_tune.cpp", "lt_field.hip", "lt_update.hip",
           "lt_obs.hip", "lt_smear.hip", "lt_hmc.hip", "lt_adjoint.hip", "lt_batch.hip", "lt_io.cpp", "lt_bmw.hip",
           "lt_cb4.hip"]

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]
DEVICE = [
    f"--offload-arch={ARCH}",
    # no packed FP32 instructions in any device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def _compile(src, defines=(), objdir=OBJ):
    path = os.path.join(CSRC, src)
    obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + DEVICE + [f"-D{d}" for d in defines] + ["-x", "hip", "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj, r.stderr


def build(verbose=False, defines=(), out=LIB, objdir=OBJ):
    """Compile every source and link ``out``; ``defines`` are extra NAME=VALUE macros
    (launch-shape knobs of lt_device.hpp)."""
    os.makedirs(objdir, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        results = list(ex.map(lambda s: _compile(s, defines, objdir), SOURCES))
    if verbose:
        for obj, err in results:
            if err.strip():
                print(err)
    objs = [o for o, _ in results]
    tmp = out + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
