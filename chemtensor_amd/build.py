"""Build libchemtensor_gpu.so in-tree for gfx950 with hipcc (no JIT cache, no torch).

    python -m chemtensor_amd.build        (or chemtensor_amd.build.build())
"""

import json
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libchemtensor_gpu.so")

HIP_SOURCES = ["kernels.hip", "fused.hip", "pair.hip", "runtime.hip", "engine.hip", "lanczos.hip", "split.hip", "dmrg.hip", "serve.hip", "capi.hip"]
CPP_SOURCES = ["host_tensor.cpp", "host_linalg.cpp", "host_models.cpp", "host_molecular.cpp", "host_bond.cpp", "host_io.cpp"]
# save_mps / load_mps use the HDF5 C library through libchemtensor_h5.so, built only where HDF5's
# header and library are found (HDF5_ROOT, else the prefixes below)
H5_LIB = os.path.join(HERE, "libchemtensor_h5.so")
H5_PREFIXES = ["/opt/conda", "/usr", "/usr/local"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]
# gfx950 device code; packed FP32 instructions are disabled for all HIP code; MFMA accumulators in
# VGPRs (the AGPR form copies the accumulator VGPR <-> AGPR around every MFMA of a k-loop)
HIP_FLAGS = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-mllvm", "-amdgpu-mfma-vgpr-form"]
# host-only sources (block bookkeeping, host linear algebra): AVX2 + FMA servers
HOST_FLAGS = ["-march=x86-64-v3"]


def hipcc():
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


RESOURCES = os.path.join(BUILD, "kernel_resources.json")


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src.replace(".", "_") + ".o")
    cmd = [hipcc(), "-c", path, "-o", obj] + COMMON
    if src.endswith(".hip"):
        # the compiler's per-kernel resource report (registers, scratch, LDS, occupancy)
        cmd += HIP_FLAGS + ["-Rpass-analysis=kernel-resource-usage"]
    else:
        cmd += HOST_FLAGS
    r = subprocess.run(cmd, check=True, capture_output=True, text=True)
    return obj, (parse_resources(r.stderr) if src.endswith(".hip") else {})


def parse_resources(text):
    """{kernel symbol: {field: value}} from the kernel-resource-usage remarks of one compile"""
    out, cur = {}, None
    for line in text.splitlines():
        if "remark:" not in line or "kernel-resource-usage" not in line:
            continue
        body = line.split("remark:", 1)[1].rsplit("[-Rpass-analysis", 1)[0].strip()
        if body.startswith("Function Name:"):
            cur = body.split(":", 1)[1].strip()
            out[cur] = {}
        elif cur is not None and ":" in body:
            k, v = body.rsplit(":", 1)
            v = v.strip()
            out[cur][k.strip()] = int(v) if v.lstrip("-").isdigit() else v
    return out


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in HIP_SOURCES + CPP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    try:
        with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
            done = list(ex.map(_compile, srcs))
    except subprocess.CalledProcessError as e:
        sys.stderr.write(e.stderr or "")
        raise
    objs = [o for o, _ in done]
    resources = {}
    for _, res in done:
        resources.update(res)
    with open(RESOURCES, "w") as f:
        json.dump(resources, f, indent=1, sort_keys=True)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-o", tmp] + objs + ["--offload-arch=gfx950", "-Wl,-Bsymbolic", "-lpthread", "-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    h5 = build_h5()
    if verbose:
        print("built", h5 if h5 else "no HDF5 support (HDF5 headers or library not found)")
    return LIB


def find_hdf5():
    """(include dir, library dir) of an HDF5 C installation, or None"""
    roots = ([os.environ["HDF5_ROOT"]] if os.environ.get("HDF5_ROOT") else []) + H5_PREFIXES
    for root in roots:
        for inc in (os.path.join(root, "include"), os.path.join(root, "include", "hdf5", "serial")):
            if not os.path.exists(os.path.join(inc, "hdf5.h")):
                continue
            for lib in (os.path.join(root, "lib"), os.path.join(root, "lib", "x86_64-linux-gnu"),
                        os.path.join(root, "lib", "x86_64-linux-gnu", "hdf5", "serial")):
                if os.path.exists(os.path.join(lib, "libhdf5.so")):
                    return inc, lib
    return None


def build_h5():
    """libchemtensor_h5.so (the HDF5 calls of save_mps / load_mps) where HDF5 exists; None otherwise"""
    found = find_hdf5()
    if found is None:
        return None
    inc, libdir = found
    tmp = H5_LIB + ".tmp"
    cmd = ["g++", "-O2", "-fPIC", "-shared", "-std=c++17", "-I", inc, os.path.join(CSRC, "h5_shim.cpp"), "-o", tmp,
           "-L", libdir, "-lhdf5", "-Wl,-rpath," + libdir]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise RuntimeError("HDF5 shim build failed")
    os.replace(tmp, H5_LIB)
    return H5_LIB


if __name__ == "__main__":
    build(verbose=True)
