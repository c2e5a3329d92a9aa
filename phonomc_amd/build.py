"""Builds the HIP device library, the C++17 host facade and the CLI in-tree.

    python -m phonomc_amd.build

Outputs go to phonomc_amd/_build/: libpmc_device.so (gfx950 kernels + C ABI),
libpmc_host.so (TOML/tables/facade), pmc (command-line driver).
"""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "phonomc_amd")
CSRC = os.path.join(PKG, "csrc")
OUT = os.path.join(PKG, "_build")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PMC_OFFLOAD_ARCH", "gfx950")

DEVICE_SOURCES = ["pmc_kernels.hip", "pmc_capi.hip", "pmc_nn.hip"]
HOST_SOURCES = ["material.cpp", "config.cpp", "mesh.cpp", "domain.cpp", "boundary_tables.cpp", "heat_source.cpp", "solver.cpp",
                "host_capi.cpp"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_device():
    objs = []
    for src in DEVICE_SOURCES:
        obj = os.path.join(OUT, src.replace(".hip", ".o"))
        _run([HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-munsafe-fp-atomics",
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              "-I", INCLUDE, "-I", CSRC, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", os.path.join(OUT, "libpmc_device.so")])


def build_host():
    objs = []
    for src in HOST_SOURCES:
        obj = os.path.join(OUT, "host_" + src.replace(".cpp", ".o"))
        _run(["g++", "-O2", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-Wno-unused-function",
              "-I", INCLUDE, "-I", os.path.join(CSRC, "host"), "-c", os.path.join(CSRC, "host", src), "-o", obj])
        objs.append(obj)
    _run(["g++", "-shared", "-fopenmp", *objs, "-L", OUT, "-lpmc_device", "-Wl,-rpath,$ORIGIN",
          "-o", os.path.join(OUT, "libpmc_host.so")])
    _run(["g++", "-O2", "-std=c++17", "-I", INCLUDE, "-I", os.path.join(CSRC, "host"),
          os.path.join(CSRC, "host", "cli.cpp"), "-L", OUT, "-lpmc_host", "-lpmc_device",
          "-Wl,-rpath,$ORIGIN", "-o", os.path.join(OUT, "pmc")])


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-j8"])


def build_reference():
    ref = os.environ.get("PHONOMC_REFERENCE", "/root/reference")
    if os.path.isdir(os.path.join(ref, "src")):
        _run(["make", "-s", "-f", os.path.join(ROOT, "oracle", "ref.mk"), f"REF={ref}", "-j8"])


def main():
    os.makedirs(OUT, exist_ok=True)
    build_device()
    build_host()
    build_oracle()
    old = os.getcwd()
    os.chdir(ROOT)
    try:
        build_reference()
    finally:
        os.chdir(old)


if __name__ == "__main__":
    sys.exit(main())
