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
HOST_SOURCES = ["material.cpp", "config.cpp", "mesh.cpp", "domain.cpp", "boundary_tables.cpp", "heat_source.cpp", "amm.cpp",
                "solver.cpp", "host_capi.cpp"]


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


HDF5_PREFIX = os.environ.get("PMC_HDF5_PREFIX", "/opt/conda")


def build_importer():
    """pmc_import_material (phono3py HDF5 folder -> raw material table), where the
    HDF5 C++ library is installed."""
    inc, lib = os.path.join(HDF5_PREFIX, "include"), os.path.join(HDF5_PREFIX, "lib")
    if not os.path.exists(os.path.join(inc, "H5Cpp.h")):
        print(f"(no HDF5 C++ library under {HDF5_PREFIX}: pmc_import_material not built)", flush=True)
        return
    # Only the HDF5 libraries come from the prefix (a directory of links to them
    # on the tool's run path): the C++ runtime stays the system's.
    links = os.path.join(OUT, "hdf5")
    os.makedirs(links, exist_ok=True)
    objs = []
    for name in ("libhdf5_cpp.so", "libhdf5.so"):
        real = os.path.realpath(os.path.join(lib, name))
        soname = [f for f in os.listdir(lib) if f.startswith(name + ".") and f.count(".") == 2]
        for f in soname + [name]:
            dst = os.path.join(links, f)
            if os.path.lexists(dst):
                os.remove(dst)
            os.symlink(real, dst)
        objs.append(os.path.join(links, name))
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", INCLUDE, "-I", os.path.join(CSRC, "host"), "-I", inc,
          os.path.join(CSRC, "host", "hdf5_import.cpp"), os.path.join(CSRC, "host", "material.cpp"), *objs,
          "-Wl,-rpath,$ORIGIN/hdf5", "-o", os.path.join(OUT, "pmc_import_material")])


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
    build_importer()
    build_oracle()
    old = os.getcwd()
    os.chdir(ROOT)
    try:
        build_reference()
    finally:
        os.chdir(old)


if __name__ == "__main__":
    sys.exit(main())
