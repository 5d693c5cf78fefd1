"""Build the native libraries in-tree (gfx950 only).

    python libaccint-public_amd/build.py [--incremental] [--only device|host|python]

Outputs (next to the Python package, libaccint-public_amd/libaccint/):
  liblai_device.so   HIP kernels + the C ABI of include/lai_device.h
  libaccint_host.so  C++20 host layer (Engine, FockBuilder, GpuFockBuilder, ...)
  _core*.so          pybind11 bindings of the host layer
Object files go to libaccint-public_amd/_build/.
"""
import argparse
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
DEV = os.path.join(PKG, "csrc", "device")
HOST = os.path.join(PKG, "csrc", "host")
PYB = os.path.join(PKG, "csrc", "python")
OUT = os.path.join(PKG, "libaccint")
BUILD = os.path.join(PKG, "_build")
INC = os.path.join(REPO, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # no packed-FP32 instructions anywhere in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INC, "-I" + DEV, "-Wno-unused-command-line-argument",
]
NPARTS = 8


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "unknown target feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def _stale(target, deps):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps)


def _dev_headers():
    return [os.path.join(DEV, f) for f in os.listdir(DEV) if f.endswith((".hpp", ".inc"))] + \
           [os.path.join(INC, "lai_device.h")]


def build_device(incremental=False, jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    hdr = _dev_headers()
    tasks = []
    objs = []
    for k in range(NPARTS):
        o = os.path.join(BUILD, "eri_inst_%d.o" % k)
        src = os.path.join(DEV, "eri_inst.hip")
        objs.append(o)
        if not incremental or _stale(o, [src] + hdr):
            tasks.append([HIPCC] + HIP_FLAGS + ["-DLAI_PART=%d" % k, "-DLAI_NPARTS=%d" % NPARTS,
                                                "-c", src, "-o", o])
    for name in ("aux_kernels", "grad_inst", "wl_kernels", "df", "fock_kernels", "persist", "api"):
        src = os.path.join(DEV, name + ".hip")
        o = os.path.join(BUILD, name + ".o")
        objs.append(o)
        if not incremental or _stale(o, [src] + hdr):
            tasks.append([HIPCC] + HIP_FLAGS + ["-c", src, "-o", o])
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, tasks))
    lib = os.path.join(OUT, "liblai_device.so")
    if not incremental or tasks or _stale(lib, objs):
        _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", lib] + objs)
    return lib


def _host_sources():
    return sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp"))


def build_host(incremental=False):
    srcs = _host_sources()
    if not srcs:
        return None
    lib = os.path.join(OUT, "libaccint_host.so")
    hdrs = []
    for root, _, files in os.walk(os.path.join(INC, "libaccint")):
        hdrs += [os.path.join(root, f) for f in files]
    hdrs += [os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".hpp")]
    if incremental and not _stale(lib, srcs + hdrs + [os.path.join(INC, "lai_device.h")]):
        return lib
    _run(["g++", "-std=c++20", "-O2", "-fPIC", "-shared", "-I" + INC, "-o", lib] + srcs +
         ["-L" + OUT, "-llai_device", "-Wl,-rpath,$ORIGIN"])
    return lib


def build_python(incremental=False):
    src = os.path.join(PYB, "bindings.cpp")
    if not os.path.exists(src):
        return None
    import pybind11
    ext = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    lib = os.path.join(OUT, "_core" + ext)
    if incremental and not _stale(lib, [src, os.path.join(OUT, "libaccint_host.so")]):
        return lib
    _run(["g++", "-std=c++20", "-O2", "-fPIC", "-shared", "-I" + INC,
          "-I" + pybind11.get_include(), "-I" + sysconfig.get_paths()["include"],
          src, "-o", lib, "-L" + OUT, "-laccint_host", "-llai_device", "-Wl,-rpath,$ORIGIN"])
    return lib


def build(incremental=False, only=None):
    if only in (None, "device"):
        build_device(incremental)
    if only in (None, "host"):
        build_host(incremental)
    if only in (None, "python"):
        build_python(incremental)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--incremental", action="store_true")
    ap.add_argument("--only", choices=["device", "host", "python"])
    args = ap.parse_args()
    build(args.incremental, args.only)
    print("built", file=sys.stderr)
