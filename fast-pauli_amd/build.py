"""Build recipe for the MI355X library and its host bindings (in-tree).

    python fast-pauli_amd/build.py            # everything
    python fast-pauli_amd/build.py hip        # libfastpauli_hip.so only

Outputs land in fast-pauli_amd/fast_pauli/ next to the Python package:
  libfastpauli_hip.so   HIP kernels + C ABI (hipcc, gfx950 only)
  _fast_pauli*.so       Python surface: pybind11 over the C++ host API, which
                        reaches the GPU only through the C ABI
"""

from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PKG = os.path.join(HERE, "fast_pauli")
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
CPP_INCLUDE = os.path.join(HERE, "cpp", "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are disabled on purpose: some of their forms have given
# silently wrong results on this hardware.
HIP_FLAGS = ["-O3", f"--offload-arch={ARCH}", "-std=c++20", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-result"]

LIB_HIP = os.path.join(PKG, "libfastpauli_hip.so")


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    noise = "is not a recognized feature for this target"
    out = "\n".join(line for line in proc.stdout.splitlines() if noise not in line)
    if out.strip():
        print(out, flush=True)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}")


def build_hip() -> str:
    """Compiles every translation unit in parallel (each carries its own gfx950
    code object; no kernel is launched from another unit) and links the library."""
    from concurrent.futures import ThreadPoolExecutor

    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    objs = [os.path.join(CSRC, os.path.basename(src)[:-4] + ".o") for src in srcs]
    # plain C++ units (no HIP): the host compiler
    cpp_srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".cpp"))
    cpp_objs = [os.path.join(CSRC, os.path.basename(src)[:-4] + ".host.o") for src in cpp_srcs]
    cxx = os.environ.get("CXX", "g++")
    cmds = [[HIPCC, *HIP_FLAGS, "-c", f"-I{INCLUDE}", f"-I{CSRC}", "-o", o, src] for src, o in zip(srcs, objs)]
    cmds += [[cxx, "-O3", "-std=c++20", "-fPIC", "-Wall", "-Wextra", "-c", f"-I{INCLUDE}", f"-I{CSRC}", "-o", o, src]
             for src, o in zip(cpp_srcs, cpp_objs)]
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", "0") or 0) or (os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        list(pool.map(_run, cmds))
    objs += cpp_objs
    tmp = LIB_HIP + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *objs])
    os.replace(tmp, LIB_HIP)
    for o in objs:
        os.remove(o)
    return LIB_HIP


def ext_suffix() -> str:
    return sysconfig.get_config_var("EXT_SUFFIX") or ".so"


def build_python() -> str | None:
    src = os.path.join(HERE, "cpp", "src", "fast_pauli_py.cpp")
    if not os.path.exists(src):
        return None
    import pybind11

    out = os.path.join(PKG, "_fast_pauli" + ext_suffix())
    tmp = out + ".tmp"
    cxx = os.environ.get("CXX", "g++")
    _run([cxx, "-O3", "-std=c++20", "-fPIC", "-shared", "-Wall", "-Wextra",
          f"-I{INCLUDE}", f"-I{CPP_INCLUDE}", f"-I{pybind11.get_include()}",
          f"-I{sysconfig.get_paths()['include']}", "-o", tmp, src,
          f"-L{PKG}", "-lfastpauli_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    return out


def build_cpp_test() -> str | None:
    """The C++ host API test program (tests/cpp/test_host_api.cpp)."""
    src = os.path.join(ROOT, "tests", "cpp", "test_host_api.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(ROOT, "tests", "cpp", "test_host_api")
    cxx = os.environ.get("CXX", "g++")
    _run([cxx, "-O2", "-std=c++20", "-Wall", "-Wextra", f"-I{INCLUDE}", f"-I{CPP_INCLUDE}", "-o", out + ".tmp",
          src, f"-L{PKG}", "-lfastpauli_hip", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN/../../fast-pauli_amd/fast_pauli"])
    os.replace(out + ".tmp", out)
    return out


# address and undefined-behaviour sanitizers for the host-only table-builder test
SANITIZE = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]


def build_tile_test() -> str | None:
    """The host tile-table builder's unit test (tests/cpp/test_tile_tables.cpp), built
    from the plain C++ builder sources with the sanitizers; needs no GPU."""
    src = os.path.join(ROOT, "tests", "cpp", "test_tile_tables.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(ROOT, "tests", "cpp", "test_tile_tables")
    cxx = os.environ.get("CXX", "g++")
    _run([cxx, "-O1", "-g", "-std=c++20", "-Wall", "-Wextra", *SANITIZE, f"-I{INCLUDE}", f"-I{CSRC}",
          "-o", out + ".tmp", src, os.path.join(CSRC, "fp_tiles_host.cpp"), os.path.join(CSRC, "fp_error.cpp")])
    os.replace(out + ".tmp", out)
    return out


def build_all() -> list[str]:
    built = [build_hip()]
    for extra in (build_python(), build_cpp_test(), build_tile_test()):
        if extra:
            built.append(extra)
    return built


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    if what == "hip":
        build_hip()
    elif what == "python":
        build_python()
    elif what == "cpptest":
        build_cpp_test()
    elif what == "tiletest":
        build_tile_test()
    else:
        build_all()
