"""Build libnexus_ckks.so in-tree with hipcc for gfx950.

Every source under nexus_amd/csrc is compiled (objects in nexus_amd/_build/) and
linked into nexus_amd/libnexus_ckks.so.  Usage: ``python nexus_amd/build.py``
(run as a file: importing the package needs the library already built).
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libnexus_ckks.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ROCM_INCLUDE = "/opt/rocm/include"
ARCH = "gfx950"
COMMON = ["-O3", "-fPIC", "-std=c++17", "-I" + CSRC, "-I" + INCLUDE, "-Wall", "-Wno-unused-function",
          "-Wno-unused-variable", "-Wno-unused-but-set-variable"]
# Packed FP32 instructions are kept out of every device code object.
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
# The encoder's double-precision FFT and the double-precision modular
# arithmetic of the NTT and the base conversions must round exactly as written.
NO_CONTRACT = {"encoder.cpp", "kernels_encode.hip", "kernels_ntt.hip", "kernels_fused.hip", "kernels_mix.hip",
               "kernels_hoist.hip", "kernels_poly.hip"}


def sources():
    out = []
    for root, _, files in os.walk(CSRC):
        for f in sorted(files):
            if f.endswith((".hip", ".cpp")):
                out.append(os.path.join(root, f))
    return out


def compile_one(src):
    rel = os.path.relpath(src, CSRC).replace(os.sep, "__")
    obj = os.path.join(OBJ, rel + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + COMMON + DEVICE
    else:
        # host-only C++: no device pass
        cmd = [CXX] + COMMON + ["-D__HIP_PLATFORM_AMD__", "-I" + ROCM_INCLUDE]
    if os.path.basename(src) in NO_CONTRACT:
        cmd += ["-ffp-contract=off"]
    cmd += ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), msgs))
    return obj, msgs


def build(verbose=False, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    objs = []
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        for obj, msgs in ex.map(compile_one, srcs):
            objs.append(obj)
            if verbose and msgs.strip():
                print(msgs)
    tmp = LIB + ".tmp"
    # every symbol must resolve inside the library (or its HIP runtime)
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-Wl,--no-undefined", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), r.stdout + r.stderr))
    os.replace(tmp, LIB)
    return LIB


EXAMPLE_SRC = os.path.join(os.path.dirname(HERE), "tests", "cpp", "dropin_example.cpp")


def build_example(out):
    """the compiled C++ caller of the drop-in (tests/cpp/dropin_example.cpp),
    linked against the built library (found next to the binary's ../../nexus_amd)"""
    rel = os.path.relpath(HERE, os.path.dirname(os.path.abspath(out)))
    cmd = [CXX, "-O1", "-std=c++17", "-I" + INCLUDE, EXAMPLE_SRC, "-o", out, "-L" + HERE, "-lnexus_ckks",
           "-Wl,-rpath,$ORIGIN/" + rel, "-Wl,-rpath-link,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("example build failed: %s\n%s" % (" ".join(cmd), r.stdout + r.stderr))
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
