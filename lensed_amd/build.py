"""Build the lensed-amd shared library in-tree.

    python -m lensed_amd.build        (or: python lensed_amd/build.py)

Compiles the host runtime (C++, HIP runtime API, hipRTC) and the
statically compiled HIP kernels for gfx950 into lensed_amd/liblensed_amd.so.
The likelihood kernels themselves are generated and compiled at run time
with hipRTC from lensed_amd/kernels/ and the object plugins in
lensed_amd/objects/.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liblensed_amd.so")
ARCH = os.environ.get("LENSED_ARCH", "gfx950")

HOST_SOURCES = ["capi.cpp", "config.cpp", "context.cpp", "data.cpp", "nested.cpp", "prior.cpp", "program.cpp", "psf.cpp", "quadrature.cpp", "rtc.cpp", "tune.cpp"]
HIP_SOURCES = ["emulator.hip", "scratch.hip"]

# packed FP32 instructions stay out of every HIP code object built here
NO_PACKED_FP32 = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def clangxx():
    for c in (os.environ.get("CXX_HOST"), "/opt/rocm/llvm/bin/clang++"):
        if c and os.path.exists(c):
            return c
    return "g++"


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    raise RuntimeError("hipcc not found")


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


ASAN_LIB = os.path.join(HERE, "liblensed_amd_asan.so")
# host sanitizers for the CPU fuzz tests: AddressSanitizer and UBSan on the
# host C++ sources only (the HIP objects are the regular ones); the runtime
# is preloaded into the test process (asan_runtime())
SANITIZE = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g"]


def asan_runtime():
    """Path of clang's shared AddressSanitizer runtime (UBSan included)."""
    import glob
    for c in sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so")):
        return c
    return None


def build_sanitized():
    """liblensed_amd_asan.so: the host runtime built with ASan + UBSan, linked
    with the regular HIP objects of build()."""
    objdir = os.path.join(HERE, "_build")
    sandir = os.path.join(HERE, "_build_asan")
    os.makedirs(sandir, exist_ok=True)
    common = ["-O1", "-fPIC", "-std=c++17", "-Wall", "-Wno-unknown-pragmas", "-Wno-unused-value", "-Wno-unused-result"]
    objs = []
    for src in HOST_SOURCES:
        obj = os.path.join(sandir, src + ".o")
        run(["/opt/rocm/llvm/bin/clang++", "-c", os.path.join(CSRC, src), "-o", obj, "-D__HIP_PLATFORM_AMD__",
             "-I/opt/rocm/include"] + SANITIZE + common)
        objs.append(obj)
    objs += [os.path.join(objdir, src + ".o") for src in HIP_SOURCES]
    tmp = ASAN_LIB + ".tmp"
    run(["/opt/rocm/llvm/bin/clang++", "-shared", "-o", tmp] + objs +
        ["-L/opt/rocm/lib", "-lhiprtc", "-lamdhip64", "-ldl", "-lpthread", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, ASAN_LIB)
    return ASAN_LIB


def build(verbose=True):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    common = ["-O2", "-fPIC", "-std=c++17", "-Wall", "-Wno-unknown-pragmas", "-Wno-unused-value", "-Wno-unused-result"]
    for src in HOST_SOURCES:
        obj = os.path.join(objdir, src + ".o")
        run([clangxx(), "-c", os.path.join(CSRC, src), "-o", obj, "-D__HIP_PLATFORM_AMD__",
             "-I/opt/rocm/include"] + common)
        objs.append(obj)
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src + ".o")
        run([hipcc(), "-c", "-x", "hip", os.path.join(CSRC, src), "-o", obj,
             "--offload-arch=" + ARCH, "-O3"] + NO_PACKED_FP32 + common)
        objs.append(obj)
    tmp = LIB + ".tmp"
    run([hipcc(), "-shared", "-o", tmp] + objs + ["-L/opt/rocm/lib", "-lhiprtc", "-lamdhip64", "-ldl", "-lpthread",
                                                   "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    if "--sanitized" in sys.argv:
        build_sanitized()
    sys.exit(0)
