"""Builds libwv_hip.so (gfx950 kernels + C ABI + host coder) and the drop-in
driver wv_compress_gpu, in-tree, with hipcc.

    python -m wavelets_amd.build            # or: python wavelets_amd/build.py
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libwv_hip.so")
DRIVER = os.path.join(HERE, "wv_compress_gpu")

ARCH = os.environ.get("WV_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

# -ffp-contract=off: coefficients and quantizer arithmetic must follow the
# reference operation order (no fused multiply-add); packed FP32 ops are
# disabled on this platform.
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
          "-Wno-unused-function", "-Wno-unused-variable",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
DEVICE = ["--offload-arch=" + ARCH, "-munsafe-fp-atomics"]

SOURCES = ["wv_dwt.hip", "wv_select.hip", "wv_quant.hip", "wv_lloyd.hip", "wv_encode.hip", "wv_decode.hip", "wv_restore.hip",
           "wv_api.cpp", "wv_compress.cpp", "wv_params.cpp", "wv_huffman.cpp",
           "wv_batch.cpp", "wv_serve.cpp", "wv_cube.cpp", "wv_split.cpp", "wv_opt.hip", "wv_copy.hip"]


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    cmd = [HIPCC] + COMMON
    # every translation unit is HIP (host-only ones just carry no kernels)
    cmd += DEVICE + ["-x", "hip"]
    cmd += ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in (r.stderr or "").splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), err))
    return obj, err


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(_compile, srcs))
    objs = [o for o, _ in results]
    if verbose:
        for o, e in results:
            if e.strip():
                print(e)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, LIB)
    drv = os.path.join(CSRC, "wv_compress_gpu.cpp")
    if os.path.exists(drv):
        cmd = [HIPCC] + COMMON + ["-D__HIP_PLATFORM_AMD__", drv, "-o", DRIVER,
                                  "-L" + HERE, "-lwv_hip", "-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("driver link failed: %s\n%s" % (" ".join(cmd), r.stderr))
    return LIB


SAN_LIB = os.path.join(HERE, "libwv_hip_san.so")
SAN_BUILD = os.path.join(HERE, "_build_san")
# host code only: the device code objects of the normal build are reused
SAN_FLAGS = ["-O1", "-g", "-Xarch_host", "-fno-omit-frame-pointer",
             "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             "-Xarch_host", "-fno-sanitize-recover=undefined"]


def asan_runtime():
    """The clang AddressSanitizer runtime (it carries the UBSan handlers)
    that a process loading the sanitized library must preload."""
    r = subprocess.run([HIPCC, "-print-file-name=libclang_rt.asan-x86_64.so"],
                       capture_output=True, text=True)
    path = r.stdout.strip()
    if not os.path.isabs(path):
        import glob
        hits = glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so")
        path = hits[0] if hits else ""
    return path


def build_sanitized():
    """libwv_hip_san.so: the host translation units (the .cube reader and
    writer, the host Huffman coder and decoder, the split plan, the API
    checks) rebuilt with AddressSanitizer and UndefinedBehaviorSanitizer on
    the host side, linked with the device-code objects of the normal build.
    Loaded through WV_LIB with the ASan runtime preloaded; CPU use only."""
    hip_objs = [os.path.join(BUILD, s + ".o") for s in SOURCES if s.endswith(".hip")]
    if not all(os.path.exists(o) for o in hip_objs):
        build()
    os.makedirs(SAN_BUILD, exist_ok=True)
    host = [s for s in SOURCES if s.endswith(".cpp")]

    def one(src):
        obj = os.path.join(SAN_BUILD, src + ".o")
        cmd = [HIPCC] + [f for f in COMMON if f != "-O3"] + SAN_FLAGS + DEVICE + \
            ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("sanitized compile failed: %s\n%s" % (" ".join(cmd), r.stderr))
        return obj

    with cf.ThreadPoolExecutor(max_workers=min(8, len(host))) as ex:
        objs = list(ex.map(one, host))
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", SAN_LIB] + objs + \
        hip_objs + ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("sanitized link failed: %s\n%s" % (" ".join(cmd), r.stderr))
    return SAN_LIB


if __name__ == "__main__":
    if "--sanitize" in sys.argv:
        print(build_sanitized())
    else:
        print(build(verbose="-v" in sys.argv))
