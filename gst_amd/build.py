"""Builds the native libraries in-tree.

  gst_amd/libgst_amd.so     the gfx950 decoder (HIP kernels + host orchestration + C ABI)
  gst_amd/libgst_encode.so  CPU texture generator and GST encoder (input tooling)
  oracle/liboracle_gst.so   the CPU restatement used as the checker (test-only)
  oracle/_ref/ref_driver    the reference's encoder and CPU stage functions, only where
                            the reference sources are present (test-only)

Every call compiles; outputs are written to a temporary name and renamed, so a
reader never maps a half-written library.
"""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
REFERENCE = "/root/reference"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Wall", "-Wno-unused-command-line-argument",
    # No packed FP32 instructions anywhere in the device code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd, cwd=None):
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        raise RuntimeError("build failed (%s):\n%s" % (" ".join(cmd), proc.stdout))
    return proc.stdout


def _atomic_build(out, cmd_for):
    tmp = out + ".tmp%d" % os.getpid()
    try:
        _run(cmd_for(tmp))
        os.replace(tmp, out)
    finally:
        if os.path.exists(tmp):
            os.unlink(tmp)
    return out


def build_decoder():
    out = os.path.join(HERE, "libgst_amd.so")
    srcs = [os.path.join(CSRC, "gst_kernels.hip"), os.path.join(CSRC, "gst_decoder.cpp"),
            os.path.join(CSRC, "gst_serve.cpp"), os.path.join(CSRC, "gst_ans.cpp")]
    srcs = [s for s in srcs if os.path.exists(s)]
    return _atomic_build(out, lambda o: [HIPCC, *HIP_FLAGS, "-I" + INCLUDE, "-I" + CSRC, *srcs,
                                         "-o", o, "-lpthread"])


def build_encoder():
    out = os.path.join(HERE, "libgst_encode.so")
    cxx = shutil.which("g++") or "g++"
    # -ffp-contract=off: gste_compress_dxt repeats the reference encoder's float
    # arithmetic operation by operation, with no fused multiply-adds.
    return _atomic_build(out, lambda o: [cxx, "-std=c++17", "-O2", "-Wall", "-Wextra", "-fPIC", "-shared",
                                         "-pthread", "-ffp-contract=off", "-I" + INCLUDE, "-I" + CSRC,
                                         os.path.join(CSRC, "gst_encode.cpp"), os.path.join(CSRC, "gst_gentc.cpp"),
                                         "-o", o])


def build_cpp_tests():
    """tests/cpp/codec_test and tests/cpp/ans_ocl_test: the reference's end-to-end decoder
    test and its ANS decoder tests against the C++ API; tests/cpp/encoder_test: the
    encoder's C++ API; tests/cpp/ans_units_test: the rANS work-unit enumeration."""
    outs = []
    for name in ("codec_test", "ans_ocl_test"):
        src = os.path.join(ROOT, "tests", "cpp", name + ".cpp")
        out = os.path.join(ROOT, "tests", "cpp", name)
        outs.append(_atomic_build(out, lambda o, src=src: [
            HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-Wall", "-Wno-unused-command-line-argument",
            "-I" + INCLUDE, src, "-L" + HERE, "-lgst_amd", "-Wl,-rpath," + HERE, "-o", o]))
    # The encoder's C++ API needs no GPU: host compiler, libgst_encode.so only.
    src = os.path.join(ROOT, "tests", "cpp", "encoder_test.cpp")
    cxx = shutil.which("g++") or "g++"
    outs.append(_atomic_build(os.path.join(ROOT, "tests", "cpp", "encoder_test"), lambda o: [
        cxx, "-std=c++17", "-O2", "-Wall", "-Wextra", "-I" + INCLUDE, src, "-L" + HERE, "-lgst_encode",
        "-Wl,-rpath," + HERE, "-o", o]))
    # The rANS work-unit enumeration the host and the kernels share: host code only.
    src = os.path.join(ROOT, "tests", "cpp", "ans_units_test.cpp")
    outs.append(_atomic_build(os.path.join(ROOT, "tests", "cpp", "ans_units_test"), lambda o: [
        cxx, "-std=c++17", "-O2", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
        "-I" + CSRC, "-I" + INCLUDE, src, "-o", o]))
    return outs


def build_sanitizer_tests():
    """Host-sanitizer builds of the host code for tests/test_sanitizers.py (CPU only):
    the server's dry-run driver with ThreadSanitizer and with AddressSanitizer (the host
    side of libgst_amd's sources; device code is built as usual and never run by it), the
    encoder's C++ test and the oracle with AddressSanitizer + UndefinedBehaviorSanitizer,
    and the oracle's thread pool with ThreadSanitizer."""
    cpp = os.path.join(ROOT, "tests", "cpp")
    lib_srcs = [os.path.join(CSRC, n) for n in ("gst_serve.cpp", "gst_decoder.cpp", "gst_kernels.hip", "gst_ans.cpp")]
    outs = []
    for kind in ("thread", "address"):
        out = os.path.join(cpp, "serve_dryrun_%s" % ("tsan" if kind == "thread" else "asan"))
        outs.append(_atomic_build(out, lambda o, kind=kind: [
            HIPCC, "--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-Wno-unused-command-line-argument",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Xarch_host", "-fsanitize=%s" % kind, "-Xarch_host", "-fno-omit-frame-pointer",
            "-I" + INCLUDE, "-I" + CSRC, *lib_srcs, os.path.join(cpp, "serve_dryrun_test.cpp"), "-o", o,
            "-lpthread"]))
    gcc = shutil.which("gcc") or "gcc"
    cxx = shutil.which("g++") or "g++"
    oracle = os.path.join(ROOT, "oracle")
    for tag, flags in (("asan", ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"]),
                       ("tsan", ["-fsanitize=thread"])):
        outs.append(_atomic_build(os.path.join(cpp, "oracle_sanitize_%s" % tag), lambda o, flags=flags: [
            gcc, "-std=c11", "-O1", "-g", "-fno-omit-frame-pointer", *flags, "-pthread", "-I" + oracle,
            os.path.join(oracle, "gst_oracle.c"), os.path.join(cpp, "oracle_sanitize_test.c"), "-o", o]))
    outs.append(_atomic_build(os.path.join(cpp, "encoder_asan"), lambda o: [
        cxx, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
        "-fno-sanitize-recover=undefined", "-pthread", "-ffp-contract=off", "-I" + INCLUDE, "-I" + CSRC,
        os.path.join(cpp, "encoder_test.cpp"), os.path.join(CSRC, "gst_encode.cpp"),
        os.path.join(CSRC, "gst_gentc.cpp"), "-o", o]))
    return outs


def build_oracle():
    """The checker libraries (not part of the product)."""
    oracle = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-B", "-C", oracle, os.path.join(oracle, "liboracle_gst.so")])
    if os.path.isdir(REFERENCE):
        _run(["make", "-s", "-C", oracle, "ref"])


def build_all():
    build_decoder()
    build_encoder()
    build_cpp_tests()
    build_oracle()
    build_sanitizer_tests()


if __name__ == "__main__":
    build_all()
    print("built:", os.path.join(HERE, "libgst_amd.so"), os.path.join(HERE, "libgst_encode.so"))
