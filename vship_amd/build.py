"""Build recipe of the native parts, in-tree (no JIT cache, nothing installed).

    python -m vship_amd.build            # libvship.so + libvship_synth.so
    python -m vship_amd.build --all      # also the CPU oracle and, when the
                                         # reference sources are mounted, the
                                         # reference shim used by the tests

libvship.so is compiled for gfx950 with hipcc. Packed FP32 instructions are
disabled on the device side (-packed-fp32-ops); hipcc's host pass warns that it
does not know that feature, which is harmless.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("VSHIP_OFFLOAD_ARCH", "gfx950")

SOURCES = ["ssimu2_kernels.hip", "map_mfma.hip", "engine.hip", "butter.hip", "vship_api.hip", "serve.hip", "tune.hip",
           "vs_filter.hip", "noalloc.hip", "host_stage.hip"]
HIPFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-warning-option",
    "-I", os.path.join(ROOT, "include"),
]
# Per-source options. The matrix-core map kernel is scheduled by LLVM's
# memory-clause strategy: measured at 1920x1080, 240 pairs, 7.59-7.61 ms
# against 7.67-7.72 ms with the default (occupancy) strategy on one box, the
# same scores; the ILP strategy gave 7.53-7.64 ms and the iterative
# register-minimising one 9.26 ms.
SOURCE_FLAGS = {
    "map_mfma.hip": ["-mllvm", "-amdgpu-sched-strategy=max-memory-clause"],
}
LIB = os.path.join(HERE, "libvship.so")
SYNTH_SRC = os.path.join(HERE, "synth", "synth.c")
SYNTH_LIB = os.path.join(HERE, "synth", "libvship_synth.so")
OBJDIR = os.path.join(HERE, "csrc", "_build")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build_lib(jobs=8, verbose=False):
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = [os.path.join(OBJDIR, os.path.splitext(s)[0] + ".o") for s in srcs]

    def compile_one(src, obj):
        return _run([HIPCC, *HIPFLAGS, *SOURCE_FLAGS.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj])

    with cf.ThreadPoolExecutor(max_workers=min(jobs, len(srcs))) as ex:
        for out in ex.map(compile_one, srcs, objs):
            if verbose and out:
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


CLI_SRC = os.path.join(CSRC, "ffvship_cli.cpp")
CLI_BIN = os.path.join(HERE, "ffvship")


def build_cli():
    """The FFVship-style command line (vship_amd/ffvship) over libvship.so."""
    tmp = CLI_BIN + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-pthread", "-I", os.path.join(ROOT, "include"), "-o", tmp, CLI_SRC,
          "-L", HERE, "-lvship", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, CLI_BIN)
    return CLI_BIN


def build_synth():
    tmp = SYNTH_LIB + ".tmp"
    _run(["gcc", "-O3", "-mavx2", "-mfma", "-fPIC", "-fopenmp", "-shared", "-o", tmp, SYNTH_SRC, "-lm"])
    os.replace(tmp, SYNTH_LIB)
    return SYNTH_LIB


MOCK_SRC = os.path.join(ROOT, "tests", "vs_mock_host.cpp")
MOCK_LIB = os.path.join(ROOT, "tests", "_build", "libvs_mock_host.so")
CLI_HOOKS_LIB = os.path.join(ROOT, "tests", "_build", "libffvship_host.so")


def build_test_helpers():
    """The VapourSynth mock host the plugin tests drive (tests/vs_mock_host.cpp),
    and the command line's host logic with C entry points for the CPU tests."""
    os.makedirs(os.path.dirname(MOCK_LIB), exist_ok=True)
    tmp = MOCK_LIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-Wall", "-o", tmp, MOCK_SRC, "-ldl"])
    os.replace(tmp, MOCK_LIB)
    tmp = CLI_HOOKS_LIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-pthread", "-Wall", "-DFFVSHIP_TEST_HOOKS", "-I",
          os.path.join(ROOT, "include"), "-o", tmp, CLI_SRC, "-L", HERE, "-lvship", f"-Wl,-rpath,{HERE}"])
    os.replace(tmp, CLI_HOOKS_LIB)
    return MOCK_LIB


def build_oracle(with_reference=True):
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    ref = os.environ.get("VSHIP_REFERENCE", "/root/reference")
    if with_reference and os.path.exists(os.path.join(ref, "src", "ssimu2", "main.hpp")):
        _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", "clients", f"REFERENCE={ref}"])


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--all", action="store_true", help="also build the oracle and the reference shim")
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args(argv)
    print(build_lib(a.j, a.v))
    print(build_cli())
    print(build_synth())
    print(build_test_helpers())
    if a.all:
        build_oracle()
        print("oracle built")


if __name__ == "__main__":
    sys.exit(main())
