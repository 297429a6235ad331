"""Builds the in-tree shared libraries of the package.

libk3dom.so    the product: gfx950 kernels + C ABI (hipcc --offload-arch=gfx950)
libk3scene.so  the synthetic box-scene / LiDAR generator used by the benchmark
               and the tests (host C++)
k3dom_replay   frame-replay driver over the C++ drop-in header include/k3dom/dom.h
k3dom_replay_dsphdmib  the same over include/dsphdmib/dom.h
_k3window      libtorch autograd function over the differentiated window (links libk3dom.so)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # bit-reproducible FP32: no contraction, IEEE divide/sqrt, denormals kept
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    # packed FP32 instructions disabled
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib(verbose_asm=False):
    out = os.path.join(HERE, "libk3dom.so")
    srcs = [os.path.join(CSRC, f) for f in ("k3_kernels.hip", "k3_bgki.hip", "k3_adjoint.hip", "k3_noalloc.hip", "k3_api.cpp",
                                             "k3_params.cpp", "k3_window.cpp", "k3_batch.cpp", "k3_checkpoint.cpp",
                                             "k3_ckformat.cpp")]
    cmd = ["hipcc", *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-o", out + ".tmp", *srcs]
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build_scene():
    out = os.path.join(HERE, "libk3scene.so")
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp",
           os.path.join(CSRC, "k3_scene.cpp"), "-o", out + ".tmp"]
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build_replay():
    outs = []
    for name, defs in (("k3dom_replay", []), ("k3dom_replay_dsphdmib", ["-DK3_REPLAY_DSPHDMIB"])):
        out = os.path.join(HERE, name)
        cmd = ["g++", "-O2", "-std=c++17", *defs, "-I", os.path.join(ROOT, "include"),
               os.path.join(CSRC, "k3_replay.cpp"), "-L", HERE, "-lk3dom", "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"]
        _run(cmd)
        os.replace(out + ".tmp", out)
        outs.append(out)
    return outs


def build_torch_ext():
    import sysconfig
    import torch
    from torch.utils import cpp_extension as ce
    out = os.path.join(HERE, "_k3window" + sysconfig.get_config_var("EXT_SUFFIX"))
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", f"-D_GLIBCXX_USE_CXX11_ABI={abi}",
           "-DTORCH_EXTENSION_NAME=_k3window", "-DTORCH_API_INCLUDE_EXTENSION_H",
           *[f"-I{p}" for p in ce.include_paths()], f"-I{sysconfig.get_paths()['include']}",
           os.path.join(CSRC, "k3_torch_window.cpp"),
           *[f"-L{p}" for p in ce.library_paths()], "-lc10", "-ltorch", "-ltorch_cpu", "-ltorch_python",
           "-L", HERE, "-lk3dom", "-Wl,-rpath,$ORIGIN",
           *[f"-Wl,-rpath,{p}" for p in ce.library_paths()], "-o", out + ".tmp"]
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build_all():
    build_scene()
    build_lib()
    build_replay()
    build_torch_ext()


TARGETS = {"all": build_all, "lib": build_lib, "scene": build_scene, "replay": build_replay, "torch": build_torch_ext}

if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    bad = [w for w in what if w not in TARGETS]
    if bad:
        sys.exit(f"usage: python -m k3dom_amd.build [{'|'.join(TARGETS)}]...  (unknown: {' '.join(bad)})")
    for w in what:
        TARGETS[w]()
