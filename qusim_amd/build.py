"""Build libqsim_hip.so (HIP, gfx950) and libqsim_host.so (host C++ adapter)
in-tree.  `python -m qusim_amd.build` or __graft_entry__.build()."""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
BUILD = os.path.join(HERE, "_build")
INCLUDE = os.path.join(ROOT, "include")

HIP_SOURCES = [
    "qsim_runtime.hip",
    "qsim_ops.hip",
    "qsim_fft.hip",
    "qsim_split.hip",
    "qsim_born.hip",
    "qsim_gen.hip",
]
HOST_SOURCES = ["hip_device.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ROCM_LIB = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")
ARCH = os.environ.get("QSIM_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    # no packed FP32 (v_pk_*_f32) instructions in any kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INCLUDE, "-I" + CSRC,
]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def _gen_headers():
    sys.path.insert(0, CSRC)
    try:
        import gen_twiddles  # noqa: E402
        gen_twiddles.main()
    finally:
        sys.path.pop(0)


def build_hip(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    _gen_headers()
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = []
    cmds = []
    for s in srcs:
        o = os.path.join(BUILD, s.replace(".hip", ".o"))
        objs.append(o)
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    out = os.path.join(HERE, "libqsim_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
         + ["-L" + ROCM_LIB, "-Wl,-rpath," + ROCM_LIB, "-lrocprofiler-sdk-roctx"])
    os.replace(tmp, out)
    return out


def build_host():
    srcs = [os.path.join(HOST, s) for s in HOST_SOURCES if os.path.exists(os.path.join(HOST, s))]
    if not srcs:
        return None
    out = os.path.join(HERE, "libqsim_host.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I" + INCLUDE, "-I" + HOST,
          "-o", tmp] + srcs + ["-L" + HERE, "-lqsim_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    return out


def build_host_driver():
    """tests/cpp/hip_device_driver: the reference's call sequences on HipDevice."""
    src = os.path.join(ROOT, "tests", "cpp", "hip_device_driver.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(ROOT, "tests", "cpp", "hip_device_driver")
    _run(["g++", "-O2", "-std=c++17", "-I" + INCLUDE, "-I" + HOST, "-o", out, src,
          "-L" + HERE, "-lqsim_host", "-lqsim_hip",
          "-Wl,-rpath,$ORIGIN/../../qusim_amd"])
    return out


def build_oracle():
    """The CPU oracle (test infrastructure) and, when the reference tree is
    present, the compiled-reference checker."""
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-C", odir, "liboracle.so"])
    if os.path.isdir("/root/reference/src"):
        _run(["make", "-C", odir, "ref"])


def build_all():
    lib = build_hip()
    host = build_host()
    build_host_driver()
    build_oracle()
    return lib, host


if __name__ == "__main__":
    print(build_all())
