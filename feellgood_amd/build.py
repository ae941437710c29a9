"""Build recipe of the engine's native libraries (in-tree, under feellgood_amd/_lib).

* libfgmi.so      -- hipcc --offload-arch=gfx950, every *.hip of feellgood_amd/csrc
* libfgmi_host.so -- g++ -std=c++17, every *.cpp of feellgood_amd/host
* fgmi_run        -- the drop-in time-loop executable (host driver over both libraries)

Packed FP32 instructions are disabled for device code (-packed-fp32-ops).
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "_lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics", "-Wall", "-Wno-unused-result"]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wextra"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _jobs():
    return max(1, min(8, int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))


def build_hip():
    os.makedirs(os.path.join(LIB, "obj"), exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))
    objs = []
    procs = []
    for s in srcs:
        o = os.path.join(LIB, "obj", os.path.basename(s) + ".o")
        objs.append(o)
        cmd = [HIPCC] + HIP_FLAGS + ["-c", s, "-o", o]
        print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
        if len(procs) >= _jobs():
            for p in procs:
                if p.wait() != 0:
                    raise RuntimeError("hipcc failed")
            procs = []
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("hipcc failed")
    tmp = os.path.join(LIB, "libfgmi.so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, os.path.join(LIB, "libfgmi.so"))


def build_host_lib():
    os.makedirs(LIB, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(HERE, "host", "*.cpp")))
    drv_only = ("driver.cpp", "settings.cpp")
    lib_srcs = [s for s in srcs if not s.endswith(drv_only)]
    tmp = os.path.join(LIB, "libfgmi_host.so.tmp")
    _run(["g++"] + HOST_FLAGS + ["-shared", "-o", tmp] + lib_srcs)
    os.replace(tmp, os.path.join(LIB, "libfgmi_host.so"))


def build_driver():
    """Link the drop-in time loop against both libraries (libfgmi.so must exist)."""
    drv = os.path.join(HERE, "host", "driver.cpp")
    _run(["g++"] + HOST_FLAGS + ["-o", os.path.join(LIB, "fgmi_run"), drv,
                                 os.path.join(HERE, "host", "settings.cpp"),
                                 "-L" + LIB, "-lfgmi_host", "-lfgmi", "-Wl,-rpath,$ORIGIN",
                                 "-Wl,-rpath-link," + "/opt/rocm/lib"])


def build_host():
    build_host_lib()
    if os.path.exists(os.path.join(LIB, "libfgmi.so")):
        build_driver()


def build_all():
    build_host_lib()
    build_hip()
    build_driver()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"all": build_all, "hip": build_hip, "host": build_host}[what]()
