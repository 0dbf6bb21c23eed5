"""Build recipe for libyafel_hip.so (gfx950) and the C++ test drivers.

    python -m yafel_amd.build            # build everything in-tree

Each .hip translation unit is compiled by hipcc for gfx950 in parallel, then
linked into yafel_amd/libyafel_hip.so.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libyafel_hip.so")
ARCH = os.environ.get("YAFEL_HIP_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-fno-gpu-rdc",
    # packed FP32 instructions stay off (hipcc warns that the host pass does not know the feature)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
    "-Wno-unused-result",
]

SOURCES = ["runtime.hip", "sparse.hip", "blas.hip", "krylov.hip", "sort.hip", "triplets.hip",
           "mesh.hip", "assembly.hip", "multi.hip", "serve.hip", "dg.hip", "recovery.hip", "io.hip"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    return r.returncode, out


def build(verbose=False, sources=None):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in (sources or SOURCES) if os.path.exists(os.path.join(CSRC, s))]
    jobs = []
    for s in srcs:
        obj = os.path.join(OBJ, s.replace(".hip", ".o"))
        jobs.append((s, [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, s), "-o", obj], obj))
    objs = []
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        futs = {ex.submit(_run, cmd): (s, obj) for s, cmd, obj in jobs}
        errors = []
        for f in cf.as_completed(futs):
            s, obj = futs[f]
            rc, out = f.result()
            if rc != 0:
                errors.append(f"--- {s} ---\n{out}")
            elif verbose and out.strip():
                print(f"--- {s} ---\n{out}")
            objs.append(obj)
    if errors:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errors))
    objs = [os.path.join(OBJ, s.replace(".hip", ".o")) for s in srcs]
    tmp = LIB + ".tmp"
    rc, out = _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    if rc != 0:
        raise RuntimeError("link failed:\n" + out)
    os.replace(tmp, LIB)
    return LIB


def build_cpp_tests():
    """C++ drivers under tests/cpp that use the header-only YAFEL-named layer."""
    src_dir = os.path.join(ROOT, "tests", "cpp")
    if not os.path.isdir(src_dir):
        return []
    outs = []
    for f in sorted(os.listdir(src_dir)):
        if not f.endswith(".cpp"):
            continue
        exe = os.path.join(src_dir, f[:-4])
        cmd = ["g++", "-std=c++17", "-O2", "-I" + os.path.join(ROOT, "include"),
               "-I/opt/rocm/include", os.path.join(src_dir, f), "-o", exe,
               "-L" + HERE, "-lyafel_hip", "-Wl,-rpath," + HERE]
        rc, out = _run(cmd)
        if rc != 0:
            raise RuntimeError(f"g++ failed for {f}:\n{out}")
        outs.append(exe)
    return outs


if __name__ == "__main__":
    lib = build(verbose="-v" in sys.argv)
    print("built", lib)
    for exe in build_cpp_tests():
        print("built", exe)
