"""Build libfeddgpu.so (gfx950) in-tree with hipcc, and the CPU checker in oracle/.

Usage: python -m feddlib_amd.build  [--no-oracle]
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfeddgpu.so")

SOURCES = ["fedd_host.hip", "fedd_ns.hip", "fedd_elas.hip", "fedd_solve.hip", "fedd_mlp.hip", "fedd_mg.hip", "fedd_hdf5.hip", "fedd_nsolve.hip",
           "fedd_surface.hip"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-munsafe-fp-atomics",
    # packed FP32 instructions disabled (see the project notes on CDNA4 builds)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-missing-braces", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result",
]


def hipcc():
    for p in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(p) or p == "hipcc":
            return p
    raise RuntimeError("hipcc not found")


def build_lib(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = []
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    procs = []
    for s in srcs:
        o = os.path.join(HERE, "_build", os.path.basename(s) + ".o")
        objs.append(o)
        cmd = [hipcc()] + HIPCC_FLAGS + ["-c", s, "-o", o]
        if verbose:
            print(" ".join(cmd))
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out.decode())
            raise RuntimeError("hipcc failed: " + " ".join(cmd))
        elif verbose and out:
            sys.stderr.write(out.decode()[-4000:])
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    oracle = os.path.join(ROOT, "oracle")
    subprocess.check_call(["make", "-s", "-C", oracle, "oracle"])
    if os.path.isdir("/root/reference"):
        subprocess.check_call(["make", "-s", "-C", oracle, "ref"])


def main(argv):
    build_lib(verbose="-v" in argv)
    if "--no-oracle" not in argv:
        build_oracle()


if __name__ == "__main__":
    main(sys.argv[1:])
