"""Build recipe for libnqm.so (gfx950, wave64) -- hipcc, in-tree.

    python gpu-n-queens_amd/build.py        # -> gpu-n-queens_amd/nqm/libnqm.so

Every HIP source under csrc/ is compiled for gfx950 only. Packed FP32 is
switched off for all device code (-packed-fp32-ops); the host pass warns
that it does not know the feature, which is harmless.
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "nqm", "libnqm.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
    "-I", CSRC,
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def build(verbose=False, extra=None):
    """Compile every csrc/*.hip and link libnqm.so. Returns the .so path."""
    objs = []
    tmp = tempfile.mkdtemp(prefix="nqm_build_")
    try:
        for src in sources():
            obj = os.path.join(tmp, os.path.basename(src) + ".o")
            cmd = [HIPCC, *COMMON, *(extra or []), "-c", src, "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            subprocess.run(cmd, check=True, stderr=None if verbose else subprocess.PIPE)
            objs.append(obj)
        tmp_out = OUT + ".tmp"
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp_out, *objs,
               "-lpthread"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(tmp_out, OUT)
    except subprocess.CalledProcessError as e:
        if e.stderr:
            sys.stderr.write(e.stderr.decode(errors="replace"))
        raise
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return OUT


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(OUT)
