"""Build libwbmi.so (gfx950) in-tree from csrc/*.hip with hipcc.

    python -m wannierberri_amd.build_lib        (or call build())
"""
from concurrent.futures import ThreadPoolExecutor
import glob
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libwbmi.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions are kept out of all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wno-unused-result",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libwbmi.so")


def build(verbose=False, jobs=None):
    """Compile each source to an object (in parallel, up to `jobs` at once), then link."""
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    inc = "-I" + os.path.join(HERE, "..", "include")
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    jobs = jobs or min(len(srcs), os.cpu_count() or 1, 8)
    with tempfile.TemporaryDirectory(prefix="wbmi_build_") as tmpdir:
        objs = [os.path.join(tmpdir, os.path.basename(s)[:-4] + ".o") for s in srcs]
        cmds = [[hipcc] + compile_flags + [inc, "-c", s, "-o", o] for s, o in zip(srcs, objs)]
        with ThreadPoolExecutor(max_workers=jobs) as ex:
            for f in [ex.submit(_run, c, verbose) for c in cmds]:
                f.result()
        tmp = LIB + ".tmp"
        _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", tmp], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
