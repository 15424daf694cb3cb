"""Build libcapytaine_amd.so (the HIP kernels and the C-ABI) in place.

    python -m capytaine_amd.build_lib        # or build() from Python

hipcc cross-compiles for gfx950; no GPU is needed to build.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcapytaine_amd.so")
SOURCES = ["api.hip", "assembly.hip", "tabulation.hip", "lu.hip", "symmetry.hip", "gmres.hip", "fingreen3d.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # Packed-FP32 instructions are disabled: some of their forms have given
    # wrong results on this hardware.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
    "-Wno-unused-command-line-argument",
]


def build(verbose=False):
    """Compile each translation unit in parallel, then link."""
    import concurrent.futures as cf
    import tempfile
    objdir = tempfile.mkdtemp(prefix="cpt_build_")
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC] + [f for f in FLAGS if f != "-shared"] + ["-c", os.path.join(CSRC, src), "-o", obj])

    def run(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        return subprocess.run(cmd, capture_output=True, text=True)

    with cf.ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        results = list(ex.map(run, cmds))
    for res in results:
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError("hipcc failed building libcapytaine_amd.so")
    tmp = LIB + ".tmp"
    res = run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", tmp])
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed linking libcapytaine_amd.so")
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    os.rmdir(objdir)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
