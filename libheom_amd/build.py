"""In-tree build of the HIP library (libheom_amd/_lib/libheom_amd.so, gfx950)."""
import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs=8, quiet=True):
    """Compile every HIP/C++ source of the engine with hipcc for gfx950."""
    out = subprocess.DEVNULL if quiet else None
    subprocess.run(["make", "-C", os.path.join(_HERE, "csrc"), "-j%d" % jobs, "all"],
                   check=True, stdout=out)
    return os.path.join(_HERE, "_lib", "libheom_amd.so")


if __name__ == "__main__":
    print(build(quiet=False))
