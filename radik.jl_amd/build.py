"""Builds libradik_hip.so for gfx950 in-tree (radik.jl_amd/lib/)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def build(force=True):
    args = ["make", "-C", HERE, "-j8"]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True)
    return os.path.join(HERE, "lib", "libradik_hip.so")


if __name__ == "__main__":
    print(build(force="--incremental" not in sys.argv))
