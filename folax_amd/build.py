"""Build libfemhot.so in-tree for gfx950 (hipcc via folax_amd/csrc/Makefile)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build(force=False, jobs=8):
    csrc = os.path.join(HERE, "csrc")
    if force:
        subprocess.run(["make", "-C", csrc, "clean"], check=True, stdout=subprocess.DEVNULL)
    subprocess.run(["make", "-C", csrc, f"-j{jobs}"], check=True)
    so = os.path.join(HERE, "libfemhot.so")
    if not os.path.exists(so):
        raise RuntimeError("libfemhot.so was not produced")
    return so


if __name__ == "__main__":
    print(build())
