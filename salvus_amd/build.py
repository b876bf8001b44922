"""Build the HIP library for gfx950 in-tree (salvus_amd/_lib/libsalvus_sem.so)."""
import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent


def build(jobs=None, force=True):
    jobs = jobs or min(16, os.cpu_count() or 4)
    csrc = HERE / "csrc"
    if force:
        # always recompile (the build check must compile every time)
        subprocess.run(["make", "-s", "-C", str(csrc), "clean"], check=True)
    subprocess.run(["make", "-s", "-C", str(csrc), "-j%d" % jobs], check=True)
    return HERE / "_lib" / "libsalvus_sem.so"


if __name__ == "__main__":
    print(build())
