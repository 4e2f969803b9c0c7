"""Build the C restatement (oracle/nh_cpu.c) as libnhcpu32.so (the CPU baseline) and
libnhcpu64.so (FP64 oracle) next to this file.  Test/baseline infrastructure only."""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "nh_cpu.c")
# -mavx2 -mfma: vector ISA present on both the build container and the GPU hosts
FLAGS = ["-O3", "-mavx2", "-mfma", "-fopenmp", "-fPIC", "-shared", "-std=c11"]


def build():
    out = []
    for real, name in (("float", "libnhcpu32.so"), ("double", "libnhcpu64.so")):
        dst = os.path.join(HERE, name)
        cmd = ["gcc", *FLAGS, f"-DREAL={real}", SRC, "-o", dst + ".tmp", "-lm"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("gcc failed on oracle/nh_cpu.c")
        os.replace(dst + ".tmp", dst)
        out.append(dst)
    return out


if __name__ == "__main__":
    print("\n".join(build()))
