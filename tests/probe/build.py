"""Build the test-only probe library tests/probe/libnhprobe.so (gfx950)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build():
    out = os.path.join(HERE, "libnhprobe.so")
    subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-Xclang",
                    "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument",
                    os.path.join(HERE, "mfma_probe.hip"), "-o", out], check=True, capture_output=True)
    return out


if __name__ == "__main__":
    print(build())
