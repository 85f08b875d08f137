"""In-tree build of libpyranda_amd.so for gfx950 (hipcc, no JIT cache)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRCS = [os.path.join(HERE, "csrc", "pyr_core.hip"), os.path.join(HERE, "csrc", "pyr_mlp.hip"),
        os.path.join(HERE, "csrc", "pyr_fuse.cpp")]
OUT = os.path.join(HERE, "libpyranda_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
    # keep packed-FP32 instructions out of the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose=False, out=OUT, defines=()):
    """Compile the library; `defines` (e.g. ["PYR_PREFETCH=0"]) set the tuning
    knobs of pyr_core.hip for a variant written to `out`."""
    tmp = out + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-D" + d for d in defines] + ["-o", tmp] + SRCS + ["-lhiprtc"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build(verbose=True)
