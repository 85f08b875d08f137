"""In-tree build of libpyranda_amd.so for gfx950 (hipcc, no JIT cache)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRCS = [os.path.join(HERE, "csrc", "pyr_core.hip"), os.path.join(HERE, "csrc", "pyr_mlp.hip"),
        os.path.join(HERE, "csrc", "pyr_fuse.cpp")]
# compile units: pyr_core.hip builds as four parts (host code and C-ABI; the
# line kernels of each solve mode), compiled in parallel
UNITS = [(SRCS[0], ["PYR_PART=%d" % k]) for k in range(4)] + [(SRCS[1], []), (SRCS[2], [])]
OUT = os.path.join(HERE, "libpyranda_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
    # keep packed-FP32 instructions out of the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # multiply-adds fused only within one source expression: the same source
    # then rounds the same way in every kernel it is inlined into (the
    # default, fast, lets the backend fuse across statements depending on
    # the surrounding code, so two launch shapes of one operator could differ
    # in the last bit)
    "-ffp-contract=on",
]


def build(verbose=False, out=OUT, defines=()):
    """Compile the library; `defines` (e.g. ["PYR_PREFETCH=0"]) set the tuning
    knobs of pyr_core.hip for a variant written to `out`. The sources compile
    to objects in parallel, then link."""
    tmp = out + ".tmp"
    defs = ["-D" + d for d in defines]
    objs = [out + ".%d.o" % i for i in range(len(UNITS))]
    flags = [f for f in FLAGS if f != "-shared"]
    procs = [subprocess.Popen([HIPCC] + flags + defs + ["-D" + d for d in udefs] + ["-c", "-o", o, src],
                              stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
             for (src, udefs), o in zip(UNITS, objs)]
    errs = []
    for p in procs:
        msg, _ = p.communicate()
        if p.returncode != 0:
            errs.append(msg)
    if errs:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errs))
    cmd = [HIPCC] + FLAGS + ["-o", tmp] + objs + ["-lhiprtc"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    for o in objs:
        os.remove(o)
    if r.returncode != 0:
        raise RuntimeError("hipcc link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build(verbose=True)
