"""Builds libdwhip.so in-tree with hipcc for gfx950.

    python -m deconwolf_amd.build [-v] [-j N] [-D NAME=VALUE ...] [-o LIB]

-D passes a define to every unit (the tile parameters DWHIP_XTHR,
DWHIP_CTHR, DWHIP_ZTHR, DWHIP_XMAXE, DWHIP_CMAXE, DWHIP_YMAXE, DWHIP_ZMAXE); -o builds that
variant elsewhere (objects in a folder beside it).

Every pass-kernel mode is its own translation unit (k_xpass.hip and
k_colpass.hip compiled once per mode), compiled in parallel, then linked.
Objects go to deconwolf_amd/_build/ (not tracked).
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdwhip.so")

N_XMODES = 7
N_CMODES = 9

# -packed-fp32-ops: no packed FP32 instructions in device code.  The host
# pass of hipcc warns that it does not know the feature; that is harmless.
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-Wall", "-Wno-unused-result", "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    return "hipcc"


def units():
    """(source, extra defines, object name) for every translation unit."""
    u = []
    for m in range(N_XMODES):
        u.append(("k_xpass.hip", ["-DXMODE=%d" % m], "k_xpass_%d.o" % m))
    for m in range(N_CMODES):
        u.append(("k_colpass.hip", ["-DCMODE=%d" % m], "k_colpass_%d.o" % m))
    u.append(("k_misc.hip", [], "k_misc.o"))
    u.append(("k_output.hip", [], "k_output.o"))
    u.append(("k_bw.hip", [], "k_bw.o"))
    u.append(("noalloc.hip", [], "noalloc.o"))
    u.append(("k_generic.hip", [], "k_generic.o"))
    for src in ("dwhip.cpp", "profile.cpp", "batch.cpp", "tiling.cpp", "graph.cpp", "generic.cpp", "tune.cpp"):
        if os.path.exists(os.path.join(CSRC, src)):
            u.append((src, [], src.replace(".cpp", ".o")))
    return u


def _filter(stderr):
    return "\n".join(l for l in stderr.splitlines() if "packed-fp32-ops' is not a recognized" not in l)


def _compile(unit, verbose, extra=(), objdir=OBJ):
    src, defs, obj = unit
    cmd = [hipcc()] + FLAGS + list(extra) + defs + ["-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-c",
                                                     os.path.join(CSRC, src), "-o", os.path.join(objdir, obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msg = _filter(r.stderr)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed on %s %s:\n%s" % (src, " ".join(defs), msg))
    if verbose and msg.strip():
        print(msg)
    return obj


def build(verbose=False, jobs=None, defines=(), lib=None):
    lib = lib or LIB
    objdir = OBJ if lib == LIB else lib + ".objs"
    os.makedirs(objdir, exist_ok=True)
    extra = ["-D" + d for d in defines]
    us = units()
    jobs = jobs or min(len(us), os.cpu_count() or 4, 16)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda u: _compile(u, verbose, extra, objdir), us))
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", lib + ".tmp"] + \
          [os.path.join(objdir, o) for o in objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + _filter(r.stderr))
    os.replace(lib + ".tmp", lib)
    return lib


if __name__ == "__main__":
    j = None
    if "-j" in sys.argv:
        j = int(sys.argv[sys.argv.index("-j") + 1])
    defs = [sys.argv[i + 1] for i, a in enumerate(sys.argv) if a == "-D"]
    out = sys.argv[sys.argv.index("-o") + 1] if "-o" in sys.argv else None
    print(build(verbose="-v" in sys.argv, jobs=j, defines=defs, lib=out))
