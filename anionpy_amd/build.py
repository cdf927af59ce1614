"""Build libionp_hip.so (gfx950) in-tree with hipcc.

    python -m anionpy_amd.build        # or: from anionpy_amd.build import build

Each translation unit under csrc/ is compiled to an object in
anionpy_amd/_build/, then linked into anionpy_amd/libionp_hip.so.  Flags:
  -ffp-contract=off     every IEEE op rounds once, as NumPy's loops do
  -packed-fp32-ops off  keep fp32 math in scalar/VOP3 form
  -mcode-object-version=5
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libionp_hip.so")
ARCH = os.environ.get("IONP_OFFLOAD_ARCH", "gfx950")

SOURCES = ["planner.cpp", "api.cpp", "elementwise.hip", "rows_select_cast.hip", "tiled_any.hip", "reduce.hip", "accumulate.hip", "batch.hip", "mlp.hip",
           "features.cpp", "advisor.cpp"]

FLAGS = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-fno-fast-math", "-mcode-object-version=5",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    return "hipcc"


def _stale(src, obj):
    """True when obj is missing or older than its source or any header."""
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [os.path.join(CSRC, src)] + [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".h")]
    deps.append(os.path.join(os.path.dirname(HERE), "include", "ionp_hip.h"))
    return any(os.path.getmtime(d) > t for d in deps if os.path.exists(d))


def _compile(src, incremental=False):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    if incremental and not _stale(src, obj):
        return obj, ""
    lang = ["-x", "hip"] if src.endswith(".cpp") else []
    cmd = [_hipcc()] + lang + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 feature; that warning is noise
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{err}")
    return obj, err


def build(verbose=False, incremental=False):
    """Compile every source (incremental: only those older than their
    object or a header) and link the library."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(lambda src: _compile(src, incremental), srcs))
    if verbose:
        for obj, err in results:
            if err:
                print(err, file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True, incremental="--incremental" in sys.argv[1:]))
