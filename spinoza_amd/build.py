"""Build the gfx950 shared library ``spinoza_amd/libspinoza_amd.so`` in-tree.

Every ``csrc/*.hip`` file is compiled by hipcc for gfx950 only, then linked
into one shared library exporting the C ABI of ``include/spinoza_amd.h``.

Flags that matter for results:
  -ffp-contract=off   every multiply and add rounds separately (the reference's
                      Rust code never contracts); explicit FMAs are written as
                      __builtin_fma where the reference calls mul_add.
  -packed-fp32-ops    disabled: the compiler must not form packed FP32 ops.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(ROOT, "_build", "spinoza_amd")
LIB = os.path.join(HERE, "libspinoza_amd.so")

ARCH = "gfx950"
COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result", "-Wno-unused-value",
    "-I", os.path.join(ROOT, "include"),
]


# Per-file flags.  The tile kernel selects the target's register slot with a
# runtime switch over template instances; LLVM's SimplifyCFG would sink the
# near-identical case bodies into one block with a computed register index,
# which forces the amplitude arrays into scratch memory (272 B/lane, ~25x
# slower).  Keeping the cases apart keeps them in VGPRs.
# The tile kernel also asks for a 6-waves-per-SIMD register budget (see its
# source); the compiler reaches 4 without spills and warns that 6 was not met.
_TILE_FLAGS = ["-mllvm", "-simplifycfg-sink-common=false", "-Wno-pass-failed"]
EXTRA = {
    "spz_tile.hip": _TILE_FLAGS,       # exact arithmetic
    "spz_tile_fast.hip": _TILE_FLAGS,  # AUTO arithmetic (compiled in parallel)
}


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the gfx950 library cannot be built")


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith(".hip") or f.endswith(".cpp"))


# Host-only sources (.cpp: the OpenQASM loader) are plain C++ built by the
# host compiler; they hold no device code.
HOST = ["-O2", "-fPIC", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include")]


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    if src.endswith(".cpp"):
        cmd = [os.environ.get("CXX", "g++"), "-c", src, "-o", obj] + HOST
    else:
        cmd = [hipcc(), "-c", src, "-o", obj] + COMMON + EXTRA.get(os.path.basename(src), [])
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    # the host pass does not know the packed-fp32 feature; that note is expected
    noise = [ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if noise:
        sys.stderr.write("\n".join(noise) + "\n")
    return obj


def _stale(src):
    """An object is stale when it is older than its source or any header."""
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    if not os.path.exists(obj):
        return True
    deps = [src] + [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    deps += [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))]
    return any(os.path.getmtime(d) > os.path.getmtime(obj) for d in deps)


def build(verbose=False, incremental=False):
    """Compile every source (or, with incremental=True, only those whose
    object is out of date) and link the library."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    todo = [s for s in srcs if _stale(s)] if incremental else srcs
    with concurrent.futures.ThreadPoolExecutor(max_workers=max(1, min(8, len(todo)))) as ex:
        list(ex.map(_compile, todo))
    objs = [os.path.join(BUILD, os.path.basename(s) + ".o") for s in srcs]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-o", tmp] + objs + [f"--offload-arch={ARCH}", "-fPIC"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True, incremental="--incremental" in sys.argv)
