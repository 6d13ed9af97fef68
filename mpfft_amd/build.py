"""Build libmpfft.so (HIP kernels for gfx950 + C++ host library) in-tree.

    python -m mpfft_amd.build            # compile everything that changed
    python -m mpfft_amd.build --force    # recompile everything

Objects go to mpfft_amd/_build/, the library to mpfft_amd/libmpfft.so.
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmpfft.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
MPFR_INC = os.environ.get("MPFFT_MPFR_INC", "/opt/conda/include")
MPLIBS = ["/lib/x86_64-linux-gnu/libmpfr.so.6", "/lib/x86_64-linux-gnu/libgmp.so.10"]

# -packed-fp32-ops: keep hipcc from forming packed FP32 instructions
# -ffp-contract=off: every double product and sum rounds on its own, as in
#                    the reference's gcc -O3 x86-64 build
COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          f"--offload-arch={ARCH}", f"-I{MPFR_INC}", f"-I{os.path.join(ROOT, 'include')}"]

SOURCES = ["mp_fft.hip", "fx_fft.hip", "fx_panel.hip", "dd_fft.hip", "d_fft.hip", "mpfft.cpp",
           "batch.cpp",
           "hexio.cpp"]
HEADERS = ["mp_arith.h", "fx_arith.h", "fx_lane.h", "kernels.h", "dd_conv.h"]


def _stale(obj, deps):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in deps)


def _compile(src, force):
    obj = os.path.join(BUILD, src + ".o")
    deps = [os.path.join(CSRC, src)] + [os.path.join(CSRC, h) for h in HEADERS] + \
        [os.path.join(ROOT, "include", "mpfft.h")]
    if not force and not _stale(obj, deps):
        return obj, None
    lang = [] if src.endswith(".hip") else ["-x", "hip"]
    cmd = [HIPCC, "-c"] + lang + COMMON + [os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return obj, f"{' '.join(cmd)}\n{r.stdout}\n{r.stderr}"
    return obj, None


def build(force=False, verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        results = list(ex.map(lambda s: _compile(s, force), SOURCES))
    errs = [e for _, e in results if e]
    if errs:
        raise RuntimeError("HIP build failed:\n" + "\n".join(errs))
    objs = [o for o, _ in results]
    if force or _stale(LIB, objs):
        cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", LIB] + objs + MPLIBS + \
            ["-Wl,-rpath,/lib/x86_64-linux-gnu"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    probe = os.path.join(HERE, "valu_probe")
    src = os.path.join(CSRC, "valu_probe.hip")
    if force or _stale(probe, [src]):
        cmd = [HIPCC, "-O3", f"--offload-arch={ARCH}", "-Xclang", "-target-feature",
               "-Xclang", "-packed-fp32-ops", src, "-o", probe]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"probe build failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_oracle(verbose=True):
    """Compile the oracle (test infrastructure) and, where the reference tree
    is present, the reference libraries under oracle/_ref."""
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-s"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print("built oracle")


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    a = ap.parse_args()
    build(force=a.force)
    build_oracle()
    sys.exit(0)
