"""The library's HIP kernels run on the CPU under the host emulation of tests/emu (test support).

tests/emu compiles the same .hip sources for x86 against an emulated HIP runtime: work-items are
fibers of one host thread per work-group, wavefront operations (shuffles, DPP, lane reads, MFMA)
gather the 64 lanes and apply the instruction's lane layout.  `emulated()` swaps the ctypes
binding of qudos_amd to that library for the duration of a test, so the GPU tests' own code
paths run unchanged at small sizes.  The AddressSanitizer variant runs in a child process
(`run_sanitized`), since the sanitizer runtime has to be loaded before the interpreter starts.
Neither library is ever loaded by the product: qudos_amd.load_library() loads only the
hipcc-built libqdgpu.so.
"""
import contextlib
import ctypes
import glob
import os
import subprocess
import sys

import qudos_amd._lib as qlib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU_DIR = os.path.join(ROOT, "tests", "emu")
EMU_LIB = os.path.join(EMU_DIR, "lib", "libqdgpu_emu.so")
ASAN_LIB = os.path.join(EMU_DIR, "lib", "libqdgpu_emu_asan.so")
DRIVER = os.path.join(EMU_DIR, "lib", "qd_pwem_emu")   # the qd_pwem driver on the emulated library


def build(target="all"):
    subprocess.run(["make", "-s", "-j8", "-C", EMU_DIR, target], check=True)


def _stale(lib):
    if not os.path.exists(lib):
        return True
    t = os.path.getmtime(lib)
    srcs = (glob.glob(os.path.join(ROOT, "qudos_amd", "csrc", "*", "*")) + glob.glob(os.path.join(EMU_DIR, "*.*"))
            + glob.glob(os.path.join(EMU_DIR, "hip", "*")) + glob.glob(os.path.join(ROOT, "qudos_amd", "driver", "*"))
            + [os.path.join(ROOT, "include", "qudos_gpu.h")])
    return any(os.path.getmtime(s) > t for s in srcs)


def ensure_built(lib=EMU_LIB):
    if _stale(lib):
        build(os.path.relpath(lib, EMU_DIR))


_EMU = None


@contextlib.contextmanager
def emulated():
    """qudos_amd calls the emulated library inside the block."""
    global _EMU
    ensure_built(EMU_LIB)
    if _EMU is None:
        _EMU = qlib.bind(ctypes.CDLL(EMU_LIB))
    saved = qlib._LIB
    qlib._LIB = _EMU
    try:
        yield _EMU
    finally:
        qlib._LIB = saved


def asan_runtime():
    out = subprocess.run(["/opt/rocm/llvm/bin/clang++", "-print-file-name=libclang_rt.asan-x86_64.so"],
                         capture_output=True, text=True, check=True).stdout.strip()
    if not os.path.exists(out):
        cands = glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so")
        out = cands[0] if cands else out
    return out


def run_sanitized(code, timeout=1800):
    """Runs `code` (Python source) in a child interpreter in which qudos_amd is bound to the
    AddressSanitizer build; the child's output is returned, and a sanitizer report fails."""
    ensure_built(ASAN_LIB)
    env = dict(os.environ)
    env["LD_PRELOAD"] = asan_runtime()
    env["ASAN_OPTIONS"] = "detect_leaks=0:detect_stack_use_after_return=0:detect_odr_violation=0:abort_on_error=1:halt_on_error=1"
    env["QD_EMU_LIB"] = ASAN_LIB
    env["PYTHONPATH"] = os.pathsep.join([ROOT, os.path.join(ROOT, "tests")] + env.get("PYTHONPATH", "").split(os.pathsep))
    prelude = ("import ctypes, os\nimport qudos_amd._lib as qlib\n"
               "qlib._LIB = qlib.bind(ctypes.CDLL(os.environ['QD_EMU_LIB']))\n")
    r = subprocess.run([sys.executable, "-c", prelude + code], env=env, capture_output=True, text=True, timeout=timeout,
                       cwd=ROOT)
    if r.returncode != 0 or "ERROR: AddressSanitizer" in r.stderr or "runtime error" in r.stderr:
        raise AssertionError(f"sanitized run failed (rc {r.returncode}):\n{r.stdout[-4000:]}\n{r.stderr[-8000:]}")
    return r.stdout
