"""Build and load libchfft_hip.so, the HIP (gfx950) engine behind the C ABI in
include/chfft_hip.h. There is no CPU fallback: if the library is missing or the
GPU is absent, calls fail loudly."""
import ctypes
import os
import subprocess

_PKG = os.path.dirname(os.path.abspath(__file__))
_ROOT = os.path.dirname(_PKG)
LIB_PATH = os.path.join(_PKG, "libchfft_hip.so")
SOURCES = ["csrc/chfft_kernels.hip", "csrc/chfft_sfft.hip", "csrc/chfft_multi.hip", "csrc/chfft_tables.hip",
           "csrc/chfft_plan.hip"]
COMPILE_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # no packed FP32 instructions (v_pk_*_f32) in any kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

OK, ERR_INVALID_LENGTH, ERR_INVALID_ARGUMENT, ERR_UNSUPPORTED, ERR_HIP, ERR_OOM, ERR_INCOMPLETE, \
    ERR_SCRATCH_LIMIT = range(8)
NOALLOC_SCRATCH_BYTES = 96 << 20
F32, F64 = 0, 1
METHODS = {
    "forward": 0, "forward0": 1, "forwardu": 2, "forwardn": 3,
    "backward": 4, "backward0": 5, "backwardu": 6, "backwardn": 7,
}
DCT_II, DCT_III = 2, 3

_LIB = None


def build(verbose=False):
    """Compile the HIP sources into LIB_PATH (cross-compiles without a GPU):
    one hipcc process per source file in parallel, then one link."""
    import concurrent.futures
    import tempfile
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    with tempfile.TemporaryDirectory(prefix="chfft_build_") as tmp:
        def compile_one(src):
            obj = os.path.join(tmp, os.path.basename(src) + ".o")
            cmd = [hipcc] + COMPILE_FLAGS + ["-c", os.path.join(_PKG, src), "-o", obj]
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"hipcc failed on {src}:\n" + r.stderr[-4000:])
            return obj
        with concurrent.futures.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
            objs = list(ex.map(compile_one, SOURCES))
        r = subprocess.run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", LIB_PATH + ".tmp"],
                           capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("hipcc link failed:\n" + r.stderr[-4000:])
    os.replace(LIB_PATH + ".tmp", LIB_PATH)
    if verbose:
        print("built", LIB_PATH)
    return LIB_PATH


# Every symbol include/chfft_hip.h declares, with its ctypes signature.
def _signatures():
    vp, sz, ci = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
    pp = ctypes.POINTER(ctypes.c_void_p)
    return {
        "chfft_plan_c2c_1d": ([pp, sz, sz, ci], ci),
        "chfft_plan_r2c_1d": ([pp, sz, sz, ci], ci),
        "chfft_plan_c2c_2d": ([pp, sz, sz, sz, ci], ci),
        "chfft_plan_dct_1d": ([pp, ci, sz, ci, sz, ci], ci),
        "chfft_plan_mdct_1d": ([pp, sz, vp, sz, sz, ci], ci),
        "chfft_plan_c2c_multi": ([pp, vp, sz, ci], ci),
        "chfft_plan_set_grid": ([vp, ci], ci),
        "chfft_plan_check": ([vp, vp], ci),
        "chfft_exec_mdct_ola": ([vp, vp, vp, ci, vp], ci),
        "chfft_plan_autotune": ([vp, vp, vp, ci, vp], ci),
        "chfft_plan_tune": ([vp, ctypes.c_char_p], ci),
        "chfft_plan_tuning": ([vp, ctypes.c_char_p, ctypes.c_size_t], ci),
        "chfft_plan_destroy": ([vp], ci),
        "chfft_exec": ([vp, vp, vp, ci, vp], ci),
        "chfft_exec_noalloc": ([vp, vp, vp, ci, vp], ci),
        "chfft_exec_host": ([vp, vp, vp, ci], ci),
        "chfft_exec_range": ([vp, vp, vp, ci, ci, ci, vp], ci),
        "chfft_plan_launch_count": ([vp, ci], ci),
        "chfft_plan_io_bytes": ([vp, ci, ctypes.POINTER(sz), ctypes.POINTER(sz)], ci),
        "chfft_plan_scratch_bytes": ([vp], sz),
        "chfft_plan_describe": ([vp, ctypes.c_char_p, sz], ci),
        "chfft_status_string": ([ci], ctypes.c_char_p),
    }


def lib():
    """Load the built library (does not build it: build() first)."""
    global _LIB
    if _LIB is None:
        # PyTorch-ROCm ships its own HIP runtime (torch/lib/libamdhip64.so,
        # SONAME libamdhip64.so.7). Importing torch first makes this library's
        # libamdhip64.so.7 dependency bind to that same copy, so that one
        # process never holds two HIP runtimes (the second one to initialise
        # would find no GPU).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP extension not built: {LIB_PATH} is missing (run chfft_amd.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _signatures().items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _LIB = L
    return _LIB


class ChfftError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        msg = lib().chfft_status_string(status).decode()
        super().__init__(f"{what}: {msg}" if what else msg)


def check(status, what=""):
    if status != OK:
        if status == ERR_INVALID_LENGTH:
            raise ValueError("invalid length")
        raise ChfftError(status, what)
