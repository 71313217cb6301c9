"""ctypes binding of libcpet_hip.so (declared in include/cpet_hip.h).

The library is loaded once per process, the way the reference loads its
native module (CPET/utils/calculator.py:18-29 -> CPET/utils/c_ops.py:8-12).
There is no CPU fallback: if the library is missing or no GPU is visible the
calls raise, so a GPU run can never silently pass on another code path.
"""
from __future__ import annotations

import ctypes
import os
import threading
from pathlib import Path

# CPET_HIP_LIB selects another build of the library (for A/B timing).
LIB_PATH = Path(os.environ.get("CPET_HIP_LIB") or
                Path(__file__).resolve().with_name("libcpet_hip.so"))

# status codes (include/cpet_hip.h)
CPET_OK = 0
CPET_ERR_ARG = -1
CPET_ERR_HIP = -2
CPET_ERR_NODEV = -3
CPET_ERR_NOMEM = -4
CPET_ERR_BUDGET = -5

# flags
F_DEVICE = 0x1
F_OUT_F64 = 0x2
F_NO_SYNC = 0x4
F_DIPOLE_PHYSICAL = 0x10
F_TOPO_RK4 = 0x20
F_COUNTS = 0x40
PREC = {"auto": 0 << 8, "naive": 1 << 8, "kahan": 2 << 8, "f64": 3 << 8,
        "tiled": 4 << 8, "mixed": 5 << 8}


class CpetError(RuntimeError):
    """A libcpet_hip call returned a non-zero status."""

    def __init__(self, status: int, message: str):
        super().__init__(f"libcpet_hip error {status}: {message}")
        self.status = status


class CpetNotBuilt(RuntimeError):
    """libcpet_hip.so is missing; build it with `python -m pycpet_amd.build`."""


_lock = threading.RLock()
_lib = None

c_int = ctypes.c_int
c_float = ctypes.c_float
c_uint32 = ctypes.c_uint32
c_void_p = ctypes.c_void_p
P = ctypes.c_void_p  # array arguments are passed as raw addresses

# name -> argtypes (restype is int for all functions but cpet_last_error)
SIGNATURES = {
    # reference-compatible (math_module.c names)
    "compute_looped_field": [c_int, c_int, P, P, P, P],
    "calc_esp_base": [P, P, c_int, P, P],
    "calc_field": [P, P, c_int, P, P],
    "calc_field_base": [P, P, c_int, P, P],
    "thread_operation": [c_int, c_int, c_float, P, P, P, P, P],
    "thread_operation_dipole": [c_int, c_int, c_float, P, P, P, P, P],
    "compute_batched_field": [c_int, c_int, c_int, P, P, P, P],
    # host helpers of the reference binding (no GPU)
    "sparse_dot": [P, P, c_int, P, c_int, P, c_int, P, c_int],
    "dot": [P, P, P, c_int, c_int],
    "vecaddn": [P, P, P, c_int],
    "einsum_operation_batch": [c_int, c_int, P, P, P, P],
    "einsum_operation": [c_int, P, P, P, P],
    "einsum_ij_i_batch": [c_int, c_int, c_int, P, P],
    "einsum_ij_i": [c_int, c_int, P, P],
    # batched / device-resident
    "cpet_device_count": [P],
    "cpet_ctx_create": [c_int, P],
    "cpet_ctx_destroy": [c_void_p],
    "cpet_ctx_set_stream": [c_void_p, c_void_p],
    "cpet_ctx_get_stream": [c_void_p, P],
    "cpet_ctx_use_own_stream": [c_void_p],
    "cpet_ctx_sync": [c_void_p],
    "cpet_field_points": [c_void_p, c_int, c_int, P, P, P, P, P, c_int, c_uint32],
    "cpet_field_grid": [c_void_p, c_int, c_int, c_int, P, P, P, c_int, P, P, P, P, c_int, c_uint32],
    "cpet_topo": [c_void_p, c_int, c_int, P, P, c_float, P, P, P, P, P, P,
                  c_int, P, c_int, P, P, c_uint32],
    "cpet_topo_noalloc": [c_void_p, c_int, c_int, P, P, c_float, P, P, P, P, P, P,
                          c_int, P, c_int, P, P, c_uint32],
    "cpet_set_tuning": [c_void_p, c_int, c_int, c_int],
    "cpet_topo_dipole": [c_void_p, c_int, c_int, P, P, c_float, P, P, P, P, P, P, c_uint32],
    "cpet_ctx_set_profiling": [c_void_p, c_int],
    "cpet_ctx_phase_times": [c_void_p, P, c_int],
    "cpet_ctx_refined_points": [c_void_p, P],
    "cpet_cluster_histograms": [c_void_p, c_int, P, P, c_int, P, c_int, P, P, c_uint32],
    "cpet_chi2_matrix": [c_void_p, c_int, c_int, P, P, c_uint32],
    "cpet_order_stats": [c_void_p, ctypes.c_int64, c_int, P, c_int, P, P, P, P, c_uint32],
    "cpet_read_rows": [ctypes.c_char_p, c_int, P, ctypes.c_int64, P],
    "cpet_pdb_scan": [ctypes.c_char_p, P],
    "cpet_pdb_parse": [ctypes.c_char_p, c_int, c_int, P, P, P, P, P, P, P],
    "cpet_pqr_parse": [ctypes.c_char_p, c_int, c_int, P, P, P, P, P, P],
    "cpet_write_rows": [ctypes.c_char_p, ctypes.c_char_p, P, c_int, ctypes.c_int64, c_int, c_int,
                        c_int],
    "cpet_batch_volume": [c_void_p, c_int, P, P, P, P, P, P, P, c_int, c_uint32],
    "cpet_batch_topo": [c_void_p, c_int, P, P, P, P, P, P, P, P, P, P, P, c_int, P, c_int, P,
                        P, c_uint32],
    "cpet_filter_radius_residue": [c_void_p, c_int, P, c_int, P, c_int, ctypes.c_double, P,
                                   c_int, P, c_uint32],
    "cpet_frame_transform": [c_void_p, c_int, P, c_int, P, c_int, P, c_int, P, c_int, P,
                             c_int, c_uint32],
    "cpet_filter_in_box": [c_void_p, c_int, P, c_int, P, c_int, P, P, c_uint32],
    "cpet_compact_charges": [c_void_p, c_int, P, P, P, P, P, P, P, c_uint32],
}


def _share_torch_hip_runtime() -> None:
    """Makes torch's bundled HIP runtime the process's HIP runtime.

    torch's libraries ask for `libamdhip64.so` from torch/lib, while this
    library asks for the `libamdhip64.so.7` soname.  If this library were
    loaded first, importing torch later would map a second HIP runtime into
    the process (its device enumeration then fails).  Loading torch's copy
    first with RTLD_GLOBAL makes both resolve to that one runtime, so device
    pointers from torch tensors are valid here.  Without torch installed the
    system ROCm runtime is used.
    """
    import importlib.util

    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.submodule_search_locations:
        return
    for loc in spec.submodule_search_locations:
        rt = Path(loc) / "lib" / "libamdhip64.so"
        if rt.exists():
            ctypes.CDLL(str(rt), mode=ctypes.RTLD_GLOBAL)
            return


def load(path: str | os.PathLike | None = None):
    """Loads (once) and returns the library with argtypes declared."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        _share_torch_hip_runtime()
        p = Path(path) if path is not None else LIB_PATH
        if not p.exists():
            raise CpetNotBuilt(
                f"{p} not found: build the HIP library first "
                "(python -m pycpet_amd.build or __graft_entry__.build())")
        lib = ctypes.CDLL(str(p))
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = c_int
        lib.cpet_last_error.argtypes = []
        lib.cpet_last_error.restype = ctypes.c_char_p
        if path is None:
            _lib = lib
        return lib


def last_error(lib=None) -> str:
    lib = lib or load()
    msg = lib.cpet_last_error()
    return msg.decode() if msg else ""


def check(rc: int, lib=None) -> None:
    if rc != CPET_OK:
        raise CpetError(rc, last_error(lib))


def device_count() -> int:
    lib = load()
    n = ctypes.c_int(0)
    check(lib.cpet_device_count(ctypes.byref(n)), lib)
    return n.value


class Context:
    """Owns a cpet_ctx (device, stream, scratch).  Not shared across threads
    without the library's own per-context lock, which serialises calls."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = ctypes.c_void_p()
        check(self.lib.cpet_ctx_create(device, ctypes.byref(h)), self.lib)
        self.handle = h
        self.device = device

    def set_stream(self, stream_handle: int | None) -> None:
        """Queue work on this HIP stream handle; 0 / None is the null stream
        (torch's default stream)."""
        check(self.lib.cpet_ctx_set_stream(self.handle, stream_handle or None), self.lib)

    def use_own_stream(self) -> None:
        check(self.lib.cpet_ctx_use_own_stream(self.handle), self.lib)

    def stream(self) -> int:
        s = ctypes.c_void_p()
        check(self.lib.cpet_ctx_get_stream(self.handle, ctypes.byref(s)), self.lib)
        return s.value or 0

    def sync(self) -> None:
        check(self.lib.cpet_ctx_sync(self.handle), self.lib)

    def set_tuning(self, kind: int, a: int, b: int) -> None:
        check(self.lib.cpet_set_tuning(self.handle, kind, a, b), self.lib)

    def set_profiling(self, on: bool) -> None:
        check(self.lib.cpet_ctx_set_profiling(self.handle, 1 if on else 0), self.lib)

    def phase_times(self, n: int = 3):
        """Durations (ms) of the last call's phases, measured with HIP events
        on the context's stream (profiling must be on)."""
        buf = (ctypes.c_float * n)()
        check(self.lib.cpet_ctx_phase_times(self.handle, buf, n), self.lib)
        return list(buf)

    def refined_points(self) -> int:
        """Points the last auto-precision field call recomputed in FP64."""
        n = ctypes.c_longlong(0)
        check(self.lib.cpet_ctx_refined_points(self.handle, ctypes.byref(n)), self.lib)
        return n.value

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.cpet_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    with _lock:
        if _default_ctx is None:
            dev = int(os.environ.get("CPET_DEVICE", "0"))
            _default_ctx = Context(dev)
        return _default_ctx
