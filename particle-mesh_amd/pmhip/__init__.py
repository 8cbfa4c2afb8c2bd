"""ctypes binding of libpmhip.so (the C ABI declared in include/pm_api.h).

Device memory comes from torch (``torch.empty(..., device="cuda")``); the
library only sees raw device pointers and the current HIP stream.  Every
call goes through the HIP library: if it is not built, ``lib()`` raises.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent.parent
# PMHIP_LIB selects another build of the same library (e.g. one compiled
# with different tile or workgroup constants).
LIB_PATH = Path(os.environ.get("PMHIP_LIB", PKG_DIR / "lib" / "libpmhip.so"))

PM_OK = 0
PM_ERR_ARG = -1
PM_ERR_HIP = -2
PM_ERR_NOMEM = -3
PM_ERR_RANGE = -4

PM_FP32 = 0
PM_FP64 = 1
PM_GREEN_REF_KH = 0
PM_GREEN_CONT_K2 = 1
PM_DEPOSIT_ATOMIC = 0
PM_DEPOSIT_SORTED = 1
PM_STEP_REUSE_BINS = 0x100
PM_NPHASES = 7
PM_FFT_SLOTS = 5
PM_FFT_VARIANTS = 9

PHASE_NAMES = ["deposit", "fft_z_fwd", "fft_y_fwd", "fft_x_green", "fft_y_inv", "fft_z_inv", "interp_kick_drift"]

# Reference constants (utils_inline.h:13-18).
DA = 0.001
START_A = 0.9

_c_int = ctypes.c_int
_c_i64 = ctypes.c_int64
_c_dbl = ctypes.c_double
_c_vp = ctypes.c_void_p
_c_vpp = ctypes.POINTER(ctypes.c_void_p)

# name -> (restype, argtypes); mirrors include/pm_api.h
SIGNATURES = {
    "pm_last_error": (ctypes.c_char_p, []),
    "pm_version": (ctypes.c_char_p, []),
    "pm_init_buf": (_c_int, [_c_vp, _c_int, _c_int, _c_int, _c_vp]),
    "pm_deposit_particles": (_c_int, [_c_vp, _c_vp, _c_vp, _c_i64, _c_vp, _c_int, _c_int, _c_int,
                                      _c_int, _c_int, _c_int, _c_dbl, _c_vp]),
    "pm_reorder_slab": (_c_int, [_c_vp, _c_vp, _c_int, _c_int, _c_vp]),
    "pm_reorder_slab_back": (_c_int, [_c_vp, _c_vp, _c_int, _c_int, _c_vp]),
    "pm_transpose_slab": (_c_int, [_c_vp, _c_vp, _c_int, _c_vp]),
    "pm_transpose_slab_back": (_c_int, [_c_vp, _c_vp, _c_int, _c_vp]),
    "pm_centered_derivatives": (_c_int, [_c_vp, _c_vp, _c_vp, _c_vp, _c_int, _c_int, _c_int, _c_vp]),
    "pm_update_particles": (_c_int, [_c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp,
                                     _c_dbl, _c_i64, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int,
                                     _c_vp]),
    "pm_range_errors": (_c_i64, []),
    "pm_function_a": (_c_dbl, [_c_dbl]),
    "pm_fft_create": (_c_int, [_c_vpp, _c_int, _c_int, _c_vp]),
    "pm_fft_destroy": (_c_int, [_c_vp]),
    "pm_fft_forward": (_c_int, [_c_vp, _c_vp]),
    "pm_fft_apply_green": (_c_int, [_c_vp, _c_dbl]),
    "pm_fft_backward": (_c_int, [_c_vp, _c_vp]),
    "pm_fft_spectrum": (_c_vp, [_c_vp]),
    "pm_ic_zeldovich": (_c_int, [_c_int, _c_int, _c_vp, _c_dbl, _c_dbl, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp,
                                 _c_vp]),
    "pm_plan_create": (_c_int, [_c_vpp, _c_int, _c_i64, _c_int, _c_int, _c_dbl, _c_vp]),
    "pm_plan_destroy": (_c_int, [_c_vp]),
    "pm_plan_set_stream": (_c_int, [_c_vp, _c_vp]),
    "pm_plan_density": (_c_vp, [_c_vp]),
    "pm_plan_potential": (_c_vp, [_c_vp]),
    "pm_deposit": (_c_int, [_c_vp, _c_vp, _c_vp, _c_vp, _c_int]),
    "pm_solve": (_c_int, [_c_vp]),
    "pm_interp_kick_drift": (_c_int, [_c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_dbl, _c_dbl]),
    "pm_interp_force": (_c_int, [_c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp]),
    "pm_split_load": (_c_int, [_c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp]),
    "pm_split_store": (_c_int, [_c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp]),
    "pm_split_step": (_c_int, [_c_vp, _c_dbl, _c_dbl]),
    "pm_split_run_kdk": (_c_int, [_c_vp, _c_int, _c_dbl]),
    "pm_step": (_c_int, [_c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_dbl, _c_dbl, _c_int]),
    "pm_run_kdk": (_c_int, [_c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_int, _c_dbl, _c_int]),
    "pm_run_cosmo": (_c_int, [_c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_vp, _c_dbl, _c_dbl, _c_int, _c_int]),
    "pm_adjoint_deposit": (_c_int, [_c_vp] * 8),
    "pm_interp_transpose": (_c_int, [_c_vp] * 7),
    "pm_adjoint_solve": (_c_int, [_c_vp]),
    "pm_plan_adjoint_mesh": (_c_vp, [_c_vp, _c_int]),
    "pm_adjoint_force": (_c_int, [_c_vp] * 10 + [_c_int]),
    "pm_adjoint_step": (_c_int, [_c_vp] * 13 + [_c_dbl, _c_dbl, _c_int]),
    "pm_run_adjoint": (_c_int, [_c_vp] * 13 + [_c_int, _c_dbl, _c_int]),
    "pm_snapshot_write": (_c_int, [ctypes.c_char_p, _c_i64, _c_int, _c_int, _c_dbl] + [_c_vp] * 7),
    "pm_snapshot_info": (_c_int, [ctypes.c_char_p, ctypes.POINTER(_c_i64), ctypes.POINTER(_c_int),
                                  ctypes.POINTER(_c_int), ctypes.POINTER(_c_dbl)]),
    "pm_snapshot_read": (_c_int, [ctypes.c_char_p, _c_i64] + [_c_vp] * 7),
    "pm_plan_enable_timing": (_c_int, [_c_vp, _c_int]),
    "pm_plan_phase_ms": (_c_int, [_c_vp, ctypes.POINTER(_c_dbl), ctypes.POINTER(_c_i64)]),
    "pm_plan_reset_timing": (_c_int, [_c_vp]),
    "pm_plan_tune": (_c_int, [_c_vp, _c_int, ctypes.POINTER(_c_dbl)]),
    "pm_plan_set_fft_shape": (_c_int, [_c_vp, ctypes.POINTER(_c_int)]),
    "pm_plan_get_fft_shape": (_c_int, [_c_vp, ctypes.POINTER(_c_int)]),
    "pm_fft_tune": (_c_int, [_c_vp, _c_vp, _c_int, ctypes.POINTER(_c_dbl)]),
    "pm_fft_get_shape": (_c_int, [_c_vp, ctypes.POINTER(_c_int)]),
    "pm_wisdom_import": (_c_int, [ctypes.c_char_p]),
    "pm_wisdom_export": (_c_int, [ctypes.c_char_p]),
    "pm_wisdom_lookup": (_c_int, [_c_int, _c_int, ctypes.POINTER(_c_int), ctypes.POINTER(_c_int)]),
    "pm_wisdom_forget": (None, []),
    "pm_copy_rate": (_c_int, [ctypes.c_size_t, _c_int, ctypes.POINTER(ctypes.c_double)]),
    "pm_deposit_sorted_noalloc": (_c_int, [_c_vp, _c_vp, _c_vp, _c_i64, _c_int, _c_int, _c_vp, _c_vp]),
}

_lib = None


class PMError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"pm error {code}: {msg}")
        self.code = code


def lib() -> ctypes.CDLL:
    """Load libpmhip.so once; raise if it is missing (no fallback path)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
        handle = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def last_error() -> str:
    return lib().pm_last_error().decode()


def check(rc: int) -> None:
    if rc != PM_OK:
        raise PMError(rc, last_error())


def _torch():
    import torch

    return torch


def dptr(t) -> int:
    """Device pointer of a contiguous CUDA tensor."""
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return t.data_ptr()


def current_stream() -> int:
    return _torch().cuda.current_stream().cuda_stream


class _DevView:
    """Minimal __cuda_array_interface__ exporter for library-owned memory."""

    def __init__(self, ptr: int, shape, typestr: str):
        self.__cuda_array_interface__ = {
            "shape": tuple(shape),
            "typestr": typestr,
            "data": (int(ptr), False),
            "version": 3,
            "strides": None,
        }


def device_view(ptr: int, shape, dtype):
    """A torch tensor aliasing library-owned device memory (no copy)."""
    torch = _torch()
    typestr = {torch.float32: "<f4", torch.float64: "<f8"}[dtype]
    return torch.as_tensor(_DevView(ptr, shape, typestr), device="cuda")


def mesh_dtype(precision: int):
    torch = _torch()
    return torch.float32 if precision == PM_FP32 else torch.float64


class Plan:
    """The fused one-box PM step (pm_plan_* of include/pm_api.h)."""

    def __init__(self, ng: int, np_: int, precision: int = PM_FP32, green: int = PM_GREEN_REF_KH,
                 green_scale: float = 1.0 / START_A, stream: int | None = None):
        self.ng, self.np, self.precision = ng, np_, precision
        self._stream = current_stream() if stream is None else stream
        h = ctypes.c_void_p()
        check(lib().pm_plan_create(ctypes.byref(h), ng, np_, precision, green, green_scale, self._stream))
        self.h = h

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().pm_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def density(self):
        n = self.ng
        return device_view(lib().pm_plan_density(self.h), (n, n, n), mesh_dtype(self.precision))

    def potential(self):
        n = self.ng
        return device_view(lib().pm_plan_potential(self.h), (n, n, n), mesh_dtype(self.precision))

    def deposit(self, pos, variant: int = PM_DEPOSIT_ATOMIC):
        check(lib().pm_deposit(self.h, dptr(pos[0]), dptr(pos[1]), dptr(pos[2]), variant))

    def solve(self):
        check(lib().pm_solve(self.h))

    def interp_kick_drift(self, pos, vel, kick: float, drift: float):
        check(lib().pm_interp_kick_drift(self.h, *(dptr(t) for t in pos), *(dptr(t) for t in vel),
                                         kick, drift))

    def interp_force(self, pos, out):
        check(lib().pm_interp_force(self.h, *(dptr(t) for t in pos), *(dptr(t) for t in out)))

    def step(self, pos, vel, kick: float, drift: float, variant: int = PM_DEPOSIT_ATOMIC):
        check(lib().pm_step(self.h, *(dptr(t) for t in pos), *(dptr(t) for t in vel), kick, drift, variant))

    def run_kdk(self, pos, vel, nsteps: int, dt: float, variant: int = PM_DEPOSIT_ATOMIC):
        check(lib().pm_run_kdk(self.h, *(dptr(t) for t in pos), *(dptr(t) for t in vel), nsteps, dt, variant))

    def run_cosmo(self, pos, mom, a0: float, da: float, nsteps: int, variant: int = PM_DEPOSIT_SORTED):
        """Cosmological KDK in the expansion factor (pm_run_cosmo)."""
        check(lib().pm_run_cosmo(self.h, *(dptr(t) for t in pos), *(dptr(t) for t in mom), a0, da, nsteps, variant))

    # ------------------------------------------- split-precision state
    def split_load(self, pos, vel):
        """Copy FP64 particles into the plan's split state (pm_split_load)."""
        check(lib().pm_split_load(self.h, *(dptr(t) for t in pos), *(dptr(t) for t in vel)))

    def split_store(self, pos, vel):
        """Write the plan's split state back into FP64 arrays (pm_split_store)."""
        check(lib().pm_split_store(self.h, *(dptr(t) for t in pos), *(dptr(t) for t in vel)))

    def split_step(self, kick: float, drift: float):
        check(lib().pm_split_step(self.h, kick, drift))

    def split_run_kdk(self, nsteps: int, dt: float):
        check(lib().pm_split_run_kdk(self.h, nsteps, dt))

    # ---------------------------------------------------------- adjoint
    def adjoint_deposit(self, pos, mesh_bar, xbar):
        """xbar += m grad_x sum_c W(x, c) mesh_bar[c] (transposed CIC)."""
        check(lib().pm_adjoint_deposit(self.h, *(dptr(t) for t in pos), dptr(mesh_bar), *(dptr(t) for t in xbar)))

    def interp_transpose(self, pos, u):
        check(lib().pm_interp_transpose(self.h, *(dptr(t) for t in pos), *(dptr(t) for t in u)))

    def adjoint_solve(self):
        check(lib().pm_adjoint_solve(self.h))

    def adjoint_mesh(self, which: int):
        n = self.ng
        ptr = lib().pm_plan_adjoint_mesh(self.h, which)
        if not ptr:
            raise PMError(PM_ERR_ARG, "adjoint meshes not allocated")
        return device_view(ptr, (n, n, n), mesh_dtype(self.precision))

    def adjoint_force(self, pos, u, xbar, variant: int = PM_DEPOSIT_SORTED):
        check(lib().pm_adjoint_force(self.h, *(dptr(t) for t in pos), *(dptr(t) for t in u),
                                     *(dptr(t) for t in xbar), variant))

    def adjoint_step(self, pos, vel, xbar, vbar, kick: float, drift: float, variant: int = PM_DEPOSIT_SORTED):
        check(lib().pm_adjoint_step(self.h, *(dptr(t) for t in pos), *(dptr(t) for t in vel),
                                    *(dptr(t) for t in xbar), *(dptr(t) for t in vbar), kick, drift, variant))

    def run_adjoint(self, pos, vel, xbar, vbar, nsteps: int, dt: float, variant: int = PM_DEPOSIT_SORTED):
        check(lib().pm_run_adjoint(self.h, *(dptr(t) for t in pos), *(dptr(t) for t in vel),
                                   *(dptr(t) for t in xbar), *(dptr(t) for t in vbar), nsteps, dt, variant))

    def set_stream(self, stream: int):
        """Run the plan's later calls on another HIP stream (pm_plan_set_stream)."""
        self._stream = stream
        check(lib().pm_plan_set_stream(self.h, stream))

    def enable_timing(self, max_steps: int):
        """Record phase events for the next max_steps steps (0 = off)."""
        check(lib().pm_plan_enable_timing(self.h, max_steps))

    def reset_timing(self):
        check(lib().pm_plan_reset_timing(self.h))

    def phase_ms(self):
        ms = (ctypes.c_double * PM_NPHASES)()
        cnt = (ctypes.c_int64 * PM_NPHASES)()
        check(lib().pm_plan_phase_ms(self.h, ms, cnt))
        return list(ms), list(cnt)

    # ------------------------------------------------- FFT shapes / tuning
    def tune(self, repeats: int = 10) -> float:
        """Time every FFT pass shape on this plan (overwrites its density,
        potential and spectra), keep the fastest and cache it (pm_plan_tune).
        Returns the summed pass milliseconds of the chosen shape."""
        ms = ctypes.c_double()
        check(lib().pm_plan_tune(self.h, repeats, ctypes.byref(ms)))
        return ms.value

    @property
    def fft_shape(self) -> list[int]:
        v = (ctypes.c_int * PM_FFT_SLOTS)()
        check(lib().pm_plan_get_fft_shape(self.h, v))
        return list(v)

    @fft_shape.setter
    def fft_shape(self, variants) -> None:
        v = (ctypes.c_int * PM_FFT_SLOTS)(*variants)
        check(lib().pm_plan_set_fft_shape(self.h, v))


class FFT:
    """The reference's FFT class at one box (pm_fft_* of include/pm_api.h)."""

    def __init__(self, ng: int, precision: int = PM_FP64, stream: int | None = None):
        self.ng, self.precision = ng, precision
        h = ctypes.c_void_p()
        check(lib().pm_fft_create(ctypes.byref(h), ng, precision, current_stream() if stream is None else stream))
        self.h = h

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().pm_fft_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, slab):
        check(lib().pm_fft_forward(self.h, dptr(slab)))

    def apply_green(self, a: float):
        check(lib().pm_fft_apply_green(self.h, a))

    def backward(self, slab):
        check(lib().pm_fft_backward(self.h, dptr(slab)))

    def tune(self, slab, repeats: int = 10) -> float:
        """pm_fft_tune: slab (ng^3 reals, overwritten) is the scratch mesh."""
        ms = ctypes.c_double()
        check(lib().pm_fft_tune(self.h, dptr(slab), repeats, ctypes.byref(ms)))
        return ms.value

    @property
    def shape(self) -> list[int]:
        v = (ctypes.c_int * PM_FFT_SLOTS)()
        check(lib().pm_fft_get_shape(self.h, v))
        return list(v)

    def spectrum(self):
        """Complex view [kz][ky][kx] (kz in [0, ng/2]) of the handle's spectrum."""
        torch = _torch()
        n = self.ng
        v = device_view(lib().pm_fft_spectrum(self.h), ((n // 2 + 1), n, n, 2), mesh_dtype(self.precision))
        return torch.view_as_complex(v)


def zeldovich_ic(nrow: int, ng: int, noise, rms_cells: float, vel_factor: float, stream: int | None = None):
    """Zel'dovich lattice ICs on the GPU (pm_ic_zeldovich); noise is an
    nrow^3 float64 device tensor.  Returns (pos, vel) lists of device tensors."""
    torch = _torch()
    n = nrow ** 3
    pos = [torch.empty(n, dtype=torch.float64, device=noise.device) for _ in range(3)]
    vel = [torch.empty(n, dtype=torch.float64, device=noise.device) for _ in range(3)]
    check(lib().pm_ic_zeldovich(nrow, ng, dptr(noise), rms_cells, vel_factor, *(dptr(t) for t in pos),
                                *(dptr(t) for t in vel), current_stream() if stream is None else stream))
    return pos, vel


def range_errors() -> int:
    n = lib().pm_range_errors()
    if n < 0:
        check(int(n))
    return int(n)


# ------------------------------------------------------------- snapshots
SNAP_MAGIC = b"PMSNAP01"
SNAP_HEADER = 64


def write_snapshot(path, pos, vel, ng: int, step: int = 0, a: float = 1.0, stream: int | None = None):
    """pm_snapshot_write of (3, np) device arrays (x, y, z and vx, vy, vz)."""
    np_ = pos[0].numel()
    st = current_stream() if stream is None else stream
    check(lib().pm_snapshot_write(str(path).encode(), np_, ng, step, a, *(dptr(t) for t in pos),
                                  *(dptr(t) for t in vel), st))


def snapshot_info(path):
    """(np, ng, step, a) from a snapshot header."""
    n, g, s, a = _c_i64(), _c_int(), _c_int(), _c_dbl()
    check(lib().pm_snapshot_info(str(path).encode(), ctypes.byref(n), ctypes.byref(g), ctypes.byref(s),
                                 ctypes.byref(a)))
    return n.value, g.value, s.value, a.value


def read_snapshot(path, pos, vel, stream: int | None = None):
    """pm_snapshot_read into (3, np) device arrays."""
    st = current_stream() if stream is None else stream
    check(lib().pm_snapshot_read(str(path).encode(), pos[0].numel(), *(dptr(t) for t in pos),
                                 *(dptr(t) for t in vel), st))


def load_snapshot_numpy(path):
    """Host-side reader of the same format (numpy only, no GPU): returns
    (header dict, positions (3, np), velocities (3, np))."""
    import numpy as np

    with open(path, "rb") as f:
        head = f.read(SNAP_HEADER)
    if len(head) != SNAP_HEADER or head[:8] != SNAP_MAGIC:
        raise ValueError(f"{path}: not a PM snapshot")
    version, hsize = np.frombuffer(head[8:16], "<u4")
    if version != 1 or hsize != SNAP_HEADER:
        raise ValueError(f"{path}: unsupported snapshot version {version}")
    n = int(np.frombuffer(head[16:24], "<i8")[0])
    ng, step = (int(v) for v in np.frombuffer(head[24:32], "<i4"))
    a = float(np.frombuffer(head[32:40], "<f8")[0])
    data = np.fromfile(path, dtype="<f8", offset=SNAP_HEADER, count=6 * n).reshape(6, n)
    return {"np": n, "ng": ng, "step": step, "a": a}, data[:3], data[3:]


# ------------------------------------------------------------ tuning cache
def wisdom_import(path) -> None:
    """Merge a wisdom file into the process-wide FFT shape cache."""
    check(lib().pm_wisdom_import(str(path).encode()))


def wisdom_export(path) -> None:
    check(lib().pm_wisdom_export(str(path).encode()))


def wisdom_lookup(ng: int, precision: int):
    """The cached shape of (ng, precision) as a list of 5 variants, or None."""
    v = (ctypes.c_int * PM_FFT_SLOTS)()
    found = ctypes.c_int()
    check(lib().pm_wisdom_lookup(ng, precision, v, ctypes.byref(found)))
    return list(v) if found.value else None


def wisdom_forget() -> None:
    lib().pm_wisdom_forget()


def copy_rate(nbytes: int = 1 << 30, repeats: int = 20) -> float:
    """The device's streaming copy rate in GB/s (bytes read + written), the
    measured ceiling of the HBM-bound kernels (pm_copy_rate)."""
    gbs = ctypes.c_double()
    check(lib().pm_copy_rate(nbytes, repeats, ctypes.byref(gbs)))
    return gbs.value
