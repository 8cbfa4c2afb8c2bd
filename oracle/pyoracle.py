"""numpy/ctypes binding of the FP64 oracle (liboracle.so) and, where it was
built, of the reference's own operators (_ref/libpmref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, as the reference that GPU results are
compared against or as the timed CPU baseline.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "liboracle.so"
REF_SO = HERE / "_ref" / "libpmref.so"

GREEN_REF_KH = 0
GREEN_CONT_K2 = 1

_orc = None
_ref = None

_D = ctypes.c_double
_I = ctypes.c_int
_I64 = ctypes.c_int64
_P = ctypes.c_void_p


def _p(a: np.ndarray) -> int:
    assert a.flags.c_contiguous
    return a.ctypes.data


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def orc() -> ctypes.CDLL:
    global _orc
    if _orc is None:
        if not ORACLE_SO.exists():
            build()
        lib = ctypes.CDLL(str(ORACLE_SO))
        sig = {
            "orc_mt_uniform": (None, [ctypes.c_uint64, _I64, _D, _D, _P]),
            "orc_mt_raw": (None, [ctypes.c_uint64, _I64, _P]),
            "orc_init_particles": (None, [_I64, _I, _P, _P, _P]),
            "orc_init_cube": (None, [_I, _P]),
            "orc_deposit_cube": (None, [_I64, _P, _P, _P, _D, _I, _P]),
            "orc_ghost_fold": (None, [_I, _P, _P]),
            "orc_halo_expand": (None, [_I, _P, _P]),
            "orc_centered_derivatives": (None, [_I, _P, _P, _P, _P]),
            "orc_update_particles": (None, [_I64, _P, _P, _P, _P, _P, _P, _I, _D, _P, _P, _P]),
            "orc_function_a": (_D, [_D]),
            "orc_fft_r2c": (_I, [_I, _P, _P]),
            "orc_fft_c2r": (_I, [_I, _P, _P]),
            "orc_green": (_D, [_I, _I, _I, _I, _I]),
            "orc_solve": (_I, [_I, _I, _D, _P, _P]),
            "orc_deposit_periodic": (None, [_I64, _P, _P, _P, _D, _I, _P]),
            "orc_interp_force": (None, [_I64, _P, _P, _P, _P, _I, _P, _P, _P]),
            "orc_interp_kick_drift": (None, [_I64, _P, _P, _P, _P, _P, _P, _P, _I, _D, _D]),
            "orc_run_kdk": (_I, [_I64, _P, _P, _P, _P, _P, _P, _I, _I, _D, _I, _D]),
            "orc_run_cosmo": (_I, [_I64, _P, _P, _P, _P, _P, _P, _I, _I, _D, _D, _I, _D]),
            "orc_max_threads": (_I, []),
            "orc_set_threads": (None, [_I]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _orc = lib
    return _orc


def ref_available() -> bool:
    return REF_SO.exists()


def ref() -> ctypes.CDLL:
    global _ref
    if _ref is None:
        lib = ctypes.CDLL(str(REF_SO))
        lib.ref_init_particles.restype = None
        lib.ref_init_particles.argtypes = [ctypes.c_longlong, _I, _P, _P, _P]
        lib.ref_step.restype = _I
        lib.ref_step.argtypes = [_I, ctypes.c_longlong, _D] + [_P] * 13
        lib.ref_fft_roundtrip.restype = _I
        lib.ref_fft_roundtrip.argtypes = [_I, _P, _P]
        _ref = lib
    return _ref


# ----------------------------------------------------------------- oracle
def mt_raw(seed: int, count: int) -> np.ndarray:
    out = np.zeros(count, dtype=np.uint64)
    orc().orc_mt_raw(seed, count, _p(out))
    return out


def init_particles(np_: int, ng: int):
    x, y, z = (np.zeros(np_) for _ in range(3))
    orc().orc_init_particles(np_, ng, _p(x), _p(y), _p(z))
    return x, y, z


def function_a(a: float) -> float:
    return orc().orc_function_a(a)


def reference_step(ng: int, pos, p, a: float = 0.9):
    """One reference kick in the reference layout (cube -> fold -> solve ->
    halo -> centred differences -> update_particles).  Returns a dict of the
    intermediate meshes; p is updated in place."""
    x, y, z = pos
    npart = x.size
    n1, n3 = ng + 1, ng + 3
    cube = np.zeros(n1 ** 3)
    orc().orc_init_cube(ng, _p(cube))
    orc().orc_deposit_cube(npart, _p(x), _p(y), _p(z), ng ** 3 / npart, ng, _p(cube))
    slab = np.zeros(ng ** 3)
    orc().orc_ghost_fold(ng, _p(cube), _p(slab))
    phi = np.zeros(ng ** 3)
    rc = orc().orc_solve(ng, GREEN_REF_KH, 1.0 / a, _p(slab), _p(phi))
    assert rc == 0
    halo = np.zeros(n3 ** 3)
    orc().orc_halo_expand(ng, _p(phi), _p(halo))
    d = [np.zeros(n1 ** 3) for _ in range(3)]
    orc().orc_centered_derivatives(ng, _p(halo), *(_p(v) for v in d))
    kick = function_a(a) * 0.001
    orc().orc_update_particles(npart, _p(x), _p(y), _p(z), *(_p(v) for v in d), ng, kick, *(_p(v) for v in p))
    return {"cube": cube, "slab": slab, "phi": phi, "halo": halo, "dx": d[0], "dy": d[1], "dz": d[2]}


def deposit_periodic(pos, ng: int, mass: float | None = None) -> np.ndarray:
    x, y, z = (np.ascontiguousarray(v, dtype=np.float64) for v in pos)
    out = np.zeros(ng ** 3)
    m = ng ** 3 / x.size if mass is None else mass
    orc().orc_deposit_periodic(x.size, _p(x), _p(y), _p(z), m, ng, _p(out))
    return out.reshape(ng, ng, ng)


def solve(delta: np.ndarray, mode: int = GREEN_REF_KH, scale: float = 1.0 / 0.9) -> np.ndarray:
    ng = delta.shape[0]
    d = np.ascontiguousarray(delta, dtype=np.float64)
    out = np.zeros_like(d)
    rc = orc().orc_solve(ng, mode, scale, _p(d), _p(out))
    assert rc == 0
    return out


def fft_r2c(real: np.ndarray) -> np.ndarray:
    ng = real.shape[0]
    r = np.ascontiguousarray(real, dtype=np.float64)
    out = np.zeros((ng, ng, ng // 2 + 1), dtype=np.complex128)
    assert orc().orc_fft_r2c(ng, _p(r), _p(out)) == 0
    return out


def fft_c2r(spec: np.ndarray) -> np.ndarray:
    ng = spec.shape[0]
    s = np.ascontiguousarray(spec, dtype=np.complex128)
    out = np.zeros((ng, ng, ng))
    assert orc().orc_fft_c2r(ng, _p(s), _p(out)) == 0
    return out


def green(mode: int, ng: int, i: int, j: int, k: int) -> float:
    return orc().orc_green(mode, ng, i, j, k)


def interp_force(pos, phi: np.ndarray):
    x, y, z = (np.ascontiguousarray(v, dtype=np.float64) for v in pos)
    ng = phi.shape[0]
    ph = np.ascontiguousarray(phi, dtype=np.float64)
    g = [np.zeros(x.size) for _ in range(3)]
    orc().orc_interp_force(x.size, _p(x), _p(y), _p(z), _p(ph), ng, *(_p(v) for v in g))
    return g


def interp_kick_drift(pos, vel, phi: np.ndarray, kick: float, drift: float):
    """In place on the float64 arrays pos and vel."""
    ng = phi.shape[0]
    ph = np.ascontiguousarray(phi, dtype=np.float64)
    orc().orc_interp_kick_drift(pos[0].size, *(_p(v) for v in pos), *(_p(v) for v in vel), _p(ph), ng, kick, drift)


def run_kdk(pos, vel, ng: int, nsteps: int, dt: float, mode: int, scale: float):
    """In place on the float64 arrays pos and vel."""
    rc = orc().orc_run_kdk(pos[0].size, *(_p(v) for v in pos), *(_p(v) for v in vel), ng, nsteps, dt, mode, scale)
    assert rc == 0


def run_cosmo(pos, mom, ng: int, nsteps: int, a0: float, da: float, mode: int, scale: float):
    """In place on the float64 arrays pos and mom (cosmological KDK in a)."""
    rc = orc().orc_run_cosmo(pos[0].size, *(_p(v) for v in pos), *(_p(v) for v in mom), ng, nsteps, a0, da, mode,
                             scale)
    assert rc == 0


def set_threads(n: int) -> None:
    orc().orc_set_threads(n)


def max_threads() -> int:
    return orc().orc_max_threads()


# ------------------------------------------------------- reference harness
def ref_init_particles(np_: int, ng: int):
    x, y, z = (np.zeros(np_) for _ in range(3))
    ref().ref_init_particles(np_, ng, _p(x), _p(y), _p(z))
    return x, y, z


def ref_step(ng: int, pos, a: float = 0.9):
    """The reference's own operators at one rank; returns outputs dict."""
    x, y, z = (np.ascontiguousarray(v, dtype=np.float64) for v in pos)
    npart = x.size
    n1, n3 = ng + 1, ng + 3
    out = {k: np.zeros(n1 ** 3) for k in ("cube", "dx", "dy", "dz")}
    out.update({"slab": np.zeros(ng ** 3), "phi": np.zeros(ng ** 3), "halo": np.zeros(n3 ** 3)})
    p = [np.zeros(npart) for _ in range(3)]
    rc = ref().ref_step(ng, npart, a, _p(x), _p(y), _p(z), *(_p(v) for v in p), _p(out["cube"]), _p(out["slab"]),
                        _p(out["phi"]), _p(out["halo"]), _p(out["dx"]), _p(out["dy"]), _p(out["dz"]))
    assert rc == 0
    out["px"], out["py"], out["pz"] = p
    return out


# ------------------------------------------------------- initial conditions
def zeldovich_ic(noise: np.ndarray, ng: int, rms_cells: float, vel_factor: float):
    """numpy restatement of pm_ic_zeldovich (FP64): displaced lattice from a
    white-noise field on the nrow^3 particle lattice."""
    nrow = noise.shape[0]
    w = np.fft.rfftn(noise)
    f = np.fft.fftfreq(nrow, 1.0 / nrow)
    kx = f[:, None, None]
    ky = f[None, :, None]
    kz = np.arange(nrow // 2 + 1, dtype=np.float64)[None, None, :]
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    psi = []
    for ka in (kx, ky, kz):
        with np.errstate(divide="ignore", invalid="ignore"):
            fac = np.where((k2 > 0) & (np.abs(ka) != nrow // 2), ka / (k2 * np.sqrt(k2)), 0.0)
        psi.append(np.fft.irfftn(1j * fac * w, s=noise.shape, axes=(0, 1, 2)).ravel())
    s = rms_cells / np.sqrt(np.mean(psi[0] ** 2 + psi[1] ** 2 + psi[2] ** 2))
    idx = (np.arange(nrow) + 0.5) * (ng / nrow)
    Q = np.meshgrid(idx, idx, idx, indexing="ij")
    pos, vel = [], []
    for q, p in zip(Q, psi):
        x = q.ravel() + s * p
        x = x - ng * np.floor(x / ng)
        x[x >= ng] = 0.0
        pos.append(x)
        vel.append(vel_factor * s * p)
    return pos, vel
