"""Python mirror of the reference's sheet-step API over libsilt_amd.so.

The functions keep the reference's names and argument meaning
(src/fluid_core.rs ``step_sheet``, ``SheetDrive``) and call the C ABI declared
in include/silt_sheet.h.  There is no fallback: if the HIP library is missing
or a call fails, they raise.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)  # dirty-gen_amd/
LIB_PATH = os.path.join(ROOT, "libsilt_amd.so")

SILT_OK = 0
SILT_EINVAL = -1
SILT_EDEVICE = -2
SILT_ENOMEM = -3
SILT_EUNSUPPORTED = -4
SILT_ETOOBIG = -5
NOALLOC_MAX_CONTACTS = 1024  # SILT_NOALLOC_MAX_CONTACTS, include/silt_rut.h
DRIVE_GRAVITY = 0
DRIVE_WIND = 1
CONFINE = {"reference": 0, "jacobi": 1}
PATH = {"exact": 0, "fast": 1, "fused_exact": 2}

_lib = None


class SiltError(RuntimeError):
    pass


class Drive(C.Structure):
    """SheetDrive (src/fluid_core.rs:204-211)."""
    _fields_ = [("kind", C.c_int32), ("x", C.c_float), ("z", C.c_float), ("vorticity", C.c_float)]

    @classmethod
    def gravity(cls, vorticity):
        return cls(DRIVE_GRAVITY, 0.0, 0.0, vorticity)

    @classmethod
    def wind(cls, x, z, vorticity):
        return cls(DRIVE_WIND, x, z, vorticity)


class SheetOpts(C.Structure):
    _fields_ = [("eps", C.c_float), ("banks", C.c_int32), ("sub_cap", C.c_float), ("pace", C.c_int32),
                ("row_split", C.c_int32), ("confine", C.c_int32), ("path", C.c_int32)]


class LayerLedger(C.Structure):
    _fields_ = [("volume_m3", C.c_double), ("initial_m3", C.c_double), ("cut_m3", C.c_double),
                ("residue_m3", C.c_double), ("gap_m3", C.c_double)]


class PairLedger(C.Structure):
    _fields_ = [("water", LayerLedger), ("slurry", LayerLedger), ("total_m3", C.c_double)]


def build() -> None:
    """Compile libsilt_amd.so for gfx950 with the package Makefile."""
    subprocess.run(["make", "-s", "-j8", "-C", ROOT], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SiltError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


_P = C.c_void_p
_Z = C.c_size_t
_F = C.c_float


def _declare(L):
    L.silt_sheet_opts_default.argtypes = [C.POINTER(SheetOpts)]
    L.silt_sheet_opts_default.restype = None
    L.silt_version.restype = C.c_char_p
    L.silt_step_sheet.argtypes = [_P, _P, _P, _P, _P, _P, _Z, _Z, _F, _F, Drive]
    L.silt_step_sheet.restype = C.c_int
    L.silt_step_sheet_ex.argtypes = [_P, _P, _P, _P, _P, _P, _Z, _Z, _F, _F, Drive, C.POINTER(SheetOpts)]
    L.silt_step_sheet_ex.restype = C.c_int
    L.silt_step_sheet_device.argtypes = [_P, _P, _P, _P, _P, _F, _P, _Z, _F, _F, Drive, C.POINTER(SheetOpts), _P]
    L.silt_step_sheet_device.restype = C.c_int
    L.silt_step_sheet_walled.argtypes = [_P, _P, _P, _P, _P, _P, _P, _Z, _Z, _F, _F, Drive, C.POINTER(SheetOpts)]
    L.silt_step_sheet_walled.restype = C.c_int
    L.silt_step_sheet_walled_device.argtypes = [_P, _P, _P, _P, _P, _F, _P, _P, _Z, _F, _F, Drive,
                                                C.POINTER(SheetOpts), _P]
    L.silt_step_sheet_walled_device.restype = C.c_int
    L.silt_displace_occupied.argtypes = [_P, _P, _P, _P, _Z, _Z]
    L.silt_displace_occupied.restype = C.c_int
    L.silt_pair_create.argtypes = [_Z, _F, C.c_int]
    L.silt_pair_create.restype = _P
    L.silt_pair_destroy.argtypes = [_P]
    L.silt_pair_destroy.restype = None
    L.silt_pair_upload.argtypes = [_P, _P, _P, _P, _P, _P, _P, _P, _P, _F]
    L.silt_pair_upload.restype = C.c_int
    L.silt_pair_download.argtypes = [_P, _P, _P, _P, _P, _P, _P]
    L.silt_pair_download.restype = C.c_int
    L.silt_pair_step.argtypes = [_P, _F]
    L.silt_pair_step.restype = C.c_int
    L.silt_pair_ledger.argtypes = [_P, C.POINTER(PairLedger)]
    L.silt_pair_ledger.restype = C.c_int
    L.silt_pair_last_kernel_ms.argtypes = [_P]
    L.silt_pair_last_kernel_ms.restype = _F
    L.silt_pair_last_fallback.argtypes = [_P]
    L.silt_pair_last_fallback.restype = C.c_int
    L.silt_pair_autotune.argtypes = [_P, _F, C.POINTER(C.c_int), _Z, C.c_int, C.POINTER(C.c_int),
                                     C.POINTER(C.c_float)]
    L.silt_pair_autotune.restype = C.c_int
    L.silt_pair_set_xcd_weights.argtypes = [_P, C.POINTER(C.c_float)]
    L.silt_pair_balance.argtypes = [_P, _F, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                    C.POINTER(C.c_float)]
    L.silt_pair_clock_records.argtypes = [_P, C.POINTER(C.c_ulonglong), _Z]
    L.silt_pair_clock_records.restype = _Z
    L.silt_pair_set_speculation.argtypes = [_P, C.c_int]
    L.silt_pair_speculation_counts.argtypes = [_P, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    L.silt_batch_create.argtypes = [_Z, C.POINTER(C.c_size_t), C.POINTER(C.c_float), C.c_int]
    L.silt_batch_create.restype = _P
    L.silt_batch_destroy.argtypes = [_P]
    L.silt_batch_destroy.restype = None
    L.silt_batch_upload.argtypes = [_P, _Z, _P, _P, _P, _P, _F, _F]
    L.silt_batch_download.argtypes = [_P, _Z, _P, _P, _P]
    L.silt_batch_step.argtypes = [_P, _F]
    L.silt_batch_set_xcd_weights.argtypes = [_P, C.POINTER(C.c_float)]
    L.silt_batch_xcd_weights_on.argtypes = [_P]
    L.silt_batch_ledger.argtypes = [_P, _Z, C.POINTER(LayerLedger)]
    L.silt_batch_last_kernel_ms.argtypes = [_P]
    L.silt_batch_last_kernel_ms.restype = C.c_float
    for n in ("silt_batch_last_fallbacks", "silt_batch_graphs_captured"):
        getattr(L, n).argtypes = [_P]
    L.silt_batch_use_graph.argtypes = [_P, C.c_int]
    L.silt_batch_set_rows.argtypes = [_P, C.c_int]
    L.silt_batch_autotune.argtypes = [_P, _F, C.POINTER(C.c_int), _Z, C.c_int, C.POINTER(C.c_int),
                                      C.POINTER(C.c_float)]
    L.silt_plan_boxes_debug.argtypes = [C.POINTER(C.c_void_p), C.POINTER(C.c_int), _Z, _F, _F, _P, _P, _Z]
    L.silt_set_chunk_rows.argtypes = [_Z, C.c_int]
    L.silt_set_chunk_rows.restype = C.c_int
    L.silt_chunk_rows.argtypes = [_Z]
    L.silt_chunk_rows.restype = C.c_int
    L.silt_mlp_create.argtypes = [C.c_uint]
    L.silt_mlp_create.restype = _P
    L.silt_mlp_destroy.argtypes = [_P]
    L.silt_mlp_destroy.restype = None
    L.silt_mlp_get_params.argtypes = [_P, _P]
    L.silt_mlp_set_params.argtypes = [_P, _P]
    L.silt_mlp_forward.argtypes = [_P, _P, _Z, _P]
    L.silt_mlp_infer.argtypes = [_P, _P, _Z, _P]
    L.silt_mlp_gradients.argtypes = [_P, _P, _P, _Z, C.POINTER(C.c_double), _P]
    L.silt_mlp_train_batch.argtypes = [_P, _P, _P, _Z, _F, C.POINTER(C.c_double)]


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _drive(drive):
    if isinstance(drive, Drive):
        return drive
    if drive[0] == "gravity":
        return Drive.gravity(drive[1])
    if drive[0] == "wind":
        return Drive.wind(drive[1], drive[2], drive[3])
    raise ValueError(drive)


def sheet_opts(*, eps=1.0e-6, banks=False, sub_cap=8.0, pace=0, row_split=True, confine="reference",
               path="exact"):
    o = SheetOpts()
    lib().silt_sheet_opts_default(C.byref(o))
    o.eps, o.banks, o.sub_cap, o.pace = eps, int(banks), sub_cap, pace
    o.row_split, o.confine, o.path = int(row_split), CONFINE[confine], PATH[path]
    return o


def _check_status(rc):
    if rc < 0:
        raise SiltError(f"silt call failed with status {rc}")
    return rc


def step_sheet(h, u, v, bed, nu, gate, res, cell, dt, drive, **opts):
    """``step_sheet`` (src/fluid_core.rs:770-785) on numpy float32 planes, in
    place.  Keyword options are those of ``step_sheet_eps``.  Returns the
    number of substeps (0 when the reference would return early)."""
    for a, nm in ((h, "h"), (u, "u"), (v, "v"), (bed, "bed"), (nu, "nu")):
        if a.dtype != np.float32 or not a.flags.c_contiguous:
            raise TypeError(f"{nm} must be C-contiguous float32")
    if gate.dtype != np.uint8 or not gate.flags.c_contiguous:
        raise TypeError("gate must be C-contiguous uint8")
    n = min(a.size for a in (h, u, v, bed, nu, gate))
    o = sheet_opts(**opts)
    rc = lib().silt_step_sheet_ex(_ptr(h), _ptr(u), _ptr(v), _ptr(bed), _ptr(nu), _ptr(gate), n, res, cell, dt,
                                  _drive(drive), C.byref(o))
    return _check_status(rc)


def step_sheet_device(h, u, v, bed, nu, gate, res, cell, dt, drive, nu_uniform=math.nan, stream=0, **opts):
    """The same step on device buffers.  Arguments are integer device
    addresses (``tensor.data_ptr()``); ``nu``/``gate`` may be 0 with
    ``nu_uniform`` finite / gate all ones."""
    o = sheet_opts(**opts)
    rc = lib().silt_step_sheet_device(h, u, v, bed, nu or None, nu_uniform, gate or None, res, cell, dt,
                                      _drive(drive), C.byref(o), stream or None)
    return _check_status(rc)


def step_sheet_walled(h, u, v, bed, nu, gate, solid, res, cell, dt, drive, **opts):
    """The walled sheet (include/silt_sheet.h): ``step_sheet_eps_in`` with a
    wall mask (uint8, nonzero = wall), on numpy float32 planes in place.
    Returns the number of substeps."""
    for a, nm in ((h, "h"), (u, "u"), (v, "v"), (bed, "bed"), (nu, "nu")):
        if a.dtype != np.float32 or not a.flags.c_contiguous:
            raise TypeError(f"{nm} must be C-contiguous float32")
    for a, nm in ((gate, "gate"), (solid, "solid")):
        if a.dtype != np.uint8 or not a.flags.c_contiguous:
            raise TypeError(f"{nm} must be C-contiguous uint8")
    n = min(a.size for a in (h, u, v, bed, nu, gate, solid))
    o = sheet_opts(**opts)
    rc = lib().silt_step_sheet_walled(_ptr(h), _ptr(u), _ptr(v), _ptr(bed), _ptr(nu), _ptr(gate), _ptr(solid), n,
                                      res, cell, dt, _drive(drive), C.byref(o))
    return _check_status(rc)


def step_sheet_walled_device(h, u, v, bed, nu, gate, solid, res, cell, dt, drive, nu_uniform=math.nan, stream=0,
                             **opts):
    """The walled step on device buffers (integer device addresses)."""
    o = sheet_opts(**opts)
    rc = lib().silt_step_sheet_walled_device(h, u, v, bed, nu or None, nu_uniform, gate or None, solid, res, cell,
                                             dt, _drive(drive), C.byref(o), stream or None)
    return _check_status(rc)


def displace_occupied(h, u, v, solid, res):
    """displace_occupied (src/fluid_core.rs:4382-4444) on numpy planes in place."""
    for a, nm in ((h, "h"), (u, "u"), (v, "v")):
        if a.dtype != np.float32 or not a.flags.c_contiguous:
            raise TypeError(f"{nm} must be C-contiguous float32")
    if solid.dtype != np.uint8 or not solid.flags.c_contiguous:
        raise TypeError("solid must be C-contiguous uint8")
    n = min(a.size for a in (h, u, v, solid))
    return _check_status(lib().silt_displace_occupied(_ptr(h), _ptr(u), _ptr(v), _ptr(solid), n, res))


def set_chunk_rows(res, rows):
    """Rows per workgroup of the fused kernel for sheets of side res (0: default)."""
    _check_status(lib().silt_set_chunk_rows(res, rows))


def chunk_rows(res):
    """The rows pinned for side res, 0 when the default applies."""
    return int(_check_status(lib().silt_chunk_rows(res)))


class Batch:
    """Independent sheet windows stepped together (include/silt_batch.h):
    window k is ``res[k]`` square cells of ``cell[k]`` metres, one layer each,
    stepped as ``step_sheet`` with a Gravity drive."""

    def __init__(self, res, cell, path="fast"):
        self.res = [int(r) for r in res]
        n = len(self.res)
        cells = [float(c) for c in (cell if hasattr(cell, "__len__") else [cell] * n)]
        self._h = lib().silt_batch_create(n, (C.c_size_t * n)(*self.res), (C.c_float * n)(*cells), PATH[path])
        if not self._h:
            raise SiltError("silt_batch_create failed")

    def close(self):
        if getattr(self, "_h", None):
            lib().silt_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, k, bed, h, u, v, nu=0.9, vorticity=0.08):
        arrs = [np.ascontiguousarray(a, np.float32) for a in (bed, h, u, v)]
        n = self.res[k] ** 2
        if any(a.size != n for a in arrs):
            raise ValueError("plane size does not match the window")
        _check_status(lib().silt_batch_upload(self._h, k, *(_ptr(a) for a in arrs), nu, vorticity))

    def download(self, k):
        out = [np.empty(self.res[k] ** 2, np.float32) for _ in range(3)]
        _check_status(lib().silt_batch_download(self._h, k, *(_ptr(a) for a in out)))
        return out

    def step(self, dt):
        return _check_status(lib().silt_batch_step(self._h, dt))

    def ledger(self, k):
        led = LayerLedger()
        _check_status(lib().silt_batch_ledger(self._h, k, C.byref(led)))
        return led

    def last_kernel_ms(self):
        return lib().silt_batch_last_kernel_ms(self._h)

    def last_fallbacks(self):
        return _check_status(lib().silt_batch_last_fallbacks(self._h))

    def graphs_captured(self):
        return _check_status(lib().silt_batch_graphs_captured(self._h))

    def use_graph(self, on):
        _check_status(lib().silt_batch_use_graph(self._h, 1 if on else 0))

    def set_rows(self, rows):
        _check_status(lib().silt_batch_set_rows(self._h, rows))

    def autotune(self, dt, rows=None, reps=3):
        best, ms = C.c_int(0), C.c_float(0.0)
        arr = None if not rows else (C.c_int * len(rows))(*rows)
        _check_status(lib().silt_batch_autotune(self._h, dt, arr, len(rows) if rows else 0, reps, C.byref(best),
                                                C.byref(ms)))
        return best.value, ms.value

    def set_xcd_weights(self, weights):
        """XCD weights for the launches (8 in [0.25, 4], e.g. from
        SheetPair.balance on this chip); None clears them.  autotune() keeps
        them only when they make the step faster."""
        arr = None if weights is None else (C.c_float * 8)(*[float(w) for w in weights])
        _check_status(lib().silt_batch_set_xcd_weights(self._h, arr))

    def xcd_weights_on(self):
        return bool(_check_status(lib().silt_batch_xcd_weights_on(self._h)))


def plan_boxes(planes, cell, eps=1.0e-6, cap=64):
    """Work boxes (i0, i1, j0, j1) of square depth planes planned together on
    the GPU, as a batch step plans them (the reference's sheet_work_boxes,
    src/fluid_core.rs:1325-1447).  One list per plane; [] for a dry plane,
    None when the plane would be planned on the host."""
    arrs = [np.ascontiguousarray(a, np.float32).ravel() for a in planes]
    res = [int(round(np.sqrt(a.size))) for a in arrs]
    for a, r in zip(arrs, res):
        if r * r != a.size:
            raise ValueError("planes must be square")
    n = len(arrs)
    hp = (C.c_void_p * n)(*[a.ctypes.data for a in arrs])
    rs = (C.c_int * n)(*res)
    counts = np.zeros(n, np.int32)
    boxes = np.zeros((n, cap, 4), np.int32)
    _check_status(lib().silt_plan_boxes_debug(hp, rs, n, cell, eps, _ptr(counts), _ptr(boxes), cap))
    out = []
    for k in range(n):
        c = int(counts[k])
        out.append(None if c < 0 else [tuple(int(v) for v in boxes[k, b]) for b in range(min(c, cap))])
    return out


class SheetPair:
    """Water and slurry over one bed, resident on the GPU (include/silt_pair.h):
    one ``step`` is the generic-res ``step_sheet`` on both layers with the
    layer constants of ``FluidCore::step_pair``."""

    def __init__(self, res, cell, path="fast"):
        self.res, self.cell = int(res), float(cell)
        self._h = lib().silt_pair_create(self.res, self.cell, PATH[path])
        if not self._h:
            raise SiltError("silt_pair_create failed")

    def close(self):
        if self._h:
            lib().silt_pair_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, bed, water, uw, vw, slurry, us, vs, nu_slurry):
        arrs = [np.ascontiguousarray(a, np.float32) for a in (bed, water, uw, vw, slurry, us, vs)]
        n = self.res * self.res
        if any(a.size != n for a in arrs):
            raise ValueError("planes must be res * res")
        if np.ndim(nu_slurry) == 0:
            nu_ptr, nu0 = None, float(nu_slurry)
        else:
            nu_arr = np.ascontiguousarray(nu_slurry, np.float32)
            nu_ptr, nu0 = _ptr(nu_arr), math.nan
        _check_status(lib().silt_pair_upload(self._h, *(_ptr(a) for a in arrs), nu_ptr, nu0))

    def download(self):
        n = self.res * self.res
        out = [np.empty(n, np.float32) for _ in range(6)]
        _check_status(lib().silt_pair_download(self._h, *(_ptr(a) for a in out)))
        return out

    def step(self, dt):
        return _check_status(lib().silt_pair_step(self._h, dt))

    def ledger(self):
        led = PairLedger()
        _check_status(lib().silt_pair_ledger(self._h, C.byref(led)))
        return led

    def last_kernel_ms(self):
        return lib().silt_pair_last_kernel_ms(self._h)

    def last_fallback(self):
        return lib().silt_pair_last_fallback(self._h)

    def set_speculation(self, on):
        """Launch the last step's plan before this step's plan is confirmed
        (default on); a different plan is undone and redone."""
        _check_status(lib().silt_pair_set_speculation(self._h, 1 if on else 0))

    def speculation_counts(self):
        h, m = C.c_longlong(0), C.c_longlong(0)
        _check_status(lib().silt_pair_speculation_counts(self._h, C.byref(h), C.byref(m)))
        return h.value, m.value

    def _tires_declared(self):
        L = lib()
        if not getattr(L, "_pair_tires_declared", False):
            L.silt_pair_set_ground.argtypes = [_P, _P, _F, _F]
            L.silt_pair_get_ground.argtypes = [_P, _P, _P, _P]
            L.silt_pair_set_surrogate.argtypes = [_P, _P]
            L.silt_pair_step_tires.argtypes = [_P, _F, C.POINTER(TirePush), _Z, _Z, C.POINTER(TireContact), _Z]
            L.silt_pair_last_tire_ms.argtypes = [_P]
            L.silt_pair_last_tire_ms.restype = _F
            L.silt_pair_set_tire_batching.argtypes = [_P, C.c_int]
            L._pair_tires_declared = True
        return L

    def set_ground(self, bed, origin):
        b = np.ascontiguousarray(bed, np.float32)
        _check_status(self._tires_declared().silt_pair_set_ground(self._h, _ptr(b), origin[0], origin[1]))

    def ground(self):
        n = self.res * self.res
        out = [np.empty(n, np.float32) for _ in range(3)]
        _check_status(self._tires_declared().silt_pair_get_ground(self._h, *(_ptr(a) for a in out)))
        return out

    def set_surrogate(self, params):
        p = None if params is None else np.ascontiguousarray(params, np.float32)
        _check_status(self._tires_declared().silt_pair_set_surrogate(self._h, None if p is None else _ptr(p)))

    def step_tires(self, dt, pushes, per_vehicle=4, contacts=()):
        """One step with vehicles (include/silt_pair.h): pushes are dicts of
        push_tire's fields, per_vehicle consecutive ones per vehicle;
        contacts are dicts of TireContact's fields."""
        L = self._tires_declared()
        parr = (TirePush * max(len(pushes), 1))(*[TirePush(**q) for q in pushes])
        carr = (TireContact * max(len(contacts), 1))(*[TireContact(**c) for c in contacts])
        return _check_status(L.silt_pair_step_tires(self._h, dt, parr, len(pushes), per_vehicle, carr, len(contacts)))

    def last_tire_ms(self):
        return self._tires_declared().silt_pair_last_tire_ms(self._h)

    def set_tire_batching(self, on):
        """One launch per tire stage for all vehicles (default) or one per
        vehicle; the results are the same."""
        _check_status(self._tires_declared().silt_pair_set_tire_batching(self._h, 1 if on else 0))

    def autotune(self, dt, rows=None, reps=5):
        """Time candidate rows per workgroup of the fused kernel on this
        handle's state (restored afterwards) and pin the fastest for this
        side; returns (rows, mean kernel ms), rows 0 being the default."""
        best, ms = C.c_int(0), C.c_float(0.0)
        arr = None if not rows else (C.c_int * len(rows))(*rows)
        _check_status(lib().silt_pair_autotune(self._h, dt, arr, len(rows) if rows else 0, reps, C.byref(best),
                                               C.byref(ms)))
        return best.value, ms.value

    def set_xcd_weights(self, weights):
        """Rows of the fused launch's tiles scaled per XCD (8 weights in
        [0.25, 4]); None restores uniform chunks.  Results do not change."""
        arr = None if weights is None else (C.c_float * 8)(*[float(w) for w in weights])
        _check_status(lib().silt_pair_set_xcd_weights(self._h, arr))

    def balance(self, dt, reps=5):
        """Calibrate the XCD weights on this chip from the workgroups' clock
        records (state restored afterwards).  Returns (weights, (uniform ms,
        kept ms), when each XCD finished with uniform chunks (us after the
        launch's start), the same in the last round); the weights are all 1
        when no round was faster than uniform chunks."""
        w, ms, us = (C.c_float * 8)(), (C.c_float * 2)(), (C.c_float * 16)()
        _check_status(lib().silt_pair_balance(self._h, dt, reps, w, ms, us))
        return list(w), (ms[0], ms[1]), list(us[:8]), list(us[8:])

    def clock_records(self):
        """The last calibration step's per-workgroup records as a (n, 4)
        uint64 array: start and end (100 MHz real-time clock), pipeline rows
        | XCC_ID << 32, tile | HW_ID << 32; unused rows are zero."""
        n = lib().silt_pair_clock_records(self._h, None, 0)
        out = np.zeros((n, 4), np.uint64)
        lib().silt_pair_clock_records(self._h, out.ctypes.data_as(C.POINTER(C.c_ulonglong)), n)
        return out


class TirePush(C.Structure):
    """TirePush (src/fluid_core.rs:357-368)."""
    _fields_ = [(n, C.c_float) for n in ("x", "z", "forward_x", "forward_z", "radius", "width", "speed", "slip")]


class SprayEvent(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("x", "y", "z", "vx", "vy", "vz", "volume")] + [("kind", C.c_int32)]


class FluidLedger(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("water_m3", "slurry_m3", "frozen_water_m3", "frozen_slurry_m3",
                                          "spray_m3", "total_m3")]


GRAIN = {"clay": 0, "silt": 1, "sand": 2, "water": 3, "other": 4}
WINDOW_RES = 200


def _declare_core(L):
    L.silt_core_create.argtypes = [_F, _F, C.c_int]
    L.silt_core_create.restype = _P
    L.silt_core_destroy.argtypes = [_P]
    L.silt_core_destroy.restype = None
    for name, at in (("silt_core_set_fine_pace", [_P, C.c_int]), ("silt_core_set_fine_sub_cap", [_P, _F]),
                     ("silt_core_set_near_field", [_P, C.c_int]), ("silt_core_set_coarse_confine", [_P, C.c_int]),
                     ("silt_core_window_origin", [_P, _P]), ("silt_core_map_water_bed", [_P, _P, _P]),
                     ("silt_core_pour", [_P, _F, _F, _F, _F, _F, C.c_int, C.c_int, _F]),
                     ("silt_core_lay_slurry", [_P, _P, C.c_int, _F]),
                     ("silt_core_read_substrate", [_P, _P, _P, _P]),
                     ("silt_core_push_tire", [_P, C.POINTER(TirePush)]),
                     ("silt_core_step", [_P, _F, C.POINTER(SprayEvent), _Z, C.POINTER(C.c_size_t)]),
                     ("silt_core_ledger", [_P, C.POINTER(FluidLedger)]),
                     ("silt_core_get_planes", [_P] + [_P] * 8), ("silt_core_set_planes", [_P] + [_P] * 8),
                     ("silt_core_publish", [_P] + [_P] * 6)):
        getattr(L, name).argtypes = at
        getattr(L, name).restype = C.c_int
    L.silt_core_near_tiles.argtypes = [_P]
    L.silt_core_fine_water_debug.argtypes = [_P, C.POINTER(TirePush), _Z, _P, _P, _P, _P]
    L.silt_core_fine_water_debug.restype = C.c_int
    L.silt_core_near_tiles.restype = C.c_long
    L.silt_core_set_ground.argtypes = [_P, _P, _P]
    L.silt_core_set_ground.restype = C.c_int
    L.silt_core_get_ground.argtypes = [_P, _P, _P, _P]
    L.silt_core_get_ground.restype = C.c_int
    L.silt_core_map_water_velocity.argtypes = [_P, _P, _P]
    L.silt_core_non_finite_count.argtypes = [_P]
    L.silt_core_non_finite_count.restype = C.c_long
    L.silt_core_set_center.argtypes = [_P, _F, _F]
    L.silt_core_center.argtypes = [_P, _P]
    L.silt_core_coarse_cell.argtypes = [_P, _Z, _Z, C.POINTER(CoarseCell)]
    L.silt_core_sample.argtypes = [_P, _F, _F, C.POINTER(GroundSample)]
    L.silt_core_ledger_full.argtypes = [_P, C.POINTER(FluidLedgerFull)]
    L.silt_core_ledger_lines.argtypes = [_P, C.POINTER(CoreLines)]
    L.silt_core_set_hydro_solids.argtypes = [_P, C.POINTER(HydroSolid), _Z]
    L.silt_core_hydro_load.argtypes = [_P, C.POINTER(HydroLoad)]
    L.silt_core_solid_mask.argtypes = [_P, C.c_int, _P]
    L.silt_core_note_surface_foam.argtypes = [_P, _F]
    L.silt_core_pond_ruts.argtypes = [_P, _P]
    L.silt_core_get_grain.argtypes = [_P, _P]
    L.silt_core_get_suspended.argtypes = [_P, _P, _P]
    L.silt_core_set_suspended.argtypes = [_P, _P]
    L.silt_core_seat_rut_water.argtypes = [_P, _P, _P, _P, _P]
    L.silt_core_seat_rut_sheet.argtypes = [_P, _P]
    L.silt_core_get_tracers.argtypes = [_P, _P, _P, _P]
    L.silt_core_set_tracers.argtypes = [_P, _P, _P]
    L.silt_core_recenter.argtypes = [_P, _F, _F]
    L.silt_core_window_bounds.argtypes = [_P, _P]
    L.silt_core_lay_slurry_on_water.argtypes = [_P, _F, C.c_int]
    L.silt_core_lay_slurry_on_water.restype = _F
    L.silt_core_parked_cells.argtypes = [_P]
    L.silt_core_parked_cells.restype = C.c_long
    L.silt_core_drain_over_bed.argtypes = [_P, _F, C.c_uint]
    L.silt_core_drain_over_bed.restype = _F
    L.silt_core_set_tire_model.argtypes = [_P, C.c_int]
    L.silt_core_set_surrogate.argtypes = [_P, _P, C.c_int, _F]
    L.silt_core_surrogate_stats.argtypes = [_P, C.POINTER(SurrogateStats)]
    L.silt_trainer_create.argtypes = [_Z, C.c_uint]
    L.silt_trainer_create.restype = _P
    L.silt_trainer_destroy.argtypes = [_P]
    L.silt_trainer_destroy.restype = None
    L.silt_trainer_record.argtypes = [_P, _P, _F, C.POINTER(SprayEvent), _Z, C.POINTER(C.c_size_t)]
    for n in ("silt_trainer_steps", "silt_trainer_samples"):
        getattr(L, n).argtypes = [_P]
        getattr(L, n).restype = C.c_long
    L.silt_trainer_fit.argtypes = [_P, _Z, C.c_int, C.c_int, _F, C.POINTER(C.c_double)]
    L.silt_trainer_eval.argtypes = [_P, _Z, _Z, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.silt_trainer_params.argtypes = [_P, _P]
    L.silt_trainer_set_params.argtypes = [_P, _P]
    L.silt_surrogate_targets_debug.argtypes = [_P, _P, _P, _P] + [C.c_int] * 4 + [_P]
    L.silt_surrogate_apply_debug.argtypes = [_P, _P, _P] + [C.c_int] * 4


TIRE_MODEL = {"nested": 0, "coarse": 1, "surrogate": 2}


class CoarseCell(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("x", "z", "bed_m", "water_m", "slurry_m", "moisture", "u", "v",
                                         "settled_dust_m", "turbidity")]


class GroundSample(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("water_m", "slurry_m", "u", "v", "settled_dust_m", "moisture", "bed_m")]


class FluidLedgerFull(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "water_m3", "slurry_m3", "dust_air_m3", "dust_settled_m3", "frozen_water_m3", "frozen_slurry_m3",
        "frozen_dust_m3", "spray_m3", "fog_air_m3", "suspended_m3", "settled_suspended_m3", "frozen_suspended_m3",
        "entrained_m3", "reach_m3", "total_m3")]


class CoreLines(C.Structure):
    _fields_ = [(n, C.c_double * 2) for n in (
        "base_m3", "cut_m3", "residue_m3", "handoff_m3", "source_m3", "total_m3", "gap_m3")]


class HydroSolid(C.Structure):
    _fields_ = [("x", _F), ("z", _F), ("yaw", _F), ("half_length", _F), ("half_width", _F), ("bottom_y", _F),
                ("top_y", _F), ("contact", C.c_int32), ("radius", _F), ("width", _F)]


class HydroLoad(C.Structure):
    _fields_ = [(n, _F) for n in ("force_x", "force_z", "buoyancy_n", "hull_buoyancy_n", "wheel_buoyancy_n",
                                  "displaced_m3", "eroded_m3", "hole_m")] + \
               [("wheel_blocks", C.c_uint32), ("belly_blocks", C.c_uint32)]


class SurrogateStats(C.Structure):
    _fields_ = [("last_tire_ms", C.c_float), ("last_judge_err_m", C.c_float), ("last_judge_rms_m", C.c_float),
                ("judged", C.c_int),
                ("fallbacks", C.c_int), ("last_mode", C.c_int)]


class FluidCore:
    """``FluidCore`` on the sheet path (include/silt_core.h), state on the GPU.
    Reference closures are arrays sampled at coarse cell centres."""

    PLANES = ("water", "slurry", "uw", "vw", "us", "vs", "bed", "moisture")

    def __init__(self, center=(0.0, 0.0), path="exact"):
        L = lib()
        if not getattr(L, "_core_declared", False):
            _declare_core(L)
            L._core_declared = True
        self._c = L.silt_core_create(center[0], center[1], PATH[path])
        if not self._c:
            raise SiltError("silt_core_create failed")

    def close(self):
        if self._c:
            lib().silt_core_destroy(self._c)
            self._c = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_fine_pace(self, pace):
        _check_status(lib().silt_core_set_fine_pace(self._c, pace))

    def set_fine_sub_cap(self, cap):
        _check_status(lib().silt_core_set_fine_sub_cap(self._c, cap))

    def set_near_field(self, on):
        _check_status(lib().silt_core_set_near_field(self._c, int(on)))

    def set_coarse_confine(self, mode):
        _check_status(lib().silt_core_set_coarse_confine(self._c, CONFINE[mode]))

    def origin(self):
        o = np.zeros(2, np.float32)
        _check_status(lib().silt_core_window_origin(self._c, _ptr(o)))
        return o

    def pour_water(self, x0, x1, z0, z1, depth):
        _check_status(lib().silt_core_pour(self._c, x0, x1, z0, z1, depth, 1, 3, 1.0))

    def pour_slurry(self, x0, x1, z0, z1, depth, grain, moisture):
        g = GRAIN[grain] if isinstance(grain, str) else int(grain)
        _check_status(lib().silt_core_pour(self._c, x0, x1, z0, z1, depth, 0, g, moisture))

    def map_water_bed(self, bed, water):
        b, w = (np.ascontiguousarray(a, np.float32) for a in (bed, water))
        _check_status(lib().silt_core_map_water_bed(self._c, _ptr(b), _ptr(w)))

    def lay_slurry(self, depth, grain, moisture):
        d = np.ascontiguousarray(depth, np.float32)
        g = GRAIN[grain] if isinstance(grain, str) else int(grain)
        _check_status(lib().silt_core_lay_slurry(self._c, _ptr(d), g, moisture))

    def read_substrate(self, bed, moisture, grain):
        b, m = (np.ascontiguousarray(a, np.float32) for a in (bed, moisture))
        g = np.ascontiguousarray(grain, np.uint8)
        _check_status(lib().silt_core_read_substrate(self._c, _ptr(b), _ptr(m), _ptr(g)))

    def push_tire(self, x, z, forward_x, forward_z, radius, width, speed, slip):
        p = TirePush(x, z, forward_x, forward_z, radius, width, speed, slip)
        _check_status(lib().silt_core_push_tire(self._c, C.byref(p)))

    def step(self, dt, max_events=4096):
        ev = (SprayEvent * max_events)()
        n = C.c_size_t(0)
        _check_status(lib().silt_core_step(self._c, dt, ev, max_events, C.byref(n)))
        return [ev[k] for k in range(min(n.value, max_events))], n.value

    def ledger(self):
        led = FluidLedger()
        _check_status(lib().silt_core_ledger(self._c, C.byref(led)))
        return {k: getattr(led, k) for k, _ in FluidLedger._fields_}

    def planes(self):
        n = WINDOW_RES * WINDOW_RES
        out = {k: np.empty(n, np.float32) for k in self.PLANES}
        _check_status(lib().silt_core_get_planes(self._c, *(_ptr(out[k]) for k in self.PLANES)))
        return out

    def set_planes(self, planes):
        arrs = [None if planes.get(k) is None else np.ascontiguousarray(planes[k], np.float32) for k in self.PLANES]
        _check_status(lib().silt_core_set_planes(self._c, *(None if a is None else _ptr(a) for a in arrs)))

    def publish(self, soil_bed=None, soil_placed=None, soil_rut=None):
        n = WINDOW_RES * WINDOW_RES
        gr, film, flow = np.empty(4 * n, np.float32), np.empty(2 * n, np.float32), np.empty(4 * n, np.float32)
        soils = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (soil_bed, soil_placed, soil_rut)]
        _check_status(lib().silt_core_publish(self._c, *(None if a is None else _ptr(a) for a in soils),
                                              _ptr(gr), _ptr(film), _ptr(flow)))
        return gr, film, flow

    def publish_surface(self, soil_wetness=None, soil_placed=None):
        """moisture, wet_edge and foam_crest maps (include/silt_core.h)."""
        L = lib()
        if not getattr(L, "_publish_surface_declared", False):
            L.silt_core_publish_surface.argtypes = [_P, _P, _P, _P, _P, _P]
            L._publish_surface_declared = True
        n = WINDOW_RES * WINDOW_RES
        moist, edge, foam = np.empty(n, np.float32), np.empty(n, np.float32), np.empty(2 * n, np.float32)
        soils = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (soil_wetness, soil_placed)]
        _check_status(L.silt_core_publish_surface(self._c, *(None if a is None else _ptr(a) for a in soils),
                                                  _ptr(moist), _ptr(edge), _ptr(foam)))
        return moist, edge, foam

    def publish_device(self, ground_rut4, film2, flow4, soil_bed=0, soil_placed=0, soil_rut=0, stream=0):
        """silt_core_publish into device buffers given as integer addresses
        (0 for none); asynchronous on ``stream``."""
        L = lib()
        if not getattr(L, "_publish_device_declared", False):
            L.silt_core_publish_device.argtypes = [_P] * 8
            L.silt_core_publish_surface_device.argtypes = [_P] * 7
            L._publish_device_declared = True
        a = [x or None for x in (soil_bed, soil_placed, soil_rut, ground_rut4, film2, flow4, stream)]
        _check_status(L.silt_core_publish_device(self._c, *a))

    def publish_surface_device(self, moisture1, wet_edge1, foam_crest2, soil_wetness=0, soil_placed=0, stream=0):
        """silt_core_publish_surface into device buffers (integer addresses)."""
        L = lib()
        if not getattr(L, "_publish_device_declared", False):
            L.silt_core_publish_device.argtypes = [_P] * 8
            L.silt_core_publish_surface_device.argtypes = [_P] * 7
            L._publish_device_declared = True
        a = [x or None for x in (soil_wetness, soil_placed, moisture1, wet_edge1, foam_crest2, stream)]
        _check_status(L.silt_core_publish_surface_device(self._c, *a))

    def near_tiles(self):
        return int(_check_status(lib().silt_core_near_tiles(self._c)))

    def map_water_velocity(self, u, v):
        """map_water_velocity (src/fluid_core.rs:5358-5372) from per-cell arrays."""
        uu, vv = (np.ascontiguousarray(a, np.float32) for a in (u, v))
        _check_status(lib().silt_core_map_water_velocity(self._c, _ptr(uu), _ptr(vv)))

    def non_finite_count(self):
        return int(_check_status(lib().silt_core_non_finite_count(self._c)))

    def set_center(self, x, z):
        _check_status(lib().silt_core_set_center(self._c, x, z))

    def center(self):
        out = np.zeros(2, np.float32)
        _check_status(lib().silt_core_center(self._c, _ptr(out)))
        return out

    def coarse_cell(self, i, j):
        cc = CoarseCell()
        if _check_status(lib().silt_core_coarse_cell(self._c, i, j, C.byref(cc))) == 0:
            return None
        return {k: getattr(cc, k) for k, _ in CoarseCell._fields_}

    def sample(self, x, z):
        gs = GroundSample()
        _check_status(lib().silt_core_sample(self._c, x, z, C.byref(gs)))
        return {k: getattr(gs, k) for k, _ in GroundSample._fields_}

    def ledger_full(self):
        led = FluidLedgerFull()
        _check_status(lib().silt_core_ledger_full(self._c, C.byref(led)))
        return {k: getattr(led, k) for k, _ in FluidLedgerFull._fields_}

    def set_hydro_solids(self, solids):
        """set_hydro_solids (include/silt_core.h): dicts of HydroSolid's
        fields (missing ones 0)."""
        arr = (HydroSolid * max(len(solids), 1))(*[HydroSolid(**{k: s.get(k, 0) for k, _ in HydroSolid._fields_})
                                                   for s in solids])
        _check_status(lib().silt_core_set_hydro_solids(self._c, arr, len(solids)))

    def hydro_load(self):
        out = HydroLoad()
        _check_status(lib().silt_core_hydro_load(self._c, C.byref(out)))
        return {k: getattr(out, k) for k, _ in HydroLoad._fields_}

    def note_surface_foam(self, dt):
        _check_status(lib().silt_core_note_surface_foam(self._c, dt))

    def pond_ruts(self):
        """pond_ruts (src/fluid_core.rs:5509-5611): cubic metres moved into
        the tire cuts."""
        moved = np.zeros(1, np.float32)
        _check_status(lib().silt_core_pond_ruts(self._c, _ptr(moved)))
        return float(moved[0])

    def seat_rut_water(self, pressed, is_rut):
        """seat_rut_water (src/fluid_core.rs:5622-5736) with the classify
        closure sampled per coarse cell: the pond surface, or None."""
        p = np.ascontiguousarray(pressed, np.float32)
        r = np.ascontiguousarray(is_rut, np.uint8)
        eta, has = np.zeros(1, np.float32), np.zeros(1, np.int32)
        _check_status(lib().silt_core_seat_rut_water(self._c, _ptr(p), _ptr(r), _ptr(eta), _ptr(has)))
        return float(eta[0]) if has[0] == 1 else None

    def suspended(self):
        """The wheel-lifted suspended solids (m, 200 x 200) and the FP64
        accounts (entrained_m3, settled_m3)."""
        s, acc = np.zeros(WINDOW_RES * WINDOW_RES, np.float32), np.zeros(2, np.float64)
        _check_status(lib().silt_core_get_suspended(self._c, _ptr(s), _ptr(acc)))
        return s, (float(acc[0]), float(acc[1]))

    def set_suspended(self, susp):
        s = np.ascontiguousarray(susp, np.float32)
        _check_status(lib().silt_core_set_suspended(self._c, _ptr(s)))

    def grain(self):
        out = np.zeros(WINDOW_RES * WINDOW_RES, np.uint8)
        _check_status(lib().silt_core_get_grain(self._c, _ptr(out)))
        return out

    def seat_rut_sheet(self):
        """seat_rut_sheet (src/world.rs:5539-5561) on the core's ground:
        True when the sheet was seated on a pond floor."""
        seated = np.zeros(1, np.int32)
        _check_status(lib().silt_core_seat_rut_sheet(self._c, _ptr(seated)))
        return bool(seated[0])

    def tracers(self):
        """(foam, turbid, (stirred_m3, settled_m3)) of the window."""
        n = WINDOW_RES * WINDOW_RES
        foam, turbid, acc = np.zeros(n, np.float32), np.zeros(n, np.float32), np.zeros(2, np.float64)
        _check_status(lib().silt_core_get_tracers(self._c, _ptr(foam), _ptr(turbid), _ptr(acc)))
        return foam, turbid, (float(acc[0]), float(acc[1]))

    def set_tracers(self, foam=None, turbid=None):
        f = None if foam is None else np.ascontiguousarray(foam, np.float32)
        t = None if turbid is None else np.ascontiguousarray(turbid, np.float32)
        _check_status(lib().silt_core_set_tracers(self._c, None if f is None else _ptr(f),
                                                  None if t is None else _ptr(t)))

    def solid_mask(self, fine=False):
        out = np.zeros((160 if fine else WINDOW_RES) ** 2, np.uint8)
        _check_status(lib().silt_core_solid_mask(self._c, int(fine), _ptr(out)))
        return out

    def ledger_lines(self):
        """The FP64 lines that close the ledger (include/silt_core.h):
        {line: (water, slurry)} for base, cut, residue, handoff, total and
        gap, in m^3."""
        out = CoreLines()
        _check_status(lib().silt_core_ledger_lines(self._c, C.byref(out)))
        return {k: (getattr(out, k)[0], getattr(out, k)[1]) for k, _ in CoreLines._fields_}

    def window_bounds(self):
        out = np.zeros(3, np.float32)
        _check_status(lib().silt_core_window_bounds(self._c, _ptr(out)))
        return tuple(float(v) for v in out)

    def lay_slurry_on_water(self, depth, grain):
        v = lib().silt_core_lay_slurry_on_water(self._c, depth, GRAIN[grain])
        if v != v:
            raise SiltError("silt_core_lay_slurry_on_water failed")
        return float(v)

    def recenter(self, x, z):
        """recenter (src/fluid_core.rs:5409-5422): move the window."""
        _check_status(lib().silt_core_recenter(self._c, x, z))

    def parked_cells(self):
        return int(_check_status(lib().silt_core_parked_cells(self._c)))

    def drain_over_bed(self, dt, steps):
        v = lib().silt_core_drain_over_bed(self._c, dt, int(steps))
        if v != v:
            raise SiltError("silt_core_drain_over_bed failed")
        return float(v)

    def set_tire_model(self, model):
        """"nested" (the reference's 2 cm patch), "coarse" (no patch) or
        "surrogate" (coarse + the tire sub-grid network)."""
        _check_status(lib().silt_core_set_tire_model(self._c, TIRE_MODEL[model]))

    def set_surrogate(self, params, judge_every=30, tol=1.0e-3):
        p = np.ascontiguousarray(params, np.float32)
        if p.size != MLP_PARAMS:
            raise ValueError("parameter vector size")
        _check_status(lib().silt_core_set_surrogate(self._c, _ptr(p), int(judge_every), tol))

    def surrogate_stats(self):
        st = SurrogateStats()
        _check_status(lib().silt_core_surrogate_stats(self._c, C.byref(st)))
        return {k: getattr(st, k) for k, _ in SurrogateStats._fields_}

    def fine_water_debug(self, pushes):
        """fine_water_debug (src/fluid_core/near.rs:341-346): the 2 cm water,
        u, v and origin of the patch laid under ``pushes`` (dicts of
        push_tire's fields); the patch is restricted again afterwards."""
        arr = (TirePush * max(len(pushes), 1))(*[TirePush(**q) for q in pushes])
        fw, fu, fv = (np.zeros(160 * 160, np.float32) for _ in range(3))
        o = np.zeros(2, np.float32)
        _check_status(lib().silt_core_fine_water_debug(self._c, arr, len(pushes), _ptr(fw), _ptr(fu), _ptr(fv),
                                                       _ptr(o)))
        return fw, fu, fv, o

    def set_ground(self, bed, placed=None):
        b = np.ascontiguousarray(bed, np.float32)
        p = None if placed is None else np.ascontiguousarray(placed, np.float32)
        _check_status(lib().silt_core_set_ground(self._c, _ptr(b), None if p is None else _ptr(p)))

    def stamp_ruts(self, contacts, dt):
        L = _rut_lib()
        if not getattr(L, "_core_rut_declared", False):
            L.silt_core_stamp_ruts.argtypes = [_P, C.POINTER(TireContact), _Z, _F]
            L.silt_core_stamp_ruts.restype = C.c_int
            L._core_rut_declared = True
        arr = (TireContact * max(len(contacts), 1))(*[TireContact(**c) for c in contacts])
        _check_status(L.silt_core_stamp_ruts(self._c, arr, len(contacts), dt))

    def ground(self):
        n = WINDOW_RES * WINDOW_RES
        out = [np.empty(n, np.float32) for _ in range(3)]
        _check_status(lib().silt_core_get_ground(self._c, *(_ptr(a) for a in out)))
        return out


class TireContact(C.Structure):
    """silt_tire_contact (include/silt_rut.h)."""
    _fields_ = [(n, C.c_float) for n in ("x", "z", "hub_y", "forward_x", "forward_z", "right_x", "right_z", "radius",
                                         "width", "sinkage", "normal", "push", "k_eq", "n")]


def _declare_rut(L):
    L.silt_bekker_load.argtypes = [_F] * 5
    L.silt_bekker_load.restype = _F
    L.silt_rut_depth.argtypes = [_F] * 6
    L.silt_rut_depth.restype = _F
    L.silt_rut_stamp_device.argtypes = [_P, _P, _P, _Z, _F, _F, _F, C.POINTER(TireContact), _Z, _F, _P]
    L.silt_rut_stamp_device.restype = C.c_int
    L.silt_rut_stamp_device_noalloc.argtypes = [_P, _P, _P, _Z, _F, _F, _F, C.POINTER(TireContact), _Z, _F, _P]
    L.silt_rut_stamp_device_noalloc.restype = C.c_int


def _rut_lib():
    L = lib()
    if not getattr(L, "_rut_declared", False):
        _declare_rut(L)
        L._rut_declared = True
    return L


def bekker_load(sinkage, radius, width, k_eq, n):
    return _rut_lib().silt_bekker_load(sinkage, radius, width, k_eq, n)


def rut_depth(normal, push, radius, width, k_eq, n):
    return _rut_lib().silt_rut_depth(normal, push, radius, width, k_eq, n)


def rut_stamp_device(bed, placed, initial, res, cell, origin, contacts, dt, stream=0):
    """Stamp contacts (dicts of TireContact fields) into device planes given
    as integer addresses."""
    arr = (TireContact * len(contacts))(*[TireContact(**c) for c in contacts])
    return _check_status(_rut_lib().silt_rut_stamp_device(bed, placed, initial, res, cell, origin[0], origin[1], arr,
                                                          len(contacts), dt, stream or None))


def rut_stamp_device_noalloc(bed, placed, initial, res, cell, origin, contacts, dt, stream=0):
    """rut_stamp_device without any memory allocation (at most
    NOALLOC_MAX_CONTACTS contacts; the library's one array per GPU is shared
    by every call)."""
    arr = (TireContact * max(len(contacts), 1))(*[TireContact(**c) for c in contacts])
    return _check_status(_rut_lib().silt_rut_stamp_device_noalloc(bed, placed, initial, res, cell, origin[0],
                                                                  origin[1], arr, len(contacts), dt, stream or None))


MLP_IN, MLP_HIDDEN, MLP_OUT, MLP_PARAMS = 1536, 128, 8, 214280


class Mlp:
    """The tire sub-grid network on its own (include/silt_mlp.h)."""

    def __init__(self, seed=1):
        self._h = lib().silt_mlp_create(int(seed))
        if not self._h:
            raise SiltError("silt_mlp_create failed")

    def close(self):
        if getattr(self, "_h", None):
            lib().silt_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.empty(MLP_PARAMS, np.float32)
        _check_status(lib().silt_mlp_get_params(self._h, _ptr(out)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != MLP_PARAMS:
            raise ValueError("parameter vector size")
        _check_status(lib().silt_mlp_set_params(self._h, _ptr(p)))

    @staticmethod
    def _x(x):
        x = np.ascontiguousarray(x, np.float32)
        if x.ndim != 2 or x.shape[1] != MLP_IN:
            raise ValueError("x must be n x 1536")
        return x

    def forward(self, x):
        x = self._x(x)
        y = np.empty((x.shape[0], MLP_OUT), np.float32)
        _check_status(lib().silt_mlp_forward(self._h, _ptr(x), x.shape[0], _ptr(y)))
        return y

    def infer(self, x):
        x = self._x(x)
        y = np.empty((x.shape[0], MLP_OUT), np.float32)
        _check_status(lib().silt_mlp_infer(self._h, _ptr(x), x.shape[0], _ptr(y)))
        return y

    def gradients(self, x, t):
        x = self._x(x)
        t = np.ascontiguousarray(t, np.float32).reshape(x.shape[0], MLP_OUT)
        g = np.empty(MLP_PARAMS, np.float32)
        sse = C.c_double()
        _check_status(lib().silt_mlp_gradients(self._h, _ptr(x), _ptr(t), x.shape[0], C.byref(sse), _ptr(g)))
        return sse.value, g

    def train_batch(self, x, t, lr=1e-3):
        x = self._x(x)
        t = np.ascontiguousarray(t, np.float32).reshape(x.shape[0], MLP_OUT)
        sse = C.c_double()
        _check_status(lib().silt_mlp_train_batch(self._h, _ptr(x), _ptr(t), x.shape[0], lr, C.byref(sse)))
        return sse.value


class Trainer:
    """Training set and trainer of the tire sub-grid network
    (include/silt_mlp.h): records nested steps of a FluidCore beside a
    coarse-only twin and fits the network on the GPU."""

    def __init__(self, max_steps=4096, seed=1):
        L = lib()
        if not getattr(L, "_core_declared", False):
            _declare_core(L)
            L._core_declared = True
        self._t = L.silt_trainer_create(int(max_steps), int(seed))
        if not self._t:
            raise SiltError("silt_trainer_create failed")

    def close(self):
        if getattr(self, "_t", None):
            lib().silt_trainer_destroy(self._t)
            self._t = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def record(self, core, dt, max_events=4096):
        ev = (SprayEvent * max_events)()
        n = C.c_size_t(0)
        return int(_check_status(lib().silt_trainer_record(self._t, core._c, dt, ev, max_events, C.byref(n))))

    def steps(self):
        return int(_check_status(lib().silt_trainer_steps(self._t)))

    def samples(self):
        return int(_check_status(lib().silt_trainer_samples(self._t)))

    def fit(self, iterations, batch=512, lr=1e-3, train_steps=0):
        mse = C.c_double()
        _check_status(lib().silt_trainer_fit(self._t, int(train_steps), int(iterations), int(batch), lr,
                                             C.byref(mse)))
        return mse.value

    def evaluate(self, first, count):
        a, b = C.c_double(), C.c_double()
        _check_status(lib().silt_trainer_eval(self._t, int(first), int(count), C.byref(a), C.byref(b)))
        return a.value, b.value

    def params(self):
        out = np.empty(MLP_PARAMS, np.float32)
        _check_status(lib().silt_trainer_params(self._t, _ptr(out)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _check_status(lib().silt_trainer_set_params(self._t, _ptr(p)))


def surrogate_targets(hn_w, hc_w, hn_s, hc_s, region):
    """Training targets of a nested / coarse-only pair of depth planes over a
    region (i0, j0, nx, nz): [cell][E, W, N, S water, E, W, N, S slurry] mm."""
    L = lib()
    if not getattr(L, "_core_declared", False):
        _declare_core(L)
        L._core_declared = True
    arrs = [np.ascontiguousarray(a, np.float32) for a in (hn_w, hc_w, hn_s, hc_s)]
    i0, j0, nx, nz = region
    out = np.empty((nx * nz, 8), np.float32)
    _check_status(L.silt_surrogate_targets_debug(*(_ptr(a) for a in arrs), i0, j0, nx, nz, _ptr(out)))
    return out


def surrogate_apply(water, slurry, y, region):
    """Apply predicted edge fluxes y ([cell][8] mm) over a region; returns the
    new water and slurry planes."""
    L = lib()
    if not getattr(L, "_core_declared", False):
        _declare_core(L)
        L._core_declared = True
    w, s_ = (np.array(a, np.float32, copy=True) for a in (water, slurry))
    yy = np.ascontiguousarray(y, np.float32)
    i0, j0, nx, nz = region
    _check_status(L.silt_surrogate_apply_debug(_ptr(w), _ptr(s_), _ptr(yy), i0, j0, nx, nz))
    return w, s_
