"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py load this module as the checker.  The HIP
product path never imports it.

Every function mirrors a reference function of src/fluid_core.rs (cited on the
C side in oracle/sheet.c) and works in place on numpy arrays, like the
reference's ``&mut [f32]`` slices.  ``precision`` 32 is the reference's FP32
arithmetic restated; 64 is the same algorithm in FP64.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
PAR_LIB_PATH = os.path.join(HERE, "liboracle_par.so")

CONFINE = {"reference": 0, "jacobi": 1, "gauss_seidel": 2}
GRAIN = {"clay": 0, "silt": 1, "sand": 2, "water": 3, "other": 4}

_lib = None
_lib_par = None


def build() -> None:
    """Compile oracle/liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


_P = C.c_void_p
_Z = C.c_size_t


def lib_par():
    """The CPU baseline build (oracle/liboracle_par.so): the FP32 step at AVX2
    width in row stripes of at most 4 threads, water || slurry."""
    global _lib_par
    if _lib_par is None:
        if not os.path.exists(PAR_LIB_PATH):
            build()
        _lib_par = C.CDLL(PAR_LIB_PATH)
        f = _lib_par.oracle_step_pair_par_f32
        R = C.c_float
        f.argtypes = [_P, _P, _P, _P, R, _P, _P, _P, _P, R, _P, _P, _Z, R, R]
        f.restype = C.c_int
    return _lib_par


def step_pair_par(water, slurry, bed, nu_w, nu_s, gate, res, cell, dt, vort_w=0.08, vort_s=0.05):
    """One two-layer step on the CPU baseline build: water = (h, u, v) and
    slurry = (h, u, v) FP32 planes in place, each a step_sheet with its own
    viscosity plane and vorticity, the two layers on two threads."""
    n = res * res
    for a, nm in ((bed, "bed"), (nu_w, "nu_w"), (nu_s, "nu_s")):
        _check(a, np.float32, n, nm)
    for a in tuple(water) + tuple(slurry):
        _check(a, np.float32, n, "plane")
    _check(gate, np.uint8, n, "gate")
    return lib_par().oracle_step_pair_par_f32(*(_ptr(a) for a in water), _ptr(nu_w), vort_w,
                                              *(_ptr(a) for a in slurry), _ptr(nu_s), vort_s, _ptr(bed),
                                              _ptr(gate), res, cell, dt)


def _declare(L):
    for sfx, R in (("_f32", C.c_float), ("_f64", C.c_double)):
        f = getattr(L, "oracle_step_sheet_ex" + sfx)
        f.argtypes = [_P, _P, _P, _P, _P, _P, _Z, R, R, C.c_int, R, R, R, R, C.c_int, R,
                      C.c_int, C.c_int, C.c_int, _P, _P]
        f.restype = C.c_int
        for name in ("oracle_transport_open_ref", "oracle_transport_open_rows"):
            f = getattr(L, name + sfx)
            f.argtypes = [_P, _P, _P, _Z, R, R, _Z, _Z, _Z, _Z, C.c_int, _P, R, _P, _P]
            f.restype = None
        for name in ("oracle_sheet_work_boxes", "oracle_sheet_work_boxes_flood"):
            f = getattr(L, name + sfx)
            f.argtypes = [_P, _Z, R, C.c_int, _P, _Z]
            f.restype = _Z
        f = getattr(L, "oracle_awake_work_boxes" + sfx)
        f.argtypes = [_P, _P, _P, _P, _Z, R, R, _P, _Z]
        f.restype = _Z
        f = getattr(L, "oracle_dynamic_viscosity" + sfx)
        f.argtypes = [C.c_int, R]
        f.restype = R
        for name in ("oracle_sheet_mobility", "oracle_sheet_damp"):
            f = getattr(L, name + sfx)
            f.argtypes = [R]
            f.restype = R
        f = getattr(L, "oracle_column_sum" + sfx)
        f.argtypes = [_P, _Z, R]
        f.restype = R
        f = getattr(L, "oracle_surface_pressure" + sfx)
        f.argtypes = [_P, _P, _Z, _Z, _Z, _Z, _Z, _P]
        f.restype = None
        f = getattr(L, "oracle_force_open" + sfx)
        f.argtypes = [_P, _P, _P, _P, _P, _Z, R, R, R, R, _Z, _Z, _Z, _Z, C.c_int, C.c_int]
        f.restype = None
        f = getattr(L, "oracle_soften_border" + sfx)
        f.argtypes = [_P, _P, C.c_int]
        f.restype = None
        f = getattr(L, "oracle_mlp_forward" + sfx)
        f.argtypes = [_P, _P, _Z, C.c_int, _P, _P, _P, _P, _P]
        f.restype = None
        f = getattr(L, "oracle_step_sheet_walled" + sfx)
        f.argtypes = [_P, _P, _P, _P, _P, _P, _P, _Z, R, R, C.c_int, R, R, R, R, R, C.c_int]
        f.restype = C.c_int
        f = getattr(L, "oracle_lift_solid_surface" + sfx)
        f.argtypes = [_P, _P, _P, _Z, _Z, _Z, _Z, _Z, R, _P]
        f.restype = None
        f = getattr(L, "oracle_noslip_ghost" + sfx)
        f.argtypes = [_P, _P, _P, _P, _Z, _Z, _Z, _Z, _Z, R, _P, _P]
        f.restype = None
        f = getattr(L, "oracle_block_into_solid" + sfx)
        f.argtypes = [_P, _P, _P, _P, _Z, _Z, _Z, _Z, _Z, R]
        f.restype = None
        f = getattr(L, "oracle_displace_occupied" + sfx)
        f.argtypes = [_P, _P, _P, _P, _Z, _P]
        f.restype = None
        f = getattr(L, "oracle_redirect_into_walls" + sfx)
        f.argtypes = [_P, _P, _P, _P, _P, _Z, R, R, _Z, _Z, _Z, _Z, _P, _P]
        f.restype = None
        f = getattr(L, "oracle_transport_walled" + sfx)
        f.argtypes = [_P, _P, _P, _P, _P, _P, _Z, R, R, _Z, _Z, _Z, _Z, C.c_int, _P, _P]
        f.restype = None
        f = getattr(L, "oracle_mlp_gradients" + sfx)
        f.argtypes = [_P, _P, _P, _Z, C.c_int, _P]
        f.restype = C.c_double


def _dt(precision):
    return np.float32 if precision == 32 else np.float64


def _sfx(precision):
    if precision not in (32, 64):
        raise ValueError("precision must be 32 or 64")
    return "_f32" if precision == 32 else "_f64"


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _check(a, dtype, n, name):
    if not isinstance(a, np.ndarray) or a.dtype != dtype or not a.flags.c_contiguous or a.size < n:
        raise TypeError(f"{name}: need a C-contiguous {np.dtype(dtype).name} array of >= {n} items")


def step_sheet(h, u, v, bed, nu, gate, res, cell, dt, drive=("gravity", 0.0), *, eps=1.0e-6,
               banks=False, sub_cap=8.0, pace=0, inner_par=True, confine="reference",
               precision=32, want_boxes=False):
    """Reference ``step_sheet_eps`` (src/fluid_core.rs:856-889) on numpy planes.

    ``drive`` is ``("gravity", vorticity)`` or ``("wind", x, z, vorticity)``.
    ``inner_par`` False is the layer stepped beside the other one inside
    ``FluidCore::step_pair`` (row splits off, so the coarse confinement is the
    in-place Gauss-Seidel sweep).  Returns the substep count (and the boxes).
    """
    dtype = _dt(precision)
    n = res * res
    for a, nm in ((h, "h"), (u, "u"), (v, "v"), (bed, "bed"), (nu, "nu")):
        _check(a, dtype, n, nm)
    _check(gate, np.uint8, n, "gate")
    if drive[0] == "gravity":
        kind, wx, wz, vort = 0, 0.0, 0.0, float(drive[1])
    elif drive[0] == "wind":
        kind, wx, wz, vort = 1, float(drive[1]), float(drive[2]), float(drive[3])
    else:
        raise ValueError(drive)
    boxes = np.zeros(32, dtype=np.uintp)
    nb = C.c_size_t(0)
    f = getattr(lib(), "oracle_step_sheet_ex" + _sfx(precision))
    nsub = f(_ptr(h), _ptr(u), _ptr(v), _ptr(bed), _ptr(nu), _ptr(gate), res, cell, dt, kind, wx,
             wz, vort, eps, int(banks), sub_cap, int(pace), int(inner_par), CONFINE[confine],
             _ptr(boxes), C.byref(nb))
    if want_boxes:
        k = min(nb.value, 8)
        return nsub, [tuple(int(x) for x in boxes[4 * i:4 * i + 4]) for i in range(k)]
    return nsub


def step_sheet_walled(h, u, v, bed, nu, gate, solid, res, cell, dt, drive=("gravity", 0.0), *, eps=1.0e-6,
                      sub_cap=8.0, inner_par=True, precision=32):
    """The walled sheet: ``step_sheet_eps_in`` with a wall mask
    (src/fluid_core.rs:856-889, substep :1879-2082 with solid = Some).
    ``solid`` is a uint8 plane, nonzero = wall.  Returns the substep count."""
    dtype = _dt(precision)
    n = res * res
    for a, nm in ((h, "h"), (u, "u"), (v, "v"), (bed, "bed"), (nu, "nu")):
        _check(a, dtype, n, nm)
    _check(gate, np.uint8, n, "gate")
    _check(solid, np.uint8, n, "solid")
    if drive[0] == "gravity":
        kind, wx, wz, vort = 0, 0.0, 0.0, float(drive[1])
    elif drive[0] == "wind":
        kind, wx, wz, vort = 1, float(drive[1]), float(drive[2]), float(drive[3])
    else:
        raise ValueError(drive)
    f = getattr(lib(), "oracle_step_sheet_walled" + _sfx(precision))
    return f(_ptr(h), _ptr(u), _ptr(v), _ptr(bed), _ptr(nu), _ptr(gate), _ptr(solid), res, cell, dt, kind, wx, wz,
             vort, eps, sub_cap, int(inner_par))


def _box_all(res, box):
    return (0, res - 1, 0, res - 1) if box is None else box


def lift_solid_surface(eta, h, solid, res, box=None, eps=1.0e-6, precision=32):
    """lift_solid_surface (src/fluid_core.rs:4225-4276) in place on eta."""
    tmp = np.zeros(res * res, _dt(precision))
    getattr(lib(), "oracle_lift_solid_surface" + _sfx(precision))(_ptr(eta), _ptr(h), _ptr(solid), res,
                                                                    *_box_all(res, box), eps, _ptr(tmp))


def noslip_ghost(u, v, h, solid, res, box=None, eps=1.0e-6, precision=32):
    """write_noslip_ghost (src/fluid_core.rs:4278-4326) in place."""
    gu, gv = np.zeros(res * res, _dt(precision)), np.zeros(res * res, _dt(precision))
    getattr(lib(), "oracle_noslip_ghost" + _sfx(precision))(_ptr(u), _ptr(v), _ptr(h), _ptr(solid), res,
                                                              *_box_all(res, box), eps, _ptr(gu), _ptr(gv))


def block_into_solid(u, v, h, solid, res, box=None, eps=1.0e-6, precision=32):
    """block_into_solid (src/fluid_core.rs:4328-4380) in place."""
    getattr(lib(), "oracle_block_into_solid" + _sfx(precision))(_ptr(u), _ptr(v), _ptr(h), _ptr(solid), res,
                                                                  *_box_all(res, box), eps)


def displace_occupied(h, u, v, solid, res, precision=32):
    """displace_occupied (src/fluid_core.rs:4382-4444) in place."""
    give = np.zeros(res * res, _dt(precision))
    getattr(lib(), "oracle_displace_occupied" + _sfx(precision))(_ptr(h), _ptr(u), _ptr(v), _ptr(solid), res,
                                                                   _ptr(give))


def redirect_into_walls(h, u, v, solid, bed, res, cell, dt, box=None, precision=32):
    """redirect_into_walls (src/fluid_core.rs:4446-4563) in place on h."""
    rem, add = np.zeros(res * res, _dt(precision)), np.zeros(res * res, _dt(precision))
    getattr(lib(), "oracle_redirect_into_walls" + _sfx(precision))(_ptr(h), _ptr(u), _ptr(v), _ptr(solid),
                                                                     _ptr(bed), res, cell, dt, *_box_all(res, box),
                                                                     _ptr(rem), _ptr(add))


def transport_walled(h, u, v, gate, solid, bed, res, cell, dt, along_x, box=None, precision=32):
    """transport with solids (src/fluid_core.rs:3894-3977) in place on h."""
    rem, add = np.zeros(res * res, _dt(precision)), np.zeros(res * res, _dt(precision))
    getattr(lib(), "oracle_transport_walled" + _sfx(precision))(_ptr(h), _ptr(u), _ptr(v), _ptr(gate), _ptr(solid),
                                                                  _ptr(bed), res, cell, dt, *_box_all(res, box),
                                                                  int(along_x), _ptr(rem), _ptr(add))


def surface_pressure(h, bed, res, box, precision=32):
    """The substep's pressure plane d = 0.45 (0.62 G5(eta) + 0.38 S5(eta)) +
    0.55 eta on box (i0, i1, j0, j1) (src/fluid_core.rs:1903-1935); zero
    outside the box."""
    dtype = _dt(precision)
    n = res * res
    _check(h, dtype, n, "h")
    _check(bed, dtype, n, "bed")
    d = np.zeros(n, dtype)
    getattr(lib(), "oracle_surface_pressure" + _sfx(precision))(_ptr(h), _ptr(bed), res, *box, _ptr(d))
    return d


def force_open(h, u, v, nu, p, res, cell, dt, vorticity, box, eps=1.0e-6, inner_par=True, confine="reference",
               precision=32):
    """Gravity kick from pressure plane p and the confinement sweep on the wet
    cells of box (force_gravity_open and the Jacobi sweeps,
    src/fluid_core.rs:2204-2595), in place on u, v."""
    dtype = _dt(precision)
    n = res * res
    for a, nm in ((h, "h"), (u, "u"), (v, "v"), (nu, "nu"), (p, "p")):
        _check(a, dtype, n, nm)
    getattr(lib(), "oracle_force_open" + _sfx(precision))(_ptr(h), _ptr(u), _ptr(v), _ptr(nu), _ptr(p), res, cell,
                                                          dt, vorticity, eps, *box, int(inner_par),
                                                          CONFINE[confine])


def soften_border(u, v, res, precision=32):
    """soften_border (src/fluid_core.rs:9594-9620) with no open reach, in
    place."""
    dtype = _dt(precision)
    _check(u, dtype, res * res, "u")
    _check(v, dtype, res * res, "v")
    getattr(lib(), "oracle_soften_border" + _sfx(precision))(_ptr(u), _ptr(v), res)


def transport(h, vel, gate, res, cell, dt, i0, i1, j0, j1, along_x, bed=None, frac_cap=0.45,
              form="rows", precision=32):
    """One open transport pass; returns (removed, added) scratch planes.

    ``form`` "serial" is the reference's edge loop (transport_open_ref),
    "rows" the branch-free row form (transport_open_rows_body).
    """
    dtype = _dt(precision)
    n = res * res
    removed = np.zeros(n, dtype=dtype)
    added = np.zeros(n, dtype=dtype)
    name = "oracle_transport_open_ref" if form == "serial" else "oracle_transport_open_rows"
    f = getattr(lib(), name + _sfx(precision))
    f(_ptr(h), _ptr(vel), _ptr(gate), res, cell, dt, i0, i1, j0, j1, int(along_x), _ptr(bed),
      frac_cap, _ptr(removed), _ptr(added))
    return removed, added


def work_boxes(h, res, eps, allow_split=True, form="runs", precision=32):
    out = np.zeros(4 * 4096, dtype=np.uintp)
    name = "oracle_sheet_work_boxes" if form == "runs" else "oracle_sheet_work_boxes_flood"
    f = getattr(lib(), name + _sfx(precision))
    k = f(_ptr(h), res, eps, int(allow_split), _ptr(out), 4096)
    return [tuple(int(x) for x in out[4 * i:4 * i + 4]) for i in range(min(k, 4096))]


def awake_boxes(h, u, v, bed, res, cell, eps, precision=32):
    out = np.zeros(4 * 4096, dtype=np.uintp)
    f = getattr(lib(), "oracle_awake_work_boxes" + _sfx(precision))
    k = f(_ptr(h), _ptr(u), _ptr(v), _ptr(bed), res, cell, eps, _ptr(out), 4096)
    return [tuple(int(x) for x in out[4 * i:4 * i + 4]) for i in range(min(k, 4096))]


def dynamic_viscosity(grain, moisture, precision=32):
    g = GRAIN[grain] if isinstance(grain, str) else int(grain)
    return getattr(lib(), "oracle_dynamic_viscosity" + _sfx(precision))(g, moisture)


def sheet_mobility(nu, precision=32):
    return getattr(lib(), "oracle_sheet_mobility" + _sfx(precision))(nu)


def sheet_damp(nu, precision=32):
    return getattr(lib(), "oracle_sheet_damp" + _sfx(precision))(nu)


def column_sum(h, cell, precision=32):
    return getattr(lib(), "oracle_column_sum" + _sfx(precision))(_ptr(h), h.size, cell)


# --- FluidCore restatement (oracle/core.c) ------------------------------------

WINDOW_RES = 200
FINE_RES = 160


def _declare_core(L):
    for sfx, R in (("_f32", C.c_float), ("_f64", C.c_double)):
        g = lambda n: getattr(L, n + sfx)  # noqa: E731
        g("oracle_core_new").argtypes = [R, R]
        g("oracle_core_new").restype = _P
        g("oracle_core_free").argtypes = [_P]
        for n, at in (("oracle_core_set_fine_pace", [_P, C.c_int]), ("oracle_core_set_fine_sub_cap", [_P, R]),
                      ("oracle_core_set_coarse_confine", [_P, C.c_int]), ("oracle_core_set_near_field", [_P, C.c_int]),
                      ("oracle_core_push_tire", [_P, _P]), ("oracle_core_map_water_bed", [_P, _P, _P]),
                      ("oracle_core_paint", [_P, R, R, R, R, R, C.c_int, C.c_int, R]),
                      ("oracle_core_lay_slurry", [_P, _P, C.c_int, R]),
                      ("oracle_core_read_substrate", [_P, _P, _P, _P]), ("oracle_core_get", [_P, _P]),
                      ("oracle_core_set", [_P, _P]), ("oracle_core_origin", [_P, _P]),
                      ("oracle_core_ledger", [_P, _P]), ("oracle_core_publish", [_P, _P, _P]),
                      ("oracle_core_publish_surface", [_P, _P, _P, _P, _P, _P]),
                      ("oracle_core_refine", [_P, R]), ("oracle_core_restrict", [_P]),
                      ("oracle_core_fine_get", [_P, _P, _P]), ("oracle_recentre", [_P]),
                      ("oracle_core_fine_water_debug", [_P, _P, C.c_int, _P, _P, _P, _P])):
            g(n).argtypes = at
            g(n).restype = None
        g("oracle_core_step").argtypes = [_P, R, _P]
        g("oracle_core_step").restype = C.c_int
        g("oracle_core_drain_over_bed").argtypes = [_P, R, C.c_uint]
        g("oracle_core_drain_over_bed").restype = R
        g("oracle_core_recenter").argtypes = [_P, R, R]
        g("oracle_core_recenter").restype = None
        g("oracle_core_parked_cells").argtypes = [_P]
        g("oracle_core_parked_cells").restype = C.c_long
        g("oracle_core_near_tiles").argtypes = [_P]
        g("oracle_core_near_tiles").restype = _Z
        g("oracle_core_fine_on").argtypes = [_P]
        g("oracle_core_fine_on").restype = C.c_int
        g("oracle_floor_scale").argtypes = [R, _P]
        g("oracle_floor_scale").restype = R
        g("oracle_core_push_on").argtypes = [_P, _P, C.c_int]
        g("oracle_core_push_on").restype = None
        g("oracle_core_peel").argtypes = [_P, C.c_int, C.POINTER(C.c_int)]
        g("oracle_core_peel").restype = R
        g("oracle_breaking_foam").argtypes = [R, R]
        g("oracle_breaking_foam").restype = R
        g("oracle_sample_foam").argtypes = [_P, C.c_int, R, R]
        g("oracle_sample_foam").restype = R
        g("oracle_core_pond_ruts").argtypes = [_P]
        g("oracle_core_pond_ruts").restype = R
        g("oracle_core_seat_rut_water").argtypes = [_P, _P, _P, _P]
        g("oracle_core_seat_rut_water").restype = C.c_int
        g("oracle_core_grain").argtypes = [_P, _P]
        g("oracle_core_grain").restype = None
        for n, at in (("oracle_core_entrain", [_P, R]), ("oracle_core_suspended_step", [_P, R]),
                      ("oracle_core_set_susp", [_P, _P]), ("oracle_core_susp", [_P, _P, _P])):
            g(n).argtypes = at
            g(n).restype = None
        for n, at in (("oracle_core_set_hydro_solids", [_P, _P, C.c_int]),
                      ("oracle_core_solid_mask", [_P, C.c_int, _P]), ("oracle_core_hydro_load", [_P, _P]),
                      ("oracle_core_bed_drop", [_P, _P]), ("oracle_core_scour", [_P, R]),
                      ("oracle_core_note_surface_foam", [_P, R]), ("oracle_core_tracers", [_P, _P, _P, _P]),
                      ("oracle_core_set_tracers", [_P, _P, _P]), ("oracle_core_tracer_step", [_P, R]),
                      ("oracle_shore_cell", [_P, _P, _P, C.c_int, C.c_int, _P])):
            g(n).argtypes = at
            g(n).restype = None


def breaking_foam(speed, thickness, precision=32):
    """breaking_foam (src/fluid_core.rs:9676-9686)."""
    _declare_core_once()
    return getattr(lib(), "oracle_breaking_foam" + _sfx(precision))(speed, thickness)


def sample_foam(src, res, x, z, precision=32):
    """sample_foam (src/fluid_core.rs:9693-9714) at index-space (x, z)."""
    _declare_core_once()
    return getattr(lib(), "oracle_sample_foam" + _sfx(precision))(_ptr(src), res, x, z)


def shore_cell(water, uw, vw, res, idx, precision=32):
    """shore_cell (src/fluid_core.rs:9826-9874): (foam production, silt rate)."""
    _declare_core_once()
    out = np.zeros(2, _dt(precision))
    getattr(lib(), "oracle_shore_cell" + _sfx(precision))(_ptr(water), _ptr(uw), _ptr(vw), res, idx, _ptr(out))
    return float(out[0]), float(out[1])


def _declare_core_once():
    L = lib()
    if not getattr(L, "_core_declared", False):
        _declare_core(L)
        L._core_declared = True


SOLID_FIELDS = ("x", "z", "yaw", "half_length", "half_width", "bottom_y", "top_y", "contact", "radius", "width")
HYDRO_LOAD_FIELDS = ("force_x", "force_z", "buoyancy_n", "hull_buoyancy_n", "wheel_buoyancy_n", "displaced_m3",
                     "eroded_m3", "hole_m", "wheel_blocks", "belly_blocks")


class Core:
    """FluidCore on the sheet path (oracle/core.c).  Closures of the reference
    API are arrays sampled at coarse cell centres."""

    def __init__(self, center=(0.0, 0.0), precision=32):
        self.precision = precision
        self.dtype = _dt(precision)
        self.sfx = _sfx(precision)
        L = lib()
        if not getattr(L, "_core_declared", False):
            _declare_core(L)
            L._core_declared = True
        self._c = self._f("oracle_core_new")(center[0], center[1])

    def _f(self, name):
        return getattr(lib(), name + self.sfx)

    def __del__(self):
        try:
            self._f("oracle_core_free")(self._c)
        except Exception:
            pass

    def _arr(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        return a

    def origin(self):
        o = np.zeros(2, self.dtype)
        self._f("oracle_core_origin")(self._c, _ptr(o))
        return o

    def cell_centres(self):
        o = self.origin().astype(np.float64)
        c = (np.arange(WINDOW_RES) + 0.5) * 0.2
        return np.meshgrid((o[0] + c).astype(self.dtype), (o[1] + c).astype(self.dtype))

    def set_fine_pace(self, pace):
        self._f("oracle_core_set_fine_pace")(self._c, pace)

    def set_fine_sub_cap(self, cap):
        self._f("oracle_core_set_fine_sub_cap")(self._c, cap)

    def set_coarse_confine(self, mode):
        self._f("oracle_core_set_coarse_confine")(self._c, CONFINE[mode])

    def set_near_field(self, on):
        self._f("oracle_core_set_near_field")(self._c, int(on))

    def pour_water(self, x0, x1, z0, z1, depth):
        self._f("oracle_core_paint")(self._c, x0, x1, z0, z1, depth, 1, 3, 1.0)

    def pour_slurry(self, x0, x1, z0, z1, depth, grain, moisture):
        g = GRAIN[grain] if isinstance(grain, str) else int(grain)
        self._f("oracle_core_paint")(self._c, x0, x1, z0, z1, depth, 0, g, moisture)

    def map_water_bed(self, bed, water):
        b, w = self._arr(bed), self._arr(water)
        self._f("oracle_core_map_water_bed")(self._c, _ptr(b), _ptr(w))

    def set_hydro_solids(self, solids):
        """set_hydro_solids (src/fluid_core.rs:5257-5260): dicts of
        HydroSolid's fields."""
        a = np.array([[float(s.get(k, 0.0)) for k in SOLID_FIELDS] for s in solids] or [[0.0] * 10], self.dtype)
        self._f("oracle_core_set_hydro_solids")(self._c, _ptr(a), len(solids))

    def solid_mask(self, fine=False):
        n = (160 if fine else WINDOW_RES) ** 2
        out = np.zeros(n, np.uint8)
        self._f("oracle_core_solid_mask")(self._c, int(fine), _ptr(out))
        return out

    def hydro_load(self):
        out = np.zeros(10, self.dtype)
        self._f("oracle_core_hydro_load")(self._c, _ptr(out))
        return dict(zip(HYDRO_LOAD_FIELDS, (float(x) for x in out)))

    def tracer_step(self, dt):
        """step_tracers alone (src/fluid_core.rs:6210-6239) with the pushes
        given since the last step."""
        self._f("oracle_core_tracer_step")(self._c, dt)

    def note_surface_foam(self, dt):
        self._f("oracle_core_note_surface_foam")(self._c, dt)

    def tracers(self):
        """(foam, turbid, stirred_m3, settled_m3) of the window."""
        foam, turbid = np.zeros(WINDOW_RES ** 2, self.dtype), np.zeros(WINDOW_RES ** 2, self.dtype)
        acc = np.zeros(2, self.dtype)
        self._f("oracle_core_tracers")(self._c, _ptr(foam), _ptr(turbid), _ptr(acc))
        return foam, turbid, float(acc[0]), float(acc[1])

    def set_tracers(self, foam=None, turbid=None):
        f = None if foam is None else self._arr(foam)
        t = None if turbid is None else self._arr(turbid)
        self._f("oracle_core_set_tracers")(self._c, _ptr(f), _ptr(t))

    def pond_ruts(self):
        """pond_ruts (src/fluid_core.rs:5509-5611): cubic metres moved into
        the tire cuts."""
        return float(self._f("oracle_core_pond_ruts")(self._c))

    def seat_rut_water(self, pressed, is_rut):
        """seat_rut_water (src/fluid_core.rs:5622-5736) with the classify
        closure sampled per coarse cell: the pond surface, or None."""
        p = self._arr(pressed)
        r = np.ascontiguousarray(is_rut, np.uint8)
        eta = np.zeros(1, self.dtype)
        ok = self._f("oracle_core_seat_rut_water")(self._c, _ptr(p), _ptr(r), _ptr(eta))
        return float(eta[0]) if ok else None

    def entrain(self, dt):
        """entrain_wake (src/fluid_core.rs:8758-8808) of the pushes given
        with push_tire, alone."""
        self._f("oracle_core_entrain")(self._c, dt)

    def suspended_step(self, dt):
        """step_suspended (src/fluid_core.rs:8750-9003) alone."""
        self._f("oracle_core_suspended_step")(self._c, dt)

    def set_suspended(self, susp):
        s = self._arr(susp)
        self._f("oracle_core_set_susp")(self._c, _ptr(s))

    def suspended(self):
        """(plane, {column_m3, settled_m3, frozen_m3, entrained_m3})."""
        s, acc = np.zeros(WINDOW_RES * WINDOW_RES, self.dtype), np.zeros(4, self.dtype)
        self._f("oracle_core_susp")(self._c, _ptr(s), _ptr(acc))
        return s, dict(zip(("column_m3", "settled_m3", "frozen_m3", "entrained_m3"), (float(x) for x in acc)))

    def grain(self):
        out = np.zeros(WINDOW_RES * WINDOW_RES, np.uint8)
        self._f("oracle_core_grain")(self._c, _ptr(out))
        return out

    def scour(self, dt):
        self._f("oracle_core_scour")(self._c, dt)

    def bed_drop(self):
        out = np.zeros(WINDOW_RES * WINDOW_RES, self.dtype)
        self._f("oracle_core_bed_drop")(self._c, _ptr(out))
        return out

    def lay_slurry(self, depth, grain, moisture):
        d = self._arr(depth)
        g = GRAIN[grain] if isinstance(grain, str) else int(grain)
        self._f("oracle_core_lay_slurry")(self._c, _ptr(d), g, moisture)

    def read_substrate(self, bed, moisture, grain):
        b, m = self._arr(bed), self._arr(moisture)
        g = np.ascontiguousarray(grain, np.uint8)
        self._f("oracle_core_read_substrate")(self._c, _ptr(b), _ptr(m), _ptr(g))

    def push_tire(self, x, z, forward_x, forward_z, radius, width, speed, slip):
        p = np.array([x, z, forward_x, forward_z, radius, width, speed, slip], self.dtype)
        self._f("oracle_core_push_tire")(self._c, _ptr(p))

    def push_on(self, x, z, forward_x, forward_z, radius, width, speed, slip=0.0, fine=False):
        """push_on (src/fluid_core.rs:8169-8239) of one push on the coarse
        (or fine) planes, nothing else."""
        p = np.array([x, z, forward_x, forward_z, radius, width, speed, slip], self.dtype)
        self._f("oracle_core_push_on")(self._c, _ptr(p), int(fine))

    def peel(self, fine=False):
        """peel_spray (src/fluid_core.rs:8241-8314) of the pushes given with
        push_tire: (volume peeled, event count)."""
        ev = C.c_int(0)
        vol = self._f("oracle_core_peel")(self._c, int(fine), C.byref(ev))
        return float(vol), ev.value

    def step(self, dt):
        vol = np.zeros(1, self.dtype)
        ev = self._f("oracle_core_step")(self._c, dt, _ptr(vol))
        return ev, float(vol[0])

    def planes(self):
        out = np.zeros(8 * WINDOW_RES * WINDOW_RES, self.dtype)
        self._f("oracle_core_get")(self._c, _ptr(out))
        n = WINDOW_RES * WINDOW_RES
        names = ("water", "slurry", "uw", "vw", "us", "vs", "bed", "moisture")
        return {k: out[i * n:(i + 1) * n].copy() for i, k in enumerate(names)}

    def set_planes(self, planes):
        n = WINDOW_RES * WINDOW_RES
        buf = np.zeros(8 * n, self.dtype)
        for i, k in enumerate(("water", "slurry", "uw", "vw", "us", "vs", "bed", "moisture")):
            buf[i * n:(i + 1) * n] = planes[k]
        self._f("oracle_core_set")(self._c, _ptr(buf))

    def ledger(self):
        out = np.zeros(6, self.dtype)
        self._f("oracle_core_ledger")(self._c, _ptr(out))
        keys = ("water_m3", "slurry_m3", "frozen_water_m3", "frozen_slurry_m3", "spray_m3", "total_m3")
        return {k: float(v) for k, v in zip(keys, out)}

    def publish(self):
        n = WINDOW_RES * WINDOW_RES
        film = np.zeros(2 * n, self.dtype)
        flow = np.zeros(4 * n, self.dtype)
        self._f("oracle_core_publish")(self._c, _ptr(film), _ptr(flow))
        return film, flow

    def publish_surface(self, wetness=None, placed=None):
        """moisture, wet_edge and foam_crest maps of the render contract
        (src/render_contract.rs:457-500, :526-569); wetness / placed are the
        soil texels under the window (None: no soil)."""
        n = WINDOW_RES * WINDOW_RES
        w = None if wetness is None else np.ascontiguousarray(wetness, self.dtype)
        p = None if placed is None else np.ascontiguousarray(placed, self.dtype)
        moist, edge, foam = np.zeros(n, self.dtype), np.zeros(n, self.dtype), np.zeros(2 * n, self.dtype)
        self._f("oracle_core_publish_surface")(self._c, None if w is None else _ptr(w), None if p is None else _ptr(p),
                                               _ptr(moist), _ptr(edge), _ptr(foam))
        return moist, edge, foam

    def near_tiles(self):
        return int(self._f("oracle_core_near_tiles")(self._c))

    def refine(self, dt):
        self._f("oracle_core_refine")(self._c, dt)

    def restrict(self):
        self._f("oracle_core_restrict")(self._c)

    def fine_water_debug(self, pushes):
        """fine_water_debug (near.rs:341-346): 2 cm water, u, v and origin of
        the patch laid under ``pushes`` (dicts of push_tire's fields)."""
        keys = ("x", "z", "forward_x", "forward_z", "radius", "width", "speed", "slip")
        p = np.array([[q[k] for k in keys] for q in pushes], self.dtype).reshape(-1)
        fw, fu, fv = (np.zeros(FINE_RES * FINE_RES, self.dtype) for _ in range(3))
        o = np.zeros(2, self.dtype)
        self._f("oracle_core_fine_water_debug")(self._c, _ptr(p), len(pushes), _ptr(fw), _ptr(fu), _ptr(fv), _ptr(o))
        return fw, fu, fv, o

    def fine_water(self):
        fw = np.zeros(FINE_RES * FINE_RES, self.dtype)
        o = np.zeros(2, self.dtype)
        self._f("oracle_core_fine_get")(self._c, _ptr(fw), _ptr(o))
        return fw, o

    def recenter(self, x, z):
        """recenter (src/fluid_core.rs:5409-5422): move the window, parking
        cells that slide off and restoring those that slide on."""
        self._f("oracle_core_recenter")(self._c, x, z)

    def parked_cells(self):
        return int(self._f("oracle_core_parked_cells")(self._c))

    def drain_over_bed(self, dt, steps):
        """drain_over_bed (src/fluid_core.rs:5123-5214); the water volume."""
        return float(self._f("oracle_core_drain_over_bed")(self._c, dt, int(steps)))

    def sample_water(self, x, z):
        """Coarse water depth under a world point (sample_coarse when the patch
        is packed, src/fluid_core.rs:6806-6813)."""
        o = self.origin().astype(np.float64)
        i = int(np.floor((np.float32(x) - o[0]) / 0.2))
        j = int(np.floor((np.float32(z) - o[1]) / 0.2))
        if not (0 <= i < WINDOW_RES and 0 <= j < WINDOW_RES):
            return 0.0
        return float(self.planes()["water"][j * WINDOW_RES + i])


def floor_scale(base, r, precision=32):
    L = lib()
    if not getattr(L, "_core_declared", False):
        _declare_core(L)
        L._core_declared = True
    a = np.ascontiguousarray(r, _dt(precision))
    return getattr(L, "oracle_floor_scale" + _sfx(precision))(base, _ptr(a))


def recentre(r, precision=32):
    L = lib()
    if not getattr(L, "_core_declared", False):
        _declare_core(L)
        L._core_declared = True
    a = np.ascontiguousarray(r, _dt(precision)).copy()
    getattr(L, "oracle_recentre" + _sfx(precision))(_ptr(a))
    return a


# --- Bekker rut coupling (oracle/rut.c) ----------------------------------------

def _declare_rut(L):
    for sfx, R in (("_f32", C.c_float), ("_f64", C.c_double)):
        g = lambda n: getattr(L, n + sfx)  # noqa: E731
        g("oracle_bekker_load").argtypes = [R, R, R, R, R]
        g("oracle_bekker_load").restype = R
        g("oracle_sinkage_with").argtypes = [R, R, R, R, R]
        g("oracle_sinkage_with").restype = R
        g("oracle_rut_depth_from_force").argtypes = [R, R, R, R, R, R]
        g("oracle_rut_depth_from_force").restype = R
        g("oracle_contact_length").argtypes = [R, R]
        g("oracle_contact_length").restype = R
        g("oracle_terramech").argtypes = [R, _P, _P]
        g("oracle_terramech").restype = None
        g("oracle_softness_of").argtypes = [R, R, R]
        g("oracle_softness_of").restype = R
        g("oracle_stamp_tread").argtypes = [_P, _P, _P, C.c_int, R, R, R, _P, R]
        g("oracle_stamp_tread").restype = R


def _rut_lib():
    L = lib()
    if not getattr(L, "_rut_declared", False):
        _declare_rut(L)
        L._rut_declared = True
    return L


def bekker_load(sinkage, radius, width, k_eq, n, precision=32):
    return getattr(_rut_lib(), "oracle_bekker_load" + _sfx(precision))(sinkage, radius, width, k_eq, n)


def sinkage_with(load, radius, width, k_eq, n, precision=32):
    return getattr(_rut_lib(), "oracle_sinkage_with" + _sfx(precision))(load, radius, width, k_eq, n)


def rut_depth_from_force(normal, push, radius, width, k_eq, n, precision=32):
    return getattr(_rut_lib(), "oracle_rut_depth_from_force" + _sfx(precision))(normal, push, radius, width, k_eq, n)


def terramech(softness, precision=32):
    dt = _dt(precision)
    k = np.zeros(1, dt)
    n = np.zeros(1, dt)
    getattr(_rut_lib(), "oracle_terramech" + _sfx(precision))(softness, _ptr(k), _ptr(n))
    return float(k[0]), float(n[0])


def softness_of(wetness, loose, water, precision=32):
    return getattr(_rut_lib(), "oracle_softness_of" + _sfx(precision))(wetness, loose, water)


CONTACT_FIELDS = ("x", "z", "hub_y", "forward_x", "forward_z", "right_x", "right_z", "radius", "width", "sinkage",
                  "normal", "push", "k_eq", "n")


def stamp_tread(bed, placed, initial, res, cell, origin, contact, dt, precision=32):
    """stamp_wheel's loaded lane tread branch on numpy planes, in place;
    returns the cut volume (m^3)."""
    dtype = _dt(precision)
    for a in (bed, placed, initial):
        _check(a, dtype, res * res, "plane")
    cv = np.array([contact[k] for k in CONTACT_FIELDS], dtype)
    return getattr(_rut_lib(), "oracle_stamp_tread" + _sfx(precision))(_ptr(bed), _ptr(placed), _ptr(initial), res,
                                                                         cell, origin[0], origin[1], _ptr(cv), dt)


# Tire sub-grid network (oracle/mlp.c).
MLP_IN, MLP_HIDDEN, MLP_OUT = 1536, 128, 8
MLP_PARAMS = MLP_HIDDEN * MLP_IN + MLP_HIDDEN + MLP_HIDDEN * MLP_HIDDEN + MLP_HIDDEN + MLP_OUT * MLP_HIDDEN + MLP_OUT


def mlp_forward(params, x, bf16=True, precision=32):
    """y[n][8] of the network with parameter vector params on x[n][1536]."""
    dt = _dt(precision)
    p = np.ascontiguousarray(params, dtype=dt)
    xx = np.ascontiguousarray(x, dtype=dt)
    n = xx.shape[0]
    y = np.zeros((n, MLP_OUT), dtype=dt)
    getattr(lib(), "oracle_mlp_forward" + _sfx(precision))(_ptr(p), _ptr(xx), n, int(bf16), _ptr(y), None, None,
                                                            None, None)
    return y


def mlp_gradients(params, x, t, bf16=True, precision=32):
    """(sum of squared errors, gradient vector) of the mean squared error."""
    dt = _dt(precision)
    p = np.ascontiguousarray(params, dtype=dt)
    xx = np.ascontiguousarray(x, dtype=dt)
    tt = np.ascontiguousarray(t, dtype=dt)
    g = np.zeros(MLP_PARAMS, dtype=dt)
    sse = getattr(lib(), "oracle_mlp_gradients" + _sfx(precision))(_ptr(p), _ptr(xx), _ptr(tt), xx.shape[0],
                                                                   int(bf16), _ptr(g))
    return sse, g
