"""Pins of oracle functions that no reference fixture covers: each check is a
closed form, a known answer worked out cell by cell, or a brute-force sweep on
a tiny grid, chosen so that a plausible slip in the oracle (a swapped weight,
a wrong sign or neighbour, the wrong sweep order, a footprint off by a cell)
fails it.

* pressure blend (src/fluid_core.rs:1923-1935): on eta = a x^2 + b z^2 + c x z
  the separable 5-tap filters give closed forms fixed by the taps' sums and
  second moments and by the weights 0.62 / 0.38 and 0.45 / 0.55;
* vorticity confinement (:2380-2503, :3442-3471): the kick on a Gaussian
  vortex against the continuous force eps dt w (-d|w|/dz, d|w|/dx), and the
  in-place (Gauss-Seidel) and snapshot (Jacobi) sweeps against brute-force
  sweeps on an 8 x 8 grid;
* push_on (:8169-8239), peel_spray (:8241-8314), soften_border
  (:9594-9620) and one transport face (open_edge_dh, :2887-2911): known
  answers on single cells.
"""
import math

import numpy as np
import pytest

import pyoracle as O

G_TAPS = np.array([0.06, 0.24, 0.40, 0.24, 0.06], np.float32).astype(np.float64)
S_TAPS = np.array([0.02, 0.08, 0.80, 0.08, 0.02], np.float32).astype(np.float64)
M = np.arange(-2, 3, dtype=np.float64)


def f32(x):
    """A constant as the reference writes it: an f32 literal."""
    return float(np.float32(x))


def _moments(k):
    return k.sum(), (M * k).sum(), (M * M * k).sum()


@pytest.mark.parametrize("a,b,c", [(0.8, 0.0, 0.0), (0.0, -0.5, 0.0), (0.3, 0.7, 0.0), (0.0, 0.0, 0.9)])
def test_pressure_blend_on_quadratic_surfaces(a, b, c):
    """d = 0.45 (0.62 G(eta) + 0.38 S(eta)) + 0.55 eta with G, S separable.
    For eta = a x^2 + b z^2 + c x z (x, z in metres from the box centre) and
    taps k of sum s, first moment 0 and second moment m2 (cell units):
    K(x^2) = s^2 x^2 + s m2 h^2, K(z^2) likewise, K(x z) = s^2 x z."""
    res, cell = 24, 0.2
    i = np.arange(res, dtype=np.float64) - 11.5
    X, Z = np.meshgrid(i * cell, i * cell)
    eta = a * X * X + b * Z * Z + c * X * Z
    h = np.full(res * res, 0.05, np.float64)
    bed = (eta.ravel() - h).astype(np.float64)
    d = O.surface_pressure(h, bed, res, (0, res - 1, 0, res - 1), precision=64).reshape(res, res)
    sg, m1g, m2g = _moments(G_TAPS)
    ss, m1s, m2s = _moments(S_TAPS)
    assert m1g == 0.0 and m1s == 0.0
    h2 = cell * cell

    def filt(s, m2):
        return s * s * (a * X * X + b * Z * Z + c * X * Z) + s * m2 * h2 * (a + b)

    want = f32(0.45) * (f32(0.62) * filt(sg, m2g) + f32(0.38) * filt(ss, m2s)) + f32(0.55) * eta
    inner = (slice(2, res - 2), slice(2, res - 2))
    err = np.abs(d[inner] - want[inner]).max()
    assert err <= 1e-12 * max(1.0, np.abs(want[inner]).max()), err
    # The offset from eta is fixed by the weights and the second moments:
    # a slip such as 0.38 / 0.62 swapped or a dropped 0.55 eta term moves it.
    if c == 0.0:
        off = (d - eta)[inner].mean()
        assert off == pytest.approx(0.45 * (0.62 * m2g + 0.38 * m2s) * h2 * (a + b), rel=1e-6, abs=1e-12)
        assert abs(0.45 * (0.62 * m2g + 0.38 * m2s) - 0.32256) < 1e-6


def _vortex(res, cell, amp, sigma):
    """Stream function psi = amp exp(-r^2 / 2 sigma^2): u = -dpsi/dz,
    v = dpsi/dx, vorticity w = dv/dx - du/dz = laplacian(psi)."""
    c = (np.arange(res, dtype=np.float64) - (res - 1) / 2.0) * cell
    X, Z = np.meshgrid(c, c)
    r2 = X * X + Z * Z
    psi = amp * np.exp(-r2 / (2 * sigma * sigma))
    u = Z / sigma ** 2 * psi
    v = -X / sigma ** 2 * psi
    w = psi * (r2 / sigma ** 4 - 2.0 / sigma ** 2)
    # d|w| by hand: dw/dr = psi r (4 / sigma^4 - r^2 / sigma^6), then the chain rule.
    r = np.sqrt(r2)
    dwdr = psi * r * (4.0 / sigma ** 4 - r2 / sigma ** 6)
    with np.errstate(invalid="ignore", divide="ignore"):
        nx = np.sign(w) * dwdr * np.where(r > 0, X / r, 0.0)
        nz = np.sign(w) * dwdr * np.where(r > 0, Z / r, 0.0)
    return u, v, w, nx, nz, np.sqrt(r2)


@pytest.mark.parametrize("confine", ["jacobi", "gauss_seidel"])
def test_confinement_on_a_gaussian_vortex(confine):
    res, cell, dt, eps = 96, 0.2, 1.0 / 120.0, 0.08
    sigma = 12 * cell
    u, v, w, nx, nz, r = _vortex(res, cell, amp=1.5, sigma=sigma)
    want_du = eps * dt * (-nz * w)
    want_dv = eps * dt * (nx * w)
    n = res * res
    uu, vv = u.ravel().copy(), v.ravel().copy()
    O.force_open(np.full(n, 0.05), uu, vv, np.full(n, 0.9), np.zeros(n), res, cell, dt, eps,
                 (0, res - 1, 0, res - 1), confine=confine, precision=64)
    du = (uu - u.ravel()).reshape(res, res)
    dv = (vv - v.ravel()).reshape(res, res)
    # Away from the grid sides, the far field (|w| < 1e-4 gets no kick) and
    # the ring where w changes sign (|w| has a kink there).
    use = (r < 2.5 * sigma) & (np.abs(r - math.sqrt(2.0) * sigma) > 3 * cell) & (np.abs(w) > 1e-3)
    scale = max(np.abs(want_du[use]).max(), np.abs(want_dv[use]).max())
    e_u = np.abs(du - want_du)[use].max() / scale
    e_v = np.abs(dv - want_dv)[use].max() / scale
    # Central differences of central differences: O((cell / sigma)^2).
    assert e_u < 0.03 and e_v < 0.03, (e_u, e_v)
    # The kick turns with the vortex: at (x, 0) it is along z only.
    assert np.sign(du[res // 2, res // 2 + 6]) == np.sign(want_du[res // 2, res // 2 + 6])


def _vort(u, v, res, i, j, cell):
    ip, im = min(i + 1, res - 1), max(i - 1, 0)
    jp, jm = min(j + 1, res - 1), max(j - 1, 0)
    return (v[j, ip] - v[j, im]) / (2 * cell) - (u[jp, i] - u[jm, i]) / (2 * cell)


def _brute_sweep(u, v, res, cell, dt, eps, in_place):
    """The confinement sweep by hand: raster order, vorticity of the live
    planes (in place) or of a snapshot taken first."""
    u, v = u.copy(), v.copy()
    su, sv = (u, v) if in_place else (u.copy(), v.copy())
    for j in range(res):
        for i in range(res):
            w = _vort(su, sv, res, i, j, cell)
            if not math.isfinite(w) or abs(w) < 1e-4:
                continue
            wxp = abs(_vort(su, sv, res, min(i + 1, res - 1), j, cell))
            wxm = abs(_vort(su, sv, res, max(i - 1, 0), j, cell))
            wzp = abs(_vort(su, sv, res, i, min(j + 1, res - 1), cell))
            wzm = abs(_vort(su, sv, res, i, max(j - 1, 0), cell))
            du = eps * dt * (-((wzp - wzm) / (2 * cell)) * w)
            dv = eps * dt * (((wxp - wxm) / (2 * cell)) * w)
            if du * du + dv * dv <= 0.0625:
                u[j, i] += du
                v[j, i] += dv
    return u, v


def test_sweeps_in_place_and_from_a_snapshot_match_brute_force():
    """Strong confinement on a random 8 x 8 field, so the in-place sweep
    (later cells see earlier kicks) differs visibly from the snapshot one."""
    res, cell, dt, eps = 8, 0.2, 1.0 / 30.0, 25.0
    rng = np.random.default_rng(17)
    u = rng.uniform(-0.3, 0.3, (res, res))
    v = rng.uniform(-0.3, 0.3, (res, res))
    n = res * res
    outs = {}
    for mode, in_place in (("gauss_seidel", True), ("jacobi", False)):
        uu, vv = u.ravel().copy(), v.ravel().copy()
        O.force_open(np.full(n, 0.05), uu, vv, np.full(n, 0.9), np.zeros(n), res, cell, dt, eps,
                     (0, res - 1, 0, res - 1), confine=mode, precision=64)
        bu, bv = _brute_sweep(u, v, res, cell, dt, eps, in_place)
        assert np.allclose(uu.reshape(res, res), bu, rtol=0, atol=1e-14), mode
        assert np.allclose(vv.reshape(res, res), bv, rtol=0, atol=1e-14), mode
        outs[mode] = uu
    assert np.abs(outs["gauss_seidel"] - outs["jacobi"]).max() > 1e-3
    # Some kicks hit the 0.25 m/s cap and are dropped; some pass.
    kicked = np.abs(outs["jacobi"] - u.ravel()) > 0
    assert 0 < kicked.sum() < n


def _core_with_water(depth=0.05, slurry=0.05):
    c = O.Core((0.0, 0.0), precision=64)
    n = O.WINDOW_RES ** 2
    c.set_planes({"water": np.full(n, depth), "slurry": np.full(n, slurry), "uw": np.zeros(n), "vw": np.zeros(n),
                  "us": np.zeros(n), "vs": np.zeros(n), "bed": np.zeros(n), "moisture": np.zeros(n)})
    return c


def test_push_on_single_cells():
    """A tire on the face between cells (99, 100) and (100, 100) (x = 0, z at
    row 100's centre; the window's origin is -20 m) rolling along +x at 2 m/s, radius 0.38 m, width 0.28 m: the footprint
    reaches 1.4 r behind and 1.1 r ahead along the axis and 0.65 w + cell to
    the sides; the shove is speed (1 - flank^2) along the heading (0.85 of it
    on water; slurry 1 ahead of the tire centre and 0.72 at or behind it)
    plus a lateral part away from the axis."""
    c = _core_with_water()
    assert np.allclose(c.origin(), [-20.0, -20.0], rtol=0, atol=1e-6)  # the f32 cell: -20.0000003
    c.push_on(0.0, 0.1, 1.0, 0.0, 0.38, 0.28, 2.0)
    p = {k: a.reshape(200, 200) for k, a in c.planes().items()}
    # On the axis: lateral offset 0 (to the f32 cell's rounding), full shove
    # 2.0; cell 99 + k sits (k - 0.5) 0.2 m along the heading.
    for k in range(-2, 3):
        assert p["uw"][100, 100 + k - 1] == pytest.approx(1.7, rel=1e-6), k
        assert abs(p["vw"][100, 100 + k - 1]) < 1e-6
        assert p["us"][100, 100 + k - 1] == pytest.approx(2.0 * (1.0 if k > 0 else 0.72), rel=1e-6), k
    # Behind: -0.7 m < -0.532 m; ahead: 0.5 m > 0.418 m.  Outside.
    for k in (-3, 3):
        assert p["uw"][100, 100 + k - 1] == 0.0 and p["us"][100, 100 + k - 1] == 0.0
    # One cell to +z: lat = -0.2 m, flank = 0.2 / 0.382, pushed further to +z.
    flank = 0.2 / (0.28 * 0.65 + 0.2)
    shove = 2.0 * (1.0 - flank * flank)
    assert p["uw"][101, 100] == pytest.approx(0.85 * shove, rel=1e-6)
    assert p["vw"][101, 100] == pytest.approx(0.18 * shove * flank, rel=1e-6)
    assert p["vw"][99, 100] == pytest.approx(-0.18 * shove * flank, rel=1e-6)
    assert p["vs"][101, 100] == pytest.approx(0.28 * shove * flank, rel=1e-6)
    # Two cells to the side: 0.4 m > 0.382 m, outside.
    assert p["uw"][102, 100] == 0.0 and p["uw"][98, 100] == 0.0
    # Heading along +z the axis vector is (1, 0): a cell at +x sits at
    # lat = +0.2 m and is pushed further to +x on both layers.
    c3 = _core_with_water()
    c3.push_on(0.1, 0.0, 0.0, 1.0, 0.38, 0.28, 2.0)
    q = {k: a.reshape(200, 200) for k, a in c3.planes().items()}
    assert q["uw"][100, 101] == pytest.approx(0.18 * shove * flank, rel=1e-6)
    assert q["uw"][100, 99] == pytest.approx(-0.18 * shove * flank, rel=1e-6)
    assert q["vw"][100, 101] == pytest.approx(0.85 * shove, rel=1e-6)
    assert q["us"][100, 101] == pytest.approx(0.28 * shove * flank, rel=1e-6)
    # A thin film (<= 4 mm) is not shoved.
    c2 = _core_with_water(depth=0.004, slurry=0.0)
    c2.push_on(0.0, 0.1, 1.0, 0.0, 0.38, 0.28, 2.0)
    q = c2.planes()
    assert not q["uw"].any() and not q["us"].any()


def test_peel_spray_known_answers():
    """peel_spray takes min(4 % of the depth, 4 mm) from each cell whose
    centre lies within 1.3 r of a tire moving at >= 0.8 m/s and holding at
    least 1 cm, if that is at least 1e-6 m^3."""
    c = _core_with_water(depth=0.05, slurry=0.0)
    c.push_tire(0.1, 0.1, 1.0, 0.0, 0.38, 0.28, 2.0, 0.2)
    vol, events = c.peel()
    # Cell centres within 0.494 m of (0.1, 0.1): offsets (a, b) cells with
    # 0.04 (a^2 + b^2) <= 0.244036, i.e. a^2 + b^2 <= 6: 21 cells.
    area = f32(0.2) ** 2  # the window's f32 cell
    assert events == 21
    assert vol == pytest.approx(21 * 0.002 * area, rel=1e-6)
    w = c.planes()["water"].reshape(200, 200)
    assert w[100, 100] == pytest.approx(0.048, rel=1e-6) and w[100, 102] == pytest.approx(0.048, rel=1e-6)
    assert w[102, 102] == 0.05  # 8 > 6: outside
    # A wider tire (r = 0.5 m, reach 0.65 m): a^2 + b^2 <= 10.56, 37 cells.
    c = _core_with_water(depth=0.05, slurry=0.0)
    c.push_tire(0.1, 0.1, 1.0, 0.0, 0.5, 0.28, 2.0, 0.2)
    vol, events = c.peel()
    assert events == 37 and vol == pytest.approx(37 * 0.002 * area, rel=1e-6)
    # Deep water: the 4 mm cap.
    c = _core_with_water(depth=0.5, slurry=0.0)
    c.push_tire(0.1, 0.1, 1.0, 0.0, 0.38, 0.28, 2.0, 0.2)
    vol, events = c.peel()
    assert events == 21 and vol == pytest.approx(21 * 0.004 * area, rel=1e-6)
    # Slow tires and shallow water peel nothing.
    for depth, speed in ((0.05, 0.79), (0.0099, 2.0)):
        c = _core_with_water(depth=depth, slurry=0.0)
        c.push_tire(0.1, 0.1, 1.0, 0.0, 0.38, 0.28, speed, 0.2)
        assert c.peel() == (0.0, 0)


def test_soften_border_known_answers():
    """u, v scale by smoothstep(d / edge), d the distance in cells to the
    nearest window side, edge = max(res / 12, 2)."""
    res = 200
    u = np.ones(res * res)
    v = np.full(res * res, 2.0)
    O.soften_border(u, v, res, precision=64)
    U = u.reshape(res, res)
    V = v.reshape(res, res)
    edge = 16

    def ss(d):
        t = d / edge
        return t * t * (3 - 2 * t)

    for (j, i), d in (((100, 0), 0), ((100, 4), 4), ((100, 8), 8), ((100, 15), 15), ((100, 16), 16),
                      ((100, 199), 0), ((195, 100), 4), ((2, 4), 2), ((100, 100), 100)):
        want = ss(d) if d < edge else 1.0
        assert U[j, i] == pytest.approx(want, rel=1e-12, abs=0), (j, i)
        assert V[j, i] == pytest.approx(2 * want, rel=1e-12, abs=0), (j, i)
    assert U[100, 8] == 0.5
    u = np.ones(24 * 24)
    O.soften_border(u, np.ones(24 * 24), 24, precision=64)
    assert u.reshape(24, 24)[12, 1] == 0.5  # edge max(24 / 12, 2) = 2: t = 1 / 2, smoothstep 0.5


def _face(h0, h1, u0, u1, cell=0.2, dt=1.0 / 30.0, gate1=1):
    """Two cells of a 4 x 4 grid, (1, 1) and (2, 1), one face between them
    along x; every other cell dry and still."""
    res = 4
    h = np.zeros(res * res)
    vel = np.zeros(res * res)
    gate = np.ones(res * res, np.uint8)
    h[1 * res + 1], h[1 * res + 2] = h0, h1
    vel[1 * res + 1], vel[1 * res + 2] = u0, u1
    gate[1 * res + 2] = gate1
    O.transport(h, vel, gate, res, cell, dt, 1, 2, 1, 1, True, precision=64)
    return h[1 * res + 1], h[1 * res + 2]


def test_one_transport_face_known_answers():
    """open_edge_dh: the donor gives h_d min(|u_face| dt / cell, 0.45), the
    face speed the mean of the two cells' speeds; nothing below 1e-5 m/s, into
    a dry closed gate, or from a dry donor."""
    a, b = _face(0.05, 0.0, 0.8, 0.4)  # face speed 0.6: frac 0.1
    assert a == pytest.approx(0.045, rel=1e-12) and b == pytest.approx(0.005, rel=1e-12)
    a, b = _face(0.05, 0.02, 3.0, 3.0)  # frac 0.5 capped at 0.45
    assert a == pytest.approx(0.05 * 0.55, rel=1e-12) and b == pytest.approx(0.02 + 0.05 * 0.45, rel=1e-12)
    a, b = _face(0.05, 0.03, -0.6, -0.6)  # flows toward -x: donor is (2, 1)
    assert a == pytest.approx(0.05 + 0.03 * 0.1, rel=1e-12) and b == pytest.approx(0.03 * 0.9, rel=1e-12)
    assert _face(0.05, 0.0, 1.0e-5, 0.9e-5) == (0.05, 0.0)  # |u_face| < 1e-5
    assert _face(0.05, 0.0, 0.8, 0.4, gate1=0) == (0.05, 0.0)  # dry receiver behind a closed gate
    a, b = _face(0.05, 0.01, 0.8, 0.4, gate1=0)  # a closed gate with water behind it still fills
    assert a == pytest.approx(0.045, rel=1e-12)
    assert _face(0.0, 0.0, 0.8, 0.4) == (0.0, 0.0)
