"""A vehicle on a seeded steering path through the 40 m window, as the
nested-patch workload drives it: four tires (hubs at +-0.7 m across and
+-1.2 m along the heading; radius 0.38 m, width 0.28 m, src/vehicle.rs:15-16)
at a steady speed, 3.5 kN on each tire, Bekker soil k_eq = k_c / b + k_phi
(n = 1; src/soil.rs:724-738) cutting ruts that feed the sheet bed each step.

The steering is an Ornstein-Uhlenbeck turn rate drawn from the seed that
turns back toward the window centre near the edge, so the patch stays inside
the window."""
import math

import numpy as np

RADIUS, WIDTH = 0.38, 0.28
LOAD = 3500.0
K_EQ = 30e3 / WIDTH + 1.5e6
HUBS = [(-0.7, -1.2), (-0.7, 1.2), (0.7, -1.2), (0.7, 1.2)]  # (across, along)


class Drive:
    """One vehicle.  It keeps within `half` metres of `centre` (the window
    centre by default)."""

    def __init__(self, seed, speed=8.0, half=13.0, slip=0.2, centre=(0.0, 0.0)):
        self.rng = np.random.default_rng(seed)
        self.speed, self.half, self.slip = speed, half, slip
        self.cx, self.cz = centre
        side = self.rng.integers(4)
        a = self.rng.uniform(-half * 0.6, half * 0.6)
        starts = [(-half, a, 0.0), (half, a, math.pi), (a, -half, 0.5 * math.pi), (a, half, -0.5 * math.pi)]
        x, z, self.heading = starts[side]
        self.x, self.z = x + self.cx, z + self.cz
        self.heading += self.rng.uniform(-0.4, 0.4)
        self.rate = 0.0

    def forward(self):
        return math.cos(self.heading), math.sin(self.heading)

    def advance(self, dt):
        self.rate += -1.5 * self.rate * dt + 1.2 * math.sqrt(dt) * self.rng.standard_normal()
        r = math.hypot(self.x - self.cx, self.z - self.cz)
        if r > self.half:
            # Turn toward the centre.
            want = math.atan2(self.cz - self.z, self.cx - self.x)
            err = math.atan2(math.sin(want - self.heading), math.cos(want - self.heading))
            self.rate = 2.0 * err
        self.heading += self.rate * dt
        fx, fz = self.forward()
        self.x += fx * self.speed * dt
        self.z += fz * self.speed * dt

    def _hub(self, across, along):
        fx, fz = self.forward()
        rx, rz = fz, -fx
        return self.x + across * rx + along * fx, self.z + across * rz + along * fz

    def pushes(self):
        fx, fz = self.forward()
        out = []
        for across, along in HUBS:
            x, z = self._hub(across, along)
            out.append(dict(x=x, z=z, forward_x=fx, forward_z=fz, radius=RADIUS, width=WIDTH, speed=self.speed,
                            slip=self.slip))
        return out

    def contacts(self):
        fx, fz = self.forward()
        out = []
        for across, along in HUBS:
            x, z = self._hub(across, along)
            out.append(dict(x=x, z=z, hub_y=RADIUS, forward_x=fx, forward_z=fz, right_x=fz, right_z=-fx,
                            radius=RADIUS, width=WIDTH, sinkage=0.02, normal=LOAD, push=400.0, k_eq=K_EQ, n=1.0))
        return out


def nested_scene_core(S, scenes, seed, path="exact"):
    """A FluidCore on the window scene of `seed` (fBm bed, Gaussian puddle,
    slurry band) with its ground set so ruts feed the sheet bed."""
    sc = scenes.window_scene(seed=seed)
    core = S.FluidCore((0.0, 0.0), path=path)
    core.set_ground(sc["bed"])
    core.map_water_bed(sc["bed"], sc["water"])
    core.lay_slurry(sc["slurry"], "clay", 0.6)
    return core


def drive_step(core, drive, dt, step_fn=None, ponds=False):
    """Ruts, tires, one step (step_fn(core, dt) or core.step), with ponds
    the pond and seat passes ThroughDrive runs after the step
    (src/world.rs:4686-4694), then the vehicle moves."""
    core.stamp_ruts(drive.contacts(), dt)
    for p in drive.pushes():
        core.push_tire(**p)
    out = step_fn(core, dt) if step_fn else core.step(dt)
    if ponds:
        core.pond_ruts()
        core.seat_rut_sheet()
    drive.advance(dt)
    return out


class Fleet:
    """n vehicles on a large sheet: each keeps to its own disc of radius
    `half` around a centre drawn from the seed, the discs at least
    2 half + 8 m apart so no two vehicles' network regions meet."""

    def __init__(self, seed, n, extent, half=10.0, speed=8.0):
        rng = np.random.default_rng(seed)
        lo, hi = -0.5 * extent + half + 6.0, 0.5 * extent - half - 6.0
        centres = []
        while len(centres) < n:
            c = rng.uniform(lo, hi, 2)
            if all(math.hypot(c[0] - d[0], c[1] - d[1]) >= 2 * half + 8.0 for d in centres):
                centres.append((float(c[0]), float(c[1])))
        self.drives = [Drive(seed * 1000 + k, speed=speed, half=half, centre=centres[k]) for k in range(n)]

    def advance(self, dt):
        for d in self.drives:
            d.advance(dt)

    def pushes(self):
        return [p for d in self.drives for p in d.pushes()]

    def contacts(self):
        return [c for d in self.drives for c in d.contacts()]
