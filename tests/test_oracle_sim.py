"""The oracle's simulation against the reference test suite
(tests/test_mojo.mojo:48-108), closed forms and brute force."""
import numpy as np
import pytest

import oracle as O
from helpers import corridor_state, ray_circle_range, track

G, GRIP, STEER_MAX, ACCEL_MAX, SPEED_MAX = 9.81, 1.05, 0.4189, 9.51, 5.0


def test_physics_closed_forms():
    """tests/test_mojo.mojo:48-55"""
    d = O.derivative
    assert np.array_equal(d([0, 0, 0, 0], 0, 0, GRIP, 0, 0), np.zeros(4, np.float32)), "a car at rest stays at rest"
    assert d([0, 0, 0, 2], 0, 0, GRIP, 0.3, 0)[3] < 0, "an uphill slope slows the car"
    assert d([0, 0, 0, 2], 0, 0, GRIP, 0, 0)[3] == 0, "coasting on the flat holds speed"
    limit = d([0, 0, 0, 5], STEER_MAX, ACCEL_MAX, GRIP, 0, 0)
    assert abs(limit[2] * 5) <= GRIP * G + 1e-4 and limit[3] <= 1e-4, "the friction circle caps the tyres"
    assert d([0, 0, 0, 2], 0.2, 0, GRIP, 0, 0)[2] > 0.5, "steering yaws the car"


def test_derivative_matches_float64_formula():
    rng = np.random.default_rng(3)
    for _ in range(200):
        p = rng.uniform([-5, -5, -3, -5], [5, 5, 3, 5]).astype(np.float32)
        steer, accel, grip = rng.uniform(-0.42, 0.42), rng.uniform(-9.5, 9.5), rng.uniform(0.9, 1.2)
        along, across = rng.uniform(-0.3, 0.3, 2)
        got = O.derivative(p, steer, accel, grip, along, across)
        nz = 1 / np.sqrt(1 + along**2 + across**2)
        load = grip * G * nz
        v = max(abs(p[3]), 0.5) * (1 if p[3] >= 0 else -1)
        slide = G * across * nz
        lat = np.clip(v * p[3] * np.tan(steer) / 0.3302 + slide, -load, load)
        yaw = (lat - slide) / v
        lon = np.sqrt(max(load**2 - lat**2, 0))
        dv = np.clip(accel, -lon, lon) - G * along * nz
        ref = [p[3] * np.cos(p[2]), p[3] * np.sin(p[2]), yaw, dv]
        np.testing.assert_allclose(got, ref, rtol=2e-5, atol=2e-5)


def test_corridor_suite():
    """tests/test_mojo.mojo:58-92: lidar, crash, respawn, speed cap, progress, steering."""
    beams = 108
    sim = O.Sim(track("corridor"), 4, 1, beams=beams)
    s, actions = corridor_state()
    sim.state[:] = s
    sim.actions[:] = actions
    sim.step()
    assert abs(sim.obs[4 + beams // 2, 0] * 10 - 7.1) <= 0.1, "the centre beam stops on the ramp face"
    assert abs(sim.obs[4 + beams // 2, 1] * 10 - 0.73) <= 0.1, "the centre beam reads the wall"
    crashed = False
    for _ in range(60):
        sim.step()
        d = sim.done
        assert d[0] == 0 and d[2] == 0 and d[3] == 0, "the open road does not crash"
        if d[1] == 1 and not crashed:
            assert sim.state[6, 1] == 0 and sim.state[8, 1] == 1, "a crash respawns at once"
            crashed = True
    assert crashed, "driving into the wall crashes"
    st, o = sim.state, sim.obs
    assert st[0, 0] == 1 and st[4, 0] == 0 and st[3, 0] == 0, "a car at rest stays at rest"
    assert st[4, 2] == SPEED_MAX and st[0, 2] > 3.5 and abs(st[1, 2]) < 1e-4, "full throttle"
    assert abs(st[10, 2] - 10 * (st[0, 2] - 1)) <= 0.01, "continuous progress reward"
    assert st[1, 3] > 0.1 and st[0, 3] < st[0, 2], "steering turns the car left"
    assert o[1, 2] == 1 and o[beams + 4, 2] == 1, "speed and bias observations"
    assert np.isfinite(o).all(), "observations stay finite"


@pytest.mark.parametrize("name,deck", [("bridge", 0.9), ("bridge8", 0.5)])
def test_bridge_layers(name, deck):
    """tests/test_mojo.mojo:95-108 on the reference's bridge and the 8 m-lobe bridge."""
    m = O.Map(track(name))
    assert m.layers == 2 and m.probe(0, 0, 0)[0] == 0 and m.probe(0, 0, deck * 0.6)[0] == 1, "layer by height"
    assert m.probe(0, 0, 0)[1] > 1 and m.probe(0, 0, deck)[1] > 1, "both roads cross at the origin"
    for w in range(0, m.waypoints, 2):
        x, y, z = m.waypoint(w, 0), m.waypoint(w, 1), m.waypoint(w, 2)
        layer, clearance, _ = m.probe(x, y, z)
        assert clearance > 0.9, "the route centre is clear of walls"
        assert abs(m.surface(layer, x, y)[0] - z) <= 0.03, "the surface follows the route"
        assert abs(m.progress(layer, x, y) - w) <= 1e-3, "progress counts waypoints"


def _place(sim, x, y, z, heading, speed=0.0):
    sim.state[:] = 0
    sim.state[0], sim.state[1], sim.state[2], sim.state[3], sim.state[4] = x, y, z, heading, speed
    sim.state[9] = 1


def test_flat_ring_lidar_closed_form():
    """On the flat R = 10 m ring (half width 1.3 m) a car on the centreline sees the
    walls where rays meet the circles of radius R +- (1.3 - 1.5 cells), give or
    take the nearest-cell quantisation of the clearance grid."""
    tr = track("ring10")
    beams = 1081
    sim = O.Sim(tr, 1, 0, beams=beams)
    cell = float(tr[4])
    R, hw = 10.0, 1.3
    _place(sim, R, 0.0, 0.0, np.pi / 2)
    sim.actions[:] = 0
    sim.step()
    heading = float(sim.state[3, 0])
    x, y = float(sim.state[0, 0]), float(sim.state[1, 0])
    ox, oy = x + 0.27 * np.cos(heading), y + 0.27 * np.sin(heading)
    radii = [R + hw - 1.5 * cell, R - hw + 1.5 * cell]
    checked = 0
    for b in range(0, beams, 20):
        az = -4.71238898 / 2 + b * 4.71238898 / (beams - 1)
        dx, dy = np.cos(heading + az), np.sin(heading + az)
        expect = min(ray_circle_range(ox, oy, dx, dy, radii), 10.0)
        got = float(sim.obs[4 + b, 0]) * 10
        # incidence-angle bound on the along-ray error of a 1-cell clearance error
        hit_x, hit_y = ox + expect * dx, oy + expect * dy
        nrm = np.hypot(hit_x, hit_y)
        sin_inc = abs((hit_x * dx + hit_y * dy) / nrm) if expect < 10 else 1.0
        tol = 2.5 * cell / max(sin_inc, 0.15)
        assert abs(got - expect) <= tol, (b, got, expect, tol)
        checked += 1
    assert checked > 50


def test_lidar_never_passes_a_wall_brute_force():
    """Sphere tracing is conservative: on the corridor every beam stops no later
    than a dense 2 mm march first sees clearance below 1.5 cells or the floor."""
    tr = track("corridor")
    m = O.Map(tr)
    beams = 54
    sim = O.Sim(tr, 1, 0, beams=beams)
    for heading in (0.0, 0.7, 2.0, -2.5):
        _place(sim, 2.0, 0.3, 0.0, heading)
        sim.actions[:] = 0
        sim.step()
        h = float(sim.state[3, 0])
        along, across = float(sim.obs[2, 0]), float(sim.obs[3, 0])
        fwd = np.array([np.cos(h), np.sin(h), along])
        left = np.array([-np.sin(h), np.cos(h), across])
        origin = np.array([sim.state[0, 0], sim.state[1, 0], sim.state[2, 0]], float) + 0.27 * fwd + [0, 0, 0.23]
        for b in range(beams):
            az = -4.71238898 / 2 + b * 4.71238898 / (beams - 1)
            d = np.cos(az) * fwd + np.sin(az) * left
            d /= np.linalg.norm(d)
            t = 0.0
            while t < 10:
                p = origin + t * d
                _, c, hh = m.probe(*p)
                if c < 1.5 * m.cell or p[2] - hh - 0.03 < 0:
                    break
                t += 0.002
            got = float(sim.obs[4 + b, 0]) * 10
            assert got <= min(t, 10) + 2 * m.cell, (heading, b, got, t)


def test_rk4_matches_float64_fine_integration():
    """One env step (4 RK4 substeps) on the flat ring against a float64 RK4 with
    256 substeps of the same ODE (steer ramps at the commanded rate)."""
    tr = track("ring10")
    sim = O.Sim(tr, 1, 0, beams=8)
    _place(sim, 10.0, 0.0, 0.0, np.pi / 2, speed=3.0)
    sim.state[5] = 0.05
    a0, a1 = 0.5, 0.3
    sim.actions[:] = np.array([[a0], [a1]], np.float32)
    p0 = np.array([10.0, 0.0, np.pi / 2, 3.0])
    sim.step()
    rate, accel, grip = a0 * 3.2, a1 * 9.51, 1.05 * 1.0

    def f(p, steer):
        v = max(abs(p[3]), 0.5) * (1 if p[3] >= 0 else -1)
        lat = np.clip(v * p[3] * np.tan(steer) / 0.3302, -grip * G, grip * G)
        lon = np.sqrt(max((grip * G) ** 2 - lat**2, 0))
        return np.array([p[3] * np.cos(p[2]), p[3] * np.sin(p[2]), lat / v, np.clip(accel, -lon, lon)])

    p, steer, k = p0.copy(), 0.05, 256
    dt = (1 / 60) / k
    for _ in range(k):
        k1 = f(p, steer)
        k2 = f(p + dt / 2 * k1, steer + rate * dt / 2)
        k3 = f(p + dt / 2 * k2, steer + rate * dt / 2)
        k4 = f(p + dt * k3, steer + rate * dt)
        p = p + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        steer += rate * dt
    got = np.array([sim.state[0, 0], sim.state[1, 0], sim.state[3, 0], sim.state[4, 0]], float)
    np.testing.assert_allclose(got, p, atol=2e-4)
    assert abs(sim.state[5, 0] - steer) < 1e-6


def test_timeout_and_respawn_bookkeeping():
    """A car at STEPS = HORIZON - 1 times out (done = 2) and respawns with
    EPISODE + 1 and STEPS = 0 (racer/sim.mojo:220-229)."""
    sim = O.Sim(track("ring10"), 2, 5, beams=16)
    sim.reset()
    sim.state[6, 0] = 2999
    ep = sim.state[8].copy()
    sim.actions[:] = 0
    sim.step()
    assert sim.done[0] == 2 and sim.done[1] == 0
    assert sim.state[6, 0] == 0 and sim.state[8, 0] == ep[0] + 1 and sim.state[6, 1] == 2
