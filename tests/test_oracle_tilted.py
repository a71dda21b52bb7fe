"""Closed-form pins of the oracle's tilted-surface arithmetic: the slope terms
of the angle of incidence (solar_kernel.cu:68), the Perez sky's tilt terms and
coefficient table (solar_kernel.cu:42-47, device_globals.cu:50-57) and the
receiver normal of the reflection weight (reflection_kernels.cu:55-63).

Every expected value here comes from geometry or from the published Perez
(1990) model, never from the oracle's own expressions:
  * the sun vector and the plane normals are written from compass directions
    (x = east, y = south on the north-up raster, z = up);
  * the Perez coefficients are the rounded "all sites composite" table of
    Perez, Ineichen, Seals, Michalsky and Stewart, Solar Energy 44 (1990),
    laid out bin-major as published (8 clearness bins x f11 f12 f13 f21 f22 f23);
  * the oracle's F1 and F2 are recovered per hour from its diffuse output on
    two vertical walls and fitted per clearness bin, then compared with that
    table, so a typo or transposition in any coefficient row fails.
A misreading of the aspect or azimuth in degrees, a swapped or negated normal
component, or a wrong sign of the tilt term each fail at least one test."""
import numpy as np

import oracle_lib as orc
from scenes import scene_from_heights, sky

# Perez et al. 1990, all-sites composite; row = clearness bin, columns f11 f12 f13 f21 f22 f23.
PEREZ_1990 = np.array([
    [-0.008, 0.588, -0.062, -0.060, 0.072, -0.022],
    [0.130, 0.683, -0.151, -0.019, 0.066, -0.029],
    [0.330, 0.487, -0.221, 0.055, -0.064, -0.026],
    [0.568, 0.187, -0.295, 0.109, -0.152, -0.014],
    [0.873, -0.392, -0.362, 0.226, -0.462, 0.001],
    [1.132, -1.237, -0.412, 0.288, -0.823, 0.056],
    [1.060, -1.600, -0.359, 0.264, -1.127, 0.131],
    [0.678, -0.327, -0.250, 0.156, -1.377, 0.251],
])
EPS_EDGES = [1.065, 1.230, 1.500, 1.950, 2.800, 4.500, 6.200]
EPS_MID = [1.03, 1.15, 1.36, 1.72, 2.35, 3.6, 5.3, 9.0]
KAPPA = 1.041


def sun_vector(az_deg, alt_deg):
    """Unit vector towards the sun: azimuth clockwise from north, x east, y south, z up."""
    az, alt = np.radians(az_deg), np.radians(alt_deg)
    return np.array([np.cos(alt) * np.sin(az), -np.cos(alt) * np.cos(az), np.sin(alt)])


def plane_normal(tilt_deg, facing_az_deg):
    """Outward normal of a plane tilted by tilt_deg towards the compass azimuth facing_az_deg."""
    t = np.radians(tilt_deg)
    horiz = sun_vector(facing_az_deg, 0.0)  # horizontal unit vector towards that azimuth
    return np.cos(t) * np.array([0.0, 0.0, 1.0]) + np.sin(t) * horiz


def perez_terms(zenith_rad, dhi, dni, day):
    """Published Perez quantities for one hour: clearness bin, brightness delta, F1, F2."""
    z = zenith_rad
    am = 1.0 / (np.cos(z) + 0.50572 * (96.07995 - np.degrees(z)) ** -1.6364)   # Kasten-Young air mass
    B = 2 * np.pi * day / 365.0
    e0 = 1367 * (1.00011 + 0.034221 * np.cos(B) + 0.00128 * np.sin(B) + 0.000719 * np.cos(2 * B)
                 + 0.000077 * np.sin(2 * B))                                 # Spencer extraterrestrial
    eps = max(1.0, ((dhi + dni / np.cos(z)) / dhi + KAPPA * z ** 3) / (1 + KAPPA * z ** 3))
    k = int(np.searchsorted(EPS_EDGES, eps, side="right"))
    delta = am * dhi / e0
    f = PEREZ_1990[k]
    return k, delta, max(0.0, f[0] + f[1] * delta + f[2] * z), f[3] + f[4] * delta + f[5] * z


# Voxels of the test plane: (tilt in degrees, facing relative to the sun's azimuth or absolute).
PLANES = [
    ("wall facing the sun", 90.0, "sun", 0.0),
    ("wall facing away", 90.0, "sun", 180.0),
    ("30 deg towards the sun", 30.0, "sun", 0.0),
    ("45 deg towards the sun", 45.0, "sun", 0.0),
    ("60 deg towards the sun", 60.0, "sun", 0.0),
    ("45 deg facing south", 45.0, "abs", 180.0),
    ("wall facing east", 90.0, "abs", 90.0),
    ("wall facing west", 90.0, "abs", 270.0),
    ("horizontal", 0.0, "abs", 0.0),
]


def _plane_scene():
    sc = scene_from_heights(np.full((4, 4), 200.0, np.float32))
    assert sc.num_total == 16 and sc.num_top == 16
    return sc


def _orient(sc, az_deg):
    """Slopes and aspects (radians; aspect = the compass azimuth the plane faces) of PLANES."""
    for v, (_, tilt, mode, face) in enumerate(PLANES):
        facing = (az_deg + face) % 360.0 if mode == "sun" else face
        sc.slopes[v] = np.float32(np.radians(tilt))
        sc.aspects[v] = np.float32(np.radians(facing))
    return sc


def _hours(count=60):
    _, _, _, bins = sky()
    hours = orc.sunlit_hours(bins["nudged"])
    alt = bins["nudged"][hours, 1]
    keep = hours[(alt > 4.0) & (alt < 80.0)]
    return bins, keep[np.linspace(0, keep.size - 1, count).astype(int)]


def test_direct_beam_on_tilted_planes_is_dni_times_the_geometric_incidence():
    sc = _plane_scene()
    bins, hours = _hours()
    n_dir = 0
    for hi in hours:
        if bins["sun_dense_idx"][hi] < 0:
            continue
        az, alt = (float(x) for x in bins["nudged"][hi])
        dni = np.zeros(8760, np.float32)
        dhi = np.zeros(8760, np.float32)
        dni[hi], dhi[hi] = 700.0, 90.0
        _orient(sc, az)
        lit = orc.solar_hour(sc, bins, dni, dhi, np.zeros(bins["svf_bytes"] * 16, np.uint8), int(hi))
        dark = orc.solar_hour(sc, bins, dni, dhi, np.full(bins["svf_bytes"] * 16, 255, np.uint8), int(hi))
        direct = (lit - dark).astype(np.float64)
        s = sun_vector(az, alt)
        zen = np.radians(90.0 - alt)
        for v, (name, tilt, mode, face) in enumerate(PLANES):
            facing = (az + face) % 360.0 if mode == "sun" else face
            cos_i = float(plane_normal(tilt, facing) @ s)
            if name == "wall facing the sun":
                assert abs(cos_i - np.cos(np.radians(alt))) < 1e-12        # geometry: cos AOI = cos(alt)
            if mode == "sun" and tilt < 90.0:
                assert abs(cos_i - np.cos(zen - np.radians(tilt))) < 1e-12  # cos AOI = cos(zenith - tilt)
            want = 700.0 * cos_i if cos_i > 1e-6 else 0.0
            if abs(cos_i) <= 1e-6:
                continue                                                  # grazing: either side is right
            assert abs(direct[v] - want) <= 2e-4 * 700.0, (hi, name, direct[v], want)
        assert direct[1] == 0.0                                           # facing away: no beam at all
        n_dir += 1
    assert n_dir >= 40


def test_south_facing_45_degree_slope_at_solar_noon():
    """Hours whose sun azimuth is within 2 degrees of south: the equator-facing 45-degree
    plane sees the beam at cos(zenith - 45 deg), up to the small azimuth offset."""
    sc = _plane_scene()
    _, _, _, bins = sky()
    hours = orc.sunlit_hours(bins["nudged"])
    az = bins["nudged"][hours, 0]
    noon = hours[(np.abs(az - 180.0) < 2.0) & (bins["sun_dense_idx"][hours] >= 0)]
    assert noon.size > 150
    for hi in noon[::7]:
        a, alt = (float(x) for x in bins["nudged"][hi])
        dni = np.zeros(8760, np.float32)
        dhi = np.zeros(8760, np.float32)
        dni[hi], dhi[hi] = 800.0, 80.0
        _orient(sc, a)
        lit = orc.solar_hour(sc, bins, dni, dhi, np.zeros(bins["svf_bytes"] * 16, np.uint8), int(hi))
        dark = orc.solar_hour(sc, bins, dni, dhi, np.full(bins["svf_bytes"] * 16, 255, np.uint8), int(hi))
        want = 800.0 * np.cos(np.radians(90.0 - alt - 45.0))
        offset = 800.0 * np.sin(np.radians(90.0 - alt)) * np.sin(np.radians(45.0)) * \
            (1 - np.cos(np.radians(a - 180.0)))
        assert abs(float(lit[5] - dark[5]) - want) <= offset + 2e-4 * 800.0, (hi, lit[5] - dark[5], want)


def _craft_weather(bins, hours, rng):
    """DNI/DHI per chosen hour so that each Perez clearness bin is hit mid-bin by several
    hours of different zenith and brightness."""
    dni = np.zeros(8760, np.float32)
    dhi = np.zeros(8760, np.float32)
    target = {}
    for j, hi in enumerate(hours):
        k = j % 8
        z = np.radians(90.0 - float(bins["nudged"][hi, 1]))
        for _ in range(200):  # keep F1 > 0 and the vertical walls' diffuse unclamped
            d = float(rng.uniform(30.0, 400.0))
            n = d * np.cos(z) * (EPS_MID[k] - 1.0) * (1 + KAPPA * z ** 3)
            _, _, F1, F2 = perez_terms(z, d, n, int(hi) // 24)
            if F1 > 0.02 and (1 - F1) / 2 + F2 > 0.05:
                break
        else:
            raise AssertionError(f"no unclamped brightness for bin {k} at zenith {z:.3f}")
        dhi[hi], dni[hi] = d, n
        target[int(hi)] = k
    return dni, dhi, target


def _recover_f1_f2(sc, bins, dni, dhi, hi):
    """F1 and F2 of hour hi from the oracle's diffuse on the two vertical walls (shaded, so
    diffuse only): facing away, D = DHI((1 - F1)/2 + F2); facing the sun, the circumsolar
    term adds DHI F1 cos(alt) / max(cos 85 deg, cos zenith)."""
    az, alt = (float(x) for x in bins["nudged"][hi])
    _orient(sc, az)
    dark = orc.solar_hour(sc, bins, dni, dhi, np.full(bins["svf_bytes"] * 16, 255, np.uint8), int(hi))
    d_sun, d_away = float(dark[0]), float(dark[1])
    z = np.radians(90.0 - alt)
    b = max(np.cos(np.radians(85.0)), np.cos(z))
    f1 = (d_sun - d_away) * b / (float(dhi[hi]) * np.cos(np.radians(alt)))
    f2 = d_away / float(dhi[hi]) - (1.0 - f1) / 2.0
    return f1, f2, dark


def test_perez_tilt_terms_and_coefficient_table_match_the_published_model():
    sc = _plane_scene()
    bins, hours = _hours(count=96)
    dni, dhi, target = _craft_weather(bins, hours, np.random.default_rng(11))
    rows = {k: [] for k in range(8)}
    for hi in hours:
        hi = int(hi)
        day = hi // 24
        az, alt = (float(x) for x in bins["nudged"][hi])
        z = np.radians(90.0 - alt)
        k, delta, F1, F2 = perez_terms(z, float(dhi[hi]), float(dni[hi]), day)
        assert k == target[hi]
        f1, f2, dark = _recover_f1_f2(sc, bins, dni, dhi, hi)
        # The published table has three decimals; implementations of the model carry longer
        # values that agree with it within 1e-3 per coefficient (1.1326 for 1.132, say).
        tol = 1e-3 * (1.0 + delta + z) + 1e-5
        assert abs(f1 - F1) <= tol, (hi, k, f1, F1)
        assert abs(f2 - F2) <= tol, (hi, k, f2, F2)
        # Every plane: DHI((1 - F1)(1 + cos s)/2 + F1 max(0, cos AOI)/b + F2 sin s).
        s = sun_vector(az, alt)
        b = max(np.cos(np.radians(85.0)), np.cos(z))
        for v, (name, tilt, mode, face) in enumerate(PLANES):
            facing = (az + face) % 360.0 if mode == "sun" else face
            cos_i = float(plane_normal(tilt, facing) @ s)
            t = np.radians(tilt)
            a_b = max(cos_i, 0.0) / b
            want = float(dhi[hi]) * ((1 - F1) * (1 + np.cos(t)) / 2 + F1 * a_b + F2 * np.sin(t))
            # coefficient rounding enters through F1 (times the circumsolar ratio a/b) and F2
            bound = tol * float(dhi[hi]) * ((1 + np.cos(t)) / 2 + a_b + np.sin(t))
            assert abs(float(dark[v]) - max(want, 0.0)) <= bound + 1e-3, (hi, name)
        if f1 > 1e-3:
            rows[k].append((delta, z, f1, f2))
    for k in range(8):
        r = np.array(rows[k])
        assert len(r) >= 6, k
        A = np.stack([np.ones(len(r)), r[:, 0], r[:, 1]], axis=1)
        c1 = np.linalg.lstsq(A, r[:, 2], rcond=None)[0]
        c2 = np.linalg.lstsq(A, r[:, 3], rcond=None)[0]
        np.testing.assert_allclose(np.concatenate([c1, c2]), PEREZ_1990[k], atol=1e-3, err_msg=f"bin {k}")


def _reflection_case(tilt_deg, facing_deg):
    """One receiver (cell (4, 4) of a 9 x 9 plane at 200 m, tilted and facing as given) with
    five neighbour cells: east, south, west and north at 3 cells, and a raised cell to the
    north-east.  Returns the oracle's reflected irradiance for the receiver and the
    closed form from the plane normal and the Lambertian single bounce."""
    h = np.full((9, 9), 200.0, np.float32)
    cells = {"E": (7, 4, 198.0), "S": (4, 7, 199.0), "W": (1, 4, 196.0), "N": (4, 1, 198.0), "NE": (6, 2, 212.0)}
    for x, y, z in cells.values():
        h[y, x] = z
    sc = scene_from_heights(h)
    t = 4 * 9 + 4
    sc.slopes[t] = np.float32(np.radians(tilt_deg))
    sc.aspects[t] = np.float32(np.radians(facing_deg))
    hourly = np.zeros(sc.num_top, np.float32)
    ids = []
    for j, (x, y, z) in enumerate(cells.values()):
        hourly[y * 9 + x] = 100.0 * (j + 1)
        ids.append(y * 9 + x + 1)
    off = np.zeros(sc.num_total + 1, np.int64)
    off[t + 1:] = len(ids)
    got = float(orc.reflect_hour_csr(sc, off, np.array(ids, np.int32), hourly)[t])
    n = plane_normal(tilt_deg, facing_deg)
    want = 0.0
    seen = []
    for j, (key, (x, y, z)) in enumerate(cells.items()):
        v = np.array([x - 4.0, y - 4.0, z - 200.0])
        d = np.linalg.norm(v)
        cos_r = float(n @ v) / d
        if cos_r <= 1e-9:  # behind the plane (or in it, to rounding)
            continue
        seen.append(key)
        want += 0.2 * 100.0 * (j + 1) * cos_r * max(abs(v[2]) / d, 0.1) / (np.pi * d * d)
    return got, want, seen


def test_reflection_receiver_normal_from_compass_directions():
    cases = {  # (tilt, facing) -> neighbours in front of the plane
        (90.0, 90.0): {"E", "NE"},
        (90.0, 180.0): {"S"},
        (90.0, 270.0): {"W"},
        (90.0, 0.0): {"N", "NE"},
        (45.0, 90.0): {"E", "NE"},
        (60.0, 270.0): {"W", "NE"},
        (0.0, 0.0): {"NE"},
    }
    for (tilt, face), front in cases.items():
        got, want, seen = _reflection_case(tilt, face)
        assert set(seen) == front, (tilt, face, seen)
        assert abs(got - want) <= 2e-6 * want + 1e-9, (tilt, face, got, want)
