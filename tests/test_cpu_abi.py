"""Host-side checks of the C ABI library (no GPU): it loads, exports every symbol the
header declares, and its host-only entry points (RNG, scenes, mass properties,
status strings) match the reference's conventions and golden streams."""
import json
import math
import os
import re

import numpy as np
import pytest
from scipy import integrate

from helpers import seed

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "axiom", "axm.h")
GOLDEN = os.path.join(ROOT, "tests", "golden", "pcg_streams.json")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(axm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(axm):
    L = axm.lib()
    declared = header_functions()
    assert len(declared) >= 20
    missing = [f for f in declared if not hasattr(L, f)]
    assert not missing, missing
    assert sorted(axm.EXPORTS) == declared


def test_no_device_is_an_error_not_a_fallback(axm):
    if axm.device_count() > 0:
        pytest.skip("a GPU is visible")
    b, d = axm.scene(0, seed(0), 100)
    with pytest.raises(axm.AxmError) as e:
        axm.World(d)
    assert e.value.code in (503, 505)


# ErrorCode values and messages of include/axiom/core/error_code.hpp:21-57 and
# src/core/error_code.cpp:5-65 (the status codes the ABI returns).
REFERENCE_CODES = {
    0: "Success", 202: "Unexpected null pointer", 300: "Invalid collision shape",
    301: "GJK algorithm failed to converge", 302: "EPA algorithm failed to converge",
    400: "Invalid rigid body properties", 401: "Constraint solver failed to converge",
    502: "GPU buffer allocation failed", 503: "Invalid GPU operation or state", 504: "GPU operation timed out",
    505: "GPU operation failed", 600: "Invalid parameter", 601: "Value out of range",
}


def test_status_strings_match_reference_messages(axm):
    L = axm.lib()
    for code, msg in REFERENCE_CODES.items():
        assert L.axm_status_string(code).decode() == msg, code
    assert L.axm_status_string(12345).decode() == "Unknown error"


def test_rng_matches_reference_golden_streams(axm):
    """The scene RNG reproduces DeterministicRNG (include/axiom/math/random.hpp:19-68) bit for bit:
    raw u32 draws and nextFloat(min, max), including the draw that rounds to exactly `max`."""
    doc = json.load(open(GOLDEN))
    assert len(doc["streams"]) >= 10
    for e in doc["streams"]:
        u, f = axm.rng_stream(e["seed"], doc["draws"], e["float_lo"], e["float_hi"])
        assert u.tolist() == e["u32"], e["seed"]
        assert f.view(np.uint32).tolist() == e["float_bits"], e["seed"]
    q = [e for e in doc["streams"] if e["seed"] == doc["quirk_seed"]][0]
    assert q["u32"][0] >= 0xFFFFFF80
    _, f = axm.rng_stream(q["seed"], 1, 0.0, 1.0)
    assert f[0] == np.float32(1.0)  # nextFloat() == 1.0f: the reference's [0, 1) is not strict


def _capsule_inertia_quadrature(r, h, rho):
    """Solid of revolution about Y: slices of radius R(y); I_yy = int rho pi R^4 / 2,
    I_xx = int rho pi R^2 (R^2 / 4 + y^2)."""
    def R2(y):
        a = abs(y)
        return r * r if a <= h else max(r * r - (a - h) ** 2, 0.0)
    lim = h + r
    pts = [-h, h] if h > 0 else None
    m = integrate.quad(lambda y: rho * math.pi * R2(y), -lim, lim, points=pts, epsabs=0, epsrel=1e-12)[0]
    iyy = integrate.quad(lambda y: rho * math.pi * R2(y) ** 2 / 2, -lim, lim, points=pts, epsabs=0, epsrel=1e-12)[0]
    ixx = integrate.quad(lambda y: rho * math.pi * R2(y) * (R2(y) / 4 + y * y), -lim, lim, points=pts, epsabs=0,
                         epsrel=1e-12)[0]
    return m, ixx, iyy


@pytest.mark.parametrize("r,h", [(0.2, 0.2), (0.35, 0.6), (0.5, 1.0), (0.4, 0.0)])
def test_capsule_mass_properties_by_quadrature(axm, r, h):
    im, ii = axm.mass_properties(axm.SHAPE_CAPSULE, [r, h], 1000.0)
    m, ixx, iyy = _capsule_inertia_quadrature(r, h, 1000.0)
    assert abs(1 / im - m) / m < 1e-6
    assert abs(1 / ii[0] - ixx) / ixx < 1e-6 and abs(1 / ii[2] - ixx) / ixx < 1e-6
    assert abs(1 / ii[1] - iyy) / iyy < 1e-6


def test_sphere_and_box_mass_properties(axm):
    im, ii = axm.mass_properties(axm.SHAPE_SPHERE, [0.5], 1000.0)
    m = 1000 * 4 / 3 * math.pi * 0.125
    assert abs(1 / im - m) / m < 1e-6 and np.allclose(1 / ii, 0.4 * m * 0.25, rtol=1e-6)
    im, ii = axm.mass_properties(axm.SHAPE_BOX, [0.25, 0.5, 0.75], 1000.0)
    m = 1000 * 8 * 0.25 * 0.5 * 0.75
    assert abs(1 / im - m) / m < 1e-6
    # I_xx = m/3 (b^2 + c^2) on half extents (a, b, c)
    assert np.allclose(1 / ii, [m / 3 * (0.25 + 0.5625), m / 3 * (0.0625 + 0.5625), m / 3 * (0.0625 + 0.25)],
                       rtol=1e-6)
    with pytest.raises(axm.AxmError) as e:
        axm.mass_properties(axm.SHAPE_SPHERE, [-1.0])
    assert e.value.code == 300
    with pytest.raises(axm.AxmError) as e:
        axm.mass_properties(9, [1.0])
    assert e.value.code == 300


def test_desc_defaults(axm):
    d = axm.default_desc()
    assert list(d.gravity) == [0.0, np.float32(-9.81), 0.0]  # Y-up (include/axiom/gui/physics_panel.hpp:34-43)
    assert d.warm_start == axm.WARM_CACHE and d.auto_grow == 1
    assert d.num_planes == 1 and list(d.planes[0]) == [0.0, 1.0, 0.0, 0.0]  # ground plane y = 0


def test_scene_config0_matches_its_description(axm):
    """1k bodies, even index sphere r~U(0.25,0.75), odd index box h~U(0.25,0.75)^3, density 1000,
    inside a 30 m cube above the ground, unit quaternions, mu 0.5."""
    b, d = axm.scene(0, seed(0))
    assert b.n == 1000
    assert np.all(b.shape_type[0::2] == axm.SHAPE_SPHERE) and np.all(b.shape_type[1::2] == axm.SHAPE_BOX)
    assert np.all((b.shape_params[0::2, 0] >= 0.25) & (b.shape_params[0::2, 0] <= 0.75))
    assert np.all((b.shape_params[1::2] >= 0.25) & (b.shape_params[1::2] <= 0.75))
    assert np.allclose(np.linalg.norm(b.quat, axis=1), 1.0, atol=1e-6)
    assert np.all(b.pos[:, 1] > 0) and np.ptp(b.pos[:, 0]) <= 30.0 and np.ptp(b.pos[:, 2]) <= 30.0
    for k in range(0, 1000, 97):
        im, ii = axm.mass_properties(int(b.shape_type[k]), b.shape_params[k])
        assert im == b.inv_mass[k] and np.array_equal(ii, b.inv_inertia[k])
    assert abs(d.friction - 0.5) < 1e-7 and abs(d.restitution - 0.1) < 1e-7
    # the same seed gives the same scene; another seed another one
    b2, _ = axm.scene(0, seed(0))
    b3, _ = axm.scene(0, seed(0, 1))
    assert np.array_equal(b.pos, b2.pos) and not np.array_equal(b.pos, b3.pos)
    # seeds differing only in bit 0 share a stream (state = seed | 1, random.hpp:25)
    b4, _ = axm.scene(0, seed(0) + 1)
    assert np.array_equal(b.pos, b4.pos)


def test_scene_config1_pile(axm):
    b, d = axm.scene(1, seed(1), 3000)
    assert b.n == 3000
    t = b.shape_type
    assert np.all(t[0::3] == 0) and np.all(t[1::3] == 1) and np.all(t[2::3] == 2)
    caps = b.shape_params[2::3]
    assert np.all((caps[:, 0] >= 0.2) & (caps[:, 0] <= 0.5) & (caps[:, 1] >= 0.2) & (caps[:, 1] <= 1.0))
    assert np.all(np.abs(b.pos[:, 0]) <= 10) and np.all(np.abs(b.pos[:, 2]) <= 10)
    assert abs(d.friction - 0.6) < 1e-7 and d.num_planes == 5


def test_scene_config3_batch(axm):
    b, d = axm.scene(3, seed(3), 0, 64)
    counts = np.bincount(b.world, minlength=64)
    assert counts.size == 64 and np.all(counts >= 16) and np.all(counts <= 4096)
    # no two worlds share a stream
    firsts = {tuple(b.pos[np.nonzero(b.world == w)[0][0]]) for w in range(64)}
    assert len(firsts) == 64
    assert np.all(np.diff(b.world.astype(np.int64)) >= 0)  # worlds are contiguous
    # each world is a config-0 scene with its own seed
    w0 = np.nonzero(b.world == 5)[0]
    s, _ = axm.scene(0, seed(3, 5 + 1), len(w0))
    assert np.array_equal(s.pos, b.pos[w0]) and np.array_equal(s.shape_params, b.shape_params[w0])


def test_training_log_round_trip(tmp_path):
    """The contact-impulse training log (warmstart.save_log / load_log): rows and description come
    back exactly; rows of the wrong width are refused."""
    import warmstart
    rng = np.random.default_rng(2)
    X = rng.normal(size=(1000, 24)).astype(np.float32)
    Y = rng.normal(size=(1000, 3)).astype(np.float32)
    p = str(tmp_path / "log.npz")
    warmstart.save_log(p, X, Y, "seeded piles, steps 150-600")
    X2, Y2, src = warmstart.load_log(p)
    assert np.array_equal(X, X2) and np.array_equal(Y, Y2) and src == "seeded piles, steps 150-600"
    import pytest
    with pytest.raises(ValueError):
        warmstart.save_log(p, X[:, :20], Y)
