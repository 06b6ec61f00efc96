"""Per-body material, damping, kinematic bodies and collision filters on the GPU against the CPU
oracle, through the C ABI: a 1k-body mixed scene with random per-body friction, restitution and
damping, a few kinematic bodies sweeping through it and collision groups / masks; pair lists,
manifolds, colours, friction bounds and the state match over 120 steps."""
import numpy as np
import pytest

from helpers import seed, rel_err
from test_gpu_parity import DT, _compare_step

pytestmark = pytest.mark.gpu


def _extras(b, rng):
    n = b.n
    b.material = np.stack([rng.uniform(0.1, 0.9, n), rng.uniform(0.0, 0.6, n), rng.uniform(0.0, 0.02, n),
                           rng.uniform(0.0, 0.05, n)], 1).astype(np.float32)
    flags = np.zeros(n, np.uint32)
    kin = rng.choice(n, 8, replace=False)
    flags[kin] = 1
    b.inv_mass[kin] = 0.0
    b.inv_inertia[kin] = 0.0
    b.lin_vel[kin] = rng.normal(0, 3.0, (8, 3)).astype(np.float32)
    b.flags = flags
    cat = (1 << rng.integers(0, 3, n)).astype(np.uint32)
    mask = np.full(n, 0xFFFFFFFF, np.uint32)
    mask[rng.random(n) < 0.2] = 0x1          # some bodies only meet category 1
    group = np.zeros(n, np.int32)
    group[rng.random(n) < 0.1] = -3           # a group that never collides with itself
    group[rng.random(n) < 0.05] = 2           # a group that always collides with itself
    b.filter = np.stack([cat, mask, group.view(np.uint32)], 1).astype(np.uint32)
    return kin


def test_body_extras_match_oracle(axm, orc, gpu):
    rng = np.random.default_rng(21)
    b, d = axm.scene(0, seed(0), 1000)
    kin = _extras(b, rng)
    w = axm.World(d)
    w.set_bodies(b)
    ow = orc.OracleWorld(orc.params_from_desc(d))
    ow.set_bodies(b)
    x0 = b.pos[kin].copy()
    for k in range(120):
        _compare_step(axm, w, ow, 10, check_state=False)
        assert np.array_equal(w.friction_bounds(), ow.friction_bounds()), k
        g = w.get_bodies()
        pos, quat, v, av = ow.get_state()
        assert rel_err(g.lin_vel, v) <= 1e-4, k
        assert rel_err(g.pos, pos) <= 1e-5, k
    g = w.get_bodies()
    # kinematic bodies kept their velocity and moved by it
    assert np.array_equal(g.lin_vel[kin], b.lin_vel[kin])
    assert np.allclose(g.pos[kin], x0 + 120 * DT * b.lin_vel[kin], atol=1e-3)
