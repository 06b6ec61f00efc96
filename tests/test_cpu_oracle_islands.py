"""Islands and sleeping on the CPU oracle: labels of hand-built scenes (components of the awake
dynamic bodies over the manifolds, labelled by their smallest body index), the rest-time rule,
an island falling asleep after sleep_time seconds, and a sleeping island woken the step after an
awake body touches it."""
import numpy as np

from helpers import seed
from test_cpu_oracle_world import DT, make, world

NONE = 0x7FFFFFFF


def test_labels_are_smallest_index_of_each_component(orc, axm):
    """Two separate stacks (bodies 3-1-4 and 0-2) on the ground and one free body 5: labels
    1 for {1, 3, 4}, 0 for {0, 2} and 5 for {5}; the ground does not join the stacks."""
    specs = [(1, [0.5, 0.5, 0.5], [10, 1.49, 0], [0, 0, 0], [0, 0, 0], False),   # 0: stack B, top
             (1, [0.5, 0.5, 0.5], [0, 1.49, 0], [0, 0, 0], [0, 0, 0], False),    # 1: stack A, middle
             (1, [0.5, 0.5, 0.5], [10, 0.5, 0], [0, 0, 0], [0, 0, 0], False),    # 2: stack B, bottom
             (1, [0.5, 0.5, 0.5], [0, 0.5, 0], [0, 0, 0], [0, 0, 0], False),     # 3: stack A, bottom
             (1, [0.5, 0.5, 0.5], [0, 2.48, 0], [0, 0, 0], [0, 0, 0], False),    # 4: stack A, top
             (0, [0.5], [-10, 5.0, 0], [0, 0, 0], [0, 0, 0], False)]             # 5: falling alone
    b = make(axm, specs)
    ow, _ = world(orc, axm, b)
    st = ow.step(DT, 20)
    lab, asleep, rest = ow.islands()
    assert lab.tolist() == [0, 1, 0, 1, 1, 5]
    assert st.islands == 3 and st.sleeping == 0 and not asleep.any()


def test_static_and_plane_partners_do_not_connect(orc, axm):
    """Two spheres both resting on one static box: each is its own island."""
    specs = [(1, [2.0, 0.5, 2.0], [0, 0.5, 0], [0, 0, 0], [0, 0, 0], True),
             (0, [0.5], [-1, 1.49, 0], [0, 0, 0], [0, 0, 0], False),
             (0, [0.5], [1, 1.49, 0], [0, 0, 0], [0, 0, 0], False)]
    b = make(axm, specs)
    ow, _ = world(orc, axm, b)
    st = ow.step(DT, 10)
    lab, asleep, rest = ow.islands()
    assert lab[0] == NONE and lab[1] == 1 and lab[2] == 2 and st.islands == 2


def test_resting_island_falls_asleep_and_wakes_on_contact(orc, axm):
    """A sphere resting on the ground: its rest time grows by dt per still step (|v|, |w| below the
    thresholds) and it falls asleep once that reaches sleep_time (velocities zeroed).  A second sphere
    released 5 m above it lands about a second later: in the step of their first manifold the lower
    sphere is still asleep (it acts as static), in the next step it is awake again."""
    r = 0.5
    b = make(axm, [(0, [r], [0, r, 0], [0, 0, 0], [0, 0, 0], False),
                   (0, [0.4], [0.0, 2 * r + 0.4 + 5.0, 0.0], [0, 0, 0], [0, 0, 0], False)])
    ow, d = world(orc, axm, b, allow_sleep=1, sleep_time=0.25)
    slept_at = contact_at = None
    prev_rest = 0.0
    for k in range(240):
        st = ow.step(DT, 10)
        lab, asleep, rest = ow.islands()
        if slept_at is None:
            if asleep[0]:
                slept_at = k
                assert rest[0] >= 0.25 and st.sleeping == 1
                pos, quat, v, w = ow.get_state()
                assert np.all(v[0] == 0) and np.all(w[0] == 0)
            else:
                # the rest time either restarts at 0 or grows by exactly dt
                assert rest[0] == 0.0 or abs(rest[0] - (prev_rest + np.float32(DT))) < 1e-6
                prev_rest = rest[0]
            continue
        m = ow.manifolds()
        if contact_at is None and (0, 1) in set(zip(m["a"].tolist(), m["b"].tolist())):
            contact_at = k
            assert asleep[0]  # acted as static in this step
            continue
        if contact_at is not None:
            assert not asleep[0] and lab[0] == lab[1] == 0
            break
    assert slept_at is not None and slept_at < 60 and contact_at is not None


def _settled_oracle(orc, axm, config, n, steps):
    b, d = axm.scene(config, seed(config), n)
    ow = orc.OracleWorld(orc.params_from_desc(d))
    ow.set_bodies(b)
    ow.set_threads(4)
    for _ in range(steps):
        ow.step(DT, 20)
    pos, quat, v, w = ow.get_state()
    s = b.copy()
    s.pos, s.quat, s.lin_vel, s.ang_vel = pos, quat, v, w
    return s, d


def test_fixed_sweeps_unchanged_by_island_split(orc, axm):
    """Without a tolerance every island runs every sweep, and islands share no dynamic body, so each
    body sees the same visits in the same order: bit-identical steps with and without the split."""
    s, d = _settled_oracle(orc, axm, 0, 1000, 100)
    out = []
    for split in (0, 1):
        d.island_solve = split
        ow = orc.OracleWorld(orc.params_from_desc(d))
        ow.set_bodies(s)
        for _ in range(3):
            st = ow.step(DT, 15)
        out.append((ow.get_state(), st.residual, st.iterations))
    assert out[0][1] == out[1][1] and out[0][2] == out[1][2] == 15
    for x, y in zip(out[0][0], out[1][0]):
        assert np.array_equal(x, y)


def test_island_stops_never_need_more_sweeps(orc, axm):
    """With a tolerance, an island solved on its own stops at the first sweep that leaves its own
    residual <= tol; solved together, the step stops at the first sweep where every island is below
    it, which no island reaches before its own first.  So from the same state the split step runs
    at most as many sweeps, and both reach the tolerance."""
    s, d = _settled_oracle(orc, axm, 0, 1000, 100)
    res = {}
    for split in (0, 1):
        d.island_solve = split
        ow = orc.OracleWorld(orc.params_from_desc(d))
        ow.set_bodies(s)
        ow.set_tolerance(1e-4, 300, 1)
        st = ow.step(DT, 300)
        res[split] = st
    assert res[0].residual <= 1e-4 and res[1].residual <= 1e-4
    assert res[1].iterations <= res[0].iterations
    assert res[0].manifolds == res[1].manifolds > 50
