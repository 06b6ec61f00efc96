"""World snapshots and warm-start network files through the C ABI.

A world saved after some steps and loaded into a fresh world (created with another
description) steps on bit for bit like the original, eagerly and from replayed graphs, with
sleeping islands and residual-stopped sweeps; a network saved mid-training and loaded
elsewhere trains on bit for bit.  Damaged or foreign files are refused with
InvalidParameter and leave the target as it was."""
import numpy as np
import pytest

from helpers import seed

pytestmark = pytest.mark.gpu

DT = 1.0 / 120.0


def _arrays(w):
    g = w.get_bodies()
    lab, asleep, rest = w.islands()
    return [g.pos.copy(), g.quat.copy(), g.lin_vel.copy(), g.ang_vel.copy(), lab.copy(), asleep.copy(), rest.copy()]


def _same(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("graph", [False, True])
def test_world_snapshot_resumes_bit_identically(axm, gpu, tmp_path, graph):
    b, d = axm.scene(0, seed(0), 1000)
    d.allow_sleep = 1
    d.residual_tol = 1e-4
    d.max_iterations = 300
    w = axm.World(d)
    w.set_bodies(b)
    w.run(DT, 300, 90, graph=graph)
    path = str(tmp_path / "world.axm")
    w.save(path)
    its = [w.step(DT, 300).iterations for _ in range(5)]
    st = w.run(DT, 300, 40, graph=graph)
    ref = _arrays(w)
    w.close()
    w2 = axm.World(axm.default_desc())
    w2.load(path)
    its2 = [w2.step(DT, 300).iterations for _ in range(5)]
    st2 = w2.run(DT, 300, 40, graph=graph)
    assert its == its2 and st.iterations == st2.iterations and st.residual == st2.residual
    _same(ref, _arrays(w2))
    assert w2.get_desc().allow_sleep == 1 and np.float32(w2.get_desc().residual_tol) == np.float32(1e-4)
    w2.close()


def test_world_snapshot_of_a_fresh_world_and_a_batch(axm, gpu, tmp_path):
    """No step yet (no warm-start cache in the file), and a batch of worlds."""
    b, d = axm.scene(3, seed(3), 0, 32)
    w = axm.World(d)
    w.set_bodies(b)
    path = str(tmp_path / "batch.axm")
    w.save(path)
    w.run(DT, 10, 30, graph=True)
    ref = _arrays(w)
    w.close()
    w2 = axm.World(axm.default_desc())
    w2.load(path)
    w2.run(DT, 10, 30, graph=True)
    _same(ref, _arrays(w2))
    w2.close()


def test_damaged_snapshot_is_refused(axm, gpu, tmp_path):
    b, d = axm.scene(0, seed(0), 1000)
    w = axm.World(d)
    w.set_bodies(b)
    w.run(DT, 10, 20, graph=False)
    path = tmp_path / "w.axm"
    w.save(str(path))
    raw = bytearray(path.read_bytes())
    raw[len(raw) // 2] ^= 0x40
    bad = tmp_path / "bad.axm"
    bad.write_bytes(bytes(raw))
    short = tmp_path / "short.axm"
    short.write_bytes(bytes(raw[: len(raw) - 9]))
    before = _arrays(w)
    for p in (bad, short, tmp_path / "missing.axm"):
        with pytest.raises(axm.AxmError) as e:
            w.load(str(p))
        assert e.value.code == axm.INVALID_PARAMETER
    _same(before, _arrays(w))
    w.step(DT, 10)  # still usable
    w.close()


def test_network_file_resumes_training(axm, gpu, tmp_path):
    rng = np.random.default_rng(5)
    X = rng.normal(size=(20000, axm.MLP_INPUTS)).astype(np.float32)
    Y = rng.normal(size=(20000, axm.MLP_OUTPUTS)).astype(np.float32)
    net = axm.Mlp(0, seed=9)
    net.train(X, Y, epochs=1, batch=1024, lr=1e-3)
    path = str(tmp_path / "net.axmmlp")
    net.save(path)
    net.train(X, Y, epochs=1, batch=1024, lr=1e-3)
    ref = net.params
    net2 = axm.Mlp(0, seed=1)
    net2.load(path)
    net2.train(X, Y, epochs=1, batch=1024, lr=1e-3)
    assert np.array_equal(ref, net2.params)
    raw = bytearray(open(path, "rb").read())
    raw[100] ^= 1
    bad = tmp_path / "bad.axmmlp"
    bad.write_bytes(bytes(raw))
    with pytest.raises(axm.AxmError) as e:
        net2.load(str(bad))
    assert e.value.code == axm.INVALID_PARAMETER


def test_reload_into_the_same_world_replays_its_graphs(axm, gpu, tmp_path):
    """Reloading a world's own snapshot keeps its captured step graphs (nothing they hold
    changed): the replayed steps after the reload end bit for bit where the first ones did, also
    with profiling (eagerly launched sweep kernels between two graphs) on."""
    b, d = axm.scene(2, seed(2), 200000)
    d.residual_tol = 1e-4
    d.max_iterations = 300
    w = axm.World(d)
    w.set_bodies(b)
    w.run(DT, 300, 12, graph=True)
    path = str(tmp_path / "same.axm")
    w.save(path)
    outs = []
    for prof in (False, True, True):
        w.load(path)
        w.profile(prof)
        st = w.run(DT, 300, 15, graph=True)
        if prof:
            p = w.profile(False)
            assert p.steps == 15 and p.bundle_launches == 15 and p.bundle_point_sweeps > 0
        outs.append((st.iterations, _arrays(w)))
    for it, arrs in outs[1:]:
        assert it == outs[0][0]
        _same(outs[0][1], arrs)
    w.close()
