"""Learned warm start: training-data logging from the plain solver, training of the
network on the GPU (axm_mlp_train) and the iterations-to-tolerance comparison.

The network predicts each contact point's change of impulse from the previous
step's (in units of mN * 9.81 * dt) from 24 per-point inputs computed by the
step's preparation kernel; see include/axiom/axm.h (axm_mlp_*) for the inputs
and the packing of the weights.
"""
import numpy as np

import axm


def collect(bodies, desc, steps, sweeps, dt=1.0 / 120.0, every=1, tol=0.0, max_sweeps=400):
    """Runs the plain solver (previous-step warm start) and returns (X [n, 24], Y [n, 3]): the
    inputs of every contact point and its change of impulse / u.  With tol > 0 every step is
    iterated to residual <= tol (at most max_sweeps), so the targets are the converged impulses;
    otherwise each step runs `sweeps` sweeps."""
    d = axm.default_desc()
    for f, _ in axm.WorldDesc._fields_:
        setattr(d, f, getattr(desc, f))
    d.warm_start = axm.WARM_CACHE
    d.residual_tol = tol
    d.max_iterations = max_sweeps
    w = axm.World(d)
    w.set_bodies(bodies)
    w.set_feature_logging(True)
    xs, ys = [], []
    for k in range(steps):
        w.step(dt, sweeps if tol <= 0 else max_sweeps)
        if k % every:
            continue
        X, Y = w.features()
        m = w.manifolds()
        valid = np.arange(4)[None, :] < m["npts"][:, None]
        xs.append(X[valid])
        ys.append(Y[valid])
    w.close()
    if not xs:
        return np.zeros((0, axm.MLP_INPUTS), np.float32), np.zeros((0, axm.MLP_OUTPUTS), np.float32)
    return np.concatenate(xs), np.concatenate(ys)


def fold_normalisation(params, mu, sigma, yscale):
    """Network trained on (X - mu) / sigma with targets Y / yscale -> the same network on raw X and Y:
    W1 <- W1 / sigma, b1 <- b1 - W1 (mu / sigma), W3 <- yscale W3, b3 <- yscale b3."""
    p = np.array(params, np.float64)
    W1 = p[0:3072].reshape(128, 24)
    b1 = p[3072:3200]
    W3 = p[19712:20096].reshape(3, 128)
    b3 = p[20096:20099]
    b1[:] = b1 - W1 @ (mu / sigma)
    W1[:] = W1 / sigma[None, :]
    W3[:] = W3 * yscale[:, None]
    b3[:] = b3 * yscale
    return p.astype(np.float32)


def train(X, Y, epochs=5, batch=4096, lr=1e-3, seed=1, device=0, shuffle_seed=7, clip=None):
    """Trains a fresh network on (X, Y) with minibatch Adam on the GPU and returns (Mlp, per-epoch loss).
    Inputs are standardised and targets scaled per output for training; the scalings are folded into
    the first and last layers afterwards, so the returned network takes raw inputs.  `clip` bounds
    |Y| (in units of u) to keep rare impact impulses from dominating the squared loss."""
    X = np.asarray(X, np.float64)
    Y = np.asarray(Y, np.float64)
    if clip is not None:
        Y = np.clip(Y, -clip, clip)
    mu = X.mean(0)
    sigma = np.maximum(X.std(0), 1e-3)
    ys = np.maximum(Y.std(0), 1e-3)
    Xn = ((X - mu) / sigma).astype(np.float32)
    Yn = (Y / ys).astype(np.float32)
    rng = np.random.default_rng(shuffle_seed)
    perm = rng.permutation(len(Xn))
    net = axm.Mlp(device, seed=seed)
    losses = net.train(Xn[perm], Yn[perm], epochs=epochs, batch=batch, lr=lr)
    net.params = fold_normalisation(net.params, mu, sigma, ys)
    return net, losses


def iterations_to_tolerance(bodies, desc, steps, tol, max_iters, warm, net=None, warmup=0, dt=1.0 / 120.0,
                            every=1, timing=None):
    """Steps a world with the residual-stopped solver (residual checked every `every` sweeps) and
    returns the sweeps each of the `steps` counted steps needed (capped at max_iters) and whether it
    reached `tol`; `warmup` earlier steps run the same way uncounted.  `warm` is axm.WARM_CACHE or
    axm.WARM_MLP.  With a dict as `timing`, the wall time of the counted steps goes to timing["s"]."""
    import time
    d = axm.default_desc()
    for f, _ in axm.WorldDesc._fields_:
        setattr(d, f, getattr(desc, f))
    d.warm_start = warm
    d.residual_tol = tol
    d.residual_every = every
    d.max_iterations = max_iters
    w = axm.World(d)
    w.set_bodies(bodies)
    if net is not None:
        w.set_mlp(net)
    iters, reached = [], []
    t0 = None
    for k in range(warmup + steps):
        if k == warmup:
            w.sync()
            t0 = time.perf_counter()
        st = w.step(dt, max_iters, axm.STEP_RESIDUAL)
        if k >= warmup:
            iters.append(st.iterations)
            reached.append(st.residual <= tol)
    w.sync()
    if timing is not None and t0 is not None:
        timing["s"] = time.perf_counter() - t0
    w.close()
    return np.array(iters), np.array(reached)


def train_on_config0(seeds=16, steps=240, epochs=10, batch=2048, lr=1e-3, tol=1e-5):
    """The small training run: contacts logged from the plain solver iterated to residual <= tol on
    `seeds` seeded 1k-body mixed scenes (world seeds 11..10+seeds of configuration 0), then
    minibatch Adam on the GPU."""
    base = 0x4158494F4D000000
    xs, ys = [], []
    for sd in range(1, seeds + 1):
        b, d = axm.scene(0, base + 2 * (10 + sd), 1000)
        X, Y = collect(b, d, steps, 0, tol=tol)
        xs.append(X)
        ys.append(Y)
    X, Y = np.concatenate(xs), np.concatenate(ys)
    net, losses = train(X, Y, epochs=epochs, batch=batch, lr=lr, clip=64.0)
    return net, losses, len(X)


def collect_pile(seed, rows_wanted, steps=600, first=150, every=5, tol=1e-5, max_sweeps=400, rng=None,
                 dt=1.0 / 120.0, settle_sweeps=20):
    """Contacts logged on one seeded dense pile (100k spheres / boxes / capsules, the pile
    configuration): the pile is stepped with `settle_sweeps` sweeps, and every `every`-th step from
    step `first` on is solved to residual <= tol (the converged impulses are the targets); each
    logged step's contact points are sampled uniformly down to the remaining rows' even share
    over the remaining logged steps."""
    b, d = axm.scene(1, seed, 0)
    d.warm_start = axm.WARM_CACHE
    w = axm.World(d)
    w.set_bodies(b)
    d_plain = axm.WorldDesc.from_buffer_copy(d)
    d_tol = axm.WorldDesc.from_buffer_copy(d)
    d_tol.residual_tol = tol
    d_tol.max_iterations = max_sweeps
    logged = [k for k in range(first, steps) if (k - first) % every == 0]
    rng = rng or np.random.default_rng(seed & 0xFFFF)
    xs, ys = [], []
    w.run(dt, settle_sweeps, first, graph=True)
    w.set_feature_logging(True)
    k = first
    left = rows_wanted
    for i, nxt in enumerate(logged):
        if nxt > k:
            w.run(dt, settle_sweeps, nxt - k, graph=True)
        w.configure(d_tol)
        w.step(dt, max_sweeps)
        w.configure(d_plain)
        k = nxt + 1
        X, Y = w.features()
        m = w.manifolds()
        valid = np.arange(4)[None, :] < m["npts"][:, None]
        X, Y = X[valid], Y[valid]
        share = -(-left // (len(logged) - i))
        if len(X) > share:
            pick = np.sort(rng.choice(len(X), share, replace=False))
            X, Y = X[pick], Y[pick]
        left -= len(X)
        xs.append(X)
        ys.append(Y)
    w.close()
    return np.concatenate(xs), np.concatenate(ys)


TRAIN_ROWS = 20_000_000
TRAIN_DATA = ("20M contact points from 4 seeded 100k-body dense piles (steps 150-600, every 5th step solved "
              "to residual <= 1e-5, the converged impulses as targets), Adam lr 1e-3, 5 epochs, batch 4096")


def train_config1(rows=TRAIN_ROWS, epochs=5, batch=4096, lr=1e-3, seeds=4, log_path=None):
    """The warm-start network of the learned form: trained on `rows` contact points logged on
    `seeds` seeded dense piles (held-out seeds are used for evaluation).  With `log_path` the
    training log is also written there (save_log)."""
    base = 0x4158494F4D000000 + 1000
    xs, ys = [], []
    for k in range(seeds):
        X, Y = collect_pile(base + 2 * (100 + k), (rows - sum(len(x) for x in xs)) // (seeds - k))
        xs.append(X)
        ys.append(Y)
    X, Y = np.concatenate(xs), np.concatenate(ys)
    if log_path:
        save_log(log_path, X, Y, TRAIN_DATA)
    net, losses = train(X, Y, epochs=epochs, batch=batch, lr=lr, clip=64.0)
    return net, losses, len(X)


# Training log file: a numpy .npz (no pickled objects) holding
#   version  int32 [1]                      format version, 1
#   X        float32 [n, 24]                the step's network inputs per contact point
#   Y        float32 [n, 3]                 (final - previous impulse) / u per contact point
#   source   uint8 [k]                      UTF-8 description of how the rows were logged
LOG_VERSION = 1


def save_log(path, X, Y, source=""):
    X = np.ascontiguousarray(X, np.float32)
    Y = np.ascontiguousarray(Y, np.float32)
    if X.ndim != 2 or X.shape[1] != axm.MLP_INPUTS or Y.shape != (X.shape[0], axm.MLP_OUTPUTS):
        raise ValueError("training log rows must be [n, 24] inputs and [n, 3] targets")
    with open(path, "wb") as f:
        np.savez(f, version=np.array([LOG_VERSION], np.int32), X=X, Y=Y,
                 source=np.frombuffer(source.encode("utf-8"), np.uint8))


def load_log(path):
    """(X, Y, source) of a training log written by save_log."""
    with np.load(path, allow_pickle=False) as z:
        if int(z["version"][0]) != LOG_VERSION:
            raise ValueError("training log version %d not supported" % int(z["version"][0]))
        X, Y = z["X"], z["Y"]
        if X.ndim != 2 or X.shape[1] != axm.MLP_INPUTS or Y.shape != (X.shape[0], axm.MLP_OUTPUTS):
            raise ValueError("training log rows must be [n, 24] inputs and [n, 3] targets")
        return X, Y, bytes(z["source"]).decode("utf-8")


def train_default():
    """The training run the benchmark uses."""
    return train_config1()
