"""Warm-start network: the oracle's backpropagation pinned by central finite
differences (CPU, FP64), BF16 rounding pinned by known values; the device
forward / gradient against the oracle and a training run (GPU)."""
import numpy as np
import pytest

from oracle import mlp as OM

DIMS = [6, 32, 32, 24, 10]


def _params(dims, seed):
    rng = np.random.default_rng(seed)
    n = sum(dims[l] * dims[l + 1] for l in range(len(dims) - 1)) + sum(dims[1:])
    p = np.zeros(n)
    off = 0
    for l in range(len(dims) - 1):
        k = dims[l] * dims[l + 1]
        p[off:off + k] = rng.standard_normal(k) * np.sqrt(2.0 / dims[l])
        off += k
    p[off:] = 0.1 * rng.standard_normal(n - off)
    return p


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-3], dtype=np.float32)
    r = OM.bf16(x)
    # ties go to even: 1 + 2^-8 -> 1, 1 + 3 2^-8 -> 1 + 2^-6; 1 + 2^-7 exact
    assert list(r[:4]) == [1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6]
    assert r[4] == -2.5
    assert abs(r[5] - 3.0e-3) < 3.0e-3 * 2 ** -8


def test_oracle_gradient_finite_differences():
    rng = np.random.default_rng(1)
    p = _params(DIMS, 2)
    x = rng.standard_normal((7, DIMS[0]))
    t = rng.standard_normal((7, DIMS[-1]))
    loss, g = OM.loss_and_grad(p, DIMS, x, t, rounding=False)
    idx = rng.choice(p.size, 60, replace=False)
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (OM.loss_and_grad(pp, DIMS, x, t, rounding=False)[0] -
              OM.loss_and_grad(pm, DIMS, x, t, rounding=False)[0]) / (2 * h)
        assert abs(fd - g[i]) < 1e-7 + 1e-5 * abs(fd)


@pytest.mark.gpu
@pytest.mark.parametrize("rows", [1, 7, 64, 200])
def test_device_forward_and_gradient(rows):
    from cfft_hip.mlp import MLP
    dims = [6, 512, 512, 512, 256]
    net = MLP(dims, seed=3)
    p = net.params()
    rng = np.random.default_rng(rows)
    x = rng.standard_normal((rows, 6)).astype(np.float32)
    t = rng.standard_normal((rows, 256)).astype(np.float32)
    y = net.forward(x)
    yr = OM.forward(p, dims, x)[0][-1]
    scale = max(1.0, float(np.max(np.abs(yr))))
    # activations that land within an FP32 rounding of a BF16 rounding boundary
    # can round the other way (one BF16 ulp, 2^-8): a few outputs move by that
    # much, the bulk agrees to FP32 accumulation error
    assert np.max(np.abs(y - yr)) < 5e-3 * scale
    assert np.mean(np.abs(y - yr)) < 2e-5 * scale
    loss, g = net.gradient(x, t)
    lr, gr = OM.loss_and_grad(p, dims, x, t)
    assert abs(loss - lr) < 1e-4 * lr
    off = 0
    for l in range(len(dims) - 1):
        k = dims[l] * dims[l + 1]
        a, b = g[off:off + k], gr[off:off + k]
        assert np.linalg.norm(a - b) < 2e-3 * np.linalg.norm(b) + 1e-12, f"W{l}"
        off += k
    assert np.linalg.norm(g[off:] - gr[off:]) < 2e-3 * np.linalg.norm(gr[off:])


@pytest.mark.gpu
def test_device_training_fits_smooth_map():
    """Adam on a smooth 6 -> 16 map: the loss falls by 20x and the trained
    network predicts held-out points."""
    from cfft_hip.mlp import MLP
    rng = np.random.default_rng(0)
    A = rng.standard_normal((6, 16))
    X = rng.uniform(-1, 1, (1200, 6)).astype(np.float32)
    Y = np.sin(X @ A).astype(np.float32)
    net = MLP([6, 128, 128, 128, 16], seed=1)
    hist = net.train(X[:1000], Y[:1000], epochs=60, batch=64, lr=2e-3, seed=2)
    assert hist[-1] < hist[0] / 20
    err = np.mean((net.forward(X[1000:]) - Y[1000:]) ** 2)
    assert err < 3 * hist[-1] + 1e-3


def test_committed_warm_start_model_file():
    """The trained warm-start network ships as plain arrays (no pickle): a
    4-layer MLP 6 -> 512 -> 512 -> 512 -> 256 trained on 2000 solutions at 64^3,
    whose recorded held-out evaluation met the 0.5x wall-time target."""
    import json
    import os
    path = os.path.join(os.path.dirname(__file__), "..", "crystallographicfft.jl_amd", "models",
                        "warmstart.npz")
    d = np.load(path, allow_pickle=False)
    dims = [int(v) for v in d["dims"]]
    assert dims == [6, 512, 512, 512, 256]
    n_params = sum(a * b + b for a, b in zip(dims[:-1], dims[1:]))
    assert d["params"].size == n_params and d["params"].dtype == np.float32
    assert np.all(np.isfinite(d["params"])) and d["sd"].min() > 0
    meta = json.loads(str(d["meta"]))
    assert meta["grid"] == 64 and meta["n_train"] == 2000 and meta["n_stars"] == 64
    ev = meta["eval"]
    assert ev["n"] == 200 and ev["fallbacks"] == 0 and ev["time_ratio"] <= 0.5
