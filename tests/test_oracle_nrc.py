"""Pins of the radiance-cache network oracle (oracle/nrc.cpp).

* gradient: central finite differences (eps 1e-3, fp64 evaluation) of the loss with its
  denominators frozen at the base point, which is the objective the stop-gradient loss
  differentiates;
* forward: an identity-like network whose output is the encoding itself;
* encoding: hand-evaluated triangle-wave and one-blob values;
* rounding emulation: torch's bfloat16 and numpy's float16 conversions (round to nearest even).
"""
import numpy as np
import pytest

import oracle_lib

P = 21504
HIDDEN = 5


def w_index(l, i, o):
    return l * 4096 + i * 64 + o if l < HIDDEN else HIDDEN * 4096 + i * 16 + o


def random_params(seed=1, scale=1.0):
    """Xavier-uniform hidden layers and a 64x3 output layer (padding columns zero)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(P, np.float32)
    s = scale * np.sqrt(6.0 / 128)
    p[:HIDDEN * 4096] = rng.uniform(-s, s, HIDDEN * 4096)
    out = np.zeros((64, 16), np.float32)
    out[:, :3] = rng.uniform(-scale * np.sqrt(6.0 / 67), scale * np.sqrt(6.0 / 67), (64, 3))
    p[HIDDEN * 4096:] = out.ravel()
    return p


def random_raw(n, seed=2):
    rng = np.random.default_rng(seed)
    raw = np.zeros((n, 16), np.float32)
    raw[:, 0:8] = rng.uniform(0, 1, (n, 8))
    raw[:, 8:14] = rng.uniform(0, 1, (n, 6)) * (rng.uniform(0, 1, (n, 1)) < 0.7)
    return raw


def test_param_count():
    assert oracle_lib.nrc_param_count() == P


def test_encoding_hand_values():
    raw = np.zeros(16, np.float32)
    raw[0:3] = [0.3, 0.8, 0.0]
    raw[3], raw[4] = 0.125, 0.625
    raw[7] = 0.875
    raw[8:14] = [0.1, 0.2, 0.3, 0.4, 0.5, 0.6]
    e = oracle_lib.nrc_encode(raw)[0]
    # triangle waves |2 frac(2^(f-1) x) - 1|
    assert e[0] == pytest.approx(0.7, abs=1e-6)      # x = 0.15
    assert e[1] == pytest.approx(0.4, abs=1e-6)      # x = 0.3
    assert e[2] == pytest.approx(0.2, abs=1e-6)      # x = 0.6
    assert e[12] == pytest.approx(0.2, abs=1e-6)     # 0.8 / 2 = 0.4
    assert e[24] == 1.0 and e[25] == 1.0              # coordinate 0: frac = 0
    # one-blob with 4 bins of width 1/4 and a triweight kernel (1 - d^2/9)^3: s = 0.125 sits on bin 0
    assert e[36] == 1.0
    assert e[37] == pytest.approx((8 / 9) ** 3, rel=1e-6)
    assert e[38] == pytest.approx((5 / 9) ** 3, rel=1e-6)
    assert e[39] == 0.0
    assert e[40 + 2] == 1.0                            # s = 0.625 sits on bin 2
    assert e[52 + 3] == 1.0                            # roughness 0.875 on bin 3
    assert list(e[56:62]) == pytest.approx([0.1, 0.2, 0.3, 0.4, 0.5, 0.6])
    assert e[62] == 1.0 and e[63] == 1.0


def identity_params():
    p = np.zeros(P, np.float32)
    for l in range(HIDDEN):
        for i in range(64):
            p[w_index(l, i, i)] = 1.0
    for c in range(3):
        p[w_index(HIDDEN, 56 + c, c)] = 1.0
    return p


def test_forward_identity_network_returns_albedo():
    raw = random_raw(16, seed=5)
    y = oracle_lib.nrc_forward(identity_params(), raw)
    np.testing.assert_allclose(y, raw[:, 8:11], rtol=0, atol=1e-7)


def test_rounding_emulation_matches_library_conversions():
    torch = pytest.importorskip("torch")
    raw = random_raw(64, seed=6)
    p = identity_params()
    y_bf = oracle_lib.nrc_forward(p, raw, "bf16")
    y_h = oracle_lib.nrc_forward(p, raw, "fp16")
    ref_bf = torch.from_numpy(raw[:, 8:11].copy()).to(torch.bfloat16).float().numpy()
    ref_h = raw[:, 8:11].astype(np.float16).astype(np.float32)
    np.testing.assert_array_equal(y_bf.astype(np.float32), ref_bf)
    np.testing.assert_array_equal(y_h.astype(np.float32), ref_h)
    assert not np.array_equal(ref_bf, raw[:, 8:11])  # the emulation actually rounds


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(7)
    n = 12
    params = random_params(seed=3, scale=1.5)
    raw = random_raw(n, seed=4)
    target = rng.uniform(0, 2, (n, 3)).astype(np.float32)
    loss, grad, y = oracle_lib.nrc_grad(params, raw, target)
    denom = y * y
    # every output-layer weight plus a random sample of hidden weights (all layers)
    live = [w_index(HIDDEN, k, c) for k in range(64) for c in range(3)]
    idx = np.concatenate([np.array(live), rng.choice(HIDDEN * 4096, 600, replace=False)])
    eps = 1e-3
    fd = np.zeros(idx.size)
    for m, j in enumerate(idx):
        pp, pm = params.copy(), params.copy()
        pp[j] += eps
        pm[j] -= eps
        lp, _, _ = oracle_lib.nrc_grad(pp, raw, target, denom_ref=denom, want_grad=False)
        lm, _, _ = oracle_lib.nrc_grad(pm, raw, target, denom_ref=denom, want_grad=False)
        fd[m] = (lp - lm) / (float(pp[j]) - float(pm[j]))
    err = np.abs(fd - grad[idx]) / np.abs(grad).max()
    # central differences are O(eps^2) accurate away from ReLU kinks; a kink inside [w-eps, w+eps]
    # can spoil a handful of entries
    assert np.quantile(err, 0.99) < 1e-5, np.sort(err)[-10:]
    assert np.abs(grad).max() > 0
    # padding columns of the output layer get no gradient
    pad = [w_index(HIDDEN, k, c) for k in range(64) for c in range(3, 16)]
    assert not grad[pad].any()


def test_loss_is_relative_l2():
    raw = random_raw(5, seed=8)
    params = random_params(seed=9)
    y = oracle_lib.nrc_forward(params, raw)
    t = np.ones((5, 3), np.float32)
    loss, _, y2 = oracle_lib.nrc_grad(params, raw, t, want_grad=False)
    np.testing.assert_allclose(y2, y, rtol=1e-15)
    assert loss == pytest.approx(np.sum((y - 1.0) ** 2 / (y * y + 0.01)) / 5, rel=1e-12)
