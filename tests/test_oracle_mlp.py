"""Pins of the FP64 network restatement (oracle/mlp_oracle.py): torch's own
float64 layers and autograd on the CPU, central finite differences, torch's
BF16 rounding."""
import numpy as np
import pytest
import torch

import mlp_oracle as mo


def _torch_net(p, nin, wid, depth):
    d = mo.unpack(np.asarray(p, dtype=np.float64), nin, wid, depth)
    layers = []
    fin = nin
    for l in range(1, depth + 1):
        lin = torch.nn.Linear(fin, wid, dtype=torch.float64)
        with torch.no_grad():
            lin.weight.copy_(torch.from_numpy(d[f"W{l}"].T.copy()))
            lin.bias.copy_(torch.from_numpy(d[f"b{l}"].copy()))
        layers += [lin, torch.nn.SiLU()]
        fin = wid
    out = torch.nn.Linear(wid, 1, dtype=torch.float64)
    with torch.no_grad():
        out.weight.copy_(torch.from_numpy(d["w_out"][None, :].copy()))
        out.bias.copy_(torch.from_numpy(d["b_out"].copy()))
    return torch.nn.Sequential(*layers, out)


def _params(rng, nin, wid, depth):
    p = rng.normal(0.0, 1.0, mo.nparams(nin, wid, depth))
    d = mo.unpack(p, nin, wid, depth)
    for name, s in mo.shapes(nin, wid, depth):
        d[name][...] /= np.sqrt(s[0]) if len(s) == 2 or name == "w_out" else 10.0
    return p


@pytest.mark.parametrize("wid", [16, 256])
def test_forward_and_gradient_match_torch(wid):
    rng = np.random.default_rng(5)
    nin, depth, n = 8, 4, 37
    p = _params(rng, nin, wid, depth)
    x = rng.normal(size=(n, nin))
    t = rng.normal(size=n)
    net = _torch_net(p, nin, wid, depth)
    xt = torch.from_numpy(x)
    yt = net(xt)[:, 0]
    np.testing.assert_allclose(mo.forward(p, x, depth), yt.detach().numpy(), rtol=1e-12, atol=1e-12)
    lt = torch.mean((yt - torch.from_numpy(t)) ** 2)
    lt.backward()
    L, g = mo.gradient(p, x, t, depth)
    assert abs(L - lt.item()) < 1e-12 * max(1.0, L)
    gd = mo.unpack(g, nin, wid, depth)
    mods = [m for m in net if isinstance(m, torch.nn.Linear)]
    for l in range(1, depth + 1):
        np.testing.assert_allclose(gd[f"W{l}"], mods[l - 1].weight.grad.numpy().T, rtol=1e-10, atol=1e-13)
        np.testing.assert_allclose(gd[f"b{l}"], mods[l - 1].bias.grad.numpy(), rtol=1e-10, atol=1e-13)
    np.testing.assert_allclose(gd["w_out"], mods[-1].weight.grad.numpy()[0], rtol=1e-10, atol=1e-13)
    np.testing.assert_allclose(gd["b_out"], mods[-1].bias.grad.numpy(), rtol=1e-10, atol=1e-13)


def test_gradient_finite_differences():
    rng = np.random.default_rng(11)
    nin, wid, depth, n = 8, 12, 4, 23
    p = _params(rng, nin, wid, depth)
    x = rng.normal(size=(n, nin))
    t = rng.normal(size=n)
    _, g = mo.gradient(p, x, t, depth)
    h = 1e-6
    for k in rng.choice(p.size, size=60, replace=False):
        q = p.copy()
        q[k] += h
        lp = mo.loss(q, x, t, depth)
        q[k] -= 2 * h
        lm = mo.loss(q, x, t, depth)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1e-3, abs(g[k])) + 1e-9, (k, fd, g[k])


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(3)
    a = np.concatenate([rng.normal(size=5000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 5000),
                        np.array([1.0, 1.00390625, 1.01171875, -3.0e-3, 0.0], dtype=np.float32)])
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(mo.bf16_round(a), ref)


def test_device_weights_round_only_the_bf16_blocks():
    rng = np.random.default_rng(2)
    p = rng.normal(size=mo.nparams()).astype(np.float32)
    q = mo.device_weights(p)
    dp, dq = mo.unpack(p.astype(np.float64)), mo.unpack(q)
    for name, _ in mo.shapes():
        if name in ("W2", "W3", "W4", "w_out"):
            assert np.array_equal(dq[name], mo.bf16_round(dp[name].astype(np.float32)))
            assert not np.array_equal(dq[name], dp[name])
        else:
            assert np.array_equal(dq[name], dp[name])


def test_product_parameter_layout():
    """mcmax_amd.mlp's start parameters follow the flat layout: weights set,
    biases zero"""
    from mcmax_amd import mlp
    assert mlp.NPARAMS == mo.nparams()
    d = mo.unpack(mlp.init_params(4))
    for name, s in mo.shapes():
        if name.startswith("b"):
            assert np.all(d[name] == 0.0), name
        else:
            bound = np.sqrt(6.0 / s[0]) / 1.1
            assert 0.9 * bound < np.abs(d[name]).max() <= bound, name
