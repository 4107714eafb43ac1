"""FP64 restatement of the warm-start network (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU leg use this file, as
the checker of the device network in mcmax_amd/csrc/mlp.hip; nothing in the
product imports it.

The network has no counterpart in the reference (MCMax starts the structure
iteration from the fixed guess of Init.f:3935); it maps 8 cell features to
log T: 8 -> 4 x 256 SiLU -> 1.  Parameters are one flat vector in the layout
of include/mcx.h: W1 [8][256], b1, then per hidden layer l = 2..4
W_l [256 in][256 out] and b_l, then w_out [256] and b_out.

`forward` / `gradient` evaluate the network in FP64.  `device_weights`
rounds the weights the device reads in BF16 (W_2..W_4 and w_out) to BF16
(round to nearest even), so that a comparison with the device measures only
the rounding of its activations.

Pinned by (tests/test_oracle_mlp.py): torch's own float64 Linear/SiLU layers
and autograd on the CPU (an independent implementation of the same
function and of its gradient), and central finite differences of the loss.
"""
import numpy as np

NIN, WID, DEPTH = 8, 256, 4


def shapes(nin=NIN, wid=WID, depth=DEPTH):
    """(name, shape) of every parameter block in the flat vector's order."""
    out = [("W1", (nin, wid)), ("b1", (wid,))]
    for l in range(2, depth + 1):
        out += [(f"W{l}", (wid, wid)), (f"b{l}", (wid,))]
    out += [("w_out", (wid,)), ("b_out", (1,))]
    return out


def nparams(nin=NIN, wid=WID, depth=DEPTH):
    return sum(int(np.prod(s)) for _, s in shapes(nin, wid, depth))


def unpack(p, nin=NIN, wid=WID, depth=DEPTH):
    """dict name -> view of the flat vector p."""
    p = np.asarray(p)
    out, o = {}, 0
    for name, s in shapes(nin, wid, depth):
        k = int(np.prod(s))
        out[name] = p[o:o + k].reshape(s)
        o += k
    assert o == p.size, (o, p.size)
    return out


def bf16_round(a):
    """float32 -> nearest BF16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def device_weights(p, nin=NIN, wid=WID, depth=DEPTH):
    """the parameters as the device evaluates them: W_2..W_depth and w_out in
    BF16, layer 1 and the biases in FP32"""
    q = np.asarray(p, dtype=np.float32).copy()
    d = unpack(q, nin, wid, depth)
    for name in [f"W{l}" for l in range(2, depth + 1)] + ["w_out"]:
        d[name][...] = bf16_round(d[name])
    return q.astype(np.float64)


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s + z * s * (1.0 - s)


def forward(p, x, depth=DEPTH, keep=False, act_bf16=False):
    """y [n] for x [n][nin] in float64; keep=True also returns the
    pre-activations of the hidden layers.  act_bf16=True rounds every hidden
    layer's output to BF16 as the device stores it (everything else stays
    float64), which leaves only the device's FP32 accumulation as a
    difference."""
    x = np.asarray(x, dtype=np.float64)
    nin = x.shape[1]
    wid = _width(len(p), nin, depth)
    d = unpack(np.asarray(p, dtype=np.float64), nin, wid, depth)
    h, zs = x, []
    for l in range(1, depth + 1):
        z = h @ d[f"W{l}"] + d[f"b{l}"]
        zs.append(z)
        h = _silu(z)
        if act_bf16:
            h = bf16_round(h.astype(np.float32)).astype(np.float64)
    y = h @ d["w_out"] + d["b_out"][0]
    return (y, zs) if keep else y


def _width(npar, nin, depth):
    # npar = nin*w + w + (depth-1)(w*w + w) + w + 1
    a, b, c = depth - 1, nin + 2 + (depth - 1), 1 - npar
    w = int(round((-b + np.sqrt(b * b - 4 * a * c)) / (2 * a)))
    assert nparams(nin, w, depth) == npar
    return w


def loss(p, x, t, depth=DEPTH):
    """mean squared error sum_r (y_r - t_r)^2 / n"""
    y = forward(p, x, depth)
    return float(np.mean((y - np.asarray(t, dtype=np.float64)) ** 2))


def gradient(p, x, t, depth=DEPTH):
    """(loss, dloss/dp) by backpropagation in float64, dp in the flat layout"""
    x = np.asarray(x, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    n, nin = x.shape
    wid = _width(len(p), nin, depth)
    d = unpack(np.asarray(p, dtype=np.float64), nin, wid, depth)
    y, zs = forward(p, x, depth, keep=True)
    g = np.zeros(len(p))
    gd = unpack(g, nin, wid, depth)
    e = y - t
    dy = 2.0 * e / n
    hs = [x] + [_silu(z) for z in zs]
    gd["w_out"][...] = hs[depth].T @ dy
    gd["b_out"][...] = dy.sum()
    dh = np.outer(dy, d["w_out"])
    for l in range(depth, 0, -1):
        dz = dh * _dsilu(zs[l - 1])
        gd[f"W{l}"][...] = hs[l - 1].T @ dz
        gd[f"b{l}"][...] = dz.sum(axis=0)
        dh = dz @ d[f"W{l}"].T
    return float(np.mean(e * e)), g
