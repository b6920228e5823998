"""Oracle (test infrastructure only) for the learned initial guess.

CPU restatement in numpy of lywell_amd/csrc/lw_mlp.hip: the pointwise MLP
(in -> hidden x depth (tanh) -> out), its backward pass, and the guess head
    guess = (1 + o0 + j o1) xt1 + (o2 + j o3) (xt1 - xt2)
with xt_a = x_a exp(-j k (d - d_a) . e_centro), the reference's warm start
(main.f90:1017-1022) applied to the two nearest solved angles.

The reference has no learned part, so there is no reference code to follow
beyond that warm start; this module is pinned by finite differences and
closed forms in tests/test_mlp.py.  With bf16=True every GEMM operand is
rounded to bfloat16 (round to nearest even) exactly where the GPU kernels
round it, and products are summed in float64.
"""
import numpy as np


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def unpack(p, shapes):
    """Flat parameter vector -> [(W, b), ...] (W row-major (out, in))."""
    out, at = [], 0
    for (ws, bs) in shapes:
        nw = ws[0] * ws[1]
        W = p[at:at + nw].reshape(ws)
        at += nw
        b = p[at:at + bs[0]]
        at += bs[0]
        out.append((W, b))
    assert at == p.size
    return out


def shapes_of(in_dim, hidden, depth, out_dim):
    dims = [in_dim] + [hidden] * depth + [out_dim]
    return [((dims[l + 1], dims[l]), (dims[l + 1],)) for l in range(len(dims) - 1)]


def _r(x, on):
    return bf16(x).astype(np.float64) if on else np.asarray(x, np.float64)


def forward(p, shapes, x, bf=True):
    """Activations [x, y_0, ..., y_L]; hidden layers tanh, last linear."""
    layers = unpack(np.asarray(p, np.float64 if not bf else np.float32), shapes)
    acts = [np.asarray(x, np.float32 if bf else np.float64)]
    for l, (W, b) in enumerate(layers):
        z = _r(acts[-1], bf) @ _r(W, bf).T + np.asarray(b, np.float64)
        y = np.tanh(z) if l < len(layers) - 1 else z
        acts.append(y.astype(np.float32) if bf else y)
    return acts


def backward(p, shapes, x, dy, bf=True):
    """Gradients of sum(dy * MLP(x)): flat parameter gradient and dx."""
    layers = unpack(np.asarray(p, np.float32 if bf else np.float64), shapes)
    acts = forward(p, shapes, x, bf)
    grads = []
    d = np.asarray(dy, np.float64)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        if l < len(layers) - 1:
            y = np.asarray(acts[l + 1], np.float64)
            d = d * (1.0 - y * y)
            if bf:
                d = d.astype(np.float32)
        dW = _r(d, bf).T @ _r(acts[l], bf)
        db = np.asarray(d, np.float64).sum(axis=0)
        grads.append((dW, db))
        d = _r(d, bf) @ _r(W, bf)
    flat = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in reversed(grads)])
    return flat, d


def guess_head(o, xt1, xt2):
    """guess = (1 + o0 + j o1) xt1 + (o2 + j o3) (xt1 - xt2) per row."""
    o = np.asarray(o, np.float64)
    a = (1.0 + o[:, 0]) + 1j * o[:, 1]
    c = o[:, 2] + 1j * o[:, 3]
    return a * xt1 + c * (xt1 - xt2)


def guess_loss(o, xt1, xt2, xtrue, w):
    """Mean over rows of w |guess - xtrue|^2, and its gradient wrt o."""
    g = guess_head(o, xt1, xt2)
    e = g - xtrue
    rows = o.shape[0]
    loss = float(np.sum(w * np.abs(e) ** 2) / rows)
    s = 2.0 * w / rows
    df = xt1 - xt2
    do = np.stack([s * np.real(np.conj(e) * xt1), s * np.real(np.conj(e) * 1j * xt1),
                   s * np.real(np.conj(e) * df), s * np.real(np.conj(e) * 1j * df)], axis=1)
    return loss, do


def phase_corrected(x, e_centro, k0, d, d_prev):
    """x * exp(-j k (d - d_prev) . e_centro) per unknown (main.f90:1017-1022)."""
    ph = np.exp(-1j * k0 * (np.asarray(e_centro).reshape(-1, 3) @ (np.asarray(d) - np.asarray(d_prev))))
    return np.asarray(x) * ph
