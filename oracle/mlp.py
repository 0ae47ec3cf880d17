"""CPU restatement of the warm-start network (include/espice_mlp.h) for the
tests: test infrastructure only, never a product path.

The network: standardized inputs -> `depth` tanh layers of `hidden` units ->
linear outputs, loss = mean over the batch and outputs of the squared error
of the standardized targets. With `bf16=True` every operand the GPU's BF16
MFMA kernels read is rounded to bfloat16 (round to nearest even) at the same
places: the standardized inputs, the weights, the hidden activations and the
back-propagated output gradients; sums are FP64 here and FP32 on the GPU, so
the two agree to a tolerance, not bitwise. With `bf16=False` the whole
restatement is FP64, and its gradient is pinned by central finite
differences (tests/test_mlp.py).

Parameters are packed layer by layer: W (fan_out x fan_in, row-major), then b.

The reference has no such network; parity of what it produces is judged by
Newton convergence of the lanes it starts, not by this file.
"""
import numpy as np


def bf16(a):
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


def shapes(n_in, hidden, depth, n_out):
    """(fan_out, fan_in) of every layer."""
    dims = [n_in] + [hidden] * depth + [n_out]
    return [(dims[i + 1], dims[i]) for i in range(len(dims) - 1)]


def unpack(params, n_in, hidden, depth, n_out):
    out, at = [], 0
    for fo, fi in shapes(n_in, hidden, depth, n_out):
        W = np.asarray(params[at:at + fo * fi], dtype=np.float64).reshape(fo, fi)
        at += fo * fi
        b = np.asarray(params[at:at + fo], dtype=np.float64)
        at += fo
        out.append((W, b))
    assert at == len(params)
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.reshape(-1), b]) for W, b in layers])


def standardize(v, mean, sd):
    v = np.asarray(v, dtype=np.float64)
    sd = np.asarray(sd, dtype=np.float64)
    safe = np.where(sd != 0, sd, 1.0)
    return np.where(sd != 0, (v - mean) / safe, 0.0)


def _round(a, on):
    return bf16(np.asarray(a, dtype=np.float32)).astype(np.float64) if on else a


def _f32(a, on):
    return np.asarray(a, dtype=np.float32).astype(np.float64) if on else a


def forward(layers, x, in_mean, in_sd, bf16_ops=True):
    """Standardized outputs and the cache for backward."""
    h = _round(_f32(standardize(x, in_mean, in_sd), bf16_ops), bf16_ops)
    hs, zs = [h], []
    for k, (W, b) in enumerate(layers):
        z = _f32(h @ _round(_f32(W, bf16_ops), bf16_ops).T + _f32(b, bf16_ops), bf16_ops)
        zs.append(z)
        if k + 1 < len(layers):
            t = _f32(np.tanh(z), bf16_ops)
            h = _round(t, bf16_ops)
            hs.append(h)
    return zs[-1], (hs, zs)


def loss_and_grads(layers, x, y, in_mean, in_sd, out_mean, out_sd, bf16_ops=True):
    """(loss, packed gradient) of the mean squared standardized error."""
    yhat, (hs, zs) = forward(layers, x, in_mean, in_sd, bf16_ops)
    t = _f32(standardize(y, out_mean, out_sd), bf16_ops)
    B, n_out = yhat.shape
    d = yhat - t
    loss = float(np.mean(d * d))
    dz = _f32(2.0 * d / (B * n_out), bf16_ops)
    grads = [None] * len(layers)
    for k in range(len(layers) - 1, -1, -1):
        W, _ = layers[k]
        dzb = _round(dz, bf16_ops)
        dW = _f32(dzb.T @ hs[k], bf16_ops)
        db = _f32(dz.sum(0), bf16_ops)
        grads[k] = (dW, db)
        if k == 0:
            break
        dh = _f32(dzb @ _round(_f32(W, bf16_ops), bf16_ops), bf16_ops)
        tz = np.tanh(zs[k - 1])
        dz = _f32(dh * (1.0 - tz * tz), bf16_ops)
    return loss, pack(grads)


def predict(layers, x, in_mean, in_sd, out_mean, out_sd, bf16_ops=True):
    """De-standardized outputs (a zero std gives exactly the mean)."""
    z, _ = forward(layers, x, in_mean, in_sd, bf16_ops)
    out_sd = np.asarray(out_sd, dtype=np.float64)
    return np.where(out_sd != 0, out_mean + out_sd * z, out_mean)
