"""Oracle (test infrastructure only): float64 forward/backward of the
domain-hit predictor MLP, with the inputs of every bf16 matrix product
rounded to bf16 exactly where the GPU kernels round them, so that the GPU
gradients can be compared at fp32-accumulation tolerance.

Parameter layout (flat): W1[128][22], b1[128], W2[128][128], b2[128],
w3[128], b3[1], emb[nd][16].  Sample layout: origin(3), direction(3),
domain, label.
"""
import numpy as np

HID, NIN, NEMB, NRAY = 128, 22, 16, 6


def bf16(x):
    """Round float32 values to bfloat16 (round to nearest even), as float64."""
    a = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    a = (a + 0x7FFF + ((a >> 16) & 1)) & 0xFFFF0000
    return a.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p, nd):
    o = 0
    W1 = p[o:o + HID * NIN].reshape(HID, NIN); o += HID * NIN
    b1 = p[o:o + HID]; o += HID
    W2 = p[o:o + HID * HID].reshape(HID, HID); o += HID * HID
    b2 = p[o:o + HID]; o += HID
    w3 = p[o:o + HID]; o += HID
    b3 = p[o]; o += 1
    emb = p[o:o + nd * NEMB].reshape(nd, NEMB)
    return W1, b1, W2, b2, w3, b3, emb


def features(samples, emb, center, inv_half):
    s = np.asarray(samples, np.float32)
    o = (s[:, 0:3] - center.astype(np.float32)) * inv_half.astype(np.float32)
    d = s[:, 3:6]
    e = emb[s[:, 6].astype(np.int64)].astype(np.float32)
    return np.concatenate([o.astype(np.float32), d, e], axis=1)


def forward_backward(params, samples, nd, center, inv_half):
    """Returns (summed BCE, flat gradient of the mean BCE, logits)."""
    p = np.asarray(params, np.float32)
    W1, b1, W2, b2, w3, b3, emb = unpack(p, nd)
    n = len(samples)
    y = np.asarray(samples, np.float64)[:, 7]
    X = bf16(features(samples, emb, center, inv_half))
    W1b, W2b = bf16(W1), bf16(W2)
    h1 = np.maximum(X @ W1b.T + b1, 0.0)
    h1b = bf16(h1)
    h2 = np.maximum(h1b @ W2b.T + b2, 0.0)
    z = h2 @ w3.astype(np.float64) + float(b3)
    loss = float(np.sum(np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z)))))
    dz = (1.0 / (1.0 + np.exp(-z)) - y) / n
    g_w3 = h2.T @ dz
    g_b3 = dz.sum()
    dh2 = dz[:, None] * w3[None, :] * (h2 > 0)
    g_b2 = dh2.sum(0)
    dh2b = bf16(dh2)
    g_W2 = dh2b.T @ h1b
    dh1 = (dh2b @ W2b) * (h1 > 0)
    g_b1 = dh1.sum(0)
    dh1b = bf16(dh1)
    g_W1 = dh1b.T @ X
    demb = dh1b @ bf16(W1[:, NRAY:NRAY + NEMB])
    g_emb = np.zeros((nd, NEMB))
    np.add.at(g_emb, np.asarray(samples)[:, 6].astype(np.int64), demb)
    g = np.concatenate([g_W1.ravel(), g_b1, g_W2.ravel(), g_b2, g_w3, [g_b3], g_emb.ravel()])
    return loss, g, z


def numerical_grad(params, samples, nd, center, inv_half, idx, eps=1e-3):
    """Central differences of the mean BCE without bf16 rounding (float64),
    for the parameters in idx: pins the analytic backward of the oracle."""
    def loss_of(pv):
        W1, b1, W2, b2, w3, b3, emb = unpack(pv, nd)
        s = np.asarray(samples, np.float64)
        o = (s[:, 0:3] - center) * inv_half
        X = np.concatenate([o, s[:, 3:6], emb[s[:, 6].astype(np.int64)]], axis=1)
        h1 = np.maximum(X @ W1.T + b1, 0.0)
        h2 = np.maximum(h1 @ W2.T + b2, 0.0)
        z = h2 @ w3 + b3
        y = s[:, 7]
        return np.mean(np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z))))
    p = np.asarray(params, np.float64).copy()
    out = []
    for i in idx:
        p[i] += eps
        lp = loss_of(p)
        p[i] -= 2 * eps
        lm = loss_of(p)
        p[i] += eps
        out.append((lp - lm) / (2 * eps))
    return np.array(out)


def analytic_grad_f64(params, samples, nd, center, inv_half):
    """Backward without bf16 rounding (float64), same formulas as above."""
    p = np.asarray(params, np.float64)
    W1, b1, W2, b2, w3, b3, emb = unpack(p, nd)
    s = np.asarray(samples, np.float64)
    n = len(s)
    X = np.concatenate([(s[:, 0:3] - center) * inv_half, s[:, 3:6], emb[s[:, 6].astype(np.int64)]], axis=1)
    h1 = np.maximum(X @ W1.T + b1, 0.0)
    h2 = np.maximum(h1 @ W2.T + b2, 0.0)
    z = h2 @ w3 + b3
    dz = (1.0 / (1.0 + np.exp(-z)) - s[:, 7]) / n
    dh2 = dz[:, None] * w3[None, :] * (h2 > 0)
    dh1 = (dh2 @ W2) * (h1 > 0)
    g_emb = np.zeros((nd, NEMB))
    np.add.at(g_emb, s[:, 6].astype(np.int64), dh1 @ W1[:, NRAY:NRAY + NEMB])
    return np.concatenate([(dh1.T @ X).ravel(), dh1.sum(0), (dh2.T @ h1).ravel(), dh2.sum(0),
                           h2.T @ dz, [dz.sum()], g_emb.ravel()])
