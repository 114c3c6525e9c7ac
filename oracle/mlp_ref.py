"""FP64 numpy reference of the warm-start MLP (16 -> 64 -> 64 -> 1, GELU).

TEST INFRASTRUCTURE ONLY (the checker of the gfx950 MLP kernels).

Parameter layout (FP32 vector, see include/bubbles_amd.h):
  W1[64][16] b1[64] W2[64][64] b2[64] W3[64] b3[1] | mean[16] inv_std[16] out_scale[1]
Forward:  xn = clip((x - mean) * inv_std, -8, 8);  z1 = W1 xn + b1;  h1 = gelu(z1)
          z2 = W2 h1 + b2;  h2 = gelu(z2);  y = W3 . h2 + b3
Loss:     0.5 / B * sum w (y - t / out_scale)^2, w = 2(1 - tau) above the
          target and 2 tau below (expectile tau; 0.5 = mean squared error)
With ``emulate=True`` the values the device feeds to BF16 MFMA operands are
rounded to bfloat16 (round to nearest even) at the same points: xn, W1, h1,
W2 in the forward pass; dz2 (for dH1 and dW2), h1, dz1 and xn in the backward
pass.  Everything else stays FP64, so remaining differences are FP32
accumulation order and the device's exp-based tanh.
"""
from __future__ import annotations

import numpy as np

H, F = 64, 16
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * F
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
N_TRAIN = OFF_B3 + 1
OFF_MEAN = N_TRAIN
OFF_ISTD = OFF_MEAN + F
OFF_OSCALE = OFF_ISTD + F
OFF_ALPHA = OFF_OSCALE + 1
OFF_PCAP = OFF_ALPHA + 1
OFF_RESID = OFF_PCAP + 1
N_PARAMS = OFF_RESID + 1
F_PPREV = 10  # feature index of the previous substep's pressure

K0 = 0.7978845608028654
K1 = 0.044715


def bf16(a) -> np.ndarray:
    """Round to bfloat16 (nearest even) and return as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(K0 * (z + K1 * z ** 3)))


def dgelu(z):
    t = np.tanh(K0 * (z + K1 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * K0 * (1.0 + 3.0 * K1 * z * z)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return {
        "W1": p[OFF_W1:OFF_B1].reshape(H, F), "b1": p[OFF_B1:OFF_W2],
        "W2": p[OFF_W2:OFF_B2].reshape(H, H), "b2": p[OFF_B2:OFF_W3],
        "W3": p[OFF_W3:OFF_B3], "b3": p[OFF_B3],
        "mean": p[OFF_MEAN:OFF_ISTD], "istd": p[OFF_ISTD:OFF_OSCALE], "oscale": p[OFF_OSCALE],
        "alpha": p[OFF_ALPHA], "pcap": p[OFF_PCAP], "resid": p[OFF_RESID],
    }


FEAT_CLIP = 8.0  # normalised features are clipped to +-8 standard deviations


def _normalise(P, x, emulate):
    if emulate:
        x32 = np.asarray(x, dtype=np.float32)
        xn = (x32 - P["mean"].astype(np.float32)) * P["istd"].astype(np.float32)
        return bf16(np.clip(xn, np.float32(-FEAT_CLIP), np.float32(FEAT_CLIP)))
    return np.clip((np.asarray(x, dtype=np.float64) - P["mean"]) * P["istd"], -FEAT_CLIP, FEAT_CLIP)


def forward(params, x, emulate=True):
    """Returns (y, cache); y is the normalised output (before out_scale)."""
    P = unpack(params)
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    xn = _normalise(P, x, emulate)                      # (n, 16)
    z1 = xn @ r(P["W1"]).T + P["b1"]                    # (n, 64)
    h1 = gelu(z1)
    z2 = r(h1) @ r(P["W2"]).T + P["b2"]
    h2 = gelu(z2)
    y = h2 @ P["W3"] + P["b3"]
    return y, {"xn": xn, "z1": z1, "h1": h1, "z2": z2, "h2": h2, "P": P}


def predict(params, x, emulate=True):
    """Prediction in target units: y * out_scale + resid * x[PPREV]."""
    y, c = forward(params, x, emulate)
    p = np.asarray(params, dtype=np.float64)
    return y * c["P"]["oscale"] + p[OFF_RESID] * np.asarray(x, dtype=np.float64)[:, F_PPREV]


def warm_pressure(params, x, emulate=True):
    """What the solver starts from, in target units: alpha * clamp(pred, 0, pcap)."""
    p = np.asarray(params, dtype=np.float64)
    return p[OFF_ALPHA] * np.clip(predict(params, x, emulate), 0.0, p[OFF_PCAP])


def loss_and_grad(params, x, t, emulate=True, expectile=0.5):
    """Expectile loss 0.5 mean(w (y - t)^2), w = 2(1 - tau) where y > t and
    2 tau where y < t (tau = 0.5: mean squared error), and its gradient."""
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    y, c = forward(params, x, emulate)
    P = c["P"]
    B = y.shape[0]
    tn = np.asarray(t, dtype=np.float64) / P["oscale"]
    diff = y - tn
    w = np.where(diff > 0, 2.0 * (1.0 - expectile), 2.0 * expectile)
    loss = 0.5 * np.sum(w * diff * diff) / B
    dy = w * diff / B                                   # (n,)
    g = np.zeros(N_PARAMS)
    g[OFF_W3:OFF_B3] = c["h2"].T @ dy
    g[OFF_B3] = dy.sum()
    dz2 = np.outer(dy, P["W3"]) * dgelu(c["z2"])        # (n, 64)
    g[OFF_B2:OFF_W3] = dz2.sum(axis=0)
    g[OFF_W2:OFF_B2] = (r(dz2).T @ r(c["h1"])).ravel()
    dh1 = r(dz2) @ r(P["W2"])                           # (n, 64)
    dz1 = dh1 * dgelu(c["z1"])
    g[OFF_B1:OFF_W2] = dz1.sum(axis=0)
    g[OFF_W1:OFF_B1] = (r(dz1).T @ c["xn"]).ravel()
    return loss, g


def adam_step(params, grad, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update of the trainable part (t = step count after increment)."""
    p = np.array(params, dtype=np.float64)
    gg = np.asarray(grad, dtype=np.float64)[:N_TRAIN]
    m = b1 * m + (1 - b1) * gg
    v = b2 * v + (1 - b2) * gg * gg
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    p[:N_TRAIN] -= lr * mh / (np.sqrt(vh) + eps)
    return p, m, v
