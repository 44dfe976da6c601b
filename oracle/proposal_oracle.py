"""CPU oracle of the block proposal network (test infrastructure only).

The network has no counterpart in lacker/seticore; this restates the HIP
kernels' definition (seticore_amd/csrc/pn_kernels.hip) in numpy so the GPU
forward pass and gradient can be checked:

* `forward(z, params, bf16=True)` mirrors the GPU: BF16 weights for the
  convolutions, FP32 head weights, activations rounded to BF16
  (round-to-nearest-even) after each ReLU; sums accumulate in float64 here
  (FP32 on the GPU).
* `gradient(z, labels, params, pos_weight)` is the analytic gradient of the
  mean weighted binary cross-entropy of the block logits, through the max
  (argmax position only), the ReLUs (where the stored activation is > 0) and
  the convolutions, with the same BF16 weights the forward pass used.
* With bf16=False everything stays float64, which the tests use to pin the
  analytic gradient against central finite differences.

Pooling (`pool`) restates the GPU tile construction: z = 2 * (sum of 4
channels - median/4) / std, clamped to [-4, 32], rounded to BF16.
"""

import numpy as np

ROWS, POOLED, BLOCK = 16, 1024, 4096
C1, H1, W1 = 16, 16, 512
C2, H2, W2 = 32, 8, 256
C3, H3, W3 = 32, 4, 128

OFF_W1 = 0
OFF_B1 = OFF_W1 + C1 * 15
OFF_W2 = OFF_B1 + C1
OFF_B2 = OFF_W2 + C2 * 15 * C1
OFF_W3 = OFF_B2 + C2
OFF_B3 = OFF_W3 + C3 * 15 * C2
OFF_WH = OFF_B3 + C3
OFF_BH = OFF_WH + C3
NUM_PARAMS = OFF_BH + 1  # 23393


def bf16_round(x):
    """float32 -> nearest-even bfloat16, returned as float32 values."""
    f = np.ascontiguousarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unpack(params, bf16=True):
    p = np.asarray(params, np.float64 if not bf16 else np.float32)
    r = (lambda a: bf16_round(a).astype(np.float64)) if bf16 else (lambda a: a.astype(np.float64))
    w1 = r(p[OFF_W1:OFF_B1]).reshape(C1, 15, 1)
    w2 = r(p[OFF_W2:OFF_B2]).reshape(C2, 15, C1)
    w3 = r(p[OFF_W3:OFF_B3]).reshape(C3, 15, C2)
    b1 = p[OFF_B1:OFF_W2].astype(np.float64)
    b2 = p[OFF_B2:OFF_W3].astype(np.float64)
    b3 = p[OFF_B3:OFF_WH].astype(np.float64)
    wh = p[OFF_WH:OFF_BH].astype(np.float64)
    bh = float(p[OFF_BH])
    return w1, b1, w2, b2, w3, b3, wh, bh


def _conv(x, w, b, sh, sw, ho, wo):
    """x [B][H][W][CI]; w [CO][15][CI] (taps row-major over 3x5); pad (1, 2)."""
    B, H, W, CI = x.shape
    xp = np.zeros((B, H + 2, W + 4, CI), np.float64)
    xp[:, 1:H + 1, 2:W + 2, :] = x
    out = np.zeros((B, ho, wo, w.shape[0]), np.float64)
    for t in range(15):
        kh, kw = divmod(t, 5)
        sl = xp[:, kh:kh + sh * (ho - 1) + 1:sh, kw:kw + sw * (wo - 1) + 1:sw, :]
        out += sl @ w[:, t, :].T
    return out + b


def forward(z, params, bf16=True):
    """z [B][16][1024] -> (logits [B], argmax [B], (h1, h2, h3))."""
    w1, b1, w2, b2, w3, b3, wh, bh = unpack(params, bf16)
    rnd = (lambda a: bf16_round(a.astype(np.float32)).astype(np.float64)) if bf16 else (lambda a: a)
    x = np.asarray(z, np.float64)[..., None]
    h1 = rnd(np.maximum(_conv(x, w1, b1, 1, 2, H1, W1), 0.0))
    h2 = rnd(np.maximum(_conv(h1, w2, b2, 2, 2, H2, W2), 0.0))
    h3 = rnd(np.maximum(_conv(h2, w3, b3, 2, 2, H3, W3), 0.0))
    s = (h3 @ wh + bh).reshape(len(x), -1)  # [B][512], position = y*128 + x
    arg = np.argmax(s, axis=1)  # first maximum
    return s[np.arange(len(x)), arg], arg, (h1, h2, h3)


def loss_and_dz(logits, labels, pos_weight):
    z = np.asarray(logits, np.float64)
    y = np.asarray(labels, np.float64)
    w = np.where(y > 0.5, pos_weight, 1.0)
    loss = w * (np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z))))
    dz = w * (1.0 / (1.0 + np.exp(-z)) - y) / len(z)
    return loss.mean(), dz


def gradient(z, labels, params, pos_weight, bf16=True):
    """(loss, grad [NUM_PARAMS]) of the mean weighted BCE of the block logits."""
    w1, b1, w2, b2, w3, b3, wh, bh = unpack(params, bf16)
    logits, arg, (h1, h2, h3) = forward(z, params, bf16)
    loss, dz = loss_and_dz(logits, labels, pos_weight)
    B = len(logits)
    g = np.zeros(NUM_PARAMS, np.float64)
    x = np.asarray(z, np.float64)
    for b in range(B):
        y3, x3 = divmod(int(arg[b]), W3)
        d = dz[b]
        a3 = h3[b, y3, x3]
        g[OFF_WH:OFF_BH] += d * a3
        g[OFF_BH] += d
        dy3 = np.where(a3 > 0, d * wh, 0.0)
        g[OFF_B3:OFF_WH] += dy3
        # conv3 -> h2 positions of the 15 taps
        dh2 = {}
        gw3 = np.zeros((C3, 15, C2))
        for t in range(15):
            yy, xx = 2 * y3 + t // 5 - 1, 2 * x3 + t % 5 - 2
            if 0 <= yy < H2 and 0 <= xx < W2:
                gw3[:, t, :] = np.outer(dy3, h2[b, yy, xx])
                v = dy3 @ w3[:, t, :]
                dh2[(yy, xx)] = dh2.get((yy, xx), 0.0) + np.where(h2[b, yy, xx] > 0, v, 0.0)
        g[OFF_W3:OFF_B3] += gw3.ravel()
        # conv2
        dh1 = {}
        gw2 = np.zeros((C2, 15, C1))
        for (yy, xx), dy2 in dh2.items():
            g[OFF_B2:OFF_W3] += dy2
            for t in range(15):
                y1, x1 = 2 * yy + t // 5 - 1, 2 * xx + t % 5 - 2
                if 0 <= y1 < H1 and 0 <= x1 < W1:
                    gw2[:, t, :] += np.outer(dy2, h1[b, y1, x1])
                    dh1[(y1, x1)] = dh1.get((y1, x1), 0.0) + dy2 @ w2[:, t, :]
        g[OFF_W2:OFF_B2] += gw2.ravel()
        # conv1
        gw1 = np.zeros((C1, 15))
        for (y1, x1), v in dh1.items():
            dy1 = np.where(h1[b, y1, x1] > 0, v, 0.0)
            g[OFF_B1:OFF_W2] += dy1
            for t in range(15):
                zy, zx = y1 + t // 5 - 1, 2 * x1 + t % 5 - 2
                if 0 <= zy < ROWS and 0 <= zx < POOLED:
                    gw1[:, t] += dy1 * x[b, zy, zx]
        g[OFF_W1:OFF_B1] += gw1.ravel()
    return loss, g


def pool(data, median, std):
    """data [16][N] float32 -> BF16-valued float32 tiles [N/4096][16][1024]."""
    d = np.asarray(data, np.float32)
    T, N = d.shape
    q = d.reshape(T, N // 4, 4)
    s = (q[:, :, 0] + q[:, :, 1]) + (q[:, :, 2] + q[:, :, 3])
    v = np.float32(2.0) * (s - np.float32(0.25) * np.float32(median)) / np.float32(std)
    v = np.clip(v, np.float32(-4.0), np.float32(32.0)).astype(np.float32)
    v = bf16_round(v)
    return v.reshape(T, N // BLOCK, POOLED).transpose(1, 0, 2).copy()
