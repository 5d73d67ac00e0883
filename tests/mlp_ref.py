"""CPU FP32 reference of the subtree-size MLP (test helper, numpy).

Parameter layout: W1[16][64], b1[64], W2[64][64], b2[64], w3[64], b3.
`match_device=True` rounds the values the device feeds to FP16 MFMA inputs
(weights W1/W2, layer-1 activations, layer-2 output gradients, layer-1
gradients) the same way; otherwise everything stays FP32."""
import numpy as np

IN, HID = 16, 64


def unpack(p):
    p = np.asarray(p)
    o = 0
    W1 = p[o:o + IN * HID].reshape(IN, HID); o += IN * HID
    b1 = p[o:o + HID]; o += HID
    W2 = p[o:o + HID * HID].reshape(HID, HID); o += HID * HID
    b2 = p[o:o + HID]; o += HID
    w3 = p[o:o + HID]; o += HID
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def h16(a, on):
    """FP16 rounding where the device rounds; otherwise the input precision."""
    return a.astype(np.float16).astype(np.float32) if on else a


def features(n, prefixes):
    """Feature rows (float32) for prefixes (count x depth column bytes)."""
    prefixes = np.asarray(prefixes)
    out = np.zeros((len(prefixes), IN), np.float32)
    mask = (1 << n) - 1
    for i, pre in enumerate(prefixes):
        cols = ld = rd = 0
        for c in pre:
            b = 1 << int(c)
            cols |= b
            ld = ((ld | b) << 1) & mask
            rd = (rd | b) >> 1
        row = len(pre)
        f = [row, bin(~cols & mask).count("1"), bin(ld & mask).count("1"), bin(rd & mask).count("1")]
        for j in range(8):
            if row + j < n:
                att = cols | ((ld << j) & mask) | (rd >> j)
                f.append(bin(~att & mask).count("1"))
            else:
                f.append(0)
        out[i, :12] = np.array(f, np.float32) / 16.0
    return out


def forward(p, x, match_device=True):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    x = h16(x, match_device)
    h1 = np.maximum(x @ h16(W1, match_device) + b1, 0)
    h1h = h16(h1, match_device)
    h2 = np.maximum(h1h @ h16(W2, match_device) + b2, 0)
    y = h2 @ w3 + b3
    return y, (x, h1, h1h, h2)


def gradient(p, x, t, match_device=True):
    W1, b1, W2, b2, w3, b3 = unpack(p)
    y, (x, h1, h1h, h2) = forward(p, x, match_device)
    B = len(t)
    err = y - t
    dy = 2 * err / B
    gw3 = h2.T @ dy
    gb3 = dy.sum()
    dh2 = (dy[:, None] * w3[None, :]) * (h2 > 0)
    gb2 = dh2.sum(0)
    dh2h = h16(dh2, match_device)
    gW2 = h1h.T @ dh2h
    dh1 = (dh2h @ h16(W2, match_device).T) * (h1 > 0)
    gb1 = dh1.sum(0)
    gW1 = x.T @ h16(dh1, match_device)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]])
    return g, float((err ** 2).mean())
