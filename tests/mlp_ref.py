"""CPU restatement of the proposal network (ever_amd/csrc/ever_mlp.hip) for the tests: encoding, forward, pinball
loss and backward, with bf16 rounding at the points the GPU rounds (rnd=True) or none (rnd=False, FP64 for
finite differences)."""
import numpy as np

IN, HID = 64, 128


def layout():
    off, o = {}, 0
    for name, size in (("w1", IN * HID), ("b1", HID), ("w2", HID * HID), ("b2", HID), ("w3", HID * HID),
                       ("b3", HID), ("w4", HID), ("b4", 1)):
        off[name] = (o, size)
        o += size
    return off, o


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def encode(rayo, rayd, tmin=0.0):
    """(R,64) features: o/4, p/4 + 5 octaves of sin/cos(pi 2^j p/4), d + 4 octaves of sin/cos(pi 2^j d), 0."""
    rd = np.asarray(rayd, np.float64)
    d = rd / np.linalg.norm(rd, axis=1, keepdims=True)
    o = np.asarray(rayo, np.float64) + tmin * rd
    p = o - (o * d).sum(1, keepdims=True) * d
    cols = [0.25 * o, 0.25 * p]
    for j in range(5):
        for a in range(3):
            v = np.pi * (2.0 ** j) * 0.25 * p[:, a]
            cols.append(np.stack([np.sin(v), np.cos(v)], 1))
    cols.append(d)
    for j in range(4):
        for a in range(3):
            v = np.pi * (2.0 ** j) * d[:, a]
            cols.append(np.stack([np.sin(v), np.cos(v)], 1))
    x = np.concatenate(cols, 1)
    assert x.shape[1] == 63
    return np.concatenate([x, np.zeros((x.shape[0], 1))], 1)


def unpack(params):
    off, _ = layout()
    g = lambda n: params[off[n][0]:off[n][0] + off[n][1]]
    return (g("w1").reshape(IN, HID), g("b1"), g("w2").reshape(HID, HID), g("b2"), g("w3").reshape(HID, HID),
            g("b3"), g("w4"), g("b4")[0])


def forward(params, x, rnd=True):
    """Returns y and the activations (x, h1, h2, h3) as used by the backward."""
    W1, b1, W2, b2, W3, b3, w4, b4 = unpack(np.asarray(params, np.float64))
    R = (lambda a: bf16(a).astype(np.float64)) if rnd else (lambda a: a)
    xs = R(x)
    hs = [xs]
    h = xs
    for W, b in ((W1, b1), (W2, b2), (W3, b3)):
        h = R(np.maximum(h @ R(W) + b, 0.0))
        hs.append(h)
    return h @ w4 + b4, hs


def loss_and_grad(params, x, target, tau, rnd=True):
    """Pinball loss of y against log(target) (rows with target <= 0 ignored, normalised by the row count) and its
    gradient w.r.t. the parameter block."""
    params = np.asarray(params, np.float64)
    W1, b1, W2, b2, W3, b3, w4, b4 = unpack(params)
    R = (lambda a: bf16(a).astype(np.float64)) if rnd else (lambda a: a)
    y, (xs, h1, h2, h3) = forward(params, x, rnd)
    n = x.shape[0]
    valid = target > 0
    u = np.where(valid, np.log(np.where(valid, target, 1.0)) - y, 0.0)
    ind = (u < 0).astype(np.float64)
    loss = float((u * (tau - ind) * valid).sum() / n)
    dy = np.where(valid, -(tau - ind) / n, 0.0)
    off, count = layout()
    grad = np.zeros(count)
    put = lambda name, v: grad.__setitem__(slice(off[name][0], off[name][0] + off[name][1]), np.ravel(v))
    put("w4", h3.T @ dy)
    put("b4", [dy.sum()])
    dz = R(np.outer(dy, w4) * (h3 > 0))
    for l, (W, hin, hname) in enumerate(((W3, h2, "3"), (W2, h1, "2"), (W1, xs, "1"))):
        put("w" + hname, hin.T @ dz)
        put("b" + hname, dz.sum(0))
        if hname == "1":
            break
        dh = dz @ R(W).T
        dz = R(dh * (hin > 0))
    return loss, grad
