"""FP64 numpy reference of the HypoCS-surrogate MLP (test infrastructure).

14 -> 128 -> 128 -> 128 -> 7, GELU in tanh form on the hidden layers, loss
mean((y_n - t_n)^2) over all outputs of all rows in normalised units.  The
parameter layout matches the library: W1[14][128] b1 W2[128][128] b2
W3[128][128] b3 W4[128][7] b4 (W[k][n] maps input k to output n)."""
import numpy as np

NIN, NH, NOUT = 14, 128, 7
SIZES = [(NIN, NH), (NH,), (NH, NH), (NH,), (NH, NH), (NH,), (NH, NOUT), (NOUT,)]
NPARAM = sum(int(np.prod(s)) for s in SIZES)
C = np.sqrt(2.0 / np.pi)


def unpack(p):
    out, o = [], 0
    for s in SIZES:
        n = int(np.prod(s))
        out.append(np.asarray(p[o:o + n], np.float64).reshape(s))
        o += n
    return out


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(C * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    t = np.tanh(C * (z + 0.044715 * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * C * (1.0 + 3 * 0.044715 * z * z)


def normalise(a, mean, sd):
    return (np.asarray(a, np.float64) - mean) / sd


def forward_n(p, xn):
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    z1 = xn @ W1 + b1
    a1 = gelu(z1)
    z2 = a1 @ W2 + b2
    a2 = gelu(z2)
    z3 = a2 @ W3 + b3
    a3 = gelu(z3)
    return a3 @ W4 + b4, (xn, z1, a1, z2, a2, z3, a3)


def forward(p, norm, x):
    """Raw inputs -> raw outputs."""
    xn = normalise(x, norm[:NIN], norm[NIN:2 * NIN])
    yn, _ = forward_n(p, xn)
    return yn * norm[2 * NIN + NOUT:] + norm[2 * NIN:2 * NIN + NOUT]


def loss_grad(p, norm, x, t):
    xn = normalise(x, norm[:NIN], norm[NIN:2 * NIN])
    tn = normalise(t, norm[2 * NIN:2 * NIN + NOUT], norm[2 * NIN + NOUT:])
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(p)
    y, (xn, z1, a1, z2, a2, z3, a3) = forward_n(p, xn)
    m = y.size
    loss = float(np.sum((y - tn) ** 2) / m)
    dy = 2.0 * (y - tn) / m
    gW4, gb4 = a3.T @ dy, dy.sum(0)
    d3 = (dy @ W4.T) * gelu_grad(z3)
    gW3, gb3 = a2.T @ d3, d3.sum(0)
    d2 = (d3 @ W3.T) * gelu_grad(z2)
    gW2, gb2 = a1.T @ d2, d2.sum(0)
    d1 = (d2 @ W2.T) * gelu_grad(z1)
    gW1, gb1 = xn.T @ d1, d1.sum(0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3, gW4, gb4)])
    return loss, g


def init_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    out = []
    for s in SIZES:
        if len(s) == 2:
            a = scale * np.sqrt(6.0 / (s[0] + s[1]))
            out.append(rng.uniform(-a, a, s))
        else:
            out.append(rng.uniform(-0.1, 0.1, s))
    return np.concatenate([o.ravel() for o in out])
