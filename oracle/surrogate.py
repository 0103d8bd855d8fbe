"""CPU oracle of the learned surrogate network -- TEST INFRASTRUCTURE ONLY.

The surrogate has no counterpart in pylbm.  This is the plain numpy statement
of the network the HIP kernels (pylbm_amd/csrc/surrogate.hip) implement, used
to check their outputs and gradients:

  x (N, H, W, 4) -> conv3x3(4 -> 32) -> SiLU -> conv3x3(32 -> 32) -> SiLU
    -> conv3x3(32 -> 32) -> SiLU -> conv3x3(32 -> 3) -> increment (N, H, W, 3)

with periodic padding, weights W[tap][ci][co] (tap = 3*(dy+1) + (dx+1)), biases
b[co], loss = mean over all elements of (increment - target)^2, and Adam.

Pinned by: central finite differences of the loss with respect to inputs,
weights and biases (tests/test_oracle_surrogate.py), SiLU' against a finite
difference, and a direct loop convolution on a tiny image.
"""
import numpy as np

TAPS = [(dy, dx) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
CHANNELS = [4, 32, 32, 32, 3]


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def conv3x3(x, w, b):
    """x (N, H, W, Ci), w (9, Ci, Co), b (Co): periodic 3x3 convolution."""
    out = np.zeros(x.shape[:3] + (w.shape[2],), dtype=np.float64)
    for t, (dy, dx) in enumerate(TAPS):
        xs = np.roll(x, shift=(-dy, -dx), axis=(1, 2))       # xs[p] = x[p + (dy, dx)]
        out += xs @ w[t]
    return out + b


def conv3x3_input_grad(g, w):
    """d/dx of sum(g * conv3x3(x, w, b))."""
    out = np.zeros(g.shape[:3] + (w.shape[1],), dtype=np.float64)
    for t, (dy, dx) in enumerate(TAPS):
        out += np.roll(g @ w[t].T, shift=(dy, dx), axis=(1, 2))
    return out


def conv3x3_weight_grad(x, g):
    """d/dw[t] of sum(g * conv3x3(x, w, b)) -> (9, Ci, Co); and d/db."""
    dw = np.zeros((9, x.shape[3], g.shape[3]))
    for t, (dy, dx) in enumerate(TAPS):
        xs = np.roll(x, shift=(-dy, -dx), axis=(1, 2))
        dw[t] = xs.reshape(-1, x.shape[3]).T @ g.reshape(-1, g.shape[3])
    return dw, g.reshape(-1, g.shape[3]).sum(axis=0)


def init_params(seed=0, scale_last=0.1):
    rng = np.random.default_rng(seed)
    params = []
    for li in range(4):
        ci, co = CHANNELS[li], CHANNELS[li + 1]
        std = np.sqrt(2.0 / (9 * ci))
        w = rng.normal(0.0, std, (9, ci, co))
        if li == 3:
            w *= scale_last
        params.append([w, np.zeros(co)])
    return params


def forward(params, x):
    """Returns (increment, cache)."""
    acts, pre = [x], []
    a = x
    for li, (w, b) in enumerate(params):
        z = conv3x3(a, w, b)
        pre.append(z)
        a = silu(z) if li < 3 else z
        acts.append(a)
    return a, (acts, pre)


def loss_and_grads(params, x, target):
    y, (acts, pre) = forward(params, x)
    diff = y - target
    loss = float(np.mean(diff ** 2))
    g = 2.0 * diff / diff.size
    grads = [None] * 4
    for li in range(3, -1, -1):
        if li < 3:
            g = g * silu_grad(pre[li])
        dw, db = conv3x3_weight_grad(acts[li], g)
        grads[li] = [dw, db]
        if li > 0:
            g = conv3x3_input_grad(g, params[li][0])
    return loss, grads


class Adam:
    def __init__(self, params, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.lr, self.b1, self.b2, self.eps = lr, b1, b2, eps
        self.m = [[np.zeros_like(p) for p in layer] for layer in params]
        self.v = [[np.zeros_like(p) for p in layer] for layer in params]
        self.t = 0

    def step(self, params, grads):
        self.t += 1
        c1 = 1.0 - self.b1 ** self.t
        c2 = 1.0 - self.b2 ** self.t
        for L, G, M, V in zip(params, grads, self.m, self.v):
            for i in range(len(L)):
                M[i][...] = self.b1 * M[i] + (1 - self.b1) * G[i]
                V[i][...] = self.b2 * V[i] + (1 - self.b2) * G[i] ** 2
                L[i][...] = L[i] - self.lr * (M[i] / c1) / (np.sqrt(V[i] / c2) + self.eps)
