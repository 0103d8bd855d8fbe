"""CPU oracle of the learned surrogate network -- TEST INFRASTRUCTURE ONLY.

The surrogate has no counterpart in pylbm.  This is the plain numpy statement
of the network the HIP kernels (pylbm_amd/csrc/surrogate.hip) implement, used
to check their outputs and gradients.

Fine state (N, Hf, Wf, 4) = (rho, ux, uy, 0) is normalised per channel and
unshuffled into 4x4 patches: coarse pixel (Y, X) channel (4 py + px) * 4 + ch
holds fine pixel (4 Y + py, 4 X + px) channel ch.  Then

  x (N, H, W, 64) -> conv3x3(64 -> C) -> SiLU -> conv3x3(C -> C) -> SiLU
    -> conv3x3(C -> C) -> SiLU -> conv3x3(C -> 64) -> y (N, H, W, 64)

with periodic padding, weights W[tap][ci][co] (tap = 3*(dy+1) + (dx+1)),
biases b[co].  The loss is the mean of (y - target)^2 over the channels with
ch = c % 4 < 3 (the slot ch = 3 carries nothing); y is the increment of the
normalised state in units of the increment scale.

Pinned by: central finite differences of the loss with respect to inputs,
weights and biases, SiLU' against a finite difference, a direct loop
convolution on a tiny image, and unshuffle/shuffle round trips
(tests/test_oracle_surrogate.py); Adam against torch.optim.Adam on the CPU and
its first step -lr sign(g) (tests/test_oracle_pins.py).
"""
import numpy as np

TAPS = [(dy, dx) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
PATCH = 4
IN_CH = PATCH * PATCH * 4


def channels(hidden=64):
    return [IN_CH, hidden, hidden, hidden, IN_CH]


def loss_mask():
    """(64,) 1.0 on the channels that carry (rho, ux, uy), 0.0 on the empty slot."""
    return (np.arange(IN_CH) % 4 < 3).astype(np.float64)


def unshuffle(fine):
    """(N, Hf, Wf, 4) -> (N, Hf/4, Wf/4, 64)."""
    n, hf, wf, c = fine.shape
    t = fine.reshape(n, hf // PATCH, PATCH, wf // PATCH, PATCH, c)
    return t.transpose(0, 1, 3, 2, 4, 5).reshape(n, hf // PATCH, wf // PATCH, PATCH * PATCH * c)


def shuffle(coarse):
    """Inverse of :func:`unshuffle`."""
    n, h, w, cc = coarse.shape
    c = cc // (PATCH * PATCH)
    t = coarse.reshape(n, h, w, PATCH, PATCH, c)
    return t.transpose(0, 1, 3, 2, 4, 5).reshape(n, h * PATCH, w * PATCH, c)


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


def init_params(seed=0, hidden=64, scale_last=0.1):
    rng = np.random.default_rng(seed)
    ch = channels(hidden)
    params = []
    for li in range(4):
        ci, co = ch[li], ch[li + 1]
        std = np.sqrt(2.0 / (9 * ci))
        w = rng.normal(0.0, std, (9, ci, co))
        if li == 3:
            w *= scale_last
        params.append([w, np.zeros(co)])
    return params


def forward(params, x):
    """Returns (y, cache)."""
    acts, pre = [x], []
    a = x
    for li, (w, b) in enumerate(params):
        z = conv3x3(a, w, b)
        pre.append(z)
        a = silu(z) if li < 3 else z
        acts.append(a)
    return a, (acts, pre)


def loss_and_grads(params, x, target):
    """Masked mean-square loss over the fine (rho, ux, uy) values and its gradients."""
    y, (acts, pre) = forward(params, x)
    mask = loss_mask()
    diff = (y - target) * mask
    count = diff[..., 0].size * mask.sum()
    loss = float(np.sum(diff ** 2) / count)
    g = 2.0 * diff / count
    grads = [None] * 4
    for li in range(3, -1, -1):
        if li < 3:
            g = g * silu_grad(pre[li])
        dw, db = conv3x3_weight_grad(acts[li], g)
        grads[li] = [dw, db]
        if li > 0:
            g = conv3x3_input_grad(g, params[li][0])
    return loss, grads


def step_state(params, fine, shift, scale, dscale):
    """One surrogate step of a fine state (N, Hf, Wf, 4): state + increment * dscale."""
    sh = np.array(list(shift) + [0.0])
    sc = np.array(list(scale) + [1.0])
    ds = np.array(list(dscale) + [0.0])
    x = (fine - sh) / sc
    x[..., 3] = 0.0
    y, _ = forward(params, unshuffle(x))
    return fine + shuffle(y) * ds


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
