"""Learned surrogate of a periodic D2Q9 run (no counterpart in pylbm).

A 4-layer periodic 3x3 convolution network on 4x4 patches maps the normalised
moment fields (rho, ux, uy) at t to their increment over K lattice steps.  The
fine state is unshuffled into coarse pixels of 64 channels (16 fine pixels x 4
fields), so the convolutions run on a grid 16 times smaller and see 16 fine
cells in every direction (farther than sound travels in K = 16 steps); the
network is 64 -> C -> C -> C -> 64 channels with SiLU, C = 32 or 64 (train_surrogate
trains C = 32 by default; the network classes default to 64).
Training and inference run on the project's own HIP kernels (FP16 or BF16
MFMA operands, FP32 accumulation, FP32 master weights, Adam); training data
are trajectories the HIP lattice Boltzmann run itself produces.

The moments are read through the same route as ``Simulation.m``
(pylbm/simulation.py:215-224: f2m over the halo domain) and the lattice
Boltzmann run is restarted from a surrogate state with pylbm's moment
initialisation, conserved moments -> equilibrium -> m2f
(pylbm/simulation.py:257-320).

:class:`SurrogateRun` advances a simulation by intervals of K steps: the
network does every interval except every ``judge_every``-th one, which the
plain lattice Boltzmann run computes from the current surrogate state; the
rel-L2 distance of the velocity between that interval's lattice Boltzmann
result and the network's prediction for it is the drift.  Past ``tol`` (or on
a non-finite drift) the run falls back to the plain lattice Boltzmann method
for the rest of the advance.
"""
import ctypes
import time

import numpy as np

from . import runtime as rtm
from .cases import QX, QY, RHO, decaying_d2q9, fourier_velocity
from .simulation import Simulation

PATCH = 4
IN_CH = 64
EPI_ACT, EPI_LIN, EPI_BWD = 0, 1, 2
# 16-bit element type of activations, gradients and packed weights (LBM_SUR_F16 / _BF16)
ELEMS = {"float16": 0, "bfloat16": 1}


def channels(hidden=64):
    if hidden not in (32, 64):
        raise ValueError("hidden channels must be 32 or 64 (got %r)" % hidden)
    return (IN_CH, hidden, hidden, hidden, IN_CH)


def _f3(v):
    return (ctypes.c_float * 3)(*[float(x) for x in v])


def init_params(seed=0, hidden=64, scale_last=0.1):
    """He-normal weights w[9][cin][cout] (last layer scaled down), zero biases."""
    rng = np.random.default_rng(seed)
    ch = channels(hidden)
    out = []
    for li in range(4):
        std = np.sqrt(2.0 / (9 * ch[li]))
        w = rng.normal(0.0, std, (9, ch[li], ch[li + 1]))
        if li == 3:
            w *= scale_last
        out.append([w.astype(np.float32), np.zeros(ch[li + 1], np.float32)])
    return out


class SurrogateNet:
    """Device-resident network for batches of up to ``batch`` fine fields of Hf x Wf;
    ``elem`` ('float16' or 'bfloat16') is the type of its 16-bit MFMA operands."""

    def __init__(self, rt, Hf, Wf, batch=1, hidden=64, train=True, params=None, seed=0, elem="float16"):
        if Hf % PATCH or Wf % (PATCH * 32):
            raise ValueError("the surrogate needs Hf % 4 == 0 and Wf % 128 == 0 (got %dx%d)" % (Hf, Wf))
        if elem not in ELEMS:
            raise ValueError("elem must be one of %s (got %r)" % (sorted(ELEMS), elem))
        self.elem_name, self.elem = elem, ELEMS[elem]
        self.rt, self.lib, self.batch, self.train = rt, rt.lib, batch, train
        self.stream = rt.stream  # every launch of the network goes to this stream
        self.Hf, self.Wf = Hf, Wf
        self.H, self.W = Hf // PATCH, Wf // PATCH
        self.hidden = hidden
        self.ch = channels(hidden)
        ch = self.ch
        npix = batch * self.H * self.W
        self._bufs = []
        al = self._alloc
        self.xin = al(npix * IN_CH * 2)
        # second input buffer of the fused inference step; run() exchanges it with xin
        # after an odd number of steps, so it has xin's size
        self.xalt = al(npix * IN_CH * 2)
        self.act = [self.xin] + [al(npix * hidden * 2) for _ in range(3)]
        self.y = al(npix * IN_CH * 4)
        self.w = [al(9 * ch[i] * ch[i + 1] * 4) for i in range(4)]
        self.b = [al(ch[i + 1] * 4) for i in range(4)]
        self.afwd = [al(9 * ch[i] * ch[i + 1] * 2) for i in range(4)]
        if train:
            self.z = [al(npix * hidden * 2) for _ in range(3)]
            self.tgt = al(npix * IN_CH * 4)
            self.dy = al(npix * IN_CH * 2)
            self.g = [al(npix * hidden * 2) for _ in range(2)]
            self.abwd = [None] + [al(9 * ch[i] * ch[i + 1] * 2) for i in range(1, 4)]
            nblk = self.lib.lbm_sur_wgrad_blocks()
            self.partial = al(nblk * (9 * 64 * 64 + 64) * 4)
            self.dwt = [al(9 * ch[i] * ch[i + 1] * 4) for i in range(4)]
            self.db = [al(ch[i + 1] * 4) for i in range(4)]
            nparam = [9 * ch[i] * ch[i + 1] + ch[i + 1] for i in range(4)]
            self.m = [al(n * 4) for n in nparam]
            self.v = [al(n * 4) for n in nparam]
            for i in range(4):
                self._zero(self.m[i], nparam[i] * 4)
                self._zero(self.v[i], nparam[i] * 4)
            self.lpart = al(256 * 8)
            self.step = 0
        self.set_params(params if params is not None else init_params(seed, hidden))

    # ------------------------------------------------------------ memory
    def _alloc(self, nbytes):
        p = self.rt.malloc(max(int(nbytes), 256))
        self._bufs.append(p)
        return p

    def _zero(self, ptr, nbytes):
        self.rt.check(self.lib.lbm_memset(ctypes.c_void_p(ptr), 0, int(nbytes), self.stream), "memset")

    def free(self):
        for p in self._bufs:
            self.rt.free(p)
        self._bufs = []

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    # ------------------------------------------------------------ params
    def set_params(self, params):
        ch = self.ch
        for i, (w, b) in enumerate(params):
            w = np.ascontiguousarray(w, dtype=np.float32).reshape(9, ch[i], ch[i + 1])
            b = np.ascontiguousarray(b, dtype=np.float32).reshape(ch[i + 1])
            self.rt.h2d(self.w[i], w)
            self.rt.h2d(self.b[i], b)
        self.pack()

    def get_params(self):
        ch = self.ch
        out = []
        for i in range(4):
            w = np.empty((9, ch[i], ch[i + 1]), np.float32)
            b = np.empty(ch[i + 1], np.float32)
            self.rt.d2h(w, self.w[i])
            self.rt.d2h(b, self.b[i])
            out.append([w, b])
        self.rt.sync()
        return out

    def pack(self):
        ch = self.ch
        for i in range(4):
            abwd = self.abwd[i] if self.train else None
            self.rt.check(self.lib.lbm_sur_pack(ctypes.c_void_p(self.w[i]), ch[i], ch[i + 1], ch[i], ch[i + 1],
                                                ctypes.c_void_p(self.afwd[i]), ctypes.c_void_p(abwd), self.elem,
                                                self.stream), "lbm_sur_pack")

    # ------------------------------------------------------------ passes
    def _conv(self, src, cinp, wpk, bias, zin, out, zout, coutp, epi, n):
        self.rt.check(self.lib.lbm_sur_conv(ctypes.c_void_p(src), cinp, ctypes.c_void_p(wpk),
                                            ctypes.c_void_p(bias), ctypes.c_void_p(zin), ctypes.c_void_p(out),
                                            ctypes.c_void_p(zout), coutp, epi, n, self.H, self.W, self.elem,
                                            self.stream), "lbm_sur_conv")

    def hidden_layers(self, n=1, keep=False):
        """Layers 0..2 from xin; keep=True stores the pre-activations for backward."""
        if n > self.batch:
            raise ValueError("batch %d exceeds the allocated %d" % (n, self.batch))
        if keep and not self.train:
            raise ValueError("network allocated without training buffers")
        ch = self.ch
        for i in range(3):
            self._conv(self.act[i], ch[i], self.afwd[i], self.b[i], None, self.act[i + 1],
                       self.z[i] if keep else None, ch[i + 1], EPI_ACT, n)

    def forward(self, n=1, keep=False):
        """y (FP32 coarse [n][H][W][64]) from xin."""
        self.hidden_layers(n, keep)
        self._conv(self.act[3], self.hidden, self.afwd[3], self.b[3], None, self.y, None, IN_CH, EPI_LIN, n)
        return self.y

    def apply(self, d_state, norm_args, n=1):
        """Inference step: d_state += y * dscale and xin = the next input (one fused kernel)."""
        dscale, shift, scale = norm_args
        self.hidden_layers(n)
        self.rt.check(self.lib.lbm_sur_apply_step(ctypes.c_void_p(self.act[3]), self.hidden,
                                                  ctypes.c_void_p(self.afwd[3]), ctypes.c_void_p(self.b[3]),
                                                  ctypes.c_void_p(d_state), ctypes.c_void_p(self.xin), dscale,
                                                  shift, scale, n, self.H, self.W, self.elem, self.stream),
                      "lbm_sur_apply_step")

    def run(self, d_state, norm_args, nsteps, src=None):
        """nsteps inference steps of one field in one library call: d_state = the state
        nsteps steps after src (default d_state); xin must hold src's input and then
        holds the input of the new d_state (xin and xalt trade places after an odd
        nsteps)."""
        dscale, shift, scale = norm_args
        if not hasattr(self, "_cnet"):
            c = rtm.LbmSurNet()
            for i in range(4):
                c.act[i] = self.act[i]
                c.wpk[i] = self.afwd[i]
                c.bias[i] = self.b[i]
            c.xalt = self.xalt
            c.hidden = self.hidden
            c.elem = self.elem
            self._cnet = c
        c = self._cnet
        self.rt.check(self.lib.lbm_sur_run(ctypes.byref(c), ctypes.c_void_p(src), ctypes.c_void_p(d_state), dscale,
                                           shift, scale, int(nsteps), 1, self.H, self.W, self.stream), "lbm_sur_run")
        if c.act[0] != self.xin:
            self.xin, self.xalt = c.act[0], c.xalt
            self.act[0] = self.xin

    def loss(self, n, gscale=1.0, sync=True):
        """Masked MSE of y against tgt; writes dY = 2 gscale (y - t). Returns the loss if sync."""
        out = ctypes.c_double(0.0)
        self.rt.check(self.lib.lbm_sur_mse(ctypes.c_void_p(self.y), ctypes.c_void_p(self.tgt), n * self.H * self.W,
                                           float(gscale), ctypes.c_void_p(self.dy), ctypes.c_void_p(self.lpart),
                                           ctypes.byref(out) if sync else None, self.elem, self.stream),
                      "lbm_sur_mse")
        return out.value if sync else None

    def backward(self, n):
        """Weight and bias gradients (dwt, db) of every layer from dY."""
        lib, st, ch = self.lib, self.stream, self.ch
        grad_in = self.dy
        for i in range(3, -1, -1):
            self.rt.check(lib.lbm_sur_wgrad(ctypes.c_void_p(self.act[i]), ch[i], ctypes.c_void_p(grad_in), ch[i + 1],
                                            n, self.H, self.W, ctypes.c_void_p(self.partial),
                                            ctypes.c_void_p(self.dwt[i]), ctypes.c_void_p(self.db[i]),
                                            self.elem, st), "lbm_sur_wgrad")
            if i > 0:
                out = self.g[i % 2]
                self._conv(grad_in, ch[i + 1], self.abwd[i], None, self.z[i - 1], out, None, ch[i], EPI_BWD, n)
                grad_in = out

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, eps_scale=1.0):
        """One Adam step on every layer, then repack the 16-bit operands."""
        self.step += 1
        ch = self.ch
        for i in range(4):
            self.rt.check(self.lib.lbm_sur_adam(ctypes.c_void_p(self.w[i]), ctypes.c_void_p(self.b[i]),
                                                ctypes.c_void_p(self.m[i]), ctypes.c_void_p(self.v[i]),
                                                ctypes.c_void_p(self.dwt[i]), ctypes.c_void_p(self.db[i]),
                                                ch[i], ch[i + 1], ch[i], ch[i + 1], float(lr), float(b1), float(b2),
                                                float(eps * eps_scale), self.step, self.stream), "lbm_sur_adam")
        self.pack()

    def train_step(self, n, lr, sync=False):
        """forward, loss-scaled MSE gradient, backward, Adam (same step as the unscaled loss)."""
        self.forward(n, keep=True)
        loss = self.loss(n, gscale=1.0, sync=sync)
        self.backward(n)
        self.adam(lr=lr, eps_scale=float(n * self.H * self.W * 48))
        return loss

    def grads(self):
        """Host copies of (dW[9][cin][cout], db[cout]) per layer, as produced by backward."""
        ch = self.ch
        out = []
        for i in range(4):
            dwt = np.empty((9, ch[i + 1], ch[i]), np.float32)
            db = np.empty(ch[i + 1], np.float32)
            self.rt.d2h(dwt, self.dwt[i])
            self.rt.d2h(db, self.db[i])
            self.rt.sync()
            out.append([dwt.transpose(0, 2, 1).astype(np.float64), db.astype(np.float64)])
        return out


class Normalization:
    """Per-channel input shift/scale and increment scale of the network."""

    def __init__(self, shift, scale, dscale):
        self.shift = np.asarray(shift, np.float64)
        self.scale = np.asarray(scale, np.float64)
        self.dscale = np.asarray(dscale, np.float64)

    @classmethod
    def from_states(cls, states, rho0=1.0):
        """states: host array (T, H, W, 4) of one trajectory at interval spacing."""
        s = np.asarray(states, np.float64)
        d = s[1:] - s[:-1]
        drho = np.sqrt(np.mean((s[..., 0] - rho0) ** 2))
        u = np.sqrt(np.mean(s[..., 1] ** 2 + s[..., 2] ** 2) / 2.0)
        ddrho = np.sqrt(np.mean(d[..., 0] ** 2))
        du = np.sqrt(np.mean(d[..., 1] ** 2 + d[..., 2] ** 2) / 2.0)
        return cls((rho0, 0.0, 0.0), (max(drho, 1e-12), u, u), (max(ddrho, 1e-12), du, du))


class LatticeState:
    """One periodic D2Q9 simulation plus the moment <-> state couplings the surrogate needs."""

    def __init__(self, n=512, nu=1e-3, dtype="float32", seed=0, rt=None):
        self.sim = Simulation(decaying_d2q9(n, nu=nu, seed=seed), dtype=dtype)
        if not self.sim.periodic_only:
            raise ValueError("the surrogate handles fully periodic runs only")
        self.rt = self.sim.rt
        self.lib = self.rt.lib
        self.H, self.W = self.sim.domain.shape_in
        c = self.sim.scheme.consm
        self.irho, self.iqx, self.iqy = c[RHO], c[QX], c[QY]
        self.npix = self.H * self.W
        M = self.sim.spec.M
        rows = [[float(v) for v in M[i]] for i in (self.irho, self.iqx, self.iqy)]
        self._mc = (ctypes.c_double * (3 * self.sim.ns))(*[v for r in rows for v in r])
        self._direct = self.sim.storage == rtm.LBM_F32
        self.blocked = True

    def read_state(self, d_state):
        """Conserved moments of the current populations into an FP32 state buffer
        (plain FP32 populations: straight from F; other formats: f2m then the moments)."""
        sim = self.sim
        F = sim.container.F
        if self._direct:
            self.rt.check(self.lib.lbm_sur_populations_to_state(ctypes.c_void_p(F.ptr), ctypes.byref(F.layout),
                                                                self._mc, self.H, self.W, ctypes.c_void_p(d_state),
                                                                self.rt.stream), "lbm_sur_populations_to_state")
            return
        sim._f2m(F, sim.container.m)
        sim._update_m = False
        m = sim.container.m
        self.rt.check(self.lib.lbm_sur_moments_to_state(ctypes.c_void_p(m.ptr), ctypes.byref(m.layout), self.irho,
                                                        self.iqx, self.iqy, self.H, self.W, ctypes.c_void_p(d_state),
                                                        self.rt.stream), "lbm_sur_moments_to_state")

    def _bind(self):
        """ctypes arguments of the judge path, built again whenever F and Fnew have
        exchanged their buffers (an odd number of launches)."""
        sim = self.sim
        c = sim.container
        m = c.m
        self._m_ptr, self._m_lay = ctypes.c_void_p(m.ptr), ctypes.byref(m.layout)
        self._f_lay = ctypes.byref(c.F.layout)
        self._params = ctypes.byref(sim.params)
        self._vel = ctypes.byref(sim.vel)
        self._parity = ctypes.c_int(0)
        self._bound = (c.F.ptr, c.Fnew.ptr)

    def restart(self, d_state):
        """Populations from a state: conserved moments -> equilibrium -> m2f (pylbm moment init),
        the last two in one kernel."""
        sim = self.sim
        c = sim.container
        if getattr(self, "_bound", None) != (c.F.ptr, c.Fnew.ptr):
            self._bind()
        st, lib, mod = self.rt.stream, self.lib, sim.module
        self.rt.check(lib.lbm_sur_state_to_moments(ctypes.c_void_p(d_state), self.H, self.W, self._m_lay, self.irho,
                                                   self.iqx, self.iqy, self._m_ptr, st), "lbm_sur_state_to_moments")
        self.rt.check(mod.lbm_gen_eq_m2f(self._m_ptr, self._m_lay, ctypes.c_void_p(c.F.ptr), self._f_lay,
                                         self._params, st), "eq_m2f")
        # the fold step writes every cell of Fnew and reads no halo of F
        sim._update_m = True

    def run(self, nsteps):
        """nsteps plain steps (fully periodic, no links): the temporally blocked run
        (several steps per launch in LDS, bit-identical to one launch per step) when
        ``blocked`` is set and the grid allows it, else one folded fused launch per
        step."""
        if nsteps % 2 or (self.blocked and self.sim._blocking_shape() is not None):
            self.sim.run(nsteps)
            return
        sim = self.sim
        c = sim.container
        if getattr(self, "_bound", None) != (c.F.ptr, c.Fnew.ptr):
            self._bind()
        sim.params.t = sim.t
        self.rt.check(sim.module.lbm_gen_run(ctypes.c_void_p(c.F.ptr), ctypes.c_void_p(c.Fnew.ptr), self._f_lay,
                                             self._vel, self._params, sim._link_sets, 0, int(nsteps), 0, 1,
                                             ctypes.byref(self._parity), self.rt.stream), "run")
        sim._update_m = True
        sim.nt += nsteps
        sim.t = sim.t + nsteps * sim.dt

    def initial_state(self, seed, rms=0.05, kmin=4.0, kmax=16.0):
        """Host FP32 state of the random Fourier velocity of ``seed`` (rho = 1)."""
        ux, uy = fourier_velocity(self.H, kmin=kmin, kmax=kmax, rms=rms, seed=seed)
        s = np.zeros((self.H, self.W, 4), np.float32)
        s[..., 0] = 1.0
        s[..., 1] = ux
        s[..., 2] = uy
        return s


class Trajectories:
    """Device-resident pool of trajectories (states every K steps) for training."""

    def __init__(self, lattice, K=16, steps=2000, pool=16, kband=(4.0, 16.0)):
        self.lat, self.K, self.kband = lattice, K, kband
        self.T = steps // K + 1
        self.pool = pool
        self.stride = lattice.npix * 16
        self.base = lattice.rt.malloc(pool * self.T * self.stride)
        self.filled = 0
        self.next = 0

    def state_ptr(self, slot, t):
        return self.base + (slot * self.T + t) * self.stride

    def generate(self, seed):
        """Run one trajectory from the Fourier state of ``seed``; returns its pool slot."""
        lat, rt = self.lat, self.lat.rt
        slot = self.next
        s0 = self.state_ptr(slot, 0)
        rt.h2d(s0, lat.initial_state(seed, kmin=self.kband[0], kmax=self.kband[1]))
        lat.restart(s0)
        lat.read_state(s0)
        for t in range(1, self.T):
            lat.run(self.K)
            lat.read_state(self.state_ptr(slot, t))
        self.next = (self.next + 1) % self.pool
        self.filled = min(self.filled + 1, self.pool)
        return slot

    def download(self, slot, every=1):
        out = []
        for t in range(0, self.T, every):
            h = np.empty((self.lat.H, self.lat.W, 4), np.float32)
            self.lat.rt.d2h(h, self.state_ptr(slot, t))
            out.append(h)
        self.lat.rt.sync()
        return np.stack(out)

    def free(self):
        if self.base:
            self.lat.rt.free(self.base)
            self.base = None


class Trainer:
    """Adam training of a SurrogateNet on pairs (state_t, state_{t+K}) of a pool.

    pushforward: probability that a sample's input is the network's own one-step
    prediction from state_{t-K} (no gradient through it) instead of state_t, so
    the network learns on the distribution its rollouts produce; noise: standard
    deviation of Gaussian noise added to the input state, in units of the input
    scale (the target is the increment from the perturbed state)."""

    def __init__(self, net, data, norm, seed=0, pushforward=0.0, noise=0.0):
        self.net, self.data, self.norm = net, data, norm
        self.rng = np.random.default_rng(seed)
        self.lib, self.rt = net.lib, net.rt
        self._shift, self._scale, self._dscale = _f3(norm.shift), _f3(norm.scale), _f3(norm.dscale)
        self._sigma = _f3(np.asarray(norm.scale) * noise)
        self.pushforward, self.noise = pushforward, noise
        self.scratch = None
        if pushforward > 0 or noise > 0:
            self.scratch = net.rt.malloc(net.batch * data.stride)
        self.losses = []
        self._draw = 0

    def free(self):
        if self.scratch:
            self.rt.free(self.scratch)
            self.scratch = None

    def _load_batch(self, n):
        net, d, lat = self.net, self.data, self.data.lat
        npix = lat.npix
        norm_args = (self._dscale, self._shift, self._scale)
        picks = []
        for b in range(n):
            slot = int(self.rng.integers(d.filled))
            pf = self.pushforward > 0 and self.rng.random() < self.pushforward
            t = int(self.rng.integers(1 if pf else 0, d.T - 1))
            picks.append((slot, t, pf))
        # pushforward inputs first: they run the network on the batch-0 buffers
        for b, (slot, t, pf) in enumerate(picks):
            sb = self.scratch + b * d.stride if self.scratch else None
            if pf:
                self.rt.d2d(sb, d.state_ptr(slot, t - 1), d.stride)
                self.rt.check(self.lib.lbm_sur_make_input(ctypes.c_void_p(sb), 1, lat.H, lat.W, self._shift,
                                                          self._scale, ctypes.c_void_p(net.xin), net.elem,
                                                          self.rt.stream), "lbm_sur_make_input")
                net.run(sb, norm_args, 1)
            elif self.noise > 0:
                self.rt.d2d(sb, d.state_ptr(slot, t), d.stride)
            if self.noise > 0:
                self._draw += 1
                self.rt.check(self.lib.lbm_sur_perturb(ctypes.c_void_p(sb), npix, self._sigma,
                                                       int(self._draw * 1000003 + b), self.rt.stream),
                              "lbm_sur_perturb")
        for b, (slot, t, pf) in enumerate(picks):
            s0 = self.scratch + b * d.stride if (pf or self.noise > 0) else d.state_ptr(slot, t)
            s1 = d.state_ptr(slot, t + 1)
            self.rt.check(self.lib.lbm_sur_make_input(ctypes.c_void_p(s0), 1, lat.H, lat.W, self._shift,
                                                      self._scale, ctypes.c_void_p(net.xin + b * npix * 8),
                                                      net.elem, self.rt.stream), "lbm_sur_make_input")
            self.rt.check(self.lib.lbm_sur_make_target(ctypes.c_void_p(s0), ctypes.c_void_p(s1), 1, lat.H, lat.W,
                                                       self._dscale, ctypes.c_void_p(net.tgt + b * npix * 16),
                                                       self.rt.stream), "lbm_sur_make_target")

    def steps(self, nsteps, lr, batch=None, log_every=0):
        n = batch or self.net.batch
        for i in range(nsteps):
            self._load_batch(n)
            sync = bool(log_every) and (i % log_every == 0 or i == nsteps - 1)
            loss = self.net.train_step(n, lr(self.net.step) if callable(lr) else lr, sync=sync)
            if sync:
                self.losses.append((self.net.step, loss))


class SurrogateRun:
    """Advance a LatticeState by intervals of K steps with the network, judged by the plain run."""

    def __init__(self, net, lattice, norm, K=16, judge_every=8, tol=1e-2, overlap=False):
        if net.Hf != lattice.H or net.Wf != lattice.W:
            raise ValueError("network and lattice sizes differ")
        self.net, self.lat, self.norm = net, lattice, norm
        self.K, self.judge_every, self.tol = K, judge_every, tol
        self.overlap = overlap
        self.rt, self.lib = lattice.rt, lattice.lib
        self.state = self.rt.malloc(lattice.npix * 16)
        self.pred = self.rt.malloc(lattice.npix * 16)
        self.ckpt = [self.rt.malloc(lattice.npix * 16) for _ in range(3)]
        # overlapped judging: the network runs on its own stream; per judge parity the
        # judge's input state and the network's prediction of that interval
        self.jin = [self.rt.malloc(lattice.npix * 16) for _ in range(2)]
        self.jpred = [self.rt.malloc(lattice.npix * 16) for _ in range(2)]
        self.jout = self.rt.malloc(lattice.npix * 16)
        sa = ctypes.c_void_p()
        self.rt.check(self.lib.lbm_stream_create(ctypes.byref(sa)), "lbm_stream_create")
        self.net_stream = sa
        self.ev = {k: [self.rt.event(), self.rt.event()] for k in ("in", "pred", "free", "drift")}
        self.lpart = self.rt.malloc(514 * 8)
        hp = ctypes.c_void_p()
        self.rt.check(self.lib.lbm_host_malloc(ctypes.byref(hp), 2 * 2 * 8), "lbm_host_malloc")
        self.hdrift = hp.value
        self.events = [self.rt.event(), self.rt.event()]
        self._shift, self._scale, self._dscale = _f3(norm.shift), _f3(norm.scale), _f3(norm.dscale)
        self.judges = []
        self.fallback_at = None

    def free(self):
        for p in [self.state, self.pred, self.lpart, self.jout] + self.ckpt + self.jin + self.jpred:
            self.rt.free(p)
        self.state = self.pred = self.lpart = self.jout = None
        self.ckpt, self.jin, self.jpred = [], [], []
        for pair in self.ev.values():
            for e in pair:
                self.lib.lbm_event_destroy(e)
        self.ev = {}
        if self.net_stream:
            self.lib.lbm_stream_destroy(self.net_stream)
            self.net_stream = None
        if self.hdrift:
            self.rt.check(self.lib.lbm_host_free(ctypes.c_void_p(self.hdrift)), "lbm_host_free")
            self.hdrift = None
        for e in self.events:
            self.lib.lbm_event_destroy(e)
        self.events = []

    def _make_input(self, d_state):
        self.rt.check(self.lib.lbm_sur_make_input(ctypes.c_void_p(d_state), 1, self.lat.H, self.lat.W, self._shift,
                                                  self._scale, ctypes.c_void_p(self.net.xin), self.net.elem,
                                                  self.rt.stream), "lbm_sur_make_input")

    def _net_step(self, d_state):
        """d_state += increment(net(xin)); xin must hold d_state's input and then holds the next one."""
        self.net.apply(d_state, (self._dscale, self._shift, self._scale))

    def rel_l2(self, a, b):
        out = ctypes.c_double(0.0)
        self.rt.check(self.lib.lbm_sur_rel_l2(ctypes.c_void_p(a), ctypes.c_void_p(b), self.lat.npix,
                                              ctypes.c_void_p(self.lpart), ctypes.byref(out), self.rt.stream),
                      "lbm_sur_rel_l2")
        return out.value

    def start_from_lattice(self):
        self.lat.read_state(self.state)
        self.judges = []
        self.fallback_at = None

    def _drift_async(self, slot):
        """Start the drift of pred against state into pinned slot `slot` (no host wait)."""
        host = self.hdrift + 16 * slot
        self.rt.check(self.lib.lbm_sur_rel_l2_async(ctypes.c_void_p(self.pred), ctypes.c_void_p(self.state),
                                                    self.lat.npix, ctypes.c_void_p(self.lpart), ctypes.c_void_p(host),
                                                    self.rt.stream), "lbm_sur_rel_l2_async")
        self.rt.record(self.events[slot])

    def _judge_read(self, slot, ckpt):
        """After the lattice ran a judged interval: its state into self.state and ckpt,
        the network input of that state into xin, and the drift of pred against it
        into pinned slot `slot` (one kernel for plain FP32 populations)."""
        lat = self.lat
        if not lat._direct:
            lat.read_state(self.state)
            self._drift_async(slot)
            self.rt.d2d(ckpt, self.state, lat.npix * 16)
            self._make_input(self.state)
            return
        F = lat.sim.container.F
        self.rt.check(self.lib.lbm_sur_judge_read(ctypes.c_void_p(F.ptr), ctypes.byref(F.layout), lat._mc, lat.H,
                                                  lat.W, ctypes.c_void_p(self.pred), ctypes.c_void_p(self.state),
                                                  ctypes.c_void_p(ckpt), self._shift, self._scale,
                                                  ctypes.c_void_p(self.net.xin), self.net.elem,
                                                  ctypes.c_void_p(self.lpart), ctypes.c_void_p(self.hdrift + 16 * slot),
                                                  self.rt.stream), "lbm_sur_judge_read")
        self.rt.record(self.events[slot])

    def _drift_wait(self, slot):
        self.rt.check(self.lib.lbm_event_sync(self.events[slot]), "lbm_event_sync")
        d, n = (ctypes.c_double * 2).from_address(self.hdrift + 16 * slot)
        return 0.0 if d == 0.0 and n == 0.0 else float(np.sqrt(d / n)) if n > 0 else float("nan")

    # ------------------------------------------------------------ helpers
    def _copy(self, dst, src, stream):
        self.rt.check(self.lib.lbm_memcpy_d2d(ctypes.c_void_p(dst), ctypes.c_void_p(src), self.lat.npix * 16,
                                              stream), "lbm_memcpy_d2d")

    def _record(self, event, stream):
        self.rt.check(self.lib.lbm_event_record(event, stream), "lbm_event_record")

    def _wait(self, stream, event):
        self.rt.check(self.lib.lbm_stream_wait_event(stream, event), "lbm_stream_wait_event")

    def advance(self, nintervals):
        """nintervals intervals of K lattice steps (see the class docstring); the result
        ends in self.state and in the lattice."""
        if self.overlap:
            self._advance_overlapped(nintervals)
        else:
            self._advance_anchored(nintervals)

    def _advance_overlapped(self, nintervals):
        """The network runs every interval on its own stream; every judge_every-th
        interval the lattice (on its stream) re-runs that interval from the same
        state, concurrently with the network's next intervals.  Judge j's lattice
        result is checkpoint j % 3; its drift is read on the host after judge j + 1
        is queued.  A failed judge returns to the last checkpoint whose judge passed
        (or the start) and runs the plain lattice method from there to the end."""
        lat, K, je = self.lat, self.K, self.judge_every
        A, B = self.net_stream, self.rt.stream
        norm_args = (self._dscale, self._shift, self._scale)
        ev = self.ev
        saved = self.net.stream
        self.net.stream = A
        confirmed, confirmed_at = 2, 0
        self.rt.d2d(self.ckpt[confirmed], self.state, lat.npix * 16)   # on B, before A starts
        self._record(ev["free"][0], B)
        self._wait(A, ev["free"][0])
        self.rt.check(self.lib.lbm_sur_make_input(ctypes.c_void_p(self.state), 1, lat.H, lat.W, self._shift,
                                                  self._scale, ctypes.c_void_p(self.net.xin), self.net.elem, A),
                      "make_input")
        pending = None
        i = j = 0

        def check(p):
            drift = self._drift_wait(p[0] % 2)
            self.judges.append(drift)
            return drift <= self.tol  # False for NaN as well

        def fall_back(at):
            self.fallback_at = at
            self.rt.check(self.lib.lbm_stream_sync(A), "lbm_stream_sync")
            self.rt.d2d(self.state, self.ckpt[confirmed], lat.npix * 16)
            lat.restart(self.state)
            lat.run(K * (nintervals - confirmed_at))
            lat.read_state(self.state)

        try:
            while i < nintervals:
                nsur = min(je - 1 - (i % je), nintervals - i)
                if nsur > 0:
                    self.net.run(self.state, norm_args, nsur)
                    i += nsur
                    continue
                q = j % 2
                if j >= 2:
                    self._wait(A, ev["free"][q])           # B is done with jin[q], jpred[q]
                self._copy(self.jin[q], self.state, A)
                self._record(ev["in"][q], A)
                self.net.run(self.state, norm_args, 1)   # the network's own step of the judged interval
                self._copy(self.jpred[q], self.state, A)
                self._record(ev["pred"][q], A)
                # the lattice re-runs the interval on B
                self._wait(B, ev["in"][q])
                lat.restart(self.jin[q])
                lat.run(K)
                lat.read_state(self.jout)
                self._wait(B, ev["pred"][q])
                self.rt.check(self.lib.lbm_sur_rel_l2_async(ctypes.c_void_p(self.jpred[q]), ctypes.c_void_p(self.jout),
                                                            lat.npix, ctypes.c_void_p(self.lpart),
                                                            ctypes.c_void_p(self.hdrift + 16 * q), B),
                              "lbm_sur_rel_l2_async")
                self._record(self.events[q], B)
                self._copy(self.ckpt[j % 3], self.jout, B)
                self._record(ev["free"][q], B)
                i += 1
                if pending is not None:
                    if not check(pending):
                        fall_back(pending[1] - 1)
                        return
                    confirmed, confirmed_at = pending[0] % 3, pending[1]
                pending = (j, i)
                j += 1
            if pending is not None and not check(pending):
                fall_back(pending[1] - 1)
                return
            # the lattice takes the final state
            self._record(ev["in"][0], A)
            self._wait(B, ev["in"][0])
            lat.restart(self.state)
        finally:
            self.rt.check(self.lib.lbm_stream_sync(A), "lbm_stream_sync")
            self.net.stream = saved

    def _advance_anchored(self, nintervals):
        """nintervals intervals of K lattice steps; state ends in self.state (and in the lattice).

        Every judge_every-th interval is computed by the plain lattice run from the
        surrogate state; its result is a checkpoint.  The drift of a judge (rel-L2
        of the velocity between the network's prediction for that interval and the
        lattice result) is read on the host only after the next cycle has been
        queued, so the stream never drains.  When a drift exceeds tol (or is not
        finite), the run falls back: it returns to the last checkpoint whose judge
        passed (or the start) and runs the plain lattice Boltzmann method from
        there to the end.  Judge j uses drift slot j % 2 and checkpoint j % 3; the
        start is checkpoint 2 (judge -1)."""
        lat, K, je = self.lat, self.K, self.judge_every
        norm_args = (self._dscale, self._shift, self._scale)
        nbytes = lat.npix * 16
        confirmed, confirmed_at = 2, 0          # checkpoint whose judge passed, its interval
        self.rt.d2d(self.ckpt[confirmed], self.state, nbytes)
        pending = None                          # (judge index, interval after it)
        self._make_input(self.state)
        j = 0
        i = 0

        def check(p):
            drift = self._drift_wait(p[0] % 2)
            self.judges.append(drift)
            return drift <= self.tol  # False for NaN as well

        def fall_back(at):
            self.fallback_at = at
            self.rt.d2d(self.state, self.ckpt[confirmed], nbytes)
            lat.restart(self.state)
            lat.run(K * (nintervals - confirmed_at))
            lat.read_state(self.state)

        while i < nintervals:
            nsur = min(je - 1 - (i % je), nintervals - i)
            if nsur > 0:
                self.net.run(self.state, norm_args, nsur)
                i += nsur
                continue
            # the plain run computes this interval from the surrogate state; the
            # network's prediction for it goes to pred
            self.net.run(self.pred, norm_args, 1, src=self.state)
            lat.restart(self.state)
            lat.run(K)
            # lattice state -> state and checkpoint j % 3, its network input, drift
            self._judge_read(j % 2, self.ckpt[j % 3])
            i += 1
            if pending is not None:
                if not check(pending):
                    fall_back(pending[1] - 1)
                    return
                confirmed, confirmed_at = pending[0] % 3, pending[1]
            pending = (j, i)
            j += 1
        if pending is not None and not check(pending):
            fall_back(pending[1] - 1)
            return
        if nintervals % je:
            # the network did the last intervals: the lattice takes its state (after a
            # judge the lattice already holds the state, with its own populations)
            lat.restart(self.state)


def train_surrogate(lattice, ntraj=200, steps=2000, K=16, pool=8, batch=8, steps_per_traj=100, lr=2e-3,
                    lr_min=1e-5, seed=0, kband=(4.0, 16.0), hidden=32, pushforward=0.0, noise=0.002, log=None,
                    elem="float16"):
    """Generate ntraj trajectories with the HIP lattice Boltzmann run and train a network on them.

    Returns (net, norm, info). Trajectory seeds are seed * 100000 + j."""
    rt = lattice.rt
    data = Trajectories(lattice, K=K, steps=steps, pool=pool, kband=kband)
    t0 = time.perf_counter()
    slot = data.generate(seed * 100000)
    norm = Normalization.from_states(data.download(slot, every=4))
    net = SurrogateNet(rt, lattice.H, lattice.W, batch=batch, hidden=hidden, train=True, seed=seed, elem=elem)
    trainer = Trainer(net, data, norm, seed=seed, pushforward=pushforward, noise=noise)
    total = ntraj * steps_per_traj

    def sched(step):
        f = min(step / float(total), 1.0)
        return lr_min + 0.5 * (lr - lr_min) * (1.0 + np.cos(np.pi * f))

    t_gen = time.perf_counter() - t0
    t_train = 0.0
    for j in range(ntraj):
        if j > 0:
            ta = time.perf_counter()
            data.generate(seed * 100000 + j)
            rt.sync()
            t_gen += time.perf_counter() - ta
        ta = time.perf_counter()
        trainer.steps(steps_per_traj, sched, batch=batch, log_every=steps_per_traj)
        rt.sync()
        t_train += time.perf_counter() - ta
        if log and (j % 20 == 0 or j == ntraj - 1):
            log("traj %d/%d loss %.4e (gen %.1fs train %.1fs)" % (j + 1, ntraj, trainer.losses[-1][1], t_gen, t_train))
    data.free()
    trainer.free()
    info = {"trajectories": ntraj, "steps_per_trajectory": steps, "K": K, "optimizer_steps": net.step,
            "batch": batch, "hidden": hidden, "elem": elem, "pushforward": pushforward, "noise": noise,
            "final_loss": trainer.losses[-1][1], "generate_s": t_gen, "train_s": t_train,
            "losses": trainer.losses}
    return net, norm, info
