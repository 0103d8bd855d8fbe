"""Host logic of the judged surrogate run (pylbm_amd/surrogate.py SurrogateRun.advance)
on the CPU: device buffers, the network and the lattice are replaced by a model
that tracks which simulated interval each buffer holds and by whom it was
produced, so the scheduling of judges, checkpoints, deferred drift reads and the
fall-back can be checked exactly, for the anchored and the overlapped schedule."""
import ctypes

import pytest

from pylbm_amd.surrogate import SurrogateRun


class _Rt:
    def __init__(self, bufs):
        self.bufs = bufs

    def d2d(self, dst, src, nbytes):
        self.bufs[dst] = self.bufs[src]

    def record(self, event):
        pass


class _Net:
    def __init__(self, bufs, log):
        self.bufs, self.log = bufs, log

    def run(self, ptr, norm_args, nsteps, src=None):
        t, hist = self.bufs[ptr if src is None else src]
        self.bufs[ptr] = (t + nsteps, hist + "s" * nsteps)
        self.log.append(("net", nsteps))


class _Lat:
    """The lattice holds one state; restart loads a buffer, read_state stores it."""

    def __init__(self, bufs, K, log):
        self.bufs, self.K, self.log = bufs, K, log
        self.cur = None
        self.npix = 1

    def restart(self, ptr):
        self.cur = self.bufs[ptr]

    def run(self, nsteps):
        assert nsteps % self.K == 0
        t, hist = self.cur
        self.cur = (t + nsteps // self.K, hist + "L" * (nsteps // self.K))
        self.log.append(("lbm", nsteps // self.K))

    def read_state(self, ptr):
        self.bufs[ptr] = self.cur


def _run(nintervals, je, bad=()):
    bufs = {}
    log = []
    run = object.__new__(SurrogateRun)
    run.overlap = False
    run.K, run.judge_every, run.tol = 16, je, 1e-2
    run.rt = _Rt(bufs)
    run.net = _Net(bufs, log)
    run.lat = _Lat(bufs, 16, log)
    run.state, run.pred = "state", "pred"
    run.ckpt = ["c0", "c1", "c2"]
    run._dscale = run._shift = run._scale = None
    run.judges, run.fallback_at = [], None
    judged = []

    def drift_async(slot):
        judged.append((slot, bufs["state"][0]))

    def drift_wait(slot):
        # the judge whose drift sits in this slot: the latest one written there
        k = max(i for i, (s, _) in enumerate(judged) if s == slot)
        return 1.0 if k in bad else 1e-3

    def judge_read(slot, ckpt):
        run.lat.read_state("state")
        drift_async(slot)
        run.rt.d2d(ckpt, "state", 0)

    run._drift_async = drift_async
    run._judge_read = judge_read
    run._drift_wait = drift_wait
    run._make_input = lambda ptr: None
    bufs["state"] = (0, "")
    run.advance(nintervals)
    return run, bufs, log


@pytest.mark.parametrize("nint,je", [(16, 8), (125, 8), (7, 8), (13, 4), (1, 1)])
def test_all_judges_pass(nint, je):
    run, bufs, log = _run(nint, je)
    t, hist = bufs["state"]
    assert t == nint and run.fallback_at is None
    assert len(run.judges) == nint // je
    expect = "".join("L" if (i + 1) % je == 0 else "s" for i in range(nint))
    assert hist == expect
    # the lattice ends holding the state (restarted from it, or its own judged result)
    assert run.lat.cur == bufs["state"]


@pytest.mark.parametrize("nint,je,bad", [(64, 8, 0), (64, 8, 1), (64, 8, 3), (64, 8, 7), (20, 4, 2), (5, 1, 0)])
def test_failed_judge_falls_back_to_last_passed_checkpoint(nint, je, bad):
    run, bufs, log = _run(nint, je, bad=(bad,))
    t, hist = bufs["state"]
    # judge `bad` covers interval (bad + 1) * je - 1; its predecessor's checkpoint holds
    # the state after interval bad * je (the start for bad = 0)
    assert run.fallback_at == (bad + 1) * je - 1
    start = bad * je
    assert t == nint
    prefix = "".join("L" if (i + 1) % je == 0 else "s" for i in range(start))
    assert hist == prefix + "L" * (nint - start)
    assert len(run.judges) == bad + 1
    assert run.lat.cur == bufs["state"]


# ------------------------------------------------------------ overlapped schedule
def _v(x):
    return x.value if isinstance(x, ctypes.c_void_p) else x


class _Lib:
    """The entry points the overlapped schedule calls, on the buffer model."""

    def __init__(self, bufs, judged):
        self.bufs, self.judged = bufs, judged

    def lbm_memcpy_d2d(self, dst, src, nbytes, stream):
        self.bufs[_v(dst)] = self.bufs[_v(src)]
        return 0

    def lbm_sur_rel_l2_async(self, a, b, npix, part, host, stream):
        self.judged.append(((_v(host) - 1000) // 16, self.bufs[_v(b)]))
        return 0

    def lbm_sur_make_input(self, *args):
        return 0

    def lbm_event_record(self, e, s):
        return 0

    def lbm_stream_wait_event(self, s, e):
        return 0

    def lbm_stream_sync(self, s):
        return 0


class _RtO(_Rt):
    def check(self, rc, what):
        assert rc == 0, what


def _run_overlapped(nintervals, je, bad=()):
    bufs = {}
    log = []
    judged = []
    run = object.__new__(SurrogateRun)
    run.overlap = True
    run.K, run.judge_every, run.tol = 16, je, 1e-2
    run.rt = _RtO(bufs)
    run.rt.stream = "B"
    run.lib = _Lib(bufs, judged)
    run.net = _Net(bufs, log)
    run.net.stream = "B"
    run.net.xin = 0
    run.net.elem = 0
    run.lat = _Lat(bufs, 16, log)
    run.lat.H = run.lat.W = 4
    run.state, run.pred, run.lpart, run.jout, run.hdrift = 1, 2, 50, 40, 1000
    run.ckpt, run.jin, run.jpred = [10, 11, 12], [20, 21], [30, 31]
    run.net_stream = "A"
    run.ev = {k: [None, None] for k in ("in", "pred", "free", "drift")}
    run.events = [None, None]
    run._dscale = run._shift = run._scale = None
    run.judges, run.fallback_at = [], None

    def drift_wait(slot):
        k = max(i for i, (s, _) in enumerate(judged) if s == slot)
        return 1.0 if k in bad else 1e-3

    run._drift_wait = drift_wait
    bufs[1] = (0, "")
    run.advance(nintervals)
    assert run.net.stream == "B"
    return run, bufs, judged


@pytest.mark.parametrize("nint,je", [(16, 8), (125, 8), (7, 8), (13, 4), (1, 1), (3, 1)])
def test_overlapped_all_judges_pass(nint, je):
    run, bufs, judged = _run_overlapped(nint, je)
    t, hist = bufs[1]
    # the network does every interval; the lattice only re-ran the judged ones
    assert t == nint and hist == "s" * nint and run.fallback_at is None
    assert len(run.judges) == nint // je
    # judge k compared the lattice re-run of interval (k + 1) je - 1 from the network state
    for k, (slot, (tj, hj)) in enumerate(judged):
        assert slot == k % 2 and tj == (k + 1) * je and hj == "s" * ((k + 1) * je - 1) + "L"


@pytest.mark.parametrize("nint,je,bad", [(64, 8, 0), (64, 8, 1), (64, 8, 3), (64, 8, 7), (20, 4, 2), (5, 1, 0),
                                         (6, 1, 4)])
def test_overlapped_failed_judge_falls_back_to_last_passed_checkpoint(nint, je, bad):
    run, bufs, judged = _run_overlapped(nint, je, bad=(bad,))
    t, hist = bufs[1]
    assert run.fallback_at == (bad + 1) * je - 1
    assert t == nint
    # the last passed judge (bad - 1) holds the lattice re-run of its interval
    if bad == 0:
        prefix = ""
    else:
        prefix = "s" * (bad * je - 1) + "L"
    assert hist == prefix + "L" * (nint - len(prefix))
    assert len(run.judges) == bad + 1
