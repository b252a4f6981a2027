"""Serving schedule and admission logic (host only, CPU)."""
import numpy as np

from chicago_amd import serve


def test_schedule_keeps_live_count_and_mix():
    lives = serve.make_schedule(3, n_live=50, duration=30.0, mean_life=5.0)
    for t in (0.0, 7.3, 15.0, 29.9):
        assert sum(1 for l in lives if l.t_join <= t < l.t_leave) == 50
    sizes = [(l.width, l.height) for l in lives]
    share_720 = sum(1 for s in sizes if s == (1280, 720)) / len(sizes)
    assert 0.35 < share_720 < 0.65
    again = serve.make_schedule(3, n_live=50, duration=30.0, mean_life=5.0)
    assert [(l.width, l.t_join, l.t_leave) for l in lives] == \
        [(l.width, l.t_join, l.t_leave) for l in again]


def test_arrivals_rate_and_order():
    lives = serve.make_schedule(5, n_live=10, duration=2.0, mean_life=100.0)
    arr = serve.frame_arrivals(lives, 6, fps=60.0, jitter=0.002, duration=2.0)
    t = np.array([a[0] for a in arr])
    assert np.all(np.diff(t) >= 0)
    # about 60 frames per second per live stream, frame 0 (intra) excluded
    assert abs(len(arr) - 10 * 2.0 * 60) <= 10 * 2
    per = {}
    for _, sid, k in arr:
        per.setdefault(sid, []).append(k)
    for ks in per.values():
        assert ks == list(range(1, len(ks) + 1))
    # jitter only delays frames, by 2 ms on average
    life = {l.sid: l for l in lives}
    delays = [t0 - (life[sid].t_join + (life[sid].phase + k) / 60.0) for t0, sid, k in arr]
    assert min(delays) >= 0 and 0.001 < np.mean(delays) < 0.003


class _FakeStream:
    def __init__(self, w, h):
        self.n_mbs = (w // 16) * (h // 16)
        self.resets = 0
        self.uploads = 0

    def upload(self, *a):
        self.uploads += 1

    def reset_temporal(self):
        self.resets += 1


class _FakeCtx:
    """Context stand-in with an HBM budget counted in streams."""

    def __init__(self, max_streams):
        self.max_streams, self.opened = max_streams, 0

    def open_stream(self, w, h, n):
        from chicago_amd import evxa
        if self.opened >= self.max_streams:
            raise evxa.EvxaError(7, "stream does not fit the HBM budget")
        self.opened += 1
        return _FakeStream(w, h)


def test_admission_memory_and_compute():
    frames = {(1280, 720): [(None, None, None)] * 3}
    adm = serve.Admission(_FakeCtx(2), 3, frames, capacity_mbps=None)
    a = adm.acquire(1280, 720)
    b = adm.acquire(1280, 720)
    assert a is not None and b is not None and a.uploads == 3
    assert adm.acquire(1280, 720) is None and adm.rejected_memory == 1
    adm.release(1280, 720, a)
    c = adm.acquire(1280, 720)               # reuses the pooled device stream
    assert c is a and adm.pools[(1280, 720)].opened == 2
    per = adm.demand(1280, 720)
    adm2 = serve.Admission(_FakeCtx(10), 3, frames, capacity_mbps=2.5 * per)
    assert adm2.acquire(1280, 720) and adm2.acquire(1280, 720)
    assert adm2.acquire(1280, 720) is None and adm2.rejected_compute == 1
