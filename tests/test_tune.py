"""The autotuner keeps the configured pad when the pad is physical (QA or
PID on) and sweeps it otherwise (host logic only: the batch is a stand-in
that records what it was asked to build)."""
from pas_cg_amd import params as P
from pas_cg_amd import tune
from pas_cg_amd import workloads as W


class FakeBatch:
    pads = []

    def __init__(self, p, models, **kw):
        FakeBatch.pads.append(round(p["nl_pad"] / P.ANGSTROM, 6))
        self.pad = p["nl_pad"]

    def run(self, n):
        pass

    def autotune(self, n):
        return 256, 0.0

    def last_ms(self):
        return 100.0 / (1.0 + self.pad), 1

    def close(self):
        pass


def _tune(monkeypatch, p):
    FakeBatch.pads = []
    monkeypatch.setattr(tune.G, "Batch", FakeBatch)
    return tune.tune(p, [None], trial_steps=10)


def test_pad_fixed_with_qa_or_pid(monkeypatch):
    p = W.full_params(pad_A=2.0)
    assert tune.pad_is_physical(p)
    r = _tune(monkeypatch, p)
    assert FakeBatch.pads == [2.0] and r["pad_A"] == 2.0 and r["pad_fixed"]
    p.update(qa_on=0, pid_on=1)
    r = _tune(monkeypatch, p)
    assert FakeBatch.pads == [2.0] and r["pad_fixed"]


def test_pad_swept_without_qa_and_pid(monkeypatch):
    p = W.base_params()
    assert not tune.pad_is_physical(p)
    r = _tune(monkeypatch, p)
    assert FakeBatch.pads == [4.0, 6.0, 8.0, 10.0, 12.0]
    assert r["pad_A"] == 12.0 and not r["pad_fixed"]
