"""The batch autotuner's choice rule on the host (pypty_amd/autotune.py): the fastest candidate wins,
and a larger candidate within the tie tolerance of it wins over it.  HIP events are replaced by a
scripted clock, so no device is needed."""
import pytest


class _Clock:
    """Stands in for torch.cuda.Event: elapsed_time returns the scripted time of the candidate
    whose pass ran between the two records."""

    now = 0.0

    def __init__(self, enable_timing=False):
        self.t = None

    def record(self):
        self.t = _Clock.now

    def synchronize(self):
        pass

    def elapsed_time(self, other):
        return other.t - self.t


@pytest.fixture()
def tuner(monkeypatch):
    import torch

    from pypty_amd import autotune

    monkeypatch.setattr(torch.cuda, "Event", _Clock)
    monkeypatch.setattr(autotune, "_cache", {})
    monkeypatch.setattr(autotune, "_loaded", True)
    monkeypatch.delenv("PYPTY_AMD_TUNE_CACHE", raising=False)
    return autotune


def _passes(ms_per_pass):
    def make_pass(batch):
        def run():
            _Clock.now += ms_per_pass[batch]
        return run
    return make_pass


def test_fastest_candidate_wins(tuner):
    assert tuner.choose_batch("a", (64, 128, 256, 512), _passes({64: 9.0, 128: 7.0, 256: 5.0, 512: 6.0})) == 256


def test_larger_candidate_within_the_tie_wins(tuner):
    # 512 is 1% slower than 256: within the 2% tie, so the larger batch is kept
    assert tuner.choose_batch("b", (64, 128, 256, 512), _passes({64: 9.0, 128: 7.0, 256: 5.0, 512: 5.05})) == 512
    # 3% slower: outside the tie
    assert tuner.choose_batch("c", (64, 128, 256, 512), _passes({64: 9.0, 128: 7.0, 256: 5.0, 512: 5.15})) == 256
    # a smaller candidate within the tie does not displace the fastest
    assert tuner.choose_batch("d", (64, 128, 256, 512), _passes({64: 9.0, 128: 5.05, 256: 5.0, 512: 6.0})) == 256


def test_choice_is_cached_per_key(tuner):
    calls = []

    def make_pass(batch):
        calls.append(batch)
        return _passes({4: 2.0, 8: 1.0})(batch)

    assert tuner.choose_batch("e", (4, 8), make_pass) == 8
    assert tuner.choose_batch("e", (4, 8), make_pass) == 8 and calls == [4, 8]
