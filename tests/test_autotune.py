"""The multi-track shape autotuner (europaraypaths_amd/autotune.py) on the
host: it keeps the fastest candidate, sets it on the context and persists
it per device name; apply() restores it in a later context.  A stand-in
context times each candidate from a table (no GPU)."""
import json

import pytest

from europaraypaths_amd import autotune
from europaraypaths_amd._abi import ERP_TUNE_PAIRS_PER_LAUNCH, ERP_TUNE_TRACK_ROWS


class _Ctx:
    """erp_set_tuning / run_tracks_dev stand-in: a run lasts the time the
    table gives the current (rows, pairs) and records the calls."""

    def __init__(self, cost):
        self.cost, self.tune, self.runs = cost, {}, 0

    def set_tuning(self, key, value):
        self.tune[key] = value

    def run_tracks_dev(self, jobs):
        import time
        self.runs += 1
        time.sleep(self.cost[(self.tune[ERP_TUNE_TRACK_ROWS], self.tune[ERP_TUNE_PAIRS_PER_LAUNCH])])


@pytest.fixture
def no_gpu_buffers(monkeypatch):
    import europaraypaths_amd._native as nat

    class _Buf:
        def __init__(self, ns, nr):
            self.ns, self.nr = ns, nr

        def free(self):
            pass

    monkeypatch.setattr(nat, "DeviceTraces", _Buf)
    monkeypatch.setattr(autotune, "device_name", lambda device=0: "test-device")


def test_tune_picks_the_fastest_and_persists(tmp_path, no_gpu_buffers):
    cost = {(r, c): 0.004 for r in autotune.ROWS for c in autotune.PAIRS}
    cost[(32, 4096)] = 0.001
    ctx = _Ctx(cost)

    class _Par:
        nr = 1024

    jobs = [(_Par(), None, None, list(range(100)), None)] * 3
    rec = autotune.tune(ctx, autotune.sample_jobs(jobs, 2, 10), repeats=2)
    assert (rec["rows"], rec["pairs"]) == (32, 4096)
    assert ctx.tune == {ERP_TUNE_TRACK_ROWS: 32, ERP_TUNE_PAIRS_PER_LAUNCH: 4096}
    assert len(rec["candidates"]) == len(autotune.ROWS) * len(autotune.PAIRS)
    assert ctx.runs == 3 * len(rec["candidates"])   # one untimed, two timed each
    assert rec["sample"] == {"tracks": 2, "shots": 20}
    path = str(tmp_path / "tuning.json")
    autotune.save(rec, path=path)
    assert json.load(open(path))["test-device"]["rows"] == 32
    other = _Ctx(cost)
    assert autotune.apply(other, path=path)["pairs"] == 4096
    assert other.tune == {ERP_TUNE_TRACK_ROWS: 32, ERP_TUNE_PAIRS_PER_LAUNCH: 4096}
    assert autotune.apply(other, path=str(tmp_path / "none.json")) is None
