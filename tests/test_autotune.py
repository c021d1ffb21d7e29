"""Host logic of the tile-shape autotuner on the CPU: the per-shape cache, its
keys, and the choice and reuse of a batch shape (a stand-in batch records the
calls; the device timing itself runs in the GPU tests and the benchmark)."""

from __future__ import annotations

import json

import pytest

from pcbhip import autotune


class FakeBatch:
    """Pass time as a function of the shape; records every call."""

    def __init__(self, times):
        self.times = times
        self.shape = None
        self.calls = []
        self.slots = {}

    def set_shape(self, rows, target):
        self.shape = (rows, target)
        self.calls.append(("shape", rows, target))

    def set(self, i, rhs, x0, config):
        self.slots[i] = (rhs, x0, config)
        self.calls.append(("set", i))

    def run(self):
        self.calls.append(("run",))

    def time_pass(self, launches):
        self.calls.append(("time", self.shape))
        return self.times[self.shape]


def _times(best):
    return {shape: (1.0 if shape == best else 2.0 + shape[1] / 4096) for shape in autotune.pass_candidates()}


@pytest.fixture(autouse=True)
def _no_device_config(monkeypatch):
    monkeypatch.setattr(autotune, "one_step_config", lambda: "one-step")


def test_keys_ignore_board_order_but_not_repeats_or_precision():
    a = autotune.board_set_key("batch", [(4, 32, 40), (8, 100, 90)])
    assert a == autotune.board_set_key("batch", [(8, 100, 90), (4, 32, 40)])
    assert a != autotune.board_set_key("batch", [(8, 100, 90), (4, 32, 40), (4, 32, 40)])
    assert a != autotune.board_set_key("batch", [(8, 100, 90), (4, 32, 40)], inner_fp64=True)
    assert a != autotune.board_set_key("serve", [(8, 100, 90), (4, 32, 40)])


def test_cache_round_trip_and_unreadable_files(tmp_path):
    cache = autotune.ShapeCache(tmp_path / "sub" / "shapes.json")
    assert cache.get("k") is None
    cache.put("k", {"rows": 8, "target": 2048, "us": 3.5})
    assert autotune.ShapeCache(tmp_path / "sub" / "shapes.json").get("k")["target"] == 2048
    # A shape outside the candidates (an old or edited file) is not used.
    cache.put("bad", {"rows": 6, "target": 2048})
    assert cache.get("bad") is None
    garbled = tmp_path / "garbled.json"
    garbled.write_text("{not json")
    assert autotune.ShapeCache(garbled).get("k") is None


def test_cache_path_follows_the_environment(tmp_path, monkeypatch):
    monkeypatch.setenv(autotune.CACHE_ENV, str(tmp_path / "env.json"))
    assert autotune.ShapeCache().path == tmp_path / "env.json"


def test_tune_batch_times_every_candidate_restores_the_slots_and_caches(tmp_path):
    cache = autotune.ShapeCache(tmp_path / "shapes.json")
    best = (8, 2048)
    batch = FakeBatch(_times(best))
    slots = [("rhs0", None, "cfg0"), ("rhs1", "x1", "cfg1")]
    chosen = autotune.tune_batch(batch, "key", slots=slots, cache=cache)
    assert (chosen["rows"], chosen["target"]) == best and not chosen["cached"]
    assert batch.shape == best
    timed = [c[1] for c in batch.calls if c[0] == "time"]
    assert sorted(timed) == sorted(autotune.pass_candidates())
    # The probe run happens with the one-step configuration; afterwards every
    # slot holds its own right-hand side, start and configuration again.
    run_at = batch.calls.index(("run",))
    assert all(c[0] != "time" for c in batch.calls[:run_at])
    assert batch.slots == {0: slots[0], 1: slots[1]}
    assert json.loads((tmp_path / "shapes.json").read_text())["key"]["rows"] == 8


def test_a_cached_shape_is_set_without_timing(tmp_path):
    cache = autotune.ShapeCache(tmp_path / "shapes.json")
    cache.put("key", {"rows": 4, "target": 512, "us": 1.0})
    batch = FakeBatch(_times((8, 4096)))
    chosen = autotune.tune_batch(batch, "key", slots=[("r", None, "c")], cache=cache)
    assert chosen["cached"] and batch.shape == (4, 512)
    assert batch.calls == [("shape", 4, 512)]
