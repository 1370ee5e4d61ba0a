"""Host logic of the work-queue autotuner (crvhip/autotune.py): the choice rule, the cache key and the cache file.
The device side is tested in test_gpu_batched.py."""
import math

from crvhip import autotune as AT


def test_pick_takes_the_smallest_median_and_the_earlier_on_ties():
    a, b, c = {"tile": 64}, {"tile": 128}, {"tile": 256}
    best, t = AT.pick([(a, [10.0, 30.0, 11.0]), (b, [9.0, 12.0, 50.0]), (c, [12.0, 12.0, 12.0])])
    assert best == a and t == 11.0                    # medians 11, 12, 12: an outlier does not decide
    best, t = AT.pick([(a, [5.0, 7.0]), (b, [6.0, 6.0])])
    assert best == a and t == 6.0                     # even counts: mean of the middle two; tie -> earlier
    assert AT.pick([]) == (None, math.inf)


def test_signature_buckets_steps_per_event():
    assert AT.signature("MI355X", 1000, 12000) == AT.signature("MI355X", 1000, 15000)      # 12 and 15: 2^3
    assert AT.signature("MI355X", 1000, 12000) != AT.signature("MI355X", 1000, 33000)      # 33: 2^5
    assert AT.signature("MI355X", 1000, 12000) != AT.signature("other", 1000, 12000)
    assert AT.signature("MI355X", 0, 0).endswith("spe2^0")


def test_cache_round_trip(tmp_path):
    path = str(tmp_path / "sub" / "tune.json")
    assert AT.load_cache(path) == {}
    AT.store_cache({"k": {"opts": {"tile": 128, "sort_tasks": 1, "grid": 0}, "median_ms": 3.5}}, path)
    assert AT.load_cache(path)["k"]["opts"]["tile"] == 128
    (tmp_path / "bad.json").write_text("{not json")
    assert AT.load_cache(str(tmp_path / "bad.json")) == {}
