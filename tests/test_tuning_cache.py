"""The persistent variant cache (passivetracerflows/tuning.py), on the host.

A stand-in object with the Problem/Ensemble variant interface records what the cache
applies: a matching entry is applied without tuning (five-field entries from before the
workgroup-lines flag read it as off; a fixed-shape slot, R = 0, is left alone); a missing
or stale entry runs the tuner and stores its choice.
"""
import json

from passivetracerflows import tuning


class FakeSlots:
    def __init__(self, names, variants):
        self._names = names
        self.variants = list(variants)
        self.set_calls = []
        self.tuned = 0

    def slots(self):
        return [f"{n} [kernel variant]" for n in self._names]

    def set_variant(self, i, R, W, occ3=False, par=False, wg=False):
        if R == 0:
            raise AssertionError("a fixed-shape slot takes no variant")
        self.set_calls.append((i, R, W, occ3, par, wg))
        self.variants[i] = (R, W, occ3, par, wg)

    def get_variant(self, i):
        return self.variants[i]

    def autotune(self):
        self.tuned += 1
        self.variants = [(8, 4, True, False, True) if v[0] else v for v in self.variants]


def test_cache_applies_stored_choice_and_skips_fixed_slots(tmp_path):
    path = str(tmp_path / "tuning.json")
    names = ["step_member", "stage1_x_advect", "stage1_y_forward"]
    data = {"key": [["step_member", 0, 0, False, False, False],
                    ["stage1_x_advect", 4, 8, True, False, False],
                    ["stage1_y_forward", 8, 16, False, False]]}   # five fields: wg off
    with open(path, "w") as f:
        json.dump(data, f)
    obj = FakeSlots(names, [(0, 0, False, False, False), (16, 1, False, False, False),
                            (16, 1, False, False, False)])
    assert tuning.apply_or_tune(obj, "key", path) == "cache"
    assert obj.tuned == 0
    assert obj.set_calls == [(1, 4, 8, True, False, False), (2, 8, 16, False, False, False)]


def test_cache_tunes_and_stores_when_missing_or_stale(tmp_path):
    path = str(tmp_path / "tuning.json")
    names = ["step_member", "stage1_x_advect"]
    obj = FakeSlots(names, [(0, 0, False, False, False), (16, 1, False, False, False)])
    assert tuning.apply_or_tune(obj, "key", path) == "tuned"
    assert obj.tuned == 1
    stored = tuning.load(path)["key"]
    assert stored == [["step_member", 0, 0, False, False, False],
                      ["stage1_x_advect", 8, 4, True, False, True]]
    # a stale entry (other slot names) is replaced
    obj2 = FakeSlots(["stage1_x_advect"], [(16, 1, False, False, False)])
    assert tuning.apply_or_tune(obj2, "key", path) == "tuned"
    assert [e[0] for e in tuning.load(path)["key"]] == ["stage1_x_advect"]
