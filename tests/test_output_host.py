"""Host logic of Diagnostic/increment!/Output (no GPU): FourierFlows semantics of
the record schedule and the group layout of the output directory."""
import os

import numpy as np

from passivetracerflows import output as OUT


class _Clock:
    def __init__(self):
        self.t, self.step, self.dt = 0.0, 0, 0.5


class _FakeProb:
    """Only what Diagnostic/increment_ read: clock and calc's argument."""

    def __init__(self):
        self.clock = _Clock()
        self.value = 1.0

    def advance(self):
        self.clock.t += self.clock.dt
        self.clock.step += 1
        self.value *= 2.0


def test_diagnostic_schedule_matches_fourierflows():
    p = _FakeProb()
    d = OUT.Diagnostic(lambda q: q.value, p, freq=3, nsteps=10)
    assert d.ndata == 4 and d.i == 1 and d.data == [1.0] and d.steps == [0]
    for _ in range(10):
        p.advance()
        OUT.increment_(d)
    # records at construction and after steps 3, 6, 9
    assert d.steps == [0, 3, 6, 9]
    assert d.t == [0.0, 1.5, 3.0, 4.5]
    assert d.data == [1.0, 8.0, 64.0, 512.0]
    assert d.value == 512.0
    # more records than ndata: the arrays grow (push!)
    for _ in range(6):
        p.advance()
        OUT.increment_([d])
    assert d.steps[-2:] == [12, 15] and d.ndata == d.i == 6


def test_uniquepath_and_jldopen(tmp_path):
    base = str(tmp_path / "out.jld2")
    os.makedirs(base)
    p1 = OUT._uniquepath(base)
    assert p1.endswith("out_1.jld2")
    OUT._write(p1, "snapshots/t/50", 1.0)
    OUT._write(p1, "snapshots/c/50", np.arange(4.0))
    OUT._write(p1, "grid/Lx", 2 * np.pi)
    f = OUT.jldopen(p1)
    assert f.keys("snapshots") == ["c", "t"]
    assert f.keys("snapshots/t") == ["50"]
    assert f["snapshots/t/50"] == 1.0
    assert np.array_equal(f["snapshots/c/50"], np.arange(4.0))
    assert f["grid/Lx"] == 2 * np.pi
