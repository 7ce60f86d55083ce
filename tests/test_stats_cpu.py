"""Statistics files and schedule on the CPU (loma_phd_amd.stats).

* write_sta / write_spe produce the reference's record layout
  (escru.f:351-460): the record sizes of the files that the reference wrote
  in tests/golden/make_ref_stats.py (header 28 bytes, y/fmap 16 per row,
  33 float64 per row; .spe header 36 bytes, jsptot + y/fmap, one float32
  record per kx plane), and read_sta / read_spe give back what was written.
* advance_with_stats follows cross1's schedule (cross.f:234-239, 1145-1161):
  for nhist 3 and ntimes 2 over 7 steps the window covers steps 2-3 and 5-6
  and is read after steps 3 and 6; a start in the middle of a window does
  not read a window that was never opened.
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def test_sta_spe_round_trip_and_layout(tmp_path):
    from loma_phd_amd import stats as S
    z = np.load(os.path.join(HERE, "golden", "ref_stats_192x453x96.npz"), allow_pickle=False)
    my = int(z["grid"][1])
    sta = z["sta_cols_0"]
    nsp, nz1 = len(z["jsp"]), 31
    rng = np.random.default_rng(3)
    spe = rng.random((4, 12, nsp, nz1 + 1)).astype(np.float32)
    st = dict(sta=sta, spe=spe, nacum=2, nacumsp=2)
    y, fmap = z["yf"][:, 0], z["yf"][:, 1]
    S.write_sta(tmp_path / "a.sta", st, y=y, fmap=fmap, time=0.095, Re=160.0, alp=0.25,
                bet=0.25)
    S.write_spe(tmp_path / "a.spe", st, jsptot=z["jsp"], y=y, fmap=fmap, time=0.095, Re=160.0,
                alp=0.25, bet=0.25, mx=128, mz=63)
    rec = S.read_records(tmp_path / "a.sta")
    assert [len(r) for r in rec] == [28, 16 * my, 33 * 8 * my]
    rec = S.read_records(tmp_path / "a.spe")
    assert [len(r) for r in rec] == [36, 4 * nsp + 16 * my] + [12 * nsp * (nz1 + 1) * 4] * 4
    a = S.read_sta(tmp_path / "a.sta")
    assert a["my"] == my and a["nacum"] == 2
    assert np.array_equal(a["sta"], sta) and np.array_equal(a["fmap"], fmap)
    b = S.read_spe(tmp_path / "a.spe")
    assert np.array_equal(b["spe"], spe) and np.array_equal(b["jsptot"], z["jsp"])
    assert (b["mx"], b["my"], b["mz"], b["nacumsp"]) == (128, my, 63, 2)


class _Log:
    """A stand-in run that records what the schedule asks of it."""

    def __init__(self):
        self.istep, self.on, self.events = 0, False, []

    def advance(self, n):
        for _ in range(n):
            self.istep += 1
            self.events.append(("step", self.istep, self.on))

    def stats_accumulate(self, on=True):
        self.on = bool(on)

    def stats_read(self):
        self.events.append(("read", self.istep))
        return dict(nacum=sum(1 for e in self.events if e[0] == "step" and e[2]))

    def stats_reset(self):
        self.events = [e for e in self.events if e[0] != "step"]


def test_schedule_follows_cross1():
    from loma_phd_amd import stats as S
    run = _Log()
    wins = S.advance_with_stats(run, 7, nhist=3, ntimes=2)
    assert [(s, w["nacum"]) for s, w in wins] == [(3, 2), (6, 2)]
    acc = [e[1] for e in run.events if e[0] == "step" and e[2]]
    assert acc == []  # both windows were read and reset
    run = _Log()
    opened = []
    S.advance_with_stats(run, 7, nhist=3, ntimes=2,
                         on_window=lambda r, s, w: opened.append(s))
    assert opened == [3, 6]


def test_schedule_started_inside_a_window():
    from loma_phd_amd import stats as S
    run = _Log()
    run.istep = 2  # the next step is 3, inside the first window
    wins = S.advance_with_stats(run, 4, nhist=3, ntimes=2, first_step=3)
    # step 3 ends a window that was never opened; steps 5-6 form the next
    assert [(s, w["nacum"]) for s, w in wins] == [(6, 2)]
