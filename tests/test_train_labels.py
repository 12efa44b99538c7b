"""Proposer training's host logic (CPU): a CVO-labelled set written in parts
with --label-only and merged back with --labels.  The GPU labelling and the
model are replaced by stand-ins; the checks are on which pairs and labels
reach training: a part's file regenerates exactly its kept pairs from their
seeds, and the merged set is this run's kept pairs followed by the file's."""
import json
import sys

import numpy as np

import rkhs_ba_amd
import rkhs_ba_amd.proposer as proposer
import rkhs_ba_amd.train_proposer as tp


def _fake_labels(a, b, t, starts, iters, **kw):
    # keep every third pair; label = the generating transform plus a marker
    keep = np.arange(0, a.shape[0], 3)
    lab = t[keep].copy()
    lab[:, 6] += 100.0
    return keep, lab, 0.5


class _FakeCvo:
    def __init__(self, *a, **k):
        pass


class _FakeProp:
    seen = []

    def __init__(self, cvo, seed=0):
        pass

    def propose(self, a, b):
        n = a.shape[0]
        qt = np.zeros((n, 8, 7), np.float32)
        qt[:, :, 0] = 1.0
        return qt, np.zeros((n, 8), np.float32), None

    def train_step(self, a, b, t, lr):
        _FakeProp.seen.append((a.copy(), t.copy()))
        return 0.0

    def save(self, path):
        np.save(path, np.zeros(1, np.float32))


def test_label_parts_merge(tmp_path, monkeypatch):
    monkeypatch.setattr(tp, "cvo_labels", _fake_labels)
    monkeypatch.setattr(rkhs_ba_amd, "CvoGPU", _FakeCvo)
    monkeypatch.setattr(proposer, "PoseProposer", _FakeProp)
    part = str(tmp_path / "part1.npz")
    common = ["--npts", "64", "--val", "4", "--workers", "2", "--batch", "8"]
    monkeypatch.setattr(sys, "argv", ["train", "--seed", "1", "--pairs", "30", "--label-only", part] + common)
    tp.main()
    z = np.load(part, allow_pickle=False)
    assert int(z["first_seed"]) == 2_000_000 and int(z["n"]) == 30
    np.testing.assert_array_equal(z["keep"], np.arange(0, 30, 3))
    assert z["labels"].shape == (10, 7) and np.all(z["labels"][:, 6] > 50.0)

    out = str(tmp_path / "w.npy")
    _FakeProp.seen = []
    monkeypatch.setattr(sys, "argv", ["train", "--seed", "2", "--pairs", "24", "--epochs", "1", "--labels", part,
                                      "--out", out] + common)
    tp.main()
    summary = json.load(open(str(tmp_path / "w.json")))
    assert summary["pairs"] == 8 + 10
    lab = summary["labels"]
    assert (lab["kept"], lab["of"]) == (18, 54)
    assert [p["first_seed"] for p in lab["parts"]] == [3_000_000, 2_000_000]
    # every trained pair is a kept pair of one of the two parts, with its label
    a1, _, t1 = tp.dataset(2_000_000, 30, 64, 2)
    a2, _, t2 = tp.dataset(3_000_000, 24, 64, 2)
    want = {a2[i].tobytes(): t2[i, 6] + 100.0 for i in range(0, 24, 3)}
    want.update({a1[i].tobytes(): t1[i, 6] + 100.0 for i in range(0, 30, 3)})
    n = 0
    for a, t in _FakeProp.seen:
        for k in range(a.shape[0]):
            assert abs(want[a[k].tobytes()] - t[k, 6]) < 1e-4
            n += 1
    assert n == 16  # 18 pairs in batches of 8: two full batches an epoch
