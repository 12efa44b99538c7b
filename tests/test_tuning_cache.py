"""The autotuner's JSON cache (no device needed): store / lookup round trip,
replacement of an entry, several keys in one file, and rejection of bad
values, bad keys and damaged files."""
import json

import pytest

from rkhs_ba_amd.runtime import CvoError, tuning_cache_lookup, tuning_cache_store

KEY_A = "gfx950/256cu|src16.0|tgt16.0|F5|K19|geo1|col1|sem1|gt0|ell0.5-0.05|sp0.001|nn512|knn0"
KEY_B = "gfx950/256cu|src14.5|tgt14.5|F5|K0|geo1|col1|sem0|gt0|ell0.5-0.05|sp0.001|nn512|knn0"


def test_round_trip_and_replace(tmp_path):
    path = str(tmp_path / "a" / "b" / "tune.json")
    assert tuning_cache_lookup(path, KEY_A) is None
    tuning_cache_store(path, KEY_A, 16, 1, 0.7, lanes_small_ell=8, split_ell=0.11, cell_scale_small_ell=1.0,
                       score_us=21.5, n_tried=48)
    e = tuning_cache_lookup(path, KEY_A)
    assert e == {"lanes_per_row": 16, "lanes_small_ell": 8, "split_ell": 0.11, "xcd_remap": 1, "cell_scale": 0.7,
                 "cell_scale_small_ell": 1.0, "source": "cached", "score_us": 21.5, "n_tried": 48}
    tuning_cache_store(path, KEY_B, 16, 0, 1.0, score_us=5.0, n_tried=48)
    assert tuning_cache_lookup(path, KEY_B)["lanes_small_ell"] == 16
    assert tuning_cache_lookup(path, KEY_B)["cell_scale_small_ell"] == 1.0
    tuning_cache_store(path, KEY_A, 16, 1, 0.5, score_us=19.0, n_tried=48)
    assert tuning_cache_lookup(path, KEY_A)["cell_scale"] == 0.5
    assert tuning_cache_lookup(path, KEY_B)["lanes_per_row"] == 16
    doc = json.load(open(path))          # the file stays valid JSON
    assert doc["version"] == 1 and set(doc["entries"]) == {KEY_A, KEY_B}
    assert tuning_cache_lookup(path, "other") is None


def test_rejects_bad_values_and_keys(tmp_path):
    path = str(tmp_path / "tune.json")
    for bad in [(2, 0, 1.0), (32, 0, 1.0), (16, 2, 1.0), (16, 0, 0.2), (8, 0, 1.2)]:
        with pytest.raises(CvoError):
            tuning_cache_store(path, KEY_A, *bad)
    with pytest.raises(CvoError):
        tuning_cache_store(path, KEY_A, 16, 0, 1.0, lanes_small_ell=12, split_ell=0.1)
    with pytest.raises(CvoError):
        tuning_cache_store(path, KEY_A, 16, 0, 1.0, lanes_small_ell=8, split_ell=-0.1)
    with pytest.raises(CvoError):
        tuning_cache_store(path, KEY_A, 16, 0, 1.0, lanes_small_ell=8, split_ell=0.1, cell_scale_small_ell=0.3)
    for key in ["", 'a"b', "a{b", "a\nb"]:
        with pytest.raises(CvoError):
            tuning_cache_store(path, key, 8, 0, 1.0)


def test_damaged_entries_are_parse_errors(tmp_path):
    path = tmp_path / "tune.json"
    path.write_text('{\n"version": 1,\n"entries": {\n"%s": {"lanes_per_row": 12, "xcd_remap": 0, "cell_scale": 1.0},\n'
                    '"%s": {"xcd_remap": 0}\n}\n}\n' % (KEY_A, KEY_B))
    # an entry without the lane policy fields reads as one lane count
    other = tmp_path / "old.json"
    other.write_text('{\n"version": 1,\n"entries": {\n"%s": {"lanes_per_row": 8, "xcd_remap": 1, "cell_scale": 0.7}\n'
                     '}\n}\n' % KEY_A)
    e = tuning_cache_lookup(str(other), KEY_A)
    assert (e["lanes_per_row"], e["lanes_small_ell"], e["split_ell"], e["cell_scale_small_ell"]) == (8, 8, 0.0, 0.7)
    with pytest.raises(CvoError):
        tuning_cache_lookup(str(path), KEY_A)   # 12 lanes is not a valid setting
    with pytest.raises(CvoError):
        tuning_cache_lookup(str(path), KEY_B)   # fields missing
