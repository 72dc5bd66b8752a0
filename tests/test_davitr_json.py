"""exec.davitr JSON in the layout of output_ev_and_time
(Wstat/davidson_diago.f90:1296-1332), read the way the reference's test
harness reads it (test-suite/main_test.py:36-51: exec.davitr[-1].ev)."""
import json

from west_amd.wstat import _human_time, write_davitr


def test_davitr_layout_and_append(tmp_path):
    f = tmp_path / "wstat.json"
    f.write_text(json.dumps({"input": {"wstat_control": {"n_pdep_eigen": 2}}}))
    recs = [{"dav_iter": 0, "ev": [-1.5, -1.2], "conv": [False, False], "notcnv": 2, "time_elap:sec": 2.63,
             "time_iter:sec": 2.63, "sternop_ncalls": 41, "dfpt_tr2": 1e-4, "dfpt_dim": 2, "diago_dim": 2},
            {"dav_iter": 1, "ev": [-1.6, -1.25], "conv": [True, False], "notcnv": 1, "time_elap:sec": 75.25,
             "time_iter:sec": 72.62, "sternop_ncalls": 44, "dfpt_tr2": 1e-16, "dfpt_dim": 2, "diago_dim": 4}]
    write_davitr(f, recs[:1])
    write_davitr(f, recs[1:])
    doc = json.loads(f.read_text())
    assert doc["input"]["wstat_control"]["n_pdep_eigen"] == 2       # the rest of the log is kept
    assert doc["exec"]["ndav"] == 2
    last = doc["exec"]["davitr"][-1]
    assert last["ev"] == [-1.6, -1.25] and last["conv"] == [True, False]
    assert set(last) == {"dav_iter", "ev", "conv", "notcnv", "time_elap:sec", "time_elap:hum", "time_iter:sec",
                         "time_iter:hum", "sternop_ncalls", "dfpt_tr2", "dfpt_dim", "diago_dim"}
    assert doc["exec"]["davitr"][0]["time_elap:hum"] == "02.6s"      # as test001/ref/wstat.json:293
    assert last["time_elap:hum"] == "01m-15.2s"


def test_human_readable_time():
    # Tools/human_readable_time.f90: tenths truncated, two-digit fields
    assert _human_time(0.05) == "< 00.1s"
    assert _human_time(6.1) == "06.1s"
    assert _human_time(3725.45) == "01h-02m-05.4s"
    assert _human_time(90061.7) == "1d-01h-01m-01.6s"
