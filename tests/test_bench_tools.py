"""The trace summaries of bench.py on small synthetic rocprofv3-shaped databases (no GPU)."""
import os
import sqlite3
import sys

sys.path.insert(0, os.path.join(os.path.dirname(__file__), ".."))
import bench  # noqa: E402


def _kernel_db(path, names):
    con = sqlite3.connect(path)
    con.execute("create table kernels (name text, start integer, end integer)")
    t = 0
    for n in names:
        con.execute("insert into kernels values (?, ?, ?)", (n, t, t + 1000))
        t += 1000
    con.commit()
    con.close()


def test_replay_summary_finds_the_whole_replay(tmp_path, capsys):
    # one replay = a different group, then three length groups with the same kernel names: the last 12 kernels are a
    # period of 4 repeated three times, but the replay is 13 kernels long
    group = ["void k_pass<1>", "void k_fused2<16>", "void k_pass<2>", "void k_crt<3, 8>"]
    replay = ["void k_r3_fwd"] + group * 3
    db = str(tmp_path / "t.db")
    _kernel_db(db, ["void k_warm"] + replay * 4)          # one warm and three timed replays
    bench.replay_summary(db, 3)
    out = capsys.readouterr().out
    assert out.startswith("13 kernels per run")
    assert "| void k_pass<1> | 3 | 3.0 |" in out


def test_counter_summary_averages_per_dispatch(tmp_path, capsys):
    db = str(tmp_path / "c.db")
    con = sqlite3.connect(db)
    con.execute("create table counters_collection (dispatch_id integer, kernel_name text, counter_name text, value real)")
    rows = [(1, "k_a", "SQ_WAVES", 10), (1, "k_a", "SQ_WAVES", 6), (2, "k_a", "SQ_WAVES", 20), (3, "k_b", "SQ_WAVES", 4)]
    con.executemany("insert into counters_collection values (?, ?, ?, ?)", rows)
    con.commit()
    con.close()
    bench.counter_summary(db)
    out = capsys.readouterr().out.splitlines()
    assert out[0] == "| kernel | dispatches | SQ_WAVES |"
    assert out[2] == "| k_a | 2 | 18 |"                      # (10 + 6 + 20) / 2 dispatches
    assert out[3] == "| k_b | 1 | 4 |"
