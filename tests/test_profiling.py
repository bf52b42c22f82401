"""The rocprofv3 summary helpers on a small database of the same schema."""
import sqlite3

import pytest

import profiling


@pytest.fixture()
def prof_dir(tmp_path):
    d = tmp_path / "run"
    d.mkdir()
    con = sqlite3.connect(str(d / "run_results.db"))
    con.execute("create table counters_collection (kernel_name text, dispatch_id int, "
                "counter_name text, value real, duration int)")
    rows = [("void fpho::eval_kernel<8, true>(args)", k, "FETCH_SIZE", 100.0 + k, 115_000_000)
            for k in range(5)]
    rows += [("void fpho::eval_kernel<4, true>(args)", 10 + k, "FETCH_SIZE", 1.0, 500_000)
             for k in range(3)]
    con.executemany("insert into counters_collection values (?, ?, ?, ?, ?)", rows)
    con.execute("create table top_kernels (name text, total_calls int, total_duration real, "
                "average real, percentage real)")
    con.execute("insert into top_kernels values ('k', 2, 10.0, 5.0, 100.0)")
    con.execute("create table regions (name text, duration int)")
    con.executemany("insert into regions values (?, ?)",
                    [("hipGraphLaunch", 2000), ("hipGraphLaunch", 4000),
                     ("hipLaunchKernel", 500), ("hipStreamBeginCapture", 1000)])
    con.commit()
    con.close()
    return str(d)


def test_counter_medians_filters_short_dispatches(prof_dir):
    med = profiling.counter_medians(prof_dir, min_ms=10)
    assert list(med) == [("void fpho::eval_kernel<8, true>", "FETCH_SIZE")]
    n, value, dur = med[("void fpho::eval_kernel<8, true>", "FETCH_SIZE")]
    assert n == 5 and value == 102.0 and dur == 115_000_000
    assert len(profiling.counter_medians(prof_dir)) == 2


def test_kernel_stats_and_summary_round_trip(prof_dir, tmp_path):
    assert profiling.kernel_stats(prof_dir) == [("k", 2, 10.0, 5.0, 100.0)]
    f = tmp_path / "s.txt"
    f.write_text("# comment: x\nlaunch: 8,1,0,8\ntraffic_bytes_per_launch: 1.5e+10\n")
    kv = profiling.read_summary(str(f))
    assert kv == {"launch": "8,1,0,8", "traffic_bytes_per_launch": "1.5e+10"}
    assert len(profiling.kernel_source_digest()) == 16


def test_kernel_digest_ignores_comments_and_whitespace(tmp_path):
    a = tmp_path / "a.hip"
    b = tmp_path / "b.hip"
    c = tmp_path / "c.hip"
    a.write_text("// header\nint f(int x) { return x + 1; }  /* note */\n")
    b.write_text("/* other\n   note */ int f(int x) {\n    return x + 1;   // same code\n}\n")
    c.write_text("int f(int x) { return x + 2; }\n")
    d = profiling.kernel_source_digest
    assert d([str(a)]) == d([str(b)])
    assert d([str(a)]) != d([str(c)])


def test_api_stats_counts_calls_per_name(prof_dir):
    rows = profiling.api_stats(prof_dir)
    assert rows[0] == ("hipGraphLaunch", 2, 6.0, 3.0)
    assert {r[0] for r in rows} == {"hipGraphLaunch", "hipLaunchKernel", "hipStreamBeginCapture"}
    assert [r[0] for r in profiling.api_stats(prof_dir, "Graph")] == ["hipGraphLaunch"]
