"""Synthetic workloads: deterministic and shaped as the configurations describe."""
import numpy as np

import mhm2_staging_amd as m


def test_deterministic_across_thread_counts():
    a = m.generate(2, npairs=3000, threads=1)
    b = m.generate(2, npairs=3000, threads=7)
    for f in ("reads", "read_off", "read_len", "ctgs", "ctg_off", "ctg_len"):
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f))


def test_config_shapes():
    c1 = m.generate(1)
    assert c1.npairs == 10000 and (c1.read_len == 150).all() and (c1.ctg_len == 350).all()
    c2 = m.generate(2, npairs=20000)
    assert c2.read_len.min() >= 100 and c2.read_len.max() <= 250
    assert c2.ctg_len.min() >= 150 and c2.ctg_len.max() <= 1024
    frac_n = (c2.ctgs[:-1] == ord("N")).mean()
    assert 0.003 < frac_n < 0.007
    c3 = m.generate(3, npairs=20000)
    assert set(np.unique(c3.read_len)) == {150, 250}
    assert 0.65 < (c3.read_len == 150).mean() < 0.75
    assert 300 < np.median(c3.ctg_len) < 500 and c3.ctg_len.min() >= 150 and c3.ctg_len.max() <= 1024
    c4 = m.generate(4, npairs=4000)
    assert (c4.ctg_len == 1024).all() and set(np.unique(c4.read_len)) == {250, 4000}
    c5 = m.generate(5, npairs=2000)
    assert (c5.read_len == 150).all() and c5.ctg_len.max() <= 1024


def test_from_strings_roundtrip():
    b = m.from_strings(["ACGT", "", "NNA"], ["AACGT", "G", ""])
    assert b.npairs == 3 and b.read(0) == "ACGT" and b.ctg(0) == "AACGT" and b.read(2) == "NNA" and b.ctg(2) == ""


def test_host_worker_pool(tmp_path):
    """The host worker pool used for packing and staging copies runs every index
    of every region exactly once, with regions issued from several threads."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = tmp_path / "thread_pool_test"
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", os.path.join(root, "tests", "cpp", "thread_pool_test.cpp"),
                    "-o", str(exe)], check=True)
    p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
