"""The GPUDriver call protocol (Python mirror and the C++ drop-in) on the GPU."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import mhm2_staging_amd as m
from golden.kat_cases import KAT_CASES
from oracle_lib import oracle_pairs

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_python_gpudriver_protocol_klign_mapping():
    """klign's call sequence and Aln mapping (src/klign.cpp:361-385) give SSW's coordinates."""
    b = m.generate(1, npairs=2000, seed=21)
    reads = [b.read(i) for i in range(b.npairs)]
    ctgs = [b.ctg(i) for i in range(b.npairs)]
    drv = m.GPUDriver()
    drv.init(0, 1, 1, -3, -5, -2, 300, ambiguity=2)
    drv.run_kernel_forwards(reads, ctgs, 150, 350)
    drv.kernel_block()
    drv.run_kernel_backwards(reads, ctgs, 150, 350)
    drv.kernel_block()
    r = drv.get_aln_results()
    o = oracle_pairs(reads, ctgs, (1, 3, 5, 2, 2))
    np.testing.assert_array_equal(r.top_scores, o["score"])
    np.testing.assert_array_equal(r.query_begin, o["ref_begin"])   # read coordinates
    np.testing.assert_array_equal(r.query_end, o["ref_end"])
    np.testing.assert_array_equal(r.ref_begin, o["query_begin"])   # contig coordinates
    np.testing.assert_array_equal(r.ref_end, o["query_end"])


def test_cpp_dropin_reference_adept_test_sequence(tmp_path):
    """The reference's TEST(MHMTest, AdeptSW) sequence (test/ssw-test.cpp:240-431) through the C++ header."""
    exe = os.path.join(ROOT, "tests", "cpp", "adept_kat")
    assert os.path.exists(exe), "tests/cpp/adept_kat not built (run __graft_entry__.build())"
    print("[gpu] running", exe, flush=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    got = [list(map(int, line.split())) for line in out.stdout.strip().splitlines()]
    # the driver test passes reads = the KAT queries and contigs = the KAT refs; with klign's
    # orientation the read is SSW-ref
    o = oracle_pairs([c[0] for c in KAT_CASES], [c[1] for c in KAT_CASES], (1, 1, 1, 1, 1))
    want = [[int(r["score"]), int(r["ref_begin"]), int(r["ref_end"]), int(r["query_begin"]), int(r["query_end"])]
            for r in o]
    assert got == want


def test_cpp_dropin_get_device_name():
    """get_device_name returns the device's PCI bus id, as the reference's does (src/adept-sw/driver.cpp:35-39)."""
    exe = os.path.join(ROOT, "tests", "cpp", "adept_kat")
    out = subprocess.run([exe, "--device-name"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    name = out.stdout.strip()
    assert re.fullmatch(r"[0-9a-fA-F]{4}:[0-9a-fA-F]{2}:[0-9a-fA-F]{2}\.[0-7]", name), name
    buf = ctypes.create_string_buffer(64)
    assert m.lib().mhmsw_device_name(0, buf, 64) == 0 and buf.value.decode() == name


def test_cpp_dropin_short_narrowing_throws():
    """An end coordinate past 32767 does not fit AlignmentResults' short fields: the driver throws a GPU assert
    instead of wrapping (the reference's 1024-base limit never produced one)."""
    exe = os.path.join(ROOT, "tests", "cpp", "adept_kat")
    out = subprocess.run([exe, "--narrowing"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("GPU assert pair 0:") and "short fields" in out.stdout


def test_python_gpudriver_short_narrowing_raises():
    rng = np.random.default_rng(5)
    seq = "".join(np.array(list("ACGT"))[rng.integers(0, 4, 40000)])
    drv = m.GPUDriver()
    drv.init(0, 1, 1, -3, -5, -2, 300)
    drv.run_kernel_forwards([seq], [seq[33000:38000]])
    with pytest.raises(RuntimeError, match="GPU assert pair 0"):
        drv.kernel_block()
