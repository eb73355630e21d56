"""klign's batching loop (tests/cpp/mini_klign.cpp) over the drop-in driver.

The harness replays mhm2's klign path for seeded reads: the window rule of
compute_alns_for_read (src/klign.cpp:672-739), the perfect-match shortcut and
dispatch rule of align_read (:218-249), the CPU work-steal of
kernel_align_block (:552-606) and the Aln mapping of both sides (:251-273 CPU,
:356-388 GPU).  The expected records come from the same harness with every
pair aligned by the reference SSW on the CPU (klign without a GPU).

Whichever side aligns a pair, the Aln must be the same in every field the GPU
path fills (coordinates, score1, identity); score2 and mismatches are the
values klign's GPU path writes (0) for GPU-aligned pairs and SSW's for the
others, as in klign.
"""
import json
import os
import subprocess

import pytest

from oracle_lib import have_ref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "tests", "cpp", "mini_klign")
SSW_LIB = os.path.join(ROOT, "oracle", "_ref", "libssw_ref.so")
ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")

# Aln::to_string columns, then identity, mismatches, side
COORD_COLS = (0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 11)  # all but score2 (10) and mismatches (12)


def _need_harness():
    if not os.path.exists(HARNESS):
        pytest.fail("tests/cpp/mini_klign is not built (run __graft_entry__.build())")


def _cpu_lib_args():
    if have_ref():
        return ["--ssw", SSW_LIB]
    return ["--oracle", ORACLE_LIB]


def run_harness(tmp_path, name, *args):
    out = str(tmp_path / (name + ".tsv"))
    cmd = [HARNESS, "--out", out] + _cpu_lib_args() + [str(a) for a in args]
    print("[mini_klign]", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert p.returncode == 0, p.stderr
    summary = json.loads(p.stdout.strip().splitlines()[-1])
    with open(out) as f:
        rows = [line.rstrip("\n").split("\t") for line in f]
    return summary, rows


def compare(expected, got, gpu_score2=False):
    assert len(expected) == len(got)
    bad = []
    for e, g in zip(expected, got):
        if [e[c] for c in COORD_COLS] != [g[c] for c in COORD_COLS]:
            bad.append((e, g))
            continue
        if g[13] == "G" and gpu_score2:
            if (g[10], g[12]) != (e[10], e[12]):   # every field as the CPU path writes it
                bad.append((e, g))
        elif g[13] == "G":
            if g[10] != "0" or g[12] != "0":
                bad.append((e, g))
        elif e[10] != "-1" and g[10] != "-1" and (e[10], e[12]) != (g[10], g[12]):
            bad.append((e, g))
    assert not bad, "%d records differ, first: %r" % (len(bad), bad[:3])


# ---- CPU ----

def test_cpu_harness_ssw_vs_restatement(tmp_path):
    """All-CPU klign with the reference SSW and with the scalar restatement
    give the same coordinates, score1 and identity for every record."""
    _need_harness()
    if not have_ref():
        pytest.skip("reference SSW library not built")
    out_a = str(tmp_path / "a.tsv")
    out_b = str(tmp_path / "b.tsv")
    common = [HARNESS, "--mode", "cpu", "--reads", "6000", "--seed", "5"]
    pa = subprocess.run(common + ["--out", out_a, "--ssw", SSW_LIB], capture_output=True, text=True, timeout=300)
    pb = subprocess.run(common + ["--out", out_b, "--oracle", ORACLE_LIB], capture_output=True, text=True,
                        timeout=300)
    assert pa.returncode == 0 and pb.returncode == 0, pa.stderr + pb.stderr
    ra = [l.split("\t") for l in open(out_a).read().splitlines()]
    rb = [l.split("\t") for l in open(out_b).read().splitlines()]
    assert len(ra) == len(rb) > 5000
    for x, y in zip(ra, rb):
        assert [x[c] for c in COORD_COLS] == [y[c] for c in COORD_COLS], (x, y)
    sa = json.loads(pa.stdout)
    assert sa["perfect"] > 0 and sa["stolen"] + sa["perfect"] == sa["alns"]


def test_cpu_harness_without_gpu_reports_no_gpu(tmp_path):
    _need_harness()
    p = subprocess.run([HARNESS, "--mode", "gpu", "--out", str(tmp_path / "x.tsv"), "--oracle", ORACLE_LIB,
                        "--reads", "10"], capture_output=True, text=True, timeout=120)
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    assert p.returncode == 3 and "no GPU" in p.stderr


# ---- GPU ----

@pytest.mark.gpu
@pytest.mark.parametrize("scoring", ["1,3,5,2,2", "1,1,1,1,1"])
def test_klign_gpu_blocks_match_cpu(tmp_path, scoring):
    """Default block size (KLIGN_GPU_BLOCK_SIZE = 20000) and work-steal on."""
    _need_harness()
    exp_s, exp = run_harness(tmp_path, "cpu", "--mode", "cpu", "--reads", 40000, "--scoring", scoring)
    got_s, got = run_harness(tmp_path, "gpu", "--mode", "gpu", "--reads", 40000, "--scoring", scoring)
    print(json.dumps(exp_s), json.dumps(got_s), flush=True)
    assert got_s["gpu"] > 0 and got_s["blocks"] >= 2
    compare(exp, got)


@pytest.mark.gpu
def test_klign_gpu_work_steal_many_blocks(tmp_path):
    """Small blocks and a slow GPU side: many pairs are stolen by the CPU, the
    rest go to the GPU; records are identical to the all-CPU run."""
    _need_harness()
    exp_s, exp = run_harness(tmp_path, "cpu", "--mode", "cpu", "--reads", 30000, "--seed", 7)
    got_s, got = run_harness(tmp_path, "gpu", "--mode", "gpu", "--reads", 30000, "--seed", 7, "--block", 700,
                             "--gpu-delay-ms", 15)
    print(json.dumps(got_s), flush=True)
    assert got_s["stolen"] > 0 and got_s["gpu"] > 0
    compare(exp, got)


@pytest.mark.gpu
def test_klign_gpu_memory_dispatch_rule(tmp_path):
    """A small memory budget makes align_read dispatch on its memory estimate
    (tot_mem_est >= gpu_mem_avail) long before the block-size limit."""
    _need_harness()
    exp_s, exp = run_harness(tmp_path, "cpu", "--mode", "cpu", "--reads", 20000, "--seed", 9)
    got_s, got = run_harness(tmp_path, "gpu", "--mode", "gpu", "--reads", 20000, "--seed", 9, "--mem", 200000,
                             "--no-steal")
    print(json.dumps(got_s), flush=True)
    assert got_s["stolen"] == 0 and got_s["blocks"] > 20
    compare(exp, got)


@pytest.mark.gpu
@pytest.mark.parametrize("scoring", ["1,3,5,2,2", "1,1,1,1,1"])
def test_klign_gpu_score2_records_identical(tmp_path, scoring):
    """With the driver's score2 on, a GPU-aligned Aln equals the CPU-aligned
    one in every field (SSW reports 0 mismatches without a CIGAR)."""
    _need_harness()
    if not have_ref():
        pytest.skip("needs the reference SSW for the CPU side (its score2)")
    exp_s, exp = run_harness(tmp_path, "cpu", "--mode", "cpu", "--reads", 30000, "--seed", 13, "--scoring", scoring)
    got_s, got = run_harness(tmp_path, "gpu", "--mode", "gpu", "--reads", 30000, "--seed", 13, "--scoring", scoring,
                             "--gpu-score2", "--block", 4000, "--gpu-delay-ms", 5)
    print(json.dumps(got_s), flush=True)
    assert got_s["gpu"] > 0 and got_s["stolen"] > 0
    compare(exp, got, gpu_score2=True)


def test_cpu_harness_report_cigar_sam_lines(tmp_path):
    """klign's CIGAR-reporting run on the CPU: every record carries
    set_sam_string's line; perfect matches report "<n>=" and NM 0; SSW's
    CIGARs span the whole contig window."""
    _need_harness()
    if not have_ref():
        pytest.skip("reference SSW library not built")
    s, rows = run_harness(tmp_path, "cpu", "--mode", "cpu", "--reads", 4000, "--seed", 3, "--report-cigar")
    assert s["perfect"] > 0 and s["stolen"] > 0
    import re
    for r in rows:
        sam = r[14:]  # read, flag, contig, pos, mapq, cigar, * 0 0 * *, AS, NM, RG
        assert len(sam) == 14, r
        assert sam[0] == r[0] and sam[1] == ("0" if r[8] == "Plus" else "16") and sam[2] == r[4] and sam[3] == r[5]
        assert sam[6:11] == ["*", "0", "0", "*", "*"]
        assert sam[11] == f"AS:i:{r[9]}" and sam[12] == f"NM:i:{r[12]}" and sam[13] == "RG:Z:0"
        cig = re.findall(r"(\d+)([MIDNSHP=X])", sam[5])
        assert "".join(n + o for n, o in cig) == sam[5]
        if r[13] == "P":
            assert sam[5] == f"{int(r[6]) - int(r[5]) + 1}=" and r[12] == "0"
        else:
            # SSW's CIGAR walks the whole SSW-query (the contig window) with soft clips
            qlen = sum(int(n) for n, o in cig if o in "SM=XI")
            assert qlen > 0


@pytest.mark.gpu
@pytest.mark.parametrize("scoring", ["1,3,5,2,2", "1,1,1,1,1"])
def test_klign_gpu_cigar_records_identical(tmp_path, scoring):
    """klign's CIGAR-reporting run with blocks on the GPU (the driver's
    CIGARs, mismatches and score2) writes exactly the records and SAM lines
    of the all-CPU run, with work-steal mixing both sides."""
    _need_harness()
    if not have_ref():
        pytest.skip("needs the reference SSW for the CPU side")
    exp_s, exp = run_harness(tmp_path, "cpu", "--mode", "cpu", "--reads", 30000, "--seed", 17, "--scoring", scoring,
                             "--report-cigar")
    got_s, got = run_harness(tmp_path, "gpu", "--mode", "gpu", "--reads", 30000, "--seed", 17, "--scoring", scoring,
                             "--report-cigar", "--gpu-cigar", "--block", 4000, "--gpu-delay-ms", 5)
    print(json.dumps(got_s), flush=True)
    assert got_s["gpu"] > 0 and got_s["stolen"] > 0
    assert len(exp) == len(got)
    bad = [(e, g) for e, g in zip(exp, got) if e[:13] + e[14:] != g[:13] + g[14:]]
    assert not bad, "%d records differ, first: %r" % (len(bad), bad[:2])


@pytest.mark.gpu
def test_klign_report_cigar_keeps_blocks_on_cpu(tmp_path):
    """Without the driver's CIGARs, klign sends CIGAR-reporting blocks to
    ssw_align_block (src/klign.cpp:589-599): no pair is aligned on the GPU."""
    _need_harness()
    exp_s, exp = run_harness(tmp_path, "cpu", "--mode", "cpu", "--reads", 5000, "--seed", 19, "--report-cigar")
    got_s, got = run_harness(tmp_path, "gpu", "--mode", "gpu", "--reads", 5000, "--seed", 19, "--report-cigar")
    assert got_s["gpu"] == 0
    assert [r[:13] + r[14:] for r in exp] == [r[:13] + r[14:] for r in got]
