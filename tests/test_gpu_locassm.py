"""Local assembly walks on the GPU (include/mhmla.h) against the oracle
(oracle/locassm_oracle.c, the reference kernel's semantics restated): every
extension byte, every length and every walk-termination count, on the
closed-form cases of tests/test_oracle_locassm.py, generated batches at
several mer lengths and walk limits, and edge cases."""
import numpy as np
import pytest

from mhm2_staging_amd import locassm
from mhm2_staging_amd import locassm_data as ld
from oracle_lib import la_extend_batch, la_params

pytestmark = pytest.mark.gpu

COMP = str.maketrans("ACGTN", "TGCAN")


def rc(s):
    return s.translate(COMP)[::-1]


def rand_seq(rng, n):
    return "".join(rng.choice(list("ACGT"), n))


@pytest.fixture(scope="module")
def assembler():
    assert locassm.lib().mhmla_device_count() > 0, "no GPU visible: GPU tests must run on the GPU machine"
    a = locassm.LocalAssembler(0)
    yield a
    a.close()



def check(assembler, b, **kw):
    print(f"[gpu] locassm nctgs={b.nctgs} nreads={b.nreads} params={kw}", flush=True)
    got = assembler.extend(b, locassm.params(**kw))
    want = la_extend_batch(la_params(**kw), b, with_terms=True)
    ll, le, rl, re, tm = want
    for name, g, w in (("left_len", got.left_len, ll), ("right_len", got.right_len, rl)):
        bad = np.nonzero(g != w)[0]
        assert len(bad) == 0, f"{name} differs at contig {bad[0]}: got {g[bad[0]]}, want {w[bad[0]]}"
    for c in range(b.nctgs):
        assert got.left_str(c) == bytes(le[c, :ll[c]]).decode(), f"left extension of contig {c}"
        assert got.right_str(c) == bytes(re[c, :rl[c]]).decode(), f"right extension of contig {c}"
    bad = np.nonzero((got.terms != tm).any(axis=(1, 2)))[0]
    assert len(bad) == 0, f"walk terminations differ at contig {bad[0]}: got {got.terms[bad[0]]}, want {tm[bad[0]]}"
    return got


def closed_form_batch():
    """The cases of tests/test_oracle_locassm.py as one batch (right sides),
    plus the left-side case."""
    ctgs = []
    rng = np.random.default_rng(12)
    g = rand_seq(rng, 3000)
    c = 1000
    alt = g[:c + 50] + ("A" if g[c + 50] != "A" else "C") + g[c + 51:]
    reads = []
    for s in (c - 140, c - 130, c - 125):
        reads.append((g[s:s + 300], "I" * 300))
        reads.append((alt[s:s + 300], "I" * 300))
    ctgs.append((g[:c], 10.0, [], reads))                       # fork at every mer length
    rng = np.random.default_rng(13)
    unit = rand_seq(rng, 25)
    per = unit * 40
    ctgs.append((rand_seq(rng, 200) + per[:150], 10.0, [], [(per[:300], "I" * 300)]))  # repeat after 25
    rng = np.random.default_rng(11)
    g = rand_seq(rng, 2000)
    ctgs.append((g[500:1500], 10.0, [(rc(g[440:600]), "I" * 160), (rc(g[400:560]), "I" * 160)], []))
    for seed in range(6):
        rng = np.random.default_rng(seed)
        g = rand_seq(rng, 3000)
        reads = []
        for _ in range(int(rng.integers(1, 12))):
            s = int(rng.integers(1200 - 150, 1200 + 400))
            L = int(rng.integers(60, 151))
            reads.append((g[s:s + L], "I" * L))
        ctgs.append((g[:1200], 10.0, [], reads))
    return ld.from_contigs(ctgs)


def test_closed_form_cases(assembler):
    b = closed_form_batch()
    got = check(assembler, b, kmer_len=33, walk_len_limit=300)
    assert got.right_len[0] == 50 and tuple(got.terms[0, 1]) == (0, 12, 0)
    assert got.right_len[1] == 25 and tuple(got.terms[1, 1]) == (0, 0, 12)
    assert got.left_len[2] == 100


@pytest.mark.parametrize("kmer_len,walk,seed", [(21, 300, 1), (33, 360, 2), (55, 360, 3), (77, 120, 4),
                                                (99, 500, 5), (121, 360, 6)])
def test_generated_batches(assembler, kmer_len, walk, seed):
    b = ld.generate(1500, seed=seed)
    got = check(assembler, b, kmer_len=kmer_len, walk_len_limit=walk)
    assert (got.right_len > 0).mean() > 0.2


def test_other_rules_and_qualities(assembler):
    b = ld.generate(800, seed=9, qual_offset=64)
    check(assembler, b, kmer_len=33, walk_len_limit=300, qual_offset=64)
    check(assembler, b, kmer_len=41, walk_len_limit=200, qual_offset=64, min_qual=15, min_hi_qual=30,
          min_viable_depth=0.1, min_expected_depth=0.5, rating_thres=2, shift=4, min_kmer_len=17, max_kmer_len=65)


def test_edge_cases(assembler):
    rng = np.random.default_rng(21)
    g = rand_seq(rng, 2000)
    c = 600
    ctgs = [
        ("", 5.0, [], []),                                            # empty contig
        (g[:20], 5.0, [(g[:100], "I" * 100)], [(g[:100], "I" * 100)]),  # shorter than the smallest mer
        (g[:c], 5.0, [], []),                                         # no reads
        (g[:c], 5.0, [], [("ACGT", "IIII"), ("", "")]),               # reads shorter than any mer
        (g[:c - 10] + "N" + g[c - 9:c], 7.9, [], [(g[c - 60:c - 10] + "N" + g[c - 9:c + 90], "I" * 150)]),
        (g[:c - 5] + "R" + g[c - 4:c], 7.0, [(rc(g[:60]), "I" * 60)], []),  # IUPAC: N on the left side
        (g[:c], 30.0, [], [(g[c - 50:c + 100].lower(), "I" * 150)]),  # lowercase bases: no votes
        (g[:c], 2.0, [], [(g[c - 40:c + 10], "I" * 50)] * 3000 + [(g[c - 40:c + 110], "I" * 150)]),  # 3001 reads
        (g[:c], 20.0, [], [(g[c - 200:c + 300], "I" * 500)]),         # 500-base read
    ]
    # quality bytes below the offset (negative quality) and mixed qualities
    q = "".join(chr(33 + int(x)) for x in rng.integers(0, 41, 150))
    ctgs.append((g[:c], 6.0, [], [(g[c - 50:c + 100], q), (g[c - 50:c + 100], "!" * 150)]))
    b = ld.from_contigs(ctgs)
    for kw in (dict(kmer_len=21, walk_len_limit=300), dict(kmer_len=33, walk_len_limit=0),
               dict(kmer_len=33, walk_len_limit=1), dict(kmer_len=25, walk_len_limit=50, max_reads=2)):
        check(assembler, b, **kw)


def test_tables_resized_when_mers_are_distinct(assembler):
    """Sides whose reads share almost no mers overflow a table sized for
    overlapping reads: the count is redone at the allocated size, and the
    walks are the oracle's."""
    rng = np.random.default_rng(55)
    ctgs = []
    for _ in range(80):
        g = rand_seq(rng, 1400)
        c = 600
        true_reads = [(g[c - 100 + 13 * j:c + 50 + 13 * j], "I" * 150) for j in range(int(rng.integers(1, 8)))]
        junk = [(rand_seq(rng, 150), "I" * 150) for _ in range(int(rng.integers(3, 70)))]
        right = true_reads + junk
        rng.shuffle(right)
        left = [(rand_seq(rng, 120), "I" * 120) for _ in range(int(rng.integers(0, 30)))]
        ctgs.append((g[:c], float(rng.uniform(2, 30)), left, right))
    b = ld.from_contigs(ctgs)
    for k in (33, 21):
        check(assembler, b, kmer_len=k, walk_len_limit=300)
        assert assembler.stats()["table_resizes"] > 0


def test_restaged_and_rerun(assembler):
    """A staged batch runs again with other parameters; a new batch replaces it."""
    b1 = ld.generate(600, seed=31)
    b2 = ld.generate(300, seed=32, ctg_len=(60, 400))
    assembler.stage(b1)
    for k in (33, 21):
        assembler.run(locassm.params(kmer_len=k, walk_len_limit=300))
        assembler.wait()
        got = assembler.results()
        ll, le, rl, re = la_extend_batch(la_params(kmer_len=k, walk_len_limit=300), b1)
        assert np.array_equal(got.left_len, ll) and np.array_equal(got.right_len, rl)
    check(assembler, b2, kmer_len=33, walk_len_limit=300)
    st = assembler.stats()
    assert st["contigs"] == 300 and st["walks"] > 0 and st["mers"] > 0


def test_errors(assembler):
    b = ld.from_contigs([("ACGTXACGT", 1.0, [], [])])
    with pytest.raises(locassm.LocAssmError):
        assembler.stage(b)
    good = ld.from_contigs([("ACGT" * 10, 1.0, [], [])])
    assembler.stage(good)
    with pytest.raises(locassm.LocAssmError):
        assembler.run(locassm.params(shift=0))


def driver_input(b, mer_len, walk, qual_offset=33):
    lines = [f"{b.nctgs} {mer_len} {walk} {qual_offset}"]
    for c in range(b.nctgs):
        left, right = b.reads_of(c, "left"), b.reads_of(c, "right")
        lines.append(f"{int(b.cid[c])} {float(b.depth[c])!r} {b.ctg(c) or '-'} {len(left)} {len(right)}")
        lines += [f"{r or '-'} {q or '-'}" for r, q in left + right]
    return "\n".join(lines) + "\n"


def test_reference_driver_interface(assembler):
    """locassm_driver::local_assem_driver, called through the reference's
    header by a C++ program (tests/cpp/locassm_driver_test.cpp), extends every
    contig exactly as the oracle's walks say (left + contig + right)."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "locassm_driver_test")
    assert os.path.exists(exe), "tests/cpp/locassm_driver_test not built (run __graft_entry__.build())"
    b = ld.generate(400, seed=41)
    for mer_len, walk in ((33, 300), (21, 150)):
        out = subprocess.run([exe], input=driver_input(b, mer_len, walk), capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr
        ll, le, rl, re = la_extend_batch(la_params(kmer_len=mer_len, walk_len_limit=walk), b)
        got = out.stdout.split("\n")
        for c in range(b.nctgs):
            want = bytes(le[c, :ll[c]]).decode() + b.ctg(c) + bytes(re[c, :rl[c]]).decode()
            cid, seq = got[c].split(" ")
            assert int(cid) == int(b.cid[c]) and seq == (want or "-"), f"contig {c}"


@pytest.mark.parametrize("seed", range(4))
def test_randomized_shapes(assembler, seed):
    """Generated batches with random read lengths, contig lengths, error and
    fork/repeat rates, and random parameters, against the oracle."""
    rng = np.random.default_rng(100 + seed)
    b = ld.generate(int(rng.integers(200, 900)), seed=200 + seed,
                    ctg_len=(int(rng.integers(20, 200)), int(rng.integers(300, 3000))),
                    read_len=int(rng.choice([60, 100, 150, 250])), sub_rate=float(rng.uniform(0, 0.05)),
                    n_rate=float(rng.uniform(0, 0.02)), fork_rate=float(rng.uniform(0, 0.5)),
                    repeat_rate=float(rng.uniform(0, 0.3)))
    kw = dict(kmer_len=int(rng.integers(15, 80)), walk_len_limit=int(rng.integers(1, 600)),
              shift=int(rng.integers(1, 12)), min_kmer_len=int(rng.integers(9, 25)),
              max_kmer_len=int(rng.integers(60, 140)), max_reads=int(rng.choice([5, 40, 3000])))
    check(assembler, b, **kw)
