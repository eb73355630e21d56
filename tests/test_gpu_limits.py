"""Length limits of the batch entry points: a pair outside them fails the whole batch with MHMSW_E_ARG and a
message naming it, before anything runs on the device (mhmsw.cpp pack_batch); the context stays usable.

SSW itself aligns any lengths (its 16-bit pass, src/ssw/ssw_core.cpp:566-852): these are limits of this
engine's 32-bit form, which keeps one boundary row of the shorter sequence in LDS (8192 bases), and of its
score2 pass (SSW-query sequences of at most 8177 bases).
"""
import numpy as np
import pytest

import mhm2_staging_amd as m
from mhm2_staging_amd import _native

pytestmark = pytest.mark.gpu


def rand(n, seed):
    return "".join(np.array(list("ACGT"))[np.random.default_rng(seed).integers(0, 4, n)])


def expect_arg_error(al, batch, text):
    with pytest.raises(_native.MhmswError) as e:
        al.stage(batch)
    assert e.value.code == _native.MHMSW_E_ARG and text in str(e.value), str(e.value)


def test_rejections_then_context_still_aligns():
    good = m.generate(1, npairs=500)
    with m.Aligner(device=0) as al:
        want = al.align(good)
        # both sequences longer than the 32-bit form's LDS row
        big = rand(8193, 1)
        expect_arg_error(al, m.from_strings(["ACGT", big], ["ACGT", big[::-1]]), "pair 1: both sequences exceed 8192")
        # one sequence of 2^20 bases
        expect_arg_error(al, m.from_strings([rand(1 << 20, 2)], ["ACGTACGT"]), "longer than 2^20")
        # a negative length
        b = m.from_strings(["ACGT", "ACGT"], ["ACGT", "ACGT"])
        b.read_len[1] = -1
        expect_arg_error(al, b, "negative length at pair 1")
        # score2 with an SSW-query (the window, in klign's orientation) past 8177 bases
        al.set_score2(True)
        expect_arg_error(al, m.from_strings(["ACGTACGT"], [rand(8178, 3)]), "score2 needs SSW-query sequences")
        al.set_score2(False)
        # the largest accepted shapes run: 8192 x 8192 and an 8177-base window with score2
        edge = rand(8192, 4)
        r = al.align(m.from_strings([edge], [edge]))
        assert r["score"][0] == 8192 and r["read_end"][0] == 8191
        al.set_score2(True)
        w = rand(8177, 5)
        r = al.align(m.from_strings([w[100:400]], [w]))
        assert r["score"][0] == 300 and r["ctg_begin"][0] == 100
        al.set_score2(False)
        # and the context still aligns a normal batch bit-identically
        assert (al.align(good) == want).all()
