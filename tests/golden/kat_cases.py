"""The 18 known-answer cases of mhm2's test/ssw-test.cpp:151-238 (TEST(MHMTest, ssw)).

Each case: (query, ref, qstart, qend, rstart, rend) as asserted there with
``Align(query, ref, ...)``: ``query`` is SSW-query (striped), ``ref`` is
SSW-ref (outer).  ``None`` expectations mark the "no match" case, which the
reference checks with a relaxed predicate (span <= 2, score <= 4).
"""

KAT_CASES = [
    ("ACGT", "ACGT", 0, 3, 0, 3),
    ("ACGT", "AACGT", 0, 3, 1, 4),
    ("ACGT", "ACGTT", 0, 3, 0, 3),
    ("TACGT", "ACGT", 1, 4, 0, 3),
    ("TTACGT", "ACGT", 2, 5, 0, 3),
    ("ACGTT", "ACGT", 0, 3, 0, 3),
    ("ACGTTT", "ACGT", 0, 3, 0, 3),
    ("TACGTT", "ACGT", 1, 4, 0, 3),
    ("TTACGTT", "ACGT", 2, 5, 0, 3),
    ("TACGTTT", "ACGT", 1, 4, 0, 3),
    ("TTACGTTT", "ACGT", 2, 5, 0, 3),
    ("AAAATTTTCCCCGGGG", "AAAATTTTCCCCGGGG", 0, 15, 0, 15),
    ("AAAATTTTACCCGGGG", "AAAATTTTCCCCGGGG", 0, 15, 0, 15),   # 1 substitution
    ("AAAATTTTACCCCGGGG", "AAAATTTTCCCCGGGG", 0, 16, 0, 15),  # 1 insertion
    ("AAAATTTCCCCGGGG", "AAAATTTTCCCCGGGG", 0, 14, 0, 15),    # 1 deletion
    ("GCTAGCTAGCTAGCTA", "AAAATTTTCCCCGGGG", None, None, None, None),  # no match
    ("GCTAAAATTTTCCCCGGGG", "AAAATTTTCCCCGGGG", 3, 18, 0, 15),  # soft clip start
    ("AAAATTTTCCCCGGGGACT", "AAAATTTTCCCCGGGG", 0, 15, 0, 15),  # soft clip end
]

# the three aligners of test/ssw-test.cpp:103-112 (match, mismatch, gap open, gap extend, ambiguity)
KAT_SCORINGS = {
    "ssw_default": (2, 2, 3, 1, 2),   # StripedSmithWaterman::Aligner() defaults, ssw.cpp:274-287
    "mhm2_tree": (1, 1, 1, 1, 1),     # ALN_* as set in CMakeDefinitions.txt:113-117
    "cigar": (2, 4, 4, 2, 1),         # cigar_aln_scoring, ssw-test.cpp:103
}
# mhm2's default scoring (CMakeDefinitions.txt:118-122); the reference test does not use it
MHM2_DEFAULT_SCORING = (1, 3, 5, 2, 2)
