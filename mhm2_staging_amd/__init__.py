"""MI355X-native batched Smith-Waterman for mhm2's klign read-to-contig stage.

The compute path is libmhmsw.so (hand-written HIP for gfx950 behind the C ABI
in include/mhmsw.h); this package only binds it and mirrors the reference's
adept_sw::GPUDriver interface.
"""
from ._native import MODE_FAST, MODE_PLAIN, RESULT_DTYPE, MhmswError, lib
from .aligner import (MHM2_SCORING, MHM2_TREE_SCORING, Aligner, AlignmentResults, GPUDriver, ResidentContigs,
                      Server, cigar_string, get_avail_gpu_mem_per_rank, get_num_node_gpus, get_tot_gpu_mem)
from .synth import ContigSet, PairBatch, WindowBatch, contigs, from_strings, generate, window_pairs

__all__ = ["MODE_FAST", "MODE_PLAIN", "RESULT_DTYPE", "MhmswError", "lib", "MHM2_SCORING", "MHM2_TREE_SCORING",
           "Aligner", "AlignmentResults", "GPUDriver", "get_avail_gpu_mem_per_rank", "get_num_node_gpus",
           "get_tot_gpu_mem", "PairBatch", "from_strings", "generate", "ResidentContigs", "Server", "ContigSet",
           "WindowBatch", "contigs", "window_pairs", "cigar_string"]
