"""The C-ABI library loads on the CPU and exports every symbol the headers declare."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "mhm2_staging_amd", "libmhmsw.so")


def declared_c_functions(headers=("mhmsw.h", "mhmsw_synth.h"), prefix="mhmsw_"):
    names = []
    for h in headers:
        with open(os.path.join(ROOT, "include", h)) as f:
            txt = f.read()
        names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(" + prefix + r"\w+)\s*\(", txt, re.M)
    return sorted(set(names))


def exported():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    return set(line.split()[-1] for line in out.splitlines() if line.strip())


def test_library_built():
    assert os.path.exists(LIB), "libmhmsw.so not built: run __graft_entry__.build()"


def test_exports_every_declared_c_symbol():
    decl = declared_c_functions()
    assert len(decl) >= 18
    missing = [d for d in decl if d not in exported()]
    assert not missing, missing


def test_exports_every_local_assembly_symbol():
    decl = declared_c_functions(("mhmla.h",), "mhmla_")
    assert len(decl) == 11
    missing = [d for d in decl if d not in exported()]
    assert not missing, missing
    from mhm2_staging_amd import locassm
    assert sorted(locassm.MHMLA_EXPORTS) == decl
    L = locassm.lib()
    p = locassm.params()
    assert (p.kmer_len, p.min_kmer_len, p.max_kmer_len, p.shift, p.min_qual, p.min_hi_qual, p.max_reads) == \
        (33, 21, 121, 8, 10, 20, 3000)
    assert L.mhmla_device_count() >= 0


def test_exports_adept_sw_driver_symbols():
    out = subprocess.run(["nm", "-DC", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    for sym in ["adept_sw::GPUDriver::init(int, int, short, short, short, short, int)",
                "adept_sw::GPUDriver::run_kernel_forwards(",
                "adept_sw::GPUDriver::run_kernel_backwards(",
                "adept_sw::GPUDriver::kernel_is_done()",
                "adept_sw::GPUDriver::kernel_block()",
                "adept_sw::GPUDriver::~GPUDriver()",
                "adept_sw::get_tot_gpu_mem()",
                "adept_sw::get_avail_gpu_mem_per_rank(int, int)",
                "adept_sw::get_device_name[abi:cxx11](int)",
                "adept_sw::get_num_node_gpus()",
                "adept_sw::initialize_gpu()",
                "adept_sw::initialize_gpu(double&, int&, unsigned long&)"]:
        assert sym in out, sym


def test_exports_locassm_driver_symbols():
    """The local assembly drop-in (include/gpu_loc_assem/driver.hpp) is in the library."""
    out = subprocess.run(["nm", "-DC", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    for sym in ["locassm_driver::local_assem_driver(std::vector<CtgWithReads, std::allocator<CtgWithReads> >&, "
                "unsigned int, unsigned int, unsigned int, unsigned int, int, int, locassm_driver::accum_data&, "
                "int, int, int, int, int)",
                "locassm_driver::get_gpu_per_node()",
                "locassm_driver::get_device_mem(int, int)",
                "locassm_driver::ctg_bucket::clear()"]:
        assert sym in out, sym


def test_python_binding_lists_all_exports():
    from mhm2_staging_amd import _native
    assert sorted(_native.EXPORTS) == declared_c_functions()
    L = _native.lib()
    for name in _native.EXPORTS:
        assert hasattr(L, name)


def test_no_device_paths_fail_cleanly_on_cpu():
    import ctypes
    import mhm2_staging_amd as m
    if m.get_num_node_gpus() > 0:
        pytest.skip("a GPU is visible here")
    L = m.lib()
    ctx = ctypes.c_void_p()
    assert L.mhmsw_create(0, ctypes.byref(ctx)) == -4  # MHMSW_E_NODEVICE
    with pytest.raises(m.MhmswError):
        m.Aligner()
    from mhm2_staging_amd import locassm
    assert L.mhmla_create(0, ctypes.byref(ctx)) == -4  # MHMLA_E_NODEVICE
    with pytest.raises(locassm.LocAssmError):
        locassm.LocalAssembler()
    drv = m.GPUDriver()
    with pytest.raises(RuntimeError):
        drv.init(0, 1, 1, -3, -5, -2, 300)


def test_drop_in_header_compiles_with_a_klign_style_caller(tmp_path):
    """A caller written against the reference header's API compiles and links against libmhmsw.so."""
    src = tmp_path / "caller.cpp"
    src.write_text(r'''
#include "adept_sw/driver.hpp"
#include <cstdio>
int main() {
  adept_sw::GPUDriver drv;
  if (adept_sw::get_num_node_gpus() == 0) { std::puts("nogpu"); return 0; }
  drv.init(0, 1, (short)1, (short)-3, (short)-5, (short)-2, 300);
  std::vector<std::string> reads{"ACGT"}, ctgs{"AACGT"};
  drv.run_kernel_forwards(reads, ctgs, 4, 5);
  drv.kernel_block();
  drv.run_kernel_backwards(reads, ctgs, 4, 5);
  drv.kernel_block();
  adept_sw::AlignmentResults &r = drv.get_aln_results();
  std::printf("%d %d %d %d %d\n", r.top_scores[0], r.query_begin[0], r.query_end[0], r.ref_begin[0], r.ref_end[0]);
  return 0;
}
''')
    exe = tmp_path / "caller"
    subprocess.run(["g++", "-std=c++14", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe), LIB,
                    f"-Wl,-rpath,{os.path.dirname(LIB)}"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    assert out.strip() in ("nogpu", "4 0 3 1 4")


def test_serving_entry_points_fail_cleanly_on_cpu():
    import ctypes
    import mhm2_staging_amd as m
    if m.get_num_node_gpus() > 0:
        pytest.skip("a GPU is visible here")
    L = m.lib()
    h = ctypes.c_void_p()
    buf = (ctypes.c_char * 8)(*b"ACGTACGT")
    off = (ctypes.c_int64 * 1)(0)
    ln = (ctypes.c_int32 * 1)(8)
    assert L.mhmsw_contigs_load(0, 1, buf, off, ln, ctypes.byref(h)) == -4
    srv = ctypes.c_void_p()
    assert L.mhmsw_server_create(0, None, None, 1, 0, ctypes.byref(srv)) == -1


def test_host_packer_simd_matches_scalar():
    """The SSSE3 and AVX-512 2-bit packers used for staging and resident contigs
    write what the per-base table packer writes, on every byte value and every
    length up to 300 (the AVX-512 form where the CPU has it)."""
    exe = os.path.join(ROOT, "tests", "cpp", "pack_test")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/pack_test is not built (run __graft_entry__.build())")
    p = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0 and p.stdout.startswith("ok"), p.stdout + p.stderr
