"""CPU checks of the boundary: the libraries load, export every function the
public headers declare, and the Python mirror refuses to run without them.
No compute call is made (there may be no GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIB = os.path.join(ROOT, "sympiler_amd", "libsympiler_amd.so")
GEN = os.path.join(ROOT, "sympiler_amd", "libsymamd_gen.so")


def declared_c_functions(header):
    text = open(os.path.join(INC, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sym(?:amd|gen)_[a-z0-9_]+)\s*\(", text)))


def exported(path):
    out = subprocess.check_output(["nm", "-D", "--defined-only", path], text=True)
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_library_loads():
    ctypes.CDLL(LIB)
    ctypes.CDLL(GEN)


def test_every_declared_entry_point_is_exported():
    names = declared_c_functions("sympiler_amd.h")
    assert len(names) >= 25
    missing = [n for n in names if n not in exported(LIB)]
    assert not missing, missing
    gnames = declared_c_functions("sympiler_amd_gen.h")
    assert not [n for n in gnames if n not in exported(GEN)]


def test_sym_lib_dropins_exported_with_reference_mangling():
    # mangled names of the reference's declarations in
    # include/sympiler/sparse_blas_lib.h (namespace sym_lib, C++ linkage)
    want = {
        "_ZN7sym_lib10sptrsv_csrEiPiS0_PdS1_",
        "_ZN7sym_lib19sptrsv_csr_levelsetEiPKiS1_PKdPdiS1_S1_",
        "_ZN7sym_lib14sptrsv_csr_lbcEiPiS0_PdS1_iS0_S0_S0_",
        "_ZN7sym_lib10sptrsv_cscEiPiS0_PdS1_",
        "_ZN7sym_lib19sptrsv_csc_levelsetEiPiS0_PdS1_iPKiS3_",
        "_ZN7sym_lib14sptrsv_csc_lbcEiPiS0_PdS1_iS0_S0_S0_",
        "_ZN7sym_lib7ltsolveEiPiS0_PdS1_",
        "_ZN7sym_lib8spmv_csrEiPKiS1_PKdS3_Pd",
        "_ZN7sym_lib8spmv_csrEiPKiS1_PKdS3_S3_Pddd",
        "_ZN7sym_lib17spmv_csr_parallelEiPKiS1_PKdS3_Pd",
        "_ZN7sym_lib14spmv_csr_blockEiPKiS1_PKdiS1_S3_Pd",
        "_ZN7sym_lib8spmv_cscEiPKiS1_PKdS3_Pd",
        "_ZN7sym_lib8spmv_cscEiPKiS1_PKdS3_S3_Pddd",
        "_ZN7sym_lib8spmv_cscEiPKiS1_PKdS3_S3_PdS4_S4_",
        "_ZN7sym_lib12spmv_csc_lbcEiPKiS1_iS1_S1_PKdS3_Pd",
        "_ZN7sym_lib12spmv_csc_lbcEPKiS1_S1_iS1_S1_PKdS3_Pd",
        "_ZN7sym_lib17spmv_csc_parallelEiPKiS1_PKdS3_Pd",
        "_ZN7sym_lib14spmv_csc_blockEiPKiS1_PKdS3_PdiS1_",
        "_ZN7sym_lib18build_levelSet_CSCEiPKiS1_RPiS3_",
        "_ZN7sym_lib24sptrsv_csr_lbc_redundantEiPiS0_PdS1_S1_iS0_S0_S0_S1_",
        "_ZN7sym_lib14sptrsv_csr_iluEiPKiS1_PKdS1_Pd",
        "_ZN7sym_lib18sptrsv_csr_ilu_lbcEiPKiS1_PKdS1_PdiS1_S1_S1_",
        "_ZN7sym_lib9spic0_csrEiPdPiS1_",
        "_ZN7sym_lib13spic0_csr_lbcEiPdPiS1_iS1_S1_S1_",
        "_ZN7sym_lib14spic0_csc_leftEiPKiS1_PKdS1_S1_Pd",
        "_ZN7sym_lib18spic0_csc_left_LBCEiPKiS1_PKdS1_S1_iS1_S1_S1_Pd",
        "_ZN7sym_lib10spilu0_csrEiiPKiS1_PKdS1_Pd",
        "_ZN7sym_lib14spilu0_csr_lbcEiiPKiS1_PKdS1_PdiS1_S1_S1_Pi",
        "_ZN7sym_lib16sptrsv_csc_blockEiPKiS1_S1_PdS2_iS1_",
        "_ZN7sym_lib5parsy19getCoarseLevelSet_6EmPKiS2_RiRPiS5_S3_S5_S5_iiiPd",
        "_ZN7sym_lib32get_coarse_levelSet_DAG_CSC_treeEmPKiS1_iRiRPiS2_S4_S4_iiiPd",
        "_ZN7sym_lib5parsy13blockedLsolveEiPmPiPdiS1_S2_S2_iS3_",
        "_ZN7sym_lib5parsy22H2LeveledBlockedLsolveEiPmPiPdiS1_S2_S2_iS3_iS2_S2_iS2_S2_i",
        "_ZN7sym_lib5parsy18blockedLsolve_mrhsEiPmPiPdiS1_S2_S2_iS3_ii",
        "_ZN7sym_lib5parsy14blockedLTsolveEiPmPiPdiS1_S2_S2_iS3_",
        "_ZN7sym_lib5parsy19blockedLTsolve_mrhsEiPmPiPdiS1_S2_S2_iS3_ii",
        "_ZN7sym_lib5parsy27H2LeveledBlockedLsolve_mrhsEiPmPiPdiS1_S2_S2_iS3_iiS2_S2_iS2_S2_ii",
        "_ZN7sym_lib5parsy23H2LeveledBlockedLTsolveEiPmPiPdiS1_S2_S2_iS3_iS2_S2_iS2_S2_i",
        "_ZN7sym_lib5parsy28H2LeveledBlockedLTsolve_mrhsEiPmPiPdiS1_S2_S2_iS3_iiS2_S2_iS2_S2_ii",
    }
    missing = want - exported(LIB)
    assert not missing, missing


def test_mangled_names_match_the_reference_objects():
    # every sym_lib function of the solve path that the reference's own build
    # defines (when present) is exported under the same mangled name
    import oracle
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built")
    on_path = {"sptrsv_csr", "sptrsv_csr_levelset", "sptrsv_csr_lbc",
               "sptrsv_csr_lbc_redundant", "sptrsv_csr_ilu", "sptrsv_csr_ilu_lbc",
               "sptrsv_csc", "sptrsv_csc_levelset", "sptrsv_csc_lbc", "ltsolve",
               "spmv_csr", "spic0_csc_left", "spic0_csc_left_LBC", "spic0_csr",
               "spic0_csr_lbc", "sptrsv_csc_block", "sptrsv_csc_levelset_block",
               "spmv_csr_parallel", "spmv_csr_block", "spmv_csc", "spmv_csc_lbc",
               "spmv_csc_parallel", "spmv_csc_block"}

    def fname(sym):  # the length-prefixed function name after "sym_lib"
        m = re.match(r"_ZN7sym_lib(\d+)", sym)
        return sym[m.end():m.end() + int(m.group(1))] if m else None

    ref = {s for s in exported(oracle.REF_PATH) if fname(s) in on_path}
    assert len(ref) == 27, sorted(ref)  # spmv_csr: two overloads, spmv_csc: three, _lbc: two
    # the tree-level coarsening of parsy (InspectionLevel_06.cpp:17)
    ref |= {s for s in exported(oracle.REF_PATH)
            if s.startswith("_ZN7sym_lib5parsy19getCoarseLevelSet_6")}
    assert len(ref) == 28, sorted(ref)
    missing = ref - exported(LIB)
    assert not missing, missing


def test_parsy_solves_match_the_reference_declarations():
    # oracle/_ref/parsy_decls.o takes the address of every supernodal solve
    # declared in the reference's parsy/triangularSolve/Triangular_BCSC.h
    # (built from that header by oracle/Makefile): its undefined symbols are
    # the reference's mangled names, all exported by the library
    obj = os.path.join(ROOT, "oracle", "_ref", "parsy_decls.o")
    if not os.path.exists(obj):
        pytest.skip("oracle/_ref/parsy_decls.o not built")
    out = subprocess.run(["nm", "-u", obj], capture_output=True, text=True, check=True).stdout
    ref = {ln.split()[-1] for ln in out.splitlines() if "_ZN7sym_lib5parsy" in ln}
    assert len(ref) == 8, sorted(ref)
    missing = ref - exported(LIB)
    assert not missing, missing


def test_python_mirror_fails_loudly_without_library(monkeypatch, tmp_path):
    import sympiler_amd as sa
    monkeypatch.setattr(sa, "_lib", None)
    monkeypatch.setattr(sa, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        sa.lib()


def test_product_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "sympiler_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "from oracle" not in text, f
                assert "liboracle" not in text and "libsymref" not in text, f
