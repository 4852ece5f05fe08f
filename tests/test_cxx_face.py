"""The C++ drop-in face (include/msgseqtk/, libmsgs_cxx.so) and its test program
tests/cxx/msgseqtk_api_test.cpp: the reference's SMEM_test / SeedChain_test / Alignment_test
cases and invariants, and the PE driver's pairing (AlignmentPE, PairingScheme) checked against
the device pairing entry, through EGriceLab::MSGseqTK classes backed by the device."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "msgseqtk_amd", "libmsgs_cxx.so")
PROG = os.path.join(ROOT, "tests", "cxx", "msgseqtk_api_test")


def _defined_symbols():
    out = subprocess.run(["nm", "-D", "-C", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    return out


def test_cxx_library_exports_the_face():
    """Every out-of-line member the header declares is defined in the library."""
    syms = _defined_symbols()
    for name in ["EGriceLab::MSGseqTK::FMDIndex::FMDIndex(", "EGriceLab::MSGseqTK::FMDIndex::load(",
                 "EGriceLab::MSGseqTK::FMDIndex::accessSA(", "EGriceLab::MSGseqTK::FMDIndex::locateAll(",
                 "EGriceLab::MSGseqTK::FMDIndex::backExt(", "EGriceLab::MSGseqTK::FMDIndex::fwdExt(",
                 "EGriceLab::MSGseqTK::FMDIndex::count(", "EGriceLab::MSGseqTK::MetaGenome::load(",
                 "EGriceLab::MSGseqTK::MetaGenome::loadSeq(", "EGriceLab::MSGseqTK::MetaGenome::getBDSeq(",
                 "EGriceLab::MSGseqTK::SMEM_LIST::findAllSMEMS(", "EGriceLab::MSGseqTK::SMEM_LIST::findMEMS(",
                 "EGriceLab::MSGseqTK::SMEM_LIST::findSeeds(", "EGriceLab::MSGseqTK::SMEM::getSeeds(",
                 "EGriceLab::MSGseqTK::SeedChain::getChains(", "EGriceLab::MSGseqTK::Alignment::buildAlignments(",
                 "EGriceLab::MSGseqTK::Alignment::calcMapQ(", "EGriceLab::MSGseqTK::Alignment::sort(",
                 "EGriceLab::MSGseqTK::Alignment::filter(", "EGriceLab::MSGseqTK::Alignment::mdTag2alnQLen(",
                 "EGriceLab::MSGseqTK::dna::revcom(", "EGriceLab::MSGseqTK::PrimarySeq::revcom(",
                 "EGriceLab::MSGseqTK::AlignmentPE::getPairs(", "EGriceLab::MSGseqTK::AlignmentPE::filter(",
                 "EGriceLab::MSGseqTK::AlignmentPE::calcMapQ(", "EGriceLab::MSGseqTK::AlignmentPE::sort(",
                 "EGriceLab::MSGseqTK::AlignmentPE::getInsertSize(", "EGriceLab::MSGseqTK::PairingScheme::pr("]:
        assert re.search(re.escape(name[:-1]) + r"(\[abi:cxx11\])?\(", syms), name


def test_cxx_header_namespace():
    hdr = open(os.path.join(ROOT, "include", "msgseqtk", "msgseqtk_api.h")).read()
    assert "namespace EGriceLab" in hdr and "namespace MSGseqTK" in hdr
    for h in ["FMDIndex.h", "SMEM.h", "SeedChain.h", "Alignment.h", "MetaGenome.h", "PrimarySeq.h", "PairingScheme.h"]:
        assert os.path.exists(os.path.join(ROOT, "include", "msgseqtk", h))


def test_cxx_program_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    r = subprocess.run([PROG], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "no usable GPU" in r.stderr


@pytest.mark.gpu
def test_cxx_program_on_gpu():
    print("msgseqtk_api_test", flush=True)
    r = subprocess.run([PROG], capture_output=True, text=True, timeout=110)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0 and r.stdout.rstrip().endswith("PASS")
    assert re.search(r"Alignment_test 5: \d+ alignments, best 11M1D10M MD:Z:21", r.stdout)
    assert re.search(r"pairing: 4 candidates, 2 in range, insert size 260, mapQ 3", r.stdout)
