"""libmsgs_hip.so loads without a GPU and exports every symbol include/msgs_hip.h declares."""
import os
import re

import numpy as np

import msgseqtk_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_symbols_exported():
    hdr = open(os.path.join(ROOT, "include", "msgs_hip.h")).read()
    declared = set(re.findall(r"\b(msgs_[a-z_0-9]+)\s*\(", hdr))
    assert declared, "no declarations found"
    L = M.lib()
    for name in sorted(declared):
        assert hasattr(L, name), name
    assert declared == set(M.EXPORTS)


def test_params_layout_matches_oracle():
    import ctypes
    import oracle_lib as O
    assert ctypes.sizeof(M.Params) == ctypes.sizeof(O.Params)
    for (a, ta), (b, tb) in zip(M.Params._fields_, O.Params._fields_):
        assert a == b and getattr(M.Params, a).offset == getattr(O.Params, b).offset


def test_defaults():
    p = M.default_params()
    assert (p.minLen, p.maxEvalue, p.maxNSeed, p.mode, p.maxReport) == (17, 0.01, 500, 0, 1)
    assert (p.match, p.mismatch, p.gapO, p.gapE, p.clip, p.minScoreRate) == (1, 3, 5, 2, 4, 0.65)
    assert (p.insMean, p.insSD, p.insMin, p.insMax) == (200, 20, 0, 400)


def test_no_gpu_context_fails_loudly():
    import torch
    if torch.cuda.is_available():
        return
    try:
        M.Mapper(0)
    except RuntimeError as e:
        assert "no usable GPU" in str(e)
    else:
        raise AssertionError("context without a GPU")
