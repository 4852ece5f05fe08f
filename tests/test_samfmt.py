"""The SAM writer's "%g" float formatter (XH:f / XP:f tags: htslib prints the float
widened to double with %g, src/SAMtools/BAM.h:618-620) against Python's %-formatting,
which rounds the exact binary value to six significant digits like C's printf.  CPU only:
the library is loaded but no device call is made."""
import ctypes as C

import numpy as np

import msgseqtk_amd as M


def _fmt(vals):
    v = np.ascontiguousarray(vals, np.float32)
    out = C.create_string_buffer(16 * max(1, len(v)))
    M.lib().msgs_util_format_g6(v.ctypes.data_as(C.POINTER(C.c_float)), C.c_int64(len(v)), out)
    raw = out.raw
    return [raw[16 * i:16 * i + 16].split(b"\0")[0].decode() for i in range(len(v))]


def _check(vals):
    vals = np.asarray(vals, np.float32)
    got = _fmt(vals)
    for x, g in zip(vals, got):
        assert g == "%g" % float(x), (float(x), g)


def test_g6_special_values_and_ties():
    _check([0.0, -0.0, 1.0, -1.0, 0.5, 1e-4, 9.99999e-5, 1e-5, 123456.0, 999999.0, 999999.5, 1e6, 1e7,
            123456.5, 123457.5, 12345.25, 12345.75, 0.1, 0.2, 1.0 / 3, -2.5, 100000.5, 99999.95,
            np.finfo(np.float32).max, np.finfo(np.float32).tiny, 1e-38, 1e-45, np.inf, -np.inf])


def test_g6_random_floats():
    rng = np.random.default_rng(5)
    # log10 likelihoods and posteriors as the tags carry them, then every exponent
    a = -rng.uniform(0, 60, 200000)
    b = rng.uniform(0, 1, 200000)
    c = np.ldexp(rng.uniform(1, 2, 200000), rng.integers(-140, 128, 200000)) * rng.choice([-1, 1], 200000)
    bits = rng.integers(0, 2 ** 32, 200000, dtype=np.uint64).astype(np.uint32).view(np.float32)
    bits = bits[np.isfinite(bits)]
    for v in (a, b, c, bits):
        _check(v.astype(np.float32))
