"""The host e4m3 encoder of the fp8 inference weights (csrc/fp8.h) against
torch's float8_e4m3fn conversion (round to nearest even) with the same
saturation at +-448, over every code, every midpoint between neighbouring
codes and one float either side of it, and random values across the range."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "webgpuraytracer_amd", "csrc")

SRC = r"""
#include "fp8.h"
extern "C" void encode(const float* x, unsigned char* q, long n) { for (long i = 0; i < n; ++i) q[i] = pt::e4m3_encode(x[i]); }
extern "C" void decode(const unsigned char* q, float* x, long n) { for (long i = 0; i < n; ++i) x[i] = pt::e4m3_decode(q[i]); }
"""


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("fp8")
    (d / "fp8.cpp").write_text(SRC)
    so = d / "libfp8.so"
    subprocess.run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
                    "-I" + CSRC, str(d / "fp8.cpp"), "-o", str(so)], check=True)
    return C.CDLL(str(so))


def _inputs():
    codes = torch.arange(256, dtype=torch.int32).to(torch.uint8).view(torch.float8_e4m3fn).float().numpy()
    fin = np.unique(codes[np.isfinite(codes)])
    mids = ((fin[:-1] + fin[1:]) / 2).astype(np.float32)
    rng = np.random.default_rng(5)
    rnd = (rng.standard_normal(200000) * np.exp2(rng.uniform(-14, 11, 200000))).astype(np.float32)
    special = np.array([448, 449, 463.99, 464, 1e6, 3e38, np.inf, -np.inf, 2 ** -9, 2 ** -10, 1.5 * 2 ** -10,
                        2 ** -11, 2 ** -6, np.nextafter(np.float32(2 ** -6), np.float32(0)), 0.0, -0.0], np.float32)
    xs = [fin, mids, np.nextafter(mids, np.float32(np.inf)), np.nextafter(mids, np.float32(-np.inf)), rnd, special]
    x = np.concatenate(xs).astype(np.float32)
    return np.concatenate([x, -x])


def test_encoder_matches_torch_with_saturation(lib):
    x = _inputs()
    q = np.empty(x.size, np.uint8)
    lib.encode(x.ctypes.data_as(C.c_void_p), q.ctypes.data_as(C.c_void_p), C.c_long(x.size))
    want = torch.from_numpy(np.clip(x, -448, 448)).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    bad = np.nonzero(q != want)[0]
    assert bad.size == 0, [(float(x[i]), int(q[i]), int(want[i])) for i in bad[:10]]


def test_decoder_inverts_every_code(lib):
    q = np.arange(256, dtype=np.uint8)
    x = np.empty(256, np.float32)
    lib.decode(q.ctypes.data_as(C.c_void_p), x.ctypes.data_as(C.c_void_p), C.c_long(256))
    want = torch.from_numpy(q.copy()).view(torch.float8_e4m3fn).float().numpy()
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(x), nan) and nan.sum() == 2
    assert np.array_equal(x[~nan], want[~nan])
    back = np.empty(256, np.uint8)
    lib.encode(x.ctypes.data_as(C.c_void_p), back.ctypes.data_as(C.c_void_p), C.c_long(256))
    assert np.array_equal(back[~nan], q[~nan])  # -0 and +0 keep their signs
