"""CPU-side checks of the library's host pieces (no GPU needed).

  * hypot_ref (spai-gmres-ir_amd/csrc/hypot.hpp), the function every SPAI norm
    chain uses on the GPU, equals the host libm's hypot bit for bit;
  * libbspai_hip.so loads and exports every symbol include/bspai_hip.h
    declares, and the ctypes structure layouts equal the C header's;
  * the problem generators produce the documented sizes and canonical CSR.
"""
import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import REPO

HDR = os.path.join(REPO, "include", "bspai_hip.h")
CSRC = os.path.join(REPO, "spai-gmres-ir_amd", "csrc")

HYPOT_PROBE = r"""
#include "hypot.hpp"
#include <cmath>
#include <cstdint>
#include <cstdio>
static uint64_t s = 0x9e3779b97f4a7c15ULL;
static uint64_t rnd() { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static double rd(int lo, int hi) {
    double m = 1.0 + (double)(rnd() >> 12) / 4503599627370496.0;
    double v = std::ldexp(m, lo + (int)(rnd() % (uint64_t)(hi - lo + 1)));
    return (rnd() & 1) ? -v : v;
}
int main(int argc, char **argv) {
    long n = atol(argv[1]), bad = 0;
    for (long i = 0; i < n; ++i) {
        double x = rd(-40, 12), y;
        switch (i % 6) {
        case 0: y = rd(-40, 12); break;
        case 1: y = x * (1 + std::ldexp((double)(rnd() % 4096), -42)); break;
        case 2: y = rd(-80, -30); break;
        case 3: y = rd(-1074, -1000); break;
        case 4: x = rd(500, 1023); y = rd(400, 1023); break;
        default: y = 0.0; break;
        }
        if (bspai_hip::hypot_ref(x, y) != std::hypot(x, y)) ++bad;
    }
    // chains as the SPAI norms use them
    double acc = 0, ref = 0;
    for (int i = 0; i < 100000; ++i) {
        double v = rd(-20, 2);
        acc = bspai_hip::hypot_ref(acc, v);
        ref = std::hypot(ref, v);
        if (acc != ref) { ++bad; acc = ref; }
    }
    std::printf("%ld\n", bad);
    return 0;
}
"""


def test_hypot_matches_host_libm():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "probe.cpp")
        exe = os.path.join(d, "probe")
        with open(src, "w") as f:
            f.write(HYPOT_PROBE)
        subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", CSRC, src, "-o",
                        exe], check=True)
        out = subprocess.run([exe, "3000000"], check=True, capture_output=True, text=True).stdout
    assert int(out.strip()) == 0


def declared_symbols():
    text = open(HDR).read()
    return sorted(set(re.findall(r"\b(bspai_hip_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import bspai_hip

    L = bspai_hip.load()
    names = declared_symbols()
    assert len(names) >= 25
    for name in names:
        assert hasattr(L, name), name
    from bspai_hip._lib import SIGNATURES

    assert sorted(SIGNATURES) == names


def test_struct_layouts_match_header():
    from bspai_hip import _lib

    prog = r"""
#include "bspai_hip.h"
#include <stdio.h>
#include <stddef.h>
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(bspai_hip_spai_options), sizeof(bspai_hip_spai_info),
         sizeof(bspai_hip_gmres_options), sizeof(bspai_hip_gmres_info),
         sizeof(bspai_hip_ir_options), sizeof(bspai_hip_ir_info));
  printf("%zu %zu %zu\n", offsetof(bspai_hip_ir_info, history), offsetof(bspai_hip_ir_info, residual_inf),
         offsetof(bspai_hip_ir_info, message));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "s.c")
        exe = os.path.join(d, "s")
        with open(src, "w") as f:
            f.write(prog)
        subprocess.run(["gcc", "-I", os.path.dirname(HDR), src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    sizes = [int(x) for x in out]
    assert sizes[:6] == [C.sizeof(_lib.SpaiOptions), C.sizeof(_lib.SpaiInfo),
                         C.sizeof(_lib.GmresOptions), C.sizeof(_lib.GmresInfo),
                         C.sizeof(_lib.IrOptions), C.sizeof(_lib.IrInfo)]
    assert sizes[6:] == [_lib.IrInfo.history.offset, _lib.IrInfo.residual_inf.offset,
                         _lib.IrInfo.message.offset]


def test_context_without_gpu_fails_loudly():
    """No device here: creating a context must raise, never fall back."""
    import bspai_hip
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(Exception):
        bspai_hip.Context(0)


def _canonical(a):
    assert a.ptr[0] == 0 and a.ptr[-1] == a.col.size == a.val.size
    assert np.all(np.diff(a.ptr) >= 0)
    for r in range(0, a.n, max(1, a.n // 257)):
        c = a.col[a.ptr[r]:a.ptr[r + 1]]
        assert np.all(np.diff(c) > 0)
    assert np.all(a.val != 0)


def test_generators_sizes():
    from bspai_hip import problems as P

    c1 = P.convection_diffusion_2d()
    assert (c1.n, c1.nnz) == (4096, 20224)
    _canonical(c1)
    c2 = P.convection_diffusion_3d()
    assert (c2.n, c2.nnz) == (2097152, 14581760)
    _canonical(c2)
    # interior row of the 3-D stencil: -6 (upwind x3), 21, -1 (downwind x3)
    r = (64 * 128 + 64) * 128 + 64
    assert list(c2.val[c2.ptr[r]:c2.ptr[r + 1]]) == [-6, -6, -6, 21, -1, -1, -1]
    b = P.band(64, 5.0)
    _canonical(b)
    m, theta, xs, rhs = P.diffusion_edges_2d(side=32)
    A = m.pattern.to_scipy()
    assert abs(A - A.T).max() == 0
    assert np.allclose(m.values(theta), m.pattern.val)
    assert abs(np.linalg.norm(rhs) - 1) < 1e-14


def test_power_law_statistics():
    from bspai_hip import problems as P

    a = P.power_law(n=20000)
    _canonical(a)
    lens = np.diff(a.ptr) - 1
    assert lens.min() >= 3 and lens.max() <= 200
    assert 10 < lens.mean() < 16  # truncated power law 2.2 on [4,200] has mean ~13


def test_matrix_market_round_trip():
    """The writer's output parses back exactly through the reference reader
    (oracle/_ref/cpu_fp64_ir), when that build is present."""
    from bspai_hip import problems as P

    exe = os.path.join(REPO, "oracle", "_ref", "cpu_fp64_ir")
    if not os.path.exists(exe):
        pytest.skip("reference build absent")
    a = P.band(16, 8.0)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "b.mtx")
        P.write_matrix_market(path, a)
        r = subprocess.run([exe, path], capture_output=True, text=True)
    assert r.returncode == 0 and "n=16 nnz=46" in r.stdout
