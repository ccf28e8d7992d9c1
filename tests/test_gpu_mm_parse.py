"""The Matrix Market text parsed on the GPU (bspai_hip_mm_parse_gpu) against
the library's host reader (bspai_hip_mm_parse), which the CPU tests pin to
the compiled reference reader (src/cpu.cpp:105-161) and glibc's strtod: the
same entries bit for bit and in file order, or the same message.  The
reference's own accept/reject fixtures run through both parsers in
test_gpu_reference_cases.py."""
import numpy as np
import pytest

from conftest import bits

pytestmark = pytest.mark.gpu


def _parse(text, gpu, ctx):
    from bspai_hip import api

    try:
        return api.parse_matrix_market(text, gpu=gpu, ctx=ctx)
    except api.MatrixMarketError as e:
        return str(e)


def _same(a, b):
    if isinstance(a, str) or isinstance(b, str):
        return a == b
    return (a[0], a[1]) == (b[0], b[1]) and all(
        np.array_equal(np.asarray(x).view(np.uint64), np.asarray(y).view(np.uint64))
        for x, y in zip(a[2:], b[2:]))


def _big_text(n_lines, seed, symmetric=False, exotic=False):
    rng = np.random.default_rng(seed)
    n = 5000
    lines = ["%%MatrixMarket matrix coordinate real " + ("symmetric" if symmetric else "general"),
             "% generated", f"{n} {n} {n_lines}"]
    for i in range(n_lines):
        r = int(rng.integers(1, n + 1))
        c = int(rng.integers(1, r + 1)) if symmetric else int(rng.integers(1, n + 1))
        v = rng.standard_normal() * 10.0 ** int(rng.integers(-5, 6))
        tok = repr(v)
        if exotic and i % 7 == 0:
            tok = rng.choice(["%.25e" % v, "%.3e" % v, "0x1.8p-3", "1e-310", "%.40f" % v,
                              "2.2250738585072012e-308", "+0", "-0.0e-0", "9007199254740993"])
        sep = rng.choice([" ", "\t", "  "]) if exotic else " "
        lines.append(sep.join([str(r), str(c), tok]) + ("\r" if exotic and i % 5 == 0 else ""))
        if i % 97 == 0:
            lines.append("% comment inside")
        if i % 131 == 0:
            lines.append("   \t")
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("symmetric", [False, True])
@pytest.mark.parametrize("exotic", [False, True])
def test_large_text_same_entries(ctx, symmetric, exotic):
    text = _big_text(50000, 5 + symmetric + 2 * exotic, symmetric, exotic)
    assert _same(_parse(text, True, ctx), _parse(text, False, ctx))


def test_first_error_and_counts(ctx):
    """Errors in different places: the first failing data line decides; a
    bad line after the last counted entry is trailing data; a short file is
    truncated (the same messages as the host reader)."""
    text = _big_text(40000, 3)
    lines = text.split("\n")
    n = len(lines)

    def plant(t, k, new):
        ls = t.split("\n")
        ls[k] = new
        return "\n".join(ls)

    cases = [plant(plant(text, n // 5, "1 1 1 1"), (4 * n) // 5, "0 1 1"),
             plant(text, (4 * n) // 5, "1 1 zz"),
             plant(text, 10, "1 1 1e999"),
             plant(text, 11, "3 x 1"),
             plant(text, 12, "99999999999999999999999 1 1"),
             text.rstrip("\n") + "\n1 1 x\n",
             "\n".join(l for l in lines[:-3]) + "\n",
             text.replace("\n", "\r\n", 50)]
    for t in cases:
        got, want = _parse(t, True, ctx), _parse(t, False, ctx)
        assert _same(got, want), (got if isinstance(got, str) else "entries",
                                  want if isinstance(want, str) else "entries")


TOKENS_INDEX = ["1", "2", "3", "0", "4", "01", "+1", "-1", "1.0", "99999999999999999999", "x"]
TOKENS_VALUE = ["1", "-2.5", "0", "-0.0", "1e-3", "1e308", "1e309", "1e-320", "nan", "inf",
                "0x1p-2", "1.5e", "+7", "3.", ".5", "1e+2", "abc", "1_0", "-",
                "2.4703282292062327e-324", "4.9e-324", "0x1.00000000000008p-1023",
                "123456789012345678901234567890e-330", "1.00000000000000011102230246251565404"]


def test_random_small_texts(ctx):
    """Seeded random texts from the rules' corner cases (fields, signs,
    separators, blank and comment lines, counts off by one, exotic values)."""
    rng = np.random.default_rng(77)
    for trial in range(300):
        field = rng.choice(["real", "integer", "real", "complex"])
        sym = rng.choice(["general", "symmetric", "general"])
        rows, cols = int(rng.integers(1, 4)), int(rng.integers(1, 4))
        if sym == "symmetric" and rng.random() < 0.8:
            cols = rows
        count = int(rng.integers(0, 5))
        lines = [f"%%MatrixMarket matrix coordinate {field} {sym}"]
        if rng.random() < 0.3:
            lines.append("% comment")
        lines.append(f"{rows} {cols} {count}" + (" 1" if rng.random() < 0.05 else ""))
        for _ in range(count + int(rng.integers(-1, 2))):
            r = rng.choice(TOKENS_INDEX) if rng.random() < 0.2 else str(rng.integers(1, rows + 1))
            c = rng.choice(TOKENS_INDEX) if rng.random() < 0.2 else str(rng.integers(1, cols + 1))
            v = rng.choice(TOKENS_VALUE)
            sep = rng.choice([" ", "\t", "  ", " \v "])
            extra = " 9" if rng.random() < 0.05 else ""
            lines.append(sep.join([r, c, v]) + extra + ("\r" if rng.random() < 0.2 else ""))
            if rng.random() < 0.15:
                lines.append(rng.choice(["", " ", "%x", "\t% y", "\v"]))
        text = "\n".join(lines) + ("\n" if rng.random() < 0.8 else "")
        got, want = _parse(text, True, ctx), _parse(text, False, ctx)
        assert _same(got, want), (trial, text, got, want)


def _stencil_text(side):
    """A 3-D convection-diffusion matrix as Matrix Market text, values scaled
    off the integers so that every token carries 17 significant digits."""
    from bspai_hip import problems as P

    a = P.convection_diffusion_3d(side)
    rows = np.repeat(np.arange(a.n), np.diff(a.ptr))
    vals = a.val * (1.0 + 1.0 / 3.0)
    cells = np.char.add(np.char.add(np.char.mod("%d ", rows + 1), np.char.mod("%d ", a.col + 1)),
                        np.char.mod("%.17g", vals))
    head = "%%%%MatrixMarket matrix coordinate real general\n%d %d %d\n" % (a.n, a.n, a.nnz)
    return a, rows, vals, (head + "\n".join(cells.tolist()) + "\n").encode()


def test_stencil_text_on_gpu(ctx):
    """The 64^3 system as text (1,810,432 entries): the GPU parse gives the
    generator's entries (%.17g round-trips), and the GPU canonical CSR is the
    generator's CSR."""
    import bspai_hip

    a, rows, vals, text = _stencil_text(64)
    nr, nc, r, c, v = bspai_hip.api.parse_matrix_market(text, gpu=True, ctx=ctx)
    assert (nr, nc) == (a.n, a.n)
    assert np.array_equal(r, rows) and np.array_equal(c, a.col)
    assert np.array_equal(bits(v), bits(vals))
    ptr, col, val = bspai_hip.csr_from_entries(nr, nc, r, c, v, ctx)
    assert np.array_equal(ptr, a.ptr) and np.array_equal(col, a.col)
    assert np.array_equal(bits(val), bits(vals))
