"""The oracle's passthrough vertex shader (test/shader/PassThrough.cu:16-35,
Mat4f * Vec4f of base/Math.hpp:990-1001): pinned by exact known answers that
fix the matrix index order, by an exact-rational restatement of the fused
multiply-add chain, and by an error bound against float64."""
from fractions import Fraction

import numpy as np
import pytest


def _round_f32(q: Fraction) -> float:
    """q rounded to the nearest float32 (ties to even), exactly."""
    if q == 0:
        return 0.0
    sign = -1 if q < 0 else 1
    a = abs(q)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** e > a:
        e -= 1
    e = max(e, -126)                       # subnormals share the minimum exponent
    ulp = Fraction(2) ** (e - 23)
    k, r = divmod(a, ulp)
    k = int(k)
    if r > ulp / 2 or (r == ulp / 2 and k % 2 == 1):
        k += 1
    return sign * float(k * ulp)


def _fma_chain(m, v):
    """rr = 0; rr = fma(m[i][j], v[j], rr) for j = 0..3, exact then rounded."""
    out = []
    for i in range(4):
        rr = 0.0
        for j in range(4):
            rr = _round_f32(Fraction(float(m[i][j])) * Fraction(float(v[j])) + Fraction(rr))
        out.append(rr)
    return out


def test_known_answers_fix_index_order(oracle_lib):
    m = np.array([[1, 2, 3, 4], [5, 6, 7, 8], [9, 10, 11, 12], [13, 14, 15, 16]], np.float32)
    p = np.array([[1, -2, 3], [0, 0, 0], [7, 5, -1]], np.float32)
    got = oracle_lib.vertex_passthrough(p, m)
    exact = np.concatenate([p, np.ones((3, 1), np.float32)], 1) @ m.T
    assert np.array_equal(got, exact)
    assert not np.array_equal(got, np.concatenate([p, np.ones((3, 1), np.float32)], 1) @ m)   # transposed
    ident = oracle_lib.vertex_passthrough(p, np.eye(4, dtype=np.float32))
    assert np.array_equal(ident, np.concatenate([p, np.ones((3, 1), np.float32)], 1))


def test_fused_chain_exact_on_random_inputs(oracle_lib):
    rng = np.random.default_rng(7)
    m = (rng.standard_normal((4, 4)) * np.array([1.0, 1e-3, 30.0, 1e4])).astype(np.float32)
    p = (rng.standard_normal((300, 3)) * 50.0).astype(np.float32)
    got = oracle_lib.vertex_passthrough(p, m)
    for k in range(p.shape[0]):
        v = [p[k, 0], p[k, 1], p[k, 2], np.float32(1.0)]
        assert got[k].tolist() == _fma_chain(m, v), k


def test_error_bound_against_float64(oracle_lib):
    rng = np.random.default_rng(8)
    m = rng.standard_normal((4, 4)).astype(np.float32)
    p = (rng.standard_normal((5000, 3)) * 10.0).astype(np.float32)
    got = oracle_lib.vertex_passthrough(p, m).astype(np.float64)
    v = np.concatenate([p, np.ones((p.shape[0], 1), np.float32)], 1).astype(np.float64)
    ref = v @ m.astype(np.float64).T
    bound = 4 * 2.0 ** -24 * (np.abs(v) @ np.abs(m.astype(np.float64)).T)
    assert np.all(np.abs(got - ref) <= bound)


def test_stride_keeps_the_rest_of_each_record(oracle_lib):
    p = np.array([[1, 2, 3], [4, 5, 6]], np.float32)
    out = np.full((2, 8), 7.5, np.float32)
    got = oracle_lib.vertex_passthrough(p, np.eye(4, dtype=np.float32), out=out, stride_floats=8)
    assert np.array_equal(got[:, :4], [[1, 2, 3, 1], [4, 5, 6, 1]])
    assert np.all(got[:, 4:] == 7.5)


@pytest.mark.parametrize("mutation", ["separate_rounding", "reverse_order"])
def test_mutations_are_caught(oracle_lib, mutation):
    """What the exact restatement distinguishes: rounding the product before
    the add, or summing j = 3..0, gives different bits on these inputs."""
    rng = np.random.default_rng(9)
    m = (rng.standard_normal((4, 4)) * 1e3).astype(np.float32)
    p = rng.standard_normal((200, 3)).astype(np.float32)
    got = oracle_lib.vertex_passthrough(p, m)
    diff = 0
    for k in range(p.shape[0]):
        v = [p[k, 0], p[k, 1], p[k, 2], np.float32(1.0)]
        alt = []
        for i in range(4):
            rr = np.float32(0.0)
            order = range(4) if mutation == "separate_rounding" else range(3, -1, -1)
            for j in order:
                if mutation == "separate_rounding":
                    rr = np.float32(rr + np.float32(m[i][j] * v[j]))
                else:
                    rr = np.float32(_round_f32(Fraction(float(m[i][j])) * Fraction(float(v[j])) + Fraction(float(rr))))
            alt.append(float(rr))
        diff += got[k].tolist() != alt
    assert diff > 0
