"""Pins the CPU oracle before anything is checked against it (CPU only).

Each oracle function is checked against (a) the reference's own known-answer
vectors (tests/golden/reference_goldens.json), (b) the reference's compiled
kernels where the harness was built (tests/golden/ref_vectors.npz, written by
tests/golden/make_ref_vectors.py from oracle/_ref/libref.so), and (c)
closed forms / invariants / exact rational brute force, chosen so that a
dropped term, a wrong sign or index, or a transposed operand fails.
"""
import json
import os
import sys
from fractions import Fraction
from itertools import product

import numpy as np
import pytest
from numpy.testing import assert_allclose, assert_array_equal

from oracle import oracle as O

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_goldens.json")))
REF = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz"))


# --- (a) reference known answers -------------------------------------------

@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_ft_fma_explicit_golden(dtype):
    g = GOLD["ft_fma_explicit"]
    out = np.zeros(7, dtype=dtype)
    O.ft_fma(out, np.array(g["b"], dtype), np.array(g["c"], dtype), 2, 2)
    assert_array_equal(out, np.array(g["expected"], dtype))


def test_ft_fma_construction_levels(rng):
    # tests/compute/test_ft_fma.py:23-49 level by level
    a, b, c = (rng.uniform(-1, 1, size=7) for _ in range(3))
    exp = a.copy()
    exp[0] += b[0] * c[0]
    exp[1:3] += b[0] * c[1:3] + c[0] * b[1:3]
    exp[3:] += b[0] * c[3:] + c[0] * b[3:] + np.outer(b[1:3], c[1:3]).flatten()
    out = a.copy()
    O.ft_fma(out, b, c, 2, 2)
    assert_allclose(out, exp, rtol=0, atol=1e-15)


def test_free_tensor_mul_golden():
    g = GOLD["free_tensor_mul"]
    r = O.ft_mul(np.array(g["t1"]), np.array(g["t2"]), 2, 2)
    assert_array_equal(r, g["expected"])


@pytest.mark.parametrize("name", ["exp_zero", "exp_letter"])
def test_exp_goldens(name):
    g = GOLD[name]
    assert_array_equal(O.ft_exp(np.array(g["x"], float), 2, 2), g["expected"])


def test_log_identity_golden():
    g = GOLD["log_identity"]
    assert_array_equal(O.ft_log(np.array(g["x"], float), 2, 2), g["expected"])


@pytest.mark.parametrize("width", [2, 3, 4, 5])
def test_exp_single_letter_depth3(width):
    # tests/algebra/test_free_tensor.py:367-377
    x = np.zeros(O.tensor_size(width, 3))
    x[1] = 1.0
    exp = np.zeros_like(x)
    exp[0], exp[1], exp[1 + width], exp[1 + width + width ** 2] = 1, 1, 0.5, 1 / 6
    assert_allclose(O.ft_exp(x, width, 3), exp, rtol=0, atol=1e-16)


def test_hall_basis_golden_and_sizes():
    g = GOLD["hall_order_w2_d3"]
    db, data = O.hall_basis(2, 3)
    assert [O.hall_key_str(data, k) for k in range(1, len(data))] == g["keys"]
    for w, d, n in GOLD["lie_basis_sizes"]["cases"]:
        db, data = O.hall_basis(w, d)
        assert len(data) - 1 == n
    # Witt dimensions at the benchmark shapes (level sizes from the formula)
    for (w, d, sizes) in [(3, 4, [3, 3, 8, 18]), (5, 5, [5, 10, 40, 150, 624]),
                          (8, 4, [8, 28, 168, 1008]), (6, 5, [6, 15, 70, 315, 1554]),
                          (4, 6, [4, 6, 20, 60, 204, 670])]:
        db, _ = O.hall_basis(w, d)
        assert list(np.diff(db)[1:]) == sizes
        assert [O.witt_dim(w, k) for k in range(1, d + 1)] == sizes


def test_l2t_golden_matrix():
    g = GOLD["l2t_w2_d3"]
    assert_array_equal(O.l2t_dense(2, 3), np.array(g["matrix"], float))


@pytest.mark.parametrize("case", GOLD["to_dyadic_intervals"]["cases"])
def test_to_dyadic_intervals_goldens(case):
    got = O.to_dyadic_intervals(case["inf"], case["sup"], case["tol"])
    assert got == [tuple(p) for p in case["pieces"]]


# --- (b) against the reference's compiled kernels ---------------------------

_REF_KEYS = sorted({k.split("/")[0] for k in REF.files})


@pytest.mark.parametrize("key", _REF_KEYS)
def test_oracle_matches_reference_kernels(key):
    w = int(key.split("_")[0][1:])
    d = int(key.split("_")[1][1:])
    g = {name.split("/")[1]: REF[name] for name in REF.files if name.startswith(key + "/")}
    dt = g["a"].dtype
    tol = 1e-14 if dt == np.float64 else 2e-6
    fma = g["c"].copy()
    O.ft_fma(fma, g["a"], g["b"], w, d)
    assert_allclose(fma, g["fma"], rtol=tol, atol=tol)
    im = g["a"].copy()
    O.ft_inplace_mul(im, g["b"], w, d)
    assert_allclose(im, g["inplace_mul"], rtol=tol, atol=tol)
    assert_allclose(O.ft_exp(g["x"], w, d), g["exp"], rtol=tol, atol=tol)
    assert_allclose(O.ft_fmexp(g["a"], g["x"], w, d), g["fmexp"], rtol=tol, atol=tol)
    assert_allclose(O.ft_log(g["a"], w, d), g["log"], rtol=tol, atol=tol)
    assert_array_equal(O.antipode(g["c"], w, d), g["antipode"])
    assert_allclose(O.ft_adjoint_left_mul(g["a"], g["c"], w, d), g["adj_lmul"], rtol=tol, atol=tol)
    sig = O.signature_batch(g["sig_increments"], g["sig_offsets"], w, d)
    assert_allclose(sig, g["signature"], rtol=tol, atol=tol)
    st = g["c"].copy()
    O.st_fma(st, g["a"], g["b"], w, d)
    assert_array_equal(st, g["st_fma"])          # same loop order: exact
    assert_array_equal(O.st_adj_mul(g["a"], g["c"], w, d), g["st_adj_mul"])


sys.path.insert(0, os.path.join(os.path.dirname(__file__), "golden"))
import make_ref_vectors as MRV  # noqa: E402

DIGESTS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ref_digests.json")))


@pytest.mark.parametrize("w,d", MRV.BASELINE_SHAPES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_oracle_bit_identical_to_reference_kernels_at_benchmark_shapes(w, d, dtype):
    """At the benchmark shapes (4,6), (5,5), (8,4), (6,5) the oracle reproduces
    the reference kernels' outputs bit for bit: SHA-256 of every output equals
    the digest the reference harness wrote (tests/golden/ref_digests.json) on
    the same regenerated inputs; the sampled coefficients say where a
    mismatch would be."""
    sfx = "f64" if dtype == np.float64 else "f32"
    g = MRV.baseline_inputs(w, d, dtype)
    a, b, c, x = g["a"], g["b"], g["c"], g["x"]
    fma = c.copy()
    O.ft_fma(fma, a, b, w, d)
    im = a.copy()
    O.ft_inplace_mul(im, b, w, d)
    st = c.copy()
    O.st_fma(st, a, b, w, d)
    outs = dict(fma=fma, inplace_mul=im, exp=O.ft_exp(x, w, d), fmexp=O.ft_fmexp(a, x, w, d),
                log=O.ft_log(a, w, d), antipode=O.antipode(c, w, d),
                adj_lmul=O.ft_adjoint_left_mul(a, c, w, d), st_fma=st,
                st_adj_mul=O.st_adj_mul(a, c, w, d),
                signature=O.signature_batch(g["sig_increments"], g["sig_offsets"], w, d))
    for name, arr in outs.items():
        ref = DIGESTS[f"w{w}_d{d}_{sfx}/{name}"]
        assert list(arr.shape) == ref["shape"], name
        flat = arr.ravel()
        for i, v in ref["samples"].items():
            assert flat[int(i)] == v, (name, i, flat[int(i)], v)
        assert MRV.digest(arr) == ref["sha256"], name


# --- (c) mathematics ----------------------------------------------------------

def _frac_mul(a, b, w, d):
    db = O.tensor_degree_begin(w, d)
    out = [Fraction(0)] * int(db[-1])
    for la in range(d + 1):
        for lb in range(d + 1 - la):
            na, nb = w ** la, w ** lb
            for i in range(na):
                for j in range(nb):
                    out[db[la + lb] + i * nb + j] += a[db[la] + i] * b[db[lb] + j]
    return out


def _frac_exp_increment(x, w, d):
    # exp of a degree-one element: level k = x^(k) / k!
    db = O.tensor_degree_begin(w, d)
    out = [Fraction(0)] * int(db[-1])
    out[0] = Fraction(1)
    fact = 1
    for k in range(1, d + 1):
        fact *= k
        for idx, word in enumerate(product(range(w), repeat=k)):
            v = Fraction(1)
            for a in word:
                v *= x[a]
            out[db[k] + idx] = v / fact
    return out


def test_signature_exact_rational_bruteforce():
    w, d = 2, 3
    incs = [[Fraction(1, 2), Fraction(-3, 4)], [Fraction(5, 8), Fraction(1, 8)],
            [Fraction(-1, 4), Fraction(3, 2)]]
    s = _frac_exp_increment([Fraction(0)] * w, w, d)
    for x in incs:
        s = _frac_mul(s, _frac_exp_increment(x, w, d), w, d)
    got = O.signature(np.array([[float(v) for v in x] for x in incs]), w, d)
    assert_allclose(got, [float(v) for v in s], rtol=1e-15, atol=1e-15)


@pytest.mark.parametrize("w,d", [(2, 4), (3, 3), (4, 3)])
def test_chen_identity_and_log_exp(rng, w, d):
    x = rng.normal(size=(7, w)) / 3
    full = O.signature(x, w, d)
    left, right = O.signature(x[:3], w, d), O.signature(x[3:], w, d)
    assert_allclose(O.ft_mul(left, right, w, d), full, rtol=1e-13, atol=1e-14)
    lg = O.ft_log(full, w, d)
    assert_allclose(O.ft_exp(lg, w, d), full, rtol=1e-13, atol=1e-14)
    # log of a group-like element is a Lie element: l2t(t2l(log S)) == log S
    lie = O.tensor_to_lie(lg, w, d)
    assert_allclose(O.lie_to_tensor(lie, w, d), lg, rtol=1e-12, atol=1e-14)


def test_fmexp_and_antipode_invariants(rng):
    w, d = 3, 3
    a = O.ft_exp(np.concatenate([[0.0], rng.normal(size=w), np.zeros(9 + 27)]), w, d)
    x = np.zeros_like(a)
    x[1:1 + w] = rng.normal(size=w)
    assert_allclose(O.ft_fmexp(a, x, w, d), O.ft_mul(a, O.ft_exp(x, w, d), w, d), atol=1e-14)
    one = np.zeros_like(a)
    one[0] = 1
    assert_allclose(O.ft_fmexp(one, x, w, d), O.ft_exp(x, w, d), atol=1e-15)
    # antipode of a group-like element is its inverse; antipode is an involution
    assert_allclose(O.ft_mul(O.antipode(a, w, d), a, w, d), one, atol=1e-14)
    y = rng.normal(size=a.shape)
    assert_array_equal(O.antipode(O.antipode(y, w, d), w, d), y)


def test_adjoint_left_mul_is_the_adjoint(rng):
    # <A x, y> == <x, L_A^* y> for all x, y (brute force over a basis)
    w, d = 2, 3
    n = O.tensor_size(w, d)
    A = rng.normal(size=n)
    y = rng.normal(size=n)
    lhs = np.array([np.dot(O.ft_mul(A, e, w, d), y) for e in np.eye(n)])
    assert_allclose(O.ft_adjoint_left_mul(A, y, w, d), lhs, atol=1e-13)


def test_t2l_inverts_l2t_and_matches_dynkin(rng):
    w, d = 3, 4
    l2t, t2l = O.l2t_dense(w, d), O.t2l_dense(w, d)
    assert_allclose(t2l @ l2t, np.eye(l2t.shape[1]), atol=1e-15)
    # l2t(t2l(word u)) = r(u) / |u| for every word (Dynkin)
    db = O.tensor_degree_begin(w, d)
    for k in (2, 3):
        for idx, word in enumerate(product(range(w), repeat=k)):
            col = l2t @ t2l[:, db[k] + idx]
            assert_allclose(col[db[k]:db[k + 1]], O.dynkin_word(w, list(word)) / k, atol=1e-14)


def test_constant_direction_closed_form():
    # tests/streams/test_lie_increment_path.py:162-216
    w, d = 3, 4
    n = 5
    t = np.arange(1.0, w + 1) * (2.0 / n)
    sig = O.signature(np.tile(t, (n, 1)), w, d)
    letters = np.arange(1, w + 1, dtype=float)
    exp = [1.0]
    fact = 1.0
    for k in range(1, d + 1):
        fact *= k
        for word in product(range(w), repeat=k):
            exp.append(np.prod(letters[list(word)]) * 2.0 ** k / fact)
    assert_allclose(sig, exp, rtol=1e-13)


def test_linear_signature_golden():
    g = GOLD["linear_signature"]
    x = np.array(g["increment"])
    sig = O.signature(x[None, :], 3, 3)
    l2 = np.einsum("i,j->ij", x, x) / 2
    l3 = np.einsum("ij,k->ijk", l2, x) / 3
    assert_allclose(sig, np.hstack([[1.0], x, l2.ravel(), l3.ravel()]), rtol=1e-14)


def test_treelike_paths():
    for pair in GOLD["treelike_paths"]["pairs"]:
        a = np.diff(np.array(pair["tree"], float), axis=0)
        b = np.diff(np.array(pair["pruned"], float), axis=0)
        sa = O.signature(a, 2, 2)
        sb = O.signature(b, 2, 2) if b.size else np.eye(1, 7).ravel()
        assert_allclose(sa, sb, atol=1e-12)


# --- stream query semantics -------------------------------------------------

def test_stream_tick_accuracy_golden():
    g = GOLD["stream_tick_accuracy"]
    s = O.IncrementStream(np.array(g["data"]), 2, 2, indices=np.array(g["indices"]))
    for q in g["queries"]:
        r = s.signature(q["inf"], q["sup"], O.py_param_to_resolution(q["resolution_param"]))
        if q["expected"] == "unit":
            assert_array_equal(r, np.eye(1, 7).ravel())
        else:
            arg = np.zeros(7)
            arg[:3] = q["exp_arg"]
            assert_allclose(r, O.ft_exp(arg, 2, 2), rtol=1e-15)


def test_stream_interval_alignment_golden():
    g = GOLD["stream_interval_alignment"]
    s = O.IncrementStream(np.array(g["data"]), 3, 2)
    assert_allclose(s.log_signature(g["first"]["inf"], g["first"]["sup"])[:3],
                    g["first"]["expected_lie"])
    for a, b in g["zero_windows"]:
        assert_allclose(s.log_signature(a, b), 0.0, atol=0)
        assert_allclose(s.log_signature(a, b, g["set_resolution"]), 0.0, atol=0)


def test_stream_sequence_and_level_one():
    g = GOLD["stream_sequence_of_lies"]
    s = O.IncrementStream(np.array(g["data"]), 3, 3)
    assert_allclose(s.log_signature(g["inf"], g["sup"])[:3], g["expected_lie"])
    for c in GOLD["stream_from_array_level1"]["cases"]:
        data = np.array(c["data"], float)[:, : c["width"]]
        s = O.IncrementStream(data, c["width"], c["sig_depth"])
        lo, hi = s.support
        assert_allclose(s.signature(lo, hi), c["expected"])


# --- planted mistakes: each pin above must catch them ------------------------

def test_planted_mistakes_are_caught():
    g = GOLD["ft_fma_explicit"]
    b, c = np.array(g["b"], float), np.array(g["c"], float)
    good = O.ft_mul(b, c, 2, 2)
    transposed = O.ft_mul(c, b, 2, 2)          # operands swapped
    assert np.array_equal(good, g["expected"])
    assert not np.array_equal(transposed, g["expected"])
    x = np.zeros(O.tensor_size(3, 3))
    x[1] = 1.0
    e = O.ft_exp(x, 3, 3)
    dropped_factorial = e.copy()
    dropped_factorial[1 + 3] *= 2              # x^2 without the 1/2!
    assert not np.allclose(dropped_factorial, [1, 1, 0, 0, 0.5] + [0] * 8 + [1 / 6] + [0] * 26)
    assert np.allclose(e, [1, 1, 0, 0, 0.5] + [0] * 8 + [1 / 6] + [0] * 26)


# --- signature derivative (forward mode) -------------------------------------

def test_signature_derivative_goldens():
    # tests/streams/test_lie_increment_path.py:219-247
    x = np.array([0.2, 0.4, 0.6])
    for depth, expected in [(1, [0.0, 0.0, 1.0, 0.0]),
                            (2, GOLD["sig_derivative_w3_d2"]["expected"])]:
        n = O.tensor_size(3, depth)
        E = O.signature(x[None, :], 3, depth)
        P = np.zeros(n)
        P[2] = 1.0  # letter 2
        got = O.signature_derivative([(E, P)], 3, depth)
        assert_allclose(got, expected, atol=1e-15)


def test_dexp_series_is_the_derivative_of_exp(rng):
    # d/de exp(X + e P) = exp(X) dexp_X(P): central differences in float64
    w, d = 3, 4
    n = O.tensor_size(w, d)
    X = np.zeros(n)
    X[1:] = rng.normal(size=n - 1) * 0.3
    X = O.tensor_to_lie(X, w, d) @ O.l2t_dense(w, d).T  # a Lie element
    P = np.zeros(n)
    P[1:] = rng.normal(size=n - 1)
    P = O.tensor_to_lie(P, w, d) @ O.l2t_dense(w, d).T
    h = 1e-6
    fd = (O.ft_exp(X + h * P, w, d) - O.ft_exp(X - h * P, w, d)) / (2 * h)
    an = O.ft_mul(O.ft_exp(X, w, d), O.dexp_series(X, P, w, d), w, d)
    assert_allclose(an, fd, atol=1e-8)


def test_signature_derivative_composition(rng):
    # test_lie_increment_path.py:266-287: the list form equals the running
    # product of single-interval derivatives
    w, d = 2, 3
    Es = [O.signature(rng.normal(size=(3, w)) / 3, w, d) for _ in range(3)]
    Ps = [O.lie_to_tensor(rng.uniform(0, 5, size=(O.tensor_size(w, d) * 0 + 5,)), w, d)
          for _ in range(3)]
    whole = O.signature_derivative(list(zip(Es, Ps)), w, d)
    exp_ = np.zeros_like(whole)
    for E, P in zip(Es, Ps):
        exp_ = O.ft_mul(exp_, E, w, d) + O.signature_derivative([(E, P)], w, d)
    assert_allclose(whole, exp_, atol=1e-13)


# --- shuffle product ----------------------------------------------------------

def _flat(w, *letters):
    """Flat index of a word (0-based letters), first letter most significant."""
    db = O.tensor_degree_begin(w, len(letters))
    i = 0
    for a in letters:
        i = i * w + a
    return int(db[len(letters)]) + i


def test_shuffle_goldens():
    """tests/compute/test_shuffle.py:58-112: letter (sh) letter, and a degree-2
    word (sh) a letter, at width 4, depth 3."""
    w, d = 4, 3
    size = O.tensor_size(w, d)
    lhs = np.zeros(size)
    rhs = np.zeros(size)
    lhs[_flat(w, 0)] = 1.0
    rhs[_flat(w, 1)] = 1.0
    expected = np.zeros(size)
    expected[_flat(w, 0, 1)] = 1.0
    expected[_flat(w, 1, 0)] = 1.0
    assert_array_equal(O.st_mul(lhs, rhs, w, d), expected)
    lhs = np.zeros(size)
    rhs = np.zeros(size)
    lhs[_flat(w, 0, 1)] = 1.0
    rhs[_flat(w, 2)] = 1.0
    expected = np.zeros(size)
    for word in [(0, 1, 2), (0, 2, 1), (2, 0, 1)]:
        expected[_flat(w, *word)] = 1.0
    assert_array_equal(O.st_mul(lhs, rhs, w, d), expected)


def test_shuffle_commutative_with_unit(rng):
    """tests/compute/test_shuffle.py:36-55."""
    w, d = 4, 3
    a = rng.uniform(-1, 1, size=O.tensor_size(w, d))
    b = rng.uniform(-1, 1, size=O.tensor_size(w, d))
    assert_allclose(O.st_mul(a, b, w, d), O.st_mul(b, a, w, d), atol=1e-14)
    unit = np.zeros_like(a)
    unit[0] = 1.0
    assert_allclose(O.st_mul(a, unit, w, d), a, atol=1e-15)
    assert_allclose(O.st_mul(unit, a, w, d), a, atol=1e-15)


def _shuffles(u, v):
    """All interleavings of the words u and v, by the recursion
    ua (sh) vb = (u (sh) vb) a + (ua (sh) v) b."""
    if not u:
        return [v]
    if not v:
        return [u]
    return [x + u[-1:] for x in _shuffles(u[:-1], v)] + [x + v[-1:] for x in _shuffles(u, v[:-1])]


def test_shuffle_matches_the_recursive_definition(rng):
    """Brute force by the recursive definition of the shuffle of words,
    expanded bilinearly, with exact small-integer coefficients."""
    w, d = 2, 4
    size = O.tensor_size(w, d)
    a = rng.integers(-3, 4, size=size).astype(np.float64)
    b = rng.integers(-3, 4, size=size).astype(np.float64)
    ref = np.zeros(size)
    words = [()] + [wd for k in range(1, d + 1) for wd in product(range(w), repeat=k)]
    for u in words:
        for v in words:
            if len(u) + len(v) > d:
                continue
            for x in _shuffles(list(u), list(v)):
                ref[_flat(w, *x)] += a[_flat(w, *u)] * b[_flat(w, *v)]
    assert_array_equal(O.st_mul(a, b, w, d), ref)


def test_shuffle_identity_on_signatures(rng):
    """<S, x (sh) y> = <S, x> <S, y> for a group-like S (a signature) when
    deg x + deg y <= depth."""
    w, d = 3, 5
    S = O.signature(rng.normal(size=(9, w)) / 3, w, d)
    db = O.tensor_degree_begin(w, d)
    x = np.zeros(O.tensor_size(w, d))
    y = np.zeros_like(x)
    x[:db[3]] = rng.normal(size=int(db[3]))      # degrees 0..2
    y[:db[4]] = rng.normal(size=int(db[4]))      # degrees 0..3
    assert abs(np.dot(S, O.st_mul(x, y, w, d)) - np.dot(S, x) * np.dot(S, y)) < 1e-12


def test_shuffle_adjoint_is_the_adjoint(rng):
    """<st_adj_mul(op, arg), x> = <arg, op (sh) x> over a basis of x."""
    w, d = 3, 3
    size = O.tensor_size(w, d)
    op = rng.normal(size=size)
    arg = rng.normal(size=size)
    adj = O.st_adj_mul(op, arg, w, d)
    E = np.eye(size)
    brute = np.array([np.dot(arg, O.st_mul(op, E[i], w, d)) for i in range(size)])
    assert_allclose(adj, brute, atol=1e-13)


# ---------------------------------------------------------------------------
# CBH, piecewise abelian streams and Stream.simplify
# ---------------------------------------------------------------------------

def test_cbh_closed_forms(rng):
    """cbh of one element is that element; of x and -x is 0; at width 2,
    depth 2 (Hall basis e1, e2, [e1, e2]) cbh(a e1, b e2) = (a, b, ab/2), the
    Baker-Campbell-Hausdorff series to second order; parallel level-one
    elements add."""
    w, d = 2, 2
    x = rng.normal(size=3)
    assert_array_equal(O.cbh([x], w, d), x)
    assert_allclose(O.cbh([x, -x], w, d), 0, atol=1e-15)
    a, b = 0.7, -1.3
    assert_allclose(O.cbh([np.array([a, 0, 0.]), np.array([0, b, 0.])], w, d),
                    [a, b, a * b / 2], atol=1e-15)
    assert_allclose(O.cbh([np.array([a, 2 * a, 0.]), np.array([b, 2 * b, 0.])], w, d),
                    [a + b, 2 * (a + b), 0], atol=1e-15)
    assert_allclose(O.cbh([], w, d), 0, atol=0)


def test_simplified_stream_keeps_the_signature(rng):
    """Chen's identity: when every dyadic interval of a query holds whole
    pieces, the simplified stream's signature over the support equals the
    signature of the increments (unit pieces at resolution 0, and two pieces
    inside the single dyadic interval [0, 8))."""
    w, d, N = 3, 4, 8
    inc = rng.normal(size=(N, w)) * 0.3
    st = O.IncrementStream(inc, w, d)
    full = O.signature(inc, w, d)
    pas = O.simplify(st, [(float(i), i + 1.0) for i in range(N)], 0)
    assert_allclose(pas.signature(0, N, 0), full, atol=1e-13)
    pas2 = O.simplify(st, [(0.0, 3.0), (3.0, 8.0)], 0)
    assert_allclose(pas2.signature(0, 8, 0), full, atol=1e-13)
    # a sub-range that is a union of dyadic intervals at resolution 0
    assert_allclose(pas.signature(2, 6, 0), O.signature(inc[2:6], w, d), atol=1e-13)


def test_piecewise_abelian_partial_overlaps(rng):
    """Known answers of the reference's multipliers for one piece [0, 1)
    holding L: [0.5, 1) at resolution 1 gives (1 - 0.5) L; [0.25, 0.5) at
    resolution 2 gives (1 - 0.25) L (the upper multiplier for an interior
    interval); [0, 0.5) at resolution 1 gives (1 - 0) L (a piece sticking out
    on the right takes the same branch); outside the support gives 0."""
    w, d = 2, 3
    L = rng.normal(size=5)
    pas = O.PiecewiseAbelianStream([((0.0, 1.0), L)], w, d, 0, (0.0, 1.0))
    assert_allclose(pas.log_signature(0.5, 1.0, 1), 0.5 * L, atol=1e-15)
    assert_allclose(pas.log_signature(0.25, 0.5, 2), 0.75 * L, atol=1e-15)
    assert_allclose(pas.log_signature(0.0, 0.5, 1), L, atol=1e-15)
    assert_array_equal(pas.log_signature(2.0, 3.0, 1), np.zeros(5))
    # two pieces in one dyadic interval combine by CBH
    M = rng.normal(size=5)
    pas2 = O.PiecewiseAbelianStream([((0.0, 0.5), L), ((0.5, 1.0), M)], w, d, 0, (0.0, 1.0))
    assert_allclose(pas2.log_signature(0.0, 1.0, 0), O.cbh([L, M], w, d), atol=1e-15)
