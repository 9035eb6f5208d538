"""Coded (Reed-Solomon) source delays on the host: the oracle's decoder
(Peterson-Gorenstein-Zierler, exhaustive root search, Vandermonde error
values over a polynomial-arithmetic field) pinned by brute force -- every
error pattern of weight <= t decodes -- and the engine's Berlekamp-Massey +
Forney decoder (qsft_amd.reed_solomon, whose tables the device decoder uses)
held to it on the same delay matrices and on random syndromes, failures
included. The delay matrix follows qsft/ReedSolomon.py:52-65 with the
field polynomial stated in qsft_amd.reed_solomon (galois, which the
reference uses, is absent: parity of the matrix itself is unpinned)."""
import itertools

import numpy as np
import pytest

from oracle import qsft_oracle as O
from qsft_amd.query import get_D, get_reed_solomon_dec
from qsft_amd.reed_solomon import ReedSolomon, primitive_poly


@pytest.mark.parametrize("n,t,q", [(8, 2, 3), (10, 2, 2), (12, 2, 5), (9, 3, 3), (6, 2, 7)])
def test_oracle_decoder_corrects_every_pattern_up_to_t(n, t, q):
    D = O.rs_delay_matrix(n, t, q)
    s, nc = O.rs_params(n, t, q)
    assert D.shape == (2 * t * s + 1, n) and not D[0].any()
    count = 0
    for w in range(t + 1):
        for pos in itertools.combinations(range(n), w):
            for vals in itertools.product(range(1, q), repeat=w):
                k = np.zeros(n, dtype=np.int64)
                k[list(pos)] = vals
                dec, nerr = O.rs_decode((D[1:] @ k) % q, n, t, q)
                assert np.array_equal(dec, k) and nerr == w, (pos, vals)
                count += 1
    assert count > 50


def test_oracle_decoder_refuses_patterns_it_cannot_correct():
    """Weight t + 1 patterns never decode to themselves; a decode that
    succeeds returns a pattern of weight <= t with the same syndrome."""
    n, t, q = 8, 2, 3
    D = O.rs_delay_matrix(n, t, q)
    for pos in itertools.combinations(range(n), t + 1):
        k = np.zeros(n, dtype=np.int64)
        k[list(pos)] = 1
        syn = (D[1:] @ k) % q
        dec, nerr = O.rs_decode(syn, n, t, q)
        assert not np.array_equal(dec, k)
        if nerr >= 0:
            assert np.count_nonzero(dec) <= t and np.array_equal((D[1:] @ dec) % q, syn)


@pytest.mark.parametrize("n,t,q", [(8, 2, 3), (10, 2, 2), (12, 2, 5), (40, 4, 3), (50, 3, 5), (20, 4, 29)])
def test_engine_decoder_matches_oracle(n, t, q):
    rs = ReedSolomon(n, t, q)
    D = rs.get_delay_matrix()
    assert np.array_equal(D, O.rs_delay_matrix(n, t, q))
    rng = np.random.default_rng(n * t * q)
    for _ in range(120):                                          # planted weight <= t
        k = np.zeros(n, dtype=np.int64)
        w = int(rng.integers(0, t + 1))
        k[rng.choice(n, w, replace=False)] = rng.integers(1, q, w)
        dec, nerr = rs.decode((D[1:] @ k) % q)
        assert np.array_equal(dec, k) and nerr == w
    for _ in range(120):                                          # arbitrary syndromes, failures included
        syn = rng.integers(0, q, D.shape[0] - 1)
        a, na = rs.decode(syn)
        b, nb = O.rs_decode(syn, n, t, q)
        assert np.array_equal(a, b) and na == nb


def test_field_polynomials_are_primitive_and_smallest():
    # GF(2^4): x^4 + x + 1 is the smallest primitive quartic; GF(3^2): x^2 + x + 2
    assert primitive_poly(2, 4) == [1, 1, 0, 0]
    assert primitive_poly(3, 2) == [2, 1]
    for q, s in [(2, 7), (3, 4), (5, 3), (7, 2)]:
        rs_field = ReedSolomon(q ** s - 3, 1, q).field
        assert len(set(rs_field.exp[:q ** s - 1].tolist())) == q ** s - 1      # alpha generates the group


def test_coded_delays_take_no_rng_draws_and_decoder_api():
    st = np.random.get_state()
    D = get_D(40, q=3, t=4, delays_method_source="coded", delays_method_channel="identity")
    assert np.random.get_state()[1].tolist() == st[1].tolist()
    assert D[0].shape == (33, 40)
    dec = get_reed_solomon_dec(40, 4, 3)
    k = np.zeros(40, dtype=np.int64)
    k[[3, 17, 39]] = [1, 2, 2]
    out, nerr = dec(list((D[0][1:] @ k) % 3))
    assert np.array_equal(out[0], k) and nerr == 3
    with pytest.raises(NotImplementedError):
        get_reed_solomon_dec(40, 4, 4)
