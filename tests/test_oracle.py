"""Pins for the CPU oracle (oracle/qsft_oracle.py) against closed forms, brute
force and hand-built decision cases. CPU only."""
import itertools

import numpy as np
import pytest

from oracle import qsft_oracle as O


# ----------------------------------------------------------------- helpers


@pytest.mark.parametrize("q,m", [(2, 3), (3, 2), (4, 3), (20, 2), (5, 1)])
def test_qary_ints_matches_itertools_product(q, m):
    ref = np.array(list(itertools.product(range(q), repeat=m))).T
    np.testing.assert_array_equal(O.qary_ints(m, q), ref)


@pytest.mark.parametrize("q,n", [(4, 50), (3, 40), (20, 16), (2, 70)])
def test_dec_roundtrip_and_value(q, n):
    rng = np.random.default_rng(1)
    dig = rng.integers(0, q, size=(n, 37))
    dec = O.qary_vec_to_dec(dig, q)
    for c in range(dig.shape[1]):
        # independent decimal value: Horner over Python ints
        v = 0
        for d in dig[:, c]:
            v = v * q + int(d)
        assert dec[c] == v
    np.testing.assert_array_equal(O.dec_to_qary_vec(dec, q, n), dig)


def test_sort_and_hamming():
    rng = np.random.default_rng(2)
    v = rng.integers(0, 3, size=(40, 5))
    ref = np.array(sorted(map(tuple, v)))
    np.testing.assert_array_equal(O.sort_qary_vecs(v), ref)
    np.testing.assert_array_equal(O.calc_hamming_weight(v), [sum(x != 0 for x in r) for r in v])


@pytest.mark.parametrize("q,b", [(2, 4), (3, 3), (4, 3), (5, 2)])
def test_gwht_is_dft_definition(q, b):
    """Brute-force multidimensional DFT with e^{-2 pi i <j,l>/q} and 1/q^b."""
    rng = np.random.default_rng(3)
    x = rng.normal(size=q ** b) + 1j * rng.normal(size=q ** b)
    L = O.qary_ints(b, q)                       # (b, Q)
    F = np.exp(-2j * np.pi * ((L.T @ L) % q) / q) / q ** b
    np.testing.assert_allclose(O.gwht(x, q, b), F @ x, rtol=0, atol=1e-12)


# ------------------------------------------------------------ query matrices


def test_Ms_complex_and_simple():
    np.random.seed(5)
    Ms = O.get_Ms(12, 3, 4, num_to_get=3, method="complex")
    assert len(Ms) == 3 and all(M.shape == (12, 3) for M in Ms)
    assert all((M >= 0).all() and (M < 4).all() for M in Ms)
    Ms = O.get_Ms(9, 3, 4, num_to_get=3, method="simple")
    # group 0 holds the identity in the LAST block (reverse order)
    np.testing.assert_array_equal(Ms[0][6:9], np.eye(3))
    np.testing.assert_array_equal(Ms[2][0:3], np.eye(3))
    with pytest.raises(ValueError):
        O.get_Ms(8, 3, 4, num_to_get=3, method="simple")


@pytest.mark.parametrize("q,n,R", [(4, 10, 2), (3, 7, 3), (20, 5, 1)])
def test_D_nso_structure(q, n, R):
    np.random.seed(6)
    Ds = O.get_D(n, q=q, delays_method_source="identity", delays_method_channel="nso", num_repeat=R)
    assert len(Ds) == R
    for D in Ds:
        assert D.shape == (n + 1, n)
        for i in range(1, n + 1):
            e = np.zeros(n, int)
            e[i - 1] = 1
            np.testing.assert_array_equal((D[0] - D[i]) % q, e)


def test_D_identity_and_shared_list():
    np.random.seed(7)
    Ms, Ds = O.get_Ms_and_Ds(6, 4, query_method="complex", num_subsample=3, b=2,
                             delays_method_source="identity", delays_method_channel="identity")
    assert len(Ds) == 3 and Ds[0] is Ds[1] is Ds[2]
    np.testing.assert_array_equal(Ds[0][0], np.vstack([np.zeros(6), np.eye(6)]))


def test_coset_digits_brute_force():
    rng = np.random.default_rng(8)
    q, n, b = 3, 5, 2
    M = rng.integers(0, q, size=(n, b))
    D = rng.integers(0, q, size=(4, n))
    dig = O.coset_digits(M, D, q, b)
    for p in range(4):
        for c, l in enumerate(itertools.product(range(q), repeat=b)):
            np.testing.assert_array_equal(dig[p, c], (M @ np.array(l) + D[p]) % q)


# ----------------------------------------------------- evaluation + sub-DFT


@pytest.mark.parametrize("q,n,S", [(3, 5, 6), (4, 4, 9), (2, 7, 5)])
def test_eval_matches_dense_inverse_dft(q, n, S):
    rng = np.random.default_rng(9)
    locq = rng.integers(0, q, size=(n, S))
    a = rng.normal(size=S) + 1j * rng.normal(size=S)
    dense = O.dense_time_signal(locq, a, q, n)
    m = rng.integers(0, q, size=(50, n))
    got = O.eval_synthetic(m, locq, a, q)
    np.testing.assert_allclose(got, dense[tuple(m.T)], rtol=0, atol=1e-11)


@pytest.mark.parametrize("q,n,b,S,chan,R", [(4, 6, 3, 10, "nso", 2), (3, 5, 2, 7, "identity", 1),
                                           (20, 3, 2, 12, "nso", 1), (2, 9, 4, 9, "nso", 3)])
def test_subsampled_transform_matches_closed_form(q, n, b, S, chan, R):
    np.random.seed(10)
    qa = dict(query_method="complex", num_subsample=3, b=b, delays_method_source="identity",
              delays_method_channel=chan, num_repeat=R, subsampling_method="qsft")
    sig = O.get_random_subsampled_signal(n, q, 0.0, S, 1, 10, qa)
    for i in range(3):
        for j in range(R):
            U = np.array(sig.Us[i][j][b])
            cf = O.closed_form_U(sig.Ms[i], sig.Ds[i][j], sig.locq, sig.strengths, q)
            np.testing.assert_allclose(U, cf, rtol=0, atol=1e-10)


def test_all_bs_uses_leading_columns():
    np.random.seed(11)
    q, n, b = 4, 7, 3
    qa = dict(query_method="complex", num_subsample=2, b=b, all_bs=[2, 3],
              delays_method_source="identity", delays_method_channel="identity")
    sig = O.get_random_subsampled_signal(n, q, 0.0, 6, 1, 10, qa)
    U2 = np.array(sig.Us[0][0][2])
    cf = O.closed_form_U(sig.Ms[0][:, :2], sig.Ds[0][0], sig.locq, sig.strengths, q)
    np.testing.assert_allclose(U2, cf, rtol=0, atol=1e-10)


def test_faithful_indices_identical():
    np.random.seed(12)
    qa = dict(query_method="complex", num_subsample=2, b=2, delays_method_source="identity",
              delays_method_channel="identity")
    st = np.random.get_state()
    a = O.get_random_subsampled_signal(9, 3, 0.0, 5, 1, 10, dict(qa))
    np.random.set_state(st)
    b = O.get_random_subsampled_signal(9, 3, 0.0, 5, 1, 10, dict(qa), faithful_indices=True)
    np.testing.assert_array_equal(np.array(a.Us[1][0][2]), np.array(b.Us[1][0][2]))


def test_noise_variance_per_bin():
    np.random.seed(13)
    q, n, b = 4, 8, 4
    qa = dict(query_method="complex", num_subsample=3, b=b, delays_method_source="identity",
              delays_method_channel="nso", num_repeat=2)
    sig = O.get_random_subsampled_signal(n, q, 3.0, 1, 1, 1, qa)
    base = {(i, j): np.array(sig.Us[i][j][b]) for i in range(3) for j in range(2)}
    np.random.seed(14)
    Ms, Ds, Us = sig.get_MDU(3, 2, b)
    # stored transforms untouched, noise has variance noise_sd^2 / q^b per bin
    for i in range(3):
        for j in range(2):
            np.testing.assert_array_equal(np.array(sig.Us[i][j][b]), base[(i, j)])
    allnoise = np.concatenate([(Us[i][j] - base[(0, 0)] * 0).ravel() for i in range(3) for j in range(2)])
    # subtract the signal part using the permutation drawn inside get_MDU
    np.random.seed(14)
    si = np.random.choice(3, 3, replace=False)
    di = np.random.choice(2, 2, replace=False)
    noise = np.concatenate([(Us[a][c] - base[(si[a], di[c])]).ravel() for a in range(3) for c in range(2)])
    assert abs(np.var(noise) / (9.0 / q ** b) - 1) < 0.05
    assert allnoise.size == noise.size


# --------------------------------------------------------------- detection


def _nso_D(q, n, R, rng):
    src = np.vstack([np.zeros(n), np.eye(n)]).astype(int)
    return np.vstack([(rng.integers(0, q, n) - src) % q for _ in range(R)])


@pytest.mark.parametrize("q", [2, 3, 4, 20])
def test_detect_isolated_singleton(q):
    rng = np.random.default_rng(15)
    n = 9
    k = rng.integers(0, q, n)
    a = 3.0 * np.exp(1j * 0.7)
    D = _nso_D(q, n, 2, rng)
    col = a * np.exp(2j * np.pi * ((D @ k) % q) / q)
    np.testing.assert_array_equal(O.detect_nso1(col, q, n + 1), k)
    D0 = np.vstack([np.zeros(n), np.eye(n)]).astype(int)
    col0 = a * np.exp(2j * np.pi * ((D0 @ k) % q) / q)
    np.testing.assert_array_equal(O.detect_noiseless(col0, q), k)


def test_nso1_tie_takes_first_root():
    # angle exactly pi/4 for q=4 sits halfway between roots 0 and pi/2 -> t=0
    q, n = 4, 1
    col = np.array([1 + 0j, (1 - 1j) / np.sqrt(2)])     # U0 conj(U1) has angle pi/4
    z = np.mean(col[0::2] * np.conj(col[1::2]))
    assert np.angle(z) == np.pi / 4
    assert O.detect_nso1(col, q, n + 1)[0] == 0


def test_noiseless_round_half_even():
    # q (dphi)/2pi == 0.5 exactly -> numpy rounds to 0 (half-even), not 1
    q = 4
    col = np.array([1 + 0j, np.exp(-1j * np.pi / 4)])
    v = q * (np.angle(col)[1:] - np.angle(col)[0]) / (2 * np.pi)
    assert v[0] == -0.5
    assert O.detect_noiseless(col, q)[0] == 0


# ------------------------------------------------------------------ peeling


def _group_U(M, D, ks, amps, q, Q):
    """U (P, Q) for one group holding balls ks with amplitudes amps."""
    b = M.shape[1]
    U = np.zeros((D.shape[0], Q), complex)
    w = q ** np.arange(b - 1, -1, -1)
    for k, a in zip(ks, amps):
        j = int(w @ ((M.T @ k) % q))
        U[:, j] += a * np.exp(2j * np.pi * ((D @ k) % q) / q)
    return U


def _find_config(q, n, b, C, want, rng, tries=20000):
    """Random Ms and balls satisfying a predicate on their bins."""
    w = q ** np.arange(b - 1, -1, -1)
    for _ in range(tries):
        Ms = [rng.integers(0, q, size=(n, b)) for _ in range(C)]
        ks = [rng.integers(0, q, n) for _ in range(want["balls"])]
        bins = [[int(w @ ((M.T @ k) % q)) for k in ks] for M in Ms]
        if want["pred"](bins):
            return Ms, ks, bins
    raise RuntimeError("no configuration found")


def test_jacobi_scan_order():
    """k1 alone in group 0; k1 and k3 share a bin in group 1; k3 alone in group 2.
    Under the reference's frozen-U scan the group-1 bin is a multiton in pass 1."""
    q, n, b, C = 4, 8, 2, 3
    rng = np.random.default_rng(16)

    def pred(bins):
        (a0, c0), (a1, c1), (a2, c2) = bins
        return a0 != c0 and a1 == c1 and a2 != c2
    Ms, ks, _ = _find_config(q, n, b, C, dict(balls=2, pred=pred), rng)
    D = _nso_D(q, n, 2, rng)
    amps = [2.0 + 1j, -1.5 + 0.5j]
    Us = np.array([_group_U(M, D, ks, amps, q, q ** b) for M in Ms])
    trace = []
    _, gw, it = O.peel_decode(Ms, [D] * C, Us, q, n, 1e-9, trace=trace)
    assert trace[0][1] == 1                      # one multiton in the first pass
    assert trace[0][0] == 4                      # k1 in groups 0, 2; k3 in groups 0, 2
    assert set(gw) == {tuple(k) for k in ks}
    for k, a in zip(ks, amps):
        assert abs(gw[tuple(k)] - a) < 1e-12
    assert np.abs(Us).max() < 1e-12


def test_last_writer_rho_and_running_mean():
    """The same ball detected in two groups with different noise: the peel uses
    the rho of the later (i, j); gwht is the mean of both detections."""
    q, n, b, C = 4, 6, 2, 2
    rng = np.random.default_rng(17)
    Ms, ks, bins = _find_config(q, n, b, C, dict(balls=1, pred=lambda b_: True), rng)
    D = _nso_D(q, n, 2, rng)
    a = 2.0 - 1.0j
    Us = np.array([_group_U(M, D, ks, [a], q, q ** b) for M in Ms])
    Us[0, :, bins[0][0]] += 0.01
    Us[1, :, bins[1][0]] -= 0.02j
    sig = np.exp(2j * np.pi * ((D @ ks[0]) % q) / q)
    rho0 = np.dot(np.conj(sig), Us[0, :, bins[0][0]]) / D.shape[0]
    rho1 = np.dot(np.conj(sig), Us[1, :, bins[1][0]]) / D.shape[0]
    before = Us.copy()
    result, gw, it = O.peel_decode(Ms, [D] * C, Us, q, n, 1e-2)
    assert it == 1                               # no multitons -> stop after this pass
    assert len(result) == 2
    # group 0 bin was peeled with rho1 (the last writer), not its own rho0
    np.testing.assert_allclose(Us[0, :, bins[0][0]], before[0, :, bins[0][0]] - rho1 * sig, atol=1e-14)
    np.testing.assert_allclose(gw[tuple(ks[0])], (rho0 + rho1) / 2, rtol=1e-15)


def test_bin_mismatch_is_multiton_and_zeroton_skipped():
    q, n, b = 4, 6, 2
    rng = np.random.default_rng(18)
    M = rng.integers(0, q, size=(n, b))
    D = _nso_D(q, n, 2, rng)
    k = rng.integers(0, q, n)
    w = q ** np.arange(b - 1, -1, -1)
    jtrue = int(w @ ((M.T @ k) % q))
    jwrong = (jtrue + 1) % q ** b
    U = np.zeros((1, D.shape[0], q ** b), complex)
    U[0, :, jwrong] = 1.5 * np.exp(2j * np.pi * ((D @ k) % q) / q)
    trace = []
    _, gw, it = O.peel_decode([M], [D], U, q, n, 1e-9, trace=trace)
    assert trace[0][:2] == (0, 1) and gw == {} and it == 1


def test_residual_rule_two_balls_same_bin():
    q, n, b = 4, 7, 2
    rng = np.random.default_rng(19)

    def pred(bins):
        return bins[0][0] == bins[0][1]
    Ms, ks, _ = _find_config(q, n, b, 1, dict(balls=2, pred=pred), rng)
    D = _nso_D(q, n, 2, rng)
    U = np.array([_group_U(Ms[0], D, ks, [1.0, 0.9j], q, q ** b)])
    trace = []
    O.peel_decode(Ms, [D], U, q, n, 1e-9, trace=trace)
    assert trace[0][:2] == (0, 1)


def test_iteration_cap_is_15():
    """Ball k sits in group 0 with amplitude 2 and in group 1 with amplitude 1.
    Peeling with the last writer's rho leaves (2-1) sig_k behind in one group,
    the next pass moves it to the other, and so on: one singleton every pass.
    A bin-mismatched column in group 0 is a multiton in every pass. The loop
    must stop at the 15-iteration cap."""
    q, n, b = 4, 6, 2
    rng = np.random.default_rng(20)
    w = q ** np.arange(b - 1, -1, -1)
    M0, M1 = rng.integers(0, q, size=(n, b)), rng.integers(0, q, size=(n, b))
    D = _nso_D(q, n, 2, rng)
    k = rng.integers(0, q, n)
    A, B = int(w @ ((M0.T @ k) % q)), int(w @ ((M1.T @ k) % q))
    sig = np.exp(2j * np.pi * ((D @ k) % q) / q)
    U = np.zeros((2, D.shape[0], q ** b), complex)
    U[0, :, A] = 2.0 * sig
    U[1, :, B] = 1.0 * sig
    k2 = rng.integers(0, q, n)
    X = (int(w @ ((M0.T @ k2) % q)) + 1) % q ** b
    if X == A:
        X = (X + 1) % q ** b
    U[0, :, X] = 1.5 * np.exp(2j * np.pi * ((D @ k2) % q) / q)
    trace = []
    result, gw, it = O.peel_decode([M0, M1], [D, D], U, q, n, 1e-3, trace=trace)
    assert it == 15 and len(trace) == 15
    assert all(t[0] >= 1 and t[1] >= 1 for t in trace)
    assert len(result) == 2 + 14


# -------------------------------------------------------------- end to end


def test_noiseless_identity_recovers_ground_truth():
    np.random.seed(21)
    q, n, S, b = 4, 10, 20, 4
    qa = dict(query_method="complex", num_subsample=3, b=b, delays_method_source="identity",
              delays_method_channel="identity", subsampling_method="qsft")
    sig = O.get_random_subsampled_signal(n, q, 0.0, S, 1, 10, qa)
    assert len(sig.signal_w) == S
    res = O.qsft_transform(sig, 3, 1, b, "identity", "identity", report=True, sort=True)
    assert set(res["gwht"]) == set(sig.signal_w)
    for k, v in sig.signal_w.items():
        assert abs(res["gwht"][k] - v) <= 1e-10 * abs(v)
    assert res["n_samples"] == 3 * (n + 1) * q ** b
    assert isinstance(res["locations"], np.ndarray)


def test_nso_low_noise_recovers_ground_truth():
    np.random.seed(22)
    q, n, S, b = 4, 12, 30, 4
    qa = dict(query_method="complex", num_subsample=3, b=b, delays_method_source="identity",
              delays_method_channel="nso", num_repeat=2, subsampling_method="qsft")
    sig = O.get_random_subsampled_signal(n, q, 0.05, S, 1, 10, qa)
    gw = O.qsft_transform(sig, 3, 2, b, "identity", "nso")
    assert set(gw) == set(sig.signal_w)
    assert O.nmse(gw, sig.signal_w) < 1e-4


def test_dense_spectrum_brute_force():
    """q^n = 4096: the dense forward DFT of the full time signal equals the
    planted spectrum, and the recovered coefficients match it."""
    np.random.seed(23)
    q, n, S, b = 4, 6, 8, 3
    qa = dict(query_method="complex", num_subsample=3, b=b, delays_method_source="identity",
              delays_method_channel="identity")
    sig = O.get_random_subsampled_signal(n, q, 0.0, S, 1, 10, qa)
    x = O.dense_time_signal(sig.locq, sig.strengths, q, n)
    spec = O.gwht(x.ravel(), q, n).reshape([q] * n)
    for k, v in sig.signal_w.items():
        assert abs(spec[k] - v) < 1e-10
    gw = O.qsft_transform(sig, 3, 1, b, "identity", "identity")
    for k, v in gw.items():
        assert abs(spec[k] - v) < 1e-9


# ---------------------------------------------------------- held-out NMSE


def test_heldout_nmse_hand_case():
    """q = 2, n = 1: beta = {0: 1, 1: 1} predicts y(0) = 2, y(1) = 0; against
    samples (2, 1) the error is (0, -1): NMSE = 1 / 5."""
    assert O.heldout_nmse([0, 1], [2.0, 1.0], {(0,): 1.0, (1,): 1.0}, 2, 1) == pytest.approx(0.2, abs=1e-15)
    assert O.heldout_nmse([1, 0], [0.0, 2.0], {(0,): 1.0, (1,): 1.0}, 2, 1) == pytest.approx(0.0, abs=1e-30)


@pytest.mark.parametrize("q,n,S", [(4, 12, 30), (3, 9, 20), (20, 4, 15)])
def test_heldout_nmse_exact_and_doubled(q, n, S):
    """Predictions from the true spectrum give NMSE ~0; doubling every
    coefficient gives exactly 1 (y_hat = 2 y); the empty spectrum gives 1.
    Samples come from a brute-force sum with phases reduced mod q."""
    rng = np.random.default_rng(q + n)
    locq = rng.integers(0, q, size=(n, S))
    a = rng.uniform(1, 10, S) * np.exp(2j * np.pi * rng.uniform(size=S))
    beta = {}
    for s in range(S):
        beta[tuple(locq[:, s])] = beta.get(tuple(locq[:, s]), 0) + a[s]
    idx = [int(v) for v in rng.integers(0, q ** n, size=2500)]
    dig = O.dec_to_qary_vec(idx, q, n)
    w = np.exp(2j * np.pi * np.arange(q) / q)
    y = np.array([sum(v * w[int(np.dot(dig[:, i], k)) % q] for k, v in beta.items()) for i in range(len(idx))])
    assert O.heldout_nmse(idx, y, beta, q, n, batch=1000) < 1e-24
    assert O.heldout_nmse(idx, y, {k: 2 * v for k, v in beta.items()}, q, n) == pytest.approx(1.0, rel=1e-12)
    assert O.heldout_nmse(idx, y, {}, q, n) == 1
