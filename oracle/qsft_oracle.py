"""CPU oracle for the q-SFT hot path: a numpy restatement of basics-lab/qsft.

TEST INFRASTRUCTURE ONLY. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker or as the timed CPU baseline. The product package ``qsft_amd`` never
imports it, links it or falls back to it.

Every function cites the reference file:line whose behaviour it restates
(paths relative to the reference repository root). The reference cannot be
imported in this image: ``qsft/utils.py:6`` imports the absent ``group_lasso``,
``qsft/query.py:11`` -> ``qsft/ReedSolomon.py:1-2`` import the absent
``galois``, and the hot path calls ``np.complex`` (removed in numpy 1.24) at
``qsft/input_signal_subsampled.py:136`` and
``synt_exp/synt_src/synthetic_signal.py:129``. It is therefore read, not run.

How each part is pinned (tests/test_oracle.py):
  * index/digit helpers: brute force over itertools.product, exact big-int
    round trips;
  * M/D generation: structural invariants (ranges, D_r[0]-D_r[i] == e_i) and
    NEP 19 stream stability of the legacy RandomState; the *call order* is
    pinned by restatement only ("parity unpinned" for RNG call order);
  * sampling + sub-DFTs: the closed form U_cp[j] = sum_{M^T k == j} a_k w^<k,d_p>
    and numpy.fft.fftn on reshaped tensors;
  * detection + peeling: ground-truth recovery on noiseless decodable
    instances, hand-built bins that exercise each decision rule (zeroton,
    bin-mismatch multiton, residual multiton, Jacobi scan, last-writer rho,
    append-order running mean, stop rule, 15-iteration cap), and dense
    brute-force spectra for q^n <= 2^16;
  * held-out NMSE: a hand-computed q = 2 case, exact predictions (NMSE 0),
    doubled coefficients (NMSE 1) and the empty spectrum.
"""
from __future__ import annotations

import itertools
import os
import time

import numpy as np
import scipy.fft

# ---------------------------------------------------------------------------
# q-ary helpers (qsft/utils.py)
# ---------------------------------------------------------------------------


def qary_ints(m, q):
    """All length-m q-ary vectors as columns, MSB-first (qsft/utils.py:107-108).

    Column c holds the base-q digits of c with digit 0 most significant, the
    order itertools.product(range(q), repeat=m) enumerates.
    """
    idx = np.arange(q ** m, dtype=np.int64)
    powers = q ** np.arange(m - 1, -1, -1, dtype=np.int64)
    return (idx[None, :] // powers[:, None]) % q


def qary_vec_to_dec(x, q):
    """Digits (n, N) -> Python-int decimals, MSB-first (qsft/utils.py:74-76)."""
    x = np.asarray(x)
    n = x.shape[0]
    w = np.array([q ** (n - 1 - i) for i in range(n)], dtype=object)
    return w @ np.array(x, dtype=object)


def dec_to_qary_vec(idx, q, n):
    """Decimals -> digits (n, N) int64, MSB-first (qsft/utils.py:79-84)."""
    rem = np.array(list(idx), dtype=object)
    out = []
    for i in range(n):
        p = q ** (n - 1 - i)
        d = np.array([v // p for v in rem], dtype=object)
        rem = rem - p * d
        out.append(d)
    return np.array(out, dtype=np.int64).reshape(n, len(rem))


def sort_qary_vecs(vecs):
    """Lexicographic sort of row vectors, column 0 primary (qsft/utils.py:180-183)."""
    vecs = np.array(vecs)
    order = np.lexsort(vecs.T[::-1, :])
    return vecs[order]


def calc_hamming_weight(vecs):
    """Number of nonzero digits per row (qsft/utils.py:185-187)."""
    return np.sum(np.array(vecs) != 0, axis=1)


def gwht(x, q, n):
    """Forward-scaled n-dim q-point DFT of a flat signal (qsft/utils.py:31-36)."""
    t = np.reshape(x, [q] * n)
    return np.reshape(scipy.fft.fftn(t) / (q ** n), [q ** n])


def random_signal_strength_model(S, a_min, a_max):
    """|a| ~ U[a_min, a_max], phase ~ U[0, 2pi) (qsft/utils.py:161-164)."""
    mag = np.random.uniform(a_min, a_max, S)
    ph = np.random.uniform(0, 2 * np.pi, S)
    return mag * np.exp(1j * ph)


# ---------------------------------------------------------------------------
# Query construction (qsft/query.py)
# ---------------------------------------------------------------------------


def get_Ms(n, b, q, num_to_get=None, method="simple"):
    """C subsampling matrices M_c (n, b) (qsft/query.py:13-70)."""
    if num_to_get is None:
        num_to_get = max(n // b, 3)
    if method == "simple" and num_to_get > n // b:
        raise ValueError("When query_method is 'simple', the number of M matrices "
                         "to return cannot be larger than n // b")
    if method == "simple":
        Ms = []
        for c in range(num_to_get - 1, -1, -1):          # reverse block order, query.py:18
            M = np.zeros((n, b), dtype=np.int32)
            M[b * c:b * (c + 1), :] = np.eye(b)
            Ms.append(M)
        return Ms
    if method == "complex":
        return [np.random.randint(q, size=(n, b)) for _ in range(num_to_get)]
    raise NotImplementedError(method)


# ---------------------------------------------------------------------------
# Coded source delays: Reed-Solomon over GF(q^s) (qsft/ReedSolomon.py,
# qsft/query.py:88-92,231-243, qsft/reconstruct.py:34-51). The reference takes
# the code from `galois` (absent here); this restates the construction with
# plain polynomial arithmetic: GF(q)[x] / f, f the smallest primitive
# polynomial (galois's default is the Conway polynomial -- parity of the
# delay matrix with the reference is unpinned), H[j, i] =
# alpha^{(1+j)(n-1-i)}, digits of each entry highest degree first. The
# decoder is Peterson-Gorenstein-Zierler (locator by linear solves over
# GF(q^s)), roots by exhaustive search over all code positions and error
# values by solving the Vandermonde system -- a different algorithm from
# the engine's Berlekamp-Massey + Forney.
# ---------------------------------------------------------------------------


class OracleGF:
    """GF(q^s) elements as coefficient tuples (c_0 .. c_{s-1}, low degree first)."""

    def __init__(self, q, s):
        self.q, self.s = q, s
        self.f = None
        for v in range(q ** s):                              # candidates ordered by c_{s-1} .. c_0
            c = [(v // q ** i) % q for i in range(s)]
            if c[0] == 0:
                continue
            self.f = c
            if self._order_of_x() == q ** s - 1:
                break
            self.f = None
        if self.f is None:
            raise ValueError("no primitive polynomial")
        self.zero = tuple([0] * s)
        self.one = tuple([1] + [0] * (s - 1))

    def _mulx(self, a):
        a = list(a)
        top = a[-1]
        a = [0] + a[:-1]
        return tuple((a[i] - top * self.f[i]) % self.q for i in range(self.s))

    def _order_of_x(self):
        a = self._mulx(tuple([1] + [0] * (self.s - 1)))
        k = 1
        while a != tuple([1] + [0] * (self.s - 1)):
            a = self._mulx(a)
            k += 1
            if k > self.q ** self.s:
                return -1
        return k

    def add(self, a, b):
        return tuple((x + y) % self.q for x, y in zip(a, b))

    def sub(self, a, b):
        return tuple((x - y) % self.q for x, y in zip(a, b))

    def mul(self, a, b):
        acc = self.zero
        cur = a
        for bi in b:                                         # sum_i b_i x^i a
            if bi:
                acc = self.add(acc, tuple((bi * c) % self.q for c in cur))
            cur = self._mulx(cur)
        return acc

    def power(self, a, e):
        r = self.one
        for _ in range(e):
            r = self.mul(r, a)
        return r

    def inv(self, a):
        return self.power(a, self.q ** self.s - 2)

    def alpha_pow(self, e):
        return self.power(tuple([0, 1] + [0] * (self.s - 2)) if self.s > 1 else (self._x1(),), e % (self.q ** self.s - 1))

    def _x1(self):
        # s = 1: alpha is the root of x - g, i.e. the element g = -c_0
        return (-self.f[0]) % self.q

    def digits_msb(self, a):
        return list(reversed(a))

    def from_digits_msb(self, d):
        return tuple(int(x) % self.q for x in reversed(list(d)))


def rs_params(n, t, q):
    """s, code length q^s - 1 (qsft/ReedSolomon.py:18-25)."""
    s = int(np.ceil(np.log(n) / np.log(q)))
    if n > q ** s - 1:
        s += 1
    return s, q ** s - 1


def rs_delay_matrix(n, t, q):
    """(2 t s + 1) x n coded source delays (qsft/ReedSolomon.py:52-65)."""
    s, nc = rs_params(n, t, q)
    F = OracleGF(q, s)
    D = np.zeros((2 * t * s + 1, n), dtype=np.int64)
    for j in range(2 * t):
        for i in range(n):
            D[s * j + 1: s * (j + 1) + 1, i] = F.digits_msb(F.alpha_pow((1 + j) * (n - 1 - i)))
    return D


def _gf_solve(F, A, y):
    """Solve A x = y over GF(q^s) by Gauss-Jordan; None if singular."""
    m = len(A)
    M = [list(row) + [y[i]] for i, row in enumerate(A)]
    for col in range(m):
        piv = next((r for r in range(col, m) if M[r][col] != F.zero), None)
        if piv is None:
            return None
        M[col], M[piv] = M[piv], M[col]
        iv = F.inv(M[col][col])
        M[col] = [F.mul(iv, v) for v in M[col]]
        for r in range(m):
            if r != col and M[r][col] != F.zero:
                f = M[r][col]
                M[r] = [F.sub(a, F.mul(f, b)) for a, b in zip(M[r], M[col])]
    return [M[i][m] for i in range(m)]


def rs_decode(digits, n, t, q):
    """k (length n) from the 2 t s syndrome digits by PGZ; (k, errors), errors = -1
    and k = 0 when no error pattern of weight <= t in GF(q) has that syndrome."""
    s, nc = rs_params(n, t, q)
    F = OracleGF(q, s)
    S = [F.from_digits_msb(digits[s * j: s * (j + 1)]) for j in range(2 * t)]     # S_1 .. S_2t
    zero_k = np.zeros(n, dtype=np.int64)
    if all(x == F.zero for x in S):
        return zero_k, 0
    for nu in range(t, 0, -1):
        A = [[S[i + j] for j in range(nu)] for i in range(nu)]     # S_{i+j+1}
        rhs = [F.sub(F.zero, S[nu + i]) for i in range(nu)]         # -S_{nu+i+1}
        sol = _gf_solve(F, A, rhs)
        if sol is None:
            continue
        lam = [F.one] + list(reversed(sol))                        # Lambda_1 .. Lambda_nu
        locs = []
        for pos in range(nc):
            xinv = F.alpha_pow(-(nc - 1 - pos))
            acc, pw = F.zero, F.one
            for c in lam:
                acc = F.add(acc, F.mul(c, pw))
                pw = F.mul(pw, xinv)
            if acc == F.zero:
                locs.append(pos)
        if len(locs) != nu:
            return zero_k, -1
        X = [F.alpha_pow(nc - 1 - p) for p in locs]
        V = [[F.power(x, j + 1) for x in X] for j in range(nu)]     # S_j = sum_l e_l X_l^j
        e = _gf_solve(F, V, S[:nu])
        if e is None or any(any(c != 0 for c in v[1:]) for v in e):
            return zero_k, -1
        k = zero_k.copy()
        for p, v in zip(locs, e):
            if p >= nc - n:
                k[p - (nc - n)] = v[0]
        # the solution must reproduce every syndrome
        for j in range(2 * t):
            acc = F.zero
            for x, v in zip(X, e):
                acc = F.add(acc, F.mul(v, F.power(x, j + 1)))
            if acc != S[j]:
                return zero_k, -1
        return k, nu
    return zero_k, -1


def get_D(n, **kw):
    """Delay matrices (qsft/query.py:73-143)."""
    q = kw.get("q")
    source = kw.get("delays_method_source", "random")
    if source == "random":
        D = np.random.choice(q, (kw.get("num_delays"), n))
    elif source == "identity":
        D = np.vstack((np.zeros(n), np.eye(n))).astype(int)
    elif source == "coded":
        D = rs_delay_matrix(n, kw.get("t"), q)
    else:
        raise NotImplementedError(f"delays_method_source={source!r}")
    channel = kw.get("delays_method_channel", "identity")
    if channel == "nso":
        offsets = np.random.choice(q, (kw.get("num_repeat"), n))
        return [(row - D) % q for row in offsets]
    if channel == "identity":
        return [D % q]
    raise NotImplementedError(f"delays_method_channel={channel!r}")


def get_Ms_and_Ds(n, q, **query_args):
    """M list and the shared D list, RNG order M then D (qsft/query.py:182-203)."""
    Ms = get_Ms(n, query_args.get("b"), q, method=query_args.get("query_method"),
                num_to_get=query_args.get("num_subsample"))
    D = get_D(n, q=q, **query_args)
    return Ms, [D for _ in Ms]


def coset_digits(M, D_rows, q, b):
    """Query digits (rows, Q, n) of the cosets {M l + d : l in Z_q^b}
    (qsft/input_signal_subsampled.py:183-206, before the decimal conversion)."""
    L = qary_ints(b, q)
    ML = (M @ L) % q                                        # (n, Q)
    return np.stack([((ML + d[:, None]) % q).T for d in D_rows])


# ---------------------------------------------------------------------------
# Synthetic sparse signal (synt_exp/synt_src/synthetic_signal.py)
# ---------------------------------------------------------------------------


def generate_signal_w(n, q, S, a_min, a_max, max_weight=None):
    """Sparse spectrum, `full=False` branch (synthetic_signal.py:10-38).

    Returns (signal_w dict, locq (n, S), strengths (S,)). Support is not made
    distinct, as in the reference.
    """
    max_weight = n if max_weight is None else max_weight
    if max_weight == n:
        locq = sort_qary_vecs(np.random.randint(q, size=(n, S)).T).T
    else:
        vals = np.random.randint(q - 1, size=(max_weight, S)) + 1
        pos = np.random.choice(a=n, size=(S, max_weight))
        locq = np.zeros((n, S), dtype=int)
        for s in range(S):
            locq[pos[s, :], s] = vals[:, s]
        locq = sort_qary_vecs(locq.T).T
    strengths = random_signal_strength_model(S, a_min, a_max)
    signal_w = dict(zip(list(map(tuple, locq.T)), strengths))
    return signal_w, locq, strengths


def _eval_batch(args):
    """x(m) = exp(m @ (2 pi i locq / q)) @ a for one batch (synthetic_signal.py:100-102)."""
    digits, freq, strengths = args
    return np.exp(digits @ freq) @ strengths


def eval_synthetic(digits, locq, strengths, q, pool=None, batch=10000):
    """Evaluate the synthetic signal at query digits (N, n) in complex128.

    Same expression as synthetic_signal.py:97-118; batches of 10 000 queries
    (np.array_split into len//10000+1 parts), optionally on a process pool.
    """
    freq = 2j * np.pi * np.asarray(locq) / q
    digits = np.asarray(digits)
    parts = np.array_split(digits, len(digits) // batch + 1)
    work = [(p, freq, strengths) for p in parts]
    outs = pool.map(_eval_batch, work) if pool is not None else [_eval_batch(w) for w in work]
    return np.concatenate(outs) if outs else np.zeros(0, complex)


class SyntheticSignal:
    """Restatement of SubsampledSignal + SyntheticSubsampledSignal.

    Construction samples and sub-transforms, as
    qsft/input_signal_subsampled.py:47-155 does in __init__;
    ``get_MDU`` restates input_signal_subsampled.py:225-262 plus the noise
    injection of synthetic_signal.py:120-130.

    ``faithful_indices`` routes the query digits through Python-int decimals
    and back, exactly as the reference does (used to time the CPU baseline);
    the digits are identical either way.
    """

    def __init__(self, n, q, locq, strengths, noise_sd, query_args, signal_w=None,
                 pool=None, faithful_indices=False):
        self.n, self.q = n, q
        self.locq = np.asarray(locq)
        self.strengths = np.asarray(strengths)
        self.noise_sd = noise_sd
        self.signal_w = signal_w
        self.query_args = query_args
        self.b = query_args.get("b")
        self.all_bs = query_args.get("all_bs", [self.b])
        self.num_subsample = query_args.get("num_subsample")
        if "num_repeat" not in query_args:
            query_args["num_repeat"] = 1
        self.num_repeat = query_args.get("num_repeat")
        self.pool = pool
        self.faithful_indices = faithful_indices
        self.Ms, self.Ds = get_Ms_and_Ds(n, q, **query_args)
        self._subsample_qsft()

    def subsample_digits(self, digits):
        return eval_synthetic(digits, self.locq, self.strengths, self.q, pool=self.pool)

    def _subsample_qsft(self):
        """input_signal_subsampled.py:107-155 without the on-disk cache."""
        q, b = self.q, self.b
        Q = q ** b
        self.Us = [[{} for _ in self.Ds[i]] for i in range(len(self.Ms))]
        self.transformTimes = [[{} for _ in self.Ds[i]] for i in range(len(self.Ms))]
        for i in range(len(self.Ms)):
            for j in range(len(self.Ds[i])):
                dig = coset_digits(self.Ms[i], self.Ds[i][j], q, b)   # (rows, Q, n)
                rows = dig.shape[0]
                if self.faithful_indices:
                    dec = [qary_vec_to_dec(dig[r].T, q) for r in range(rows)]
                    dig = np.stack([dec_to_qary_vec(d, q, self.n).T for d in dec])
                samples = np.zeros((rows, Q), dtype=complex)
                if Q > 10000:
                    for r in range(rows):
                        samples[r] = self.subsample_digits(dig[r])
                else:
                    allv = self.subsample_digits(dig.reshape(rows * Q, self.n))
                    for r in range(rows):
                        samples[r] = allv[r * Q:(r + 1) * Q]
                for bb in self.all_bs:
                    t0 = time.time()
                    self.Us[i][j][bb] = [gwht(row[::q ** (b - bb)], q, bb) for row in samples]
                    self.transformTimes[i][j][bb] = time.time() - t0

    def get_source_parity(self):
        """Rows of one delay block (input_signal_subsampled.py:268-269)."""
        return self.Ds[0][0].shape[0]

    def get_MDU(self, ret_num_subsample, ret_num_repeat, b, trans_times=False):
        """Random group/repeat selection, then per-(i, j) transform-domain noise."""
        if not (ret_num_subsample <= self.num_subsample and ret_num_repeat <= self.num_repeat
                and b <= self.b):
            raise ValueError("There are not enough Ms or Ds.")
        sub_idx = np.random.choice(self.num_subsample, ret_num_subsample, replace=False)
        del_idx = np.random.choice(self.num_repeat, ret_num_repeat, replace=False)
        Ms, Ds, Us, Ts = [], [], [], []
        for i in sub_idx:
            Ms.append(self.Ms[i][:, :b])
            Ds.append([self.Ds[i][j] for j in del_idx])
            Us.append([self.Us[i][j][b] for j in del_idx])
            Ts.append([self.transformTimes[i][j][b] for j in del_idx])
        nu = self.noise_sd / np.sqrt(2 * self.q ** b)
        for i in range(len(Us)):
            for j in range(len(Us[i])):
                shape = np.array(Us[i][j]).shape
                noise = np.random.normal(0, nu, size=shape + (2,)).view(complex).reshape(shape)
                Us[i][j] = np.array(Us[i][j]) + noise                # a new array; stored U untouched
        return (Ms, Ds, Us, Ts) if trans_times else (Ms, Ds, Us)


def get_random_subsampled_signal(n, q, noise_sd, sparsity, a_min, a_max, query_args,
                                 max_weight=None, **kw):
    """synthetic_signal.py:70-85: support, strengths, then the signal."""
    signal_w, locq, strengths = generate_signal_w(n, q, sparsity, a_min, a_max, max_weight)
    return SyntheticSignal(n, q, locq, strengths, noise_sd, query_args, signal_w=signal_w, **kw)


# ---------------------------------------------------------------------------
# Singleton detection (qsft/reconstruct.py)
# ---------------------------------------------------------------------------


def detect_noiseless(col, q):
    """k_i = round_half_even(q (angle U_i - angle U_0) / 2 pi) mod q (reconstruct.py:12-31)."""
    ang = np.angle(col)
    v = q * (ang[1:] - ang[0]) / (2 * np.pi)
    return v.round().astype(int) % q


def detect_nso1(col, q, p1):
    """Soft NSO decoding (reconstruct.py:100-113).

    theta_i = angle(mean_r U[r p1] conj(U[r p1 + i])) mod 2 pi, then the first
    root 2 pi t / q (t = 0..q) nearest to theta_i, reduced mod q.
    """
    roots = 2 * np.pi / q * np.arange(q + 1)
    base = col[0::p1]
    k = np.zeros((p1 - 1,), dtype=int)
    for i in range(1, p1):
        z = np.mean(base * np.conjugate(col[i::p1]))
        theta = np.angle(z) % (2 * np.pi)
        k[i - 1] = np.abs(roots - theta).argmin() % q
    return k


def singleton_detection(col, method_source, method_channel, q, source_parity, code=None):
    """Dispatcher (reconstruct.py:132-167): channel decoding, then for coded
    sources the syndrome decoding of reconstruct.py:34-51 (code = (n, t))."""
    if method_channel == "nso":
        k = detect_nso1(col, q, source_parity)
    elif method_channel == "identity":
        k = detect_noiseless(col, q)
    else:
        raise NotImplementedError(f"reconstruct_method_channel={method_channel!r}")
    if method_source == "coded":
        n, t = code
        k, _ = rs_decode(k, n, t, q)
    elif method_source != "identity":
        raise NotImplementedError(f"reconstruct_method_source={method_source!r}")
    return k


# ---------------------------------------------------------------------------
# Peeling decoder (qsft/qsft.py)
# ---------------------------------------------------------------------------


def peel_decode(Ms, Ds, Us, q, n, cutoff, method_source="identity", method_channel="nso",
                source_parity=None, max_iter=15, trace=None, t=None):
    """The detection + peeling loop of QSFT.transform (qsft.py:142-255).

    Ms: list of C (n, b) ints; Ds: list of C (P, n) ints; Us: (C, P, Q)
    complex128 (modified in place). Returns (result list of (k, rho) in append
    order, gwht dict by running mean in append order, iterations run).
    ``trace`` (a list) receives per-iteration (n_singletons, n_multitons, n_balls).
    """
    C = len(Us)
    b = Ms[0].shape[1]
    omega = np.exp(2j * np.pi / q)
    if source_parity is None:
        source_parity = n + 1
    peeling_max = q ** n
    result = []
    it = 0
    cont = True
    num_peeling = 0
    while cont and num_peeling < peeling_max and it < max_iter:
        it += 1
        singletons = {}
        multitons = []
        for i in range(C):
            U, M, D = Us[i], Ms[i], Ds[i]
            for j, col in enumerate(U.T):
                if np.linalg.norm(col) ** 2 > cutoff * len(col):
                    k = singleton_detection(col, method_source, method_channel, q, source_parity, code=(n, t))
                    sig = omega ** (D @ k)
                    rho = np.dot(np.conjugate(sig), col) / D.shape[0]
                    res = col - rho * sig
                    j_digits = dec_to_qary_vec([j], q, b).T[0]
                    match = np.all((M.T @ k) % q == j_digits)
                    if (not match) or np.linalg.norm(res) ** 2 > cutoff * len(col):
                        multitons.append((i, j))
                    else:
                        singletons[(i, j)] = (k, rho)
        if len(multitons) == 0 or len(singletons) == 0:
            cont = False
        balls = set()
        ball_values = {}
        for (i, j) in singletons:
            k, rho = singletons[(i, j)]
            ball = tuple(k)
            balls.add(ball)
            ball_values[ball] = rho
            result.append((k, ball_values[ball]))
        if trace is not None:
            trace.append((len(singletons), len(multitons), len(balls)))
        for ball in balls:
            num_peeling += 1
            kk = np.array(ball)[..., np.newaxis]
            for l, M in enumerate(Ms):
                jl = qary_vec_to_dec(M.T.dot(kk) % q, q)[0]
                sig = omega ** (Ds[l] @ kk)
                Us[l][:, jl] -= np.array(ball_values[ball] * sig.reshape(-1, 1))[:, 0]
    gwht_d, counts = {}, {}
    for k, v in result:
        key = tuple(k)
        if key in counts:
            gwht_d[key] = (gwht_d[key] * counts[key] + v) / (counts[key] + 1)
            counts[key] += 1
        else:
            gwht_d[key] = v
            counts[key] = 1
    return result, gwht_d, it


def qsft_transform(signal, num_subsample, num_repeat, b, method_source="identity",
                   method_channel="nso", report=False, cutoff=None, sort=False, trace=None, t=None):
    """QSFT(...).transform(signal) (qsft.py:47-281) on an oracle signal."""
    q, n = signal.q, signal.n
    Ms, Ds, Us, Ts = signal.get_MDU(num_subsample, num_repeat, b, trans_times=True)
    for i in range(len(Ds)):
        Us[i] = np.vstack(Us[i])
        Ds[i] = np.vstack(Ds[i])
    transform_time = np.sum(Ts)
    Us = np.array(Us)
    if cutoff is None:
        cutoff = 1e-9 + 1.5 * (signal.noise_sd ** 2) / (q ** b)
    t0 = time.time()
    _, gw, _ = peel_decode(Ms, Ds, Us, q, n, cutoff, method_source, method_channel,
                           signal.get_source_parity(), trace=trace,
                           t=t if t is not None else signal.query_args.get("t"))
    peel_time = time.time() - t0
    if not report:
        return gw
    loc = list(gw.keys())
    if loc:
        if sort:
            loc = sort_qary_vecs(loc)                        # an ndarray, as qsft.py:268
        hw = calc_hamming_weight(loc)
        avg_hw, max_hw = np.mean(hw), np.max(hw)
    else:
        loc, avg_hw, max_hw = [], 0, 0
    return {"gwht": gw, "runtime": transform_time + peel_time,
            "n_samples": int(np.prod(Us.shape)), "locations": loc,
            "avg_hamming_weight": avg_hw, "max_hamming_weight": max_hw}


# ---------------------------------------------------------------------------
# Independent references used to pin the restatement
# ---------------------------------------------------------------------------


def closed_form_U(M, D_rows, locq, strengths, q):
    """U_{p}[j] = sum_{s: M^T k_s == j mod q} a_s w^{<k_s, d_p>} for one group.

    This is the algebraic identity fftn((M l + d) samples)/q^b satisfies; it
    is derived from the DFT definition, not from the reference code.
    """
    locq = np.asarray(locq)
    b = M.shape[1]
    Q = q ** b
    jvec = (M.T @ locq) % q                                  # (b, S)
    jdec = (q ** np.arange(b - 1, -1, -1)) @ jvec            # (S,)
    out = np.zeros((len(D_rows), Q), dtype=complex)
    for p, d in enumerate(D_rows):
        ph = np.exp(2j * np.pi * ((d @ locq) % q) / q)
        np.add.at(out[p], jdec, strengths * ph)
    return out


def dense_time_signal(locq, strengths, q, n):
    """Full q^n time signal x[m] = sum_s a_s w^{<m,k_s>} by inverse DFT (small n only)."""
    spec = np.zeros([q] * n, dtype=complex)
    for s in range(np.asarray(locq).shape[1]):
        spec[tuple(np.asarray(locq)[:, s])] += strengths[s]
    return scipy.fft.ifftn(spec) * (q ** n)


def nmse(gwht_est, signal_w):
    """sum |w - w_hat|^2 / sum |w|^2 over the union of supports (quick_example.py:76-80)."""
    diff = dict(signal_w)
    for k, v in gwht_est.items():
        diff[k] = diff.get(k, 0) - v
    num = np.sum(np.abs(np.array(list(diff.values()))) ** 2)
    den = np.sum(np.abs(np.array(list(signal_w.values()))) ** 2)
    return float(num / den)


def heldout_nmse(sample_idx_dec, samples, beta, q, n, batch=10000):
    """Held-out NMSE of a recovered spectrum (qsft/test_helper.py:235-260,
    TestHelper._test_qary): predict every test sample as
    sum_k beta[k] exp(2 pi i <m, k> / q) with the unreduced integer phase, in
    batches of 10000 queries, and return ||y_hat - y||^2 / ||y||^2; an empty
    beta gives 1."""
    if len(beta.keys()) == 0:
        return 1
    keys = np.array(list(beta.keys()))
    vals = np.array(list(beta.values()))
    idx = list(sample_idx_dec)
    y_hat = []
    for i in range(0, len(idx), batch):
        m = dec_to_qary_vec(idx[i:i + batch], q, n)
        freqs = np.array(m).T @ keys.T
        y_hat.append(np.exp(2j * np.pi * freqs / q) @ vals)
    y_hat = np.concatenate(y_hat)
    samples = np.asarray(list(samples))
    return np.linalg.norm(y_hat - samples) ** 2 / np.linalg.norm(samples) ** 2


def cpu_count():
    return os.cpu_count() or 1
