"""Reed-Solomon source code for coded delays (qsft/ReedSolomon.py,
qsft/query.py:88-92,231-243, qsft/reconstruct.py:34-51).

The reference builds its code with the `galois` package: a Reed-Solomon code
of length n_c = q^s - 1 over GF(q^s) (s = ceil(log_q n), one more if
n > q^s - 1) with 2t parity symbols, shortened to the last n positions. Its
delay matrix has a zero first row and then, for each of the 2t parity-check
rows j, the s GF(q) coefficients (highest degree first) of H[j, i] for each
of the n kept positions i. A channel decoder that returns the p = 2ts
delay-row phases of a singleton returns exactly the syndrome of its index k
viewed as an error pattern, so syndrome decoding recovers every k of Hamming
weight <= t.

This module restates that construction without galois (absent here):

  * GF(q^s) as GF(q)[x] / f(x), f the lexicographically smallest primitive
    polynomial of degree s (galois's default field polynomial is the Conway
    polynomial, which can differ: the delay matrix is then a different but
    equally valid parity-check matrix of the same code family);
  * H[j, i] = alpha^{(1 + j)(n - 1 - i)}, alpha = x (narrow-sense roots
    alpha^1 .. alpha^{2t}, codeword coefficients highest degree first, the
    shortened code keeping the last n positions);
  * syndrome decoding by Berlekamp-Massey (error locator), a Chien search over
    all n_c positions and Forney's formula for the error values; a decode is a
    failure -- k = 0, the received all-zero word unchanged -- unless the
    locator has exactly deg(Lambda) roots and every error value lies in GF(q).

The same tables drive the device decoder inside the peeling kernels.
"""
import math

import numpy as np

PRIMES = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29)


def _is_primitive(coeffs, q):
    """coeffs: monic f of degree s, coefficients c_0 .. c_{s-1} (low first).
    True if x has multiplicative order q^s - 1 modulo f."""
    s = len(coeffs)
    order = q ** s - 1
    v = [0] * s
    v[0] = 1                                                   # the polynomial 1
    for k in range(1, order + 1):
        # v <- v * x mod f
        top = v[s - 1]
        v = [0] + v[:s - 1]
        if top:
            v = [(v[i] - top * coeffs[i]) % q for i in range(s)]
        if v[0] == 1 and not any(v[1:]):
            return k == order
    return False


def primitive_poly(q, s):
    """Lexicographically smallest monic primitive polynomial of degree s over
    GF(q), as c_0 .. c_{s-1} (low degree first; x^s is implicit). Candidates
    are ordered by the integer with digits c_{s-1} .. c_0 (highest first)."""
    for v in range(q ** s):
        coeffs = [(v // q ** i) % q for i in range(s)]         # c_0 = lowest digit
        if coeffs[0] == 0:
            continue
        if _is_primitive(coeffs, q):
            return coeffs
    raise ValueError(f"no primitive polynomial of degree {s} over GF({q})")


class GF:
    """GF(q^s): elements are integers whose base-q digits (highest degree
    first) are the polynomial coefficients; exp/log tables of alpha = x."""

    def __init__(self, q, s):
        self.q, self.s = q, s
        self.order = q ** s
        self.poly = primitive_poly(q, s)
        n1 = self.order - 1
        self.exp = np.zeros(2 * n1, dtype=np.int64)
        self.log = np.full(self.order, -1, dtype=np.int64)
        v = [0] * s
        v[0] = 1
        for k in range(n1):
            val = sum(v[i] * q ** i for i in range(s))
            self.exp[k] = val
            self.log[val] = k
            top = v[s - 1]
            v = [0] + v[:s - 1]
            if top:
                v = [(v[i] - top * self.poly[i]) % q for i in range(s)]
        self.exp[n1:] = self.exp[:n1]
        self.digits_tab = np.array([[(e // q ** (s - 1 - r)) % q for r in range(s)] for e in range(self.order)],
                                   dtype=np.int64)

    def add(self, a, b):
        q = self.q
        r, p = 0, 1
        for _ in range(self.s):
            r += ((a % q + b % q) % q) * p
            a //= q
            b //= q
            p *= q
        return r

    def neg(self, a):
        q = self.q
        r, p = 0, 1
        for _ in range(self.s):
            r += ((-(a % q)) % q) * p
            a //= q
            p *= q
        return r

    def sub(self, a, b):
        return self.add(a, self.neg(b))

    def mul(self, a, b):
        if a == 0 or b == 0:
            return 0
        return int(self.exp[self.log[a] + self.log[b]])

    def inv(self, a):
        if a == 0:
            raise ZeroDivisionError("inverse of 0 in GF(q^s)")
        return int(self.exp[(self.order - 1 - self.log[a]) % (self.order - 1)])

    def pow_alpha(self, e):
        return int(self.exp[e % (self.order - 1)])

    def digits(self, a):
        """Coefficients of a, highest degree first (galois's vector order)."""
        return self.digits_tab[a]

    def from_digits(self, d):
        v = 0
        for x in d:
            v = v * self.q + int(x)
        return v

    def scalar(self, c):
        """The prime-field element c (0 <= c < q) as a field element."""
        return int(c) % self.q


class ReedSolomon:
    """The coded-source construction of qsft/ReedSolomon.py:18-25 (s, n_c, t)."""

    def __init__(self, n, t, q):
        if q not in PRIMES:
            raise NotImplementedError("q is not a prime number under 30!")
        self.q, self.t, self.ns = q, t, n
        s = math.ceil(math.log(n) / math.log(q))
        if n > q ** s - 1:
            s += 1
        self.s = s
        self.n = q ** s - 1                                    # full code length n_c
        self.k = self.n - 2 * t
        if self.k < 1:
            raise ValueError(f"t = {t} leaves no message symbols for q^s - 1 = {self.n}")
        self.field = GF(q, s)
        self.prime_field = q

    def get_parity_length(self):
        """p = 2 t s parity digits (qsft/ReedSolomon.py:67-74)."""
        return 2 * self.t * self.s

    def h_exponent(self, j, i):
        """H[j, i] = alpha^{(1 + j)(n - 1 - i)} for kept position i (0 .. n-1)."""
        return ((1 + j) * (self.ns - 1 - i)) % self.n

    def get_delay_matrix(self):
        """(p + 1) x n delay rows over GF(q): row 0 zero, rows s j + 1 .. s (j + 1)
        the digits (highest degree first) of H[j, i] (qsft/ReedSolomon.py:52-65)."""
        p, s, F = self.get_parity_length(), self.s, self.field
        D = np.zeros((p + 1, self.ns), dtype=np.int64)
        for j in range(2 * self.t):
            for i in range(self.ns):
                D[s * j + 1: s * (j + 1) + 1, i] = F.digits(F.pow_alpha(self.h_exponent(j, i)))
        return D

    def syndromes(self, digits):
        """The 2t syndromes S_1 .. S_2t from the p channel digits (s per symbol)."""
        s, F = self.s, self.field
        return [F.from_digits(digits[s * j: s * (j + 1)]) for j in range(2 * self.t)]

    def decode(self, digits):
        """Source digits k (length n) from the p channel digits, and the number
        of errors (-1: decoding failure, k = 0)."""
        F, t = self.field, self.t
        S = self.syndromes(digits)
        if not any(S):
            return np.zeros(self.ns, dtype=np.int64), 0
        # Berlekamp-Massey: Lambda(x) = 1 + l_1 x + ..., S_m = -sum_i l_i S_{m-i}
        lam = [1] + [0] * (2 * t)
        bpoly = [1] + [0] * (2 * t)
        L, m, bdisc = 0, 1, 1
        for r in range(2 * t):
            disc = S[r]
            for i in range(1, L + 1):
                disc = F.add(disc, F.mul(lam[i], S[r - i]))
            if disc == 0:
                m += 1
                continue
            coef = F.mul(disc, F.inv(bdisc))
            new = list(lam)
            for i in range(2 * t + 1 - m):
                new[i + m] = F.sub(new[i + m], F.mul(coef, bpoly[i]))
            if 2 * L <= r:
                bpoly, L, bdisc, m = lam, r + 1 - L, disc, 1
            else:
                m += 1
            lam = new
        if L > t or any(lam[L + 1:]) or lam[L] == 0:
            return np.zeros(self.ns, dtype=np.int64), -1
        # Omega = S(x) Lambda(x) mod x^2t, S(x) = sum_m S_m x^(m-1)
        omega = [0] * (2 * t)
        for a in range(2 * t):
            for b in range(L + 1):
                if a + b < 2 * t:
                    omega[a + b] = F.add(omega[a + b], F.mul(S[a], lam[b]))

        def ev(poly, x):
            acc, p = 0, 1
            for c in poly:
                acc = F.add(acc, F.mul(c, p))
                p = F.mul(p, x)
            return acc
        dlam = [F.mul(F.scalar(i % self.q), lam[i]) for i in range(1, L + 1)]     # Lambda'(x) coefficients
        k = np.zeros(self.ns, dtype=np.int64)
        roots = 0
        for pos in range(self.n):                              # Chien search over all code positions
            xinv = F.pow_alpha(-(self.n - 1 - pos))            # X^{-1}, X = alpha^{n_c - 1 - pos}
            if ev(lam, xinv) != 0:
                continue
            roots += 1
            den = ev(dlam, xinv)
            if den == 0:
                return np.zeros(self.ns, dtype=np.int64), -1
            e = F.neg(F.mul(ev(omega, xinv), F.inv(den)))
            if e >= self.q:                                    # not in GF(q): no source digit
                return np.zeros(self.ns, dtype=np.int64), -1
            i = pos - (self.n - self.ns)
            if i >= 0:
                k[i] = e
        if roots != L:
            return np.zeros(self.ns, dtype=np.int64), -1
        return k, L

    def syndrome_decode(self, syndrome):
        """Decoder callable of get_reed_solomon_dec: (k as a (1, n) array, number of errors)."""
        k, nerr = self.decode(np.asarray(syndrome, dtype=np.int64))
        return k.reshape(1, -1), nerr

    def device_tables(self):
        """int16 tables for the device decoder: exp (2 (q^s - 1) entries), log
        (q^s entries, log 0 = -1), concatenated."""
        return np.concatenate([self.field.exp, self.field.log]).astype(np.int16)
