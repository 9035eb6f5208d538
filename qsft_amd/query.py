"""Subsampling matrices M_c and delay matrices D (qsft/query.py).

Host-side parameter generation with the reference's legacy `np.random` call
order, so a seed yields the same M and D as the reference: C calls of
randint(q, (n, b)) for "complex" M (query.py:26-35), then choice(q, (R, n))
for NSO offsets (query.py:95-106) or choice(q, (num_delays, n)) for random
source delays (query.py:79-85).
"""
import numpy as np

from .reed_solomon import ReedSolomon


def get_Ms_simple(n, b, q, num_to_get=None):
    """Block identities, filled from the last group to the first (query.py:13-23)."""
    Ms = []
    for blk in range(num_to_get - 1, -1, -1):
        M = np.zeros((n, b), dtype=np.int32)
        M[b * blk:b * (blk + 1), :] = np.eye(b)
        Ms.append(M)
    return Ms


def get_Ms_complex(n, b, q, num_to_get=None):
    """Uniform M_c in Z_q^{n x b} (query.py:26-35)."""
    return [np.random.randint(q, size=(n, b)) for _ in range(num_to_get)]


def get_Ms(n, b, q, num_to_get=None, method="simple"):
    """query.py:38-70."""
    if num_to_get is None:
        num_to_get = max(n // b, 3)
    if method == "simple" and num_to_get > n // b:
        raise ValueError("When query_method is 'simple', the number of M matrices to return cannot be "
                         "larger than n // b")
    makers = {"simple": get_Ms_simple, "complex": get_Ms_complex}
    if method not in makers:
        raise NotImplementedError(f"query_method={method!r}")
    return makers[method](n, b, q, num_to_get)


def get_D_identity(n, **kwargs):
    """[0; I], shape (n + 1, n) (query.py:73-76)."""
    return np.vstack((np.zeros(n), np.eye(n))).astype(int)


def get_D_random(n, **kwargs):
    """query.py:79-85."""
    return np.random.choice(kwargs.get("q"), (kwargs.get("num_delays"), n))


def get_D_source_coded(n, **kwargs):
    """Reed-Solomon parity-check delays, (2 t s + 1) x n (query.py:88-92; see reed_solomon)."""
    return np.array(ReedSolomon(n, kwargs.get("t"), kwargs.get("q")).get_delay_matrix(), dtype=int)


def get_D_nso(n, D_source, **kwargs):
    """One block (r - D_source) mod q per random offset r (query.py:95-106)."""
    q = kwargs.get("q")
    offsets = get_D_random(n, q=q, num_delays=kwargs.get("num_repeat"))
    return [(r - D_source) % q for r in offsets]


def get_D_channel_coded(n, D, **kwargs):
    raise NotImplementedError("One day this might be implemented")


def get_D_channel_identity(n, D, **kwargs):
    return [D % kwargs.get("q")]


def get_D(n, **kwargs):
    """Source delays, then channel expansion (query.py:116-143)."""
    src = kwargs.get("delays_method_source", "random")
    makers = {"random": get_D_random, "identity": get_D_identity, "coded": get_D_source_coded}
    if src not in makers:
        raise NotImplementedError(f"delays_method_source={src!r}")
    D = makers[src](n, **kwargs)
    chan = kwargs.get("delays_method_channel", "identity")
    expand = {"nso": get_D_nso, "coded": get_D_channel_coded, "identity": get_D_channel_identity}
    if chan not in expand:
        raise NotImplementedError(f"delays_method_channel={chan!r}")
    return expand[chan](n, D, **kwargs)


def get_Ms_and_Ds(n, q, **kwargs):
    """M list, then one D list shared by every group (query.py:182-203)."""
    Ms = get_Ms(n, kwargs.get("b"), q, method=kwargs.get("query_method"),
                num_to_get=kwargs.get("num_subsample"))
    D = get_D(n, q=q, **kwargs)
    return Ms, [D for _ in Ms]


def get_reed_solomon_dec(n, t_max, q):
    """A Reed-Solomon syndrome decoder for a t_max error-correcting code (query.py:231-243):
    the bound `syndrome_decode` of the code, as the reference returns; QSFT
    reads the code from it and decodes on the device."""
    if q in [2, 3, 5, 7, 11, 13, 15, 17, 19, 23, 29]:
        return ReedSolomon(n, t_max, q).syndrome_decode
    raise NotImplementedError("q is not a prime number under 30!")
