"""Host-side q-ary helpers with the reference's names (qsft/utils.py).

These are the small integer/ordering utilities of the public API; the
numeric work of the transform runs in the HIP library. `gwht` here runs the
device DFT kernel.
"""
import json

import numpy as np


class NpEncoder(json.JSONEncoder):
    """JSON encoder for numpy scalars and arrays (qsft/utils.py:199-207)."""

    def default(self, obj):
        if isinstance(obj, np.integer):
            return int(obj)
        if isinstance(obj, np.floating):
            return float(obj)
        if isinstance(obj, np.ndarray):
            return obj.tolist()
        return super().default(obj)


def qary_ints(m, q, dtype=int):
    """(m, q^m) array whose column c holds the MSB-first digits of c (qsft/utils.py:107-108)."""
    cols = np.arange(q ** m, dtype=np.int64)
    place = q ** np.arange(m - 1, -1, -1, dtype=np.int64)
    return ((cols[None, :] // place[:, None]) % q).astype(dtype)


def qary_vec_to_dec(x, q):
    """Digit columns (n, N) -> exact Python-int indices, MSB-first (qsft/utils.py:74-76)."""
    x = np.asarray(x)
    n = x.shape[0]
    weights = np.array([q ** (n - 1 - i) for i in range(n)], dtype=object)
    return weights @ np.array(x, dtype=object)


def dec_to_qary_vec(x, q, n):
    """Python-int indices -> digit columns (n, N) int, MSB-first (qsft/utils.py:79-84)."""
    rest = np.array(list(x), dtype=object)
    rows = []
    for i in range(n):
        place = q ** (n - 1 - i)
        digit = np.array([v // place for v in rest], dtype=object)
        rest = rest - place * digit
        rows.append(digit)
    return np.array(rows, dtype=int).reshape(n, len(rest))


def sort_qary_vecs(qary_vecs):
    """Rows sorted lexicographically, first digit most significant (qsft/utils.py:180-183)."""
    v = np.array(qary_vecs)
    return v[np.lexsort(v.T[::-1, :])]


def calc_hamming_weight(qary_vecs):
    """Nonzero digits per row (qsft/utils.py:185-187)."""
    return np.sum(np.array(qary_vecs) != 0, axis=1)


def random_signal_strength_model(sparsity, a, b):
    """Magnitudes U[a, b], phases U[0, 2 pi) from the global RNG (qsft/utils.py:161-164)."""
    magnitude = np.random.uniform(a, b, sparsity)
    phase = np.random.uniform(0, 2 * np.pi, sparsity)
    return magnitude * np.exp(1j * phase)


def gwht(x, q, n, precision="fp64"):
    """Forward-scaled n-dim q-point DFT of a flat length-q^n signal, on the GPU
    (qsft/utils.py:31-36). Returns a numpy array."""
    import torch
    from . import device as dv
    dt = torch.complex128 if precision == "fp64" else torch.complex64
    xt = torch.as_tensor(np.asarray(x).reshape(1, q ** n), dtype=dt).to(dv.current_device())
    out = dv.subtransform(xt, q, n, precision=precision)
    return out.reshape(q ** n).cpu().numpy()
