"""Device-side building blocks shared by the drop-in classes and the engine.

Layouts in HBM (all row-major, int8 digits padded to n_pad = 64*ceil(n/64)):
  Mt  int8 [C][b][n_pad]      Mt[c][t][i] = M_c[i][t]
  D   int8 [C][P][n_pad]      delay rows of every group (P = repeats * (n+1))
  K   int8 [S][n_pad]         support digits
  a   complex64/128 [S]       strengths
  x   complex64/128 [C][P][Q] samples, l MSB-first
  U   complex64/128 [C][P][Q] sub-transforms (the peeler works on a complex128 copy)
"""
import numpy as np
import torch

from . import _lib as L
from . import kernels as K

PRECISIONS = {"fp32": (L.QSFT_F32, torch.complex64), "fp64": (L.QSFT_F64, torch.complex128)}


def current_device():
    if not torch.cuda.is_available():
        raise RuntimeError("qsft_amd runs on a ROCm GPU (gfx950) and none is visible")
    return torch.device("cuda", torch.cuda.current_device())


def n_pad_of(n):
    return (n + 63) // 64 * 64


def digits_tensor(rows, n_pad, device):
    """int digit rows (R, n) -> int8 [R][n_pad] on the device (zero padded)."""
    rows = np.asarray(rows)
    out = np.zeros((rows.shape[0], n_pad), np.int8)
    out[:, :rows.shape[1]] = rows
    return torch.from_numpy(out).to(device)


def Mt_tensor(Ms, b, n_pad, device):
    """List of C (n, >=b) matrices -> Mt [C][b][n_pad]."""
    C = len(Ms)
    n = Ms[0].shape[0]
    out = np.zeros((C, b, n_pad), np.int8)
    for c, M in enumerate(Ms):
        out[c, :, :n] = np.asarray(M)[:, :b].T
    return torch.from_numpy(out).to(device)


def D_tensor(Ds, n_pad, device):
    """List of C arrays (P, n) -> D [C][P][n_pad]."""
    Ds = [np.asarray(D) for D in Ds]
    C, (P, n) = len(Ds), Ds[0].shape
    out = np.zeros((C, P, n_pad), np.int8)
    for c, D in enumerate(Ds):
        out[c, :, :n] = D
    return torch.from_numpy(out).to(device)


def strengths_tensor(a, precision, device):
    dt = PRECISIONS[precision][1]
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.complex128))).to(device).to(dt)


def evaluate_cosets(q, n, b, Mt, D, Kd, a, precision):
    """x[c][p][l] = sum_s a_s w^{<(M_c l + d_cp) mod q, k_s>} -> [C][P][q^b]."""
    C, P = Mt.shape[0], D.shape[1]
    code, dt = PRECISIONS[precision]
    x = torch.empty((C, P, q ** b), dtype=dt, device=Mt.device)
    K.eval_cosets([dict(q=q, n=n, b=b, C=C, P=P, S=Kd.shape[0], Mt=Mt, D=D, K=Kd, a=a, x=x)], code)
    return x


def evaluate_cosets_mfma(q, n, b, Mt, D, Kd, a64, dense=False):
    """The same samples through the matrix-core kernel -> complex64 [C][P][q^b]
    (dense=True: the dense-MFMA form for q in {2, 4})."""
    C, P = Mt.shape[0], D.shape[1]
    x = torch.empty((C, P, q ** b), dtype=torch.complex64, device=Mt.device)
    ev = K.EvalMfma([dict(q=q, n=n, b=b, C=C, P=P, S=Kd.shape[0], Mt=Mt, D=D, K=Kd, a64=a64, x=x)], dense=dense)
    ev.launch()
    torch.cuda.current_stream().synchronize()      # the B-fragment workspace is released with `ev`
    return x


def query_digits(q, n, b, Mt, D):
    """Digits of every coset query -> int8 [C][P][q^b][n_pad] (stage (a) kernel)."""
    C, P, n_pad = Mt.shape[0], D.shape[1], Mt.shape[-1]
    m = torch.empty((C, P, q ** b, n_pad), dtype=torch.int8, device=Mt.device)
    K.query_digits([dict(q=q, n=n, b=b, C=C, P=P, Mt=Mt, D=D, m=m)])
    return m


def evaluate_points(q, n, m, Kd, a, precision):
    """x[i] = sum_s a_s w^{<m_i, k_s>} at explicit digits m [N][n_pad]."""
    code, dt = PRECISIONS[precision]
    x = torch.empty((m.shape[0],), dtype=dt, device=m.device)
    if m.shape[0]:
        K.eval_points([dict(q=q, n=n, S=Kd.shape[0], m=m, K=Kd, a=a, x=x)], code)
    return x


def subtransform(x, q, b, in_stride=1, precision="fp64"):
    """fftn over [q]*b of every row of x ([rows][Q_row], taking every in_stride-th
    sample), scaled by q^-b -> [rows][q^b]."""
    code, dt = PRECISIONS[precision]
    rows = x.shape[0]
    xr = x.reshape(rows, -1)
    if xr.dtype != dt:
        xr = xr.to(dt)
    xr = xr.contiguous()
    out = torch.empty((rows, q ** b), dtype=dt, device=x.device)
    K.dft([dict(q=q, b=b, rows=rows, inp=xr, in_stride=in_stride, in_row_stride=xr.shape[1], out=out)], code)
    return out
