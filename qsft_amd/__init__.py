"""qsft_amd: MI355X (gfx950) engine for sparse q-ary Fourier transforms.

Drop-in for the public API of basics-lab/qsft:
    from qsft_amd.qsft import QSFT
    from qsft_amd.input_signal_subsampled import SubsampledSignal
    from qsft_amd.synthetic_signal import get_random_subsampled_signal
The numeric stages run in hand-written HIP kernels (lib/libqsft_amd.so, built
by qsft_amd.build); importing the package fails if the library is missing.
"""
from . import _lib  # noqa: F401  (loads libqsft_amd.so; raises if absent)

__all__ = ["qsft", "input_signal_subsampled", "synthetic_signal", "query", "utils"]
