"""qsft_amd: MI355X (gfx950) engine for sparse q-ary Fourier transforms.

Drop-in for the public API of basics-lab/qsft:
    from qsft_amd.qsft import QSFT
    from qsft_amd.input_signal_subsampled import SubsampledSignal
    from qsft_amd.synthetic_signal import get_random_subsampled_signal
The numeric stages run in hand-written HIP kernels (lib/libqsft_amd.so, built
by `python -m qsft_amd.build`). Every compute module imports qsft_amd._lib,
which loads that library and raises ImportError if it is missing; the
package itself imports nothing so the build step can run before a library
exists.
"""

__all__ = ["qsft", "input_signal_subsampled", "synthetic_signal", "query", "utils", "engine", "serving",
           "reed_solomon"]
