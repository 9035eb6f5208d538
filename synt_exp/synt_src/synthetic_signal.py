"""synt_exp.synt_src.synthetic_signal (synt_exp/synt_src/synthetic_signal.py)."""
from qsft_amd.synthetic_signal import (SyntheticSignal, SyntheticSubsampledSignal,  # noqa: F401
                                       generate_signal_w, get_random_subsampled_signal)
