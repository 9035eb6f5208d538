"""qsft.input_signal_subsampled (qsft/input_signal_subsampled.py): the plugin base class."""
from qsft_amd.input_signal_subsampled import SubsampledSignal  # noqa: F401
