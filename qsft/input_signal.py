"""qsft.input_signal (qsft/input_signal.py): the Signal base class."""
from qsft_amd.input_signal import Signal  # noqa: F401
