"""qsft.qsft (qsft/qsft.py): the QSFT decoder."""
from qsft_amd.qsft import QSFT  # noqa: F401
