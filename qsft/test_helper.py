"""qsft.test_helper (qsft/test_helper.py): the experiment helper."""
from qsft_amd.test_helper import TestHelper  # noqa: F401
