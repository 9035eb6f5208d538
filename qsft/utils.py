"""qsft.utils (qsft/utils.py): q-ary index helpers."""
from qsft_amd.utils import *  # noqa: F401,F403
