"""qsft.query (qsft/query.py): M_c, delay matrices, the Reed-Solomon decoder factory."""
from qsft_amd.query import *  # noqa: F401,F403
from qsft_amd.query import get_reed_solomon_dec  # noqa: F401
