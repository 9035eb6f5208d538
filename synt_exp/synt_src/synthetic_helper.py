"""synt_exp.synt_src.synthetic_helper (synt_exp/synt_src/synthetic_helper.py)."""
from qsft_amd.synthetic_helper import SyntheticHelper  # noqa: F401
