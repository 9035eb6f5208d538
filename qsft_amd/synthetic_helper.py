"""SyntheticHelper (synt_exp/synt_src/synthetic_helper.py:4-6): a TestHelper
whose signals are synthetic subsampled signals."""
from .synthetic_signal import SyntheticSubsampledSignal
from .test_helper import TestHelper


class SyntheticHelper(TestHelper):
    __test__ = False

    def generate_signal(self, signal_args):
        return SyntheticSubsampledSignal(**signal_args)
