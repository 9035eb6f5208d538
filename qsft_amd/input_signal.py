"""Signal base class with the reference's interface (qsft/input_signal.py:9-96).

The dense full-signal mode (signal_t / calc_w) is outside the accelerated
path; only the constructor protocol (_set_params, then _init_signal) and the
shape helpers are kept for subclasses.
"""
import numpy as np


class Signal:
    def __init__(self, **kwargs):
        self._set_params(**kwargs)
        self._init_signal()

    def _set_params(self, **kwargs):
        self.n = kwargs.get("n")
        self.q = kwargs.get("q")
        self.noise_sd = kwargs.get("noise_sd", 0)
        self.N = self.q ** self.n
        self.signal_t = kwargs.get("signal_t")
        self.signal_w = kwargs.get("signal_w")
        self.calc_w = kwargs.get("calc_w", False)
        self.foldername = kwargs.get("folder")
        self.is_synt = False

    def _init_signal(self):
        if self.signal_t is None:
            self.sample()

    def sample(self):
        raise NotImplementedError

    def shape(self):
        return tuple([self.q] * self.n)

    def get_time_domain(self, inds):
        inds = np.array(inds)
        if inds.ndim == 3:
            return [self.signal_t[tuple(i)] for i in inds]
        return self.signal_t[tuple(inds)]
