"""Host-side serving-loop helpers that need no GPU."""
import gc

from qsft_amd import serving as SV


def test_gc_paused_disables_and_restores():
    assert gc.isenabled()
    with SV.gc_paused():
        assert not gc.isenabled()
        with SV.gc_paused():                                       # nested: the outer state comes back
            assert not gc.isenabled()
        assert not gc.isenabled()
    assert gc.isenabled()
    gc.disable()
    try:
        with SV.gc_paused():
            assert not gc.isenabled()
        assert not gc.isenabled()                                  # it was off before: stays off
    finally:
        gc.enable()


def test_collect_young_keeps_generation_zero_small():
    with SV.gc_paused():
        junk = [[i] for i in range(3 * SV.GC_YOUNG)]              # tracked container objects
        assert gc.get_count()[0] > SV.GC_YOUNG
        SV.collect_young()
        assert gc.get_count()[0] <= SV.GC_YOUNG
        del junk
