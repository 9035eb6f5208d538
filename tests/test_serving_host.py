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


def test_arena_reuses_and_grows_slot_buffers():
    import torch
    arena = {}
    a = SV._arena_take(arena, "work", 1000, torch.device("cpu"))
    assert a.numel() == 1000 and arena["work"].numel() >= 1000
    base = arena["work"].data_ptr()
    b = SV._arena_take(arena, "work", 800, torch.device("cpu"))       # fits: the same storage, a view
    assert b.numel() == 800 and b.data_ptr() == base
    c = SV._arena_take(arena, "work", 5000, torch.device("cpu"))      # too small: replaced, with headroom
    assert c.numel() == 5000 and arena["work"].numel() >= 5000 + 5000 // 4
    d = SV._arena_take(None, "work", 10, torch.device("cpu"))         # no arena: a fresh buffer of at least 256 B
    assert d.numel() == 256
