"""Batched engine on the GPU: uneven instances in one set of launches give the
same result as each instance alone, and agree with the drop-in API path."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
    from qsft_amd import engine as E
    from qsft_amd import kernels as K


def _same(a, b, rtol):
    ka, va = a
    kb, vb = b
    da = {tuple(r): v for r, v in zip(ka, va)}
    db = {tuple(r): v for r, v in zip(kb, vb)}
    assert set(da) == set(db)
    for k, v in da.items():
        assert abs(v - db[k]) <= rtol * abs(v), k


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_uneven_batch_equals_alone(precision):
    rng = np.random.default_rng(5)
    shapes = [(4, 50, 300, 5, 3, 2, 20.0), (20, 12, 150, 2, 3, 2, 25.0), (2, 40, 60, 7, 3, 2, None),
              (3, 16, 80, 4, 2, 3, 30.0), (4, 10, 100, 4, 3, 1, None)]
    specs = []
    for q, n, S, b, C, R, snr in shapes:
        method = "identity" if R == 1 else "nso"
        specs.append(E.random_spec(q, n, S, b, C, R, snr, rng=rng, method=method))
    print("batch of", len(specs), precision, flush=True)
    mix = E.Batch(specs, precision=precision)
    mix.launch(seed_offset=7)
    got = mix.results()
    for i, sp in enumerate(specs):
        alone = E.Batch([sp], precision=precision)
        alone.launch(seed_offset=7)
        (k, v, _, _), = alone.results()
        # FP64 peel atomics may add in another order when two balls share a bin
        _same((got[i][0], got[i][1]), (k, v), rtol=1e-12)
        if sp.noise_sd == 0:
            assert len(k) == sp.S and E.nmse(k, v, sp) < 1e-10


def test_relaunch_gives_fresh_noise_and_same_support_when_noiseless():
    rng = np.random.default_rng(6)
    sp_noisy = E.random_spec(4, 30, 200, 4, 3, 2, 15.0, rng=rng)
    sp_clean = E.random_spec(4, 30, 200, 4, 3, 2, None, rng=rng)
    bt = E.Batch([sp_noisy, sp_clean], precision="fp32")
    bt.launch(seed_offset=1)
    r1 = bt.results()
    bt.launch(seed_offset=2)
    r2 = bt.results()
    bt.launch(seed_offset=1)
    r3 = bt.results()
    v1 = dict(zip(map(tuple, r1[0][0]), r1[0][1]))
    v2 = dict(zip(map(tuple, r2[0][0]), r2[0][1]))
    v3 = dict(zip(map(tuple, r3[0][0]), r3[0][1]))
    common = set(v1) & set(v2)
    assert common and any(v1[k] != v2[k] for k in common)       # fresh noise per launch
    assert set(v1) == set(v3) and all(abs(v1[k] - v3[k]) <= 1e-12 * abs(v1[k]) for k in v1)   # same seed
    assert set(map(tuple, r1[1][0])) == {tuple(c) for c in sp_clean.locq.T}


def test_fused_screen_matches_streaming_screen():
    """Energies written by the assembler drive the same first-pass screen as
    re-reading U: same support and values, transform by transform."""
    rng = np.random.default_rng(8)
    specs = [E.random_spec(4, 50, 400, 5, 3, 2, 10.0, rng=rng), E.random_spec(3, 20, 100, 4, 3, 2, 20.0, rng=rng),
             E.random_spec(4, 64, 300, 5, 3, 1, None, rng=rng, method="identity"),
             E.random_spec(2, 70, 50, 7, 3, 2, 15.0, rng=rng)]          # P = 142 > 128: streaming screen
    print("fused vs streaming screen", flush=True)
    fused = E.Batch(specs, precision="fp32", fused_screen=True)
    plain = E.Batch(specs, precision="fp32", fused_screen=False)
    assert fused.bufs[0]["energy"] is not None and fused.bufs[3]["energy"] is None
    fused.launch(seed_offset=3)
    plain.launch(seed_offset=3)
    for a, b in zip(fused.results(), plain.results()):
        _same((a[0], a[1]), (b[0], b[1]), rtol=1e-12)


@pytest.mark.parametrize("sparse_u", [False, True])
def test_assembler_energies_are_column_norms(sparse_u):
    """Energies are the norms of the columns; a sparse U holds exactly the
    columns above the zeroton cutoff (the others keep what was there)."""
    rng = np.random.default_rng(9)
    sp = E.random_spec(4, 30, 200, 4, 3, 2, 12.0, rng=rng)
    bt = E.Batch([sp], precision="fp32", sparse_u=sparse_u)
    bt.bufs[0]["U"].fill_(float("nan"))
    bt.eval.launch()                                                     # up to the assembler only:
    bt.dft.launch()                                                      # the peeler changes U
    bt.asm.launch(4)
    torch.cuda.synchronize()
    U = bt.bufs[0]["U"].cpu().numpy()                                   # [C][Q][P]
    e = bt.bufs[0]["energy"].cpu().numpy()                              # [C][Q]
    ref = np.linalg.norm(U, axis=2) ** 2
    if not sparse_u:
        np.testing.assert_allclose(e, ref, rtol=1e-13, atol=0)
        return
    nz = e > sp.threshold() * sp.P
    assert 0 < int(nz.sum()) < nz.size
    np.testing.assert_allclose(e[nz], ref[nz], rtol=1e-13, atol=0)
    assert np.all(np.isnan(U[~nz]))                                     # skipped columns untouched
    assert np.array_equal(_written_flags(bt, 0).reshape(nz.shape), nz.astype(np.uint8))


def test_host_rows_match_device_copy():
    """Rows packed straight into pinned host memory equal the copied ones."""
    rng = np.random.default_rng(10)
    specs = [E.random_spec(4, 30, 150, 4, 3, 2, 15.0, rng=rng), E.random_spec(3, 12, 60, 3, 2, 2, None, rng=rng)]
    print("host rows vs device copy", flush=True)
    a = E.Batch(specs, precision="fp32", host_rows=True)
    b = E.Batch(specs, precision="fp32", host_rows=False)
    a.launch(seed_offset=5)
    b.launch(seed_offset=5)
    for x, y in zip(a.unpack(a.fetch()), b.unpack(b.fetch(torch.cuda.Stream()))):
        _same((x[0], x[1]), (y[0], y[1]), rtol=0.0)


def test_short_pass_launch_is_finished_on_fetch():
    """A launch with fewer peeling passes than its loops need (the served batches
    launch the most recent batches needed, plus one) is finished in fetch():
    the rows equal those of a launch with every pass."""
    rng = np.random.default_rng(21)
    specs = [E.random_spec(4, 30, 200, 4, 3, 2, 15.0, rng=rng), E.random_spec(4, 20, 120, 4, 3, 2, None, rng=rng)]
    full = E.Batch(specs, precision="fp32", fuse_from=None)
    short = E.Batch(specs, precision="fp32", fuse_from=None)
    full.launch(seed_offset=3)
    ref = full.results()
    short.passes = 1
    short.launch(seed_offset=3)
    got = short.results()
    assert short.launched == short.max_iter                          # the rest of the passes ran in fetch()
    for (ka, va, oa, _), (kb, vb, ob, _) in zip(got, ref):
        assert np.array_equal(ka, kb) and np.array_equal(oa, ob)
        assert np.array_equal(va, vb)
    assert short.passes <= short.max_iter and short.peel.last_iters >= 2


@pytest.mark.parametrize("case", ["k1_mixed_snr", "p_over_128", "k1_to_cap", "tiny_and_odd_q"])
def test_fused_tail_passes_match_per_pass_launches(case):
    """Passes fuse_from .. 14 as one kernel (one workgroup per instance looping
    over the passes, qsft_peel_passes) give what one launch pair per pass
    gives: the same keys in the same order, the same bits, the same pass
    counts, for loops that stop early and loops that run to the pass cap, with
    a sparse (K1) or dense (P > 128) working transform."""
    rng = np.random.default_rng(31)
    if case == "k1_mixed_snr":
        specs = [E.random_spec(4, 50, 1000, 6, 3, 2, snr, rng=rng) for snr in (0.0, 3.0, 5.0, 10.0, 10.0, 20.0, None)]
    elif case == "p_over_128":
        specs = [E.random_spec(4, 100, 400, 5, 3, 2, snr, rng=rng) for snr in (3.0, 10.0, None)]
    elif case == "tiny_and_odd_q":
        # fewer bins than one 64-bin group (q = 2, b = 3: 24 bins), a non-power-of-two q, noiseless loops
        specs = [E.random_spec(2, 8, 6, 3, 3, 2, None, rng=rng), E.random_spec(2, 12, 10, 3, 3, 2, 10.0, rng=rng),
                 E.random_spec(3, 10, 40, 4, 3, 2, 15.0, rng=rng), E.random_spec(3, 10, 40, 4, 3, 2, None, rng=rng)]
    else:
        # the two batches of 64 K1 transforms at 10 dB that bench.py serves, at
        # its noise draws: one of them leaves a loop that still finds
        # singletons and multitons at every pass up to the cap
        rng = np.random.default_rng(1234)
        sets = [[E.random_spec(4, 50, 1000, 6, 3, 2, 10.0, rng=rng) for _ in range(64)] for _ in range(2)]
    tries = [(specs, 5)] if case != "k1_to_cap" else [(sets[s % 2], 1000 + s) for s in range(8)]
    for specs, off in tries:
        ref = E.Batch(specs, precision="fp32", fuse_from=None)
        ref.launch(seed_offset=off)
        want = ref.results()
        iters = ref.peel.last_iters
        if case != "k1_to_cap" or iters == ref.max_iter:
            break
    print(case, "seed offset", off, "passes", iters, flush=True)
    if case == "k1_to_cap":
        assert iters == ref.max_iter
    for fuse_from in (1, 2, 4):
        bt = E.Batch(specs, precision="fp32", fuse_from=fuse_from)
        bt.launch(seed_offset=off)
        got = bt.results()
        assert bt.peel.last_iters == iters and bt.peel.last_unfinished == 0
        for a, b in zip(got, want):
            _identical(a, b)
            assert np.array_equal(a[2], b[2])                         # detection order


@pytest.mark.parametrize("case", ["k1_mixed_snr", "tiny_and_odd_q", "k2_q20", "everything_rebuilt"])
def test_first_pass_four_bins_per_wave_matches_one_bin_per_wave(case):
    """The first detection pass classifies four bins per wave (16 lanes each,
    the one-bin layout's per-lane sums rebuilt lane by lane): bit for bit the
    rows, order and pass count of the one-bin-per-wave pass, for noisy and
    noiseless K1 transforms, groups smaller than a wave (q^b < 64), q = 20,
    and a first pass whose columns are all rebuilt (the one-bin path then
    runs, the quad preconditions failing)."""
    rng = np.random.default_rng(47)
    if case == "k1_mixed_snr":
        specs = [E.random_spec(4, 50, 1000, 6, 3, 2, snr, rng=rng) for snr in (0.0, 5.0, 10.0, 10.0, 20.0, None)]
    elif case == "tiny_and_odd_q":
        specs = [E.random_spec(2, 8, 6, 3, 3, 2, None, rng=rng), E.random_spec(2, 12, 10, 3, 3, 2, 10.0, rng=rng),
                 E.random_spec(3, 10, 40, 4, 3, 2, 15.0, rng=rng), E.random_spec(4, 20, 50, 2, 3, 1, 12.0, rng=rng)]
    elif case == "k2_q20":
        specs = [E.random_spec(20, 16, 2000, 3, 3, 2, 20.0, rng=rng) for _ in range(3)]
    else:
        specs = [E.random_spec(4, 50, 1000, 6, 3, 2, 10.0, rng=rng), E.random_spec(4, 30, 100, 4, 3, 2, None, rng=rng)]
    print("four bins per wave vs one,", case, flush=True)
    runs = []
    for one in (False, True):
        bt = E.Batch(specs, precision="fp32", one_bin_waves=one)
        if case == "everything_rebuilt":
            bt.asm = K.Assemble([dict(it, cutoff=float("inf")) for it in bt.asm_items])
        bt.launch(seed_offset=9)
        runs.append((bt.results(), bt.peel.last_iters))
    (a_res, a_it), (b_res, b_it) = runs
    assert a_it == b_it
    for a, b in zip(a_res, b_res):
        _identical(a, b)
        assert np.array_equal(a[2], b[2])                             # detection order


def _written_flags(bt, i):
    d, ws = bt.peel.descs[i], bt.peel.ws[i]
    off = d.written - ws.data_ptr()
    sp = bt.specs[i]
    return ws[off: off + sp.C * sp.Q].cpu().numpy()


def _identical(a, b):
    assert [tuple(r) for r in a[0]] == [tuple(r) for r in b[0]]      # same rows, same order
    assert np.array_equal(np.asarray(a[1]), np.asarray(b[1]))         # same bits


@pytest.mark.parametrize("everything", [False, True])
def test_sparse_u_matches_dense_u(everything):
    """Sparse U: the assembler writes only the columns above the zeroton
    cutoff, the peeler rebuilds any other column it reads from the samples and
    the same Philox noise. Bit-identical rows to a fully written U. With
    `everything` the assembler writes no column at all (its cutoff is +inf),
    so every column the peeler reads, from the first pass on, is a rebuilt one."""
    rng = np.random.default_rng(12)
    specs = [E.random_spec(4, 50, 1000, 6, 3, 2, snr, rng=rng) for snr in (3.0, 10.0)]
    specs += [E.random_spec(4, 30, 100, 4, 3, 2, None, rng=rng), E.random_spec(3, 20, 90, 4, 3, 2, 20.0, rng=rng)]
    print("sparse U vs dense U, everything rebuilt:", everything, flush=True)
    sparse = E.Batch(specs, precision="fp32", sparse_u=True)
    dense = E.Batch(specs, precision="fp32", sparse_u=False)
    assert all(d.flags & 8 for d in sparse.peel.descs) and not any(d.flags & 8 for d in dense.peel.descs)
    if everything:
        sparse.asm = K.Assemble([dict(it, cutoff=float("inf")) for it in sparse.asm_items])
    sparse.launch(seed_offset=5)
    dense.launch(seed_offset=5)
    for a, b in zip(sparse.results(), dense.results()):
        _identical(a, b)
    for i, sp in enumerate(specs):
        written = _written_flags(sparse, i)
        energy = sparse.bufs[i]["energy"].cpu().numpy().ravel()
        thr = sp.threshold() * sp.P
        rebuilt = int(np.sum((written == 1) & (energy <= thr)))
        if everything:
            # every first-pass non-zeroton was read (and so rebuilt and written back)
            assert np.all(written[energy > thr] == 1), i
        else:
            assert np.array_equal(written[energy > thr], np.ones(int(np.sum(energy > thr)), np.uint8)), i
        print(f"instance {i}: written {int(written.sum())} of {written.size}, rebuilt below the cutoff {rebuilt}")


def test_sparse_u_beside_an_instance_over_128_rows():
    """A batch whose largest instance has P > 128 runs the four-rows-per-lane
    peeler for every instance, while its instances with P <= 128 keep a
    sparse working transform. Columns the assembler skipped must then be
    rebuilt by that peeler too: every instance gives the rows it gives alone
    (two-rows-per-lane peeler) and the rows of the same batch with every
    column written."""
    rng = np.random.default_rng(17)
    specs = [E.random_spec(2, 70, 60, 7, 3, 2, 12.0, rng=rng)]                  # P = 142: dense U
    specs += [E.random_spec(4, 50, 1000, 6, 3, 2, snr, rng=rng) for snr in (0.0, 3.0, 5.0)]
    specs += [E.random_spec(3, 20, 90, 4, 3, 2, 2.0, rng=rng), E.random_spec(4, 30, 100, 4, 3, 2, None, rng=rng)]
    print("sparse U next to P = 142", flush=True)
    mixed = E.Batch(specs, precision="fp32", sparse_u=True)
    dense = E.Batch(specs, precision="fp32", sparse_u=False)
    assert mixed.peel.max_p > 128
    assert all(d.flags & 8 for d in mixed.peel.descs[1:]) and not mixed.peel.descs[0].flags & 8
    mixed.launch(seed_offset=4)
    dense.launch(seed_offset=4)
    got = mixed.results()
    for a, b in zip(got, dense.results()):
        _identical(a, b)
    rebuilt = 0
    for i, sp in enumerate(specs):
        alone = E.Batch([sp], precision="fp32", sparse_u=True)
        alone.launch(seed_offset=4)
        (k, v, _, _), = alone.results()
        _same((got[i][0], got[i][1]), (k, v), rtol=1e-12)
        if i > 0:
            written = _written_flags(mixed, i)
            energy = mixed.bufs[i]["energy"].cpu().numpy().ravel()
            rebuilt += int(np.sum((written == 1) & (energy <= sp.threshold() * sp.P)))
    print("columns rebuilt below the cutoff:", rebuilt, flush=True)
    assert rebuilt > 0                                                          # the case was exercised


def test_k2_served_fp32_matches_fp64_path():
    """q = 20, n = 16, S = 2000, b = 3 (8000 bins, radix-20 DFT), NSO R = 2, 20 dB:
    the served FP32 path (matrix-core evaluation, FP32 DFT, FP64 peel) against
    the FP64 path on the same Philox noise: the same NMSE to the planted
    spectrum within 2 %, supports differing in at most 0.5 % of the keys and
    shared coefficients within 1e-4 relative."""
    rng = np.random.default_rng(20)
    sp = E.random_spec(20, 16, 2000, 3, 3, 2, 20.0, rng=rng)
    print("K2 fp32 vs fp64", flush=True)
    f32, f64 = E.Batch([sp], precision="fp32"), E.Batch([sp], precision="fp64")
    f32.launch(seed_offset=3)
    f64.launch(seed_offset=3)
    (k32, v32, _, _), = f32.results()
    (k64, v64, _, _), = f64.results()
    n32, n64 = E.nmse(k32, v32, sp), E.nmse(k64, v64, sp)
    d32 = {tuple(r): v for r, v in zip(k32, v32)}
    d64 = {tuple(r): v for r, v in zip(k64, v64)}
    common = set(d32) & set(d64)
    diff = len(set(d32) ^ set(d64))
    rel = max(abs(d32[k] - d64[k]) / abs(d64[k]) for k in common)
    print(f"NMSE fp32 {n32:.4e} fp64 {n64:.4e}; keys {len(d32)} / {len(d64)}, differing {diff}, max rel {rel:.2e}")
    assert abs(n32 - n64) <= 0.02 * n64 + 1e-12
    assert diff <= 0.005 * len(d64)
    assert rel < 1e-4


def test_k4_large_job_noiseless_support_exact():
    """The large job of the serving mix: q = 4, n = 100, S = 5000, b = 7 (16384
    bins), NSO R = 2 (P = 202 rows: the four-rows-per-lane peeler, two 64-digit
    blocks in the evaluator). Noiseless through the served FP32 engine: the
    support is exact and the values match to FP32 level."""
    rng = np.random.default_rng(41)
    sp = E.random_spec(4, 100, 5000, 7, 3, 2, None, rng=rng)
    print("K4 large job, noiseless", flush=True)
    bt = E.Batch([sp], precision="fp32")
    bt.launch()
    (k, v, _, _), = bt.results()
    assert {tuple(r) for r in k} == {tuple(c) for c in sp.locq.T}
    assert E.nmse(k, v, sp) < 1e-10


def _k3_mix(count, rng):
    """Uneven instances by the K3 recipe: q in {2, 3, 4, 20}, n in [8, 64], S
    log-uniform in [10, 2000] (at most q^n / 8), b = ceil(log_q S) + 1, C = 3,
    NSO R = 2, SNR uniform in [5, 30] dB."""
    specs = []
    for _ in range(count):
        q = int(rng.choice([2, 3, 4, 20]))
        n = int(rng.integers(8, 65))
        S = int(np.exp(rng.uniform(np.log(10), np.log(2000))))
        S = max(2, min(S, q ** n // 8))
        b = int(np.ceil(np.log(S) / np.log(q))) + 1
        specs.append(E.random_spec(q, n, S, b, 3, 2, float(rng.uniform(5, 30)), rng=rng))
    return specs


def test_k3_uneven_mix_batch_equals_alone():
    """64 instances of the K3 mix in one set of launches: every instance gives
    the rows it gives alone (a sample of 10). Noiseless copies of 16 of the
    shapes recover only true keys, with their exact values, and 95 % of all
    planted keys (peeling with C = 3 groups is not certain to finish every
    instance)."""
    rng = np.random.default_rng(33)
    specs = _k3_mix(64, rng)
    print("K3 mix:", sorted({(s.q, s.b) for s in specs}), flush=True)
    mix = E.Batch(specs, precision="fp32")
    mix.launch(seed_offset=9)
    got = mix.results()
    for i in rng.choice(len(specs), size=10, replace=False):
        alone = E.Batch([specs[i]], precision="fp32")
        alone.launch(seed_offset=9)
        (k, v, _, _), = alone.results()
        _same((got[i][0], got[i][1]), (k, v), rtol=1e-12)
    clean = [E.TransformSpec(**{**s.__dict__, "noise_sd": 0.0}) for s in specs[:16]]
    bt = E.Batch(clean, precision="fp32")
    bt.launch()
    found = planted = 0
    for (k, v, _, _), sp in zip(bt.results(), clean):
        truth = {tuple(c): a for c, a in zip(sp.locq.T, sp.strengths)}
        for r, val in zip(k, v):
            assert tuple(r) in truth, (sp.q, sp.n, sp.S, sp.b)
            assert abs(val - truth[tuple(r)]) <= 1e-4 * abs(truth[tuple(r)])
        found += len(k)
        planted += sp.S
    print(f"noiseless K3 copies: {found} of {planted} keys", flush=True)
    assert found >= 0.95 * planted
