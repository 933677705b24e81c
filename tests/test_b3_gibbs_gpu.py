"""GPU parity of the device Gibbs sampler (GibbsSampler2::next,
sampled_multispot.h:681-709) against the oracle's restatement, draw for
draw: the same MT19937 stream per chain (the reference's seeding, 37
discarded draws, then `skip`), the same masks and spot tables, and the
sampled state paths must agree exactly after every step; the sample
intensities rebuilt from the device's sample must equal base + the
oracle's (within 1e-9 absolute, FP64 sums of the same terms in another
order).  The device forms each frame's on/off difference from per-spot
sums and pairwise overlaps instead of pixel sums: the same number up to
rounding, so only an exact tie could draw differently.
"""
import numpy as np
import pytest

from oracle import b3oracle as bo
from rfba_amd import b3

pytestmark = pytest.mark.gpu

A_REF = np.array([[0.16, 0.84, 0.0], [0.495, 0.495, 0.01], [0.0, 0.0, 1.0]])  # multispot5.cfg:15
PI_REF = np.array([0.5, 0.5, 0.0])


def model(var=100.0, zfactor=2.0):
    return b3.make_model(A_REF, PI_REF, var, zfactor=zfactor)


def chain_cases():
    """(ChainData, label): dense boxes, a region with holes, an ejected spot,
    no spots, a mask longer than one LDS chunk, T = 1."""
    cases = []
    data, params, _ = b3.synth_block((9, 8, 7), 12, 5, 11, A=A_REF, pi=PI_REF)
    cases.append((b3.ChainData(data.reshape(12, -1), b3.box_pixels((9, 8, 7)), params, base=100.0, seed=7,
                               skip=0), "box 9x8x7 K=5 T=12"))
    data, params, _ = b3.synth_block((10, 9, 6), 1, 3, 12, A=A_REF, pi=PI_REF)
    xyz = b3.box_pixels((10, 9, 6))
    keep = np.ones(len(xyz), bool)
    keep[::4] = False
    cases.append((b3.ChainData(data.reshape(1, -1)[:, keep], xyz[keep], params, base=100.0, seed=8, skip=5),
                  "holes K=3 T=1"))
    data, params, _ = b3.synth_block((8, 8, 5), 40, 2, 13, A=A_REF, pi=PI_REF)
    params = params.copy()
    params[1, 2] = 60.0  # far outside the region: empty mask
    cases.append((b3.ChainData(data.reshape(40, -1), b3.box_pixels((8, 8, 5)), params, base=100.0, seed=9,
                               skip=623), "ejected spot T=40"))
    data, _, _ = b3.synth_block((5, 4, 3), 5, 1, 14, A=A_REF, pi=PI_REF)
    cases.append((b3.ChainData(data.reshape(5, -1), b3.box_pixels((5, 4, 3)), np.zeros((0, 5)), base=100.0,
                               seed=10, skip=624), "no spots"))
    data, params, _ = b3.synth_block((16, 16, 8), 20, 4, 15, A=A_REF, pi=PI_REF)
    cases.append((b3.ChainData(data.reshape(20, -1), b3.box_pixels((16, 16, 8), (3, 1, 2)), params + [0, 0, 3, 1, 2],
                               base=100.0, seed=11, skip=100000), "mask > 1024 pixels, offset origin"))
    return cases


def oracle_chain(ch, m, steps, iters, variances, A=A_REF, pi=PI_REF):
    T = ch.data.shape[0]
    data = np.asarray(ch.data, np.float64).reshape(T, -1)
    P = data.shape[1]
    K = len(ch.spots)
    spot = np.stack([bo.spot(ch.xyz, p, zfactor=m.zfactor)[0] for p in ch.spots]) if K else np.zeros((0, P))
    off, idx = bo.spot_pixels(ch.xyz, ch.spots) if K else (np.zeros(1, np.int32), np.zeros(0, np.int32))
    n = steps * iters * K * T
    u = b3.mt19937_uniforms(ch.seed, max(n, 1), skip=ch.skip)
    sample = np.full((K, T), 2, np.uint8)
    intens = np.zeros((T, P))
    out, k_u = [], 0
    for s in range(steps):
        k_u += bo.gibbs_next(data - ch.base, spot, off, idx, A, pi, variances[s], iters, sample, intens, u[k_u:])
        out.append(sample.copy())
    return np.array(out).reshape(steps, K, T), intens + ch.base, spot


def test_gibbs_matches_oracle_draw_for_draw(gpu):
    m = model()
    cases = chain_cases()
    steps, iters = 3, 2
    variances = [100.0, 150.0, 80.0]  # a variance per step, as thermodynamic integration sets it
    chains = []
    for ch, _ in cases:
        ch.slots = steps
        chains.append(ch)
    g = b3.Gibbs(chains, m)
    try:
        g.run(steps, iters, variance=variances)
        for i, (ch, label) in enumerate(cases):
            smp, intens, spot = oracle_chain(ch, m, steps, iters, variances)
            K, T = smp.shape[1:]
            got = g.samples(i, steps)
            assert np.array_equal(got, smp), label
            assert np.array_equal(g.sample(i), smp[-1]), label
            if K:
                assert np.allclose(g.spot_table(i), spot, rtol=1e-13, atol=1e-300), label
            assert np.allclose(g.sample_intensities(i), intens, rtol=0, atol=1e-9), label
            assert g.draws(i) == ch.skip + steps * iters * K * T, label
        # every state occurs somewhere (the comparison is not vacuous)
        allst = np.concatenate([g.samples(i, steps).ravel() for i in range(len(cases))])
        assert {0, 1, 2} <= set(np.unique(allst).tolist())
    finally:
        g.close()


def test_gibbs_weights_beyond_double_range_match_the_log_space_oracle(gpu):
    # A model in which an off spot never turns on (A[1][0] = 0) and a tiny
    # variance (|B_on - B_off| up to ~1e9), run on data drawn with the usual
    # blink matrix: where the data turns a spot on after off frames, the
    # path that stayed on (weight e^-1e9 relative) competes with paths that
    # pay the same order of penalty elsewhere.  Messages in plain doubles
    # lose such paths (the weight becomes 0); the reference's log-space
    # forward algorithm and the device's extended-range messages keep them,
    # so the draws must equal the oracle's.
    A_Z = np.array([[0.9, 0.09, 0.01], [0.0, 0.98, 0.02], [0.0, 0.0, 1.0]])
    m = b3.make_model(A_Z, PI_REF, 0.01, zfactor=2.0)
    data, params, states = b3.synth_block((8, 8, 5), 150, 3, 31, A=A_REF, pi=PI_REF)
    on = states == 0
    assert np.any(~on[:, :-1] & on[:, 1:]), "no off -> on switch in the data"
    ch = b3.ChainData(data.reshape(150, -1), b3.box_pixels((8, 8, 5)), params, base=100.0, seed=41, skip=2, slots=2)
    g = b3.Gibbs([ch], m)
    try:
        g.run(2, 1)
        smp, intens, _ = oracle_chain(ch, m, 2, 1, [0.01, 0.01], A=A_Z)
        assert np.array_equal(g.samples(0, 2), smp)
        assert np.allclose(g.sample_intensities(0), intens, rtol=0, atol=1e-9)
    finally:
        g.close()


def test_gibbs_runs_continue_the_stream(gpu):
    m = model()
    data, params, _ = b3.synth_block((9, 9, 6), 30, 6, 21, A=A_REF, pi=PI_REF)
    mk = lambda: b3.ChainData(data.reshape(30, -1), b3.box_pixels((9, 9, 6)), params, base=100.0, seed=3, skip=17,
                              slots=5)
    g1, g2 = b3.Gibbs([mk()], m), b3.Gibbs([mk()], m)
    try:
        g1.run(5, 1)
        g2.run(2, 1)
        first = g2.samples(0, 2)
        g2.run(3, 1)
        assert np.array_equal(np.concatenate([first, g2.samples(0, 3)]), g1.samples(0, 5))
        assert np.array_equal(g1.sample_intensities(0), g2.sample_intensities(0))
        g1.reset()
        assert np.all(g1.sample(0) == 2)
        assert np.all(g1.sample_intensities(0) == 100.0)
    finally:
        g1.close()
        g2.close()


def test_gibbs_long_series_uses_large_lds(gpu):
    # T = 2500: 85 KB of LDS per chain (more than 64 KB), 40 frame groups of
    # 64 with a partial last one
    m = model()
    T = 2500
    data, params, _ = b3.synth_block((6, 6, 4), T, 3, 51, A=A_REF, pi=PI_REF)
    ch = b3.ChainData(data.reshape(T, -1), b3.box_pixels((6, 6, 4)), params, base=100.0, seed=21, skip=3, slots=2)
    g = b3.Gibbs([ch], m)
    try:
        g.run(2, 1)
        smp, intens, _ = oracle_chain(ch, m, 2, 1, [100.0, 100.0])
        assert np.array_equal(g.samples(0, 2), smp)
        assert np.allclose(g.sample_intensities(0), intens, rtol=0, atol=1e-9)
    finally:
        g.close()


def test_gibbs_samples_feed_the_batched_likelihood(gpu):
    # the device samples are the [S][K][T] state input of the batched logL
    m = model()
    m.checkpoint = 4
    data, params, _ = b3.synth_block((9, 8, 7), 16, 4, 31, A=A_REF, pi=PI_REF)
    g = b3.Gibbs([b3.ChainData(data.reshape(16, -1), b3.box_pixels((9, 8, 7)), params, base=100.0, seed=1,
                               slots=3)], m)
    try:
        g.run(3, 1)
        states = g.samples(0, 3)
    finally:
        g.close()
    Pq = b3.perturb(params, 2, 5)
    bt = b3.Batch([b3.BlockData(data, Pq, states, 100.0)], m)
    try:
        bt.run(use_graph=False)
        L, G = bt.results(0)
    finally:
        bt.close()
    Lo, Go = bo.model_eval(data, 100.0, m, Pq, states)
    assert np.max(np.abs(L - Lo) / np.abs(Lo)) < 1e-6


def test_gibbs_rejects_bad_masks_and_sizes(gpu):
    m = model()
    data, params, _ = b3.synth_block((5, 5, 5), 4, 2, 41, A=A_REF, pi=PI_REF)
    ch = b3.ChainData(data.reshape(4, -1), b3.box_pixels((5, 5, 5)), params, base=100.0)
    g = b3.Gibbs([ch], m)
    g.close()
    # hand-built chain structs with a mask index outside the region
    from rfba_amd.srrf import DeviceBuffer
    bufs = []
    try:
        P, T, K = 125, 4, 1
        dd = DeviceBuffer.from_array(data.reshape(4, -1))
        dsp = DeviceBuffer(K * P * 8)
        doff = DeviceBuffer.from_array(np.array([0, 3], np.int32))
        dsm = DeviceBuffer(K * T)
        bufs += [dd, dsp, doff, dsm]
        import ctypes
        for bad in ([0, 5, 125], [0, 7, 7], [3, 2, 1], [-1, 2, 3]):
            didx = DeviceBuffer.from_array(np.array(bad, np.int32))
            bufs.append(didx)
            arr = (b3._GibbsChain * 1)()
            a = arr[0]
            a.P, a.T, a.K = P, T, K
            a.d_data, a.d_spot, a.d_mask_off, a.d_mask = dd.ptr, dsp.ptr, doff.ptr, didx.ptr
            a.d_sample = dsm.ptr
            plan = ctypes.c_void_p()
            rc = b3.lib().rfba3b_gibbs_create(arr, 1, ctypes.byref(m), ctypes.byref(plan), None)
            assert rc == -1 and b"mask" in b3.lib().rfba3b_last_error(), bad
        # T beyond the LDS budget (32 T + 4992 bytes > 160 KiB), rejected
        # before anything is read
        arr = (b3._GibbsChain * 1)()
        a = arr[0]
        a.P, a.T, a.K = P, 6000, 0
        a.d_data, a.d_sample = dd.ptr, dsm.ptr
        plan = ctypes.c_void_p()
        assert b3.lib().rfba3b_gibbs_create(arr, 1, ctypes.byref(m), ctypes.byref(plan), None) == -1
        assert b"LDS" in b3.lib().rfba3b_last_error()
    finally:
        for b in bufs:
            b.close()
