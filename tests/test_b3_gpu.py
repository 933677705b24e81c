"""GPU parity of the 3B likelihood kernels (librfba_3b.so) against the FP64
oracle (oracle/b3_oracle.cpp).

Bars (FP32 pixel terms, FP64 sums and recursions):
  logL: |gpu - oracle| <= 1e-6 * |oracle| (the FP32 spot images carry ~3e-7
        relative rounding, which moves logL by ~1e-7 relative; measured
        values are printed)
  gradient: ||gpu - oracle|| <= 1e-5 * ||oracle|| per item, and
            ||gpu - central FD (h = 1e-3) of the FP64 oracle|| <= 1e-4 * ||FD||.
Errors are printed per case.
"""
import numpy as np
import pytest

from oracle import b3oracle as bo
from rfba_amd import b3

pytestmark = pytest.mark.gpu

LOGL_RTOL = 1e-6
GRAD_RTOL = 1e-5


def model(**kw):
    base = dict(A=b3.blink_matrix(), pi=(0.5, 0.5, 0.0), variance=100.0, zfactor=2.0, z_exact=0, checkpoint=0,
                mu_brightness=7.0, sigma_brightness=1.0, mu_blur=0.25, sigma_blur=0.1)
    base.update(kw)
    return b3.make_model(**base)


def grid(X, Y, Z):
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(np.int32)


def compare(L, G, Lo, Go, label):
    el = np.abs(L - Lo) / np.abs(Lo)
    ng = np.linalg.norm(Go, axis=-1)
    eg = np.linalg.norm(G - Go, axis=-1) / np.maximum(ng, 1e-300)
    print("%s: logL max rel err %.3g, grad max rel err %.3g (items %d)" % (label, el.max(), eg.max(), el.size))
    assert el.max() <= LOGL_RTOL
    assert eg.max() <= GRAD_RTOL


def run_batch(blocks, m, graph=False):
    bt = b3.Batch(blocks, m)
    bt.run(use_graph=graph)
    res = [bt.results(i) for i in range(len(blocks))]
    bt.close()
    return res


@pytest.mark.parametrize("ckpt", [0, 1, 3, 11])
def test_batched_vs_oracle(gpu, ckpt):
    m = model(checkpoint=ckpt)
    data, params, states = b3.synth_block((12, 11, 10), 23, 5, 21)
    P = b3.perturb(params, 3, 22)
    st = np.stack([states, np.roll(states, 3, axis=1)])
    (L, G), = run_batch([b3.BlockData(data, P, st, 100.0)], m)
    Lo, Go = bo.model_eval(data, 100.0, m, P, st)
    compare(L, G, Lo, Go, "ckpt=%d" % ckpt)


@pytest.mark.parametrize("ckpt", [0, 1, 3, 11])
def test_checkpointed_lane_hmm_vs_oracle(gpu, ckpt):
    # the lane form of the HMM (forward messages kept every ckpt frames, the
    # rest recomputed in the adjoint; stride 11 is the batched 3B config's),
    # forced at T = 120 where the default is the one-wave form; a captured
    # graph is rebuilt when the form changes
    m = model(checkpoint=ckpt)
    data, params, states = b3.synth_block((10, 9, 8), 120, 6, 41)
    P = b3.perturb(params, 2, 42)
    st = np.stack([states, np.roll(states, 7, axis=1)])
    bt = b3.Batch([b3.BlockData(data, P, st, 100.0)], m)
    bt.run(use_graph=True)
    Lw, Gw = bt.results(0)
    assert bt.set_hmm_form(1) == 0
    bt.run(use_graph=True)
    L, G = bt.results(0)
    bt.run_per_block(per_param_set=True)
    L2, G2 = bt.results(0)
    assert bt.set_hmm_form(0) == 1
    bt.run(use_graph=True)
    Lw2, Gw2 = bt.results(0)
    with pytest.raises(b3.B3Error):
        bt.set_hmm_form(2)
    bt.close()
    assert np.array_equal(L, L2) and np.array_equal(G, G2)
    assert np.array_equal(Lw, Lw2) and np.array_equal(Gw, Gw2)
    Lo, Go = bo.model_eval(data, 100.0, m, P, st)
    compare(L, G, Lo, Go, "lane form ckpt=%d" % ckpt)
    compare(Lw, Gw, Lo, Go, "wave form")


def test_series_longer_than_one_pass(gpu):
    # T = 300: three 128-frame passes of the correlation kernel, the last one
    # partial (the frame-independent sums come from the first pass); K = 10
    # gives a full and a partial group of 8 spots; P = 504 ends in a partial
    # pixel chunk; slow bleaching keeps spots blinking through all passes
    m = model(checkpoint=0)
    data, params, states = b3.synth_block((9, 8, 7), 300, 10, 31, A=b3.blink_matrix(0.2, 0.1, 0.001))
    P = b3.perturb(params, 2, 32)
    st = np.stack([states, np.roll(states, 5, axis=1)])
    (L, G), = run_batch([b3.BlockData(data, P, st, 100.0)], m)
    Lo, Go = bo.model_eval(data, 100.0, m, P, st)
    compare(L, G, Lo, Go, "T=300 K=10")


def test_checkpoint_stride_does_not_change_results(gpu):
    data, params, states = b3.synth_block((10, 9, 8), 40, 4, 23)
    P = b3.perturb(params, 2, 24)
    outs = [run_batch([b3.BlockData(data, P, states[None], 100.0)], model(checkpoint=c))[0] for c in (1, 5, 7, 40)]
    for L, G in outs[1:]:
        np.testing.assert_allclose(L, outs[0][0], rtol=1e-13)
        np.testing.assert_allclose(G, outs[0][1], rtol=1e-9, atol=1e-12 * np.abs(outs[0][1]).max())


def test_uneven_blocks_plain_graph_and_per_block_agree(gpu):
    m = model(checkpoint=11)
    rng = np.random.default_rng(7)
    blocks = []
    for i in range(5):
        dims = tuple(int(v) for v in rng.integers(8, 15, 3))
        T = int(rng.integers(12, 30))
        K = int(rng.integers(2, 6))
        data, params, states = b3.synth_block(dims, T, K, 100 + i)
        blocks.append(b3.BlockData(data, b3.perturb(params, 2, 200 + i), states[None], 100.0))
    bt = b3.Batch(blocks, m)
    res = {}
    for mode in ("plain", "graph", "per_block", "per_param_set", "graph2"):
        if mode.startswith("per_"):
            bt.run_per_block(per_param_set=(mode == "per_param_set"))
        else:
            bt.run(use_graph=mode.startswith("graph"))
        res[mode] = [bt.results(i) for i in range(len(blocks))]
    bt.close()
    for mode in ("graph", "per_block", "per_param_set", "graph2"):
        for (L, G), (L0, G0) in zip(res[mode], res["plain"]):
            assert np.array_equal(L, L0) and np.array_equal(G, G0), mode
    for b, (L, G) in zip(blocks, res["plain"]):
        Lo, Go = bo.model_eval(b.data, 100.0, m, b.params, b.states)
        compare(L, G, Lo, Go, "uneven block %s" % (b.data.shape,))


def test_gradient_vs_central_fd(gpu):
    # exact z derivative so that finite differences apply with zF = 2
    m = model(z_exact=1, checkpoint=4)
    data, params, states = b3.synth_block((11, 10, 9), 16, 3, 31)
    P = b3.perturb(params, 1, 32)
    (L, G), = run_batch([b3.BlockData(data, P, states[None], 100.0)], m)
    xyz = grid(11, 10, 9)
    d2 = data.reshape(16, -1).astype(np.float64)
    for k in range(3):
        bg = np.full_like(d2, 100.0)
        for j in range(3):
            if j != k:
                bg += (states[j] == 0)[:, None] * bo.spot(xyz, P[0, j], 2.0)[0][None]
        phi = P[0, k]
        fd = np.empty(5)
        for n in range(5):
            h = 1e-3 * (abs(phi[n]) if n < 2 else 1.0)
            e = np.zeros(5)
            e[n] = h
            fd[n] = (bo.logl_grad(xyz, d2, bg, phi + e, m)[0] - bo.logl_grad(xyz, d2, bg, phi - e, m)[0]) / (2 * h)
        err = np.linalg.norm(G[0, k, 0] - fd) / np.linalg.norm(fd)
        print("spot %d: |grad - FD| / |FD| = %.3g" % (k, err))
        assert err <= 1e-4


def test_functor_and_forward_messages(gpu):
    m = model()
    data, params, states = b3.synth_block((9, 8, 7), 14, 3, 41)
    xyz_all = grid(9, 8, 7)
    # an irregular region: a ball of pixels around spot 0 (masked list)
    c = params[0, 2:5]
    keep = np.linalg.norm((xyz_all - c) * [1, 1, 0.5], axis=1) < 4.5
    xyz = xyz_all[keep]
    d2 = data.reshape(14, -1)[:, keep]
    bgs = []
    for s in range(3):
        bg = np.full(d2.shape, 100.0)
        for j in (1, 2):
            bg += (np.roll(states[j], s) == 0)[:, None] * bo.spot(xyz, params[j], 2.0)[0][None]
        bgs.append(bg)
    bgs = np.stack(bgs)
    phis = np.stack([params[0], params[0] + [50, 0.1, 0.3, -0.2, 0.4], [0.0, 1.3, 3, 3, 3], [900, -1, 3, 3, 3]])
    g, Ls = b3.neg_grad_sampled(xyz, d2.astype(np.float32), bgs.astype(np.float32), phis, m)
    assert np.isnan(g[2:]).all()
    for i in range(2):
        go, Lo = bo.neg_grad(xyz, d2.astype(np.float32).astype(np.float64), bgs.astype(np.float32).astype(np.float64),
                             phis[i], m)
        eg = np.linalg.norm(g[i] - go) / np.linalg.norm(go)
        el = np.abs(Ls[i] - Lo) / np.abs(Lo)
        print("functor %d: grad rel err %.3g, logL rel err %.3g" % (i, eg, el.max()))
        assert eg <= GRAD_RTOL and el.max() <= LOGL_RTOL
    delta, L = b3.forward_messages(xyz, d2.astype(np.float32), bgs[0].astype(np.float32), params[0], m)
    mu, dmu = bo.spot(xyz, params[0], 2.0)
    on, off, _ = bo.emissions(d2.astype(np.float32).astype(np.float64), bgs[0].astype(np.float32).astype(np.float64),
                              mu, dmu, 100.0)
    do, Lo = bo.forward_delta(b3.blink_matrix(), (0.5, 0.5, 0.0), on, off)
    fin = do > -1e99
    assert np.array_equal(fin, delta > -1e99)
    np.testing.assert_allclose(delta[fin], do[fin], rtol=LOGL_RTOL)
    assert L == pytest.approx(Lo, rel=LOGL_RTOL)


def test_sharp_likelihood_with_zero_transitions(gpu):
    # A model in which an off spot never turns on (A[1][0] = 0) with a
    # variance of 0.01, on data drawn with the default blink matrix: the
    # per-frame log emission ratios reach ~1e6 and state weights leave the
    # double range.  The reference's log-space recursion keeps them (a path
    # that stayed on through off frames competes where the data turns the
    # spot on again); normalised doubles would lose them.
    A_Z = np.array([[0.9, 0.09, 0.01], [0.0, 0.98, 0.02], [0.0, 0.0, 1.0]])
    m = model(A=A_Z, variance=0.01)
    data, params, states = b3.synth_block((10, 9, 8), 60, 4, 61)
    on = states == 0
    assert np.any(~on[:, :-1] & on[:, 1:]), "no off -> on switch in the data"
    P = b3.perturb(params, 3, 62)
    st = np.stack([states, np.roll(states, 5, axis=1)])
    (L, G), = run_batch([b3.BlockData(data, P, st, 100.0)], m)
    Lo, Go = bo.model_eval(data, 100.0, m, P, st)
    compare(L, G, Lo, Go, "sharp, zero transitions")
    # forward messages of one spot (FFBS input), finite where the oracle's are
    xyz = grid(10, 9, 8)
    d2 = data.reshape(60, -1)
    bg = np.full(d2.shape, 100.0)
    for j in range(1, 4):
        bg += (states[j] == 0)[:, None] * bo.spot(xyz, params[j], 2.0)[0][None]
    delta, Ld = b3.forward_messages(xyz, d2.astype(np.float32), bg.astype(np.float32), params[0], m)
    mu, dmu = bo.spot(xyz, params[0], 2.0)
    on_l, off_l, _ = bo.emissions(d2.astype(np.float32).astype(np.float64), bg.astype(np.float32).astype(np.float64),
                                  mu, dmu, 0.01)
    do, Lo1 = bo.forward_delta(A_Z, (0.5, 0.5, 0.0), on_l, off_l)
    fin = do > -1e99
    assert np.array_equal(fin, delta > -1e99)
    np.testing.assert_allclose(delta[fin], do[fin], rtol=LOGL_RTOL)
    assert Ld == pytest.approx(Lo1, rel=LOGL_RTOL)


def test_config4_shape_subset(gpu):
    # 29^3 block, T = 100, 40 spots; two of the perturbed parameter sets
    m = model(checkpoint=0)
    data, params, states = b3.synth_block((29, 29, 29), 100, 40, 6)
    P = b3.perturb(params, 2, 7)
    (L, G), = run_batch([b3.BlockData(data, P, states[None], 100.0)], m, graph=True)
    Lo, Go = bo.model_eval(data, 100.0, m, P, states[None])
    compare(L, G, Lo, Go, "29^3 T=100 K=40")


def test_ffbs_matches_oracle_sampler(gpu):
    # forward messages on the GPU, then backward sampling on the GPU; the
    # oracle's sampler on the same messages and draws gives the same states
    m = model()
    data, params, states = b3.synth_block((9, 8, 7), 40, 3, 51)
    xyz = grid(9, 8, 7)
    d2 = data.reshape(40, -1)
    k = int(np.argmax((states == 0).sum(1)))  # the spot that is on most often
    assert (states[k] == 0).sum() >= 5
    bg = np.full(d2.shape, 100.0)
    for j in range(3):
        if j != k:
            bg += (states[j] == 0)[:, None] * bo.spot(xyz, params[j], 2.0)[0][None]
    u = b3.mt19937_uniforms(7, 40)
    st, delta, L = b3.ffbs(xyz, d2.astype(np.float32), bg.astype(np.float32), params[k], m, u)
    so = bo.backward_sampling(b3.blink_matrix(), delta, u)
    assert np.array_equal(st, so.astype(np.uint8))
    # the sampled path follows the data: frames where the spot is on
    agree = np.mean((st == 0) == (states[k] == 0))
    print("ffbs: %d of %d on-frames sampled, %.2f agreement with the true states"
          % ((st == 0).sum(), (states[k] == 0).sum(), agree))
    assert agree > 0.9


def test_backward_sample_device_many_chains(gpu):
    rng = np.random.default_rng(12)
    A, pi = b3.blink_matrix(), np.array([0.5, 0.5, 0.0])
    n, T = 300, 25
    deltas = np.stack([bo.forward_delta(A, pi, rng.normal(0, 2, T), rng.normal(0, 2, T))[0] for _ in range(n)])
    U = rng.uniform(size=(n, T))
    got = b3.backward_sample(deltas, U, model())
    ref = np.stack([bo.backward_sampling(A, deltas[i], U[i]) for i in range(n)]).astype(np.uint8)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("T,K", [(1, 1), (2, 3), (3, 9)])
def test_short_series_and_single_spot(gpu, T, K):
    # T = 1 (no transitions, checkpoint stride 1), one spot, more spots than
    # a correlation group's 8
    m = model(checkpoint=0)
    data, params, states = b3.synth_block((7, 6, 5), T, K, 70 + T)
    P = b3.perturb(params, 2, 71)
    (L, G), = run_batch([b3.BlockData(data, P, states[None], 100.0)], m)
    Lo, Go = bo.model_eval(data, 100.0, m, P, states[None])
    compare(L, G, Lo, Go, "T=%d K=%d" % (T, K))


def test_multi_chunk_plan_matches_single_chunk_and_oracle(gpu, monkeypatch):
    # The residual planes of a plan are split into memory chunks when they
    # exceed the plane budget (RFBA3B_PLANE_BUDGET_MB); chunk-local plane,
    # item and correlation offsets restart in every chunk.  A budget of 1 MiB
    # over uneven blocks of 0.1-0.55 MiB of planes each gives three chunks of
    # two or three blocks; results must equal the single-chunk plan bit for bit
    # and the oracle within the bars, in every launch form.
    m = model(checkpoint=11)
    rng = np.random.default_rng(17)
    blocks = []
    for i in range(7):
        dims = tuple(int(v) for v in rng.integers(7, 14, 3))
        T = int(rng.integers(20, 41))
        K = int(rng.integers(3, 12))
        data, params, states = b3.synth_block(dims, T, K, 300 + i)
        S = int(rng.integers(1, 3))
        st = np.stack([np.roll(states, 2 * s, axis=1) for s in range(S)])
        blocks.append(b3.BlockData(data, b3.perturb(params, int(rng.integers(1, 4)), 400 + i), st, 100.0))
    monkeypatch.delenv("RFBA3B_PLANE_BUDGET_MB", raising=False)
    one = b3.Batch(blocks, m)
    assert one.chunks == 1
    one.run(use_graph=False)
    ref = [one.results(i) for i in range(len(blocks))]
    one.close()
    monkeypatch.setenv("RFBA3B_PLANE_BUDGET_MB", "1")
    bt = b3.Batch(blocks, m)
    print("chunks at 1 MiB: %d" % bt.chunks)
    assert bt.chunks >= 3
    for mode in ("plain", "graph", "per_block", "per_param_set", "graph2"):
        if mode.startswith("per_"):
            bt.run_per_block(per_param_set=(mode == "per_param_set"))
        else:
            bt.run(use_graph=mode.startswith("graph"))
        for i in range(len(blocks)):
            L, G = bt.results(i)
            assert np.array_equal(L, ref[i][0]) and np.array_equal(G, ref[i][1]), (mode, i)
    bt.close()
    for b, (L, G) in zip(blocks, ref):
        Lo, Go = bo.model_eval(b.data, 100.0, m, b.params, b.states)
        compare(L, G, Lo, Go, "chunked block %s S=%d" % (b.data.shape, b.states.shape[0]))


def _functor_case(dims, T, K, S, seed, region="ball"):
    """Data [T, P], backgrounds [S, T, P] (FP64: b0 + the other spots at
    shifted sampled states) and the pixel list of a region around spot 0."""
    data, params, states = b3.synth_block(dims, T, K, seed)
    xyz_all = grid(*dims)
    if region == "ball":
        keep = np.linalg.norm((xyz_all - params[0, 2:5]) * [1, 1, 0.5], axis=1) < 4.5
    else:
        keep = np.ones(len(xyz_all), bool)
    xyz = xyz_all[keep]
    d2 = data.reshape(T, -1)[:, keep].astype(np.float64)
    spots = [bo.spot(xyz, params[j], 2.0)[0] for j in range(K)]
    bgs = np.empty((S, T, len(xyz)))
    for s in range(S):
        bgs[s] = 100.0
        for j in range(1, K):
            bgs[s] += (np.roll(states[j], 3 * s) == 0)[:, None] * spots[j][None]
    return xyz, d2, bgs, params


@pytest.mark.parametrize("region,S", [("ball", 3), ("box", 10)])
def test_resident_functor_vs_oracle(gpu, region, S):
    # SampledBackgroundData created once, then the CG functor called with
    # changing candidate lists: FP64 data and backgrounds, the residual
    # rounded once to FP32 at creation; against the FP64 oracle of
    # multispot5.cc:616-669 at the batched form's bars
    m = model()
    xyz, d2, bgs, params = _functor_case((11, 10, 9), 24, 4, S, 81, region)
    f = b3.Functor(xyz, d2, bgs, m, max_candidates=12)
    rng = np.random.default_rng(5)
    base = params[0]
    cands = base + np.concatenate([[np.zeros(5)], rng.uniform(-0.3, 0.3, (10, 5)) * [300, 0.2, 1, 1, 1]])
    calls = [cands[:1], np.concatenate([cands[1:4], [[0.0, 1.3, 3, 3, 3], [900, -1, 3, 3, 3]]]), cands[1:11],
             cands[5:6]]
    for c, phis in enumerate(calls):
        g, L = f(phis)
        for i, phi in enumerate(phis):
            if not (phi[0] > 0 and phi[1] > 0):
                assert np.isnan(g[i]).all() and np.isnan(L[i]).all()
                continue
            go, Lo = bo.neg_grad(xyz, d2, bgs, phi, m)
            eg = np.linalg.norm(g[i] - go) / np.linalg.norm(go)
            el = np.max(np.abs(L[i] - Lo) / np.abs(Lo))
            assert eg <= GRAD_RTOL and el <= LOGL_RTOL, (c, i, eg, el)
        print("%s S=%d call %d (n=%d): ok" % (region, S, c, len(phis)))
    with pytest.raises(b3.B3Error):
        f(np.tile(base, (13, 1)))  # more candidates than the handle holds
    f.close()


HESS_RTOL = 1e-4


@pytest.mark.parametrize("region,S,bs,z_exact", [("ball", 3, 5, 0), ("box", 2, 3, 1)])
def test_functor_hessian_ffbs_vs_oracle(gpu, region, S, bs, z_exact):
    # sampled_background_spot_hessian_ffbs for the handle's data: the paths
    # drawn on the device from the device's forward messages equal the
    # oracle's for the same draws; the Hessian matches the FP64 oracle within
    # 1e-4 of its largest entry (FP32 residuals and spot tables), in the
    # reference's mode (z terms without 1/zF^2, entry [4][3] unset -> 0) and
    # the exact one
    m = model(z_exact=z_exact)
    xyz, d2, bgs, params = _functor_case((11, 10, 9), 30, 4, S, 91, region)
    f = b3.Functor(xyz, d2, bgs, m, max_candidates=2)
    phi = params[0] + [40.0, 0.05, 0.2, -0.1, 0.15]
    for call, nb in enumerate((bs, bs + 2)):  # the second call grows the path buffers
        u = np.random.default_rng(30 + call).random((S, nb, 30))
        H, D, paths = f.hessian_ffbs(phi, nb, u)
        Ho, Do, po = bo.hessian_ffbs(xyz, d2, bgs, phi, m, nb, u)
        assert np.array_equal(paths, po), "sampled paths differ"
        scale = np.abs(Ho).max()
        err = np.abs(H - Ho).max() / scale
        errd = np.abs(D - Do).max() / np.abs(Do).max()
        print("%s S=%d bs=%d z_exact=%d: Hessian err %.3g (max-normalised), deriv err %.3g, %d on-frames sampled"
              % (region, S, nb, z_exact, err, errd, int((paths == 0).sum())))
        assert err <= HESS_RTOL and errd <= HESS_RTOL
        if not z_exact:
            assert H[4, 3] != H[3, 4]
        else:
            np.testing.assert_allclose(H, H.T, rtol=0, atol=1e-12 * scale)
    # the CG functor still runs on the same handle
    g, L = f(phi)
    go, Lo = bo.neg_grad(xyz, d2, bgs, phi, m)
    assert np.linalg.norm(g[0] - go) / np.linalg.norm(go) <= GRAD_RTOL
    f.close()


def test_series_longer_than_wave_frames(gpu):
    # T = 600 is past the 512 frames the one-wave-per-item HMM spreads over
    # its lanes: the HMMs run one lane per item (serial chain over the frames)
    m = model(checkpoint=0)
    data, params, states = b3.synth_block((7, 6, 5), 600, 3, 91, A=b3.blink_matrix(0.2, 0.1, 0.001))
    P = b3.perturb(params, 2, 92)
    (L, G), = run_batch([b3.BlockData(data, P, states[None], 100.0)], m)
    Lo, Go = bo.model_eval(data, 100.0, m, P, states[None])
    compare(L, G, Lo, Go, "T=600 K=3")


def test_many_items_one_launch(gpu):
    # 4200 items (105 parameter sets x 40 spots) in one launch of the
    # one-wave-per-item HMM, against one launch per parameter set (40 items)
    # bit for bit, and the oracle on two parameter sets
    m = model(checkpoint=0)
    data, params, states = b3.synth_block((8, 8, 8), 130, 40, 93)
    P = b3.perturb(params, 105, 94)
    bt = b3.Batch([b3.BlockData(data, P, states[None], 100.0)], m)
    bt.run(use_graph=False)
    L1, G1 = bt.results(0)
    bt.run_per_block(per_param_set=True)
    L2, G2 = bt.results(0)
    bt.close()
    assert np.array_equal(L1, L2) and np.array_equal(G1, G2)
    Lo, Go = bo.model_eval(data, 100.0, m, P[:2], states[None])
    compare(L1[:2], G1[:2], Lo, Go, "8^3 T=130 K=40")


@pytest.mark.parametrize("T", [130, 600])
def test_functor_long_series(gpu, T):
    # T = 130: three frames per lane in the wave HMM; T = 600: the lane form
    # (gradient and the Hessian's forward messages)
    m = model()
    xyz, d2, bgs, params = _functor_case((9, 8, 7), T, 3, 2, 95, "ball")
    f = b3.Functor(xyz, d2, bgs, m, max_candidates=2)
    phi = params[0] + [40.0, 0.05, 0.2, -0.1, 0.15]
    g, L = f(np.stack([params[0], phi]))
    for i, p in enumerate((params[0], phi)):
        go, Lo = bo.neg_grad(xyz, d2, bgs, p, m)
        eg = np.linalg.norm(g[i] - go) / np.linalg.norm(go)
        el = np.max(np.abs(L[i] - Lo) / np.abs(Lo))
        print("T=%d candidate %d: grad err %.3g, logL err %.3g" % (T, i, eg, el))
        assert eg <= GRAD_RTOL and el <= LOGL_RTOL
    u = np.random.default_rng(96).random((2, 3, T))
    H, D, paths = f.hessian_ffbs(phi, 3, u)
    Ho, Do, po = bo.hessian_ffbs(xyz, d2, bgs, phi, m, 3, u)
    assert np.array_equal(paths, po), "sampled paths differ"
    err = np.abs(H - Ho).max() / np.abs(Ho).max()
    print("T=%d Hessian err %.3g, %d on-frames sampled" % (T, err, int((paths == 0).sum())))
    assert err <= HESS_RTOL
    f.close()
