"""CPU tests of the Gibbs sampling pieces: the spot masks (product host code
and oracle restatement of get_spot_pixels, multispot5.cc:1001-1020) and the
oracle's GibbsSampler2::next (sampled_multispot.h:681-709).

The reference ships no vectors for this path.  The masks are pinned by a
brute-force restatement in numpy (every pixel within the ball's float radius
of the rounded position) and hand cases.  The sampler is pinned by what a
Gibbs sampler must satisfy: the sample intensities it carries equal the sum
of the on spots over their masks (remove_spot/add_spot bookkeeping); with
empty masks each spot's path is drawn from the HMM prior; and its stationary
distribution over the joint state paths of two overlapping spots equals the
exact posterior obtained by enumerating all 27 x 27 joint paths.
"""
import itertools

import numpy as np

from oracle import b3oracle as bo
from rfba_amd import b3

A_REF = np.array([[0.16, 0.84, 0.0], [0.495, 0.495, 0.01], [0.0, 0.0, 1.0]])  # multispot5.cfg:15
PI_REF = np.array([0.5, 0.5, 0.0])


def brute_mask(xyz, phi):
    radius = np.float32(phi[1] * 6 + 1)
    off = [int(v + 0.5) if v > 0 else int(v - 0.5) for v in phi[2:5]]
    d = xyz - np.array(off)
    m2 = (d * d).sum(1).astype(np.float32)
    return np.nonzero(m2 <= radius * radius)[0].astype(np.int32)


def test_spot_pixels_brute_force_and_oracle():
    rng = np.random.default_rng(3)
    xyz = b3.box_pixels((13, 11, 9), (5, -3, 2))
    # a region with holes: every third pixel removed
    keep = np.ones(len(xyz), bool)
    keep[::3] = False
    for pix in (xyz, xyz[keep]):
        phi = np.column_stack([rng.uniform(100, 1000, 12), rng.uniform(0.1, 1.2, 12),
                               rng.uniform(0, 18, 12), rng.uniform(-6, 9, 12), rng.uniform(0, 12, 12)])
        phi[0, 1:] = [0.5, 11.5, 2.5, 6.5]    # radius exactly 4 (ties at |d|^2 = 16), .5 rounding up
        phi[1, 1:] = [0.25, -0.5, -3.5, 2.0]  # negative .5 rounds away from zero
        off, idx = b3.spot_pixels(pix, phi)
        ooff, oidx = bo.spot_pixels(pix, phi)
        assert np.array_equal(off, ooff) and np.array_equal(idx, oidx)
        for k in range(len(phi)):
            assert np.array_equal(idx[off[k]:off[k + 1]], brute_mask(pix, phi[k])), k
    # hand case: sigma 0 -> radius 1: the centre and its 6 face neighbours
    off, idx = b3.spot_pixels(b3.box_pixels((3, 3, 3)), np.array([[1.0, 0.0, 1.2, 0.8, 1.4]]))
    assert off[1] == 7 and set(idx.tolist()) == {13, 12, 14, 10, 16, 4, 22}
    # a spot outside the region has an empty mask
    off, idx = b3.spot_pixels(b3.box_pixels((3, 3, 3)), np.array([[1.0, 0.1, 40.0, 0.0, 0.0]]))
    assert off[1] == 0


def test_spot_pixels_rejects_unsorted_region():
    xyz = b3.box_pixels((3, 3, 3))[::-1].copy()
    try:
        b3.spot_pixels(xyz, np.array([[1.0, 0.5, 1.0, 1.0, 1.0]]))
    except b3.B3Error as e:
        assert "increasing" in str(e)
    else:
        raise AssertionError("unsorted region accepted")


def small_problem(seed, T=3, I=1.13, var=1.0):
    """Two overlapping narrow spots on a 3x3x1 region with data drawn from a
    random joint path plus unit noise."""
    rng = np.random.default_rng(seed)
    xyz = b3.box_pixels((3, 3, 1))
    phi = np.array([[I, 0.3, 1.0, 1.0, 0.0], [I, 0.3, 1.4, 1.2, 0.0]])
    spot = np.stack([bo.spot(xyz, p)[0] for p in phi])
    off, idx = bo.spot_pixels(xyz, phi)
    on = rng.uniform(size=(2, T)) < 0.5
    data = (on[:, :, None] * spot[:, None, :]).sum(0) + rng.normal(0, np.sqrt(var), (T, len(xyz)))
    return xyz, spot, off, idx, data


def joint_posterior(spot, data, var, T):
    paths = list(itertools.product(range(3), repeat=T))
    prior = np.array([PI_REF[p[0]] * np.prod([A_REF[p[t], p[t + 1]] for t in range(T - 1)]) for p in paths])
    onp = np.array([[s == 0 for s in p] for p in paths], float)  # [27, T]
    logp = np.full((len(paths), len(paths)), -np.inf)
    for i, j in itertools.product(range(len(paths)), repeat=2):
        if prior[i] == 0 or prior[j] == 0:
            continue
        mu = onp[i][:, None] * spot[0] + onp[j][:, None] * spot[1]
        logp[i, j] = np.log(prior[i]) + np.log(prior[j]) - ((data - mu) ** 2).sum() / (2 * var)
    p = np.exp(logp - logp.max())
    return paths, p / p.sum()


def test_gibbs_oracle_bookkeeping():
    xyz, spot, off, idx, data = small_problem(1, T=5)
    T, P = data.shape
    sample = np.full((2, T), 2, np.uint8)
    intens = np.zeros((T, P))
    u = np.random.default_rng(2).uniform(size=50 * 2 * T)
    used = bo.gibbs_next(data, spot, off, idx, A_REF, PI_REF, 1.0, 50, sample, intens, u)
    assert used == 50 * 2 * T
    expect = np.zeros((T, P))
    for k in range(2):
        m = idx[off[k]:off[k + 1]]
        for t in range(T):
            if sample[k, t] == 0:
                expect[t, m] += spot[k, m]
    assert np.allclose(intens, expect, rtol=0, atol=1e-12)


def test_gibbs_oracle_empty_masks_draw_the_prior():
    # no pixels: each spot's path is a draw from the blink chain itself
    T = 2
    data = np.zeros((T, 4))
    spot = np.ones((1, 4))
    off, idx = np.array([0, 0], np.int32), np.zeros(0, np.int32)
    rng = np.random.default_rng(4)
    N = 20000
    u = rng.uniform(size=N * T)
    sample = np.full((1, T), 2, np.uint8)
    intens = np.zeros((T, 4))
    counts = {}
    for n in range(N):
        bo.gibbs_next(data, spot, off, idx, A_REF, PI_REF, 1.0, 1, sample, intens, u[n * T:(n + 1) * T])
        key = tuple(sample[0])
        counts[key] = counts.get(key, 0) + 1
    for (a, b), c in counts.items():
        assert abs(c / N - PI_REF[a] * A_REF[a, b]) < 0.015, (a, b, c / N)
    assert set(counts) <= {(a, b) for a in range(3) for b in range(3) if PI_REF[a] * A_REF[a, b] > 0}


def test_gibbs_oracle_stationary_distribution_is_the_posterior():
    T, var = 3, 1.0
    xyz, spot, off, idx, data = small_problem(5, T=T)
    paths, post = joint_posterior(spot, data, var, T)
    index = {p: i for i, p in enumerate(paths)}
    rng = np.random.default_rng(6)
    burn, N = 200, 30000
    u = rng.uniform(size=(burn + N) * 2 * T)
    sample = np.full((2, T), 2, np.uint8)
    intens = np.zeros(data.shape)
    emp = np.zeros_like(post)
    step = 2 * T
    for n in range(burn + N):
        bo.gibbs_next(data, spot, off, idx, A_REF, PI_REF, var, 1, sample, intens, u[n * step:(n + 1) * step])
        if n >= burn:
            emp[index[tuple(sample[0])], index[tuple(sample[1])]] += 1
    emp /= N
    # the posterior is spread out (the test discriminates) ...
    assert post.max() < 0.5
    # ... and the sampler's long-run frequencies match it
    tv = 0.5 * np.abs(emp - post).sum()
    assert tv < 0.06, tv
    onp = np.array([[s == 0 for s in p] for p in paths], float)
    for k in range(2):
        marg = post.sum(1 - k)
        emarg = emp.sum(1 - k)
        assert np.allclose(onp.T @ emarg, onp.T @ marg, atol=0.03), (k, onp.T @ emarg, onp.T @ marg)
