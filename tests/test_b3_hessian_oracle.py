"""CPU tests that pin the oracle's FFBS Hessian (oracle/b3_oracle.cpp:
spot_hess, hess_parts, oracle_b3_hessian_paths / _ffbs), the restatement of
sampled_background_spot_hessian_ffbs (multispot5.cc:902-982) with
spot_shape_hess_position (storm.h:86-150).

* The spot Hessian in exact mode equals central finite differences of the
  spot gradient (itself pinned by finite differences of the spot value).
  In reference mode, entries without z equal the exact ones; entries with
  z drop the 1/zF^2 as the reference does (closed-form relations to the
  exact entries); [4][3] is 0 (the reference leaves it uninitialised).
* With one background sample and unit variance (the reference normalises
  its data to variance 1, multispot5.cc:1440), the reference's estimator
  over the exact posterior of the state paths is the Hessian of
  logL(phi) + log prior(phi): it must equal central differences of the
  oracle's gradient (pinned by finite differences and path enumeration).
* Two identical background samples give the one-sample result; the
  sampled estimator with many draws approaches the exact-posterior value;
  the sampled paths are the backward sampler's.
"""
import itertools

import numpy as np
import pytest

from oracle import b3oracle as bo
from rfba_amd import b3


def grid(X, Y, Z):
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(np.int32)


def model(**kw):
    base = dict(A=b3.blink_matrix(), pi=(0.5, 0.5, 0.0), variance=1.0, zfactor=2.0, z_exact=1, checkpoint=0,
                mu_brightness=2.0, sigma_brightness=1.0, mu_blur=0.25, sigma_blur=0.5)
    base.update(kw)
    return b3.make_model(**base)


PHI = np.array([900.0, 1.4, 2.3, 2.7, 3.1])


def test_spot_hessian_exact_vs_fd():
    xyz = grid(6, 5, 7)
    _, _, H = bo.spot_hess(xyz, PHI, 2.0, 1)
    for n in range(5):
        h = 1e-6 * max(1.0, abs(PHI[n]))
        e = np.zeros(5)
        e[n] = h
        fd = (bo.spot(xyz, PHI + e, 2.0, 1)[1] - bo.spot(xyz, PHI - e, 2.0, 1)[1]) / (2 * h)  # d(dmu)/dphi_n
        np.testing.assert_allclose(H[:, :, n], fd, rtol=1e-5, atol=1e-7 * np.abs(fd).max())
    np.testing.assert_allclose(H, H.transpose(0, 2, 1), rtol=0, atol=0)
    mu, dmu, _ = bo.spot_hess(xyz, PHI, 2.0, 1)
    mu0, dmu0 = bo.spot(xyz, PHI, 2.0, 1)
    np.testing.assert_allclose(mu, mu0, rtol=1e-15)
    np.testing.assert_allclose(dmu, dmu0, rtol=1e-14, atol=1e-300)


@pytest.mark.parametrize("zf", [1.0, 2.0])
def test_spot_hessian_reference_mode(zf):
    xyz = grid(6, 5, 7)
    mu, _, Hr = bo.spot_hess(xyz, PHI, zf, 0)
    _, _, He = bo.spot_hess(xyz, PHI, zf, 1)
    z2 = zf * zf
    np.testing.assert_allclose(Hr[:, :4, :4], He[:, :4, :4], rtol=1e-14, atol=1e-300)
    for i in range(4):
        np.testing.assert_allclose(Hr[:, i, 4], He[:, i, 4] * z2, rtol=1e-12, atol=1e-300)
    for i in range(3):
        np.testing.assert_allclose(Hr[:, 4, i], He[:, 4, i] * z2, rtol=1e-12, atol=1e-300)
    assert np.all(Hr[:, 4, 3] == 0)  # storm.h:143 leaves [4][3] unset
    # [4][4]: I e^s (dz^2 - sigma^2)/(sigma^5 sqrt(2 pi)) = zF^4 exact + mu (zF^2 - 1) / sigma^2
    np.testing.assert_allclose(Hr[:, 4, 4], He[:, 4, 4] * z2 * z2 + mu * (z2 - 1) / PHI[1] ** 2, rtol=1e-10,
                               atol=1e-12 * np.abs(Hr[:, 4, 4]).max())


def test_hess_log_log_normal_vs_fd():
    for x, mu, sg in ((900.0, 7.0, 1.0), (1.3, 0.25, 0.1), (0.2, -1.0, 2.0)):
        h = 1e-6 * x
        fd = (bo.lib().oracle_b3_diff_log_log_normal(x + h, mu, sg)
              - bo.lib().oracle_b3_diff_log_log_normal(x - h, mu, sg)) / (2 * h)
        assert bo.hess_log_log_normal(x, mu, sg) == pytest.approx(fd, rel=1e-6)


def small_case(seed, T=4, dims=(5, 4, 4)):
    """One dim spot over a constant background with unit-variance noise, so
    that the posterior over its blink paths is spread over many paths."""
    rng = np.random.default_rng(seed)
    xyz = grid(*dims)
    phi = np.array([6.0, 1.2, 2.2, 1.6, 1.7])
    mu, _ = bo.spot(xyz, phi, 2.0, 1)
    on = rng.random(T) < 0.5
    data = 10.0 + on[:, None] * mu[None] + rng.normal(0, 1, (T, len(xyz)))
    bg = np.full((1, T, len(xyz)), 10.0)
    return xyz, data, bg, phi + [0.8, 0.05, 0.2, -0.15, 0.1]


def posterior(xyz, data, bg, phi, m):
    """Exact posterior of all 3^T paths (enumeration of the HMM joint)."""
    mu, dmu = bo.spot(xyz, phi, m.zfactor, m.z_exact)
    on, off, _ = bo.emissions(data, bg, mu, dmu, m.variance)
    A, pi = np.array(m.A[:]).reshape(3, 3), np.array(m.pi[:])
    T = len(on)
    paths, lw = [], []
    for path in itertools.product(range(3), repeat=T):
        lp = np.log(pi[path[0]]) if pi[path[0]] > 0 else -np.inf
        for t in range(1, T):
            a = A[path[t - 1], path[t]]
            lp += np.log(a) if a > 0 else -np.inf
        lp += sum(on[t] if s == 0 else off[t] for t, s in enumerate(path))
        paths.append(path)
        lw.append(lp)
    lw = np.array(lw)
    w = np.exp(lw - lw.max())
    return np.array(paths, np.uint8), w / w.sum()


def test_exact_posterior_estimator_is_the_hessian_of_logl():
    m = model()
    xyz, data, bg, phi = small_case(3)
    paths, w = posterior(xyz, data, bg[0], phi, m)
    H, D = bo.hessian_paths(xyz, data, bg, phi, m, paths[None], w[None])

    def grad(p):
        g = bo.logl_grad(xyz, data, bg[0], p, m)[1].copy()
        g[0] += bo.lib().oracle_b3_diff_log_log_normal(p[0], m.mu_brightness, m.sigma_brightness)
        g[1] += bo.lib().oracle_b3_diff_log_log_normal(p[1], m.mu_blur, m.sigma_blur)
        return g

    fd = np.empty((5, 5))
    for n in range(5):
        h = 1e-5 * max(1.0, abs(phi[n]))
        e = np.zeros(5)
        e[n] = h
        fd[:, n] = (grad(phi + e) - grad(phi - e)) / (2 * h)
    np.testing.assert_allclose(H, fd, rtol=1e-5, atol=1e-6 * np.abs(fd).max())
    np.testing.assert_allclose(D, grad(phi), rtol=1e-9, atol=1e-12 * np.abs(D).max())
    # more than one path carries weight: the outer-product term matters
    assert np.sort(w)[-2] > 1e-3


def test_identical_samples_and_sampled_estimator():
    m = model()
    xyz, data, bg, phi = small_case(5)
    paths, w = posterior(xyz, data, bg[0], phi, m)
    H1, _ = bo.hessian_paths(xyz, data, bg, phi, m, paths[None], w[None])
    bg2 = np.concatenate([bg, bg])
    H2, _ = bo.hessian_paths(xyz, data, bg2, phi, m, np.stack([paths, paths]), np.stack([w, w]))
    np.testing.assert_allclose(H2, H1, rtol=1e-12, atol=1e-12 * np.abs(H1).max())
    # the reference's sampled form: bs backward samplings per background
    bs, T = 6000, data.shape[0]
    u = np.random.default_rng(9).random((1, bs, T))
    Hs, _, st = bo.hessian_ffbs(xyz, data, bg, phi, m, bs, u)
    np.testing.assert_allclose(Hs, H1, rtol=0.05, atol=0.05 * np.abs(H1).max())
    # the paths are the backward sampler's on the spot's forward messages
    mu, dmu = bo.spot(xyz, phi, m.zfactor, m.z_exact)
    on, off, _ = bo.emissions(data, bg[0], mu, dmu, m.variance)
    delta, _ = bo.forward_delta(np.array(m.A[:]).reshape(3, 3), np.array(m.pi[:]), on, off)
    for it in (0, 17, bs - 1):
        assert np.array_equal(st[0, it], bo.backward_sampling(np.array(m.A[:]).reshape(3, 3), delta, u[0, it]))
