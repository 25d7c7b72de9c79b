"""Pins the oracle's numerical building blocks against mathematics that fixes them."""
import math

import numpy as np
import pytest


@pytest.mark.parametrize("order", [3, 4, 5, 6, 8])
def test_bspline_partition_symmetry_derivative(orc, order):
    rng = np.random.default_rng(order)
    for pos in rng.uniform(-50, 50, size=40):
        k, v, d = orc.bspline(order, pos)
        assert k == math.floor(pos)
        x = pos - k
        assert abs(v.sum() - 1.0) < 1e-14                  # partition of unity
        assert abs(d.sum()) < 1e-13                        # derivatives sum to zero
        # symmetry M_n(u) = M_n(n - u): value at x+i equals value at (n - x - i)
        _, v2, _ = orc.bspline(order, math.floor(pos) + (1.0 - x) if x > 0 else pos)
        if x > 0:
            assert np.allclose(v, v2[::-1], atol=1e-14)
        # finite-difference derivative
        h = 1e-6
        if 1e-5 < x < 1 - 1e-5:
            _, vp, _ = orc.bspline(order, pos + h)
            _, vm, _ = orc.bspline(order, pos - h)
            assert np.allclose((vp - vm) / (2 * h), d, atol=1e-7)


def test_bspline_cubic_closed_form(orc):
    # M_4(u) on [0,1): u^3/6, [1,2): (-3u^3+12u^2-12u+4)/6, [2,3): (3u^3-24u^2+60u-44)/6,
    # [3,4): (4-u)^3/6
    def m4(u):
        if u < 1: return u ** 3 / 6
        if u < 2: return (-3 * u ** 3 + 12 * u ** 2 - 12 * u + 4) / 6
        if u < 3: return (3 * u ** 3 - 24 * u ** 2 + 60 * u - 44) / 6
        return (4 - u) ** 3 / 6
    for x in np.linspace(0.01, 0.99, 17):
        _, v, _ = orc.bspline(4, 7.0 + x)
        assert np.allclose(v, [m4(x + i) for i in range(4)], atol=1e-15)


@pytest.mark.parametrize("K", [8, 16, 32, 10, 18, 26, 40, 42, 48, 50])
def test_fft3_matches_numpy(orc, K):
    rng = np.random.default_rng(K)
    q = rng.normal(size=(K, K, K))          # q[z, y, x], x fastest
    f = orc.fft3_r2c(q)
    ref = np.fft.rfftn(q)                    # same axis convention: last axis halved
    assert np.allclose(f, ref, rtol=1e-12, atol=1e-9)
    back = orc.fft3_c2r(ref, K)
    assert np.allclose(back, np.fft.irfftn(ref, s=(K, K, K), axes=(0, 1, 2)) * K ** 3, rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("K,order", [(16, 6), (48, 5), (40, 6), (40, 11)])
def test_spme_setup_closed_form(orc, K, order):
    box = np.array([30.0, 30.0, 30.0])
    alpha = 0.35
    mf, ef, vf = orc.spme_setup(box, alpha, K, order)
    H = K // 2 + 1
    # b^2(m) = 1/|sum_k M_p(k+1) exp(2 pi i m k / K)|^2 from numpy's FFT of the spline
    _, spl, _ = orc.bspline(order, 0.0)
    arr = np.zeros(K)
    arr[: order - 1] = spl[1:order]
    with np.errstate(divide="ignore"):
        b2 = 1.0 / np.abs(np.fft.ifft(arr) * K) ** 2
    if order % 2:
        b2[K // 2] = 0.0  # the spline's transform vanishes there for odd orders (set to 0)
    n = np.arange(H)
    kx = 2 * np.pi * n / box[0]
    M2 = kx[None, None, :] ** 2 + kx[None, :, None] ** 2 + kx[:, None, None] ** 2
    with np.errstate(divide="ignore", invalid="ignore"):
        ref = 4 * np.pi / box.prod() * np.exp(-M2 / (4 * alpha ** 2)) / M2
    ref *= b2[n][None, None, :] * b2[n][None, :, None] * b2[n][:, None, None]
    ref[0, 0, 0] = 0
    assert np.allclose(mf.reshape(H, H, H), ref, rtol=1e-12, atol=0)
    w = np.ones(H); w[0] = 0.5; w[-1] = 0.5
    W = w[None, None, :] * w[None, :, None] * w[:, None, None]
    assert np.allclose(ef.reshape(H, H, H), ref * W, rtol=1e-12, atol=0)


def test_heapsort_index_sorts(orc):
    rng = np.random.default_rng(0)
    for n in (1, 2, 5, 27, 100):
        v = rng.integers(0, 4, size=n).astype(float)
        idx = orc.heapsort_index(v)
        assert sorted(idx) == list(range(1, n + 1))
        assert np.all(np.diff(v[idx - 1]) >= 0)


def test_random_structure(orc):
    r = orc.Rng(-12345)
    xs = np.array([r() for _ in range(20000)])
    assert (xs > 0).all() and (xs < 1).all()
    assert abs(xs.mean() - 0.5) < 0.01
    r2 = orc.Rng(-12345)
    assert np.array_equal(xs[:50], [r2() for _ in range(50)])


def test_openmp_build_agrees_with_the_oracle():
    """liboracle_omp.so (the multi-core CPU baseline's timing build) computes what the oracle
    does; only the summation order of the shared-out loops differs."""
    import oracle
    from molsim_amd import systems
    sp = systems.small_spme_system(seed=2, n=300, L=40.0, nmesh=32, alpha=0.3, rcut=12.0)
    tab = oracle.build_table(sp.pot)
    osys = oracle.System(sp.ro, sp.ipt, sp.az, sp.box, sp.rcut, sp.pot["epsi4pi"], tab, npt=sp.npt,
                         list_mode=1, ewald=2, surf=sp.lsurf, alpha=sp.alpha, K=sp.nmesh, order=sp.order)
    a, b = osys.utotal(), osys.utotal(omp=True)
    assert oracle.threads() == 1
    scale = abs(a["u_tot"]) + abs(a["u_rec"]) + np.abs(a["u_twob"]).max()
    assert abs(a["u_tot"] - b["u_tot"]) < 1e-12 * scale
    assert np.abs(a["force"] - b["force"]).max() < 1e-12 * np.abs(a["force"]).max()
