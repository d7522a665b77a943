"""Pins for corners of the oracle that whole-path checks leave open (CPU).

Each pin is checked twice: the oracle passes it, and a copy of the oracle
built with a planted mistake (oracle.MUTANTS, NMCO_MUTANT in nmc_oracle.c)
fails it, so the pin is shown to see that mistake.

- Source models (kernels.metal:106-166 mcsub, :366-463 mcxyz): launch states
  against the distributions the sampling formulas imply - uniform disk
  (r/R)^2 ~ U(0,1), Gaussian (r/R)^2 ~ Exp(1), isotropic cos ~ U(-1,1),
  rectangle, point, manual launch, the focus point of a focused beam, Snell
  refraction and the Fresnel specular loss.
- Step-length carry-over across tissue boundaries (kernels.metal:505-685,
  `sleft` rescaled by mus): with g = 1 and mua = 0 a photon crosses each
  tissue on a straight line, so its scattering count is Poisson with mean
  sum_i mus_i L_i - exactly, whatever the voxel faces in between.
- mcsub depth bins of A(r,z) (kernels.metal:224-234): summed over r they
  estimate the absorption per depth slab, which the layered walk (continuous
  absorption, an independent estimator) gives per layer for a stack of
  slabs of the same medium.
- mcsub exit-angle bins of Rd(r,alpha) (kernels.metal:197-204): for an
  isotropically scattering semi-infinite medium with a matched boundary
  under normal incidence the escaping flux per exit cosine mu is
  proportional to mu H(mu) / (1 + mu) (Chandrasekhar, Radiative Transfer,
  1950, ch. IV), with his H-function.
"""
from __future__ import annotations

import math

import numpy as np
import pytest
from scipy import stats

from nmc_amd import params as P
from nmc_amd import workloads as W
from oracle import oracle as O

# ------------------------------------------------------------------ fixtures
@pytest.fixture(scope="module")
def mutants(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("oracle_mutants"))
    return lambda k: O.mutant(k, d)


def _ks_uniform(x, lo=0.0, hi=1.0) -> float:
    return stats.kstest((np.asarray(x, np.float64) - lo) / (hi - lo), "uniform").pvalue


def _ks_exp(x) -> float:
    return stats.kstest(np.asarray(x, np.float64), "expon").pvalue


def _xyz(kind, **kw):
    w = W.tiny_voxel_case(kind, n=16)
    for k, v in kw.items():
        setattr(w.rp, k, v)
    return w


N_SRC = 100_000


# ---------------------------------------------------------- source models ---
def flat_beam_checks(s, rp) -> dict:
    """p-values / flags of the flat (uniform disk) focused beam, mcflag 0."""
    x, y, z, ux, uy, uz = (s[:, i].astype(np.float64) for i in range(6))
    r2 = ((x - rp.xs) ** 2 + (y - rp.ys) ** 2) / rp.radius ** 2
    phi = np.arctan2(y - rp.ys, x - rp.xs)
    # the ray from the launch point reaches depth zfocus below it at the
    # focus point (xf, yf), drawn uniformly on the waist disk
    xf = x + ux / uz * rp.zfocus
    yf = y + uy / uz * rp.zfocus
    rf2 = (xf ** 2 + yf ** 2) / rp.waist ** 2
    return {"r2": _ks_uniform(r2), "phi": _ks_uniform(phi, -math.pi, math.pi),
            "rf2": _ks_uniform(rf2), "phif": _ks_uniform(np.arctan2(yf, xf), -math.pi, math.pi),
            "indep": abs(np.corrcoef(r2, rf2)[0, 1]),
            "z": bool(np.all(z == np.float32(rp.zs))),
            "unit": float(np.abs(ux ** 2 + uy ** 2 + uz ** 2 - 1).max()),
            "w": bool(np.all(s[:, 6] == 1.0) and np.all(s[:, 7] == 0.0))}


FLAT = dict(mcflag=0, radius=0.06, waist=0.02, zfocus=0.5, xs=0.01, ys=-0.02, zs=0.0201)


def test_mcxyz_flat_beam_distribution(mutants):
    w = _xyz("mismatch", **FLAT)
    c = flat_beam_checks(O.xyz_source(w, N_SRC), w.rp)
    assert min(c["r2"], c["phi"], c["rf2"], c["phif"]) > 1e-3, c
    assert c["indep"] < 0.02 and c["z"] and c["w"] and c["unit"] < 1e-6, c
    bad = flat_beam_checks(O.xyz_source(w, N_SRC, L=mutants(1)), w.rp)
    assert bad["r2"] < 1e-12, "the flat-beam pin misses r = R*xi"


def test_mcxyz_gaussian_beam_refraction_and_specular():
    """mcflag 1: (r/R)^2 ~ Exp(1); the launch direction is the refraction
    (n2 -> n1) of the ray towards a focus point uniform on the waist disk;
    the specular loss is the textbook Fresnel reflectance of that ray."""
    n1, n2 = 1.37, 1.0
    w = _xyz("gaussian", mcflag=1, radius=0.03, waist=0.02, zfocus=0.1, xs=0.0, ys=0.0, zs=0.0,
             n1=n1, n2=n2)
    s = O.xyz_source(w, N_SRC).astype(np.float64)
    x, y, z, ux, uy, uz, W_, spec = (s[:, i] for i in range(8))
    assert _ks_exp((x ** 2 + y ** 2) / 0.03 ** 2) > 1e-3
    np.testing.assert_allclose(W_ + spec, 1.0, atol=1e-6)
    # undo Snell: the incident ray has transverse part (ux, uy) * n1 / n2
    sin_t = np.hypot(ux, uy)
    sin_i = sin_t * n1 / n2
    cos_i = np.sqrt(1 - sin_i ** 2)
    xf = x + ux * n1 / n2 / cos_i * 0.1
    yf = y + uy * n1 / n2 / cos_i * 0.1
    assert _ks_uniform((xf ** 2 + yf ** 2) / 0.02 ** 2) > 1e-3
    cos_t = np.sqrt(1 - sin_t ** 2)
    np.testing.assert_allclose(uz, cos_t, atol=2e-6)
    rs = ((n2 * cos_i - n1 * cos_t) / (n2 * cos_i + n1 * cos_t)) ** 2
    rp_ = ((n2 * cos_t - n1 * cos_i) / (n2 * cos_t + n1 * cos_i)) ** 2
    np.testing.assert_allclose(spec, 0.5 * (rs + rp_), atol=2e-6)


def test_mcxyz_isotropic_rectangle_point_and_manual_sources():
    w = _xyz("iso", mcflag=2, xs=0.01, ys=-0.02, zs=0.1)
    s = O.xyz_source(w, N_SRC).astype(np.float64)
    assert np.all(s[:, 0] == np.float32(0.01)) and np.all(s[:, 1] == np.float32(-0.02))
    assert np.all(s[:, 2] == np.float32(0.1))
    assert _ks_uniform(s[:, 5], -1.0, 1.0) > 1e-3
    assert _ks_uniform(np.arctan2(s[:, 4], s[:, 3]), -math.pi, math.pi) > 1e-3
    assert np.abs((s[:, 3:6] ** 2).sum(1) - 1).max() < 1e-6

    w = _xyz("rect", mcflag=3, radius=0.08, zs=0.0201)
    s = O.xyz_source(w, N_SRC).astype(np.float64)
    assert _ks_uniform(s[:, 0], -0.08, 0.08) > 1e-3 and _ks_uniform(s[:, 1], -0.08, 0.08) > 1e-3
    assert abs(np.corrcoef(s[:, 0], s[:, 1])[0, 1]) < 0.02
    assert np.all(s[:, 3:6] == [0.0, 0.0, 1.0]) and np.all(s[:, 2] == np.float32(0.0201))

    w = _xyz("point", mcflag=4, zs=0.0301)
    s = O.xyz_source(w, 1000)
    assert np.all(s[:, :6] == np.array([0, 0, 0.0301, 0, 0, 1], np.float32))

    w = _xyz("manual", launchflag=1, xs=0.01, ys=0.0, zs=0.05, ux0=0.6, uy0=0.0, uz0=0.8)
    s = O.xyz_source(w, 1000)
    assert np.all(s[:, :6] == np.array([0.01, 0, 0.05, 0.6, 0, 0.8], np.float32))


def _sub(**kw):
    w = W.mcsub_semi_infinite()
    for k, v in kw.items():
        setattr(w.rp, k, v)
    return w


def test_mcsub_sources(mutants):
    """mcsub's launches (kernels.metal:106-160): the beam starts on the x axis
    at the surface (cylindrical symmetry), radius R sqrt(xi) (flat) or
    R sqrt(-ln xi) (Gaussian, refracted towards a Gaussian focus); the point
    source is isotropic."""
    w = _sub(mcflag=0, radius=0.02)
    s = O.sub_source(w, N_SRC).astype(np.float64)
    assert _ks_uniform((s[:, 0] / 0.02) ** 2) > 1e-3
    assert np.all(s[:, 1:3] == 0.0) and np.all(s[:, 3:6] == [0.0, 0.0, 1.0])
    np.testing.assert_allclose(s[:, 6], ((1.33 - 1.0) / 2.33) ** 2, rtol=1e-6)
    bad = O.sub_source(w, N_SRC, L=mutants(1)).astype(np.float64)
    assert _ks_uniform((bad[:, 0] / 0.02) ** 2) < 1e-12, "the mcsub flat-beam pin misses r = R*xi"

    n1, n2, R, wst, zf = 1.33, 1.0, 0.02, 0.01, 0.5
    w = _sub(mcflag=1, radius=R, waist=wst, zfocus=zf)
    s = O.sub_source(w, N_SRC).astype(np.float64)
    x, ux, uz, rsp = s[:, 0], s[:, 3], s[:, 5], s[:, 6]
    assert _ks_exp((x / R) ** 2) > 1e-3
    # the refracted ray's ux is taken positive (kernels.metal:144), so the
    # focus offset is known up to its sign: |x - xf| = zf tan(theta_i) with
    # n2 sin(theta_i) = n1 sin(theta_t); its law follows from x and xf
    # independent with (x/R)^2, (xf/w)^2 ~ Exp(1)
    sin_i = np.sqrt(1 - uz ** 2) * n1 / n2
    d = zf * sin_i / np.sqrt(1 - sin_i ** 2)
    rng = np.random.default_rng(11)
    ref = np.abs(R * np.sqrt(rng.exponential(size=400_000)) - wst * np.sqrt(rng.exponential(size=400_000)))
    # a float cosine near 1 resolves angles down to ~2.4e-4 only (uz rounds
    # to 1 below that), so the law is compared above an offset of 1e-3 and
    # the mass below it as one share
    cut = 1e-3
    assert stats.ks_2samp(d[d > cut], ref[ref > cut]).pvalue > 1e-3
    p_lo, q_lo = (d <= cut).mean(), (ref <= cut).mean()
    assert abs(p_lo - q_lo) < 5 * math.sqrt(q_lo * (1 - q_lo) * (1 / d.size + 1 / ref.size))
    assert np.all(ux >= 0.0)
    cos_i = np.sqrt(1 - sin_i ** 2)
    rs = ((n2 * cos_i - n1 * uz) / (n2 * cos_i + n1 * uz)) ** 2
    rp_ = ((n2 * uz - n1 * cos_i) / (n2 * uz + n1 * cos_i)) ** 2
    np.testing.assert_allclose(rsp, 0.5 * (rs + rp_), atol=2e-6)

    w = _sub(mcflag=2, xs=0.0, ys=0.0, zs=0.05)
    s = O.sub_source(w, N_SRC).astype(np.float64)
    assert np.all(s[:, 2] == np.float32(0.05)) and np.all(s[:, 6] == 0.0)
    assert _ks_uniform(s[:, 5], -1.0, 1.0) > 1e-3


# ---------------------------------------------- step-length carry-over ---
def _two_tissue_column(mus1, l1, mus2, l2, n=4, d=0.005):
    nz1, nz2 = round(l1 / d), round(l2 / d)
    nz = nz1 + nz2
    rp = W._rp("carry", Nx=n, Ny=n, Nz=nz, dx=d, dy=d, dz=d, Nt=2, mcflag=4, boundaryflag=1,
               zs=0.0, zsurf=-1.0, mckernelflag=P.SIM_MCXYZ)
    tp = W._tp({1: (0.0, mus1, 1.0), 2: (0.0, mus2, 1.0)})
    v = np.ones((nz, n, n), np.uint8)
    v[nz1:] = 2
    return W.Workload("carry", P.SIM_MCXYZ, rp, tp, v.reshape(-1), P.ExtConfig.default(), 0)


CARRY = [(100.0, 0.05, 20.0, 0.1), (20.0, 0.05, 100.0, 0.1)]


@pytest.mark.parametrize("mus1,l1,mus2,l2", CARRY)
def test_mcxyz_step_carry_over_poisson_mean(mutants, mus1, l1, mus2, l2):
    w = _two_tissue_column(mus1, l1, mus2, l2)
    n = 200_000
    mean = mus1 * l1 + mus2 * l2
    tol = 5 * math.sqrt(mean / n)

    def avg(L=None):
        t = O.run(w, n, threads=8, fluence=False, L=L).totals
        assert t["escapes_bottom"] == n
        return t["interactions"] / n

    assert abs(avg() - mean) < tol, (avg(), mean)
    assert abs(avg(mutants(2)) - mean) > 10 * tol, "the carry-over pin misses the unscaled step"


@pytest.mark.parametrize("mus1,l1,mus2,l2", CARRY)
def test_layered_step_carry_over_poisson_mean(mutants, mus1, l1, mus2, l2):
    cfg = P.LayeredConfig.default(nr=10, dr=1.0, records=0)
    n = 200_000
    st = W.make_stack([l1, l2], [0.0, 0.0], [mus1, mus2], [1.0, 1.0], n, g=1.0, n_below=1.0)
    mean = mus1 * l1 + mus2 * l2
    tol = 5 * math.sqrt(mean / n)

    def avg(L=None):
        t = O.run_layered(cfg, [st], threads=8, records=False, L=L).totals[0]
        assert t["escapes_bottom"] == n
        return t["interactions"] / n

    assert abs(avg() - mean) < tol, (avg(), mean)
    assert abs(avg(mutants(2)) - mean) > 10 * tol, "the carry-over pin misses the unscaled step"


# ------------------------------------------------------ mcsub depth bins ---
K_BATCH = 16


def _batches(fn, n, k=K_BATCH):
    """Per-batch results of fn(first, count) over k disjoint photon ranges."""
    m = n // k
    return np.array([fn(b * m, m) for b in range(k)])


def _zscores(a, b):
    """z of the difference of two batch-mean estimates (rows = batches)."""
    ma, mb = a.mean(0), b.mean(0)
    va, vb = a.var(0, ddof=1) / len(a), b.var(0, ddof=1) / len(b)
    return (ma - mb) / np.sqrt(va + vb)


def _sub_depth(w, L=None):
    nz_slabs = 7

    def one(first, m):
        t = O.run(w, m, first=first, threads=8, L=L)
        a = t.a_rz.reshape(w.ext.nr + 1, w.ext.nz + 1).sum(0) / m
        return np.concatenate([a[:nz_slabs], [a[nz_slabs:].sum()]])
    return one


def test_mcsub_depth_bins_against_layered_slabs(mutants):
    mua, mus, g, n_med, dz = 1.0, 100.0, 0.9, 1.33, 0.01
    w = W.mcsub_semi_infinite()
    w.ext.dz = dz
    n = 160_000
    sub = _batches(_sub_depth(w), n)
    cfg = P.LayeredConfig.default(nr=10, dr=1.0, records=0)

    def lay(first, m):
        st = W.make_stack([dz] * 7 + [0.0], [mua] * 8, [mus] * 8, [n_med] * 8, m, g=g)
        return O.run_layered(cfg, [st], first=first, threads=8, records=False).absorbed[0] / m
    layered = _batches(lay, n)
    z = _zscores(sub, layered)
    assert np.abs(z).max() < 4.5, z
    for k in (3, 4):
        zb = _zscores(_batches(_sub_depth(w, mutants(k)), n), layered)
        assert np.abs(zb).max() > 10, f"the depth-bin pin misses: {O.MUTANTS[k]}"


# ------------------------------------------------- mcsub exit-angle bins ---
def _h_function(a):
    """Chandrasekhar's H(mu) for isotropic scattering with albedo a, by
    iterating 1/H(mu) = sqrt(1-a) + a/2 int_0^1 mu' H(mu') / (mu + mu') dmu'
    on Gauss-Legendre nodes; returns a callable for any mu."""
    x, wq = np.polynomial.legendre.leggauss(200)
    mu, wq = 0.5 * (x + 1), 0.5 * wq
    H = np.ones_like(mu)
    for _ in range(1000):
        Hn = 1.0 / (math.sqrt(1 - a) + a / 2 * ((wq * mu * H)[None, :] / (mu[:, None] + mu[None, :])).sum(1))
        if np.abs(Hn - H).max() < 1e-15:
            break
        H = Hn
    return lambda m: 1.0 / (math.sqrt(1 - a) + a / 2 * ((wq * mu * H)[None, :]
                                                      / (np.atleast_1d(m)[:, None] + mu[None, :])).sum(1))


def angular_shares(a, na):
    """Expected share of the escaping flux per exit-angle bin of width 90/na
    degrees: integral of mu H(mu) / (1 + mu) over the bin's cosines."""
    H = _h_function(a)
    x, wq = np.polynomial.legendre.leggauss(40)
    edges = np.cos(np.arange(na + 1) * (math.pi / 2) / na)
    shares = []
    for j in range(na):
        hi, lo = edges[j], edges[j + 1]
        m = lo + (hi - lo) * 0.5 * (x + 1)
        shares.append(((hi - lo) * 0.5 * wq * m * H(m) / (1 + m)).sum())
    shares = np.array(shares)
    # the same law gives the total reflectance R = a/2 H(1) int mu H/(1+mu)
    # = 1 - H(1) sqrt(1-a): a check on the H-function and the law together
    H1 = float(H(1.0)[0])
    R = a / 2 * H1 * shares.sum()
    assert R == pytest.approx(1 - H1 * math.sqrt(1 - a), rel=1e-6)
    return R * shares / shares.sum()


def test_mcsub_exit_angle_bins_follow_chandrasekhar(mutants):
    a, na = 0.9, 30
    expect = angular_shares(a, na)
    w = W.mcsub_semi_infinite()
    w.rp.n1 = w.rp.n2 = 1.0
    w.tp.muav[1], w.tp.musv[1], w.tp.gv[1] = 1.0 - a, a, 0.0
    w.ext.na, w.ext.nr, w.ext.dr = na, 10, 1.0
    n = 320_000

    def rd_alpha(L=None):
        def one(first, m):
            t = O.run(w, m, first=first, threads=8, fluence=False, L=L)
            return t.rd_ra.reshape(w.ext.nr + 1, na).sum(0) / m
        return _batches(one, n)

    got = rd_alpha()
    z = (got.mean(0) - expect) / (got.std(0, ddof=1) / math.sqrt(len(got)))
    assert np.abs(z).max() < 4.5, z
    bad = rd_alpha(mutants(4))
    zb = (bad.mean(0) - expect) / (bad.std(0, ddof=1) / math.sqrt(len(bad)) + 1e-12)
    assert np.abs(zb).max() > 10, "the exit-angle pin misses 1-based bins"
