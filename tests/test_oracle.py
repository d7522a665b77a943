"""Pins for the CPU oracle (CPU only).

Every routine of the oracle is checked here against something it shares no
code with: published Philox known-answer vectors and PyTorch's Philox engine,
numpy's float64 libm, the textbook Fresnel formulas, closed-form moments of the
phase functions, Chandrasekhar's H-function for the semi-infinite isotropic
medium, and Beer-Lambert attenuation for unscattered light.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os

import numpy as np
import pytest

from nmc_amd import params as P
from nmc_amd import workloads as W
from oracle import oracle as O

from parity_util import energy_balance

HERE = os.path.dirname(os.path.abspath(__file__))


def ulp_err(y, ref):
    sp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    return np.abs(y.astype(np.float64) - ref) / sp


# ---------------------------------------------------------------- Philox ---
def test_philox_random123_kat():
    # Known-answer vectors published with Random123 (kat_vectors, philox4x32_10).
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6,
                                                            0x6D5451FD]
    assert O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) \
        == [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_philox_matches_aten_engine():
    with open(os.path.join(HERE, "golden", "philox_vectors.json")) as f:
        d = json.load(f)
    assert len(d["vectors"]) > 50
    for v in d["vectors"]:
        assert O.philox(v["ctr"], v["key"]) == v["out"], v


def test_spare_low_bytes_give_an_independent_uniform():
    """mcsub's escape test uses a fifth uniform from the low bytes of words
    0-2 of the event's block: uniform on 64 bins and independent of the
    step-length draw (top 24 bits of word 0), by chi-square."""
    from scipy import stats
    n = 40000
    low = np.empty(n)
    top = np.empty(n)
    for i in range(n):
        b = O.philox([i * 7919 & 0xFFFFFFFF, 0, 1 + (i % 13), 0], [1, 0])
        low[i] = (((b[0] & 0xFF) << 16) | ((b[1] & 0xFF) << 8) | (b[2] & 0xFF)) * 2.0 ** -24
        top[i] = (b[0] >> 8) * 2.0 ** -24
    counts = np.histogram(low, bins=64, range=(0, 1))[0]
    assert stats.chisquare(counts).pvalue > 1e-4
    joint = np.histogram2d(low, top, bins=8, range=[[0, 1], [0, 1]])[0]
    assert stats.chi2_contingency(joint)[1] > 1e-4


# --------------------------------------------------------- transcendentals ---
def test_logf_all_uniform_draws():
    """-log(u) is taken of every u = k 2^-24, k = 1..2^24: check them all."""
    u = (np.arange(1, 2 ** 24 + 1, dtype=np.float64) * 2.0 ** -24).astype(np.float32)
    y = O.logf(u)
    assert y[-1] == 0.0
    assert ulp_err(y[:-1], np.log(u[:-1].astype(np.float64))).max() < 1.0
    assert np.all(np.diff(y) >= 0)  # monotone


def test_logf_wide_range():
    x = np.random.default_rng(1).uniform(-30, 30, 200000)
    x = np.exp(x).astype(np.float32)
    assert ulp_err(O.logf(x), np.log(x.astype(np.float64))).max() < 1.0


def test_logf_tab_all_uniform_draws():
    """The voxel walk's table log: every uniform draw k 2^-24 within one ulp of
    log, monotone, log(1) = 0 exactly."""
    u = (np.arange(1, 2 ** 24 + 1, dtype=np.float64) * 2.0 ** -24).astype(np.float32)
    y = O.logf_tab(u)
    assert y[-1] == 0.0
    assert ulp_err(y[:-1], np.log(u[:-1].astype(np.float64))).max() < 1.0
    assert np.all(np.diff(y) >= 0)


def test_logf_tab_wide_range_and_near_one():
    rng = np.random.default_rng(3)
    x = np.exp(rng.uniform(-30, 30, 400000)).astype(np.float32)
    # every float in (1 - 2^-6, 1 + 2^-5): relative precision as log -> 0
    lo, hi = np.float32(1 - 2 ** -6), np.float32(1 + 2 ** -5)
    near = np.arange(lo.view(np.uint32), hi.view(np.uint32), dtype=np.uint32).view(np.float32)
    near = near[near != 1.0]
    for v in (x, near):
        assert ulp_err(O.logf_tab(v), np.log(v.astype(np.float64))).max() < 1.0


def test_log_table_rows_hold_their_properties():
    """Rows of nmc_tables.h as log_tab relies on them: rc with at most 8
    significant bits and |m' rc - 1| < 2^-7 over the row's interval (so the
    reduction is exact), L_hi a multiple of 2^-15, L_hi + L_lo = -log(rc) to
    2^-40, rc = 1 and L = 0 on the six rows next to m' = 1."""
    import re
    text = open(os.path.join(HERE, "..", "nmc_amd", "csrc", "nmc_tables.h")).read()
    rows = re.findall(r"\{(-?0x[0-9a-fp.+-]+)f, (-?0x[0-9a-fp.+-]+)f, (-?0x[0-9a-fp.+-]+)f, 0\.0f\}",
                      text)
    assert len(rows) == 128
    for i, (a, b, c) in enumerate(rows):
        rc, lhi, llo = float.fromhex(a), float.fromhex(b), float.fromhex(c)
        if i < 2 or i >= 124:
            assert (rc, lhi, llo) == (1.0, 0.0, 0.0)
            continue
        m, _ = math.frexp(rc)
        assert math.ldexp(m, 8) == int(math.ldexp(m, 8))  # <= 8 significant bits
        lo = 1 + i / 128 if i < 64 else (1 + i / 128) / 2
        hi = lo + (1 / 128 if i < 64 else 1 / 256)
        assert max(abs(lo * rc - 1), abs(hi * rc - 1)) < 2 ** -7
        assert lhi * 32768 == int(lhi * 32768)
        assert abs(lhi + llo + math.log(rc)) < 2 ** -40
        assert np.float32(llo) == llo


def test_expf_and_expm1f():
    rng = np.random.default_rng(2)
    x = rng.uniform(-87, 0, 1_000_000).astype(np.float32)
    assert ulp_err(O.expf(x), np.exp(x.astype(np.float64))).max() < 1.0
    assert O.expf(np.array([-88.0, -1000.0], np.float32)).tolist() == [0.0, 0.0]
    xs = np.concatenate([rng.uniform(-0.4, 0.4, 500000), -np.logspace(-12, 1.5, 20000),
                         rng.uniform(-60, -0.3, 200000)]).astype(np.float32)
    assert ulp_err(O.expm1f(xs), np.expm1(xs.astype(np.float64))).max() < 4.0
    # expm1 keeps full relative precision where 1 - exp would cancel
    tiny = np.float32(-3e-7)
    assert abs(O.expm1f(np.array([tiny]))[0] / math.expm1(float(tiny)) - 1) < 1e-7


def test_sincos2pi_all_uniform_draws():
    u = (np.arange(0, 2 ** 24, 7, dtype=np.float64) * 2.0 ** -24).astype(np.float32)
    s, c = O.sincos2pi(u)
    a = 2 * np.pi * u.astype(np.float64)
    assert np.abs(s - np.sin(a)).max() < 2e-7
    assert np.abs(c - np.cos(a)).max() < 2e-7
    assert np.abs(s.astype(np.float64) ** 2 + c.astype(np.float64) ** 2 - 1).max() < 5e-7
    s1, c1 = O.sincos2pi(np.array([0.0, 0.25, 0.5, 0.75, 1.0], np.float32))
    np.testing.assert_array_equal(s1, [0, 1, 0, -1, 0])
    np.testing.assert_array_equal(c1, [1, 0, -1, 0, 1])


def test_sincos2pi_bits_every_draw():
    """The walks' azimuth from the random word: every top-24-bit value (the
    low byte does not enter) within 1.2e-7 of sin and cos of 2 pi u, u = (x >>
    8) 2^-24; s^2 + c^2 = 1 to 3e-7; exact at the quarter turns; the table rows
    are sin and cos of 2 pi j/64 rounded to float."""
    k = np.arange(0, 2 ** 24, dtype=np.uint32)
    s, c = O.sincos2pi_bits(k << np.uint32(8) | np.uint32(0x5A))
    a = 2 * np.pi * (k.astype(np.float64) * 2.0 ** -24)
    assert np.abs(s - np.sin(a)).max() < 1.2e-7
    assert np.abs(c - np.cos(a)).max() < 1.2e-7
    assert np.abs(s.astype(np.float64) ** 2 + c.astype(np.float64) ** 2 - 1).max() < 3e-7
    q = np.array([0, 1 << 30, 2 << 30, 3 << 30], np.uint32)
    sq, cq = O.sincos2pi_bits(q)
    np.testing.assert_array_equal(sq, [0, 1, 0, -1])
    np.testing.assert_array_equal(cq, [1, 0, -1, 0])
    # row j is (sin, cos)(2 pi j/64): the draws with f = 0
    j = np.arange(64, dtype=np.uint32)
    sj, cj = O.sincos2pi_bits(j << np.uint32(26))
    assert np.abs(sj - np.sin(2 * np.pi * j / 64)).max() <= 3e-8
    assert np.abs(cj - np.cos(2 * np.pi * j / 64)).max() <= 3e-8


# ------------------------------------------------------------------ optics ---
def fresnel_textbook(n1, n2, ci):
    si = math.sqrt(max(0.0, 1 - ci * ci))
    st = n1 / n2 * si
    if st >= 1:
        return 1.0, 0.0
    ct = math.sqrt(1 - st * st)
    rs = (n1 * ci - n2 * ct) / (n1 * ci + n2 * ct)
    rp = (n2 * ci - n1 * ct) / (n2 * ci + n1 * ct)
    return 0.5 * (rs * rs + rp * rp), ct


@pytest.mark.parametrize("n1,n2", [(1.33, 1.0), (1.0, 1.33), (1.4, 1.0), (1.0, 1.5), (1.5, 1.37)])
def test_fresnel_textbook(n1, n2):
    for ci in np.linspace(0.001, 0.9999, 400):
        ci = float(np.float32(ci))
        r, ca2 = O.fresnel(n1, n2, ci)
        rt, ct = fresnel_textbook(n1, n2, ci)
        # float32 evaluation: allow for the conditioning near the critical angle
        h = 1e-6 * ci
        dr = abs(fresnel_textbook(n1, n2, ci + h)[0] - fresnel_textbook(n1, n2, ci - h)[0])
        dc = abs(fresnel_textbook(n1, n2, ci + h)[1] - fresnel_textbook(n1, n2, ci - h)[1])
        assert abs(r - rt) < 2e-6 + 2 * dr, (ci, r, rt)
        assert abs(ca2 - ct) < 2e-6 + 2 * dc


def test_fresnel_special_cases():
    assert O.fresnel(1.33, 1.33, 0.3) == (0.0, pytest.approx(0.3))
    r, c = O.fresnel(1.33, 1.0, 1.0)
    assert r == pytest.approx(((1.33 - 1) / 2.33) ** 2, rel=1e-6) and c == 1.0
    assert O.fresnel(1.4, 1.0, 1e-7) == (1.0, 0.0)       # grazing
    assert O.fresnel(1.4, 1.0, 0.5)[0] == 1.0            # beyond the critical angle


# -------------------------------------------------------------- scattering ---
def _grid(n=200000):
    return ((np.arange(n) + 0.5) / n).astype(np.float32)


@pytest.mark.parametrize("g", [0.0, 0.5, 0.9, 0.98, -0.3])
def test_hg_first_moment(g):
    L = O.lib()
    u = _grid()
    ct = np.array([L.nmco_hg_cos(g, float(x)) for x in u[::4]])
    assert abs(ct.mean() - g) < 2e-4
    assert ct.min() >= -1 and ct.max() <= 1


def test_hg_cdf_matches_closed_form():
    # P(cos <= t) for HG: (1-g^2)/(2g) [ (1+g^2-2gt)^(-1/2) - (1+g)^(-1) ]
    g = 0.9
    L = O.lib()
    u = _grid(50000)
    ct = np.sort([L.nmco_hg_cos(g, float(x)) for x in u])
    for t in (-0.5, 0.0, 0.5, 0.9, 0.99):
        cdf = (1 - g * g) / (2 * g) * ((1 + g * g - 2 * g * t) ** -0.5 - 1 / (1 + g))
        assert abs(np.searchsorted(ct, t) / ct.size - cdf) < 1e-3


def test_two_term_moment():
    # two HG terms with a = 1/2: mean cos = CC g1 + (1 - CC)(-g2)
    L = O.lib()
    u = _grid()[::4]
    ct = np.array([L.nmco_tt_cos(0.9, 0.5, 0.3, 0.5, 0.8, float(x)) for x in u])
    assert abs(ct.mean() - (0.8 * 0.9 - 0.2 * 0.3)) < 3e-4


def test_two_term_general_alpha_moment():
    # Gegenbauer kernel (af != 1/2): compare with numerical quadrature of its pdf
    gf, af = 0.7, 0.8
    L = O.lib()
    u = _grid()[::4]
    ct = np.array([L.nmco_tt_cos(gf, af, 0.0, 0.5, 1.0, float(x)) for x in u])
    mu = np.linspace(-1, 1, 200001)
    pdf = (1 + gf * gf - 2 * gf * mu) ** (-(af + 1))
    assert abs(ct.mean() - np.trapezoid(mu * pdf, mu) / np.trapezoid(pdf, mu)) < 5e-4


def test_spin_geometry():
    L = O.lib()
    rng = np.random.default_rng(3)
    for _ in range(2000):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        if rng.random() < 0.05:
            d = np.array([0.0, 0.0, rng.choice([-1.0, 1.0])])
        ux, uy, uz = (C.c_float(v) for v in d)
        cost = float(rng.uniform(-1, 1))
        L.nmco_spin(cost, float(rng.random()), C.byref(ux), C.byref(uy), C.byref(uz))
        nd = np.array([ux.value, uy.value, uz.value])
        assert abs(np.linalg.norm(nd) - 1) < 2e-6
        assert abs(nd @ d - cost) < 2e-6


# ------------------------------------------------------- whole-path limits ---
TINY = ["mismatch", "matched", "gaussian", "iso", "rect", "point", "manual", "infinite",
        "surface_only", "transmission", "two_term"]


@pytest.mark.parametrize("kind", TINY)
def test_energy_balance(kind):
    t = O.run(W.tiny_voxel_case(kind, n_photons=3000), threads=4).totals
    assert t["photons"] == 3000
    assert abs(energy_balance(t) - 1.0) < 1e-5


def test_energy_balance_mcsub_and_specular():
    w = W.mcsub_semi_infinite()
    t = O.run(w, 20000, threads=4).totals
    rsp = ((1.33 - 1.0) / 2.33) ** 2
    assert t["specular"] / t["photons"] == pytest.approx(rsp, rel=1e-6)
    assert abs(energy_balance(t) - 1.0) < 1e-5


def test_beer_lambert_unscattered():
    """g = 1: light is never deflected, so every photon leaves the bottom with
    weight exp(-mua L) and voxel k absorbs exp(-mua z_k) - exp(-mua z_k+1)."""
    n, d, mua = 10, 0.05, 3.0
    rp = W._rp("bl", Nx=n, Ny=n, Nz=n, dx=d, dy=d, dz=d, Nt=1, mcflag=4, boundaryflag=1,
               zs=0.0, zsurf=-1.0, mckernelflag=P.SIM_MCXYZ)
    tp = W._tp({1: (mua, 40.0, 1.0)})
    w = W.Workload("bl", P.SIM_MCXYZ, rp, tp, np.ones(n ** 3, np.uint8),
                   P.ExtConfig.default(), 500)
    t = O.run(w, threads=2)
    L = n * d
    assert t.totals["escapes_bottom"] == 500
    assert t.totals["tt"] / 500 == pytest.approx(math.exp(-mua * L), rel=2e-5)
    col = t.fluence.reshape(n, n, n)[:, n // 2, n // 2] / 500
    z = np.arange(n + 1) * d
    np.testing.assert_allclose(col, np.exp(-mua * z[:-1]) - np.exp(-mua * z[1:]), rtol=2e-5)


def chandrasekhar_R(a: float) -> float:
    """Reflectance of a semi-infinite isotropically scattering medium (albedo a,
    matched boundary) for normal incidence: R = 1 - H(1) sqrt(1 - a), with
    Chandrasekhar's H-function from 1/H(mu) = sqrt(1-a) + a/2 int mu' H(mu')/(mu+mu')."""
    x, wq = np.polynomial.legendre.leggauss(200)
    mu, wq = 0.5 * (x + 1), 0.5 * wq
    H = np.ones_like(mu)
    for _ in range(500):
        Hn = 1.0 / (math.sqrt(1 - a) + a / 2 * (mu[:, None] * 0 + (wq * mu * H)[None, :]
                                                  / (mu[:, None] + mu[None, :])).sum(1))
        if np.abs(Hn - H).max() < 1e-14:
            break
        H = Hn
    H1 = 1.0 / (math.sqrt(1 - a) + a / 2 * (wq * mu * H / (1.0 + mu)).sum())
    return 1.0 - H1 * math.sqrt(1 - a)


@pytest.mark.parametrize("albedo", [0.9, 0.99])
def test_semi_infinite_isotropic_reflectance(albedo):
    w = W.mcsub_semi_infinite()
    w.rp.n1 = w.rp.n2 = 1.0
    w.tp.muav[1], w.tp.musv[1], w.tp.gv[1] = 1.0 - albedo, albedo, 0.0
    n = 400_000
    t = O.run(w, n, threads=8)
    # per-photon escaping weight lies in [0,1]: sigma of the mean <= 0.5/sqrt(n)
    sig = float(np.sqrt(t.rd_ra.sum() / n * (1 - t.rd_ra.sum() / n) / n))
    R = chandrasekhar_R(albedo)
    assert abs(t.totals["rd"] / n - R) < 5 * sig + 1e-4, (t.totals["rd"] / n, R, sig)


def test_mirror_symmetry_of_detected_light():
    """The skin-vessel geometry and beam are symmetric under x -> -x and y -> -y:
    detected weight in mirrored halves agrees within the statistical error."""
    w = W.skin_vessel()
    n = 40000
    t = O.run(w, n, threads=8, fluence=False)
    det = t.det2d.reshape(200, 200).astype(np.float64)
    for a, b in ((det[:100].sum(), det[100:].sum()), (det[:, :100].sum(), det[:, 100:].sum())):
        assert abs(a - b) < 4 * math.sqrt(a + b)
