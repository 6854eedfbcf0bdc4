"""Pins the CPU oracle before anything is compared against it.

Every check here is chosen so that a plausible slip in the oracle (a dropped
term, a wrong sign or index, a transposed operand) fails: the reference's own
accuracy-test output, the reference's bitwise mirror-symmetry checks,
closed forms (exactness on polynomials, L.R = I, consistency F(q,q) = F(q))
and conservation / free-stream invariants.
"""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
WHF = [O.S_WHF_N, O.S_WHF_A, O.S_WHF_AS, O.S_WHF_S, O.S_WHF_LAD, O.S_AI_S, O.S_AI_A, O.S_AI_AS, O.S_AI_LADS,
       O.S_AI_AZS, O.S_AI_F002, O.S_WHF_COMPARE, O.S_CONGAA]
SCHEMES = [O.S_JS, O.S_Z, O.S_TENO5, O.S_TENO5_CT4, O.S_TENO5_CT7, O.S_TENOS, O.S_LINEAR,
           O.S_TENOS_CT4, O.S_TENOS_CT7, O.S_TENO5_CT10, O.S_TENOA, O.S_TENOSA_F002, O.S_TENOSA_F203,
           O.S_TENOSA_F202] + WHF


# ---------------------------------------------------------------- accuracy
def _g(x, k):
    return math.pow(x, k) * math.exp(0.75 * (x - 1))


def _dg(x, k):
    return math.exp(0.75 * (-1 + x)) * k * math.pow(x, -1 + k) + 0.75 * math.exp(0.75 * (-1 + x)) * math.pow(x, k)


def _accuracy_errors(h, k, scheme=O.S_TENO5):
    """CongFD1.0/AccuracyTest.cpp:14-83 (getError) driven through the oracle."""
    n = int((1.0 - (-1.0)) / h)
    gl = [_g(-1.0 - h / 2 - i * h, k) for i in range(10)]
    gr = [_g(1.0 + h / 2 + i * h, k) for i in range(10)]
    cells = [_g(-1.0 + h / 2 + i * h, k) for i in range(n)]
    hh = (1.0 - (-1.0)) / ((n + 1) - 1)
    r = O.accuracy_rhs(scheme, cells, gl, gr, hh)
    e = [abs(r[i] - _dg(-1.0 + h / 2 + i * h, k)) for i in range(n)]
    return sum(e) / n, math.sqrt(sum(x * x for x in e) / n), max(e)


def test_accuracy_test_matches_reference_output():
    rows = [l.split() for l in open(os.path.join(GOLD, "accuracyTest.dat")).read().splitlines()[2:]]
    checked = 0
    for row in rows:
        h = float(row[0])
        if int(round(2.0 / h)) > 60:
            continue  # rows n >= 70 are not reproducible (see tests/golden/README.md)
        got = _accuracy_errors(h, 2.0)
        for g, ref in zip(got, row[1:4]):
            assert abs(g / float(ref) - 1) < 2e-7, (h, got, row)
        checked += 1
    assert checked == 6


def test_accuracy_order_fifth():
    hs = [2.0 / n for n in (10, 20, 30, 40)]
    errs = [_accuracy_errors(h, 2.0)[0] for h in hs]
    orders = [math.log(errs[i] / errs[i + 1]) / math.log(hs[i] / hs[i + 1]) for i in range(3)]
    ref = [float(l.split()[1]) for l in open(os.path.join(GOLD, "accuracy_order.txt")).read().splitlines()[1:4]]
    np.testing.assert_allclose(orders, ref, rtol=1e-6)


def test_accuracy_test_detects_planted_slip():
    # the same pipeline with WENO-JS instead of TENO5 must not match the file
    row = open(os.path.join(GOLD, "accuracyTest.dat")).read().splitlines()[2].split()
    got = _accuracy_errors(float(row[0]), 2.0, scheme=O.S_JS)
    assert abs(got[0] / float(row[1]) - 1) > 1e-3


# ------------------------------------------------------------ interpolation
@pytest.mark.parametrize("scheme", SCHEMES)
def test_inter5_exact_for_quadratics(scheme):
    rng = np.random.default_rng(1)
    for _ in range(20):
        a, b, c = rng.normal(size=3)
        x = np.arange(5.0)
        q = a + b * x + c * x * x
        xe = 2.5
        want = a + b * xe + c * xe * xe
        assert abs(O.inter5(scheme, q) - want) < 1e-12 * (1 + abs(want) + abs(c) * 10)


@pytest.mark.parametrize("scheme", [O.S_LINEAR, O.S_TENO5, O.S_TENO5_CT10, O.S_TENOA, O.S_TENOSA_F002,
                                    O.S_TENOSA_F203, O.S_TENOSA_F202] + WHF)
def test_inter5_fifth_order_exact_for_quartics(scheme):
    rng = np.random.default_rng(2)
    for _ in range(20):
        co = rng.normal(size=5)
        # resolved data (spacing 0.05): TENO5 keeps every stencil and reduces
        # to the linear 5th-order interpolant, exact for quartics
        x = 0.3 + 0.05 * np.arange(5.0)
        q = np.polyval(co, x)
        want = np.polyval(co, 0.3 + 0.05 * 2.5)
        assert abs(O.inter5(scheme, q) - want) < 1e-13 * (1 + np.abs(co).sum())


def test_teno_step_selects_smooth_candidate():
    # step (0,0,0,1,1): stencils 1 and 2 cross the jump -> flag 6 -> u0 = 0
    assert O.inter5(O.S_TENO5, [0, 0, 0, 1, 1]) == 0.0
    # mirrored step (1,1,0,0,0) seen from the other side -> u2 of (0,0,0) = 0
    assert O.inter5(O.S_TENO5, [1, 1, 0, 0, 0]) == 0.0
    # all three stencils smooth -> linear 5th-order weights
    q = [1.0, 2.0, 3.0, 4.0, 5.0]
    assert O.inter5(O.S_TENO5, q) == O.inter5(O.S_LINEAR, q) == 3.5


# the eight stencils of the TENO family by flag (interScheme.hpp:421-449):
# 0 five-point linear, 1 and 4 four-point, 2/3 u2, 5 u1, 6 u0, 7 q2
def _stencils(q):
    q0, q1, q2, q3, q4 = q
    u0 = 3 / 8 * q0 - 5 / 4 * q1 + 15 / 8 * q2
    u1 = -1 / 8 * q1 + 3 / 4 * q2 + 3 / 8 * q3
    u2 = 3 / 8 * q2 + 3 / 4 * q3 - 1 / 8 * q4
    return [3 / 128 * q0 - 5 / 32 * q1 + 45 / 64 * q2 + 15 / 32 * q3 - 5 / 128 * q4,
            -1 / 16 * q1 + 9 / 16 * q2 + 9 / 16 * q3 - 1 / 16 * q4, u2, u2,
            1 / 16 * q0 - 5 / 16 * q1 + 15 / 16 * q2 + 5 / 16 * q3, u1, u0, q2]


def _cut(scheme, q):
    """Set of stencils a TENO-family scheme cut at q (bits of its flag), read
    off from which of the eight stencil values it returned."""
    v = O.inter5(scheme, q)
    st = _stencils(q)
    flags = [0, 1, 2, 4, 5, 6, 7]  # flag 3 returns flag 2's stencil: not told apart
    vals = [st[f] for f in flags]
    hits = [f for f, x in zip(flags, vals) if abs(x - v) <= 1e-12 * (1 + abs(v))]
    if len({round(x, 9) for x in vals}) < len(vals) or len(hits) != 1 or hits[0] == 2:
        return None  # ambiguous
    f = hits[0]
    return {0: set(), 1: {0}, 2: {1}, 3: {0, 1}, 4: {2}, 5: {0, 2}, 6: {1, 2}, 7: {0, 1, 2}}[f]


def test_teno_cutoff_constants_order_the_cut_stencils():
    """The cut-off constant orders the stencils cut, for every data: TENO5
    cuts stencil k when g_k < CT sum(g), TENO-S when ll < (CT tau - (1-CT)
    beta_min) beta_k -- both monotone in CT.  So CT 1e-10 cuts a subset of
    CT 1e-7, of CT 1e-5, of CT 1e-4 (TENO5), and TENO-S CT7 (0.0729) a subset
    of CT5 (0.157), of CT4 (0.2305).  A variant wired to the wrong constant
    breaks the chain."""
    rng = np.random.default_rng(12)
    chains = [[O.S_TENO5_CT10, O.S_TENO5_CT7, O.S_TENO5, O.S_TENO5_CT4],
              [O.S_TENOS_CT7, O.S_TENOS, O.S_TENOS_CT4]]
    strict = [0, 0]
    for _ in range(3000):
        k = rng.uniform(0.2, 3.0)
        q = np.sin(k * np.arange(5.0) + rng.uniform(0, 6.3)) + rng.normal(scale=rng.choice([0, 1e-3, 0.3]), size=5)
        for c, chain in enumerate(chains):
            sets = [_cut(s, q) for s in chain]
            if any(x is None for x in sets):
                continue
            for a, b in zip(sets, sets[1:]):
                assert a <= b, (chain, q, sets)
            strict[c] += sets[0] != sets[-1]
    assert min(strict) > 50  # the constants do separate the variants


def test_teno_a_adapts_its_constant():
    """TENO-A (new_whf_TCNS_A) is TENO5 with CT = 10^-beta_A: on a step the
    difference ratios are ~eps_A, so beta_A = 5 (CT 1e-5, TENO5); on resolved
    oscillations they are ~1, so beta_A = 10 (CT 1e-10)."""
    step = [0.0, 0.0, 0.0, 1.0, 1.0]
    assert O.inter5(O.S_TENOA, step) == O.inter5(O.S_TENO5, step) == 0.0
    q = np.sin(1.7488 * np.arange(5.0) + 0.7692)  # where CT 1e-5 and 1e-10 cut differently
    assert O.inter5(O.S_TENO5, q) != O.inter5(O.S_TENO5_CT10, q)
    assert O.inter5(O.S_TENOA, q) == O.inter5(O.S_TENO5_CT10, q)
    # and its output is always one of the TENO5 family's
    rng = np.random.default_rng(3)
    for _ in range(500):
        q = rng.normal(size=5)
        assert O.inter5(O.S_TENOA, q) in {O.inter5(s, q) for s in (O.S_TENO5, O.S_TENO5_CT7, O.S_TENO5_CT10)} or \
            _cut(O.S_TENOA, q) is not None


def _eta_min(q):
    """min_k of the smoothness ratios (|2 a b| + eps_A) / (a^2 + b^2 + eps_A)
    of adjacent differences a, b of the five values."""
    d = [q[k] - q[k + 1] for k in range(4)]
    eps = 2.7551e-7
    return min((abs(2 * d[k] * d[k + 1]) + eps) / (d[k] ** 2 + d[k + 1] ** 2 + eps) for k in range(3))


def test_adaptive_teno_s_laws_at_their_ends():
    """The TENO-S variants whose C_T follows eta_min, at its two ends.
    Resolved data (eta_min >= 0.24): the F203 / F202 laws take m = 0, i.e.
    C_T = 0.022010 and -0.000616 (the cubic's and quadratic's constant
    terms), the zyc law s = 1, C_T = 0.157 / 7.117073 = 0.02206 -- all below
    TENO-S CT7's 0.0729, so they cut a subset of what CT7 cuts, and F202
    (C_T < 0) cuts nothing.  Next to a jump (eta_min < 0.01): m > 0.958 gives
    C_T in [0.148, 0.161] (F203) and [0.165, 0.171] (F202), s < 0.042 gives
    C_T in [0.104, 0.122] (zyc): between CT7 and CT4 (0.2305), zyc below
    TENO-S (0.157) and F202 above it.  TENO-S cuts more as C_T grows, so a
    law wired with a wrong coefficient or sign breaks a chain."""
    rng = np.random.default_rng(21)
    seen = {"smooth": 0, "jump": 0}
    strict = 0
    for i in range(9000):
        if i % 3 == 1:
            q = rng.uniform(-1, 1) * np.sin(rng.uniform(0.05, 0.9) * np.arange(5.0) + rng.uniform(0, 6.3))
            q = q + rng.normal(scale=rng.choice([0, 1e-4, 1e-2]), size=5)
        elif i % 3 == 2:
            # kinks: adjacent differences within a factor e^2 (eta >= 0.24) of any sign
            d = np.exp(rng.uniform(-1, 1, size=4)) * rng.choice([-1.0, 1.0], size=4)
            q = np.concatenate([[0.0], np.cumsum(d)])
        else:
            j = rng.integers(1, 5)
            q = np.where(np.arange(5) >= j, rng.uniform(0.2, 2), 0.0) + rng.normal(
                scale=rng.choice([1e-9, 1e-5, 1e-3, 1e-2]), size=5)
        e = _eta_min(q)
        cuts = {s: _cut(s, q) for s in (O.S_TENOS_CT7, O.S_TENOS, O.S_TENOS_CT4, O.S_TENOSA_F002,
                                         O.S_TENOSA_F203, O.S_TENOSA_F202)}
        if any(c is None for c in cuts.values()):
            continue
        if e >= 0.24:
            seen["smooth"] += 1
            for s in (O.S_TENOSA_F002, O.S_TENOSA_F203, O.S_TENOSA_F202):
                assert cuts[s] <= cuts[O.S_TENOS_CT7], (s, q)
            assert cuts[O.S_TENOSA_F202] == set(), q
        elif e < 0.01:
            seen["jump"] += 1
            for s in (O.S_TENOSA_F002, O.S_TENOSA_F203, O.S_TENOSA_F202):
                assert cuts[O.S_TENOS_CT7] <= cuts[s] <= cuts[O.S_TENOS_CT4], (s, q, cuts)
            assert cuts[O.S_TENOSA_F002] <= cuts[O.S_TENOS] <= cuts[O.S_TENOSA_F202], (q, cuts)
            strict += cuts[O.S_TENOSA_F002] != cuts[O.S_TENOSA_F202]
    assert min(seen.values()) > 300 and strict > 20, (seen, strict)
    # F202 on resolved data never cuts: the linear five-point value, exactly
    q = np.sin(0.3 * np.arange(5.0) + 0.4)
    assert _eta_min(q) >= 0.24 and O.inter5(O.S_TENOSA_F202, q) == O.inter5(O.S_LINEAR, q)


def test_research_ct_prime_against_the_reference_constants():
    """C_T' = (1.5 C_T / (1 - C_T))^(1/6), the constant the whf_*_S / _AS
    schemes derive from TENO5's C_T, against the literal constants of
    Teno5_CongZ / CT4 / CT7 (interScheme.hpp:494, :586, :661), which are
    (1.5 C_T)^(1/6): the two differ by the factor (1 - C_T)^(-1/6) only.
    The oracle's C_T' is read off where it decides a cut: ai_S cuts exactly
    what TENO-S with that constant cuts."""
    for ct, lit in ((1e-5, 0.15704178024750198), (1e-4, 0.23050581003334941), (1e-7, 0.072892337360747711)):
        ctp = math.pow(1.5 * ct / (1.0 - ct), 1.0 / 6)
        assert abs(ctp * (1.0 - ct) ** (1.0 / 6) - lit) <= 4e-16 * lit
    # ai_S_1 (C_T' of 1e-5) and TENO-S (0.157041780...) differ in C_T by
    # 1.7e-6 relative: the same cut except where a test ties that closely
    rng = np.random.default_rng(5)
    same = sum(O.inter5(O.S_AI_S, q) == O.inter5(O.S_TENOS, q) for q in rng.normal(size=(2000, 5)))
    assert same >= 1995


def _jump(rng):
    j = rng.integers(1, 5)
    return np.where(np.arange(5) >= j, rng.uniform(0.5, 2), 0.0) + rng.normal(scale=1e-9, size=5)


def _resolved(rng):
    # adjacent differences within a factor e^0.5 and of one sign: every eta > 0.88
    d = np.exp(rng.uniform(-0.25, 0.25, size=4)) * rng.choice([-1.0, 1.0]) * rng.uniform(0.1, 1)
    return np.concatenate([[0.0], np.cumsum(d)]) + rng.uniform(-1, 1)


@pytest.mark.parametrize("scheme,ends", [
    # Teno5_Z's cut-off with C_T from beta_A: 10 on resolved data, 5 next to jumps
    (O.S_WHF_A, (O.S_TENO5_CT10, O.S_TENO5)),
    (O.S_WHF_COMPARE, (O.S_TENO5_CT10, O.S_TENO5)),
    # chi < C_TA with C_TA = 1e-10 on resolved data: whf_TcnsN's 1e-10
    (O.S_AI_A, (O.S_WHF_N, None)),
    # the eta-law TENO-S forms: resolved data give s = 1 in both F002 laws
    (O.S_AI_F002, (O.S_TENOSA_F002, None)),
    # C = min(1e-5, |b0 + b2 - 2 b1|) = 1e-5 next to jumps: ai_S's C_T'
    (O.S_AI_AZS, (None, O.S_AI_S)),
    # k = 5 next to jumps: ctp[0] = C_T' of pow(10, -5) = 1e-5, ai_S's
    (O.S_AI_AS, (None, O.S_AI_S)),
])
def test_research_variants_reduce_at_the_ends_of_eta(scheme, ends):
    """Each adaptive research variant reduces, bit for bit, to a fixed
    scheme at one end of its indicator: resolved data (eta_min > 0.88, so
    m = 0) and data next to a jump (eta_min ~ 1e-7, so m ~ 1).  A wrong
    branch, constant or eps in the adaptive part breaks an equality."""
    rng = np.random.default_rng(31 + scheme)
    for which, gen in ((0, _resolved), (1, _jump)):
        if ends[which] is None:
            continue
        diff = [q for q in (gen(rng) for _ in range(400)) if O.inter5(scheme, q) != O.inter5(ends[which], q)]
        assert not diff, (which, diff[:2])


# Integer data whose smoothness indicators (interScheme.hpp's, evaluated
# exactly for small integers) satisfy beta0 + beta2 = 2 beta1 with every beta
# positive, found by search: whf_ai_TcnsN_AZS_yuanbao's C = min(1e-5,
# |beta0 + beta2 - 2 beta1|) is then exactly 0, C_T' = 0, and its TENO-S test
# keeps every stencil, while the same test with C = 1e-5 (ai_S) cuts one.
_AZS_FLAT = [(-5, -5, 4, 2, -1), (-4, -4, 5, 3, 0), (-1, 2, 4, -5, -5), (0, -3, -5, 4, 4),
             (0, 3, 5, -4, -4), (1, -2, -4, 5, 5), (4, 4, -5, -3, 0), (5, 5, -4, -2, 1)]


def _quartic_half_node(q):
    """The quartic through q_0..q_4 (at x = 0..4) evaluated at x = 2.5: the
    5th-order linear stencil, from Lagrange's weights (exact in binary for
    small integers)."""
    from fractions import Fraction as Fr
    out = Fr(0)
    for j in range(5):
        wj = Fr(1)
        for m in range(5):
            if m != j:
                wj *= (Fr(5, 2) - m) / (j - m)
        out += wj * Fr(q[j])
    return float(out)


def test_azs_keeps_every_stencil_where_its_indicator_vanishes():
    """The |beta0 + beta2 - 2 beta1| end of the AZS law (the jump end, C = 1e-5,
    is pinned against ai_S above): on the data above, scaled by powers of two
    (indicator still exactly 0), AZS returns the linear 5th-order value and
    ai_S does not.  A slip in the indicator (another combination of the betas,
    max in place of min) gives C > 0 and a cut stencil.  Where the signed
    combination is negative and small, C is its absolute value: on the rough
    data below (tau well above the smallest beta) the positive C_T' cuts a
    stencil, where a negative C (no absolute value) would give a NaN C_T',
    fail every cut comparison and keep all stencils."""
    for q0 in _AZS_FLAT:
        for k in (0, 10, 20):
            q = [x * 2.0 ** -k for x in q0]
            lin = _quartic_half_node(q0) * 2.0 ** -k
            assert O.inter5(O.S_AI_AZS, q) == lin, (q0, k)
            assert O.inter5(O.S_AI_S, q) != lin, (q0, k)
    # integer data with beta0 + beta2 - 2 beta1 = -7/2, -39/4, -39/4 (exact),
    # scaled by 2^-10 so that |.| < 1e-5
    for q0 in [(-3, -3, 0, -1, -3), (-3, -1, 3, 3, 2), (-2, -3, -3, 1, 3)]:
        q = [x * 2.0 ** -10 for x in q0]
        v = O.inter5(O.S_AI_AZS, q)
        assert math.isfinite(v) and v != _quartic_half_node(q0) * 2.0 ** -10, q0


def test_whf_s_forms_equal_at_k5_and_lad_bounds():
    """whf_TcnsN_AS at k = 5 (eta of the raw values tiny: a zero next to a
    jump) is whf_TcnsN_S (C_T' of 1e-5); whf_TcnsN_LAD on data with
    tau = 0 (theta = 1, m = 10) is whf_TcnsN (C_T = 1e-10) and keeps every
    stencil; ai LADS at tau = 0 takes C_T' of 1e-10."""
    rng = np.random.default_rng(8)
    for _ in range(300):
        q = _jump(rng)
        q[0] = 0.0
        assert O.inter5(O.S_WHF_AS, q) == O.inter5(O.S_WHF_S, q), q
    for _ in range(300):
        a, b, c = rng.normal(size=3)
        q = np.array([a, b, c, b, a])  # beta0 == beta2 exactly: tau = 0
        assert O.inter5(O.S_WHF_LAD, q) == O.inter5(O.S_WHF_N, q) == O.inter5(O.S_LINEAR, q)
        assert O.inter5(O.S_AI_LADS, q) == O.inter5(O.S_LINEAR, q)


def test_muscl5_minmod_values():
    """musclIn5: q2 + minmod(q2 - q1, q3 - q2) / 2 of its middle three values."""
    assert O.inter5(O.S_MUSCL5, [9.0, 1.0, 2.0, 4.0, -7.0]) == 2.5   # slopes 1, 2: the smaller
    assert O.inter5(O.S_MUSCL5, [0.0, 4.0, 3.0, 1.0, 5.0]) == 2.5    # slopes -1, -2
    assert O.inter5(O.S_MUSCL5, [0.0, 1.0, 3.0, 2.0, 0.0]) == 3.0    # extremum: no slope
    assert O.inter5(O.S_MUSCL5, [5.0, 1.0, 1.0, 7.0, 2.0]) == 1.0    # one slope zero


def test_congaa_cut_is_within_teno_s_ct7():
    """Teno5_CongAA uses TENO-S's test with C_T = 10^-k, k in 5..10 -- below
    CT7's 0.0729, so it cuts a subset of what TENO-S CT7 cuts; on resolved
    data (k = 10) it keeps every stencil."""
    rng = np.random.default_rng(17)
    for _ in range(1500):
        q = rng.normal(size=5) * rng.choice([1e-3, 1.0])
        a, b = _cut(O.S_CONGAA, q), _cut(O.S_TENOS_CT7, q)
        if a is not None and b is not None:
            assert a <= b, q
    q = 1.0 + 0.1 * np.arange(5.0) ** 2  # values of one sign and scale: eta ~ 1
    assert O.inter5(O.S_CONGAA, q) == O.inter5(O.S_LINEAR, q)


def test_weno_weights_js_known_value():
    # step data: JS weights give a known convex blend (hand-evaluated formula)
    q = [0.0, 0.0, 0.0, 1.0, 1.0]
    b = [0.0, 1.25, 3.25]
    d = [1 / 16, 5 / 8, 5 / 16]
    u = [0.0, 0.375, 0.375 * 0 + 0.75 - 0.125]
    a = [d[i] / (1e-6 + b[i]) ** 2 for i in range(3)]
    want = sum(a[i] * u[i] for i in range(3)) / sum(a)
    assert abs(O.inter5(O.S_JS, q) - want) < 1e-15


# -------------------------------------------------------------- eigensystem
def _rand_prim(rng, dim):
    p = [rng.uniform(0.3, 2.0)] + list(rng.uniform(-1.5, 1.5, size=dim)) + [rng.uniform(0.3, 2.0)]
    return np.array(p)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_eigen_left_right_inverse(dim):
    rng = np.random.default_rng(3 + dim)
    for ax in range(dim):
        for _ in range(10):
            pl, pr = _rand_prim(rng, dim), _rand_prim(rng, dim)
            L, R = O.eigen(dim, pl, pr, ax)
            np.testing.assert_allclose(L @ R, np.eye(dim + 2), atol=1e-12)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_char_roundtrip_identity(dim):
    rng = np.random.default_rng(7 + dim)
    for ax in range(dim):
        for _ in range(10):
            pl, pr, p = _rand_prim(rng, dim), _rand_prim(rng, dim), _rand_prim(rng, dim)
            _, back = O.prim_char_roundtrip(dim, pl, pr, ax, p)
            np.testing.assert_allclose(back, p, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
def test_eigenvalues_of_roe_matrix(dim):
    # R diag(lambda) L must equal the Roe-averaged flux Jacobian: check that
    # L A R is diagonal with (Vn, Vn, Vn-c, Vn+c[, Vn]) where A is built from
    # finite differences of the physical flux at the Roe state (pl == pr).
    rng = np.random.default_rng(11)
    for ax in range(dim):
        p = _rand_prim(rng, dim)
        L, R = O.eigen(dim, p, p, ax)
        nv = dim + 2

        def cons(pp):
            r = pp[0]
            vel = pp[1:1 + dim]
            return np.concatenate([[r], r * vel, [pp[-1] / 0.4 + 0.5 * r * vel @ vel]])

        def flux_of_cons(u):
            return O.node_flux(dim, O.cons_to_prim(dim, u), ax)

        u0 = cons(p)
        A = np.zeros((nv, nv))
        for j in range(nv):
            e = np.zeros(nv)
            e[j] = 1e-6 * max(1.0, abs(u0[j]))
            A[:, j] = (flux_of_cons(u0 + e) - flux_of_cons(u0 - e)) / (2 * e[j])
        Lam = L @ A @ R
        c = math.sqrt(1.4 * p[-1] / p[0])
        vn = p[1 + ax]
        want = [vn, vn, vn - c, vn + c] + ([vn] if dim == 3 else [])
        np.testing.assert_allclose(np.diag(Lam), want, atol=2e-6)
        off = Lam - np.diag(np.diag(Lam))
        assert np.abs(off).max() < 2e-6


# ------------------------------------------------------- reference symmetry
def test_reference_symmetry_checks_bitwise():
    """zone_FD/symmetrytest.cpp:11-82 with its own states, bit for bit."""
    priml = np.array([0.18217903560635557, 0.8040523419013228, 1.206, 0.07123495167920027])
    primr = np.array([0.52330217813773561, 0.0097235893105526473, 1.2060000000000002, 0.294926010838852])
    prim = np.array([1.0, 0.05, 0.8, 0.8])
    sw = lambda s: np.array([s[0], s[2], s[1], s[3]])
    primlm, primrm, primm = sw(priml), sw(primr), sw(prim)

    Char, primTrans = O.prim_char_roundtrip(2, priml, priml, 0, prim)
    Charm, primTransm = O.prim_char_roundtrip(2, primlm, primlm, 1, primm)
    assert Char[0] == -Charm[0] and Char[1] == Charm[1] and Char[2] == Charm[2] and Char[3] == Charm[3]
    assert primTrans[0] == primTransm[0] and primTrans[2] == primTransm[1] and primTrans[3] == primTransm[3]

    flux = O.riemann(2, O.F_ROE, priml, primr, 0)
    neg_n = lambda s: np.array([s[0], s[1], -s[2], s[3]])
    fluxm = O.riemann(2, O.F_ROE, neg_n(primrm), neg_n(primlm), 1)
    assert flux[0] == -fluxm[0] and flux[2] == -fluxm[1] and flux[3] == -fluxm[3]
    assert abs(flux[1] - fluxm[2]) <= 1e-15 * abs(flux[1])

    fluxInv = O.riemann(2, O.F_HLLC, primr, priml, 0)
    fluxInvm = O.riemann(2, O.F_HLLC, primrm, primlm, 1)
    assert fluxInv[0] == fluxInvm[0] and fluxInv[2] == fluxInvm[1] and fluxInv[3] == fluxInvm[3]

    fluxs = O.node_flux(2, primr, 0)
    fluxsm = O.node_flux(2, primrm, 1)
    assert fluxs[0] == fluxsm[0] and fluxs[2] == fluxsm[1] and fluxs[3] == fluxsm[3]


# ---------------------------------------------------------- Riemann fluxes
@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("flux", [O.F_HLLC, O.F_ROE])
def test_riemann_consistency(dim, flux):
    rng = np.random.default_rng(13)
    for ax in range(dim):
        for _ in range(10):
            p = _rand_prim(rng, dim)
            np.testing.assert_allclose(O.riemann(dim, flux, p, p, ax), O.node_flux(dim, p, ax), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_hllc_supersonic_upwind(dim):
    rng = np.random.default_rng(17)
    for ax in range(dim):
        pl, pr = _rand_prim(rng, dim), _rand_prim(rng, dim)
        pl[1 + ax] = 6.0
        pr[1 + ax] = 6.5  # both states supersonic to the right
        np.testing.assert_allclose(O.riemann(dim, O.F_HLLC, pl, pr, ax), O.node_flux(dim, pl, ax), rtol=1e-13)
        pl[1 + ax], pr[1 + ax] = -6.5, -6.0
        np.testing.assert_allclose(O.riemann(dim, O.F_HLLC, pl, pr, ax), O.node_flux(dim, pr, ax), rtol=1e-13)


# ---------------------------------------------------------- 3D extension
@pytest.mark.parametrize("flux", [O.F_HLLC, O.F_ROE])
@pytest.mark.parametrize("scheme", [O.S_JS, O.S_Z, O.S_TENO5])
def test_3d_reduces_to_2d_bitwise(flux, scheme):
    rng = np.random.default_rng(19)
    nx, ny = 12, 10
    p2 = np.stack([rng.uniform(0.5, 1.5, (ny, nx)), rng.uniform(-0.5, 0.5, (ny, nx)),
                   rng.uniform(-0.5, 0.5, (ny, nx)), rng.uniform(0.5, 1.5, (ny, nx))], -1)
    c2 = np.empty_like(p2)
    c2[..., 0] = p2[..., 0]
    c2[..., 1] = p2[..., 0] * p2[..., 1]
    c2[..., 2] = p2[..., 0] * p2[..., 2]
    c2[..., 3] = p2[..., 3] / 0.4 + 0.5 * p2[..., 0] * (p2[..., 1] ** 2 + p2[..., 2] ** 2)
    c3 = np.zeros((1, ny, nx, 5))
    c3[0, ..., [0, 1, 2, 4]] = np.moveaxis(c2, -1, 0)
    g2 = O.make_grid(2, (nx, ny), 0.1, scheme, flux, 1, [1, 1, 1, 1])
    g3 = O.make_grid(3, (nx, ny, 1), 0.1, scheme, flux, 1, [1] * 6)
    r2, _ = O.rhs(g2, c2.reshape(-1, 4))
    r3, _ = O.rhs(g3, c3.reshape(-1, 5))
    assert np.array_equal(r3[:, [0, 1, 2, 4]], r2)
    assert np.all(r3[:, 3] == 0)


# ----------------------------------------------- whole-RHS invariants
def _smooth_state(dim, shape, rng, amp=0.1):
    grids = np.meshgrid(*[np.arange(s) * 2 * np.pi / s for s in shape], indexing="ij")
    def f():
        ph = rng.uniform(0, 2 * np.pi, dim)
        k = rng.integers(1, 3, dim)
        return np.sin(sum(k[i] * grids[i] + ph[i] for i in range(dim)))
    rho = 1 + amp * f()
    vel = [0.3 * f() for _ in range(dim)]
    p = 1 + amp * f()
    c = [rho] + [rho * v for v in vel] + [p / 0.4 + 0.5 * rho * sum(v * v for v in vel)]
    arr = np.stack(c, -1)  # indexing (x, y, z): reorder to x fastest
    arr = np.transpose(arr, list(range(dim))[::-1] + [dim])
    return np.ascontiguousarray(arr.reshape(-1, dim + 2))


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("flux", [O.F_HLLC, O.F_ROE])
@pytest.mark.parametrize("charwise", [0, 1])
def test_conservation_periodic(dim, flux, charwise):
    rng = np.random.default_rng(23)
    shape = {1: (32,), 2: (16, 12), 3: (8, 6, 10)}[dim]
    c = _smooth_state(dim, shape, rng)
    g = O.make_grid(dim, shape, 0.05, O.S_TENO5, flux, charwise, [1] * 6)
    r, _ = O.rhs(g, c)
    tot = r.sum(0)
    scale = np.abs(r).sum(0) + 1e-30
    assert np.all(np.abs(tot) / scale < 1e-13)


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("scheme", SCHEMES[:6])
def test_freestream_preserved(dim, scheme):
    shape = {1: (16,), 2: (9, 7), 3: (5, 6, 7)}[dim]
    n = int(np.prod(shape))
    vel = [0.3, -0.2, 0.1][:dim]
    c = np.array([1.2] + [1.2 * v for v in vel] + [1.0 / 0.4 + 0.5 * 1.2 * sum(v * v for v in vel)])
    cons = np.tile(c, (n, 1))
    for flux in (O.F_HLLC, O.F_ROE):
        for bc in (1, 7):
            g = O.make_grid(dim, shape, 0.1, scheme, flux, 1, [bc] * 6)
            r, _ = O.rhs(g, cons)
            assert np.all(r == 0.0)


def test_rhs_transpose_symmetry_2d():
    """Swapping x<->y (and u<->v) maps the RHS onto itself."""
    rng = np.random.default_rng(29)
    n = 14
    c = _smooth_state(2, (n, n), rng).reshape(n, n, 4)  # [y][x][v]
    ct = np.transpose(c, (1, 0, 2))[..., [0, 2, 1, 3]].copy()
    for flux in (O.F_HLLC, O.F_ROE):
        g = O.make_grid(2, (n, n), 0.1, O.S_TENO5, flux, 1, [1] * 4)
        r, _ = O.rhs(g, c.reshape(-1, 4))
        rt, _ = O.rhs(g, ct.reshape(-1, 4))
        back = np.transpose(rt.reshape(n, n, 4), (1, 0, 2))[..., [0, 2, 1, 3]]
        np.testing.assert_allclose(back, r.reshape(n, n, 4), rtol=0, atol=1e-12 * np.abs(r).max())


def test_mnd6_exact_for_quartic_flux():
    # Linear5 half nodes are exact for degree <= 4 and MND6 differentiates
    # exactly, so the derivative of a quartic is recovered to round-off.
    co = np.array([0.3, -0.2, 0.5, 1.0, 2.0])
    h = 0.05
    n = 20
    xc = -0.5 + h / 2 + h * np.arange(n)
    gl = np.polyval(co, -0.5 - h / 2 - h * np.arange(10))
    gr = np.polyval(co, 0.5 + h / 2 + h * np.arange(10))
    r = O.accuracy_rhs(O.S_LINEAR, np.polyval(co, xc), gl, gr, h)
    np.testing.assert_allclose(r, np.polyval(np.polyder(co), xc), rtol=0, atol=1e-11)


def test_sod_against_exact_riemann_solution():
    """Config S through the oracle: 400 cells, WCNS5-JS component-wise + HLLC,
    CFL 0.5, t = 0.2; density L1 error against Toro's exact solution."""
    from cfd_inwsl_amd import cases
    from tests.test_driver import _oracle_steps_cfl
    n, h = 400, 1.0 / 400
    x = -0.5 + (np.arange(n) + 0.5) * h
    cons = np.stack([np.where(x < 0, 1.0, 0.125), np.zeros(n),
                     np.where(x < 0, 1.0 / (1.4 - 1) * 1, 1.0 / (1.4 - 1) * 0.1)], -1)
    g = O.make_grid(1, (n,), h, O.S_JS, O.F_HLLC, 0, [7, 7])
    ref, t = _oracle_steps_cfl(g, cons, 0.5, 0.2, 1)
    exact = cases.sod_exact(x + 0.5, 0.2)
    l1 = np.abs(ref[:, 0] - exact[:, 0]).mean()
    assert abs(t - 0.2) < 1e-14
    assert 5e-4 < l1 < 3e-3
    # characteristic-wise (the reference's own 1D path, with its positivity limiter)
    g2 = O.make_grid(1, (n,), h, O.S_JS, O.F_HLLC, 1, [7, 7])
    ref2, _ = _oracle_steps_cfl(g2, cons, 0.5, 0.2, 1)
    assert np.abs(ref2[:, 0] - exact[:, 0]).mean() < 3e-3


# ---------------------------------------------------------------- sources and time-dependent boundaries
@pytest.mark.parametrize("scheme", [O.S_JS, O.S_TENO5])
def test_gravity_source_balances_hydrostatic_pressure(scheme):
    """Uniform density, u = (0.3, 0), p = 1 + rho*y, g = (0.5, 1): every
    scheme reproduces linear data exactly and MND6 differentiates it exactly,
    so in the interior rhs = dF/dx + dG/dy - S = (0, -0.5 rho, rho - rho,
    -rho u.g) = (0, -0.6, 0, -0.18) for rho = 1.2.  Without gravity the y
    momentum reads dp/dy = rho instead."""
    nx, ny, h, rho, u = 10, 24, 0.05, 1.2, 0.3
    y = (np.arange(ny) + 0.5) * h
    p = 1.0 + rho * y
    cons = np.zeros((ny, nx, 4))
    cons[..., 0] = rho
    cons[..., 1] = rho * u
    cons[..., 3] = (p / 0.4 + 0.5 * rho * u * u)[:, None]
    cons = cons.reshape(-1, 4)
    for flux in (O.F_HLLC, O.F_ROE):
        for gravity, want in (((0.5, 1.0, 0.0), [0.0, -0.5 * rho, 0.0, -rho * u * 0.5]),
                              ((0.0, 0.0, 0.0), [0.0, 0.0, rho, 0.0])):
            g = O.make_grid(2, (nx, ny), h, scheme, flux, 1, [1, 1, 7, 7], gravity=gravity)
            r = O.rhs(g, cons)[0].reshape(ny, nx, 4)[5:-5]
            np.testing.assert_allclose(r, np.broadcast_to(want, r.shape), rtol=0, atol=1e-11)


def _shock_front(row, x, lo, hi):
    """x where the density first falls below the mid value, scanning from
    the post-shock side (linear interpolation between the two cells)."""
    mid = 0.5 * (lo + hi)
    i = int(np.argmax(row < mid))
    return x[i - 1] + (row[i - 1] - mid) / (row[i - 1] - row[i]) * (x[i] - x[i - 1])


def test_double_mach_top_boundary_follows_the_exact_shock():
    """Double Mach reflection: the Mach-10 shock crosses y at
    x_s(y, t) = 1/6 + (y + 20 t)/sqrt(3).  Away from the wall the computed
    front sits on it, on the top row (fed by the time-dependent DoubleMachUp
    ghost cells, so a stale or mis-staged boundary time shows there) and in
    the interior; the wall's Dirichlet strip x < 1/6 keeps the post-shock
    state; the time advances as the reference's RK3 step does."""
    nx, ny = 96, 24
    h = 1.0 / ny
    x = (np.arange(nx) + 0.5) * h
    yv = (np.arange(ny) + 0.5) * h
    X, Y = np.meshgrid(x, yv, indexing="xy")
    post = [8.0, 8.25 * math.cos(math.pi / 6), -8.25 * math.sin(math.pi / 6), 116.5]
    isp = Y >= math.sqrt(3) * (X - 1.0 / 6.0)
    r = np.where(isp, post[0], 1.4)
    u = np.where(isp, post[1], 0.0)
    v = np.where(isp, post[2], 0.0)
    p = np.where(isp, post[3], 1.0)
    cons = np.stack([r, r * u, r * v, p / 0.4 + 0.5 * r * (u * u + v * v)], -1).reshape(-1, 4)
    bc_prim = np.zeros((6, 5))
    bc_prim[0, :4] = bc_prim[2, :4] = bc_prim[3, :4] = post
    g = O.make_grid(2, (nx, ny), h, O.S_TENO5, O.F_HLLC, 1, [3, 7, 103, 8], bc_prim)  # Dirichlet, outflow, wall, top
    t = 0.0
    for _ in range(30):
        dt = O.cfl_dt(g, cons, 0.5)
        g.t = t
        cons, t_new = O.rk3(g, cons, dt=dt, nsteps=1)
        assert t_new == ((t + dt) - dt / 2 - dt / 2) + dt
        t = t_new
    rho = cons.reshape(ny, nx, 4)[..., 0]
    for j in (ny - 1, ny - 4):
        xs = 1.0 / 6.0 + (yv[j] + 20 * t) / math.sqrt(3)
        assert abs(_shock_front(rho[j], x, 1.4, 8.0) - xs) < 0.5 * h, (j, xs)
    # the Dirichlet part of the wall keeps the post-shock inflow (v = -4.1)
    # next to the wall, where a reflecting wall would stop it; the supersonic
    # post-shock flow lets only numerical leakage (~1e-4) travel upstream
    st = cons.reshape(ny, nx, 4)[0, x < 1.0 / 6.0 - 2 * h]
    assert np.all(np.abs(st[:, 0] / post[0] - 1) < 1e-3)
    assert np.all(np.abs(st[:, 2] / st[:, 0] / post[2] - 1) < 1e-3)
