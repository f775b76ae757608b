"""Mathematical pins of the two oracle functions that no recorded output of
the reference covers: the transition-selective-pulse (TSP) CX at <cos> < 1
and the Bell measurement on an asymmetric state.

* TSP CX (basicaertools.py:310-392).  The documented model is the unitary
  U(a) = diag(1, 1, [[i sin a, cos a], [cos a, i sin a]]) averaged over a
  fluctuating angle a = e1 + delta with a symmetric distribution of delta.
  Every entry of the averaged channel's transfer matrix is a trigonometric
  polynomial of degree <= 2 in delta, so for symmetric distributions it is
  affine in the two moments m1 = <cos delta> and m2 = <cos 2 delta>.  The
  reference takes m1 = err[0] and substitutes m2 = 4 m1 - 3 (exact to
  second order in delta).  Two symmetric distributions with the same m1 and
  different m2 (a two-point one and a three-point one) are averaged exactly
  with dense 4x4 unitaries; their affine extrapolation to m2 = 4 m1 - 3 is the
  documented model with the reference's substitution, and the oracle must
  equal it to rounding.  A swapped c2 / s2, a sign error in cs, or a wrong
  first-order coefficient changes the result by O(1 - m1) and fails.

* Bell measurement (dm_simulator.py:721-777).  The reference reshapes the
  state to (4^(n-q2-1), 4, 4^(q2-q1-1), 4, 4^q1), q1 < q2, so its two axes
  are the qubits at strides 4^q2 and 4^q1: qubits n-1-q2 and n-1-q1 (the
  mirrored pair).  The brute force builds the four Bell projectors on that
  pair and applies err * (Bell dephasing) + (1 - err) * (pair fully
  depolarised), which scales the XX, YY, ZZ pair components by err and
  removes every other non-identity pair component.  Random mixed states at
  n = 4 and 5 have no symmetry that would hide a wrong axis, a transposed
  reduced matrix or a wrong probability sign.
"""

import numpy as np
import pytest

from oracle import dense, refsim


def random_state(n, rng):
    rho = np.zeros((2 ** n, 2 ** n), dtype=complex)
    ws = rng.random(3)
    ws /= ws.sum()
    for w in ws:
        v = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
        v /= np.linalg.norm(v)
        rho += w * np.outer(v, v.conj())
    return rho, dense.to_pauli(rho, n)


# ---------------------------------------------------------------------------
# TSP CX at <cos> < 1
# ---------------------------------------------------------------------------

def averaged_tsp(rho, c, t, n, e1, points):
    """sum_k w_k U(e1 + d_k) rho U(e1 + d_k)^dagger on (control c, target t)."""
    out = np.zeros_like(rho)
    for w, d in points:
        out += w * dense.unitary(rho, dense.tsp_cx_matrix(e1 + d), [c, t], n)
    return out


def model_with_substitution(rho, c, t, n, cav, e1):
    """The averaged U(a) channel with <cos delta> = cav and <cos 2 delta>
    replaced by 4 cav - 3, from two exact symmetric averages (Pauli vector)."""
    dA = np.arccos(cav)                      # +-dA: m1 = cav, m2 = cos 2 dA
    A = [(0.5, dA), (0.5, -dA)]
    m2A = np.cos(2 * dA)
    dB = np.arccos(2 * cav - 1)              # 0 w.p. 1/2, +-dB w.p. 1/4: m1 = cav
    B = [(0.5, 0.0), (0.25, dB), (0.25, -dB)]
    m2B = 0.5 + 0.5 * np.cos(2 * dB)
    assert abs(0.5 + 0.5 * np.cos(dB) - cav) < 1e-15
    PA = dense.to_pauli(averaged_tsp(rho, c, t, n, e1, A), n)
    PB = dense.to_pauli(averaged_tsp(rho, c, t, n, e1, B), n)
    target = 4 * cav - 3
    return PA + (target - m2A) / (m2B - m2A) * (PB - PA), PA


@pytest.mark.parametrize('c,t,cav,e1', [(0, 1, 0.9, 0.0), (1, 0, 0.9, 0.3), (0, 2, 0.75, -0.4),
                                        (2, 1, 0.97, 1.2), (2, 0, 0.6, 2.5)])
def test_tsp_cx_below_unit_cos_is_the_averaged_unitary(c, t, cav, e1):
    n = 3
    rng = np.random.default_rng(int(100 * cav) + 7 * c + t)
    rho, a = random_state(n, rng)
    want, two_point = model_with_substitution(rho, c, t, n, cav, e1)
    got = refsim.apply_cx(a.reshape(n * [4]).copy(), c, t, [cav, e1], n).reshape(-1)
    assert np.max(np.abs(got - want)) < 1e-13
    # the substitution differs from the exact two-point average by
    # 2 (1 - cav)^2 in <cos 2 delta> (O(delta^4)), never by more
    bound = 2 * (1 - cav) ** 2 * np.max(np.abs(a)) * 4 + 1e-14
    assert np.max(np.abs(got - two_point)) <= bound


@pytest.mark.parametrize('mistake', ['swap_c2_s2', 'cs_sign', 'c_sign', 's_first_order'])
def test_tsp_pin_catches_planted_mistakes(monkeypatch, mistake):
    """The comparison above fails for each plausible slip in the
    coefficients (planted in the oracle, on the CPU)."""
    orig = refsim.cx_coefficients

    def bad(err):
        c, s, c2, s2, cs = orig(err)
        if mistake == 'swap_c2_s2':
            return c, s, s2, c2, cs
        if mistake == 'cs_sign':
            return c, s, c2, s2, -cs
        if mistake == 'c_sign':
            return -c, s, c2, s2, cs
        return c, 0.5 * s, c2, s2, cs

    monkeypatch.setattr(refsim, 'cx_coefficients', bad)
    n, c, t, cav, e1 = 3, 1, 0, 0.85, 0.35
    rho, a = random_state(n, np.random.default_rng(42))
    want, _ = model_with_substitution(rho, c, t, n, cav, e1)
    got = refsim.apply_cx(a.reshape(n * [4]).copy(), c, t, [cav, e1], n).reshape(-1)
    assert np.max(np.abs(got - want)) > 1e-6


# ---------------------------------------------------------------------------
# Bell measurement on the mirrored pair
# ---------------------------------------------------------------------------

BELL_VECTORS = [np.array([1, 0, 0, 1]) / np.sqrt(2),     # Bell_1 = (|00> + |11>)/sqrt2
                np.array([1, 0, 0, -1]) / np.sqrt(2),    # Bell_2 = (|00> - |11>)/sqrt2
                np.array([0, 1, 1, 0]) / np.sqrt(2),     # Bell_3 = (|01> + |10>)/sqrt2
                np.array([0, 1, -1, 0]) / np.sqrt(2)]    # Bell_4 = (|01> - |10>)/sqrt2


def partial_trace_pair(rho, ua, ub, n):
    """Reduced 4x4 state of qubits (ua, ub), ua the first tensor factor."""
    r = rho.reshape([2] * (2 * n))
    keep = [ua, ub]
    others = [q for q in range(n) if q not in keep]
    perm = keep + others + [n + q for q in keep] + [n + q for q in others]
    r = np.transpose(r, perm).reshape(4, 2 ** (n - 2), 4, 2 ** (n - 2))
    return np.einsum('ikjk->ij', r)


def bell_brute_force(rho, n, qa, qb, err):
    q1, q2 = min(qa, qb), max(qa, qb)
    ua, ub = n - 1 - q2, n - 1 - q1          # the pair at strides 4^q2, 4^q1
    red = partial_trace_pair(rho, ua, ub, n)
    reduced = np.array([[np.real(np.trace(red @ np.kron(dense.PAULI[i], dense.PAULI[j]))) / 4
                         for j in range(4)] for i in range(4)])
    projs = [np.outer(v, v.conj()).astype(complex) for v in BELL_VECTORS]
    dephased = np.zeros_like(rho)
    for P in projs:
        F = dense.embed(P, [ua, ub], n)
        dephased += F @ rho @ F
    depolarised = dense.channel(rho, [0.25 * np.kron(dense.PAULI[i], dense.PAULI[j])
                                      for i in range(4) for j in range(4)], [ua, ub], n)
    post = err * dephased + (1 - err) * depolarised
    probs = [float(np.real(np.trace(dense.embed(P, [ua, ub], n) @ post))) for P in projs]
    return dense.to_pauli(post, n), probs, reduced


@pytest.mark.parametrize('n,qa,qb,err', [(4, 0, 1, 1.0), (4, 3, 1, 0.9), (4, 2, 0, 0.8), (5, 1, 3, 1.0),
                                         (5, 4, 0, 0.95), (5, 2, 3, 0.7)])
def test_bell_measure_is_bell_projection_on_mirrored_pair(n, qa, qb, err):
    rng = np.random.default_rng(11 * n + 3 * qa + qb)
    rho, a = random_state(n, rng)
    post_want, probs_want, red_want = bell_brute_force(rho, n, qa, qb, err)
    post, probs, red = refsim.bell_measure(a.reshape(n * [4]).copy(), n, qa, qb, err)
    assert np.max(np.abs(red - red_want)) < 1e-15
    assert np.max(np.abs(post.reshape(-1) - post_want)) < 1e-15
    got = [probs['Bell_%d' % i] for i in range(1, 5)]
    assert np.max(np.abs(np.array(got) - np.array(probs_want))) < 1e-14
    assert abs(sum(got) - 1.0) < 1e-14


def test_bell_pin_distinguishes_the_pairs():
    """On an asymmetric state the mirrored pair and the literal pair
    (qa, qb) give different outcomes, so the test above decides which one
    the oracle reads."""
    n, qa, qb = 4, 0, 1
    rho, a = random_state(n, np.random.default_rng(5))
    _, probs, red = refsim.bell_measure(a.reshape(n * [4]).copy(), n, qa, qb, 1.0)
    lit = partial_trace_pair(rho, qa, qb, n)
    lit_red = np.array([[np.real(np.trace(lit @ np.kron(dense.PAULI[i], dense.PAULI[j]))) / 4
                         for j in range(4)] for i in range(4)])
    assert np.max(np.abs(red - lit_red)) > 1e-3
