"""Brute-force density-matrix oracle: 2^n x 2^n complex rho, Kraus channels.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

This is physics, not a restatement: it pins the Pauli-basis updates of
``refsim`` (and through them the GPU kernels) against textbook operator
algebra.  Conventions shared with the reference:

* rho = sum_c a_c sigma_{c_0} (x) ... (x) sigma_{c_{n-1}}; qubit 0 is the
  first tensor factor, i.e. the most significant bit of the matrix index
  (dm_simulator.py:1230, 1250-1253);
* a_c = Tr(rho sigma_c) / 2^n;
* U3(theta, phi, lam) = [[cos t/2, -e^{i lam} sin t/2],
  [e^{i phi} sin t/2, e^{i(phi+lam)} cos t/2]] (basicaertools.py:49-66).
"""
import itertools

import numpy as np

PAULI = [np.eye(2, dtype=complex),
         np.array([[0, 1], [1, 0]], dtype=complex),
         np.array([[0, -1j], [1j, 0]], dtype=complex),
         np.array([[1, 0], [0, -1]], dtype=complex)]


def u3_matrix(theta, phi, lam):
    c, s = np.cos(theta / 2), np.sin(theta / 2)
    return np.array([[c, -np.exp(1j * lam) * s],
                     [np.exp(1j * phi) * s, np.exp(1j * (phi + lam)) * c]])


def embed(op, qubits, n):
    """Operator on ``qubits`` (listed in op's tensor order) embedded in n qubits."""
    k = len(qubits)
    full = np.zeros((2 ** n, 2 ** n), dtype=complex)
    others = [q for q in range(n) if q not in qubits]
    opr = op.reshape([2] * (2 * k))
    for rest in itertools.product([0, 1], repeat=n - k):
        for bits_out in itertools.product([0, 1], repeat=k):
            for bits_in in itertools.product([0, 1], repeat=k):
                amp = opr[bits_out + bits_in]
                if amp == 0:
                    continue
                r = [0] * n
                c = [0] * n
                for q, b in zip(others, rest):
                    r[q] = b
                    c[q] = b
                for q, bo, bi in zip(qubits, bits_out, bits_in):
                    r[q] = bo
                    c[q] = bi
                ri = int(''.join(map(str, r)), 2) if n else 0
                ci = int(''.join(map(str, c)), 2) if n else 0
                full[ri, ci] = amp
    return full


def pauli_string(c):
    m = np.array([[1.0 + 0j]])
    for ci in c:
        m = np.kron(m, PAULI[ci])
    return m


def to_pauli(rho, n):
    """Pauli coefficients a_c = Tr(rho sigma_c)/2^n, flat, qubit 0 slowest."""
    out = np.zeros(4 ** n)
    for idx, c in enumerate(itertools.product(range(4), repeat=n)):
        out[idx] = np.real(np.trace(rho @ pauli_string(c))) / 2 ** n
    return out


def from_pauli(a, n):
    rho = np.zeros((2 ** n, 2 ** n), dtype=complex)
    for idx, c in enumerate(itertools.product(range(4), repeat=n)):
        if a[idx] != 0:
            rho += a[idx] * pauli_string(c)
    return rho


def zero_state(n):
    rho = np.zeros((2 ** n, 2 ** n), dtype=complex)
    rho[0, 0] = 1
    return rho


def unitary(rho, U, qubits, n):
    F = embed(U, qubits, n)
    return F @ rho @ F.conj().T


def channel(rho, kraus, qubits, n):
    out = np.zeros_like(rho)
    for K in kraus:
        F = embed(K, qubits, n)
        out += F @ rho @ F.conj().T
    return out


def cx_matrix():
    """Control = first listed qubit."""
    m = np.zeros((4, 4), dtype=complex)
    m[0, 0] = m[1, 1] = 1
    m[2, 3] = m[3, 2] = 1
    return m


def tsp_cx_matrix(a):
    """U(a) = diag(1, 1, [[i sin a, cos a], [cos a, i sin a]]): the noisy CX
    of the transition-selective-pulse model (basicaertools.py:319-321)."""
    m = np.zeros((4, 4), dtype=complex)
    m[0, 0] = m[1, 1] = 1
    m[2, 2] = m[3, 3] = 1j * np.sin(a)
    m[2, 3] = m[3, 2] = np.cos(a)
    return m


def gad_kraus(g, p):
    """Generalised amplitude damping: excited-state decay 1-g towards the
    state with ground population p."""
    gam = 1 - g
    return [np.sqrt(p) * np.array([[1, 0], [0, np.sqrt(1 - gam)]]),
            np.sqrt(p) * np.array([[0, np.sqrt(gam)], [0, 0]]),
            np.sqrt(1 - p) * np.array([[np.sqrt(1 - gam), 0], [0, 1]]),
            np.sqrt(1 - p) * np.array([[0, 0], [np.sqrt(gam), 0]])]


def dephase_kraus(f):
    """Off-diagonals scaled by f."""
    return [np.sqrt((1 + f) / 2) * PAULI[0], np.sqrt((1 - f) / 2) * PAULI[3]]


def depol_kraus(p):
    """rho -> (1-p) rho + p I/2."""
    return [np.sqrt(1 - 3 * p / 4) * PAULI[0]] + [np.sqrt(p / 4) * PAULI[k] for k in (1, 2, 3)]


def flip_kraus(axis, err):
    """Flip by a Pauli that anticommutes with ``axis`` (1,2,3), with
    probability (1-err)/2: the ``axis`` component (and the one other
    anticommuting component) shrink by err.  After a dephasing onto ``axis``
    this is the lossy projective measurement."""
    pr = (1 - err) / 2
    flip = {3: 1, 1: 3, 2: 3}[axis]
    return [np.sqrt(1 - pr) * PAULI[0], np.sqrt(pr) * PAULI[flip]]


def measure_collapse(rho, q, axis, err, n):
    """Non-selective measurement along Pauli ``axis`` with polarisation
    retention err: the axis component scales by err, the others vanish."""
    # dephase onto the axis: keep axis component, kill the two others
    kr = [np.sqrt(0.5) * PAULI[0], np.sqrt(0.5) * PAULI[axis]]
    rho = channel(rho, kr, [q], n)
    # shrink the axis component by err: flip with a Pauli anticommuting with it
    return channel(rho, flip_kraus(axis, err), [q], n)


def measure_direction(rho, q, nvec, err, n):
    nvec = np.asarray(nvec, dtype=float)
    P = nvec[0] * PAULI[1] + nvec[1] * PAULI[2] + nvec[2] * PAULI[3]
    kr = [np.sqrt(0.5) * PAULI[0], np.sqrt(0.5) * P]
    rho = channel(rho, kr, [q], n)
    # any unit vector orthogonal to n gives an anticommuting Pauli combination
    perp = np.cross(nvec, [1.0, 0.0, 0.0])
    if np.linalg.norm(perp) < 1e-6:
        perp = np.cross(nvec, [0.0, 1.0, 0.0])
    perp = perp / np.linalg.norm(perp)
    Q = perp[0] * PAULI[1] + perp[1] * PAULI[2] + perp[2] * PAULI[3]
    pr = (1 - err) / 2
    return channel(rho, [np.sqrt(1 - pr) * PAULI[0], np.sqrt(pr) * Q], [q], n)


def reset(rho, q, n):
    k0 = np.array([[1, 0], [0, 0]], dtype=complex)
    k1 = np.array([[0, 1], [0, 0]], dtype=complex)
    return channel(rho, [k0, k1], [q], n)


def basis_probabilities(rho, n, basis='Z', err=1.0):
    """Outcome probabilities of measuring every qubit along X/Y/Z with
    polarisation retention err; key char j = qubit j."""
    r = rho
    for q in range(n):
        r = channel(r, flip_kraus({'X': 1, 'Y': 2, 'Z': 3}[basis], err), [q], n)
    if basis == 'X':
        H = np.array([[1, 1], [1, -1]]) / np.sqrt(2)
        R = np.array([[1.0]])
        for _ in range(n):
            R = np.kron(R, H)
        r = R @ r @ R.conj().T
    elif basis == 'Y':
        # |+i> -> |0>, |-i> -> |1>
        V = np.array([[1, -1j], [1, 1j]]) / np.sqrt(2)
        R = np.array([[1.0]])
        for _ in range(n):
            R = np.kron(R, V)
        r = R @ r @ R.conj().T
    d = np.real(np.diag(r))
    keys = [''.join(s) for s in itertools.product('01', repeat=n)]
    return dict(zip(keys, d))
