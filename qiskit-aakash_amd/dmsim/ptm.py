"""Pauli transfer matrices (PTMs) of the gates and channels on the path.

A k-qubit PTM acts on the 4^k Pauli components of k qubits.  Single-qubit
maps are handed to the kernels as the three non-trivial rows (X, Y, Z) of
their 4x4 PTM; the I row of a trace-preserving map is (1, 0, 0, 0).

Component order is (I, X, Y, Z) = (0, 1, 2, 3) as in the reference
(dm_simulator.py:61-68).
"""
import numpy as np

_AXIS_PAIR = {'rz': (1, 2), 'ry': (3, 1), 'rx': (2, 3)}


def rotation4(axis, angle, err=(1.0, 0.0)):
    """4x4 PTM of one noisy rotation: the pair (k0, k1) of ``axis`` becomes
    (c k0 - s k1, c k1 + s k0), c = e0 cos(angle+e1), s = e0 sin(angle+e1)
    (basicaertools.py:93-126)."""
    c = err[0] * np.cos(angle + err[1])
    s = err[0] * np.sin(angle + err[1])
    k0, k1 = _AXIS_PAIR[axis]
    R = np.eye(4)
    R[k0, k0] = c
    R[k0, k1] = -s
    R[k1, k1] = c
    R[k1, k0] = s
    return R


def gate_ptm(name, params, rotation_error=None):
    """4x4 PTM of u1 / u3 as the product of its rotations rz(lam), ry(theta),
    rz(phi), applied in that order (basicaertools.py:69-90), each with the
    rotation error of its axis (dm_simulator.py:379-381)."""
    err = rotation_error or {}
    p = [float(x) for x in params]
    if name in ('u3', 'U'):
        seq = [('rz', p[2]), ('ry', p[0]), ('rz', p[1])]
    elif name == 'u1':
        seq = [('rz', p[0])]
    else:
        raise ValueError('not a single-qubit rotation gate: %s' % name)
    M = np.eye(4)
    for axis, ang in seq:
        M = rotation4(axis, ang, err.get(axis, (1.0, 0.0))) @ M
    return M


def depolarizing1(p):
    """rho -> (1-p) rho + p I/2: X, Y, Z scaled by 1-p."""
    return np.diag([1.0, 1 - p, 1 - p, 1 - p])


def rows3(R4):
    """The X, Y, Z rows of a trace-preserving 4x4 PTM, flat (12 values)."""
    R4 = np.asarray(R4, dtype=float)
    if R4[0, 0] != 1.0 or np.any(R4[0, 1:] != 0.0):
        raise ValueError('PTM is not trace preserving (I row must be 1,0,0,0)')
    return np.ascontiguousarray(R4[1:, :]).reshape(12)


def collapse_axis(comp, err):
    """Lossy projective measurement along X/Y/Z: keep component ``comp``
    times err, drop the other two (dm_simulator.py:574-664)."""
    R = np.zeros((4, 4))
    R[0, 0] = 1.0
    R[comp, comp] = err
    return R


def collapse_direction(nvec, err):
    """Measurement along unit vector n: (x,y,z) <- err (n.v) n
    (dm_simulator.py:666-704)."""
    nv = np.asarray(nvec, dtype=float)
    R = np.zeros((4, 4))
    R[0, 0] = 1.0
    R[1:, 1:] = err * np.outer(nv, nv)
    return R


def reset():
    """x, y <- 0; z <- a_I (dm_simulator.py:810-823)."""
    R = np.zeros((4, 4))
    R[0, 0] = 1.0
    R[3, 0] = 1.0
    return R


def memory_noise(f, p, g):
    """x,y <- sqrt(g) f (x,y); z <- g z + (1-g)(2p-1) a_I
    (dm_simulator.py:397-425)."""
    od = np.sqrt(g) * f
    R = np.diag([1.0, od, od, g])
    R[3, 0] = (1 - g) * (2 * p - 1)
    return R


# Noiseless CX in (control, target) component pairs: out (c,t) <- sign * in src.
CX_PERM = {
    (0, 2): (1, (3, 2)), (3, 2): (1, (0, 2)), (0, 3): (1, (3, 3)), (3, 3): (1, (0, 3)),
    (1, 0): (1, (1, 1)), (1, 1): (1, (1, 0)), (1, 2): (1, (2, 3)), (1, 3): (-1, (2, 2)),
    (2, 0): (1, (2, 1)), (2, 1): (1, (2, 0)), (2, 2): (-1, (1, 3)), (2, 3): (1, (1, 2)),
}


def cx_ptm(tsp=(1.0, 0.0)):
    """16x16 PTM of the CX with the transition-selective-pulse error
    [<cos a>, mean a], pair index 4*i_control + i_target.

    Coefficients c, s, c2, s2, cs as in cx_gate_dm_matrix
    (basicaertools.py:328-336); the update rows follow :346-363.
    """
    cav, mean = float(tsp[0]), float(tsp[1])
    c2av = 4 * cav - 3
    c = cav * np.cos(mean)
    s = cav * np.sin(mean)
    c2 = 0.5 * (1 + c2av * np.cos(2 * mean))
    s2 = 0.5 * (1 - c2av * np.cos(2 * mean))
    cs = c2av * np.sin(mean) * np.cos(mean)
    R = np.eye(16)

    def ix(a, b):
        return 4 * a + b

    def row(out, terms):
        R[ix(*out), :] = 0.0
        for coef, src in terms:
            R[ix(*out), ix(*src)] += coef

    row((0, 2), [(s2, (0, 2)), (c2, (3, 2)), (-cs, (0, 3)), (cs, (3, 3))])
    row((3, 2), [(c2, (0, 2)), (s2, (3, 2)), (cs, (0, 3)), (-cs, (3, 3))])
    row((0, 3), [(s2, (0, 3)), (c2, (3, 3)), (cs, (0, 2)), (-cs, (3, 2))])
    row((3, 3), [(c2, (0, 3)), (s2, (3, 3)), (-cs, (0, 2)), (cs, (3, 2))])
    row((1, 0), [(c, (1, 1)), (-s, (2, 0))])
    row((1, 1), [(c, (1, 0)), (-s, (2, 1))])
    row((1, 2), [(-s, (2, 2)), (c, (2, 3))])
    row((1, 3), [(-c, (2, 2)), (-s, (2, 3))])
    row((2, 0), [(s, (1, 0)), (c, (2, 1))])
    row((2, 1), [(s, (1, 1)), (c, (2, 0))])
    row((2, 2), [(s, (1, 2)), (-c, (1, 3))])
    row((2, 3), [(c, (1, 2)), (s, (1, 3))])
    return R


def depolarizing_pair_scale(p):
    """Per-pair scale of independent depolarising channels on both qubits of a
    two-qubit gate: (1-p)^(#non-identity components)."""
    lam = 1.0 - p
    s = np.ones(16)
    for a in range(4):
        for b in range(4):
            s[4 * a + b] = lam ** ((a != 0) + (b != 0))
    return s


def embed1(R4, j, k):
    """Embed a 4x4 PTM on block digit j into a k-qubit block (4^k x 4^k);
    component index sum_j i_j 4^j."""
    eye_hi = np.eye(4 ** (k - 1 - j))
    eye_lo = np.eye(4 ** j)
    return np.kron(np.kron(eye_hi, R4), eye_lo)


def embed2(R16, ja, jb, k):
    """Embed a 16x16 PTM on block digits (ja, jb) (pair index 4*i_a + i_b)
    into a k-qubit block."""
    dim = 4 ** k
    out = np.zeros((dim, dim))
    R = np.asarray(R16).reshape(4, 4, 4, 4)  # [out_a, out_b, in_a, in_b]
    idx = np.arange(dim)
    da = (idx >> (2 * ja)) & 3
    db = (idx >> (2 * jb)) & 3
    rest = idx & ~((3 << (2 * ja)) | (3 << (2 * jb)))
    for col in range(dim):
        ia, ib, r = da[col], db[col], rest[col]
        for oa in range(4):
            for ob in range(4):
                v = R[oa, ob, ia, ib]
                if v != 0.0:
                    out[r | (oa << (2 * ja)) | (ob << (2 * jb)), col] += v
    return out
