"""A tape whose dot products group into DENSE / DENSE_REV and whose reverse
sweep meets every guard of the reference: zero adjoints (a member scaled by
0 and a max that picks its constant), zero weights, infinite adjoints (sqrt
and ln at 0) and NaN partials (ln of a negative member)."""

import numpy as np

from echidna_amd.tape import record

N = 6
# no weight is 1: x * 1 records no product (the recorder aliases it), which
# would give the members different operand kinds
W = np.array([[0.5, -0.9, 0.0, 0.25, 2.0, -0.75],
              [0.9, 0.0, 0.0, 0.5, -0.5, 0.0],
              [-0.25, 0.75, 1.5, 0.0, 0.125, 1.25],
              [0.3, 0.3, -0.3, 0.3, -0.3, 0.3],
              [2.0, 0.0, -1.1, 0.0, 1.1, 0.0],
              [0.1, 0.2, 0.3, 0.4, 0.5, 0.6]])


def guard_net(v):
    z = []
    for j in range(N):
        s = v[0] * float(W[j, 0])
        for i in range(1, N):
            s = s + v[i] * float(W[j, i])
        z.append(s)
    return (z[0] * 0.0 + z[1].sqrt() + z[2].ln() + z[3].tanh() * 1.5 + z[4].max(0.0) * 2.0
            + z[5] * z[5])


def tape():
    return record(guard_net, [0.1] * N)[0]


def points(B: int, seed: int = 0) -> np.ndarray:
    """Random rows in [-1, 1], with all-zero rows (every member exactly 0) and
    rows on the kinks of sqrt / ln / max mixed in."""
    x = np.random.default_rng(seed).uniform(-1, 1, (B, N))
    x[1::2] = np.abs(x[1::2])  # mostly positive members: finite logs
    x[::7] = 0.0
    return x
