"""Valid PLONK instances for the quotient tests: a textbook PLONK circuit
(multiplication / addition gates, one public input row, copy constraints
c_(i-1) = a_i and b_i = b_0), its permutation polynomials, a witness and
the grand product z.  A quotient computed from these inputs passes the
degree check only if the gate, permutation and boundary terms are all
right, for random alpha, beta, gamma."""
import random

from oracle.field import R_MOD, fr_inv
from oracle.ntt import coeffs_from_evals_on_roots, omega_for_size

K1, K2 = 7, 13  # coset representatives: H, 7H, 13H are disjoint for n <= 2^20


def valid_instance(n: int, seed: int, intt=None):
    """intt(evals, omega) -> coefficients; the Python oracle's by default."""
    rng = random.Random(seed)
    omega = omega_for_size(n)
    ql, qr, qm, qo, qc, pi = ([0] * n for _ in range(6))
    a, b, c = [0] * n, [0] * n, [0] * n
    b0 = rng.randrange(1, R_MOD)
    for i in range(n):
        a[i] = c[i - 1] if i else rng.randrange(1, R_MOD)
        b[i] = b0 if i % 3 == 0 else rng.randrange(1, R_MOD)
        kind = rng.randrange(3) if i else 0
        if i == n - 1:
            kind = 3
        if kind == 0:      # multiplication: a b - c = 0
            qm[i], qo[i] = 1, R_MOD - 1
            c[i] = a[i] * b[i] % R_MOD
        elif kind == 1:    # addition with constant: a + b + k - c = 0
            k = rng.randrange(R_MOD)
            ql[i], qr[i], qo[i], qc[i] = 1, 1, R_MOD - 1, k
            c[i] = (a[i] + b[i] + k) % R_MOD
        elif kind == 2:    # scaled addition: 3a - b - c = 0
            ql[i], qr[i], qo[i] = 3, R_MOD - 1, R_MOD - 1
            c[i] = (3 * a[i] - b[i]) % R_MOD
        else:              # public input: a - PI = 0 (pi = -value)
            ql[i] = 1
            pi[i] = (R_MOD - a[i]) % R_MOD
            c[i] = rng.randrange(R_MOD)
    # copy cycles over positions (column, row): (2, i-1) <-> (0, i), and all (1, 3j) together
    pos = [(col, i) for col in range(3) for i in range(n)]
    nxt = {p: p for p in pos}
    for i in range(1, n):
        nxt[(2, i - 1)], nxt[(0, i)] = (0, i), (2, i - 1)
    group = [(1, i) for i in range(0, n, 3)]
    for j, p in enumerate(group):
        nxt[p] = group[(j + 1) % len(group)]
    ks = (1, K1, K2)
    label = {(col, i): ks[col] * pow(omega, i, R_MOD) % R_MOD for (col, i) in pos}
    sig = [[label[nxt[(col, i)]] for i in range(n)] for col in range(3)]
    beta, gamma = rng.randrange(1, R_MOD), rng.randrange(1, R_MOD)
    w = (a, b, c)
    z = [1] * n
    for i in range(n - 1):
        num = den = 1
        for col in range(3):
            num = num * (w[col][i] + beta * label[(col, i)] + gamma) % R_MOD
            den = den * (w[col][i] + beta * sig[col][i] + gamma) % R_MOD
        z[i + 1] = z[i] * num % R_MOD * fr_inv(den) % R_MOD
    evals = {"a": a, "b": b, "c": c, "z": z, "pi": pi, "ql": ql, "qr": qr, "qm": qm, "qo": qo, "qc": qc,
             "s1": sig[0], "s2": sig[1], "s3": sig[2]}
    intt = intt or coeffs_from_evals_on_roots
    coeffs = {k: intt(v, omega) for k, v in evals.items()}
    alpha = rng.randrange(1, R_MOD)
    return {"coeffs": coeffs, "evals": evals, "n": n, "omega": omega, "k1": K1, "k2": K2, "alpha": alpha,
            "beta": beta, "gamma": gamma}


def shift_for(n: int, seed: int) -> int:
    rng = random.Random(seed)
    u = pow(omega_for_size(4 * n), n, R_MOD)
    bad = {1, u, u * u % R_MOD, u * u % R_MOD * u % R_MOD}
    while True:
        s = rng.randrange(1, R_MOD)
        if pow(s, n, R_MOD) not in bad:
            return s
