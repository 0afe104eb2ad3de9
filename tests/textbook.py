"""Independent textbook constructions used to pin the oracle and the engine:
dense operators from the definitions (Kronecker / Jordan-Wigner on the bit
conventions of the reference: spin up = bit 1; Hubbard up at bit 2i, down
at 2i+1), the Bethe-ansatz ground states of the XXX and XXZ rings, the
Lieb-Wu ground state of the Hubbard ring, and exact thermal and resolvent
quantities by dense diagonalisation."""
import math

import numpy as np

SPIN, FERMION, HUBBARD, TJ = 0, 1, 2, 3


def _bit(s, b):
    return (s >> b) & 1


def local_matrix(kind, op, site, nbits):
    """Matrix F of one factor in the 2^nbits computational basis (F[t, s] =
    <t|op|s>), from the operator definitions. The reference's Sy carries
    the opposite sign of the textbook S^y (spinhalf_site.cpp:104-113)."""
    dim = 1 << nbits
    F = np.zeros((dim, dim), dtype=np.complex128)
    for s in range(dim):
        if kind == SPIN:
            b = site
            up = _bit(s, b)
            if op == "Sz":
                F[s, s] = 0.5 if up else -0.5
            elif op == "Sp" and not up:
                F[s | (1 << b), s] = 1.0
            elif op == "Sm" and up:
                F[s & ~(1 << b), s] = 1.0
            elif op == "Sx":
                F[s ^ (1 << b), s] = 0.5
            elif op == "Sy":
                # textbook: S^y|up> = +i/2 |dn>, S^y|dn> = -i/2 |up>; reference: negated
                F[s ^ (1 << b), s] = -(0.5j if up else -0.5j)
            continue
        if kind == FERMION:
            bits = [site]
        else:
            bits = [2 * site, 2 * site + 1]
        if op in ("N", "Nup", "Ndn", "Nupdn", "Id"):
            if op == "Id":
                F[s, s] = 1.0
            elif op == "N":
                F[s, s] = _bit(s, site)
            elif op == "Nup":
                F[s, s] = _bit(s, 2 * site)
            elif op == "Ndn":
                F[s, s] = _bit(s, 2 * site + 1)
            else:
                F[s, s] = _bit(s, 2 * site) * _bit(s, 2 * site + 1)
            continue
        create = op in ("Cdag", "CdagUp", "CdagDn")
        b = {"C": site, "Cdag": site, "CUp": 2 * site, "CdagUp": 2 * site, "CDn": 2 * site + 1,
             "CdagDn": 2 * site + 1}[op]
        occ = _bit(s, b)
        if create == bool(occ):
            continue
        if kind == TJ and create:
            partner = b ^ 1
            if _bit(s, partner):
                continue
        sign = (-1) ** bin(s & ((1 << b) - 1)).count("1")   # Jordan-Wigner string over lower bits
        F[s ^ (1 << b), s] = sign
    return F


def term_matrix(kind, coeff, factors, nbits):
    """coeff * F_{n-1} ... F_1 F_0 (factor 0 acts first)."""
    dim = 1 << nbits
    M = np.eye(dim, dtype=np.complex128) * coeff
    for op, site in factors:
        M = local_matrix(kind, op, site, nbits) @ M
    return M


def opsum_matrix(kind, terms, nbits):
    dim = 1 << nbits
    H = np.zeros((dim, dim), dtype=np.complex128)
    for c, fs in terms:
        H += term_matrix(kind, complex(c), fs, nbits)
    return H


def sector_states(kind, n, use_sz=0, sz2=0, use_n=0, nn=0, use_nup=0, nup=0, use_ndn=0, ndn=0):
    nbits = n if kind in (SPIN, FERMION) else 2 * n
    out = []
    for s in range(1 << nbits):
        if kind == SPIN:
            if use_sz and 2 * bin(s).count("1") - n != sz2:
                continue
        elif kind == FERMION:
            if use_n and bin(s).count("1") != nn:
                continue
        else:
            ups = sum(_bit(s, 2 * i) for i in range(n))
            dns = sum(_bit(s, 2 * i + 1) for i in range(n))
            if kind == TJ and any(_bit(s, 2 * i) and _bit(s, 2 * i + 1) for i in range(n)):
                continue
            if use_nup and ups != nup:
                continue
            if use_ndn and dns != ndn:
                continue
        out.append(s)
    return np.array(out, dtype=np.int64)


def stored_matrix(kind, terms, nbits, states, site_kind=None):
    """The reference's stored matrix on the sector: row a holds
    conj(<b|O|a>), i.e. (O^dagger)[a, b] restricted to the sector."""
    H = opsum_matrix(site_kind if site_kind is not None else kind, terms, nbits)
    return H.conj().T[np.ix_(states, states)]


def bethe_xxx_ring(N):
    """Ground-state energy of the periodic spin-1/2 XXX ring
    H = sum_i S_i . S_{i+1} (N even) from the Bethe equations in
    logarithmic form, N atan(2 l_j) = pi I_j + sum_k atan(l_j - l_k),
    E = N/4 - sum_j 2 / (1 + 4 l_j^2)."""
    M = N // 2
    I = np.array([-(M - 1) / 2 + j for j in range(M)])
    lam = 0.5 * np.tan(np.pi * I / N)
    for _ in range(200000):
        diff = lam[:, None] - lam[None, :]
        new = 0.5 * np.tan((np.pi * I + np.arctan(diff).sum(axis=1)) / N)
        if np.max(np.abs(new - lam)) < 1e-15:
            lam = new
            break
        lam = 0.5 * lam + 0.5 * new
    return N / 4.0 - np.sum(2.0 / (1.0 + 4.0 * lam ** 2))


def bethe_xxz_ring(L, delta, sz2=0):
    """Ground-state energy of the periodic XXZ ring
    H = sum_i [Sx_i Sx_{i+1} + Sy_i Sy_{i+1} + delta Sz_i Sz_{i+1}] for
    0 < delta < 1 (delta = cos gamma), L even, in the sector 2 Sz = sz2:
    M = L/2 - |Sz| down spins with rapidities lam_j solving
    L theta1(lam_j) = 2 pi I_j + sum_k theta2(lam_j - lam_k),
    theta_n(x) = 2 atan(cot(n gamma / 2) tanh x), I_j = -(M-1)/2 + j;
    E = L delta / 4 - sum_j sin^2 gamma / (cosh 2 lam_j - cos gamma)."""
    if not 0.0 < delta < 1.0 or L % 2:
        raise ValueError("critical regime 0 < delta < 1 and even L only")
    M = L // 2 - abs(sz2) // 2
    g = math.acos(delta)
    I = np.array([-(M - 1) / 2 + j for j in range(M)])
    t1, cg = math.tan(g / 2), 1.0 / math.tan(g)
    lam = np.arctanh(np.tan(np.pi * I / L) * t1)
    for _ in range(200000):
        d = lam[:, None] - lam[None, :]
        x = (2 * np.pi * I + 2 * np.arctan(cg * np.tanh(d)).sum(axis=1)) / L
        new = np.arctanh(np.tan(x / 2) * t1)
        if np.max(np.abs(new - lam)) < 1e-15:
            lam = new
            break
        lam = 0.5 * lam + 0.5 * new
    return L * delta / 4 - np.sum(math.sin(g) ** 2 / (np.cosh(2 * lam) - math.cos(g)))


def lieb_wu_ring(L, nup, ndn, U, t=1.0):
    """Ground-state energy of the periodic Hubbard ring
    H = -t sum (c+_i c_j + h.c.) + U sum n_up n_dn with N = nup + ndn
    electrons, M = min(nup, ndn), from the Lieb-Wu equations (u = U / 4t)
    L k_j = 2 pi I_j - sum_a 2 atan((sin k_j - L_a) / u),
    sum_j 2 atan((L_a - sin k_j) / u) = 2 pi J_a + sum_b 2 atan((L_a - L_b) / 2u),
    with consecutive I_j (integer iff M even) and J_a (integer iff N - M
    odd) centred on zero; E = -2 t sum_j cos k_j. Used at half filling,
    where this is the ground state."""
    from scipy.optimize import root
    N, M = nup + ndn, min(nup, ndn)
    u = U / (4.0 * t)
    I = np.array([-(N - 1) / 2 + j for j in range(N)])
    if (I[0] % 1 == 0) != (M % 2 == 0):
        I = I + 0.5
    J = np.array([-(M - 1) / 2 + a for a in range(M)])
    if (J[0] % 1 == 0) != ((N - M) % 2 == 1):
        J = J + 0.5

    def F(x):
        k, lam = x[:N], x[N:]
        f1 = L * k - 2 * np.pi * I + 2 * np.arctan((np.sin(k)[:, None] - lam[None, :]) / u).sum(axis=1)
        f2 = (2 * np.arctan((lam[:, None] - np.sin(k)[None, :]) / u).sum(axis=1) - 2 * np.pi * J
              - 2 * np.arctan((lam[:, None] - lam[None, :]) / (2 * u)).sum(axis=1))
        return np.concatenate([f1, f2])

    sol = root(F, np.concatenate([2 * np.pi * I / L, 0.3 * J]), method="hybr", tol=1e-15)
    if np.max(np.abs(F(sol.x))) > 1e-11:
        raise RuntimeError("Lieb-Wu equations did not converge")
    return -2.0 * t * np.sum(np.cos(sol.x[:N]))


def thermal(evals, betas):
    """Exact Z, E, Cv for the spectrum."""
    evals = np.asarray(evals)
    out = []
    for b in betas:
        w = np.exp(-b * evals)
        Z = w.sum()
        E = (evals * w).sum() / Z
        E2 = (evals ** 2 * w).sum() / Z
        out.append((Z, E, b * b * (E2 - E * E)))
    return out


def resolvent_spectrum(H, phi, omega, E0, eta):
    """-1/pi Im <phi|(z - H)^-1|phi>, z = omega + E0 + i eta (H Hermitian)."""
    w, V = np.linalg.eigh(H)
    c = V.conj().T @ phi
    z = omega + E0 + 1j * eta
    return -1.0 / math.pi * np.imag(np.sum(np.abs(c) ** 2 / (z - w)))
