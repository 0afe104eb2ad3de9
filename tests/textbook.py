"""Independent textbook constructions used to pin the oracle and the engine:
dense operators from the definitions (Kronecker / Jordan-Wigner on the bit
conventions of the reference: spin up = bit 1; Hubbard up at bit 2i, down
at 2i+1), the Bethe-ansatz ground state of the XXX ring, and exact thermal
and resolvent quantities by dense diagonalisation."""
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
