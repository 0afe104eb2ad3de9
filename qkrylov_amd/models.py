"""Standard lattice OpSums as ordered term lists [(coeff, [(op, site), ...])].

Factor order follows the reference: the first listed factor acts first on
the ket (src/hamiltonian/matrix_free_hamiltonian.cpp:54-68), so the operator
product c_i^dagger c_j is written [("CUp", j), ("CdagUp", i)].
Spin-1/2 bonds are {J: Sz Sz}, {J/2: Sp Sm}, {J/2: Sm Sp} as in the
reference's tests (tests/test_workflow.cpp:33-35)."""
import cmath
import math


def heisenberg_bonds(bonds, J=1.0, delta=1.0):
    terms = []
    for i, j, *w in bonds:
        Jb = J * (w[0] if w else 1.0)
        terms.append((Jb * delta, [("Sz", i), ("Sz", j)]))
        terms.append((0.5 * Jb, [("Sp", i), ("Sm", j)]))
        terms.append((0.5 * Jb, [("Sm", i), ("Sp", j)]))
    return terms


def chain_bonds(L, periodic=True):
    bonds = [(i, i + 1) for i in range(L - 1)]
    if periodic and L > 2:
        bonds.append((L - 1, 0))
    return bonds


def heisenberg_chain(L, J=1.0, periodic=True):
    return heisenberg_bonds(chain_bonds(L, periodic), J)


def xxz_chain(L, delta, J=1.0, periodic=True):
    return heisenberg_bonds(chain_bonds(L, periodic), J, delta)


def square_bonds(rows, cols, periodic=True, diagonal=False):
    """Nearest-neighbour (or diagonal next-nearest) bonds of a rows x cols
    lattice with row-major sites s = r * cols + c."""
    bonds = []
    seen = set()

    def add(a, b):
        key = (min(a, b), max(a, b))
        if a != b and key not in seen:
            seen.add(key)
            bonds.append((a, b))

    for r in range(rows):
        for c in range(cols):
            s = r * cols + c
            if not diagonal:
                if c + 1 < cols or periodic:
                    add(s, r * cols + (c + 1) % cols)
                if r + 1 < rows or periodic:
                    add(s, ((r + 1) % rows) * cols + c)
            else:
                if (r + 1 < rows or periodic) and (c + 1 < cols or periodic):
                    add(s, ((r + 1) % rows) * cols + (c + 1) % cols)
                if (r + 1 < rows or periodic) and (c - 1 >= 0 or periodic):
                    add(s, ((r + 1) % rows) * cols + (c - 1) % cols)
    return bonds


def j1j2_square(rows, cols, J1=1.0, J2=0.5, periodic=True):
    return (heisenberg_bonds(square_bonds(rows, cols, periodic), J1) +
            heisenberg_bonds(square_bonds(rows, cols, periodic, diagonal=True), J2))


def hopping_terms(bonds, t):
    terms = []
    for i, j in bonds:
        for up, dn in (("CUp", "CdagUp"), ("CDn", "CdagDn")):
            terms.append((-t, [(up, j), (dn, i)]))   # -t c_i^dag c_j
            terms.append((-t, [(up, i), (dn, j)]))   # -t c_j^dag c_i
    return terms


def hubbard_chain(L, t=1.0, U=4.0, periodic=True):
    terms = hopping_terms(chain_bonds(L, periodic), t)
    terms += [(U, [("Nupdn", i)]) for i in range(L)]
    return terms


def tj_model(rows, cols, t=1.0, J=0.4, periodic=False):
    """t-J: projected hopping, J/2 (S+_i S-_j + S-_i S+_j) written as
    four-fermion strings, and J (Sz_i Sz_j - n_i n_j / 4) =
    -J/2 (Nup_i Ndn_j + Ndn_i Nup_j)."""
    bonds = square_bonds(rows, cols, periodic)
    terms = hopping_terms(bonds, t)
    for i, j in bonds:
        # S+_i S-_j = c_i_up^dag c_i_dn c_j_dn^dag c_j_up
        terms.append((0.5 * J, [("CUp", j), ("CdagDn", j), ("CDn", i), ("CdagUp", i)]))
        terms.append((0.5 * J, [("CUp", i), ("CdagDn", i), ("CDn", j), ("CdagUp", j)]))
        terms.append((-0.5 * J, [("Nup", i), ("Ndn", j)]))
        terms.append((-0.5 * J, [("Ndn", i), ("Nup", j)]))
    return terms


def szq_terms(rows, cols, q, two_species=True):
    """S^z_q = N^-1/2 sum_r exp(i q.r) Sz_r (r = (x=c, y=r)); for two-species
    sites Sz_r = (Nup_r - Ndn_r) / 2."""
    n = rows * cols
    terms = []
    for r in range(rows):
        for c in range(cols):
            s = r * cols + c
            ph = cmath.exp(1j * (q[0] * c + q[1] * r)) / math.sqrt(n)
            if two_species:
                terms.append((0.5 * ph, [("Nup", s)]))
                terms.append((-0.5 * ph, [("Ndn", s)]))
            else:
                terms.append((ph, [("Sz", s)]))
    return terms


def lattice_momenta(rows, cols):
    return [(2 * math.pi * kx / cols, 2 * math.pi * ky / rows) for ky in range(rows) for kx in range(cols)]
