"""ctypes access to the oracle (oracle/liboracle.so) and, where it was built,
the reference-compiled checker (oracle/_ref/libqkref.so).

Test infrastructure only: the product package never imports this module.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libqkref.so")

SPIN, FERMION, HUBBARD, TJ = 0, 1, 2, 3

_c_i64 = ctypes.c_int64
_c_u64 = ctypes.c_uint64
_p_dbl = ctypes.POINTER(ctypes.c_double)
_p_u64 = ctypes.POINTER(ctypes.c_uint64)
_p_i64 = ctypes.POINTER(ctypes.c_int64)
_p_int = ctypes.POINTER(ctypes.c_int)


class QoSector(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in
                ("use_sz", "sz2", "use_nup", "nup", "use_ndn", "ndn", "use_n", "n", "use_nb", "nb")]


class QoLanczosResult(ctypes.Structure):
    _fields_ = [("energy", ctypes.c_double), ("iterations", ctypes.c_int),
                ("converged", ctypes.c_int), ("n_alpha", ctypes.c_int), ("n_beta", ctypes.c_int)]


def _dp(a):
    return a.ctypes.data_as(_p_dbl)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = ctypes.CDLL(ORACLE_SO)
        L.qo_site_apply.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int, _c_u64, _p_u64, _p_dbl, _p_dbl]
        L.qo_basis_enumerate.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(QoSector), _p_u64, _c_i64]
        L.qo_basis_enumerate.restype = _c_i64
        for f in ("qo_count_dimension",):
            getattr(L, f).argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(QoSector)]
            getattr(L, f).restype = _c_i64
        L.qo_count_contains.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(QoSector), _c_u64]
        L.qo_count_rank.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(QoSector), _c_u64]
        L.qo_count_rank.restype = _c_i64
        L.qo_count_unrank.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(QoSector), _c_i64]
        L.qo_count_unrank.restype = _c_u64
        L.qo_model_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(QoSector), ctypes.c_int, ctypes.c_int]
        L.qo_model_create.restype = ctypes.c_void_p
        L.qo_model_destroy.argtypes = [ctypes.c_void_p]
        L.qo_model_add_term.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_char_p), _p_int]
        L.qo_model_dim.argtypes = [ctypes.c_void_p]
        L.qo_model_dim.restype = _c_i64
        L.qo_model_state.argtypes = [ctypes.c_void_p, _c_i64]
        L.qo_model_state.restype = _c_u64
        L.qo_model_index.argtypes = [ctypes.c_void_p, _c_u64]
        L.qo_model_index.restype = _c_i64
        L.qo_build_csr.argtypes = [ctypes.c_void_p]
        L.qo_build_csr.restype = _c_i64
        L.qo_build_csr_rows.argtypes = [ctypes.c_void_p, _c_i64, _c_i64]
        L.qo_build_csr_rows.restype = _c_i64
        L.qo_csr_nnz.argtypes = [ctypes.c_void_p]
        L.qo_csr_nnz.restype = _c_i64
        L.qo_csr_export.argtypes = [ctypes.c_void_p, _p_i64, _p_i64, _p_dbl, _p_dbl]
        L.qo_row_entries.argtypes = [ctypes.c_void_p, _c_i64, _p_i64, _p_dbl, ctypes.c_int]
        L.qo_apply.argtypes = [ctypes.c_void_p, _p_dbl, _p_dbl]
        L.qo_apply_rows.argtypes = [ctypes.c_void_p, _p_dbl, _p_dbl, _c_i64, _c_i64]
        L.qo_lanczos.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, _p_dbl, ctypes.c_int,
                                 ctypes.POINTER(QoLanczosResult), _p_dbl, _p_dbl, _p_dbl]
        L.qo_tridiag_ground.argtypes = [_p_dbl, _p_dbl, ctypes.c_int, _p_dbl]
        L.qo_tridiag_ground.restype = ctypes.c_double
        L.qo_tridiag_components.argtypes = [_p_dbl, _p_dbl, ctypes.c_int, _p_dbl, _p_dbl]
        L.qo_compute_tridiag.argtypes = [ctypes.c_void_p, _p_dbl, ctypes.c_int, _p_dbl, _p_dbl, _p_int]
        L.qo_ftlm.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int, _p_dbl,
                              _p_dbl, _p_dbl, _p_dbl]
        L.qo_cf_coeffs.argtypes = [ctypes.c_void_p, _p_dbl, ctypes.c_int, _p_dbl, _p_dbl, _p_int, _p_int, _p_dbl]
        L.qo_spectral.argtypes = [_p_dbl, _p_dbl, ctypes.c_size_t, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_double]
        L.qo_spectral.restype = ctypes.c_double
        L.qo_davidson.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double, _p_dbl, _p_dbl,
                                  _p_int, _p_int]
        L.qo_rng_u64.argtypes = [_c_u64, _c_u64]
        L.qo_rng_u64.restype = _c_u64
        L.qo_rng_uniform.argtypes = [_c_u64, _c_u64]
        L.qo_rng_uniform.restype = ctypes.c_double
        _lib = L
    return _lib


def sector(use_sz=0, sz2=0, use_nup=0, nup=0, use_ndn=0, ndn=0, use_n=0, n=0):
    return QoSector(use_sz, sz2, use_nup, nup, use_ndn, ndn, use_n, n, 0, 0)


def site_apply(kind, op, site, state):
    ns = _c_u64(0)
    re = ctypes.c_double(0)
    im = ctypes.c_double(0)
    v = lib().qo_site_apply(kind, op.encode(), site, state, ctypes.byref(ns), ctypes.byref(re), ctypes.byref(im))
    return v, ns.value, complex(re.value, im.value)


def enumerate_basis(kind, n, sec):
    L = lib()
    d = L.qo_basis_enumerate(kind, n, ctypes.byref(sec), None, 0)
    if d < 0:
        return None
    out = np.zeros(max(d, 1), dtype=np.uint64)
    L.qo_basis_enumerate(kind, n, ctypes.byref(sec), out.ctypes.data_as(_p_u64), d)
    return out[:d]


class Model:
    """Oracle model: basis + site + opsum; terms are (coeff, [(op, site), ...])."""

    def __init__(self, bkind, n, sec, skind=None, terms=(), enumerate=True):
        self.L = lib()
        self.sec = sec
        self.h = self.L.qo_model_create(bkind, n, ctypes.byref(sec), bkind if skind is None else skind,
                                        1 if enumerate else 0)
        if not self.h:
            raise ValueError("oracle basis creation failed")
        self.dim = self.L.qo_model_dim(self.h)
        for c, fs in terms:
            self.add_term(c, fs)
        self._csr = False

    def __del__(self):
        if getattr(self, "h", None):
            self.L.qo_model_destroy(self.h)
            self.h = None

    def add_term(self, c, factors):
        c = complex(c)
        ops = (ctypes.c_char_p * len(factors))(*[f[0].encode() for f in factors])
        sites = (ctypes.c_int * len(factors))(*[f[1] for f in factors])
        self.L.qo_model_add_term(self.h, c.real, c.imag, len(factors), ops, sites)

    def state(self, i):
        return self.L.qo_model_state(self.h, i)

    def index(self, s):
        return self.L.qo_model_index(self.h, s)

    def build_csr(self):
        nnz = self.L.qo_build_csr(self.h)
        if nnz < 0:
            raise ValueError("unknown operator")
        self._csr = True
        return nnz

    def csr(self):
        if not self._csr:
            self.build_csr()
        nnz = self.L.qo_csr_nnz(self.h)
        ro = np.zeros(self.dim + 1, dtype=np.int64)
        cols = np.zeros(max(nnz, 1), dtype=np.int64)
        vals = np.zeros(2 * max(nnz, 1), dtype=np.float64)
        diag = np.zeros(2 * max(self.dim, 1), dtype=np.float64)
        self.L.qo_csr_export(self.h, ro.ctypes.data_as(_p_i64), cols.ctypes.data_as(_p_i64), _dp(vals), _dp(diag))
        return ro, cols[:nnz], vals[:2 * nnz].view(np.complex128), diag[:2 * self.dim].view(np.complex128)

    def dense(self):
        """Dense matrix A with A[row, col] = stored value (so y = A @ x)."""
        ro, cols, vals, _ = self.csr()
        A = np.zeros((self.dim, self.dim), dtype=np.complex128)
        for r in range(self.dim):
            for j in range(ro[r], ro[r + 1]):
                A[r, cols[j]] += vals[j]
        return A

    def row_entries(self, row, cap=4096):
        cols = np.zeros(cap, dtype=np.int64)
        vals = np.zeros(2 * cap, dtype=np.float64)
        c = self.L.qo_row_entries(self.h, row, cols.ctypes.data_as(_p_i64), _dp(vals), cap)
        if c < 0:
            raise ValueError("unknown operator")
        return cols[:c], vals[:2 * c].view(np.complex128)

    def apply(self, x):
        if not self._csr:
            self.build_csr()
        x = np.ascontiguousarray(x, dtype=np.complex128)
        y = np.zeros(self.dim, dtype=np.complex128)
        self.L.qo_apply(self.h, _dp(x.view(np.float64)), _dp(y.view(np.float64)))
        return y

    def lanczos(self, maxiter=200, tol=1e-12, start=None, reorth=True, want_vec=False):
        if not self._csr:
            self.build_csr()
        if start is None:
            start = start_vector(self.dim, 1234)
        start = np.ascontiguousarray(start, dtype=np.complex128)
        res = QoLanczosResult()
        al = np.zeros(maxiter + 1)
        be = np.zeros(maxiter + 1)
        vec = np.zeros(self.dim, dtype=np.complex128) if want_vec else None
        self.L.qo_lanczos(self.h, maxiter, tol, _dp(start.view(np.float64)), 1 if reorth else 0,
                          ctypes.byref(res), _dp(al), _dp(be),
                          _dp(vec.view(np.float64)) if want_vec else None)
        return dict(energy=res.energy, iterations=res.iterations, converged=bool(res.converged),
                    alphas=al[:res.n_alpha].copy(), betas=be[:res.n_beta].copy(), eigenvector=vec)

    def compute_tridiag(self, v, n_steps):
        if not self._csr:
            self.build_csr()
        v = np.ascontiguousarray(v, dtype=np.complex128)
        al = np.zeros(n_steps + 1)
        be = np.zeros(n_steps + 1)
        nb = ctypes.c_int(0)
        na = self.L.qo_compute_tridiag(self.h, _dp(v.view(np.float64)), n_steps, _dp(al), _dp(be), ctypes.byref(nb))
        return al[:na].copy(), be[:nb.value].copy()

    def ftlm(self, beta, rvecs, n_steps):
        if not self._csr:
            self.build_csr()
        rv = np.ascontiguousarray(rvecs, dtype=np.complex128)
        Z = ctypes.c_double()
        E = ctypes.c_double()
        Cv = ctypes.c_double()
        self.L.qo_ftlm(self.h, beta, rv.shape[0], n_steps, _dp(rv.view(np.float64)),
                       ctypes.byref(Z), ctypes.byref(E), ctypes.byref(Cv))
        return Z.value, E.value, Cv.value

    def cf(self, phi0, n_iter):
        if not self._csr:
            self.build_csr()
        phi = np.ascontiguousarray(phi0, dtype=np.complex128)
        al = np.zeros(n_iter + 1)
        be = np.zeros(n_iter + 1)
        na = ctypes.c_int()
        nb = ctypes.c_int()
        nrm = ctypes.c_double()
        self.L.qo_cf_coeffs(self.h, _dp(phi.view(np.float64)), n_iter, _dp(al), _dp(be),
                            ctypes.byref(na), ctypes.byref(nb), ctypes.byref(nrm))
        return al[:na.value].copy(), be[:nb.value].copy(), nrm.value

    def davidson(self, n_eig=1, max_subspace=20, tol=1e-8, want_vecs=False):
        if not self._csr:
            self.build_csr()
        ev = np.zeros(n_eig)
        vecs = np.zeros((n_eig, self.dim), dtype=np.complex128) if want_vecs else None
        it = ctypes.c_int()
        cv = ctypes.c_int()
        self.L.qo_davidson(self.h, n_eig, max_subspace, tol, _dp(ev),
                           _dp(vecs.view(np.float64)) if want_vecs else None, ctypes.byref(it), ctypes.byref(cv))
        return dict(eigenvalues=ev, eigenvectors=vecs, iterations=it.value, converged=bool(cv.value))


def spectral(alphas, betas, norm_phi0, omega, E0, eta):
    a = np.ascontiguousarray(alphas, dtype=np.float64)
    b = np.ascontiguousarray(betas if len(betas) else [0.0], dtype=np.float64)
    return lib().qo_spectral(_dp(a), _dp(b), len(a), norm_phi0, omega, E0, eta)


def tridiag_ground(alphas, betas):
    a = np.ascontiguousarray(alphas, dtype=np.float64)
    b = np.ascontiguousarray(betas if len(betas) else [0.0], dtype=np.float64)
    v = np.zeros(len(a))
    e = lib().qo_tridiag_ground(_dp(a), _dp(b), len(a), _dp(v))
    return e, v


def start_vector(dim, seed):
    """Real start vector uniform[-1, 1) from the shared counter RNG."""
    L = lib()
    return np.array([L.qo_rng_uniform(seed, i) for i in range(dim)], dtype=np.float64)


# ---------------------------------------------------------------- reference-compiled checker
_ref = None


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        R = ctypes.CDLL(REF_SO)
        R.qkref_basis_states.argtypes = [ctypes.c_int, ctypes.c_int, _p_int, _p_u64, _c_i64]
        R.qkref_basis_states.restype = _c_i64
        R.qkref_site_apply.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int, _c_u64, _p_u64, _p_dbl, _p_dbl]
        R.qkref_csr.argtypes = [ctypes.c_int, ctypes.c_int, _p_int, ctypes.c_int, ctypes.c_int, _p_dbl, _p_dbl,
                                _p_int, ctypes.POINTER(ctypes.c_char_p), _p_int, _p_i64, _p_i64, _p_dbl, _p_dbl]
        R.qkref_csr.restype = _c_i64
        _ref = R
    return _ref


def _sec_arr(sec):
    return (ctypes.c_int * 8)(sec.use_sz, sec.sz2, sec.use_nup, sec.nup, sec.use_ndn, sec.ndn, sec.use_n, sec.n)


def ref_basis_states(kind, n, sec):
    R = ref()
    d = R.qkref_basis_states(kind, n, _sec_arr(sec), None, 0)
    if d < 0:
        return None
    out = np.zeros(max(d, 1), dtype=np.uint64)
    R.qkref_basis_states(kind, n, _sec_arr(sec), out.ctypes.data_as(_p_u64), d)
    return out[:d]


def ref_site_apply(kind, op, site, state):
    ns = _c_u64(0)
    re = ctypes.c_double(0)
    im = ctypes.c_double(0)
    v = ref().qkref_site_apply(kind, op.encode(), site, state, ctypes.byref(ns), ctypes.byref(re), ctypes.byref(im))
    return v, ns.value, complex(re.value, im.value)


def ref_csr(bkind, n, sec, skind, terms):
    R = ref()
    nt = len(terms)
    cre = np.array([complex(c).real for c, _ in terms] or [0.0])
    cim = np.array([complex(c).imag for c, _ in terms] or [0.0])
    nf = (ctypes.c_int * max(nt, 1))(*[len(f) for _, f in terms])
    flat = [f for _, fs in terms for f in fs]
    ops = (ctypes.c_char_p * max(len(flat), 1))(*[f[0].encode() for f in flat])
    sites = (ctypes.c_int * max(len(flat), 1))(*[f[1] for f in flat])
    sa = _sec_arr(sec)
    nnz = R.qkref_csr(bkind, n, sa, skind, nt, _dp(cre), _dp(cim), nf, ops, sites, None, None, None, None)
    if nnz < 0:
        return None
    dim = len(ref_basis_states(bkind, n, sec))
    ro = np.zeros(dim + 1, dtype=np.int64)
    cols = np.zeros(max(nnz, 1), dtype=np.int64)
    vals = np.zeros(2 * max(nnz, 1))
    diag = np.zeros(2 * max(dim, 1))
    R.qkref_csr(bkind, n, sa, skind, nt, _dp(cre), _dp(cim), nf, ops, sites,
                ro.ctypes.data_as(_p_i64), cols.ctypes.data_as(_p_i64), _dp(vals), _dp(diag))
    return ro, cols[:nnz], vals[:2 * nnz].view(np.complex128), diag[:2 * dim].view(np.complex128)
