"""Python mirror of the reference's Python package (bindings/python/qkrylov:
basis.py, site.py, operators.py, hamiltonian.py, solvers.py) over the
engine's C ABI. The engine computes in float64 whatever `dtype` says;
`dtype=np.float32` (the reference's default there) only makes the returned
vectors complex64 as the reference's are. The default here is float64.
Device memory and streams are the engine's own; no torch."""
import ctypes
import enum

import numpy as np

from . import _lib
from ._lib import check


# ---------------------------------------------------------------- operators
class Op(str, enum.Enum):
    """Local operator names (bindings/python/qkrylov/operators.py:6-22)."""
    Sz = "Sz"
    Sp = "Sp"
    Sm = "Sm"
    Sx = "Sx"
    Sy = "Sy"
    CdagUp = "CdagUp"
    CUp = "CUp"
    CdagDn = "CdagDn"
    CDn = "CDn"
    Nup = "Nup"
    Ndn = "Ndn"
    Nupdn = "Nupdn"
    Bdag = "Bdag"
    B = "B"
    N = "N"


def _op_name(o):
    return o.value if isinstance(o, Op) else str(o)


def _vec_dtype(dtype):
    """Complex dtype of returned vectors: complex64 for float32 handles."""
    return np.complex64 if np.dtype(dtype) == np.float32 else np.complex128


_SCALARS = (int, float, complex, np.number)


def _expand(x):
    """Operator expression or scalar -> tuple of (coefficient, factors)."""
    if isinstance(x, OpProduct):
        return ((x.coeff, x.factors),)
    if isinstance(x, OpTermSum):
        return x.terms
    if isinstance(x, _SCALARS):
        return ((complex(x), ()),)
    return None


def _combine(terms, single):
    return OpProduct(*terms[0]) if single and len(terms) == 1 else OpTermSum(terms)


class OpProduct:
    """An ordered product coeff * f0 * f1 * ... of local operators, each
    factor a (name, site) pair; the first factor acts first, as in the
    reference's OperatorTerm. A local operator such as Sz(0) is the
    one-factor product with coefficient one. Products multiply by
    concatenating factors and multiplying coefficients; + and - make an
    OpTermSum (bindings/python/qkrylov/operators.py:25-105 define the same
    algebra)."""

    __slots__ = ("coeff", "factors")

    def __init__(self, coeff, factors):
        self.coeff = complex(coeff)
        self.factors = tuple((_op_name(n), int(i)) for n, i in factors)

    @property
    def name(self):
        return self.factors[0][0]

    @property
    def site(self):
        return self.factors[0][1]

    def __mul__(self, other):
        rhs = _expand(other)
        if rhs is None:
            return NotImplemented
        out = tuple((self.coeff * c, self.factors + f) for c, f in rhs)
        return _combine(out, not isinstance(other, OpTermSum))

    def __rmul__(self, other):
        if not isinstance(other, _SCALARS):
            return NotImplemented
        return OpProduct(complex(other) * self.coeff, self.factors)

    def __neg__(self):
        return OpProduct(-self.coeff, self.factors)

    def __add__(self, other):
        rhs = _expand(other)
        if rhs is None or isinstance(other, _SCALARS):
            return NotImplemented
        return OpTermSum(((self.coeff, self.factors),) + rhs)

    def __sub__(self, other):
        rhs = _expand(other)
        if rhs is None or isinstance(other, _SCALARS):
            return NotImplemented
        return OpTermSum(((self.coeff, self.factors),) + tuple((-c, f) for c, f in rhs))

    def __repr__(self):
        return f"OpProduct({self.coeff}, {list(self.factors)})"


class OpTermSum:
    """An ordered sum of OpProduct terms, kept as (coefficient, factors)
    pairs; terms are never merged or reordered."""

    __slots__ = ("terms",)

    def __init__(self, terms):
        self.terms = tuple((complex(c), tuple(f)) for c, f in terms)

    def __add__(self, other):
        rhs = _expand(other)
        if rhs is None or isinstance(other, _SCALARS):
            return NotImplemented
        return OpTermSum(self.terms + rhs)

    def __sub__(self, other):
        rhs = _expand(other)
        if rhs is None or isinstance(other, _SCALARS):
            return NotImplemented
        return OpTermSum(self.terms + tuple((-c, f) for c, f in rhs))

    def __neg__(self):
        return OpTermSum(tuple((-c, f) for c, f in self.terms))

    def __mul__(self, other):
        rhs = _expand(other)
        if rhs is None:
            return NotImplemented
        return OpTermSum(tuple((c1 * c2, f1 + f2) for c1, f1 in self.terms for c2, f2 in rhs))

    def __rmul__(self, other):
        if not isinstance(other, _SCALARS):
            return NotImplemented
        return OpTermSum(tuple((complex(other) * c, f) for c, f in self.terms))

    def __len__(self):
        return len(self.terms)

    def __repr__(self):
        return f"OpTermSum({len(self.terms)} terms)"


def _mk(name):
    def f(i):
        return OpProduct(1.0, ((name, i),))
    f.__name__ = name
    return f


Sz, Sp, Sm, Sx, Sy = (_mk(n) for n in ("Sz", "Sp", "Sm", "Sx", "Sy"))
CdagUp, CUp, CdagDn, CDn = (_mk(n) for n in ("CdagUp", "CUp", "CdagDn", "CDn"))
Nup, Ndn, Nupdn = (_mk(n) for n in ("Nup", "Ndn", "Nupdn"))
Bdag, B, N = (_mk(n) for n in ("Bdag", "B", "N"))
Cdag, C = _mk("Cdag"), _mk("C")


class OpSum:
    """Ordered list of terms (coefficient, [(op, site), ...]); the first
    factor acts first, as in the reference."""

    def __init__(self, dtype=np.float64, terms=None):
        if terms is None and not isinstance(dtype, (type, np.dtype, str)):
            dtype, terms = np.float64, dtype      # OpSum([(coeff, factors), ...])
        self.dtype = dtype
        self.terms = []
        for t in terms or []:
            self.add(*t)

    def add(self, coeff, factors):
        """Appends coeff * f0 * f1 * ... with factors [(name, site), ...]."""
        self.terms.append((complex(coeff), [(_op_name(o), int(s)) for o, s in factors]))
        return self

    def add_term(self, coeff, *ops):
        """add_term(coeff, name0, site0, name1, site1, ...) as in
        bindings/python/qkrylov/operators.py:151-161."""
        if len(ops) % 2 != 0:
            raise ValueError("Operators must be provided in (name, site) pairs.")
        self.add(coeff, [(ops[i], ops[i + 1]) for i in range(0, len(ops), 2)])

    def __iadd__(self, other):
        """Accepts (coeff, name0, site0, ...) tuples, algebraic terms and
        sums of them, and other OpSums (operators.py:163-179)."""
        if isinstance(other, tuple):
            if len(other) < 3 or len(other) % 2 == 0:
                raise ValueError("Term must be a tuple of (coeff, op1, site1, ...)")
            self.add_term(other[0], *other[1:])
            return self
        if isinstance(other, (OpProduct, OpTermSum)):
            for c, fs in _expand(other):
                self.add(c, fs)
            return self
        if isinstance(other, OpSum):
            self.terms += other.terms
            return self
        raise ValueError("Unsupported type for += on OpSum")

    def clear(self):
        self.terms = []

    @property
    def size(self):
        return len(self.terms)

    def __len__(self):
        return len(self.terms)

    def __repr__(self):
        return f"OpSum(terms={self.size})"

    def _handle(self):
        L = _lib.load()
        h = L.qkrylov_opsum_create()
        for c, fs in self.terms:
            ops = (ctypes.c_char_p * len(fs))(*[f[0].encode() for f in fs])
            sites = (ctypes.c_int * len(fs))(*[f[1] for f in fs])
            check(L.qkrylov_opsum_add_term_nbody_f64(h, c.real, c.imag, len(fs), ops, sites), "add_term")
        return h


# ---------------------------------------------------------------- sites
class Site:
    _create = None

    def __init__(self, dtype=np.float64):
        self.dtype = dtype
        self._h = getattr(_lib.load(), self._create)()

    def __repr__(self):
        return f"{type(self).__name__}()"

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.load().qkrylov_site_destroy(self._h)
            self._h = None


class SpinHalfSite(Site):
    _create = "qkrylov_spinhalf_site_create"


class FermionSite(Site):
    _create = "qkrylov_fermion_site_create"


class HubbardSite(Site):
    _create = "qkrylov_hubbard_site_create"


class TJSite(Site):
    _create = "qkrylov_tj_site_create"


# ---------------------------------------------------------------- bases
class Basis:
    _create = None
    site_class = None

    def _build(self, N, dtype, sz2=None, n=None, nup=None, ndn=None):
        self.dtype = dtype
        L = _lib.load()
        sec = L.qkrylov_sector_create()
        if sz2 is not None:
            L.qkrylov_sector_set_sz(sec, int(sz2))
        if n is not None:
            L.qkrylov_sector_set_n(sec, int(n))
        if nup is not None:
            L.qkrylov_amd_sector_set_species(sec, 0, int(nup))
        if ndn is not None:
            L.qkrylov_amd_sector_set_species(sec, 1, int(ndn))
        self._h = getattr(L, self._create)(int(N), sec)
        L.qkrylov_sector_destroy(sec)
        if not self._h:
            raise ValueError(f"basis creation failed: {_lib.last_error()}")

    def _sector_repr(self):
        return ""

    def __repr__(self):
        return f"{type(self).__name__}(N={self.nsites}, dim={self.size}{self._sector_repr()})"

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.load().qkrylov_basis_destroy(self._h)
            self._h = None

    @property
    def size(self):
        return int(_lib.load().qkrylov_basis_dimension(self._h))

    @property
    def nsites(self):
        return int(_lib.load().qkrylov_basis_nsites(self._h))

    def state(self, i):
        return int(_lib.load().qkrylov_basis_state(self._h, int(i)))

    def index(self, s):
        return int(_lib.load().qkrylov_basis_index(self._h, int(s)))

    def contains(self, s):
        return bool(_lib.load().qkrylov_basis_contains(self._h, int(s)))


class SpinHalfBasis(Basis):
    """Spin-1/2 basis; `sz` (or conserve_sz, which means sz = 0) fixes total
    Sz (bindings/python/qkrylov/basis.py:49-84). `sz2` = 2 Sz is accepted
    too."""
    _create = "qkrylov_spinhalf_basis_create"
    site_class = SpinHalfSite

    def __init__(self, N, conserve_sz=False, sz=None, dtype=np.float64, sz2=None):
        if sz2 is None and (sz is not None or conserve_sz):
            sz2 = int(2 * (sz or 0))
        self.sz2 = sz2
        self._build(N, dtype, sz2=sz2)

    def _sector_repr(self):
        return "" if self.sz2 is None else f", sz={self.sz2 / 2:g}"


class FermionBasis(Basis):
    """Spinless fermions; n != 0 or conserve_n fixes the particle number
    (basis.py:87-114)."""
    _create = "qkrylov_fermion_basis_create"
    site_class = FermionSite

    def __init__(self, N, conserve_n=False, n=0, dtype=np.float64):
        self.n = n if (conserve_n or n) else None
        self._build(N, dtype, n=self.n)

    def _sector_repr(self):
        return "" if self.n is None else f", n={self.n}"


class HubbardBasis(Basis):
    """Spin-1/2 electrons, up at bit 2i and down at bit 2i+1; each species
    is fixed on its own by nup != 0 / conserve_nup and ndn != 0 /
    conserve_ndn (basis.py:117-155)."""
    _create = "qkrylov_hubbard_basis_create"
    site_class = HubbardSite

    def __init__(self, N, conserve_nup=False, nup=0, conserve_ndn=False, ndn=0, dtype=np.float64):
        self.nup = nup if (conserve_nup or nup) else None
        self.ndn = ndn if (conserve_ndn or ndn) else None
        self._build(N, dtype, nup=self.nup, ndn=self.ndn)

    def _sector_repr(self):
        parts = [f"{k}={v}" for k, v in (("nup", self.nup), ("ndn", self.ndn)) if v is not None]
        return ", " + ", ".join(parts) if parts else ""


class TJBasis(HubbardBasis):
    """Hubbard layout without double occupancy (basis.py:158-187)."""
    _create = "qkrylov_tj_basis_create"
    site_class = TJSite


# ---------------------------------------------------------------- Hamiltonian
def _dp(a):
    return a.ctypes.data_as(_lib.p_dbl)


class MatrixFreeHamiltonian:
    """y = H x with the reference's conventions (row a sums
    conj(<b|term|a>) x[b]); float64 on the device.

    `device` and `dtype` are accepted as in
    bindings/python/qkrylov/hamiltonian.py:24-70. The engine has only the
    MI355X backend: "cpu" (the reference's default), "gpu", "hip" and "cuda"
    all run it on the current HIP device, and a ":N" suffix selects device
    N first. dtype float32 returns complex64 vectors as the reference's FP32
    build does; the arithmetic stays float64."""

    def __init__(self, basis, site=None, ops=None, device="cpu", dtype=None, split_bits=-1, operator=False):
        if ops is None and isinstance(site, (OpSum, OpTermSum, OpProduct)):
            site, ops = None, site
        if site is None:
            site = basis.site_class()
        if not isinstance(ops, OpSum):
            o = OpSum()
            o += ops
            ops = o
        if dtype is None:
            dtype = getattr(ops, "dtype", np.float64)
        L = _lib.load()
        self.basis, self.site, self.ops = basis, site, ops
        self.device, self.dtype = str(device), dtype
        self._split_bits, self._operator = int(split_bits), bool(operator)
        dev = self.device.rpartition(":")[2] if ":" in self.device else None
        if dev is not None:
            check(L.qkrylov_amd_set_device(int(dev)), f"device {self.device}")
        oh = ops._handle()
        try:
            if operator:
                self._h = L.qkrylov_operator_create(basis._h, site._h, oh)
            else:
                self._h = L.qkrylov_hamiltonian_create_ex(basis._h, site._h, oh, int(split_bits))
        finally:
            L.qkrylov_opsum_destroy(oh)
        if not self._h:
            raise ValueError(f"Hamiltonian creation failed: {_lib.last_error()}")
        self._refresh_info()

    def to(self, device=None, dtype=None):
        """New handle over the same basis, site and terms
        (hamiltonian.py:113-117)."""
        return MatrixFreeHamiltonian(self.basis, self.site, self.ops, device=device or self.device,
                                     dtype=dtype or self.dtype, split_bits=self._split_bits,
                                     operator=self._operator)

    def _refresh_info(self):
        info = _lib.HamInfo()
        check(_lib.load().qkrylov_hamiltonian_info(self._h, ctypes.byref(info)), "info")
        self.info = {f: getattr(info, f) for f, _ in info._fields_}

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.load().qkrylov_hamiltonian_destroy(self._h)
            self._h = None

    @property
    def dimension(self):
        return int(self.info["dim"])

    @property
    def is_real(self):
        return bool(self.info["is_real"])

    def apply(self, x):
        """y = H x for one vector of `dimension` entries (complex128, or
        complex64 for a float32 handle); a wrong length raises ValueError."""
        x = np.ascontiguousarray(x, dtype=np.complex128)
        if x.ndim != 1 or x.shape[0] != self.dimension:
            raise ValueError(f"input vector has shape {x.shape}, expected ({self.dimension},)")
        y = np.zeros(self.dimension, dtype=np.complex128)
        check(_lib.load().qkrylov_hamiltonian_apply_complex_f64(self._h, _dp(x.view(np.float64)),
                                                               _dp(y.view(np.float64))), "apply")
        return y.astype(_vec_dtype(self.dtype), copy=False)

    def __matmul__(self, x):
        return self.apply(x)

    def diagonal(self):
        d = np.zeros(self.dimension)
        check(_lib.load().qkrylov_hamiltonian_diagonal_f64(self._h, _dp(d)), "diagonal")
        return d.astype(np.float32) if np.dtype(self.dtype) == np.float32 else d

    def to_sparse(self):
        """Explicit CSR matrix (hamiltonian.py:147-180: column i holds H e_i),
        built from each row's entries instead of D applies; for small
        systems."""
        import scipy.sparse as sp
        D = self.dimension
        rows, cols, vals = [], [], []
        step = 4096
        for r0 in range(0, D, step):
            for i, (c, v) in enumerate(self.row_entries(r0, min(step, D - r0))):
                rows.append(np.full(len(c), r0 + i, dtype=np.int64))
                cols.append(c)
                vals.append(v)
        rows = np.concatenate(rows) if rows else np.zeros(0, np.int64)
        cols = np.concatenate(cols) if cols else np.zeros(0, np.int64)
        vals = np.concatenate(vals) if vals else np.zeros(0, np.complex128)
        keep = vals != 0
        m = sp.csr_matrix((vals[keep], (rows[keep], cols[keep])), shape=(D, D), dtype=np.complex128)
        return m.astype(_vec_dtype(self.dtype))

    def apply_device(self, x_ptr, y_ptr, nvec=1, stream=None):
        check(_lib.load().qkrylov_hamiltonian_apply_device_f64(self._h, x_ptr, y_ptr, nvec, stream), "apply_device")

    def apply_columns_device(self, x_ptr, y_ptr, diag_scale, stream=None):
        """nv = len(diag_scale) in (1, 2, 4, 8) real row-major columns
        (x[row * nv + c]); column c is multiplied by H with its diagonal
        scaled by diag_scale[c]."""
        sc = np.ascontiguousarray(diag_scale, dtype=np.float64)
        check(_lib.load().qkrylov_hamiltonian_apply_columns_device_f64(self._h, x_ptr, y_ptr, len(sc), _dp(sc),
                                                                        stream), "apply_columns_device")

    def apply_device_complex(self, x_ptr, y_ptr, stream=None):
        check(_lib.load().qkrylov_hamiltonian_apply_device_complex_f64(self._h, x_ptr, y_ptr, stream),
              "apply_device_complex")

    def row_entries(self, row0, nrows, cap=None):
        cap = cap or (self.info["n_offdiag_branches"] + 1)
        cols = np.zeros(nrows * cap, dtype=np.int64)
        vals = np.zeros(2 * nrows * cap)
        cnt = np.zeros(nrows, dtype=np.int32)
        check(_lib.load().qkrylov_hamiltonian_row_entries(self._h, row0, nrows, cap, cols.ctypes.data_as(_lib.p_i64),
                                                          _dp(vals), cnt.ctypes.data_as(_lib.p_int)), "row_entries")
        vals = vals.view(np.complex128).reshape(nrows, cap)
        cols = cols.reshape(nrows, cap)
        return [(cols[i, :cnt[i]].copy(), vals[i, :cnt[i]].copy()) for i in range(nrows)]

    def states(self, row0, n):
        out = np.zeros(n, dtype=np.uint64)
        check(_lib.load().qkrylov_hamiltonian_states(self._h, row0, n, out.ctypes.data_as(_lib.p_u64)), "states")
        return out

    TUNING_KNOBS = {"schedule": 0, "block_schedule": 1, "apply_waves": 2, "block_waves": 3, "block_rows": 4,
                    "split_schedule": 5, "split_waves": 6}

    def set_tuning(self, **knobs):
        """Launch tuning of the apply kernels (same apply output, different
        speed): schedule / block_schedule (0 one tile segment per XCD, 1 one
        front over all XCDs, 2 one front of XCD-contiguous chunks),
        apply_waves (4, 8, 16), block_waves (16, 8, 4), block_rows (64, 32,
        16 rows per work item of the multi-vector apply); split_schedule and
        split_waves set both passes of the split apply."""
        for k, v in knobs.items():
            check(_lib.load().qkrylov_hamiltonian_set_tuning(self._h, self.TUNING_KNOBS[k], int(v)), f"set_tuning {k}")

    def autotune(self, nvec=1, allow_split=True):
        """Times the apply under candidate launch tunings (and lo-bit splits
        when allow_split) and keeps the fastest; returns the report."""
        r = _lib.TuneReport()
        check(_lib.load().qkrylov_hamiltonian_autotune(self._h, int(nvec), 1 if allow_split else 0, ctypes.byref(r)),
              "autotune")
        self._refresh_info()
        return {f: getattr(r, f) for f, _ in r._fields_}

    def tuning(self):
        out = {}
        for k, i in self.TUNING_KNOBS.items():
            if i >= 5:   # the split apply's knobs: see split_tuning()
                continue
            v = ctypes.c_int()
            check(_lib.load().qkrylov_hamiltonian_get_tuning(self._h, i, ctypes.byref(v)), "get_tuning")
            out[k] = v.value
        return out

    def split_info(self, any_size=True):
        """(NL, nB, nC) of the two-pass apply's site groups, or None when the
        model admits no split (below 2^24 rows only with any_size)."""
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(_lib.load().qkrylov_hamiltonian_split_info(self._h, 1 if any_size else 0, ctypes.byref(a),
                                                         ctypes.byref(b), ctypes.byref(c)), "split_info")
        return (a.value, b.value, c.value) if a.value else None

    def apply_split_device(self, x_ptr, y_ptr, stream=None):
        """y = H x through the two-pass apply (public row order in and out)."""
        check(_lib.load().qkrylov_hamiltonian_apply_split_device_f64(self._h, x_ptr, y_ptr, stream), "apply_split")

    def time_split_apply(self, warmup=2, reps=10):
        ms = ctypes.c_double()
        check(_lib.load().qkrylov_hamiltonian_time_split_apply(self._h, warmup, reps, ctypes.byref(ms)),
              "time_split_apply")
        return ms.value

    def time_apply(self, nvec=1, warmup=2, reps=10):
        ms = ctypes.c_double()
        check(_lib.load().qkrylov_hamiltonian_time_apply(self._h, nvec, warmup, reps, ctypes.byref(ms)), "time_apply")
        return ms.value

    def aslinearoperator(self):
        import scipy.sparse.linalg as sla
        return sla.LinearOperator((self.dimension, self.dimension), matvec=self.apply, dtype=_vec_dtype(self.dtype))

    def __repr__(self):
        return f"MatrixFreeHamiltonian(dim={self.dimension})"


# ---------------------------------------------------------------- solvers
class LanczosResult:
    def __init__(self, energy, iterations, converged, eigenvector=None, alphas=None, betas=None, report=None):
        self.energy, self.iterations, self.converged = energy, iterations, converged
        self.eigenvector, self.alphas, self.betas, self.report = eigenvector, alphas, betas, report

    def __repr__(self):
        return f"LanczosResult(energy={self.energy:.12f}, iterations={self.iterations}, converged={self.converged})"


_START_MODES = {"counter": 0, "reference": 1, "reference-gcc": 1, "reference-clang": 2}


def _start_mode_code(mode):
    if isinstance(mode, str):
        if mode not in _START_MODES:
            raise ValueError(f"start mode {mode!r}: expected one of {sorted(_START_MODES)}")
        return _START_MODES[mode]
    if mode not in (0, 1, 2):
        raise ValueError(f"start mode {mode!r}: expected 0, 1 or 2")
    return int(mode)


def set_start_mode(mode):
    """Start vector of lanczos_ground_state when its call names none
    (start_mode=None), process-wide, and of the C ABI's reference-named
    Lanczos calls: "counter" (default) draws uniform [-1, 1) from the
    engine's counter RNG; "reference" (= "reference-gcc") the reference's
    std::mt19937(seed) complex draws (src/solvers/lanczos.cpp:115-119), the
    imaginary part of each entry drawn first as GCC evaluates its
    KComplex(dist(rng), dist(rng)); "reference-clang" the real part first."""
    check(_lib.load().qkrylov_amd_set_start_mode(_start_mode_code(mode)), "set_start_mode")


def get_start_mode():
    code = int(_lib.load().qkrylov_amd_get_start_mode())
    return {0: "counter", 1: "reference", 2: "reference-clang"}[code]


def reference_start_vector(n, seed=1234, mode="reference"):
    """The reference's Lanczos start vector of n entries (complex128), as
    start mode `mode` draws it; computed on the host."""
    code = _start_mode_code(mode)
    if code == 0:
        raise ValueError("reference_start_vector: mode must be 'reference' or 'reference-clang'")
    out = np.zeros(2 * int(n))
    check(_lib.load().qkrylov_amd_reference_start_vector(int(n), int(seed), code, _dp(out)),
          "reference_start_vector")
    return out.view(np.complex128)


def reference_normal_vector(n, seed=42, mode="reference"):
    """The reference's FTLM draws (std::mt19937(seed) and
    std::normal_distribution(0, 1), one stream): n complex entries
    (complex128) in the argument order of start mode `mode`; ftlm() uses
    the first n_random * dim of them for seed 42, vector by vector."""
    code = _start_mode_code(mode)
    if code == 0:
        raise ValueError("reference_normal_vector: mode must be 'reference' or 'reference-clang'")
    out = np.zeros(2 * int(n))
    check(_lib.load().qkrylov_amd_reference_normal_vector(int(n), int(seed), code, _dp(out)),
          "reference_normal_vector")
    return out.view(np.complex128)


def lanczos_ground_state(H, maxiter=200, tol=1e-12, reorthogonalize=True, return_state=True, seed=1234,
                         fixed_steps=False, start=None, check_every=0, graph=None, split=None, start_mode=None):
    """Ground state energy and, as in bindings/python/qkrylov/solvers.py:24-51,
    the eigenvector (complex, `H.dimension` entries; return_state=False
    skips it). graph: None automatic (replay a captured step graph for
    fixed-step runs below 2^22 rows when no Krylov basis is stored), True
    always, False never.
    split: True runs the two-pass apply over a relabelled basis (real
    spin-1/2 models that admit it); None (automatic) and False the
    single-pass apply; report["split"] tells which ran.
    start_mode: None the process default (set_start_mode), else as there;
    the reference modes run complex vectors for real H too, and a given
    `start` is then read as complex."""
    L = _lib.load()
    gm = 0 if graph is None else (1 if graph else 2)
    sm = 0 if split is None else (1 if split else 2)
    stm = -1 if start_mode is None else _start_mode_code(start_mode)
    o = _lib.LanczosOptions(int(maxiter), float(tol), 1 if reorthogonalize else 0, 1 if fixed_steps else 0,
                            1 if return_state else 0, int(seed), int(check_every), gm, sm, stm)
    al = np.zeros(max(maxiter, 1))
    be = np.zeros(max(maxiter, 1))
    # room for a complex vector in every case: the process default may pick one
    vec = np.zeros(2 * H.dimension) if return_state else None
    cplx_start = (not H.is_real) or (stm if stm >= 0 else int(L.qkrylov_amd_get_start_mode())) != 0
    st = None
    if start is not None:
        st = np.ascontiguousarray(start, dtype=np.complex128 if cplx_start else np.float64).view(np.float64)
    rep = _lib.LanczosReport()
    check(L.qkrylov_lanczos_run_f64(H._h, ctypes.byref(o), _dp(st) if st is not None else None, _dp(al), _dp(be),
                                     _dp(vec) if vec is not None else None, ctypes.byref(rep)), "lanczos")
    if vec is not None:
        vec = vec.view(np.complex128) if rep.vector_complex else vec[:H.dimension]
        vec = vec.astype(_vec_dtype(H.dtype), copy=False)
    return LanczosResult(rep.energy, rep.iterations, bool(rep.converged), vec, al[:rep.n_alpha].copy(),
                         be[:rep.n_beta].copy(), {f: getattr(rep, f) for f, _ in rep._fields_})


class DavidsonResult:
    def __init__(self, eigenvalues, eigenvectors, iterations, converged):
        self.eigenvalues, self.eigenvectors = eigenvalues, eigenvectors
        self.iterations, self.converged = iterations, converged

    def __repr__(self):
        return f"DavidsonResult(energies={self.eigenvalues})"


def davidson_lowest(H, n_eig=1, max_subspace=20, tol=1e-8):
    ev = np.zeros(n_eig)
    vecs = np.zeros((n_eig, 2 * H.dimension))
    info = _lib.DavidsonResultF64()
    check(_lib.load().qkrylov_davidson_lowest_complex_f64(H._h, n_eig, max_subspace, tol, _dp(ev), _dp(vecs),
                                                         ctypes.byref(info)), "davidson")
    return DavidsonResult(ev, [v.view(np.complex128).astype(_vec_dtype(H.dtype), copy=False) for v in vecs],
                          info.iterations, bool(info.converged))


class DynamicsResult:
    def __init__(self, alphas, betas, norm_phi0):
        self.alphas, self.betas, self.norm_phi0 = np.asarray(alphas), np.asarray(betas), norm_phi0


def continued_fraction_coeffs(H, phi0, n_iter=100):
    phi = np.ascontiguousarray(phi0, dtype=np.complex128)
    if phi.ndim != 1 or phi.shape[0] != H.dimension:
        raise ValueError(f"phi0 has shape {phi.shape}, expected ({H.dimension},)")
    al, be = np.zeros(n_iter), np.zeros(n_iter)
    nrm = ctypes.c_double()
    na = ctypes.c_int()
    check(_lib.load().qkrylov_continued_fraction_coeffs_complex_f64(H._h, _dp(phi.view(np.float64)), n_iter,
                                                                   _dp(al), _dp(be), ctypes.byref(nrm),
                                                                   ctypes.byref(na)), "continued_fraction")
    n = na.value
    return DynamicsResult(al[:n].copy(), be[:max(n - 1, 0)].copy() if n else be[:0], nrm.value)


def evaluate_spectral_function(res, omega, E0, eta=0.1):
    a = np.ascontiguousarray(res.alphas, dtype=np.float64)
    b = np.ascontiguousarray(res.betas if len(res.betas) else [0.0], dtype=np.float64)
    return _lib.load().qkrylov_evaluate_spectral_function_f64(_dp(a), _dp(b), len(a), res.norm_phi0, omega, E0, eta)


class FTLMResult:
    def __init__(self, beta, Z, E, Cv):
        self.beta, self.partition_function, self.internal_energy, self.specific_heat = beta, Z, E, Cv


def ftlm(H, beta, n_random=50, n_steps=100):
    r = _lib.FTLMResultF64()
    check(_lib.load().qkrylov_ftlm_f64(H._h, beta, n_random, n_steps, ctypes.byref(r)), "ftlm")
    return FTLMResult(r.beta, r.partition_function, r.internal_energy, r.specific_heat)


def ftlm_batched(H, R, M, betas, seed=42, sequential=False):
    """R Rademacher vectors (counter RNG) as one multi-vector recurrence;
    observables for every inverse temperature in `betas`."""
    bT = np.ascontiguousarray(betas, dtype=np.float64)
    nT = len(bT)
    Z, E, Cv = np.zeros(nT), np.zeros(nT), np.zeros(nT)
    al, be = np.zeros(R * M), np.zeros(R * M)
    ns = np.zeros(R, dtype=np.int32)
    sec = ctypes.c_double()
    check(_lib.load().qkrylov_ftlm_batched_f64(H._h, R, M, seed, _dp(bT), nT, 1 if sequential else 0, _dp(Z), _dp(E),
                                              _dp(Cv), _dp(al), _dp(be), ns.ctypes.data_as(_lib.p_int),
                                              ctypes.byref(sec)), "ftlm_batched")
    al, be = al.reshape(R, M), be.reshape(R, M)
    return dict(Z=Z, E=E, Cv=Cv, alphas=[al[r, :ns[r]].copy() for r in range(R)],
                betas=[be[r, :max(ns[r] - 1, 0)].copy() for r in range(R)], nsteps=ns, seconds=sec.value)


def continued_fraction_batched(H, seeds, n_iter=100):
    seeds = np.ascontiguousarray(np.asarray(seeds, dtype=np.complex128))
    ns = seeds.shape[0]
    al, be = np.zeros(ns * n_iter), np.zeros(ns * n_iter)
    nrm = np.zeros(ns)
    na = np.zeros(ns, dtype=np.int32)
    sec = ctypes.c_double()
    check(_lib.load().qkrylov_cf_batched_f64(H._h, ns, _dp(seeds.view(np.float64)), n_iter, _dp(al), _dp(be),
                                            _dp(nrm), na.ctypes.data_as(_lib.p_int), ctypes.byref(sec)), "cf_batched")
    al, be = al.reshape(ns, n_iter), be.reshape(ns, n_iter)
    out = [DynamicsResult(al[s, :na[s]].copy(), be[s, :max(na[s] - 1, 0)].copy(), nrm[s]) for s in range(ns)]
    return out, sec.value


def tridiag_lowest(alphas, betas):
    """Lowest eigenvalue of the symmetric tridiagonal (alphas, betas) by
    Sturm bisection (host; the Lanczos drivers' convergence test)."""
    a = np.ascontiguousarray(alphas, dtype=np.float64)
    b = np.ascontiguousarray(betas if len(betas) else [0.0], dtype=np.float64)
    if len(a) == 0:
        raise ValueError("empty tridiagonal")
    out = ctypes.c_double()
    check(_lib.load().qkrylov_amd_tridiag_lowest(len(a), _dp(a), _dp(b), ctypes.byref(out)), "tridiag_lowest")
    return out.value


def tridiag_eigen(alphas, betas, vectors="first"):
    """Eigenvalues (ascending) of the symmetric tridiagonal (alphas, betas)
    and, for vectors="first", the first eigenvector components (the FTLM
    weights), for vectors="all" the eigenvector matrix (columns), for
    vectors=None nothing (host implicit QL)."""
    a = np.ascontiguousarray(alphas, dtype=np.float64)
    b = np.ascontiguousarray(betas if len(betas) else [0.0], dtype=np.float64)
    n = len(a)
    want = {None: 0, "first": 1, "all": 2}[vectors]
    w = np.zeros(max(n, 1))
    z = np.zeros(max(n * n if want == 2 else n, 1))
    check(_lib.load().qkrylov_amd_tridiag_eigen(n, _dp(a), _dp(b), _dp(w), _dp(z), want), "tridiag_eigen")
    if want == 0:
        return w[:n]
    return w[:n], (z[:n * n].reshape(n, n) if want == 2 else z[:n])


class StructureFactorResult:
    def __init__(self, omega, spectrum, cf, seconds):
        self.omega, self.spectrum, self.cf, self.seconds = omega, spectrum, cf, seconds

    def __repr__(self):
        return f"StructureFactorResult(operators={len(self.cf)}, omega points={len(self.omega)})"


def structure_factor(H, psi0, operators, omega, E0, eta=0.1, n_iter=100):
    """S_q(omega) = -1/pi Im <phi_q|(omega + E0 + i eta - H)^-1|phi_q> with
    phi_q = O_q psi0 for every operator in `operators` (OpSum, term lists
    [(coeff, [(op, site), ...])], operator expressions, or
    MatrixFreeHamiltonian handles built with operator=True;
    O_q acts as written, not as its adjoint). The seeds are formed on the
    device and run as one block continued fraction of n_iter steps. Returns
    spectrum[q, w] and the continued fractions (DynamicsResult per q)."""
    L = _lib.load()
    handles, keep = [], []
    for o in operators:
        if isinstance(o, MatrixFreeHamiltonian):
            op = o
        else:
            if isinstance(o, (list, tuple)):
                o = OpSum(list(o))     # [(coeff, [(op, site), ...]), ...]
            op = MatrixFreeHamiltonian(H.basis, H.site, o, operator=True)
        keep.append(op)
        handles.append(op._h)
    nq = len(handles)
    if nq == 0:
        raise ValueError("no operators given")
    psi = np.ascontiguousarray(psi0, dtype=np.complex128)
    if psi.ndim != 1 or psi.shape[0] != H.dimension:
        raise ValueError(f"psi0 has shape {psi.shape}, expected ({H.dimension},)")
    om = np.ascontiguousarray(omega, dtype=np.float64).ravel()
    spec = np.zeros((nq, max(len(om), 1)))
    al, be = np.zeros(nq * n_iter), np.zeros(nq * n_iter)
    nrm = np.zeros(nq)
    na = np.zeros(nq, dtype=np.int32)
    sec = ctypes.c_double()
    harr = (ctypes.c_void_p * nq)(*handles)
    check(L.qkrylov_structure_factor_f64(H._h, harr, nq, _dp(psi.view(np.float64)), 1, float(E0), int(n_iter),
                                         _dp(om), len(om), float(eta), _dp(spec), _dp(al), _dp(be), _dp(nrm),
                                         na.ctypes.data_as(_lib.p_int), ctypes.byref(sec)), "structure_factor")
    al, be = al.reshape(nq, n_iter), be.reshape(nq, n_iter)
    cf = [DynamicsResult(al[q, :na[q]].copy(), be[q, :max(na[q] - 1, 0)].copy(), nrm[q]) for q in range(nq)]
    return StructureFactorResult(om, spec[:, :len(om)], cf, sec.value)


# ---------------------------------------------------------------- device memory
class DeviceArray:
    """Minimal device buffer of float64 owned by the engine's allocator."""

    def __init__(self, n):
        self.n = int(n)
        self.ptr = _lib.load().qkrylov_amd_malloc(max(self.n, 1) * 8)
        if not self.ptr:
            raise MemoryError(_lib.last_error())

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        d = cls(a.size)
        check(_lib.load().qkrylov_amd_memcpy_h2d(d.ptr, a.ctypes.data, a.size * 8), "h2d")
        return d

    def to_host(self):
        out = np.zeros(self.n)
        check(_lib.load().qkrylov_amd_synchronize(), "sync")
        check(_lib.load().qkrylov_amd_memcpy_d2h(out.ctypes.data, self.ptr, self.n * 8), "d2h")
        return out

    def free(self):
        if self.ptr:
            _lib.load().qkrylov_amd_free(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()
