/*
 * qk_oracle.c -- CPU restatement of the sjp95/qkrylov algorithm.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
 * library; the product (qkrylov_amd/) never does.
 *
 * Each function names the reference file:line it follows. Pinning of every
 * function is listed in qk_oracle.h and done in tests/test_oracle_*.py.
 */
#include "qk_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double _Complex cplx;

/* ------------------------------------------------------------------ */
/* bits                                                                */
/* ------------------------------------------------------------------ */
static inline int pc64(uint64_t x) { return __builtin_popcountll(x); }

static uint64_t bit_mask_below(int b) /* (1ULL << b) - 1 as in the reference */
{
    return (b >= 64) ? ~0ULL : ((1ULL << b) - 1ULL);
}

/* ------------------------------------------------------------------ */
/* site actions                                                        */
/* ------------------------------------------------------------------ */
/* SpinHalfSite::apply, src/sites/spinhalf_site.cpp:18-121 */
static int spin_apply(const char* op, int site, uint64_t st, uint64_t* ns, cplx* amp)
{
    const int up = (int)((st >> site) & 1ULL);
    if (!strcmp(op, "Sz")) { *ns = st; *amp = up ? 0.5 : -0.5; return 1; }
    if (!strcmp(op, "Sp")) { if (up) return 0; *ns = st | (1ULL << site); *amp = 1.0; return 1; }
    if (!strcmp(op, "Sm")) { if (!up) return 0; *ns = st & ~(1ULL << site); *amp = 1.0; return 1; }
    if (!strcmp(op, "Sx")) { *ns = st ^ (1ULL << site); *amp = 0.5; return 1; }
    if (!strcmp(op, "Sy")) {
        *ns = st ^ (1ULL << site);
        *amp = up ? (0.0 - 0.5 * I) : (0.0 + 0.5 * I);   /* :104-113 */
        return 1;
    }
    return -1; /* :118-120 throws */
}

/* FermionSite::phase / apply, src/sites/fermion_site.cpp:19-84 */
static double jw_phase(uint64_t st, int bit)
{
    return (pc64(st & bit_mask_below(bit)) % 2 == 0) ? 1.0 : -1.0;
}

static int fermion_apply(const char* op, int site, uint64_t st, uint64_t* ns, cplx* amp)
{
    const int occ = (int)((st >> site) & 1ULL);
    if (!strcmp(op, "N"))  { *ns = st; *amp = occ ? 1.0 : 0.0; return 1; }
    if (!strcmp(op, "Id")) { *ns = st; *amp = 1.0; return 1; }
    if (!strcmp(op, "C"))  { if (!occ) return 0; *ns = st & ~(1ULL << site); *amp = jw_phase(st, site); return 1; }
    if (!strcmp(op, "Cdag")) { if (occ) return 0; *ns = st | (1ULL << site); *amp = jw_phase(st, site); return 1; }
    return -1;
}

/* HubbardSite (src/sites/hubbard_site.cpp:15-125) and TJSite
   (src/sites/tj_site.cpp:11-113): up at bit 2i, dn at bit 2i+1 */
static int hubbard_like_apply(int tj, const char* op, int site, uint64_t st, uint64_t* ns, cplx* amp)
{
    const int bu = 2 * site, bd = 2 * site + 1;
    const int up = (int)((st >> bu) & 1ULL);
    const int dn = (int)((st >> bd) & 1ULL);
    if (!strcmp(op, "Nup")) { *ns = st; *amp = up ? 1.0 : 0.0; return 1; }
    if (!strcmp(op, "Ndn")) { *ns = st; *amp = dn ? 1.0 : 0.0; return 1; }
    if (!tj && !strcmp(op, "Nupdn")) { *ns = st; *amp = (up && dn) ? 1.0 : 0.0; return 1; }
    if (!strcmp(op, "CUp")) { if (!up) return 0; *ns = st & ~(1ULL << bu); *amp = jw_phase(st, bu); return 1; }
    if (!strcmp(op, "CdagUp")) {
        if (up || (tj && dn)) return 0;           /* tj_site.cpp:85 */
        *ns = st | (1ULL << bu); *amp = jw_phase(st, bu); return 1;
    }
    if (!strcmp(op, "CDn")) { if (!dn) return 0; *ns = st & ~(1ULL << bd); *amp = jw_phase(st, bd); return 1; }
    if (!strcmp(op, "CdagDn")) {
        if (dn || (tj && up)) return 0;           /* tj_site.cpp:103 */
        *ns = st | (1ULL << bd); *amp = jw_phase(st, bd); return 1;
    }
    return -1;
}

static int site_apply_c(int kind, const char* op, int site, uint64_t st, uint64_t* ns, cplx* amp)
{
    switch (kind) {
    case QO_SPIN: return spin_apply(op, site, st, ns, amp);
    case QO_FERMION: return fermion_apply(op, site, st, ns, amp);
    case QO_HUBBARD: return hubbard_like_apply(0, op, site, st, ns, amp);
    case QO_TJ: return hubbard_like_apply(1, op, site, st, ns, amp);
    }
    return -1;
}

int qo_site_apply(int site_kind, const char* op, int site, uint64_t state,
                  uint64_t* new_state, double* amp_re, double* amp_im)
{
    uint64_t ns = 0; cplx a = 0.0;
    int v = site_apply_c(site_kind, op, site, state, &ns, &a);
    if (v == 1) { *new_state = ns; *amp_re = creal(a); *amp_im = cimag(a); }
    return v;
}

/* ------------------------------------------------------------------ */
/* basis: enumeration (reference)                                      */
/* ------------------------------------------------------------------ */
static int nbits_of(int kind, int n) { return (kind == QO_HUBBARD || kind == QO_TJ) ? 2 * n : n; }

static uint64_t up_mask_of(int n) { uint64_t m = 0; for (int i = 0; i < n; ++i) m |= 1ULL << (2 * i); return m; }

/* validity predicates of the four build_* loops:
   spinhalf_basis.cpp:86-104, fermion_basis.cpp:72-89,
   hubbard_basis.cpp:72-100, tj_basis.cpp:56-96 */
static int valid_state(int kind, int n, const qo_sector* sec, uint64_t s)
{
    switch (kind) {
    case QO_SPIN:
        if (!sec->use_sz) return 1;
        return (2 * pc64(s) - n) == sec->sz2;
    case QO_FERMION:
        if (!sec->use_n) return 1;
        return pc64(s) == sec->n;
    case QO_HUBBARD:
    case QO_TJ: {
        const uint64_t um = up_mask_of(n), dm = um << 1;
        if (kind == QO_TJ) {
            for (int i = 0; i < n; ++i)
                if (((s >> (2 * i)) & 1ULL) && ((s >> (2 * i + 1)) & 1ULL)) return 0;
        }
        if (kind == QO_HUBBARD && !(sec->use_nup || sec->use_ndn)) return 1;
        int mu = !sec->use_nup || pc64(s & um) == sec->nup;
        int md = !sec->use_ndn || pc64(s & dm) == sec->ndn;
        return mu && md;
    }
    }
    return 0;
}

static int n_ok(int kind, int n)
{
    if (kind == QO_SPIN || kind == QO_FERMION) return n > 0 && n <= 63;
    return n > 0 && n <= 31;
}

int64_t qo_basis_enumerate(int kind, int n, const qo_sector* sec, uint64_t* out, int64_t cap)
{
    if (!n_ok(kind, n)) return -1;
    const int nb = nbits_of(kind, n);
    if (nb > 40) return -1; /* the oracle's enumeration is for checks only */
    const uint64_t lim = 1ULL << nb;
    int64_t d = 0;
    for (uint64_t s = 0; s < lim; ++s) {
        if (valid_state(kind, n, sec, s)) {
            if (out && d < cap) out[d] = s;
            ++d;
        }
    }
    return d;
}

/* ------------------------------------------------------------------ */
/* basis: counting rank (oracle's own closed form)                     */
/* ------------------------------------------------------------------ */
static uint64_t binom(int n, int k)
{
    if (k < 0 || n < 0 || k > n) return 0;
    if (k > n - k) k = n - k;
    uint64_t r = 1;
    for (int i = 1; i <= k; ++i) r = r * (uint64_t)(n - k + i) / (uint64_t)i;
    return r;
}
static uint64_t pow2u(int e) { return (e >= 64) ? 0 : (1ULL << e); }
static uint64_t pow3u(int e) { uint64_t r = 1; while (e-- > 0) r *= 3; return r; }

/* target quantum numbers; -1 = unconstrained; returns 0 if sector empty */
static int qn_targets(int kind, int n, const qo_sector* sec, int* A, int* B)
{
    *A = -1; *B = -1;
    switch (kind) {
    case QO_SPIN:
        if (sec->use_sz) {
            if (((sec->sz2 + n) & 1) != 0) return 0;
            *A = (sec->sz2 + n) / 2;
            if (*A < 0 || *A > n) return 0;
        }
        return 1;
    case QO_FERMION:
        if (sec->use_n) { *A = sec->n; if (*A < 0 || *A > n) return 0; }
        return 1;
    default:
        if (sec->use_nup) { *A = sec->nup; if (*A < 0 || *A > n) return 0; }
        if (sec->use_ndn) { *B = sec->ndn; if (*B < 0 || *B > n) return 0; }
        return 1;
    }
}

/* t-J complete sites: fillings of i sites with a ups, b dns (wildcards -1) */
static uint64_t tj_sites(int i, int a, int b)
{
    if (a < -1 || b < -1) return 0;
    if (a == -1 && b == -1) return pow3u(i);
    if (a == -1) return binom(i, b) * pow2u(i - b);
    if (b == -1) return binom(i, a) * pow2u(i - a);
    return binom(i, a) * binom(i - a, b);
}

/* number of valid fillings of bits [0, p) with remaining counts (a, b) */
static uint64_t completions(int kind, int p, int a, int b)
{
    if (kind == QO_SPIN || kind == QO_FERMION) {
        if (a == -1) return pow2u(p);
        return binom(p, a);
    }
    if (kind == QO_HUBBARD) {
        int nu = (p + 1) / 2, nd = p / 2;
        uint64_t cu = (a == -1) ? pow2u(nu) : binom(nu, a);
        uint64_t cd = (b == -1) ? pow2u(nd) : binom(nd, b);
        return cu * cd;
    }
    /* t-J */
    int i = p / 2;
    if ((p & 1) == 0) return tj_sites(i, a, b);
    if (a == -1) return 2 * tj_sites(i, -1, b);
    return tj_sites(i, a, b) + (a >= 1 ? tj_sites(i, a - 1, b) : 0);
}

static int is_up_bit(int kind, int p) { return (kind == QO_SPIN || kind == QO_FERMION) ? 1 : ((p & 1) == 0); }

int64_t qo_count_dimension(int kind, int n, const qo_sector* sec)
{
    int A, B;
    if (!n_ok(kind, n)) return -1;
    if (!qn_targets(kind, n, sec, &A, &B)) return 0;
    if (kind == QO_HUBBARD && !(sec->use_nup || sec->use_ndn)) { A = -1; B = -1; }
    return (int64_t)completions(kind, nbits_of(kind, n), A, B);
}

int qo_count_contains(int kind, int n, const qo_sector* sec, uint64_t s)
{
    const int nb = nbits_of(kind, n);
    if (nb < 64 && (s >> nb) != 0) return 0;
    return valid_state(kind, n, sec, s);
}

int64_t qo_count_rank(int kind, int n, const qo_sector* sec, uint64_t s)
{
    int A, B;
    if (!qo_count_contains(kind, n, sec, s)) return -1;
    qn_targets(kind, n, sec, &A, &B);
    if (kind == QO_HUBBARD && !(sec->use_nup || sec->use_ndn)) { A = -1; B = -1; }
    const int nb = nbits_of(kind, n);
    int64_t r = 0;
    int a = A, b = B;
    for (int p = nb - 1; p >= 0; --p) {
        if (!((s >> p) & 1ULL)) continue;
        r += (int64_t)completions(kind, p, a, b);
        if (is_up_bit(kind, p)) { if (a != -1) --a; }
        else { if (b != -1) --b; }
    }
    return r;
}

uint64_t qo_count_unrank(int kind, int n, const qo_sector* sec, int64_t r)
{
    int A, B;
    qn_targets(kind, n, sec, &A, &B);
    if (kind == QO_HUBBARD && !(sec->use_nup || sec->use_ndn)) { A = -1; B = -1; }
    const int nb = nbits_of(kind, n);
    uint64_t s = 0;
    int a = A, b = B;
    uint64_t rr = (uint64_t)r;
    for (int p = nb - 1; p >= 0; --p) {
        uint64_t c = completions(kind, p, a, b);
        if (rr >= c) {
            rr -= c;
            s |= 1ULL << p;
            if (is_up_bit(kind, p)) { if (a != -1) --a; }
            else { if (b != -1) --b; }
        }
    }
    return s;
}

/* ------------------------------------------------------------------ */
/* model                                                               */
/* ------------------------------------------------------------------ */
typedef struct { char op[16]; int site; } qo_factor;
typedef struct { cplx coeff; int nf; qo_factor* f; } qo_term;

struct qo_model {
    int bkind, n, skind;
    qo_sector sec;
    int enumerated;
    int64_t dim;
    uint64_t* states;          /* enumerated list, or NULL */
    int nterms, cap_terms;
    qo_term* terms;
    /* CSR (rows [csr_r0, csr_r0 + csr_rows)) */
    int64_t csr_r0, csr_rows, nnz;
    int64_t* ro;
    int64_t* ci;
    cplx* vals;
    cplx* diag;
};

qo_model* qo_model_create(int bkind, int n, const qo_sector* sec, int skind, int enumerate)
{
    if (!n_ok(bkind, n)) return NULL;
    qo_model* m = (qo_model*)calloc(1, sizeof(qo_model));
    m->bkind = bkind; m->n = n; m->skind = skind; m->sec = *sec;
    if (enumerate) {
        int64_t d = qo_basis_enumerate(bkind, n, sec, NULL, 0);
        if (d < 0) { free(m); return NULL; }
        m->states = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(d > 0 ? d : 1));
        qo_basis_enumerate(bkind, n, sec, m->states, d);
        m->dim = d;
        m->enumerated = 1;
    } else {
        m->dim = qo_count_dimension(bkind, n, sec);
    }
    return m;
}

void qo_model_destroy(qo_model* m)
{
    if (!m) return;
    for (int t = 0; t < m->nterms; ++t) free(m->terms[t].f);
    free(m->terms); free(m->states);
    free(m->ro); free(m->ci); free(m->vals); free(m->diag);
    free(m);
}

int qo_model_add_term(qo_model* m, double re, double im, int nf, const char** ops, const int* sites)
{
    if (m->nterms == m->cap_terms) {
        m->cap_terms = m->cap_terms ? 2 * m->cap_terms : 16;
        m->terms = (qo_term*)realloc(m->terms, sizeof(qo_term) * (size_t)m->cap_terms);
    }
    qo_term* t = &m->terms[m->nterms++];
    t->coeff = re + im * I;
    t->nf = nf;
    t->f = (qo_factor*)calloc((size_t)(nf > 0 ? nf : 1), sizeof(qo_factor));
    for (int k = 0; k < nf; ++k) {
        strncpy(t->f[k].op, ops[k], sizeof(t->f[k].op) - 1);
        t->f[k].site = sites[k];
    }
    return 0;
}

int64_t qo_model_dim(const qo_model* m) { return m->dim; }

uint64_t qo_model_state(const qo_model* m, int64_t i)
{
    if (i < 0 || i >= m->dim) return 0;
    if (m->enumerated) return m->states[i];
    return qo_count_unrank(m->bkind, m->n, &m->sec, i);
}

int64_t qo_model_index(const qo_model* m, uint64_t s)
{
    if (m->enumerated) {
        /* the reference uses an unordered_map over the ascending list;
           a binary search over the same list returns the same position */
        int64_t lo = 0, hi = m->dim - 1;
        while (lo <= hi) {
            int64_t mid = (lo + hi) / 2;
            if (m->states[mid] == s) return mid;
            if (m->states[mid] < s) lo = mid + 1; else hi = mid - 1;
        }
        return -1;
    }
    return qo_count_rank(m->bkind, m->n, &m->sec, s);
}

/* one row of Phase 1 (matrix_free_hamiltonian.cpp:49-90): entries in term
   order, value conj(amp), invalid/absent targets dropped */
static int row_entries_c(const qo_model* m, int64_t alpha, int64_t* cols, cplx* vals, int cap, int* err)
{
    const uint64_t s0 = qo_model_state(m, alpha);
    int cnt = 0;
    for (int t = 0; t < m->nterms; ++t) {
        const qo_term* tm = &m->terms[t];
        uint64_t st = s0;
        cplx amp = tm->coeff;
        int valid = 1;
        for (int k = 0; k < tm->nf; ++k) {
            uint64_t ns; cplx a;
            int v = site_apply_c(m->skind, tm->f[k].op, tm->f[k].site, st, &ns, &a);
            if (v < 0) { *err = 1; return -1; }
            if (!v) { valid = 0; break; }
            st = ns;
            amp *= a;
        }
        if (!valid) continue;
        int64_t b = qo_model_index(m, st);
        if (b < 0) continue;
        if (cnt < cap) { cols[cnt] = b; vals[cnt] = conj(amp); }
        ++cnt;
    }
    return cnt;
}

int qo_row_entries(const qo_model* m, int64_t row, int64_t* cols, double* vals, int cap)
{
    int err = 0;
    return row_entries_c(m, row, cols, (cplx*)vals, cap, &err);
}

static int64_t build_csr_range(qo_model* m, int64_t r0, int64_t nrows)
{
    free(m->ro); free(m->ci); free(m->vals); free(m->diag);
    m->ro = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nrows + 1));
    m->diag = (cplx*)calloc((size_t)(nrows > 0 ? nrows : 1), sizeof(cplx));
    const int cap = m->nterms > 0 ? m->nterms : 1;
    /* pass 1: counts (parallel), pass 2: fill; same entries as the serial loop */
    int64_t* cnt = (int64_t*)calloc((size_t)(nrows + 1), sizeof(int64_t));
    int err = 0;
    #pragma omp parallel
    {
        int64_t* cols = (int64_t*)malloc(sizeof(int64_t) * (size_t)cap);
        cplx* vals = (cplx*)malloc(sizeof(cplx) * (size_t)cap);
        int lerr = 0;
        #pragma omp for schedule(dynamic, 1024)
        for (int64_t a = 0; a < nrows; ++a) {
            int c = row_entries_c(m, r0 + a, cols, vals, cap, &lerr);
            cnt[a] = c < 0 ? 0 : c;
        }
        if (lerr) {
            #pragma omp atomic write
            err = 1;
        }
        free(cols); free(vals);
    }
    if (err) { free(cnt); return -1; }
    m->ro[0] = 0;
    for (int64_t a = 0; a < nrows; ++a) m->ro[a + 1] = m->ro[a] + cnt[a];
    free(cnt);
    const int64_t nnz = m->ro[nrows];
    m->ci = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nnz > 0 ? nnz : 1));
    m->vals = (cplx*)malloc(sizeof(cplx) * (size_t)(nnz > 0 ? nnz : 1));
    #pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t a = 0; a < nrows; ++a) {
        int e2 = 0;
        int c = row_entries_c(m, r0 + a, m->ci + m->ro[a], m->vals + m->ro[a], cap, &e2);
        for (int j = 0; j < c; ++j)
            if (m->ci[m->ro[a] + j] == r0 + a) m->diag[a] += m->vals[m->ro[a] + j]; /* :86-89 */
    }
    m->csr_r0 = r0; m->csr_rows = nrows; m->nnz = nnz;
    return nnz;
}

int64_t qo_build_csr(qo_model* m) { return build_csr_range(m, 0, m->dim); }
int64_t qo_build_csr_rows(qo_model* m, int64_t r0, int64_t nrows) { return build_csr_range(m, r0, nrows); }
int64_t qo_csr_nnz(const qo_model* m) { return m->nnz; }

void qo_csr_export(const qo_model* m, int64_t* ro, int64_t* cols, double* vals, double* diag)
{
    if (ro) memcpy(ro, m->ro, sizeof(int64_t) * (size_t)(m->csr_rows + 1));
    if (cols) memcpy(cols, m->ci, sizeof(int64_t) * (size_t)m->nnz);
    if (vals) memcpy(vals, m->vals, sizeof(cplx) * (size_t)m->nnz);
    if (diag) memcpy(diag, m->diag, sizeof(cplx) * (size_t)m->csr_rows);
}

/* gather SpMV, matrix_free_hamiltonian.cpp:156-167 (y zero-fill folded in) */
void qo_apply_rows(const qo_model* m, const double* xd, double* yd, int64_t r0, int64_t nrows)
{
    const cplx* x = (const cplx*)xd;
    cplx* y = (cplx*)yd;
    const int64_t off = r0 - m->csr_r0;
    #pragma omp parallel for schedule(static)
    for (int64_t a = 0; a < nrows; ++a) {
        cplx sum = 0.0;
        const int64_t lr = off + a;
        for (int64_t j = m->ro[lr]; j < m->ro[lr + 1]; ++j) sum += m->vals[j] * x[m->ci[j]];
        y[r0 + a] = sum;
    }
}

void qo_apply(const qo_model* m, const double* x, double* y) { qo_apply_rows(m, x, y, 0, m->dim); }

/* ------------------------------------------------------------------ */
/* BLAS-1 (include/qkrylov/linalg/vector_ops.hpp:14-97)                */
/* ------------------------------------------------------------------ */
static cplx vdot(const cplx* x, const cplx* y, int64_t n)
{
    double re = 0.0, im = 0.0;
    #pragma omp parallel for reduction(+:re, im) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        cplx p = conj(x[i]) * y[i];
        re += creal(p); im += cimag(p);
    }
    return re + im * I;
}
static double vnorm(const cplx* x, int64_t n) { return sqrt(creal(vdot(x, x, n))); }
static void vscal(cplx a, cplx* x, int64_t n)
{
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) x[i] *= a;
}
static void vaxpy(cplx a, const cplx* x, cplx* y, int64_t n)
{
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) y[i] += a * x[i];
}
static void vnormalize(cplx* x, int64_t n)
{
    double nr = vnorm(x, n);
    if (nr == 0.0) return;
    vscal(1.0 / nr, x, n);
}
static void apply_c(const qo_model* m, const cplx* x, cplx* y)
{
    qo_apply(m, (const double*)x, (double*)y);
}

/* ------------------------------------------------------------------ */
/* tridiagonal QR (lanczos.cpp:25-98 and ftlm.cpp:61-119)              */
/* ------------------------------------------------------------------ */
static void tridiag_qr(const double* alpha, const double* beta, int n, double* d, double* z /* n*n row-major, z[j*n+i] */)
{
    double* e = (double*)calloc((size_t)(n > 1 ? n : 1), sizeof(double));
    for (int i = 0; i < n; ++i) d[i] = alpha[i];
    for (int i = 0; i < n - 1; ++i) e[i] = beta[i];
    for (int i = 0; i < n * n; ++i) z[i] = 0.0;
    for (int i = 0; i < n; ++i) z[i * n + i] = 1.0;
    for (int iter = 0; iter < 1000; ++iter) {
        for (int i = 0; i < n - 1; ++i)
            if (fabs(e[i]) < 1e-14 * (fabs(d[i]) + fabs(d[i + 1]))) e[i] = 0.0;
        int mm = n - 1;
        while (mm > 0 && e[mm - 1] == 0.0) mm--;
        if (mm == 0) break;
        int l = mm - 1;
        while (l > 0 && e[l - 1] != 0.0) l--;
        double b = (d[mm - 1] - d[mm]) / 2.0;
        double c = e[mm - 1] * e[mm - 1];
        double s = sqrt(b * b + c);
        double shift = (b > 0) ? d[mm] - c / (b + s) : d[mm] - c / (b - s);
        double p = d[l] - shift;
        double g = e[l];
        for (int i = l; i < mm; ++i) {
            double r = hypot(p, g);
            double ct = p / r, st = g / r;
            if (i > l) e[i - 1] = r;
            double f = ct * d[i] + st * e[i];
            double gn = ct * e[i] + st * d[i + 1];
            double h = st * d[i] - ct * e[i];
            double k = st * e[i] - ct * d[i + 1];
            d[i] = ct * f + st * gn;
            e[i] = ct * h + st * k;
            d[i + 1] = st * h - ct * k;
            for (int j = 0; j < n; ++j) {
                double z1 = z[j * n + i], z2 = z[j * n + i + 1];
                z[j * n + i] = ct * z1 + st * z2;
                z[j * n + i + 1] = st * z1 - ct * z2;
            }
            if (i < mm - 1) {
                p = e[i];
                g = st * e[i + 1];
                e[i + 1] = -ct * e[i + 1];
            }
        }
    }
    free(e);
}

double qo_tridiag_ground(const double* alpha, const double* beta, int n, double* evec)
{
    if (n == 0) return 0.0;
    if (n == 1) { if (evec) evec[0] = 1.0; return alpha[0]; }
    double* d = (double*)malloc(sizeof(double) * (size_t)n);
    double* z = (double*)malloc(sizeof(double) * (size_t)n * (size_t)n);
    tridiag_qr(alpha, beta, n, d, z);
    int mi = 0;
    for (int i = 1; i < n; ++i) if (d[i] < d[mi]) mi = i;
    double e0 = d[mi];
    if (evec) for (int i = 0; i < n; ++i) evec[i] = z[i * n + mi];
    free(d); free(z);
    return e0;
}

void qo_tridiag_components(const double* alpha, const double* beta, int n, double* evals, double* fc)
{
    if (n == 0) return;
    if (n == 1) { evals[0] = alpha[0]; fc[0] = 1.0; return; }
    double* z = (double*)malloc(sizeof(double) * (size_t)n * (size_t)n);
    tridiag_qr(alpha, beta, n, evals, z);
    for (int i = 0; i < n; ++i) fc[i] = z[0 * n + i];
    free(z);
}

/* ------------------------------------------------------------------ */
/* Lanczos ground state (lanczos.cpp:101-202)                          */
/* ------------------------------------------------------------------ */
int qo_lanczos(const qo_model* m, int maxiter, double tol, const double* start, int reorth,
               qo_lanczos_result* res, double* alphas, double* betas, double* eigvec)
{
    const int64_t dim = m->dim;
    memset(res, 0, sizeof(*res));
    if (dim == 0) return 0;
    const int kmax = (int)((int64_t)maxiter < dim ? maxiter : dim);
    cplx* vprev = (cplx*)calloc((size_t)dim, sizeof(cplx));
    cplx* vcur = (cplx*)malloc(sizeof(cplx) * (size_t)dim);
    cplx* w = (cplx*)calloc((size_t)dim, sizeof(cplx));
    memcpy(vcur, start, sizeof(cplx) * (size_t)dim);
    vnormalize(vcur, dim);
    cplx** basis = (cplx**)calloc((size_t)(kmax + 1), sizeof(cplx*));
    int nb = 0;
    basis[nb] = (cplx*)malloc(sizeof(cplx) * (size_t)dim);
    memcpy(basis[nb++], vcur, sizeof(cplx) * (size_t)dim);
    int na = 0, nbeta = 0, conv = 0, iters = 0;
    double e_old = 1e100;
    double* tmpv = (double*)malloc(sizeof(double) * (size_t)(kmax + 1));
    for (int it = 0; it < kmax; ++it) {
        iters = it + 1;
        apply_c(m, vcur, w);
        double a = creal(vdot(vcur, w, dim));
        alphas[na++] = a;
        vaxpy(-a, vcur, w, dim);
        if (it > 0) vaxpy(-betas[nbeta - 1], vprev, w, dim);
        if (reorth)
            for (int j = 0; j < nb; ++j) vaxpy(-vdot(basis[j], w, dim), basis[j], w, dim);
        double b = vnorm(w, dim);
        if (b < 1e-15) { conv = 1; break; }
        if (it + 1 == kmax) break;
        betas[nbeta++] = b;
        memcpy(vprev, vcur, sizeof(cplx) * (size_t)dim);
        memcpy(vcur, w, sizeof(cplx) * (size_t)dim);
        vscal(1.0 / b, vcur, dim);
        basis[nb] = (cplx*)malloc(sizeof(cplx) * (size_t)dim);
        memcpy(basis[nb++], vcur, sizeof(cplx) * (size_t)dim);
        if (it > 0) {
            double e = qo_tridiag_ground(alphas, betas, na, NULL);
            if (fabs(e - e_old) < tol) { e_old = e; conv = 1; break; }
            e_old = e;
        }
    }
    double e0 = qo_tridiag_ground(alphas, betas, na, tmpv);
    res->energy = e0; res->iterations = iters; res->converged = conv;
    res->n_alpha = na; res->n_beta = nbeta;
    if (eigvec) {
        cplx* ritz = (cplx*)eigvec;
        memset(ritz, 0, sizeof(cplx) * (size_t)dim);
        for (int i = 0; i < na; ++i) vaxpy(tmpv[i], basis[i], ritz, dim);
        vnormalize(ritz, dim);
    }
    for (int j = 0; j < nb; ++j) free(basis[j]);
    free(basis); free(vprev); free(vcur); free(w); free(tmpv);
    return 0;
}

/* ------------------------------------------------------------------ */
/* FTLM (ftlm.cpp:24-170)                                              */
/* ------------------------------------------------------------------ */
int qo_compute_tridiag(const qo_model* m, const double* v_start, int n_steps,
                       double* alphas, double* betas, int* nbeta_out)
{
    const int64_t dim = m->dim;
    cplx* vcur = (cplx*)malloc(sizeof(cplx) * (size_t)dim);
    cplx* vprev = (cplx*)calloc((size_t)dim, sizeof(cplx));
    cplx* w = (cplx*)calloc((size_t)dim, sizeof(cplx));
    memcpy(vcur, v_start, sizeof(cplx) * (size_t)dim);
    vnormalize(vcur, dim);
    int na = 0, nbt = 0;
    for (int i = 0; i < n_steps; ++i) {
        apply_c(m, vcur, w);
        double a = creal(vdot(vcur, w, dim));
        alphas[na++] = a;
        vaxpy(-a, vcur, w, dim);
        if (i > 0) vaxpy(-betas[nbt - 1], vprev, w, dim);
        double b = vnorm(w, dim);
        if (b < 1e-15) break;
        betas[nbt++] = b;
        memcpy(vprev, vcur, sizeof(cplx) * (size_t)dim);
        memcpy(vcur, w, sizeof(cplx) * (size_t)dim);
        vscal(1.0 / b, vcur, dim);
    }
    free(vcur); free(vprev); free(w);
    if (nbeta_out) *nbeta_out = nbt;
    return na;
}

int qo_ftlm(const qo_model* m, double beta, int n_random, int n_steps,
            const double* rvecs, double* Zo, double* Eo, double* Cvo)
{
    const int64_t dim = m->dim;
    double Z = 0.0, E = 0.0, E2 = 0.0;
    double* al = (double*)malloc(sizeof(double) * (size_t)(n_steps + 1));
    double* be = (double*)malloc(sizeof(double) * (size_t)(n_steps + 1));
    double* ev = (double*)malloc(sizeof(double) * (size_t)(n_steps + 1));
    double* fc = (double*)malloc(sizeof(double) * (size_t)(n_steps + 1));
    for (int r = 0; r < n_random; ++r) {
        const cplx* rv = (const cplx*)rvecs + (size_t)r * (size_t)dim;
        double nrm = vnorm(rv, dim);
        int nbt = 0;
        int na = qo_compute_tridiag(m, (const double*)rv, n_steps, al, be, &nbt);
        qo_tridiag_components(al, be, na, ev, fc);
        for (int i = 0; i < na; ++i) {
            double wgt = nrm * nrm * fc[i] * fc[i] * exp(-beta * ev[i]);
            Z += wgt; E += ev[i] * wgt; E2 += ev[i] * ev[i] * wgt;
        }
    }
    Z /= n_random; E /= n_random; E2 /= n_random;
    *Zo = Z; *Eo = E / Z; *Cvo = (beta * beta) * (E2 / Z - (E / Z) * (E / Z));
    free(al); free(be); free(ev); free(fc);
    return 0;
}

/* ------------------------------------------------------------------ */
/* continued fraction (dynamics.cpp:12-89)                             */
/* ------------------------------------------------------------------ */
int qo_cf_coeffs(const qo_model* m, const double* phi0, int n_iter,
                 double* alphas, double* betas, int* n_alpha, int* n_beta, double* norm_phi0)
{
    const int64_t dim = m->dim;
    *n_alpha = 0; *n_beta = 0; *norm_phi0 = 0.0;
    if (dim == 0) return 0;
    double nphi = vnorm((const cplx*)phi0, dim);
    if (nphi < 1e-15) return 0;
    cplx* vcur = (cplx*)malloc(sizeof(cplx) * (size_t)dim);
    cplx* vprev = (cplx*)calloc((size_t)dim, sizeof(cplx));
    cplx* w = (cplx*)calloc((size_t)dim, sizeof(cplx));
    memcpy(vcur, phi0, sizeof(cplx) * (size_t)dim);
    vscal(1.0 / nphi, vcur, dim);
    *norm_phi0 = nphi;
    int na = 0, nbt = 0;
    for (int it = 0; it < n_iter; ++it) {
        apply_c(m, vcur, w);
        double a = creal(vdot(vcur, w, dim));
        alphas[na++] = a;
        vaxpy(-a, vcur, w, dim);
        if (it > 0) vaxpy(-betas[nbt - 1], vprev, w, dim);
        double b = vnorm(w, dim);
        if (b < 1e-15) break;
        betas[nbt++] = b;
        memcpy(vprev, vcur, sizeof(cplx) * (size_t)dim);
        memcpy(vcur, w, sizeof(cplx) * (size_t)dim);
        vscal(1.0 / b, vcur, dim);
    }
    *n_alpha = na; *n_beta = nbt;
    free(vcur); free(vprev); free(w);
    return 0;
}

double qo_spectral(const double* alphas, const double* betas, size_t n,
                   double norm_phi0, double omega, double E0, double eta)
{
    if (n == 0) return 0.0;
    cplx z = (omega + E0) + eta * I;
    cplx f = 0.0;
    for (int i = (int)n - 1; i >= 0; --i) {
        if (i == (int)n - 1) f = 1.0 / (z - alphas[i]);
        else f = 1.0 / (z - alphas[i] - betas[i] * betas[i] * f);
    }
    return -1.0 / M_PI * cimag(norm_phi0 * norm_phi0 * f);
}

/* ------------------------------------------------------------------ */
/* Davidson (davidson.cpp:22-227)                                      */
/* ------------------------------------------------------------------ */
/* solve_small_hermitian, davidson.cpp:22-82 (row-major n x n) */
static void small_hermitian(int n, const cplx* H, double* evals, cplx* evecs /* k-th vector at evecs[k*n + j] */)
{
    cplx* A = (cplx*)malloc(sizeof(cplx) * (size_t)n * (size_t)n);
    cplx* V = (cplx*)calloc((size_t)n * (size_t)n, sizeof(cplx));
    memcpy(A, H, sizeof(cplx) * (size_t)n * (size_t)n);
    for (int i = 0; i < n; ++i) V[i * n + i] = 1.0;
    for (int iter = 0; iter < 1000; ++iter) {
        double max_off = 0.0; int p = 0, q = 0;
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j)
                if (cabs(A[i * n + j]) > max_off) { max_off = cabs(A[i * n + j]); p = i; q = j; }
        if (max_off < 1e-15) break;
        cplx app = A[p * n + p], aqq = A[q * n + q], apq = A[p * n + q];
        double phi = 0.5 * atan2(2.0 * cabs(apq), creal(app - aqq));
        cplx c = cos(phi);
        cplx s = sin(phi) * apq / cabs(apq);
        for (int i = 0; i < n; ++i) {
            cplx vip = V[i * n + p], viq = V[i * n + q];
            V[i * n + p] = c * vip + s * viq;
            V[i * n + q] = -conj(s) * vip + c * viq;
            if (i != p && i != q) {
                cplx aip = A[i * n + p], aiq = A[i * n + q];
                A[i * n + p] = A[p * n + i] = c * aip + s * aiq;
                A[i * n + q] = A[q * n + i] = -conj(s) * aip + c * aiq;
            }
        }
        A[p * n + p] = c * c * app + s * conj(s) * aqq + 2.0 * creal(c * s * conj(apq));
        A[q * n + q] = s * conj(s) * app + c * c * aqq - 2.0 * creal(c * s * conj(apq));
        A[p * n + q] = A[q * n + p] = 0.0;
    }
    /* sort by diagonal (stable, so ties keep index order) */
    int* idx = (int*)malloc(sizeof(int) * (size_t)n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    for (int i = 1; i < n; ++i) {
        int v = idx[i], j = i - 1;
        while (j >= 0 && creal(A[idx[j] * n + idx[j]]) > creal(A[v * n + v])) { idx[j + 1] = idx[j]; --j; }
        idx[j + 1] = v;
    }
    for (int k = 0; k < n; ++k) {
        int i = idx[k];
        evals[k] = creal(A[i * n + i]);
        for (int j = 0; j < n; ++j) evecs[k * n + j] = V[j * n + i];
    }
    free(A); free(V); free(idx);
}

int qo_davidson(const qo_model* m, int n_eig, int max_subspace, double tol,
                double* evals_out, double* eigvecs_out, int* iterations, int* converged)
{
    const int64_t dim = m->dim;
    *iterations = 0; *converged = 0;
    if (dim == 0) return 0;
    if (n_eig > dim) n_eig = (int)dim;
    if (max_subspace < 2 * n_eig) max_subspace = 2 * n_eig;
    const int vcap = max_subspace + n_eig + 2;
    cplx** V = (cplx**)calloc((size_t)vcap, sizeof(cplx*));
    cplx** HV = (cplx**)calloc((size_t)vcap, sizeof(cplx*));
    int nv = 0;
    /* diag order (stable: ties keep index order) */
    int64_t* di = (int64_t*)malloc(sizeof(int64_t) * (size_t)dim);
    double* dr = (double*)malloc(sizeof(double) * (size_t)dim);
    for (int64_t i = 0; i < dim; ++i) { di[i] = i; dr[i] = creal(m->diag[i]); }
    /* partial selection of the n_eig smallest (stable) */
    for (int k = 0; k < n_eig; ++k) {
        int64_t best = k;
        for (int64_t i = k + 1; i < dim; ++i)
            if (dr[di[i]] < dr[di[best]] || (dr[di[i]] == dr[di[best]] && di[i] < di[best])) best = i;
        int64_t t = di[k]; di[k] = di[best]; di[best] = t;
    }
    for (int i = 0; i < n_eig; ++i) {
        cplx* v = (cplx*)calloc((size_t)dim, sizeof(cplx));
        v[di[i]] = 1.0;
        for (int j = 0; j < nv; ++j) vaxpy(-vdot(V[j], v, dim), V[j], v, dim);
        double nr = vnorm(v, dim);
        if (nr < 1e-10) {
            for (int64_t j = 0; j < dim; ++j)
                v[j] = qo_rng_uniform(1234, 2 * (uint64_t)j) + qo_rng_uniform(1234, 2 * (uint64_t)j + 1) * I;
            for (int j = 0; j < nv; ++j) vaxpy(-vdot(V[j], v, dim), V[j], v, dim);
            nr = vnorm(v, dim);
        }
        vscal(1.0 / nr, v, dim);
        V[nv] = v;
        HV[nv] = (cplx*)malloc(sizeof(cplx) * (size_t)dim);
        apply_c(m, v, HV[nv]);
        ++nv;
    }
    free(di); free(dr);
    int ret = 0;
    for (int iter = 0; iter < 100; ++iter) {
        const int mm = nv;
        cplx* Hs = (cplx*)malloc(sizeof(cplx) * (size_t)mm * (size_t)mm);
        for (int i = 0; i < mm; ++i)
            for (int j = 0; j <= i; ++j) {
                Hs[i * mm + j] = vdot(V[i], HV[j], dim);
                if (i != j) Hs[j * mm + i] = conj(Hs[i * mm + j]);
            }
        double* ev = (double*)malloc(sizeof(double) * (size_t)mm);
        cplx* es = (cplx*)malloc(sizeof(cplx) * (size_t)mm * (size_t)mm);
        small_hermitian(mm, Hs, ev, es);
        int all_conv = 1;
        cplx* corr[64];
        int ncorr = 0;
        for (int k = 0; k < n_eig; ++k) {
            double lambda = ev[k];
            const cplx* s = es + (size_t)k * (size_t)mm;
            cplx* r = (cplx*)calloc((size_t)dim, sizeof(cplx));
            for (int i = 0; i < mm; ++i) {
                vaxpy(s[i], HV[i], r, dim);
                vaxpy(-lambda * s[i], V[i], r, dim);
            }
            double rn = vnorm(r, dim);
            if (rn > tol) {
                all_conv = 0;
                cplx* t = (cplx*)malloc(sizeof(cplx) * (size_t)dim);
                for (int64_t i = 0; i < dim; ++i) {       /* davidson.cpp:171-175 */
                    cplx diff = m->diag[i] - lambda;
                    if (cabs(diff) < 1e-10) diff = (creal(diff) >= 0) ? 1e-10 : -1e-10;
                    t[i] = r[i] / diff;
                }
                if (ncorr < 64) corr[ncorr++] = t; else free(t);
            }
            free(r);
        }
        if (all_conv || iter == 99) {
            *iterations = iter + 1;
            *converged = all_conv;
            for (int k = 0; k < n_eig; ++k) {
                evals_out[k] = ev[k];
                if (eigvecs_out) {
                    cplx* dst = (cplx*)eigvecs_out + (size_t)k * (size_t)dim;
                    memset(dst, 0, sizeof(cplx) * (size_t)dim);
                    for (int i = 0; i < mm; ++i) vaxpy(es[(size_t)k * mm + i], V[i], dst, dim);
                }
            }
            free(Hs); free(ev); free(es);
            for (int c = 0; c < ncorr; ++c) free(corr[c]);
            ret = 0;
            goto done;
        }
        if (nv + ncorr > max_subspace) {
            cplx** nV = (cplx**)calloc((size_t)vcap, sizeof(cplx*));
            cplx** nHV = (cplx**)calloc((size_t)vcap, sizeof(cplx*));
            for (int k = 0; k < n_eig; ++k) {
                nV[k] = (cplx*)calloc((size_t)dim, sizeof(cplx));
                nHV[k] = (cplx*)calloc((size_t)dim, sizeof(cplx));
                for (int i = 0; i < mm; ++i) {
                    vaxpy(es[(size_t)k * mm + i], V[i], nV[k], dim);
                    vaxpy(es[(size_t)k * mm + i], HV[i], nHV[k], dim);
                }
            }
            for (int i = 0; i < nv; ++i) { free(V[i]); free(HV[i]); }
            free(V); free(HV);
            V = nV; HV = nHV; nv = n_eig;
        }
        for (int c = 0; c < ncorr; ++c) {
            cplx* t = corr[c];
            for (int j = 0; j < nv; ++j) vaxpy(-vdot(V[j], t, dim), V[j], t, dim);
            double nr = vnorm(t, dim);
            if (nr > 1e-10 && nv < vcap) {
                vscal(1.0 / nr, t, dim);
                V[nv] = t;
                HV[nv] = (cplx*)malloc(sizeof(cplx) * (size_t)dim);
                apply_c(m, t, HV[nv]);
                ++nv;
            } else {
                free(t);
            }
        }
        free(Hs); free(ev); free(es);
    }
done:
    for (int i = 0; i < nv; ++i) { free(V[i]); free(HV[i]); }
    free(V); free(HV);
    return ret;
}

/* ------------------------------------------------------------------ */
/* counter-based RNG (splitmix64 finaliser, Steele/Lea/Flood 2014)     */
/* ------------------------------------------------------------------ */
uint64_t qo_rng_u64(uint64_t seed, uint64_t counter)
{
    uint64_t z = seed * 0xD1B54A32D192ED03ULL + (counter + 1ULL) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

double qo_rng_uniform(uint64_t seed, uint64_t counter)
{
    return (double)(qo_rng_u64(seed, counter) >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}
