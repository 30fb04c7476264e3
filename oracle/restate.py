"""CPU restatement of RandLinearAlgebra.jl's hot-path arithmetic — the oracle.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may use this module, and only as the checker.
The product path (randlinearalgebra.jl_amd/) never imports it.

Each function restates a reference loop, citing the file:line it follows
(paths relative to the RandLinearAlgebra.jl repository).  Where the reference
arithmetic is fixed by its code (the FWHT stage order, Julia's Float64
promotion of alpha/beta/scaling), the restatement reproduces it operation for
operation, so results can be compared bit for bit.  Pinning (see the tests in
tests/test_oracle.py): every function is checked against a closed form or a
mathematical invariant that a dropped term, wrong sign or transposed operand
would break (dense Hadamard / sparse products, self-inverse FWHT, Philox
known-answer vectors, Floyd/Lemire distribution checks).

Parity unpinned: the random streams (philox4x32_10, Stream, floyd_sorted and
the *_gen recipe generators, fjlt_gen included).  The reference draws with Julia's
Xoshiro/StatsBase (sparse_sign.jl:167,252; srht.jl:150-151; gaussian.jl:107),
which cannot run here and which no reference test fixes; these functions
define this build's counter-based stream, pinned only to Philox known-answer
vectors, to the device generators (bit for bit) and to distribution tests.
"""
import math

import numpy as np

# ---------------------------------------------------------------------------
# Philox4x32-10 (Salmon et al., SC'11) and the draws built on it


M0, M1 = 0xD2511F53, 0xCD9E8D57
W0, W1 = 0x9E3779B9, 0xBB67AE85
MASK = 0xFFFFFFFF


def philox4x32_10(ctr, key):
    c0, c1, c2, c3 = ctr
    k0, k1 = key
    for _ in range(10):
        p0 = M0 * c0
        p1 = M1 * c2
        hi0, lo0 = p0 >> 32, p0 & MASK
        hi1, lo1 = p1 >> 32, p1 & MASK
        c0, c1, c2, c3 = (hi1 ^ c1 ^ k0) & MASK, lo1, (hi0 ^ c3 ^ k1) & MASK, lo0
        k0, k1 = (k0 + W0) & MASK, (k1 + W1) & MASK
    return [c0, c1, c2, c3]


DOM_INDEX, DOM_SIGN, DOM_GAUSS, DOM_PERM = 1, 2, 3, 4


class Stream:
    """32-bit word stream for (seed, epoch, domain, item): counter
    {block, item_lo, item_hi ^ (domain << 24), epoch_lo ^ epoch_hi}."""

    def __init__(self, seed, epoch, domain, item):
        self.key = [seed & MASK, (seed >> 32) & MASK]
        self.ctr = [0, item & MASK, ((item >> 32) ^ (domain << 24)) & MASK,
                    (epoch & MASK) ^ ((epoch >> 32) & MASK)]
        self.buf = None
        self.pos = 4

    def next(self):
        if self.pos == 4:
            self.buf = philox4x32_10(self.ctr, self.key)
            self.ctr[0] = (self.ctr[0] + 1) & MASK
            self.pos = 0
        v = self.buf[self.pos]
        self.pos += 1
        return v

    def below(self, n):
        """Lemire's multiply-and-reject uniform integer in [0, n)."""
        x = self.next()
        m = x * n
        lo = m & MASK
        if lo < n:
            t = ((1 << 32) - n) % n
            while lo < t:
                x = self.next()
                m = x * n
                lo = m & MASK
        return m >> 32

    def open01(self):
        a = self.next() >> 5
        b = self.next() >> 6
        return ((a * 67108864.0 + b) + 0.5) * (1.0 / 9007199254740992.0)


def sign_word(seed, epoch, word):
    return Stream(seed, epoch, DOM_SIGN, word).next()


def floyd_sorted(st, n, k):
    """Floyd's uniform k-subset of [0, n), kept sorted (StatsBase
    sample!(..., replace=false, ordered=true) as called at
    src/Compressors/sparse_sign.jl:167)."""
    out = []
    for t in range(n - k, n):
        v = st.below(t + 1)
        if v in out:
            v = t
        out.append(v)
        out.sort()
    return out


def sparse_sign_gen(seed, epoch, n_groups, s, nnz):
    """sparse_idx_update! (sparse_sign.jl:155-172) + rand([-sc, sc])
    (sparse_sign.jl:248-252): 0-based rows and signed values."""
    idx = np.zeros(n_groups * nnz, dtype=np.int32)
    val = np.zeros(n_groups * nnz)
    sc = 1.0 / math.sqrt(nnz)
    for j in range(n_groups):
        idx[j * nnz:(j + 1) * nnz] = floyd_sorted(Stream(seed, epoch, DOM_INDEX, j), s, nnz)
        for q in range(nnz):
            e = j * nnz + q
            w = sign_word(seed, epoch, e >> 5)
            val[e] = sc if (w >> (e & 31)) & 1 else -sc
    return idx, val


def count_sketch_gen(seed, epoch, n_groups, s):
    """count_sketch.jl:124-127: groups = rand(1:s), signs = rand([-1, 1])."""
    idx = np.zeros(n_groups, dtype=np.int32)
    val = np.zeros(n_groups)
    for j in range(n_groups):
        idx[j] = Stream(seed, epoch, DOM_INDEX, j).below(s)
        w = sign_word(seed, epoch, j >> 5)
        val[j] = 1.0 if (w >> (j & 31)) & 1 else -1.0
    return idx, val


def srht_gen(seed, epoch, n_pad, s):
    """srht.jl:150-151: bitrand(n_pad) and sample(1:n_pad, s, replace=false)
    (Floyd's subset, then a Fisher–Yates shuffle for a uniform order)."""
    nchunks = (n_pad + 63) // 64
    signs = np.zeros(n_pad, dtype=bool)
    for c in range(nchunks):
        lo = sign_word(seed, epoch, 2 * c)
        hi = sign_word(seed, epoch, 2 * c + 1)
        ch = lo | (hi << 32)
        for b in range(64):
            i = 64 * c + b
            if i < n_pad:
                signs[i] = bool((ch >> b) & 1)
    st = Stream(seed, epoch, DOM_INDEX, 0)
    chosen = set()
    op = []
    for t in range(n_pad - s, n_pad):
        v = st.below(t + 1)
        if v in chosen:
            v = t
        chosen.add(v)
        op.append(v)
    sp = Stream(seed, epoch, DOM_PERM, 0)
    for i in range(s, 1, -1):
        j = sp.below(i)
        op[i - 1], op[j] = op[j], op[i - 1]
    return np.array(op, dtype=np.int32), signs


# ---------------------------------------------------------------------------
# Walsh–Hadamard transform: src/Compressors/helpers/fwht.jl:21-118


def fwht(x, signs=None, scaling=1.0, dtype=None):
    """In-place FWHT of the columns of x (n x k or length n), natural order,
    stages h = 1, 2, 4, ... as fwht.jl:21-70.  Stage 1 evaluates
    x[i] * (signs[i] ? scaling : -scaling) and the sums in Float64 (scaling is
    Float64 in the reference), the rest in the element type."""
    x = np.asarray(x)
    dtype = x.dtype if dtype is None else np.dtype(dtype)
    two_d = x.ndim == 2
    X = x.reshape(x.shape[0], -1).astype(dtype, copy=True)
    n = X.shape[0]
    if n < 2 or (n & (n - 1)):
        raise ValueError("Size of vector must be power of 2.")
    if signs is None:
        s = np.full(n, float(scaling))
    else:
        signs = np.asarray(signs, dtype=bool)
        if signs.shape[0] != n:
            raise ValueError("Vector `x` and vector `signs` must be same length.")
        s = np.where(signs, float(scaling), -float(scaling))
    z = X[0::2].astype(np.float64) * s[0::2, None]
    y = X[1::2].astype(np.float64) * s[1::2, None]
    X[0::2] = (z + y).astype(dtype)
    X[1::2] = (z - y).astype(dtype)
    h = 2
    while h < n:
        V = X.reshape(n // (2 * h), 2, h, -1)
        a = V[:, 0].copy()
        b = V[:, 1].copy()
        V[:, 0] = a + b
        V[:, 1] = a - b
        h *= 2
    return X if two_d else X[:, 0]


def padded_size(n):
    """srht.jl:147 Int64(2^(ceil(log(2, n))))."""
    return int(2.0 ** math.ceil(math.log(float(n)) / math.log(2.0)))


def srht_left_mul(A, op, signs, scale, dtype, C=None, alpha=1.0, beta=0.0):
    """mul!(C, S::SRHTRecipe{Left}, A, alpha, beta) — srht.jl:220-286: zero
    padding, fwht!(col, signs; scaling=scale) per column, then
    Cv .= beta .* Cv .+ alpha .* padding[op, :] (Float64 broadcast)."""
    A = np.asarray(A)
    m = A.shape[0]
    k = A.shape[1] if A.ndim == 2 else 1
    n_pad = np.asarray(signs).shape[0]
    P = np.zeros((n_pad, k), dtype=dtype)
    P[:m] = A.reshape(m, k)
    P = fwht(P, signs, scale, dtype)
    G = P[np.asarray(op)].astype(np.float64)
    Cd = np.zeros_like(G) if C is None else np.asarray(C, dtype=np.float64).reshape(G.shape)
    return (beta * Cd + alpha * G).astype(dtype)


def srht_adjoint_mul(X, op, signs, scale, m, dtype, Y=None, alpha=1.0, beta=0.0):
    """Y = alpha*T'*X + beta*Y with T = (scale*H*D*pad)[op, :] — the A*S
    orientation of a Left recipe (srht.jl:289-363) on A' (and S*A of a Right
    recipe, srht.jl:432-504): padding[op, :] = alpha .* X, fwht!(col;
    scaling=scale) without signs, rows multiplied by +-1, first m rows,
    axpby!(1, P, beta, Y)."""
    X = np.asarray(X)
    s = X.shape[0]
    k = X.shape[1] if X.ndim == 2 else 1
    n_pad = np.asarray(signs).shape[0]
    P = np.zeros((n_pad, k), dtype=dtype)
    P[np.asarray(op)] = (alpha * X.reshape(s, k).astype(np.float64)).astype(dtype)
    P = fwht(P, None, scale, dtype)
    sg = np.where(np.asarray(signs, dtype=bool), 1, -1).astype(dtype)
    P = P * sg[:, None]
    Yd = np.zeros((m, k)) if Y is None else np.asarray(Y, dtype=np.float64).reshape(m, k)
    out = P[:m].astype(np.float64) if beta == 0.0 else P[:m].astype(np.float64) + beta * Yd
    return out.astype(dtype)


# ---------------------------------------------------------------------------
# sparse x dense: the SparseArrays loop behind sparse_sign.jl:347,359 and
# count_sketch.jl:174-211


def csr_from_dense(M):
    """CSR of a dense matrix with column indices ascending within each row."""
    M = np.asarray(M)
    ptr = [0]
    idx, val = [], []
    for i in range(M.shape[0]):
        nz = np.nonzero(M[i])[0]
        idx.extend(nz.tolist())
        val.extend(M[i, nz].tolist())
        ptr.append(len(idx))
    return np.array(ptr, dtype=np.int64), np.array(idx, dtype=np.int64), np.array(val)


def sparse_rowgather(ptr, idx, val, X, alpha=1.0, beta=0.0, C=None, dtype=None):
    """C = alpha*op(M)*X + beta*C where row i of op(M) has nonzeros
    (idx[p], val[p]) for p in ptr[i]:ptr[i+1].  Per output entry, as in
    SparseArrays' `_spmatmul!`: C is first filled with 0 (beta == 0) or scaled
    by beta, then C[i,k] += nzv*(X[j,k]*alpha) for the nonzeros in ascending
    j, each step in Float64 (alpha is a Float64) rounded to the element type."""
    X = np.asarray(X)
    dtype = X.dtype if dtype is None else np.dtype(dtype)
    X2 = X.reshape(X.shape[0], -1)
    nout = len(ptr) - 1
    k = X2.shape[1]
    if beta == 0.0 or C is None:
        acc = np.zeros((nout, k), dtype=dtype)
    else:
        acc = (beta * np.asarray(C, dtype=np.float64).reshape(nout, k)).astype(dtype)
    ptr = np.asarray(ptr)
    lens = np.diff(ptr)
    Xd = X2.astype(dtype).astype(np.float64) * alpha
    valv = np.asarray(val).astype(dtype).astype(np.float64)
    for q in range(int(lens.max()) if nout else 0):
        rows = np.nonzero(lens > q)[0]
        p = ptr[rows] + q
        term = valv[p][:, None] * Xd[np.asarray(idx)[p]]
        acc[rows] = (acc[rows].astype(np.float64) + term).astype(dtype)
    return acc if X.ndim == 2 else acc[:, 0]


def sparse_mul(M, X, alpha=1.0, beta=0.0, C=None, dtype=None):
    """C = alpha*M*X + beta*C for a dense-stored sparse M, SparseArrays order."""
    ptr, idx, val = csr_from_dense(M)
    return sparse_rowgather(ptr, idx, val, X, alpha, beta, C, dtype)


# ---------------------------------------------------------------------------
# FJLT: src/Compressors/fjlt.jl


DOM_FJLT_ROWS, DOM_FJLT_VALS = 6, 7


def _fjlt_block(seed, epoch, domain, item, b):
    ctr = (b & MASK, item & MASK, ((item >> 32) ^ (domain << 24)) & MASK,
           (epoch & MASK) ^ ((epoch >> 32) & MASK))
    return philox4x32_10(ctr, (seed & MASK, (seed >> 32) & MASK))


def fjlt_gen(seed, epoch, s, n_pad, q):
    """This build's stream for K = sprandn(s, n_pad, q) (fjlt.jl:183): entry
    (i, j) nonzero iff word i of stream (seed, epoch, 6, j) is below
    floor(q*2^32) (all for q = 1); the r-th nonzero of column j takes
    Box–Muller pair r//2 of stream (seed, epoch, 7, j).  Returns the CSC
    arrays (colptr, rowidx, val) with rows ascending.  Parity unpinned (a
    random stream; see the module header)."""
    thr = (1 << 32) if q >= 1.0 else int(math.floor(q * 4294967296.0))
    colptr, rowidx, val = [0], [], []
    for j in range(n_pad):
        r = 0
        for i in range(s):
            w = _fjlt_block(seed, epoch, DOM_FJLT_ROWS, j, i >> 2)[i & 3]
            if w >= thr:
                continue
            v = _fjlt_block(seed, epoch, DOM_FJLT_VALS, j, r >> 1)
            u1 = ((v[0] >> 5) * 67108864.0 + (v[1] >> 6) + 0.5) * (1.0 / 9007199254740992.0)
            u2 = ((v[2] >> 5) * 67108864.0 + (v[3] >> 6) + 0.5) * (1.0 / 9007199254740992.0)
            rad = math.sqrt(-2.0 * math.log(u1))
            a = 6.283185307179586 * u2
            rowidx.append(i)
            val.append(rad * math.sin(a) if r & 1 else rad * math.cos(a))
            r += 1
        colptr.append(len(rowidx))
    return (np.array(colptr, dtype=np.int64), np.array(rowidx, dtype=np.int64),
            np.array(val, dtype=np.float64))


def csc_dense(nrows, ncols, colptr, rowidx, val):
    M = np.zeros((nrows, ncols))
    for j in range(ncols):
        M[np.asarray(rowidx[colptr[j]:colptr[j + 1]]), j] = val[colptr[j]:colptr[j + 1]]
    return M


def fjlt_left_mul(A, K, signs, scale, dtype, C=None, alpha=1.0, beta=0.0):
    """mul!(C, S::FJLTRecipe{Left}, A, alpha, beta) — fjlt.jl:257-319: zero
    padding, fwht!(col, signs; scaling=scale) per column, then
    mul!(Cv, S.op, padding, alpha, beta) in the SparseArrays loop order.
    K is the dense-stored s x n_pad factor."""
    A = np.asarray(A)
    m = A.shape[0]
    k = A.shape[1] if A.ndim == 2 else 1
    n_pad = np.asarray(signs).shape[0]
    P = np.zeros((n_pad, k), dtype=dtype)
    P[:m] = A.reshape(m, k)
    P = fwht(P, signs, scale, dtype)
    return sparse_mul(K, P, alpha, beta, C, dtype)


def fjlt_adjoint_mul(X, K, signs, scale, m, dtype, Y=None, alpha=1.0, beta=0.0):
    """Y = alpha*T'*X + beta*Y with T = scale*K*H*D*pad — the A*S orientation
    of a Left recipe (fjlt.jl:321-393) on A' (and S*A of a Right recipe,
    fjlt.jl:470-535): P = alpha*K'*X (mul!(pv', Av, S.op, alpha, 0)),
    fwht!(col; scaling=scale) without signs, rows multiplied by +-1, first m
    rows, axpby!(1, P, beta, Y)."""
    X = np.asarray(X)
    s = X.shape[0]
    k = X.shape[1] if X.ndim == 2 else 1
    P = sparse_mul(np.asarray(K).T, X.reshape(s, k), alpha, 0.0, None, dtype)
    P = fwht(P, None, scale, dtype)
    sg = np.where(np.asarray(signs, dtype=bool), 1, -1).astype(dtype)
    P = P * sg[:, None]
    Yd = np.zeros((m, k)) if Y is None else np.asarray(Y, dtype=np.float64).reshape(m, k)
    out = P[:m].astype(np.float64) if beta == 0.0 else P[:m].astype(np.float64) + beta * Yd
    return out.astype(dtype)


# ---------------------------------------------------------------------------
# RangeFinder / RandSVD: src/Approximators/RangeApproximators/helpers/power_its.jl


def _q(Y):
    """Thin Q of Y with a positive-diagonal R (LAPACK's Q up to column signs)."""
    Q, Rr = np.linalg.qr(Y)
    s = np.sign(np.diag(Rr))
    s[s == 0] = 1
    return Q * s


def rand_power_it(A, Omega, power_its):
    """power_its.jl:15-36: Y = A*S; power_its x (Z = A'Y; Y = A Z); Q = qr(Y)."""
    Y = A @ Omega
    for _ in range(power_its):
        Z = A.T @ Y
        Y = A @ Z
    return _q(Y)


def rand_ortho_it(A, Omega, power_its):
    """power_its.jl:53-80: a QR after every product."""
    Q = _q(A @ Omega)
    for _ in range(power_its):
        Q = _q(A.T @ Q)
        Q = _q(A @ Q)
    return Q


def randsvd(A, Q):
    """randsvd.jl:174-189: QA = Q'A (Float64), U, S, V = svd(QA), U = Q*U."""
    Ub, S, Vt = np.linalg.svd(Q.T @ A, full_matrices=False)
    return Q @ Ub, S, Vt.T


# ---------------------------------------------------------------------------
# Kaczmarz: src/Solvers/kaczmarz.jl, SubSolvers/lq.jl, ErrorMethods/full_residual.jl


def lq_min_norm(At, r):
    """ldiv!(u, qr!(At')', r) of lq.jl:35-46: with At' = Q R (thin), the
    minimum-norm solution of At u = r is u = Q (R' \\ r)."""
    Q, Rr = np.linalg.qr(np.asarray(At, dtype=np.float64).T)
    z = np.linalg.solve(Rr.T, np.asarray(r, dtype=np.float64))
    return Q @ z


def kaczmarz_update_block(x, At, bt, alpha):
    """kaczmarz_update_block! (kaczmarz.jl:294-311): r = bt - At x, u from the
    LQ sub-solver, x + alpha u.  A zero block leaves x unchanged (:300)."""
    x = np.asarray(x, dtype=np.float64)
    At = np.asarray(At, dtype=np.float64)
    if not At.any():
        return x.copy()
    r = np.asarray(bt, dtype=np.float64) - At @ x
    return x + alpha * lq_min_norm(At, r)


def kaczmarz_update(x, a, c, alpha):
    """kaczmarz_update! (kaczmarz.jl:262-276): x - alpha (a.x - c)/||a||^2 a."""
    x = np.asarray(x, dtype=np.float64)
    a = np.asarray(a, dtype=np.float64).ravel()
    if not a.any():
        return x.copy()
    scaling = alpha * (a @ x - float(np.asarray(c).ravel()[0])) / (a @ a)
    return x - scaling * a


def full_residual(A, x, b):
    """compute_error of full_residual.jl:32-41: ||b - A x||_2 (FP64)."""
    return float(np.linalg.norm(np.asarray(b, np.float64) - np.asarray(A, np.float64) @
                                np.asarray(x, np.float64)))


def ihs_step(A, b, x, SA, alpha=1.0):
    """One iterative-Hessian-sketch update (ihs.jl:223-252) for a given
    sketched matrix SA (s x n): r = b - A x, g = A' r, R from qr(SA),
    u = R^-1 R'^-1 g, returns x + alpha u.  FP64."""
    A, SA = np.asarray(A, np.float64), np.asarray(SA, np.float64)
    b, x = np.asarray(b, np.float64), np.asarray(x, np.float64)
    g = A.T @ (b - A @ x)
    Rf = np.linalg.qr(SA, mode="r")
    y = np.linalg.solve(Rf.T, g)    # R' y = g
    u = np.linalg.solve(Rf, y)      # R u = y
    return x + alpha * u


def colproj_step(A, b, x, S, alpha=1.0):
    """One column-projection update (col_projection.jl:265-310, 356-364) for a
    given compression S (n x k): r = b - A x, v = argmin ||(A S) v - r|| (the
    minimum-norm least-squares solution, as the QR sub-solver's ldiv!),
    returns x + alpha S v.  FP64."""
    A, S = np.asarray(A, np.float64), np.asarray(S, np.float64).reshape(A.shape[1], -1)
    b, x = np.asarray(b, np.float64), np.asarray(x, np.float64)
    r = b - A @ x
    v = np.linalg.lstsq(A @ S, r, rcond=None)[0]
    return x + alpha * (S @ v)


def ls_gradient(A, x, b):
    """compute_error of ls_gradient.jl:39-51: ||-A'(b - A x)||_2 (FP64)."""
    A = np.asarray(A, np.float64)
    return float(np.linalg.norm(A.T @ (np.asarray(b, np.float64) - A @ np.asarray(x, np.float64))))


def agmon_select(v, cand, k):
    """The top-k step of sample_distribution!(::AgmonRecipe)
    (agmon.jl:400-414): partialsortperm of the candidates by
    (-|v[c]|, c); returns the chosen candidates (0-based) in that order."""
    v = np.abs(np.asarray(v, np.float64))
    cand = np.asarray(cand, np.int64)
    order = sorted(range(len(cand)), key=lambda i: (-v[cand[i]], cand[i]))
    return [int(cand[i]) for i in order[:k]]


def successive_inclusion(w, k):
    """Inclusion probability of each item in a weighted sample of k without
    replacement drawn item by item (P(next = i) = w_i / remaining weight),
    the law of StatsBase wsample!(replace=false); exact enumeration, small n."""
    w = np.asarray(w, dtype=np.float64)
    n = len(w)
    incl = np.zeros(n)

    def rec(chosen, prob, depth):
        if depth == k:
            for i in chosen:
                incl[i] += prob
            return
        rest = w.sum() - sum(w[i] for i in chosen)
        for i in range(n):
            if i not in chosen and w[i] > 0:
                rec(chosen | {i}, prob * w[i] / rest, depth + 1)

    rec(frozenset(), 1.0, 0)
    return incl
