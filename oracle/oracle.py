"""numpy front end of the CPU oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; the product never imports it.

* `lib()` loads liboracle.so, the C restatement (oracle.c, citations in
  oracle.h), built by oracle/Makefile.
* `host_dgemm` is the host BLAS dgemm the reference validates against
  (cblas_dgemm in src/Wrappers/cuBLASMp/cuBLASMpValidator.cpp:48,
  src/Wrappers/Cosma/cosmaValidators.cpp:52): here numpy's BLAS (OpenBLAS).
* `ref_lib()` loads _ref/libref_validation.so, the reference's own
  src/Utilities/Validation.cpp compiled by the Makefile, when present.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB selects another build of the same C source (the sanitizer build,
# `make -C oracle sanitize`); the default is the plain liboracle.so.
LIB_PATH = os.environ.get("ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_validation.so")

_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double
_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_ip = ctypes.POINTER(ctypes.c_int)
_chr = ctypes.c_char

_SIGS = {
    "or_numroc": (_i64, [_i64, _i64, _int, _int, _int]),
    "or_virtual_grid": (None, [_int, _ip, _ip]),
    "or_block_splits": (None, [_i64, _int, _i64p]),
    "or_copy_block": (None, [_i64, _i64, _dp, _i64, _dp, _i64]),
    "or_pack_block2d": (None, [_dp, _i64, _i64, _i64, _i64p, _int, _i64p, _int, _dp]),
    "or_unpack_block2d": (None, [_dp, _i64p, _int, _i64p, _int, _dp, _i64, _i64, _i64]),
    "or_pack_block_cyclic": (None, [_dp, _i64, _i64, _i64, _i64, _i64, _int, _int, _dp]),
    "or_unpack_block_cyclic": (None, [_dp, _i64, _i64, _int, _int, _dp, _i64, _i64, _i64]),
    "or_cannon_simulate": (_int, [_int, _int, _ip, _ip]),
    "or_dgemm": (None, [_chr, _chr, _i64, _i64, _i64, _dbl, _dp, _i64, _dp, _i64, _dbl, _dp, _i64]),
    "or_cannon25d": (None, [_chr, _chr, _i64, _i64, _i64, _dbl, _dp, _dp, _dbl, _dp, _int, _int, _int]),
    "or_summa": (None, [_i64, _i64, _i64, _dbl, _dp, _dp, _dbl, _dp, _i64, _i64, _i64, _int, _int]),
    "or_summa_op": (None, [_chr, _chr, _i64, _i64, _i64, _dbl, _dp, _dp, _dbl, _dp, _i64, _i64, _i64, _int, _int]),
    "or_layer_reduce": (None, [_dp, _dp, _i64, _int]),
    "or_block_seq_tiles": (None, [_int, _i64, _i64, _i64, _int, _int, _dp, _dp, _dp, _dp, _dp, _dp]),
    "or_Derror": (_dbl, [_dbl, _dbl]),
    "or_Dvec_diff": (ctypes.c_long, [_dp, _dp, _i64, _dbl]),
    "or_Dtest_equality": (ctypes.c_short, [_dp, _dp, _i64]),
    "or_splitmix64_next": (ctypes.c_uint64, [ctypes.POINTER(ctypes.c_uint64)]),
    "or_generate": (None, [ctypes.c_uint64, _int, _dp, _i64]),
}

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ref_lib():
    """The reference's compiled Validation.cpp, or None if not built."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        R = ctypes.CDLL(REF_PATH)
        R.ref_Dtest_equality.restype = ctypes.c_short
        R.ref_Dtest_equality.argtypes = [_dp, _dp, ctypes.c_longlong]
        R.ref_Dvec_diff.restype = ctypes.c_long
        R.ref_Dvec_diff.argtypes = [_dp, _dp, ctypes.c_longlong, _dbl]
        _ref = R
    return _ref


def _p(a):
    if a.dtype != np.float64 or not (a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"]):
        raise TypeError("oracle arrays must be contiguous float64")
    return a.ctypes.data_as(_dp)


def _offs(seq):
    return (ctypes.c_int64 * len(seq))(*[int(v) for v in seq])


def _op(c):
    return c.encode() if isinstance(c, str) else c


# ---------------------------------------------------------------- index maps

def numroc(n, nb, iproc, isrc, nprocs):
    return lib().or_numroc(n, nb, iproc, isrc, nprocs)


def virtual_grid(size):
    r, c = _int(), _int()
    lib().or_virtual_grid(size, ctypes.byref(r), ctypes.byref(c))
    return r.value, c.value


def block_splits(n, parts):
    out = (ctypes.c_int64 * (parts + 1))()
    lib().or_block_splits(n, parts, out)
    return list(out)


def cannon_simulate(p, c):
    """(mA, mB) arrays of shape (c, p, p, p//c): k-tile index held at each step."""
    steps = p // c
    a = np.zeros(c * p * p * steps, dtype=np.int32)
    b = np.zeros_like(a)
    lib().or_cannon_simulate(p, c, a.ctypes.data_as(_ip), b.ctypes.data_as(_ip))
    return a.reshape(c, p, p, steps), b.reshape(c, p, p, steps)


# ---------------------------------------------------------------- layouts
# Matrices are flat float64 arrays in column-major order.

def pack_block2d(G, ldg, rows, cols, row_off, col_off):
    out = np.zeros(rows * cols, dtype=np.float64)
    lib().or_pack_block2d(_p(G), ldg, rows, cols, _offs(row_off), len(row_off) - 1, _offs(col_off),
                          len(col_off) - 1, _p(out))
    return out


def unpack_block2d(packed, row_off, col_off, G, ldg, rows, cols):
    lib().or_unpack_block2d(_p(packed), _offs(row_off), len(row_off) - 1, _offs(col_off), len(col_off) - 1,
                            _p(G), ldg, rows, cols)
    return G


def block_cyclic_size(rows, cols, mb, nb, P, Q):
    return sum(numroc(rows, mb, pr, 0, P) * numroc(cols, nb, pc, 0, Q) for pr in range(P) for pc in range(Q))


def pack_block_cyclic(G, ldg, rows, cols, mb, nb, P, Q):
    out = np.zeros(block_cyclic_size(rows, cols, mb, nb, P, Q), dtype=np.float64)
    lib().or_pack_block_cyclic(_p(G), ldg, rows, cols, mb, nb, P, Q, _p(out))
    return out


def unpack_block_cyclic(packed, mb, nb, P, Q, G, ldg, rows, cols):
    lib().or_unpack_block_cyclic(_p(packed), mb, nb, P, Q, _p(G), ldg, rows, cols)
    return G


def block_seq_tiles(col_major, M, N, K, size, gA, gB, gC):
    """deliverMatrix's tiles of every rank (rank order, concatenated) for
    flat global matrices in column-major or row-major storage."""
    dRow, dCol = virtual_grid(size)
    sizes = [(M // dRow + (M % dRow if i == dRow - 1 else 0), N // dCol + (N % dCol if j == dCol - 1 else 0))
             for i in range(dRow) for j in range(dCol)]
    tA = np.zeros(max(1, sum(tm * K for tm, _ in sizes)))
    tB = np.zeros(max(1, sum(K * tn for _, tn in sizes)))
    tC = np.zeros(max(1, sum(tm * tn for tm, tn in sizes)))
    lib().or_block_seq_tiles(int(col_major), M, N, K, dRow, dCol, _p(gA), _p(gB), _p(gC), _p(tA), _p(tB), _p(tC))
    return tA, tB, tC, sizes


def layer_reduce(C, partial, n, nlayers):
    lib().or_layer_reduce(_p(C), _p(partial), n, nlayers)
    return C


# ---------------------------------------------------------------- arithmetic

def dgemm_loops(opA, opB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    """Textbook triple loop (long double sums), for small sizes."""
    lib().or_dgemm(_op(opA), _op(opB), m, n, k, alpha, _p(A), lda, _p(B), ldb, beta, _p(C), ldc)
    return C


def host_dgemm(opA, opB, alpha, A, B, beta, C):
    """C = alpha*op(A) op(B) + beta*C with numpy's BLAS on 2-D arrays (the
    host dgemm oracle).  beta == 0 ignores C's contents."""
    a = A.T if opA in ("T", "t", "C", "c") else A
    b = B.T if opB in ("T", "t", "C", "c") else B
    prod = alpha * (a @ b)
    if beta == 0.0:
        return prod
    return prod + beta * C


def cannon25d(opA, opB, M, N, K, alpha, Ap, Bp, beta, Cp, p, c, sched_std=False):
    lib().or_cannon25d(_op(opA), _op(opB), M, N, K, alpha, _p(Ap), _p(Bp), beta, _p(Cp), p, c, int(sched_std))
    return Cp


def summa(M, N, K, alpha, Ap, Bp, beta, Cp, mb, nb, kb, P, Q, opA="N", opB="N"):
    if opA == "N" and opB == "N":
        lib().or_summa(M, N, K, alpha, _p(Ap), _p(Bp), beta, _p(Cp), mb, nb, kb, P, Q)
    else:
        lib().or_summa_op(_op(opA), _op(opB), M, N, K, alpha, _p(Ap), _p(Bp), beta, _p(Cp), mb, nb, kb, P, Q)
    return Cp


def rel_frobenius(x, ref):
    """||x - ref||_F / ||ref||_F."""
    x = np.asarray(x, dtype=np.float64).ravel()
    ref = np.asarray(ref, dtype=np.float64).ravel()
    den = np.linalg.norm(ref)
    return float(np.linalg.norm(x - ref) / (den if den > 0 else 1.0))


# ---------------------------------------------------------------- validation

def Derror(a, b):
    return lib().or_Derror(a, b)


def Dvec_diff(a, b, eps):
    return lib().or_Dvec_diff(_p(a), _p(b), a.size, eps)


def Dtest_equality(C_comp, C):
    return lib().or_Dtest_equality(_p(C_comp), _p(C), C_comp.size)


# ---------------------------------------------------------------- generation

def splitmix64_stream(seed, n):
    s = ctypes.c_uint64(seed)
    return [lib().or_splitmix64_next(ctypes.byref(s)) for _ in range(n)]


def generate(seed, dist, n):
    out = np.empty(n, dtype=np.float64)
    lib().or_generate(seed, dist, _p(out), n)
    return out


def generate_f32(seed, dist, n):
    """FP32 form of `generate` (generateMatrixGPU<float>, cmatrix.cpp:301-309):
    the FP64 values rounded to nearest, ties to even (IEEE conversion)."""
    return generate(seed, dist, n).astype(np.float32)


# ---------------------------------------------------------------- Ozaki split DGEMM
# Ozaki, Ogita, Oishi, Rump, "Error-free transformations of matrix
# multiplication by using fast routines of matrix multiplication and its
# applications", Numer. Algorithms 59 (2012); Ootomo, Ozaki, Yokota, "DGEMM on
# integer matrix multiplication unit", IJHPCA (2024).  New relative to the
# reference, which has no split DGEMM.

OZ_INT8_BITS = 7     # |digit| <= 127: one int8 slice carries 7 bits below the row scale
OZ_BF16_BITS = 5     # |digit| <= 31: bf16 holds it exactly and products sum exactly in fp32


def oz_exponents(X, axis):
    """Power-of-two scale per row (axis=1) or column (axis=0): the frexp
    exponent e of max|x| (max|x| < 2^e), 0 for an all-zero line."""
    amax = np.max(np.abs(X), axis=axis)
    _, e = np.frexp(amax)
    e = np.where(amax == 0, 0, e).astype(np.int32)
    return e


def oz_slices(X, e, axis, s, bits):
    """Digits d_1..d_s with X ~= 2^e * sum_t d_t 2^(-bits*t), by truncation:
    r_0 = X*2^-e, d_t = trunc(r_{t-1}*2^bits), r_t = r_{t-1}*2^bits - d_t
    (every step exact in FP64).  Returns an (s, *X.shape) float64 array of
    integers."""
    shape = [1, 1]
    shape[1 - axis] = -1
    r = np.ldexp(X, -e.reshape(shape))
    out = np.empty((s,) + X.shape)
    scale = float(1 << bits)
    for t in range(s):
        r = r * scale
        d = np.trunc(r)
        out[t] = d
        r = r - d
    return out


def oz_flush_every(K, bits):
    """Slice products whose exact integer sums share one 32-bit accumulator."""
    dmax = (1 << bits) - 1
    if bits == OZ_BF16_BITS:
        return max(0, (1 << 24) // (dmax * dmax * K)) if K else 1
    return max(0, (2 ** 31 - 1) // (dmax * dmax * K)) if K else 1


def oz_kchunk(K, bits):
    """K chunk of one exact product: all of K for INT8 digits; for BF16 digits
    the largest multiple of the 64-element k-step with 31^2 * chunk < 2^24."""
    if bits != OZ_BF16_BITS:
        return K if K else 1
    dmax = (1 << bits) - 1
    cap = (1 << 24) // (dmax * dmax)
    kp = -(-K // 64) * 64
    return kp if kp <= cap else cap // 64 * 64


def ozaki_dgemm(a, b, s, alpha=1.0, beta=0.0, c=None, bits=OZ_INT8_BITS, kchunk=None):
    """C = alpha * A B + beta * C through s slices per operand and the
    triangular product set t + u <= s + 1, on 2-D float64 arrays a (M x K)
    and b (K x N).  Accumulation order (shared with the GPU kernel, so the
    results agree bit for bit): for level l = 2..s+1, for t = 1..l-1, add the
    exact integer sum of A_t B_(l-t) to a 32-bit accumulator; every
    `flush` products (and at the end of the level) add ldexp(acc, -bits*l)
    into the FP64 total.  For the BF16 digits the K loop is also cut into
    chunks of `kchunk`, each chunk flushed separately (fp32 sums stay exact
    below 2^24)."""
    M, K = a.shape
    N = b.shape[1]
    e = oz_exponents(a, 1)
    f = oz_exponents(b, 0)
    As = oz_slices(a, e, 1, s, bits)
    Bs = oz_slices(b, f, 0, s, bits)
    if kchunk is None:
        kchunk = oz_kchunk(K, bits)
    kchunk = max(1, min(kchunk, K if K else 1))
    flush = oz_flush_every(min(K, kchunk), bits)
    if flush < 1:
        raise ValueError("K too large for exact 32-bit slice sums")
    acc = np.zeros((M, N))
    for l in range(2, s + 2):
        for k0 in range(0, max(K, 1), kchunk):
            k1 = min(K, k0 + kchunk)
            S = np.zeros((M, N))
            cnt = 0
            for t in range(1, l):
                u = l - t
                S += As[t - 1][:, k0:k1] @ Bs[u - 1][k0:k1, :]   # exact: integers < 2^53
                cnt += 1
                if cnt == flush or t == l - 1:
                    acc += np.ldexp(S, -bits * l)
                    S[:] = 0.0
                    cnt = 0
    prod = np.ldexp(acc, e.reshape(-1, 1) + f.reshape(1, -1))
    out = alpha * prod
    if beta != 0.0:
        out = out + beta * c
    return out


# ---------------------------------------------------------------- Ozaki scheme II (CRT)
# Ozaki, Uchino, Imamura, "Ozaki Scheme II: A GEMM-oriented emulation of
# floating-point matrix multiplication using an integer modular technique"
# (2025).  Rows of A / columns of B are scaled by powers of two to integers
# (rounded to nearest) whose exact integer product is computed modulo N
# pairwise coprime moduli <= 256 (one INT8 GEMM each) and rebuilt by the
# Chinese remainder theorem.  The scaling follows the paper's Cauchy-Schwarz
# ("fast mode") bound: each scaled row of A and column of B has 2-norm at
# most 2^Q, so |A'B'| <= 2^(2Q) <= P/4; the norm is bounded by an exact
# integer sum (crt_bound_sums), so the GPU and this restatement pick the same
# shifts whatever their summation order.

CRT_MODULI = (256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197, 193, 191, 181, 179, 173)


def crt_Q(K, nmod):
    """Largest Q with 2^(2Q) + ceil(sqrt(K)) 2^Q + K <= P/4 (P = product of
    the first nmod moduli): rows and columns of 2-norm <= 2^Q before the
    rounding to integers (which adds at most 1/2 per entry) then keep
    |A'B'| <= P/4.  None when no Q >= 1 fits."""
    import math
    P = 1
    for p in CRT_MODULI[:nmod]:
        P *= p
    r = math.isqrt(max(K, 0))
    rk = r if r * r == K else r + 1
    Q = P.bit_length() // 2
    while Q >= 1 and 4 * (2 ** (2 * Q) + rk * 2 ** Q + K) > P:
        Q -= 1
    return Q if Q >= 1 else None


def crt_bound_sums(X, e, axis):
    """Per row (axis=1) or column (axis=0) with max-exponent e (max|x| < 2^e):
    an exact integer upper bound S of 2^32 * sum_k (x_k 2^-e)^2.  Entry
    |x| = m 2^ex (m in [0.5, 1)) counts as ceil(256 m)^2 * 2^(16 - 2d),
    d = e - ex >= 0, rounded up to an integer (so at least 1) when 2d > 16;
    zeros count 0."""
    shape = [1, 1]
    shape[1 - axis] = -1
    m, ex = np.frexp(np.abs(X))
    M = np.ceil(m * 256.0).astype(np.int64)
    M2 = M * M
    d = e.reshape(shape).astype(np.int64) - ex.astype(np.int64)
    sh = 2 * d - 16
    up = np.left_shift(M2, np.clip(-sh, 0, 16))
    mid = np.right_shift(M2 + np.left_shift(np.int64(1), np.clip(sh, 0, 16)) - 1, np.clip(sh, 0, 16))
    t = np.where(sh <= 0, up, np.where(sh >= 17, 1, mid))
    t = np.where(X == 0, 0, t)
    return t.sum(axis=axis).astype(np.int64)


CRT_KCHUNK = 64   # k-chunk of the line bound scan


def crt_line_bounds(X, axis, kc=CRT_KCHUNK):
    """Line exponent e and norm bound S per row (axis=1) or column (axis=0),
    from one pass over k-chunks of kc entries: chunk c gets its own max
    exponent e_c (max|x| < 2^e_c) and the exact integer bound S_c of
    crt_bound_sums at e_c; the line takes e = max e_c over its nonzero
    chunks and S = sum_c ceil(S_c / 4^(e - e_c)), still an integer upper
    bound of 2^32 * sum_k (x_k 2^-e)^2 whatever the summation order.  An
    all-zero line gets e = 0, S = 0.  Returns (e int32, S int64)."""
    Xl = np.asarray(X, dtype=np.float64)
    if axis == 0:
        Xl = Xl.T
    lines, K = Xl.shape
    nch = max(1, -(-K // kc))
    Xc = np.zeros((lines, nch * kc))
    Xc[:, :K] = Xl
    Xc = Xc.reshape(lines * nch, kc)
    amax = np.max(np.abs(Xc), axis=1)
    _, ec = np.frexp(amax)
    ec = ec.astype(np.int64)
    Sc = crt_bound_sums(Xc, ec, 1).reshape(lines, nch)
    nz = (amax != 0).reshape(lines, nch)
    ec = ec.reshape(lines, nch)
    e = np.where(nz.any(axis=1), np.max(np.where(nz, ec, np.iinfo(np.int64).min), axis=1), 0)
    d2 = np.where(nz, 2 * (e.reshape(-1, 1) - ec), 0)
    small = np.right_shift(Sc + np.left_shift(np.int64(1), np.minimum(d2, 62)) - 1, np.minimum(d2, 62))
    t = np.where(d2 >= 62, np.where(Sc > 0, 1, 0), small)
    S = np.where(nz, t, 0).sum(axis=1).astype(np.int64)
    return e.astype(np.int32), S


def crt_shifts(e, S, Q):
    """The scaling as an exponent e' = Q - sigma per line (the form crt_scale
    and the reconstruction take): sigma is the largest shift with
    2^(2(sigma + e)) S <= 2^(2Q + 32), i.e. sigma = Q + 16 - e - j with
    j = ceil(log4 S) exactly, capped at 53 - e (|x 2^sigma| < 2^53, integers
    exact in FP64); an all-zero line (S = 0) keeps sigma = 0."""
    Sl = [int(v) for v in np.asarray(S).ravel()]
    j = np.array([(max(v - 1, 0).bit_length() + 1) // 2 for v in Sl], dtype=np.int64)
    el = np.asarray(e).astype(np.int64).ravel()
    sig = np.minimum(53 - el, Q + 16 - el - j)
    sig = np.where(np.array(Sl, dtype=object) == 0, 0, sig).astype(np.int64)
    return (Q - sig).astype(np.int64).reshape(np.asarray(e).shape)


def crt_bal(x, p):
    """x mod p in [-floor(p/2), ceil(p/2) - 1] (numpy int64 arrays)."""
    r = np.mod(x, p)
    return np.where(2 * r >= p, r - p, r)


def crt_scale(X, e, axis, q):
    """Integers rint(x * 2^(q - e)), rounded to nearest (ties to even), as
    int64 (one exponent e per row, axis=1, or column, axis=0): a power-of-two
    scaling, exact in FP64 while |x 2^(q - e)| < 2^53."""
    shape = [1, 1]
    shape[1 - axis] = -1
    return np.rint(np.ldexp(X, (q - e).reshape(shape))).astype(np.int64)


def crt_tables(nmod):
    """R_t = prod_{u<t} p_u (Python ints), R_t mod p_u and R_t^{-1} mod p_t."""
    ps = CRT_MODULI[:nmod]
    R = [1]
    for p in ps[:-1]:
        R.append(R[-1] * p)
    rmod = [[R[t] % ps[u] for u in range(nmod)] for t in range(nmod)]
    rinv = [pow(R[t] % ps[t], -1, ps[t]) if t else 1 for t in range(nmod)]
    return ps, R, rmod, rinv


def crt_reconstruct(res, nmod):
    """Exact integers (Python ints, object array) from residues res[t] in
    [0, p_t) by Garner's algorithm with balanced mixed-radix digits."""
    ps, R, rmod, rinv = crt_tables(nmod)
    shape = res[0].shape
    s = [np.zeros(shape, dtype=np.int64) for _ in range(nmod)]  # partial sums mod p_u
    digits = []
    for t in range(nmod):
        v = crt_bal((res[t].astype(np.int64) - s[t]) * rinv[t], ps[t])
        digits.append(v)
        for u in range(t + 1, nmod):
            s[u] = np.mod(s[u] + v * rmod[t][u], ps[u])
    total = np.zeros(shape, dtype=object)
    for t in range(nmod):
        total = total + digits[t].astype(object) * R[t]
    return total


def ozaki2_dgemm(a, b, nmod, alpha=1.0, beta=0.0, c=None):
    """C = alpha * A B + beta * C through nmod INT8 modular products (Ozaki
    scheme II) on 2-D float64 arrays a (M x K), b (K x N).  The rebuilt
    integer V = A'B' is exact; the result is alpha * ldexp(double(V),
    e'_i + f'_j - 2Q) (+ beta * C), double(V) correctly rounded, with
    e' = Q - sigma the shifts of crt_shifts."""
    M, K = a.shape
    N = b.shape[1]
    q = crt_Q(K, nmod)
    if q is None:
        raise ValueError("%d moduli leave no bits for K=%d" % (nmod, K))
    e = crt_shifts(*crt_line_bounds(a, 1), q)
    f = crt_shifts(*crt_line_bounds(b, 0), q)
    Ai = crt_scale(a, e, 1, q)
    Bi = crt_scale(b, f, 0, q)
    ps = CRT_MODULI[:nmod]
    res = []
    for p in ps:
        ar = crt_bal(Ai, p).astype(np.float64)
        br = crt_bal(Bi, p).astype(np.float64)
        W = ar @ br                      # exact: |entries| <= 128^2 * K < 2^53
        res.append(np.mod(W.astype(np.int64), p))
    V = crt_reconstruct(res, nmod)
    Vd = np.vectorize(float, otypes=[np.float64])(V)     # correctly rounded
    prod = np.ldexp(Vd, (e.reshape(-1, 1) + f.reshape(1, -1) - 2 * q).astype(np.int64))
    out = alpha * prod
    if beta != 0.0:
        out = out + beta * c
    return out
