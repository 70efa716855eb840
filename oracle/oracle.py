"""ctypes front end of the CPU oracle (zg_oracle.c).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module.  It is the checker for the HIP path.

Array conventions (numpy):
  * Fr / Fq element: uint64[4] little-endian limbs; "mont" = Montgomery form.
  * param buffer: uint64[(h+2) * N * 8], line-major quotients[0..h) || basis ||
    high_basis, each point x_mont || y_mont (bamt layout).
  * output point: uint8[64] canonical x || y, all zeros = infinity.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libzg_oracle.so")
_lib = None

R_MOD = 21888242871839275222246405745257275088548364400416034343698204186575808495617
P_MOD = 21888242871839275222246405745257275088696311157297823662689037894645226208583


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        I = ctypes.c_int
        S = ctypes.c_size_t
        sig = {
            "zgo_fmul": (None, [I, P, P, P]),
            "zgo_to_mont": (None, [I, P, P]),
            "zgo_from_mont": (None, [I, P, P]),
            "zgo_finv": (None, [I, P, P]),
            "zgo_from_mont_n": (None, [I, P, S, P]),
            "zgo_g1_mul_gen": (None, [P, P]),
            "zgo_g1_mul": (None, [P, P, P]),
            "zgo_g1_add": (None, [P, P, P]),
            "zgo_msm": (None, [P, P, S, P]),
            "zgo_domain_root": (None, [I, P]),
            "zgo_coset_factor": (None, [I, I, P]),
            "zgo_fft": (None, [P, I]),
            "zgo_ifft": (None, [P, I]),
            "zgo_to_coset_blob": (None, [P, I, I, P]),
            "zgo_ecfft": (None, [P, I, I, P]),
            "zgo_params_from_tau": (I, [P, I, I, I, I, P]),
            "zgo_params_point": (I, [P, I, I, I, I, ctypes.c_uint64, P]),
            "zgo_eval_interp": (I, [P, I, P, P]),
            "zgo_amt_prove": (I, [P, I, I, P, P, P, P, P]),
            "zgo_multiple_multiexp": (I, [P, I, I, P, P]),
            "zgo_keccak256": (None, [P, S, ctypes.c_uint8, P]),
            "zgo_merkle": (I, [P, S, I, I, P, P]),
            "zgo_raw_to_blob": (I, [P, S, I, P]),
            "zgo_encode": (I, [P, I, I, I, P, S, P, P, P, P, P, P, P]),
            "zgo_recover": (I, [P, P, S, I, I, I, P, P]),
            "zgo_verify_msm": (I, [P, I, I, I, P, P, P, S, P]),
            "zgo_power_tau": (None, [P, ctypes.c_uint64, I, P]),
            "zgo_set_threads": (I, [I]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def set_threads(n):
    return lib().zgo_set_threads(int(n))


# ---- integer <-> limb helpers -------------------------------------------
def int_to_limbs(x):
    return np.array([(x >> (64 * i)) & 0xFFFFFFFFFFFFFFFF for i in range(4)], dtype=np.uint64)


def limbs_to_int(a):
    a = np.asarray(a, dtype=np.uint64).reshape(-1)
    return sum(int(a[i]) << (64 * i) for i in range(4))


def ints_to_limbs(xs):
    out = np.zeros((len(xs), 4), dtype=np.uint64)
    for k, x in enumerate(xs):
        out[k] = int_to_limbs(x)
    return out


def point_to_ints(b):
    """64-byte canonical point -> (x, y) ints, or None for infinity."""
    b = bytes(np.asarray(b, dtype=np.uint8))
    if not any(b):
        return None
    return int.from_bytes(b[:32], "little"), int.from_bytes(b[32:], "little")


def ints_to_point(pt):
    if pt is None:
        return np.zeros(64, dtype=np.uint8)
    x, y = pt
    return np.frombuffer(x.to_bytes(32, "little") + y.to_bytes(32, "little"), dtype=np.uint8).copy()


# ---- field ------------------------------------------------------------
def fmul(which, a, b):
    out = np.zeros(4, dtype=np.uint64)
    lib().zgo_fmul(which, _p(int_to_limbs(a)), _p(int_to_limbs(b)), _p(out))
    return limbs_to_int(out)


def to_mont(which, a):
    out = np.zeros(4, dtype=np.uint64)
    lib().zgo_to_mont(which, _p(int_to_limbs(a)), _p(out))
    return limbs_to_int(out)


def from_mont(which, a):
    out = np.zeros(4, dtype=np.uint64)
    lib().zgo_from_mont(which, _p(int_to_limbs(a)), _p(out))
    return limbs_to_int(out)


def finv(which, a):
    out = np.zeros(4, dtype=np.uint64)
    lib().zgo_finv(which, _p(int_to_limbs(a)), _p(out))
    return limbs_to_int(out)


def fr_to_mont_array(vals):
    """list of ints (canonical Fr) -> uint64[n,4] mont"""
    out = np.zeros((len(vals), 4), dtype=np.uint64)
    for k, v in enumerate(vals):
        out[k] = int_to_limbs(to_mont(0, v % R_MOD))
    return out


def fr_canonical_array(arr, which=0):
    """uint64[n,4] mont -> uint64[n,4] canonical (bulk); which = 1 for Fq."""
    a = np.ascontiguousarray(arr, dtype=np.uint64).reshape(-1, 4)
    out = np.zeros_like(a)
    lib().zgo_from_mont_n(which, _p(a), a.shape[0], _p(out))
    return out


def fr_from_mont_array(arr):
    arr = np.ascontiguousarray(arr, dtype=np.uint64).reshape(-1, 4)
    return [from_mont(0, limbs_to_int(arr[k])) for k in range(arr.shape[0])]


# ---- curve ------------------------------------------------------------
def g1_mul_gen(k):
    out = np.zeros(64, dtype=np.uint8)
    lib().zgo_g1_mul_gen(_p(int_to_limbs(k % R_MOD)), _p(out))
    return out


def g1_mul(p, k):
    out = np.zeros(64, dtype=np.uint8)
    lib().zgo_g1_mul(_p(np.ascontiguousarray(p, dtype=np.uint8)), _p(int_to_limbs(k % R_MOD)), _p(out))
    return out


def g1_add(p, q):
    out = np.zeros(64, dtype=np.uint8)
    lib().zgo_g1_add(_p(np.ascontiguousarray(p, dtype=np.uint8)),
                     _p(np.ascontiguousarray(q, dtype=np.uint8)), _p(out))
    return out


def msm(bases_mont, scalars_canon):
    bases_mont = np.ascontiguousarray(bases_mont, dtype=np.uint64)
    scalars_canon = np.ascontiguousarray(scalars_canon, dtype=np.uint64)
    n = scalars_canon.size // 4
    out = np.zeros(64, dtype=np.uint8)
    lib().zgo_msm(_p(bases_mont), _p(scalars_canon), n, _p(out))
    return out


# ---- NTT ----------------------------------------------------------------
def domain_root(log_n):
    out = np.zeros(4, dtype=np.uint64)
    lib().zgo_domain_root(log_n, _p(out))
    return limbs_to_int(out)


def coset_factor(log_n, idx):
    out = np.zeros(4, dtype=np.uint64)
    lib().zgo_coset_factor(log_n, idx, _p(out))
    return limbs_to_int(out)


def fft(a_mont, log_n):
    a = np.ascontiguousarray(a_mont, dtype=np.uint64).copy()
    lib().zgo_fft(_p(a), log_n)
    return a


def ifft(a_mont, log_n):
    a = np.ascontiguousarray(a_mont, dtype=np.uint64).copy()
    lib().zgo_ifft(_p(a), log_n)
    return a


def to_coset_blob(data_mont, log_n, coset):
    d = np.ascontiguousarray(data_mont, dtype=np.uint64)
    out = np.zeros_like(d)
    lib().zgo_to_coset_blob(_p(d), log_n, coset, _p(out))
    return out


def ecfft(points, log_n, inverse):
    pts = np.ascontiguousarray(points, dtype=np.uint8).reshape(-1)
    out = np.zeros_like(pts)
    lib().zgo_ecfft(_p(pts), log_n, int(inverse), _p(out))
    return out.reshape(-1, 64)


# ---- params / prover -----------------------------------------------------
def params_from_tau(tau, log_n, h, coset, method=0):
    n = 1 << log_n
    out = np.zeros((h + 2) * n * 8, dtype=np.uint64)
    rc = lib().zgo_params_from_tau(_p(int_to_limbs(tau % R_MOD)), log_n, h, coset, method, _p(out))
    if rc != 0:
        raise ValueError("params_from_tau failed: %d" % rc)
    return out


def params_point(tau, log_n, h, coset, line, pos):
    out = np.zeros(8, dtype=np.uint64)
    rc = lib().zgo_params_point(_p(int_to_limbs(tau % R_MOD)), log_n, h, coset, line, pos, _p(out))
    if rc != 0:
        raise ValueError("tau on the domain")
    return out


def eval_interp(y_mont, log_n, x):
    y = np.ascontiguousarray(y_mont, dtype=np.uint64)
    out = np.zeros(4, dtype=np.uint64)
    rc = lib().zgo_eval_interp(_p(y), log_n, _p(int_to_limbs(x % R_MOD)), _p(out))
    if rc != 0:
        raise ValueError("x on the domain")
    return limbs_to_int(out)


def amt_prove(bases, log_n, h, ri_mont):
    pts = (1 << (h + 1)) - 2
    com = np.zeros(64, dtype=np.uint8)
    ct = np.zeros(pts * 64, dtype=np.uint8)
    pt = np.zeros(pts * 64, dtype=np.uint8)
    hi = np.zeros(64, dtype=np.uint8)
    ri = np.ascontiguousarray(ri_mont, dtype=np.uint64)
    lib().zgo_amt_prove(_p(np.ascontiguousarray(bases)), log_n, h, _p(ri), _p(com), _p(ct), _p(pt), _p(hi))
    return com, ct.reshape(-1, 64), pt.reshape(-1, 64), hi


def multiple_multiexp(bases, log_n, h, ri_mont):
    nb = 1 << h
    out = np.zeros((h + 2) * nb * 64, dtype=np.uint8)
    ri = np.ascontiguousarray(ri_mont, dtype=np.uint64)
    lib().zgo_multiple_multiexp(_p(np.ascontiguousarray(bases)), log_n, h, _p(ri), _p(out))
    return out.reshape(-1, 64)


# ---- keccak / merkle -------------------------------------------------------
def keccak256(msg, pad=0x01):
    m = np.frombuffer(bytes(msg), dtype=np.uint8).copy() if len(msg) else np.zeros(1, dtype=np.uint8)
    out = np.zeros(32, dtype=np.uint8)
    lib().zgo_keccak256(_p(m), len(msg), pad, _p(out))
    return bytes(out)


def merkle(data, coset_n, levels):
    data = np.ascontiguousarray(data, dtype=np.uint8).reshape(-1)
    n_elems = data.size // 32
    top = sum(coset_n << k for k in range(levels + 1))
    layers = np.zeros(top * 32, dtype=np.uint8)
    root = np.zeros(32, dtype=np.uint8)
    nl = lib().zgo_merkle(_p(data), n_elems, coset_n, levels, _p(layers), _p(root))
    return nl, layers.reshape(-1, 32), bytes(root)


# ---- packing / encode --------------------------------------------------------
def raw_to_blob(raw, log_n):
    n = 1 << log_n
    out = np.zeros((n, 4), dtype=np.uint64)
    buf = np.frombuffer(bytes(raw), dtype=np.uint8).copy() if len(raw) else np.zeros(1, dtype=np.uint8)
    rc = lib().zgo_raw_to_blob(_p(buf), len(raw), log_n, _p(out))
    if rc == 1:
        raise OverflowError("TooLargeBlob")
    return out


class Encoded:
    """Result of the oracle's EncodedBlob::build."""

    def __init__(self, log_col, log_row, coset_n):
        n = 1 << (log_col + log_row)
        pts = (1 << (log_row + 1)) - 2
        self.log_col, self.log_row, self.coset_n = log_col, log_row, coset_n
        self.data = np.zeros(coset_n * n * 32, dtype=np.uint8)
        self.commitment = np.zeros(64, dtype=np.uint8)
        self.comm_tree = np.zeros(coset_n * pts * 64, dtype=np.uint8)
        self.proof_tree = np.zeros(coset_n * pts * 64, dtype=np.uint8)
        self.high = np.zeros(coset_n * 64, dtype=np.uint8)
        self.merkle = np.zeros(coset_n * ((1 << (log_row + 1)) - 1) * 32, dtype=np.uint8)
        self.root = np.zeros(32, dtype=np.uint8)


def encode(bases_list, log_col, log_row, raw):
    coset_n = len(bases_list)
    e = Encoded(log_col, log_row, coset_n)
    arr = (ctypes.c_void_p * coset_n)(*[b.ctypes.data for b in bases_list])
    buf = np.frombuffer(bytes(raw), dtype=np.uint8).copy() if len(raw) else np.zeros(1, dtype=np.uint8)
    rc = lib().zgo_encode(ctypes.cast(arr, ctypes.c_void_p), log_col, log_row, coset_n, _p(buf), len(raw),
                          _p(e.data), _p(e.commitment), _p(e.comm_tree), _p(e.proof_tree), _p(e.high),
                          _p(e.merkle), _p(e.root))
    if rc == 1:
        raise OverflowError("TooLargeBlob")
    if rc == 2:
        raise AssertionError("coset commitments differ")
    return e


# ---- RS recovery --------------------------------------------------------------
RECOVER_STATUS = {
    0: "Ok",
    1: "Not enough valid lines",
    2: "Cannot recover data: RowIdOverflow",
    3: "Incorrect scalar",
    4: "Incorrect length field",
    5: "Non zero in padding range",
    6: "panic: element out of range",
    7: "panic: recovered degree above N",
    8: "Cannot recover data: ExtaustiveK",
}


def recover(slices, log_col, log_row, coset_n=3):
    """recover_from_da_slice (crates/recovery/src/lib.rs:15-43) over {row id: row bytes}
    (each row 32 * 2^log_col canonical bytes, EncodedSlice::merkle_row).  Rows of the
    wrong length are dropped first, as the reference does.  Returns (status, bytes):
    the raw data for status 0, else b"" (RECOVER_STATUS names the outcome)."""
    col = 1 << log_col
    ok = sorted((int(k), bytes(v)) for k, v in slices.items() if len(v) == 32 * col)
    ids = np.array([k for k, _ in ok], dtype=np.uint32) if ok else np.zeros(1, dtype=np.uint32)
    rows = np.frombuffer(b"".join(v for _, v in ok), dtype=np.uint8).copy() if ok else np.zeros(1, dtype=np.uint8)
    n = 1 << (log_col + log_row)
    out = np.zeros(31 * n, dtype=np.uint8)
    out_len = ctypes.c_size_t(0)
    rc = lib().zgo_recover(_p(ids), _p(rows), len(ok), log_col, log_row, coset_n, _p(out),
                           ctypes.byref(out_len))
    return rc, (out[:out_len.value].tobytes() if rc == 0 else b"")


# ---- deferred verification MSM ---------------------------------------------------
def verify_msm(bases_list, log_col, log_row, ids, rows, alphas):
    """DeferredVerifier::record_msm + fast_check_msm's MSM over rows (see zg_oracle.h):
    ids strictly increasing row indices, rows their canonical bytes (one buffer),
    alphas canonical ints.  Returns the 64 B canonical point."""
    coset_n = len(bases_list)
    arr = (ctypes.c_void_p * coset_n)(*[b.ctypes.data for b in bases_list])
    n = len(ids)
    idv = np.array(ids, dtype=np.uint32) if n else np.zeros(1, dtype=np.uint32)
    rw = np.frombuffer(bytes(rows), dtype=np.uint8).copy() if n else np.zeros(1, dtype=np.uint8)
    al = np.frombuffer(b"".join(int(a).to_bytes(32, "little") for a in alphas), dtype=np.uint8).copy() \
        if n else np.zeros(1, dtype=np.uint8)
    out = np.zeros(64, dtype=np.uint8)
    rc = lib().zgo_verify_msm(ctypes.cast(arr, ctypes.c_void_p), log_col, log_row, coset_n, _p(idv), _p(rw), _p(al),
                              n, _p(out))
    if rc:
        raise ValueError("row index past the encoded blob")
    return bytes(out)


def power_tau(tau, start, log_n):
    """[tau^(start + i)] G1 for i < 2^log_n (power_tau, crates/amt/src/power_tau.rs:28-42),
    canonical affine, as one (2^log_n * 64) uint8 array."""
    out = np.zeros((1 << log_n) * 64, dtype=np.uint8)
    lib().zgo_power_tau(_p(int_to_limbs(tau % R_MOD)), start, log_n, _p(out))
    return out
