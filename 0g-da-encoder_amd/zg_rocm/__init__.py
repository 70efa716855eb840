"""ctypes binding of lib/libzgr.so (include/zgr.h) plus a thin Python mirror of the
reference encode API (crates/encoder: RawData, RawBlob, EncodedBlob; grpc:
EncoderService::process_data).

The HIP library is mandatory: importing works without it, but every call that
needs the GPU raises if lib/libzgr.so is missing or no device is present.  There
is no CPU fallback.
"""
import ctypes
import os

import numpy as np

from . import wire

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libzgr.so")

R_MOD = 21888242871839275222246405745257275088548364400416034343698204186575808495617
P_MOD = 21888242871839275222246405745257275088696311157297823662689037894645226208583  # BN254 Fq

OK = 0
ERR_INVALID_ARG = -1
ERR_TOO_LARGE = -2
ERR_OOM = -3
ERR_DEVICE = -4
ERR_NOT_LOADED = -5
ERR_INCONSISTENT = -6
ERR_CANCELLED = -7
FAULT_ISSUE = 0  # zgr_test_inject_fault sites

# parameter file kinds of zgr_write_params_file
FILE_AMT_MONT = 0
FILE_AMT = 1
FILE_AMT_VERIFY = 2
FILE_PTAU = 3
FILE_PTAU_MONT = 4
# point encodings of zgr_ppot_decode (ceremony files)
PPOT_G1_UNCOMPRESSED = 0
PPOT_G1_COMPRESSED = 1
PPOT_G2_UNCOMPRESSED = 2
PPOT_G2_COMPRESSED = 3
# type_hash::<Bn254>()[..6] of the reference's file names (crates/amt/src/utils.rs:11-52,
# README.md:73-82)
BN254_HASH = "5DWgDV"

# exported symbols and their ctypes signatures (kept in step with include/zgr.h)
P = ctypes.c_void_p
I = ctypes.c_int
S = ctypes.c_size_t
SIGNATURES = {
    "zgr_last_error": (ctypes.c_char_p, []),
    "zgr_ctx_create": (I, [I, I, I, I, ctypes.POINTER(P)]),
    "zgr_ctx_destroy": (None, [P]),
    "zgr_params_load": (I, [P, I, P, S]),
    "zgr_params_load_bamt": (I, [P, I, P, S]),
    "zgr_params_generate": (I, [P, P]),
    "zgr_params_from_pp": (I, [P, I, P, P, S]),
    "zgr_params_download": (I, [P, I, P, S]),
    "zgr_power_tau_setup": (I, [P, P]),
    "zgr_power_tau_load": (I, [P, P, P, P, P, S]),
    "zgr_power_tau_download": (I, [P, P, P, P, P]),
    "zgr_ppot_decode": (I, [P, I, P, S, P, ctypes.POINTER(S)]),
    "zgr_params_from_power_tau": (I, [P, I]),
    "zgr_params_g2": (I, [P, I, P, P, P]),
    "zgr_write_params_file": (I, [P, I, I, I, P, S, ctypes.POINTER(S)]),
    "zgr_amt_prove": (I, [P, I, P, S, P, P, P, P]),
    "zgr_multiple_multiexp": (I, [P, I, P, S, P]),
    "zgr_encode": (I, [P, P, S, P]),
    "zgr_encode_device": (I, [P, P, S, P]),
    "zgr_encode_batch_device": (I, [P, I, P, P, P]),
    "zgr_ntt": (I, [P, P, I, I, I]),
    "zgr_ntt_device": (I, [P, P, I, I, I]),
    "zgr_coset_extend": (I, [P, P, I, P]),
    "zgr_msm": (I, [P, P, P, S, P]),
    "zgr_ecfft": (I, [P, P, I, I, P]),
    "zgr_merkle": (I, [P, P, S, I, P, P]),
    "zgr_merkle_noalloc": (I, [P, P, S, I, P, P]),
    "zgr_server_create": (I, [P, S, I, ctypes.POINTER(P)]),
    "zgr_server_destroy": (None, [P]),
    "zgr_server_submit": (I, [P, P, S, P, ctypes.POINTER(ctypes.c_uint64)]),
    "zgr_server_poll": (I, [P, ctypes.c_uint64, I]),
    "zgr_server_cancel": (I, [P, ctypes.c_uint64]),
    "zgr_test_inject_fault": (I, [P, I, I]),
    "zgr_server_stats": (None, [P, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                                ctypes.POINTER(ctypes.c_uint64)]),
    "zgr_recover": (I, [P, P, P, S, P, S, ctypes.POINTER(S)]),
    "zgr_verify_msm": (I, [P, P, P, P, S, P]),
    "zgr_autotune": (I, [P, ctypes.c_char_p, S]),
    "zgr_kernel_time_ms": (ctypes.c_double, [P, ctypes.c_char_p, ctypes.POINTER(I)]),
    "zgr_kernel_time_reset": (None, [P]),
    "zgr_set_profiling": (None, [P, I]),
}

_lib = None


class ZgrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("zgr error %d: %s" % (code, msg))
        self.code = code
        self.msg = msg


class RequestCancelled(RuntimeError):
    """A served request withdrawn (zgr_server_cancel) because its caller went away."""


class TooLargeBlob(ValueError):
    """EncoderError::TooLargeBlob (crates/encoder/src/encoder/error.rs:3-6)."""

    def __init__(self, actual, expected_max):
        super().__init__("TooLargeBlob { actual: %d, expected_max: %d }" % (actual, expected_max))
        self.actual = actual
        self.expected_max = expected_max


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("HIP library %s is not built (run 0g-da-encoder_amd/build.py)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


_hip = None


def hip():
    """The HIP runtime instance libzgr.so is linked against (found through the
    process maps after the library is loaded), so device buffers made here and the
    encoder share one runtime."""
    global _hip
    if _hip is None:
        lib()
        path = None
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64.so" in line:
                    path = line.split()[-1]
                    break
        H = ctypes.CDLL(path or "libamdhip64.so")
        H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        H.hipFree.argtypes = [ctypes.c_void_p]
        H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        H.hipDeviceSynchronize.argtypes = []
        H.hipHostMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t, ctypes.c_uint]
        H.hipHostFree.argtypes = [ctypes.c_void_p]
        for f in (H.hipMalloc, H.hipFree, H.hipMemcpy, H.hipDeviceSynchronize, H.hipHostMalloc, H.hipHostFree):
            f.restype = ctypes.c_int
        _hip = H
    return _hip


class PinnedHostArray:
    """A page-locked host byte array (hipHostMalloc) as a numpy uint8 array (`.array`),
    so that rows in and raw data out cross PCIe at the link's rate."""

    def __init__(self, size):
        self.size = int(size)
        self.ptr = ctypes.c_void_p()
        rc = hip().hipHostMalloc(ctypes.byref(self.ptr), max(self.size, 1), 0)
        if rc:
            raise ZgrError(ERR_OOM, "hipHostMalloc(%d) failed: %d" % (self.size, rc))
        self.array = np.ctypeslib.as_array((ctypes.c_uint8 * max(self.size, 1)).from_address(self.ptr.value))[:self.size]

    def free(self):
        if self.ptr.value:
            self.array = None
            hip().hipHostFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceBuffer:
    """A hipMalloc'd byte buffer holding a copy of `data` (inputs resident in HBM)."""

    def __init__(self, data):
        data = bytes(data)
        self.size = len(data)
        self.ptr = ctypes.c_void_p()
        if self.size:
            rc = hip().hipMalloc(ctypes.byref(self.ptr), self.size)
            if rc:
                raise ZgrError(ERR_OOM, "hipMalloc(%d) failed: %d" % (self.size, rc))
            src = ctypes.create_string_buffer(data, self.size)
            rc = hip().hipMemcpy(self.ptr, src, self.size, 1)  # hipMemcpyHostToDevice
            if rc:
                raise ZgrError(ERR_DEVICE, "hipMemcpy failed: %d" % rc)

    @property
    def address(self):
        return self.ptr.value or 0

    def to_host(self):
        """The buffer's current bytes."""
        out = ctypes.create_string_buffer(max(self.size, 1))
        if self.size:
            rc = hip().hipMemcpy(out, self.ptr, self.size, 2)  # hipMemcpyDeviceToHost
            if rc:
                raise ZgrError(ERR_DEVICE, "hipMemcpy failed: %d" % rc)
        return out.raw[:self.size]

    def free(self):
        if self.ptr.value:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _check(rc):
    if rc != OK:
        raise ZgrError(rc, lib().zgr_last_error().decode(errors="replace"))
    return rc


def tau_from_seed(seed):
    """Deterministic insecure-setup tau: 64 bytes of numpy PCG64(seed) reduced mod r."""
    b = np.random.default_rng(seed).integers(0, 256, 64, dtype=np.uint8).tobytes()
    return int.from_bytes(b, "little") % R_MOD


def seeded_bytes(seed, length):
    """Seeded uniform random bytes (numpy PCG64)."""
    return np.random.default_rng(seed).integers(0, 256, length, dtype=np.uint8).tobytes()


class Geometry:
    """Blob geometry: crates/encoder/src/constants.rs:11-28 as runtime values."""

    def __init__(self, log_col, log_row, coset_n=3):
        self.log_col, self.log_row, self.coset_n = log_col, log_row, coset_n
        self.log_n = log_col + log_row
        self.h = log_row
        self.n = 1 << self.log_n
        self.row_n = 1 << log_row
        self.col_n = 1 << log_col
        self.row_encoded = self.row_n * coset_n
        self.max_blob_size = 31 * self.n
        self.max_raw_data_size = self.max_blob_size - 4
        self.tree_pts = (1 << (log_row + 1)) - 2
        self.merkle_nodes = coset_n * ((1 << (log_row + 1)) - 1)

    def params_bytes(self):
        return (self.h + 2) * self.n * 64


PRODUCTION = Geometry(10, 10, 3)  # BLOB_COL_LOG = BLOB_ROW_LOG = 10
TEST_SETTING = Geometry(5, 6, 3)   # BLOB_COL_LOG = 5, BLOB_ROW_LOG = 6 (TEST_SETTING)


class EncodedBlobOutputs:
    """Host buffers matching zgr_encoded."""

    class _C(ctypes.Structure):
        _fields_ = [(k, ctypes.c_void_p) for k in
                    ("data", "commitment", "comm_tree", "proof_tree", "high", "merkle", "file_root")]

    def __init__(self, geo, want_data=True):
        g = geo
        self.data = np.zeros(g.coset_n * g.n * 32, dtype=np.uint8) if want_data else None
        self.commitment = np.zeros(64, dtype=np.uint8)
        self.comm_tree = np.zeros(g.coset_n * g.tree_pts * 64, dtype=np.uint8)
        self.proof_tree = np.zeros(g.coset_n * g.tree_pts * 64, dtype=np.uint8)
        self.high = np.zeros(g.coset_n * 64, dtype=np.uint8)
        self.merkle = np.zeros(g.merkle_nodes * 32, dtype=np.uint8)
        self.file_root = np.zeros(32, dtype=np.uint8)

    def cstruct(self):
        c = self._C()
        for k, _ in self._C._fields_:
            a = getattr(self, k)
            setattr(c, k, a.ctypes.data if a is not None else None)
        return c


class Context:
    """A zgr_ctx: one device, one blob geometry, resident params."""

    def __init__(self, geo=PRODUCTION, device=0):
        self.geo = geo
        h = ctypes.c_void_p()
        _check(lib().zgr_ctx_create(device, geo.log_col, geo.log_row, geo.coset_n, ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().zgr_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- params ----
    def params_generate(self, tau):
        t = np.frombuffer((tau % R_MOD).to_bytes(32, "little"), dtype=np.uint8).copy()
        _check(lib().zgr_params_generate(self.h, _p(t)))

    def params_load(self, coset, bases):
        b = np.ascontiguousarray(bases)
        _check(lib().zgr_params_load(self.h, coset, _p(b), b.nbytes))

    def params_load_bamt(self, coset, image):
        """One coset's bamt file image (bytes / uint8 array), as AMTParams::from_dir_mont reads it."""
        b = np.frombuffer(image, dtype=np.uint8) if isinstance(image, (bytes, bytearray)) else np.ascontiguousarray(image)
        _check(lib().zgr_params_load_bamt(self.h, coset, _p(b), b.nbytes))

    def params_from_pp(self, coset, g1pp, high_g1pp):
        """zgr_params_from_pp: canonical affine powers of tau (2^log_n points each)."""
        a = np.ascontiguousarray(g1pp, dtype=np.uint8).reshape(-1)
        b = np.ascontiguousarray(high_g1pp, dtype=np.uint8).reshape(-1)
        _check(lib().zgr_params_from_pp(self.h, coset, _p(a), _p(b), a.size // 64))

    # ---- powers of tau, verifier params, parameter files ----
    def power_tau_setup(self, tau):
        """PowerTau::setup_inner with a known tau, on the device (kept in the context)."""
        t = np.frombuffer((tau % R_MOD).to_bytes(32, "little"), dtype=np.uint8).copy()
        _check(lib().zgr_power_tau_setup(self.h, _p(t)))

    def power_tau_load(self, g1pp, g2pp, high_g1pp, high_g2):
        a = [np.ascontiguousarray(x, dtype=np.uint8).reshape(-1) for x in (g1pp, g2pp, high_g1pp, high_g2)]
        _check(lib().zgr_power_tau_load(self.h, _p(a[0]), _p(a[1]), _p(a[2]), _p(a[3]), a[0].size // 64))

    def ppot_decode(self, kind, src, count):
        """zgr_ppot_decode: count encoded ceremony points (bytes-like, back to back) ->
        canonical affine points, (count, 64) for G1 or (count, 128) for G2.  A rejected
        point raises ZgrError with .first_bad set to its index."""
        a = np.frombuffer(src, dtype=np.uint8) if not isinstance(src, np.ndarray) else src.reshape(-1)
        a = np.ascontiguousarray(a, dtype=np.uint8)
        width = 128 if kind >= PPOT_G2_UNCOMPRESSED else 64
        per = (64 if kind >= PPOT_G2_UNCOMPRESSED else 32) * (1 if kind % 2 else 2)
        if a.size < count * per:
            raise ValueError("%d bytes hold fewer than %d encoded points of %d B" % (a.size, count, per))
        out = np.zeros((count, width), dtype=np.uint8)
        bad = ctypes.c_size_t(0)
        rc = lib().zgr_ppot_decode(self.h, kind, _p(a), count, _p(out), ctypes.byref(bad))
        if rc != OK:
            err = ZgrError(rc, lib().zgr_last_error().decode(errors="replace"))
            err.first_bad = bad.value
            raise err
        return out

    def power_tau_download(self):
        """-> g1pp (n, 64), g2pp (n, 128), high_g1pp (n, 64), high_g2 (128,): canonical affine."""
        n = self.geo.n
        g1, g2, hi, hg2 = (np.zeros(k, dtype=np.uint8) for k in (64 * n, 128 * n, 64 * n, 128))
        _check(lib().zgr_power_tau_download(self.h, _p(g1), _p(g2), _p(hi), _p(hg2)))
        return g1.reshape(n, 64), g2.reshape(n, 128), hi.reshape(n, 64), hg2

    def params_from_power_tau(self, coset):
        _check(lib().zgr_params_from_power_tau(self.h, coset))

    def params_g2(self, coset):
        """-> vanishes ((2^(h+1) - 2), 128), g2 (128,), high_g2 (128,): canonical affine G2."""
        v = np.zeros((((2 << self.geo.h) - 2), 128), dtype=np.uint8)
        g2, hg2 = np.zeros(128, dtype=np.uint8), np.zeros(128, dtype=np.uint8)
        _check(lib().zgr_params_g2(self.h, coset, _p(v), _p(g2), _p(hg2)))
        return v, g2, hg2

    def write_params_file(self, kind, coset=0, sub_depth=None):
        """The image (bytes) of one of the reference's parameter files (FILE_*)."""
        sd = self.geo.h if sub_depth is None else sub_depth
        n = ctypes.c_size_t(0)
        _check(lib().zgr_write_params_file(self.h, kind, coset, sd, None, 0, ctypes.byref(n)))
        out = np.empty(max(n.value, 1), dtype=np.uint8)
        _check(lib().zgr_write_params_file(self.h, kind, coset, sd, _p(out), out.size, ctypes.byref(n)))
        return out[:n.value].tobytes()

    def params_download(self, coset):
        out = np.zeros(self.geo.params_bytes() // 8, dtype=np.uint64)
        _check(lib().zgr_params_download(self.h, coset, _p(out), out.nbytes))
        return out

    # ---- prover seam ----
    def amt_prove(self, coset, ri_mont):
        g = self.geo
        ri = np.ascontiguousarray(ri_mont, dtype=np.uint64)
        com = np.zeros(64, dtype=np.uint8)
        ct = np.zeros(g.tree_pts * 64, dtype=np.uint8)
        pt = np.zeros(g.tree_pts * 64, dtype=np.uint8)
        hi = np.zeros(64, dtype=np.uint8)
        _check(lib().zgr_amt_prove(self.h, coset, _p(ri), ri.size // 4, _p(com), _p(ct), _p(pt), _p(hi)))
        return com, ct.reshape(-1, 64), pt.reshape(-1, 64), hi

    def multiple_multiexp(self, coset, ri_mont):
        g = self.geo
        ri = np.ascontiguousarray(ri_mont, dtype=np.uint64)
        out = np.zeros((g.h + 2) * (1 << g.h) * 64, dtype=np.uint8)
        _check(lib().zgr_multiple_multiexp(self.h, coset, _p(ri), ri.size // 4, _p(out)))
        return out.reshape(-1, 64)

    # ---- encode ----
    def encode(self, raw, want_data=True):
        if len(raw) > self.geo.max_raw_data_size:
            raise TooLargeBlob(len(raw), self.geo.max_raw_data_size)
        out = EncodedBlobOutputs(self.geo, want_data)
        buf = np.frombuffer(bytes(raw), dtype=np.uint8) if len(raw) else np.zeros(1, dtype=np.uint8)
        cs = out.cstruct()
        _check(lib().zgr_encode(self.h, _p(buf), len(raw), ctypes.byref(cs)))
        return out

    def encode_device(self, raw_dev_ptr, length, want_data=True, out=None):
        out = out or EncodedBlobOutputs(self.geo, want_data)
        cs = out.cstruct()
        _check(lib().zgr_encode_device(self.h, ctypes.c_void_p(raw_dev_ptr), length, ctypes.byref(cs)))
        return out

    def encode_batch_device(self, ptrs, lengths, outs):
        n = len(ptrs)
        parr = (ctypes.c_void_p * n)(*ptrs)
        larr = (ctypes.c_size_t * n)(*lengths)
        carr = (EncodedBlobOutputs._C * n)(*[o.cstruct() for o in outs])
        _check(lib().zgr_encode_batch_device(self.h, n, ctypes.cast(parr, P), ctypes.cast(larr, P),
                                             ctypes.cast(carr, P)))
        return outs

    # ---- primitives ----
    def ntt(self, rows_mont, log_n, inverse=False):
        a = np.ascontiguousarray(rows_mont, dtype=np.uint64).copy()
        batch = a.size // (4 << log_n)
        _check(lib().zgr_ntt(self.h, _p(a), log_n, batch, int(inverse)))
        return a

    def ntt_device(self, data_dev_ptr, log_n, batch, inverse=False):
        """zgr_ntt_device: batch rows of 2^log_n Fr (mont) in device memory, in place."""
        _check(lib().zgr_ntt_device(self.h, ctypes.c_void_p(data_dev_ptr), log_n, batch, int(inverse)))

    def coset_extend(self, data_mont, coset):
        d = np.ascontiguousarray(data_mont, dtype=np.uint64)
        out = np.zeros_like(d)
        _check(lib().zgr_coset_extend(self.h, _p(d), coset, _p(out)))
        return out

    def msm(self, bases_mont, scalars_mont):
        b = np.ascontiguousarray(bases_mont, dtype=np.uint64)
        s = np.ascontiguousarray(scalars_mont, dtype=np.uint64)
        out = np.zeros(64, dtype=np.uint8)
        _check(lib().zgr_msm(self.h, _p(b), _p(s), s.size // 4, _p(out)))
        return out

    def ecfft(self, points, log_n, inverse=False):
        p = np.ascontiguousarray(points, dtype=np.uint8).reshape(-1)
        out = np.zeros_like(p)
        _check(lib().zgr_ecfft(self.h, _p(p), log_n, int(inverse), _p(out)))
        return out.reshape(-1, 64)

    def merkle(self, data, levels):
        d = np.ascontiguousarray(data, dtype=np.uint8).reshape(-1)
        top = sum(self.geo.coset_n << k for k in range(levels + 1))
        layers = np.zeros(top * 32, dtype=np.uint8)
        root = np.zeros(32, dtype=np.uint8)
        _check(lib().zgr_merkle(self.h, _p(d), d.size // 32, levels, _p(layers), _p(root)))
        return layers.reshape(-1, 32), bytes(root)

    def recover(self, row_ids, rows):
        """zgr_recover: row_ids (strictly increasing) and their rows, one bytes-like of
        32 * 2^log_col bytes each (or one concatenated buffer) -> the raw data."""
        buf = rows if isinstance(rows, (bytes, bytearray, memoryview)) else b"".join(bytes(r) for r in rows)
        arr = np.frombuffer(buf, dtype=np.uint8) if len(buf) else np.zeros(1, dtype=np.uint8)
        out = np.empty(self.geo.max_raw_data_size, dtype=np.uint8)
        n = self.recover_into(row_ids, arr, out)
        return out[:n].tobytes()

    def recover_into(self, row_ids, rows, out):
        """zgr_recover over caller-owned host arrays: rows a contiguous uint8 array of the
        rows in id order, out a writable uint8 array of at least MAX_RAW_DATA_SIZE bytes;
        returns the raw data's length (its bytes are out[:length])."""
        ids = np.ascontiguousarray(row_ids, dtype=np.uint32).reshape(-1)
        if rows.dtype != np.uint8 or not rows.flags.c_contiguous:
            raise ValueError("rows must be a contiguous uint8 array")
        if out.dtype != np.uint8 or not out.flags.c_contiguous or not out.flags.writeable:
            raise ValueError("out must be a writable contiguous uint8 array")
        if rows.size < ids.size * 32 * self.geo.col_n:
            raise ValueError("rows holds %d bytes, %d rows need %d" % (rows.size, ids.size, ids.size * 32 * self.geo.col_n))
        n = ctypes.c_size_t(0)
        _check(lib().zgr_recover(self.h, _p(ids if ids.size else np.zeros(1, dtype=np.uint32)),
                                 _p(rows if rows.size else np.zeros(1, dtype=np.uint8)), ids.size, _p(out), out.size,
                                 ctypes.byref(n)))
        return n.value

    def verify_msm(self, row_ids, rows, alphas):
        """zgr_verify_msm: row_ids (strictly increasing), their rows (canonical bytes, one
        buffer or a list) and one canonical alpha (int) per row -> 64 B canonical point."""
        ids = np.ascontiguousarray(row_ids, dtype=np.uint32).reshape(-1)
        if len(alphas) != ids.size:
            raise ValueError("%d alphas for %d rows" % (len(alphas), ids.size))
        if isinstance(rows, np.ndarray):
            arr = np.ascontiguousarray(rows, dtype=np.uint8).reshape(-1)
        else:
            buf = rows if isinstance(rows, (bytes, bytearray, memoryview)) else b"".join(bytes(r) for r in rows)
            arr = np.frombuffer(buf, dtype=np.uint8)
        if arr.size < ids.size * 32 * self.geo.col_n:
            raise ValueError("rows holds %d bytes, %d rows need %d" % (arr.size, ids.size, ids.size * 32 * self.geo.col_n))
        if not arr.size:
            arr = np.zeros(1, dtype=np.uint8)
        al = np.frombuffer(b"".join(int(a).to_bytes(32, "little") for a in alphas), dtype=np.uint8) \
            if len(alphas) else np.zeros(1, dtype=np.uint8)
        out = np.zeros(64, dtype=np.uint8)
        _check(lib().zgr_verify_msm(self.h, _p(ids if ids.size else np.zeros(1, dtype=np.uint32)), _p(arr), _p(al),
                                    ids.size, _p(out)))
        return bytes(out)

    # ---- autotuning ----
    def autotune(self):
        """zgr_autotune: time the bucket kernels' chunk-group plans per density class and
        the NTT split, keep the fastest; returns the tuner's report (a dict)."""
        import json

        buf = ctypes.create_string_buffer(8192)
        _check(lib().zgr_autotune(self.h, buf, len(buf)))
        return json.loads(buf.value.decode())

    # ---- profiling ----
    def inject_fault(self, site, countdown):
        """zgr_test_inject_fault: the countdown-th next event at `site` fails (0 disarms)."""
        _check(lib().zgr_test_inject_fault(self.h, site, countdown))

    def set_profiling(self, on):
        lib().zgr_set_profiling(self.h, int(on))

    def kernel_time_ms(self, name):
        cnt = ctypes.c_int(0)
        ms = lib().zgr_kernel_time_ms(self.h, name.encode(), ctypes.byref(cnt))
        return ms, cnt.value

    def kernel_time_reset(self):
        lib().zgr_kernel_time_reset(self.h)


def params_file_name(kind, log_n, h, coset=0):
    """The reference's file names (ptau_file_name, amtp_file_name, amtp_verify_file_name,
    crates/amt/src/utils.rs:11-52)."""
    if kind == FILE_PTAU:
        return "power-tau-%s-%02d.bin" % (BN254_HASH, log_n)
    if kind == FILE_PTAU_MONT:
        return "power-tau-mont-%s-%02d.bin" % (BN254_HASH, log_n)
    prefix = {FILE_AMT_MONT: "amt-prove-coset%d-mont", FILE_AMT: "amt-prove-coset%d",
              FILE_AMT_VERIFY: "amt-verify-coset%d"}[kind] % coset
    return "%s-%s-%02d-%02d.bin" % (prefix, BN254_HASH, h, log_n)


def build_params(ctx, out_dir, tau=None, verify_depth=None):
    """bin/build_params.rs:24-53 in create mode on the device: powers of tau (a fresh
    random tau unless one is given), then for every coset the prove bases and vanishes
    (AMTParams::from_pp) and the reference's files -- the compressed powers of tau, the
    unmont and mont (bamt) AMT params and the verify params.  The prove depth is the
    context's log_row and must equal verify_depth (build_params passes its verify depth
    as the prove depth).  Returns the written paths."""
    import secrets

    g = ctx.geo
    vd = g.h if verify_depth is None else verify_depth
    if vd != g.h:
        raise ValueError("build_params uses the verify depth as the prove depth: %d != log_row %d" % (vd, g.h))
    os.makedirs(out_dir, exist_ok=True)
    ctx.power_tau_setup(secrets.randbelow(R_MOD - 1) + 1 if tau is None else tau)
    written = []

    def put(kind, coset=0):
        path = os.path.join(out_dir, params_file_name(kind, g.log_n, vd, coset))
        with open(path, "wb") as f:
            f.write(ctx.write_params_file(kind, coset, vd))
        written.append(path)

    put(FILE_PTAU)
    for c in range(g.coset_n):
        ctx.params_from_power_tau(c)
        put(FILE_AMT, c)
        put(FILE_AMT_MONT, c)
        put(FILE_AMT_VERIFY, c)
    return written


def make_context(geo=PRODUCTION, fixed_base=True, group_plan=None, device=0, wide=True, wide_segs=None):
    """Context with its MSM path chosen explicitly.  zgr_ctx_create reads
    ZGR_FIXED_BASE: the default keeps shifted copies of the resident bases (one bucket
    set per chunk, FB_W = 22 copies: 22 x the params' HBM); "0" selects the windowed
    kernels.  ZGR_GROUP_PLAN="none" gives every AMT line one bucket set per chunk, and a
    comma-separated list ("2,2,1,0,...") sets each line's chunk-group level (log2 of the
    chunks sharing one bucket set), capped at what the line's outputs allow (default:
    the largest such groups, up to 4 chunks, with the high-basis line one level lower
    when that fits the longest waves into the device's wave slots).  ZGR_WIDE="0" keeps
    every line in 12-bit windows (default: the lines whose outputs are sums over groups
    of at least 16 chunks use 16-bit windows and one bucket set per group), and
    ZGR_WIDE_SEGS sets the number of bucket-segment waves per wide line.  All give
    identical results."""
    want = {"ZGR_FIXED_BASE": "1" if fixed_base else "0", "ZGR_GROUP_PLAN": group_plan,
            "ZGR_WIDE": None if wide else "0", "ZGR_WIDE_SEGS": None if wide_segs is None else str(wide_segs)}
    old = {k: os.environ.get(k) for k in want}
    try:
        for k, v in want.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
        return Context(geo, device)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


class Server:
    """zgr_server: continuous-batching encode service over one Context.

    submit() returns immediately with a ticket; poll() waits for it.  While the
    server exists it owns the context; close it before using the context again."""

    PENDING = 1

    def __init__(self, ctx, hbm_budget=1 << 30, max_inflight=64):
        self.ctx = ctx
        h = ctypes.c_void_p()
        _check(lib().zgr_server_create(ctx.h, hbm_budget, max_inflight, ctypes.byref(h)))
        self.h = h
        self._live = {}

    def submit(self, raw, want_data=True, out=None):
        out = out or EncodedBlobOutputs(self.ctx.geo, want_data)
        buf = ctypes.create_string_buffer(bytes(raw), max(1, len(raw)))
        cs = out.cstruct()
        t = ctypes.c_uint64()
        _check(lib().zgr_server_submit(self.h, buf, len(raw), ctypes.byref(cs), ctypes.byref(t)))
        self._live[t.value] = (buf, cs, out)
        return t.value

    def poll(self, ticket, timeout_ms=-1):
        """Returns the EncodedBlobOutputs once done, None while pending; raises on error."""
        rc = lib().zgr_server_poll(self.h, ticket, timeout_ms)
        if rc == self.PENDING:
            return None
        _, _, out = self._live.pop(ticket)
        _check(rc)
        return out

    def cancel(self, ticket):
        """zgr_server_cancel: True when the request was withdrawn before reaching the GPU
        (its poll then raises ZgrError with ERR_CANCELLED), False when it is already on
        the GPU or done (it completes as usual)."""
        rc = lib().zgr_server_cancel(self.h, ticket)
        if rc == OK:
            return True
        if rc == ERR_INVALID_ARG and ticket in self._live:
            return False
        raise ZgrError(rc, lib().zgr_last_error().decode(errors="replace"))

    def stats(self):
        q, a, c = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        lib().zgr_server_stats(self.h, ctypes.byref(q), ctypes.byref(a), ctypes.byref(c))
        return {"queued": q.value, "admitted": a.value, "completed": c.value}

    def close(self):
        if self.h:
            lib().zgr_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# mirror of the reference encode API
# ---------------------------------------------------------------------------
class EncodedBlob:
    """EncodedBlob (crates/encoder/src/encoder/blob.rs:26-62) backed by GPU outputs."""

    def __init__(self, geo, out):
        self.geo = geo
        self.out = out

    @classmethod
    def build(cls, ctx, raw):
        return cls(ctx.geo, ctx.encode(raw))

    def get_affine_commitment(self):
        return bytes(self.out.commitment)

    def get_commitment(self):
        return self.get_affine_commitment()

    def get_file_root(self):
        return bytes(self.out.file_root)

    def get_roots(self):
        return [bytes(self.out.merkle[32 * i:32 * i + 32]) for i in range(self.geo.coset_n)]

    def get_data(self):
        return self.out.data

    def merkle_layer(self, d):
        """EncodedBlobMerkle.tree[d] for d <= log_row."""
        g = self.geo
        off = sum(g.coset_n << k for k in range(d))
        cnt = g.coset_n << d
        return self.out.merkle[32 * off:32 * (off + cnt)].reshape(cnt, 32)

    def _tree_level(self, tree, coset, d):
        g = self.geo
        base = coset * g.tree_pts + (1 << (d + 1)) - 2
        return tree.reshape(-1, 64)[base:base + (1 << (d + 1))]

    def get_row(self, index):
        """EncodedBlob::get_row (blob.rs:44-49): (row data, AMT proof, Merkle proof)."""
        g = self.geo
        assert index < g.row_encoded
        coset, local = divmod(index, g.row_n)
        row = self.out.data[32 * (coset * g.n + local * g.col_n):32 * (coset * g.n + (local + 1) * g.col_n)]
        # AllProofs::get_proof with reversed_index = bitrev(local) (proofs.rs:16-31)
        rev = int(format(local, "0%db" % g.log_row)[::-1], 2)
        proof = []
        for d in range(g.h):
            lv = rev >> (g.h - 1 - d)
            proof.append((bytes(self._tree_level(self.out.comm_tree, coset, d)[lv ^ 1]),
                          bytes(self._tree_level(self.out.proof_tree, coset, d)[lv ^ 1])))
        high = bytes(self.out.high[64 * coset:64 * coset + 64])
        # EncodedBlobMerkle::get_row (merkle/blob.rs:44-57)
        mproof = [bytes(self.merkle_layer(d)[(index >> (g.log_row - d)) ^ 1]) for d in range(g.log_row, 0, -1)]
        return {"index": index, "row": bytes(row), "amt_proof": proof, "high_commitment": high,
                "commitment": self.get_affine_commitment(), "merkle_root": self.get_roots(),
                "merkle_proof": mproof, "row_root": bytes(self.merkle_layer(g.log_row)[index])}

    def slice_fields(self, index):
        """get_row(index) as the EncodedSlice fields (encoder/blob.rs:44-49 with
        EncodedSliceAMT / EncodedSliceMerkle::new), G1 values in wire form."""
        g = self.geo
        r = self.get_row(index)
        g1 = ark_g1_uncompressed
        return {"index": index, "amt_index": index, "commitment": g1(r["commitment"]),
                "row_index": index % g.row_n, "row": r["row"],
                "amt_proof": [(g1(c), g1(p)) for c, p in r["amt_proof"]],
                "high_commitment": g1(r["high_commitment"]), "merkle_root": r["merkle_root"],
                "merkle_proof": r["merkle_proof"], "leaf_index": index, "leaf": r["row_root"]}

    def encoded_slice(self, index):
        """serialize_uncompressed(get_row(index)) (grpc/src/service.rs:93-95, 162-166)."""
        return wire.serialize_encoded_slice(self.slice_fields(index))

    def encoded_slices(self):
        """encoded_slice(i) for every row at once, as a (row_encoded, slice bytes) uint8
        array: the same layout gathered with whole-array indexing."""
        g = self.geo
        R, h = g.row_encoded, g.h
        idx = np.arange(R, dtype=np.uint64)
        coset, local = idx // g.row_n, idx % g.row_n
        rev = np.zeros(R, dtype=np.uint64)
        for b in range(h):
            rev |= ((local >> np.uint64(b)) & np.uint64(1)) << np.uint64(h - 1 - b)
        u64 = lambda v: np.ascontiguousarray(v.astype("<u8")).view(np.uint8).reshape(R, 8)  # noqa: E731
        ctree = ark_g1_uncompressed_array(self.out.comm_tree.reshape(-1, 64))
        ptree = ark_g1_uncompressed_array(self.out.proof_tree.reshape(-1, 64))
        cols = [u64(idx), u64(idx), np.broadcast_to(ark_g1_uncompressed_array(self.out.commitment.reshape(1, 64)), (R, 64)),
                u64(local), u64(np.full(R, g.col_n, dtype=np.uint64)),
                self.out.data.reshape(R, g.col_n * 32), u64(np.full(R, h, dtype=np.uint64))]
        for d in range(h):
            at = (coset * np.uint64(g.tree_pts) + np.uint64((2 << d) - 2) + ((rev >> np.uint64(h - 1 - d)) ^ np.uint64(1)))
            at = at.astype(np.int64)
            cols += [ctree[at], ptree[at]]
        high = ark_g1_uncompressed_array(self.out.high.reshape(-1, 64))
        cols.append(high[coset.astype(np.int64)])
        cols.append(np.broadcast_to(self.merkle_layer(0).reshape(1, -1), (R, 32 * g.coset_n)))
        cols.append(u64(np.full(R, g.log_row, dtype=np.uint64)))
        for d in range(g.log_row, 0, -1):
            cols.append(self.merkle_layer(d)[((idx >> np.uint64(g.log_row - d)) ^ np.uint64(1)).astype(np.int64)])
        cols += [u64(idx), self.merkle_layer(g.log_row)]
        return np.concatenate([np.asarray(c, dtype=np.uint8) for c in cols], axis=1)

    def light_slice(self, index):
        """serialize_uncompressed(get_row(index).into_light_slice()) (encoder/slice.rs:117-119)."""
        return wire.serialize_light_slice(self.slice_fields(index))


def ark_g1_uncompressed(p64):
    """ark-serialize 0.4 uncompressed G1 affine from a canonical x || y (64 zero bytes =
    infinity): x, then y with the SW flags in its spare top bits -- 0x40 in the last
    byte for infinity (x = y = 0), 0x80 when y > -y, i.e. y > (p - 1) / 2."""
    b = bytearray(p64)
    if not any(b):
        b[63] = 0x40
    elif int.from_bytes(bytes(b[32:]), "little") > (P_MOD - 1) // 2:
        b[63] |= 0x80
    return bytes(b)


class RecoveryError(ValueError):
    """An Err(String) of recover_from_da_slice; str() is the reference's text."""


class RecoveryPanic(RuntimeError):
    """An input on which the reference panics (an element >= r, or rows off one
    polynomial of degree <= N)."""


def recover_from_da_slice(ctx, slices):
    """recover_from_da_slice (crates/recovery/src/lib.rs:15-43) on the device.

    slices: {row index: row bytes} (EncodedSlice::merkle_row, 32 B canonical per
    element).  Rows of the wrong length are dropped (lib.rs:21-32); the rest go to
    zgr_recover in key order.  Returns RawData::as_bytes; raises RecoveryError with
    the reference's message, or RecoveryPanic where the reference panics."""
    width = 32 * ctx.geo.col_n
    ok = sorted((int(k), bytes(v)) for k, v in slices.items() if len(v) == width)
    try:
        return ctx.recover([k for k, _ in ok], b"".join(v for _, v in ok))
    except ZgrError as e:
        msg = e.msg
        if e.code == ERR_INCONSISTENT or "panics" in msg:
            raise RecoveryPanic(msg) from None
        if e.code == ERR_INVALID_ARG:
            raise RecoveryError(msg) from None
        raise


class VerifierError(ValueError):
    """A VerifierError of EncodedSlice::verify (crates/encoder/src/encoder/error.rs);
    str() is the reference's Debug text."""


class DeferredVerifier:
    """The MSM half of DeferredVerifier (crates/amt/src/deferred_verification.rs:34-222)
    on the device, for the rows of one encoded blob.

    record_slice() takes EncodedSlice bytes and runs the checks EncodedSlice::verify
    makes before its AMT proof check (crates/encoder/src/encoder/slice.rs:72-95,
    crates/encoder/src/amt/slice.rs:51-88, and verify_amt_proof's length and position
    checks, crates/amt/src/verify_params.rs:149-159), raising VerifierError with the
    reference's text; then it draws a random alpha for the row and keeps
    alpha * (commitment - sum_d proof[d].commitment), as record_msm does after
    verify_amt_proof's pairing checks (crates/amt/src/verify_params.rs:176-186).
    fast_check_msm() (fast_check_msm_gpu, deferred_verification.rs:203-222) compares the
    device MSM of the recorded rows over the resident basis with the device MSM of the
    recorded answers.  The pairing half (fast_check_pairing) and the Merkle proof stay
    with the caller."""

    def __init__(self, ctx, commitment, rng=None):
        """commitment: the authoritative commitment, canonical affine x || y (64 B, as
        EncodedBlob.get_affine_commitment returns it)."""
        import secrets

        self.ctx = ctx
        self.commitment = bytes(commitment)
        if len(self.commitment) != 64:
            raise ValueError("commitment must be 64 bytes (canonical x || y)")
        self._rand = rng if rng is not None else (lambda: secrets.randbelow(R_MOD))
        self._rows = {}

    def check_slice(self, f):
        """The reference's checks on one parsed slice, in its order; raises VerifierError."""
        g = self.ctx.geo
        idx, amt_index = f["index"], f["amt_index"]
        if idx != amt_index:
            raise VerifierError("UnmatchedAMTIndex { row_index: %d, amt_index: %d }" % (idx, amt_index))
        if idx != f["leaf_index"]:
            raise VerifierError("UnmatchedMerkleIndex { row_index: %d, merkle_index: %d }" % (idx, f["leaf_index"]))
        if _g1_from_wire(f["commitment"]) != _g1_from_canon(self.commitment):
            raise VerifierError("AMT(IncorrectCommitment)")
        if len(f["row"]) != g.col_n:
            raise VerifierError("AMT(IncorrectRowSize { actual: %d, expected: %d })" % (len(f["row"]), g.col_n))
        if amt_index >= g.row_encoded:
            raise VerifierError("AMT(RowIndexOverflow { actual: %d, expected_max: %d })" % (amt_index, g.row_encoded))
        coset, local = divmod(amt_index, g.row_n)
        if local != f["row_index"]:
            raise VerifierError("AMT(UnmatchedCosetIndex { coset_index: %d, local_index: %d, amt_index: %d })"
                                % (coset, local, f["row_index"]))
        depth = len(f["amt_proof"])
        # verify_amt_proof: batch = basis.len() >> depth must be the row length, and the
        # row's bit-reversed index must name one of the 2^depth batches
        err = None
        if depth > g.log_n or (g.n >> depth) != g.col_n:
            err = "UnexpectedDataLength"
        elif f["row_index"] >= (1 << depth):
            err = "IncorrectPosition"
        if err:
            raise VerifierError("AMT(IncorrectProof { coset_index: %d, amt_index: %d, error: %s })"
                                % (coset, f["row_index"], err))

    def record_slice(self, slice_bytes):
        f = wire.parse_encoded_slice(slice_bytes, self.ctx.geo.coset_n)
        self.check_slice(f)
        idx = f["index"]
        if idx in self._rows:
            raise ValueError("row %d recorded twice" % idx)
        self._rows[idx] = (b"".join(f["row"]), self._rand(), f["commitment"], [c for c, _ in f["amt_proof"]])

    def fast_check_msm(self):
        if not self._rows:
            return True
        ids = sorted(self._rows)
        lhs = self.ctx.verify_msm(ids, [self._rows[i][0] for i in ids], [self._rows[i][1] for i in ids])
        # sum_k alpha_k (C_k - sum_d P_(k,d)) with each distinct point once: the rows of a
        # blob share its commitment and the AMT nodes above them (a depth-d proof point
        # serves 2^(h-d) rows), so the scalars of equal points are summed first (the MSM
        # is linear: a P + b P = (a + b) P) -- about 6,000 points for a blob's 3,072 rows
        # instead of 33,800
        merged = {}
        for i in ids:
            _, alpha, com, proof = self._rows[i]
            for pt, k in [(com, alpha)] + [(p, R_MOD - alpha) for p in proof]:
                key = bytes(pt)
                merged[key] = merged.get(key, 0) + k
        bases, scal = [], []
        for key, k in merged.items():
            xy = _g1_from_wire(key)
            if xy is not None and k % R_MOD:
                bases.append(xy)
                scal.append(k % R_MOD)
        if not bases:
            return lhs == bytes(64)
        R256 = 1 << 256
        bm = b"".join((x * R256 % P_MOD).to_bytes(32, "little") + (y * R256 % P_MOD).to_bytes(32, "little")
                      for x, y in bases)
        sm = b"".join((k * R256 % R_MOD).to_bytes(32, "little") for k in scal)
        rhs = self.ctx.msm(np.frombuffer(bm, dtype=np.uint64), np.frombuffer(sm, dtype=np.uint64))
        return lhs == bytes(rhs)


def _g1_from_canon(b):
    """(x, y) ints of a canonical affine x || y, None for infinity (64 zero bytes)."""
    if not any(b):
        return None
    return int.from_bytes(bytes(b[:32]), "little"), int.from_bytes(bytes(b[32:64]), "little")


def _g1_from_wire(b):
    """(x, y) ints of an ark-uncompressed G1 point, None for infinity."""
    raw = bytearray(b)
    flags = raw[63] & 0xC0
    raw[63] &= 0x3F
    if flags & 0x40:
        return None
    return int.from_bytes(bytes(raw[:32]), "little"), int.from_bytes(bytes(raw[32:]), "little")


def ark_g1_uncompressed_array(pts):
    """ark_g1_uncompressed over an (n, 64) uint8 array of canonical points."""
    out = np.array(pts, dtype=np.uint8, copy=True).reshape(-1, 64)
    y = out[:, 32:].copy().view("<u8")  # (n, 4) little-endian limbs
    half = [((P_MOD - 1) // 2 >> (64 * k)) & ((1 << 64) - 1) for k in range(4)]
    gt = np.zeros(len(out), dtype=bool)
    eq = np.ones(len(out), dtype=bool)
    for k in range(3, -1, -1):
        hk = np.uint64(half[k])
        gt |= eq & (y[:, k] > hk)
        eq &= y[:, k] == hk
    zero = ~out.any(axis=1)
    out[gt, 63] |= 0x80
    out[zero, 63] = 0x40
    return out


class EncoderService:
    """EncoderService::process_data (grpc/src/service.rs:67-106).

    With a Server, requests from several threads (the gRPC workers) share the
    device through its admission queue and overlap; without one, each call encodes
    on the context directly."""

    VERSION = 0  # service.rs:98

    def __init__(self, ctx, server=None):
        self.ctx = ctx
        self.server = server

    POLL_MS = 20  # how often a served request checks that its caller is still there

    def _build(self, data, require_data, alive=None):
        g = self.ctx.geo
        if len(data) > g.max_raw_data_size:  # RawData::try_from (raw_data.rs:55-73)
            raise TooLargeBlob(len(data), g.max_raw_data_size)
        if self.server is None:
            return EncodedBlob(g, self.ctx.encode(data))
        # the slices need the encoded rows even when the reply omits them
        t = self.server.submit(data, True)
        while True:
            out = self.server.poll(t, self.POLL_MS if alive is not None else -1)
            if out is not None:
                return EncodedBlob(g, out)
            # the caller went away: withdraw the request unless it is already on the GPU
            if not alive() and self.server.cancel(t):
                try:
                    self.server.poll(t)
                except ZgrError as e:
                    if e.code != ERR_CANCELLED:
                        raise
                raise RequestCancelled("request withdrawn before it reached the GPU")

    def process_data(self, data, require_data):
        """-> dict with the EncodeBlobReply fields (encoded_slice as a list of bytes)
        and the EncodedBlob under "blob"."""
        blob = self._build(data, require_data)
        rows = blob.encoded_slices()
        return {
            "version": self.VERSION,
            "erasure_commitment": blob.get_affine_commitment(),  # x || y, 32 B LE each (service.rs:77-83)
            "storage_root": blob.get_file_root(),
            "encoded_data": bytes(blob.get_data()) if require_data else b"",
            "encoded_slice": [r.tobytes() for r in rows],
            "blob": blob,
        }

    def encode_blob(self, data, require_data, alive=None):
        """The EncodeBlobReply protobuf bytes for one request (the gRPC handler's
        body), serialized straight from the gathered slice array.  alive (the gRPC
        context's is_active) lets a served request be withdrawn when its client has
        gone: RequestCancelled is raised then."""
        blob = self._build(data, require_data, alive)
        return wire.encode_reply({
            "version": self.VERSION,
            "erasure_commitment": blob.get_affine_commitment(),
            "storage_root": blob.get_file_root(),
            "encoded_data": blob.get_data() if require_data else b"",
            "encoded_slice": blob.encoded_slices(),
        })
