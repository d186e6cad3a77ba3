"""The kernel file layout (include/msr.h, kernel persistence section) written
out independently in numpy: a reader and a writer used by the tests to check
that files the library writes hold the documented arrays, and that the
library accepts a file made from the description and refuses damaged ones."""
import numpy as np

MAGIC = b"MSRTOEPL"
C0, C1, S1 = 0x9E3779B97F4A7C15, 0xC2B2AE3D27D4EB4F, 0x5851F42D4C957F2D
HEADER = np.dtype([("magic", "S8"), ("version", "<u4"), ("header_bytes", "<u4"), ("d", "<u4"),
                   ("K", "<u4"), ("P", "<u4"), ("lambda_real", "<u4"), ("N", "<i8", 3),
                   ("nkm", "<i8"), ("key", "<u8", 2), ("payload_bytes", "<u8"),
                   ("payload_hash", "<u8", 2), ("header_hash", "<u8"), ("reserved", "u1", 16)])
assert HEADER.itemsize == 128


def _u64(x):
    return np.array(x, dtype=np.uint64)


def mix64(z):
    """splitmix64 finaliser (Steele, Lea, Flood; Vigna's constants), mod 2^64."""
    z = _u64(z)
    z = (z ^ (z >> _u64(30))) * _u64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> _u64(27))) * _u64(0x94D049BB133111EB)
    return z ^ (z >> _u64(31))


def word_hash(data: bytes, seed: int):
    """Two-lane sum over 64-bit little-endian words w_i (last one zero-padded)
    of mix64(w_i ^ (i*c + s)), finished with the byte count."""
    n = len(data)
    buf = data + b"\0" * ((-n) % 8)
    w = np.frombuffer(buf, dtype="<u8").astype(np.uint64)
    i = np.arange(w.size, dtype=np.uint64)
    with np.errstate(over="ignore"):
        h0 = mix64(w ^ (i * _u64(C0) + _u64(seed))).sum(dtype=np.uint64)
        h1 = mix64(w ^ (i * _u64(C1) + (_u64(seed) ^ _u64(S1)))).sum(dtype=np.uint64)
        nb = _u64(n)
        return (int(mix64(_u64(h0) ^ mix64(nb))), int(mix64(_u64(h1) + mix64(nb ^ _u64(S1)))))


def write(path, img_shape, lam, kmask, key=(0, 0)):
    """lam: (nkm, P) complex64 or float32 (the row-major view of Julia's
    (P, Nkm)); kmask: (nkm,) int64, 1-based."""
    lam = np.ascontiguousarray(lam)
    kmask = np.ascontiguousarray(kmask, dtype="<i8")
    nkm, P = lam.shape
    K = int((np.sqrt(8 * P + 1) - 1) // 2)
    real = not np.iscomplexobj(lam)
    payload = kmask.tobytes() + lam.astype(np.float32 if real else np.complex64).tobytes()
    h = np.zeros((), dtype=HEADER)
    h["magic"] = MAGIC
    h["version"] = 1
    h["header_bytes"] = 128
    h["d"] = len(img_shape)
    h["K"] = K
    h["P"] = P
    h["lambda_real"] = int(real)
    h["N"] = list(img_shape) + [1] * (3 - len(img_shape))
    h["nkm"] = nkm
    h["key"] = key
    h["payload_bytes"] = len(payload)
    h["payload_hash"] = word_hash(payload, 4)
    h["header_hash"] = word_hash(h.tobytes()[:104], 3)[0]
    with open(path, "wb") as f:
        f.write(h.tobytes())
        f.write(payload)


def read(path):
    """(header dict, lam (nkm, P), kmask (nkm,)) with both checksums checked."""
    raw = open(path, "rb").read()
    h = np.frombuffer(raw[:128], dtype=HEADER)[0]
    assert bytes(h["magic"]) == MAGIC and int(h["version"]) == 1
    assert word_hash(raw[:104], 3)[0] == int(h["header_hash"])
    payload = raw[128:]
    assert len(payload) == int(h["payload_bytes"])
    assert word_hash(payload, 4) == tuple(int(x) for x in h["payload_hash"])
    nkm, P = int(h["nkm"]), int(h["P"])
    kmask = np.frombuffer(payload[:8 * nkm], dtype="<i8")
    dt = np.float32 if int(h["lambda_real"]) else np.complex64
    lam = np.frombuffer(payload[8 * nkm:], dtype=dt).reshape(nkm, P)
    d = int(h["d"])
    info = {"img_shape": tuple(int(x) for x in h["N"][:d]), "Ncoeff": int(h["K"]),
            "lambda_real": bool(h["lambda_real"]), "nkm": nkm,
            "key": f"{int(h['key'][0]):016x}{int(h['key'][1]):016x}"}
    return info, lam, kmask
