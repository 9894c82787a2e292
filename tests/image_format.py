"""An independent writer of the database image format (the layout documented in
hyalite_amd/csrc/hy_image.cpp), used to pin the library's reader and writer."""
import struct

import numpy as np

MAGIC = b"HYDBIMG1"


def checksum(data: bytes) -> int:
    """The FNV-1a step over little-endian 64-bit words, the last one zero-padded."""
    pad = (-len(data)) % 8
    words = np.frombuffer(data + b"\0" * pad, dtype="<u8")
    h = 0xCBF29CE484222325
    for w in words.tolist():
        h = ((h ^ w) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def image_bytes(seqs, al, matrix, go, ge, mode, search_type, budget, mql) -> bytes:
    lens = [len(s) for s in seqs]
    off = np.zeros(len(seqs) + 1, dtype="<u8")
    np.cumsum(lens, out=off[1:])
    res = b"".join(bytes(bytearray(int(x) for x in s)) for s in seqs)
    body = MAGIC + struct.pack("<IIiiIIQQQQ", 1, al, go, ge, mode, search_type, budget, mql,
                               len(seqs), len(res))
    body += np.asarray(matrix, dtype="<i4").tobytes() + off.tobytes() + res
    return body + struct.pack("<Q", checksum(body))


def reseal(body_without_sum: bytes) -> bytes:
    return body_without_sum + struct.pack("<Q", checksum(body_without_sum))
