"""The database image format through the library's reader, without a device: an image written by
an independent writer (tests/image_format.py, the documented layout) parses and reaches the build
(which then stops at BackendUnavailable here, or builds on a device), and every kind of damage is
rejected before anything reaches the device."""
import os
import struct

import pytest

import hyalite_amd as hy
from common import dna_scheme
from image_format import MAGIC, image_bytes, reseal

SEQS = [[0, 1, 2, 3], [], [3, 3, 1], [2] * 9]


def good():
    al, mat, go, ge = dna_scheme()
    return image_bytes(SEQS, al, mat, go, ge, 0, 1, 0, 12)


def load(tmp_path, data):
    p = tmp_path / "db.img"
    p.write_bytes(data)
    return hy.Database.load(str(p), hy.BackendChoice.force(hy.Backend.Hip))


def expect_parsed(tmp_path, data):
    try:
        load(tmp_path, data)
    except hy.BackendUnavailable:
        pass


def test_independent_writer_image_parses(tmp_path):
    expect_parsed(tmp_path, good())


@pytest.mark.parametrize("damage", ["magic", "version", "truncate", "extra", "flip_residue",
                                    "flip_sum", "empty"])
def test_damaged_images_are_rejected(tmp_path, damage):
    data = bytearray(good())
    if damage == "magic":
        data[0:8] = b"NOTANIMG"
    elif damage == "version":
        data[8:12] = struct.pack("<I", 2)
    elif damage == "truncate":
        data = data[:-9]
    elif damage == "extra":
        data += b"\0"
    elif damage == "flip_residue":
        data[-9] ^= 1
    elif damage == "flip_sum":
        data[-1] ^= 0x80
    elif damage == "empty":
        data = bytearray()
    with pytest.raises(hy.DatabaseImageError):
        load(tmp_path, bytes(data))


def test_resealed_bad_offsets_are_rejected(tmp_path):
    """A consistent checksum does not make inconsistent offsets acceptable."""
    al, mat, go, ge = dna_scheme()
    data = bytearray(good()[:-8])
    hdr = 64 + al * al * 4
    # offsets[1] > offsets[2]: not non-decreasing
    struct.pack_into("<Q", data, hdr + 8, 9)
    with pytest.raises(hy.DatabaseImageError):
        load(tmp_path, reseal(bytes(data)))
    data = bytearray(good()[:-8])
    struct.pack_into("<Q", data, hdr + 8 * len(SEQS), 999)  # last offset != residue count
    with pytest.raises(hy.DatabaseImageError):
        load(tmp_path, reseal(bytes(data)))


def test_image_contents_are_validated_like_a_build(tmp_path):
    """Symbols, scheme and width are the build's checks, with the build's errors."""
    al, mat, go, ge = dna_scheme()
    with pytest.raises(hy.SymbolOutOfRange) as e:
        load(tmp_path, image_bytes([[0, 7]], al, mat, go, ge, 0, 1, 0, 8))
    assert (e.value.symbol, e.value.alphabet_len) == (7, 4)
    with pytest.raises(hy.GapOpenLessThanExtend):
        load(tmp_path, image_bytes([[0]], al, mat, 1, 2, 0, 1, 0, 8))
    with pytest.raises(hy.ScoreRangeTooWide):
        load(tmp_path, image_bytes([[0]], 2, [0, 0, 0, 0], 2_000_000_000, 1, 1, 0, 0, 1))
    with pytest.raises(hy.TracebackBudgetExceeded):
        load(tmp_path, image_bytes([[0] * 8], al, mat, go, ge, 1, 2, 1, 8))


def test_missing_file_is_an_os_error(tmp_path):
    with pytest.raises(OSError):
        hy.Database.load(str(tmp_path / "absent.img"))
