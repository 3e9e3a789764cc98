"""GEIG file format (sunny_gen_eigensolver_amd/geig.py): version 1 is byte-compatible
with the reference's fixture files (tests/common/lapack_reference.cpp:257-279), checked
against the oracle's independent C++ reader/writer and the reference fixtures'
SHA-256 of H and S; version 2 (uneven batches with stored info) round-trips and
uses the uneven-batch path's packing."""
import hashlib
import os

import numpy as np
import pytest

from oracle import oracle as O
from sunny_gen_eigensolver_amd import geig
from tests import helpers as T


def _rand_c(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def test_v1_matches_oracle_both_ways(tmp_path):
    rng = np.random.default_rng(3)
    n, b = 7, 5
    H, S, U = _rand_c(rng, b, n, n), _rand_c(rng, b, n, n), _rand_c(rng, b, n, n)
    W = rng.standard_normal((b, n))
    p1 = str(tmp_path / "pkg.bin")
    geig.write_v1(p1, 42, H, S, W, U)
    got = O.geig_read(p1)
    assert (got["n"], got["batch"], got["seed"]) == (n, b, 42)
    for k, a in (("H", H), ("S", S), ("W", W), ("U", U)):
        assert np.array_equal(got[k], a), k
    p2 = str(tmp_path / "oracle.bin")
    O.geig_write(p2, n, b, 42, H, S, W, U)
    assert open(p1, "rb").read() == open(p2, "rb").read()
    back = geig.read(p2)
    assert back["version"] == 1 and (back["n"], back["batch"], back["seed"]) == (n, b, 42)
    for k, a in (("H", H), ("S", S), ("W", W), ("U", U)):
        assert np.array_equal(back[k], a), k


@pytest.mark.parametrize("name", ["n16_b8_s42_random", "n54_b8_s42_nearid", "n54_b32_s123_random"])
def test_v1_sections_hash_like_reference_fixtures(tmp_path, name):
    """The H and S sections of a file written from the oracle's regenerated
    fixture inputs carry the reference fixtures' SHA-256."""
    meta, H, S = T.fixture_inputs(name)
    b, n = H.shape[0], H.shape[1]
    p = str(tmp_path / f"{name}.bin")
    geig.write_v1(p, meta["seed"], H, S, np.zeros((b, n)), np.zeros((b, n, n), dtype=np.complex128))
    raw = open(p, "rb").read()
    m = 16 * n * n * b
    assert hashlib.sha256(raw[24:24 + m]).hexdigest() == meta["sha256_H"]
    assert hashlib.sha256(raw[24 + m:24 + 2 * m]).hexdigest() == meta["sha256_S"]


def test_v2_round_trip_and_packing(tmp_path):
    rng = np.random.default_rng(9)
    orders = [3, 12, 1, 54, 7]
    Hs = [_rand_c(rng, n, n) for n in orders]
    Ss = [_rand_c(rng, n, n) for n in orders]
    Us = [_rand_c(rng, n, n) for n in orders]
    Ws = [rng.standard_normal(n) for n in orders]
    info = np.array([0, 0, 5, 0, 60], dtype=np.int32)
    pack = lambda ms: np.concatenate([m.reshape(-1) for m in ms])
    p = str(tmp_path / "v2.bin")
    geig.write_v2(p, 77, orders, pack(Hs), pack(Ss), pack(Ws), pack(Us), info)
    d = geig.read(p)
    assert d["version"] == 2 and d["seed"] == 77 and list(d["orders"]) == orders
    assert np.array_equal(d["H"], pack(Hs)) and np.array_equal(d["S"], pack(Ss))
    assert np.array_equal(d["W"], pack(Ws)) and np.array_equal(d["U"], pack(Us))
    assert np.array_equal(d["info"], info)
    # Pencil k sits at element offset sum_{j<k} n_j^2, as the uneven-batch path packs it.
    off = sum(n * n for n in orders[:3])
    assert np.array_equal(d["H"][off:off + 54 * 54].reshape(54, 54), Hs[3])
    # Inputs only.
    p2 = str(tmp_path / "v2in.bin")
    geig.write_v2(p2, 1, orders, pack(Hs), pack(Ss))
    d2 = geig.read(p2)
    assert "W" not in d2 and np.array_equal(d2["S"], pack(Ss))
    assert os.path.getsize(p2) == 24 + 4 * len(orders) + 2 * 16 * sum(n * n for n in orders)


def test_rejects_malformed(tmp_path):
    p = str(tmp_path / "bad.bin")
    open(p, "wb").write(np.array([0x12345678, 1, 2, 1, 0, 0], dtype="<i4").tobytes())
    with pytest.raises(ValueError, match="magic"):
        geig.read(p)
    geig.write_v1(p, 0, np.zeros((1, 2, 2)), np.zeros((1, 2, 2)), np.zeros((1, 2)), np.zeros((1, 2, 2)))
    raw = open(p, "rb").read()
    open(p, "wb").write(raw[:-8])
    with pytest.raises(ValueError, match="truncated"):
        geig.read(p)
    open(p, "wb").write(raw + b"\0")
    with pytest.raises(ValueError, match="trailing"):
        geig.read(p)
    with pytest.raises(ValueError):
        geig.write_v2(p, 0, [2, 3], np.zeros(4), np.zeros(13))
