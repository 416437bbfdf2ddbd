"""The framing oracle pinned by the reference's vectors and by zlib, and the
product's framed_size (a host function of the C ABI) against it."""
import numpy as np
import pytest

from helpers import EOF_BLOCK, bgzf

from oracle import bgzf_write as ow


def test_reference_vectors():
    # crates/fritillaria-core/src/compress.rs tests
    assert ow.MAX_COMPRESSIBLE_PAYLOAD == 65505
    assert ow.framed_size(ow.MAX_COMPRESSIBLE_PAYLOAD, 1 << 62) == ow.MAX_BLOCK_SIZE
    assert ow.choose_framing(65280, 148_256) == ow.STORED
    assert ow.framed_size(65280, 148_256) == 18 + 65280 + 5 + 8
    assert ow.choose_framing(1000, 1006) == ow.STORED
    assert ow.framed_size(1000, 1006) == 18 + 1005 + 8
    assert ow.choose_framing(1000, 1005) == ow.DEFLATED
    assert ow.choose_framing(1000, 400) == ow.DEFLATED
    assert ow.framed_size(1000, 400) == 18 + 400 + 8
    for claimed in [0, 2, 5, 999, 148_256]:
        assert ow.choose_framing(0, claimed) == ow.EMPTY
        assert ow.framed_size(0, claimed) == 28


def test_framing_bytes_are_what_zlib_reads():
    assert ow.empty_block() == ow.EOF_BLOCK == EOF_BLOCK
    rng = np.random.default_rng(1)
    for n in [1, 2, 100, 65505]:
        payload = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        s = ow.stored_block(payload)
        assert len(s) == ow.framed_size(n, 1 << 40)
        m = ow.members(s)
        assert len(m) == 1 and m[0][2] == payload and m[0][1] == len(s)
    # members of a zlib-written BGZF stream, and a corrupted one refused
    data = rng.integers(0, 4, 200_000, dtype=np.uint8).tobytes()
    stream = bgzf(data)
    ms = ow.members(stream)
    assert b"".join(m[2] for m in ms) == data and ms[-1][1] == 28
    bad = bytearray(stream)
    bad[ms[0][1] - 8] ^= 1  # CRC of the first member
    with pytest.raises(Exception):
        ow.members(bytes(bad))


def test_product_framed_size_matches():
    import fritillaria_amd as fa
    rng = np.random.default_rng(2)
    for _ in range(2000):
        n = int(rng.integers(0, 65506))
        d = int(rng.choice([0, n // 3, n + 4, n + 5, n + 6, 65510, 65511, 150_000]))
        assert fa.framed_size(n, d) == ow.framed_size(n, d), (n, d)
