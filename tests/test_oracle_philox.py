"""Pin oracle.philox against the Random123 known-answer vectors (philox4x32_10)."""
import numpy as np

from oracle import philox

KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


def test_known_answers():
    for ctr, key, want in KAT:
        got = philox.philox4x32_10(np.array(ctr, dtype=np.uint32), np.array(key, dtype=np.uint32))
        assert [int(v) for v in got] == want


def test_uniform_open_interval_and_float32_exact():
    x = np.array([0, 1, 511, 512, 0xFFFFFFFF], dtype=np.uint32)
    u = philox.to_uniform(x)
    assert np.all(u > 0) and np.all(u < 1)
    assert np.all(u.astype(np.float32).astype(np.float64) == u)


def test_block_moments():
    u = philox.block(7, np.arange(20000, dtype=np.uint64), 3, philox.PURPOSE_STEP)
    assert abs(u.mean() - 0.5) < 0.005
    assert abs(u.var() - 1 / 12) < 0.002
