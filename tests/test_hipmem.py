"""Host-side logic of simwave_amd/hipmem.py on the CPU: stream handles and
the pinned staging pool (its page-locked allocation replaced by plain
memory, so no device is needed)."""
import numpy as np
import pytest

from simwave_amd import hipmem


class _FakePinned:
    made = []

    def __init__(self, lib, nbytes):
        self.nbytes = int(nbytes)
        self.array = np.zeros(self.nbytes, dtype=np.uint8)
        _FakePinned.made.append(self)


@pytest.fixture
def fake_pool(monkeypatch):
    _FakePinned.made = []
    monkeypatch.setattr(hipmem, "_Pinned", _FakePinned)
    monkeypatch.setattr(hipmem, "_POOL", [])
    monkeypatch.setattr(hipmem, "_POOL_KEEP", 64 << 20)
    return hipmem


def test_stream_handles():
    class TorchLike:
        cuda_stream = 1234

    assert hipmem.stream_handle(None) is None
    assert hipmem.stream_handle(TorchLike()) == 1234
    assert hipmem.stream_handle(77) == 77


def test_staging_reuses_buffers_and_bounds_the_pool(fake_pool):
    with fake_pool.staging(None, 3 << 20) as a:
        assert a.nbytes >= 3 << 20
        assert a.nbytes == 4 << 20                 # powers of two
    with fake_pool.staging(None, 1 << 20) as b:
        assert b is a                              # the idle buffer is reused
        with fake_pool.staging(None, 1 << 20) as c:
            assert c is not b and c.nbytes == 1 << 20   # in use: a new one
    assert len(_FakePinned.made) == 2
    # smallest idle buffer that fits
    with fake_pool.staging(None, 100) as d:
        assert d.nbytes == 1 << 20
    # a buffer that would take the idle total past the limit is not kept
    with fake_pool.staging(None, 80 << 20):
        pass
    assert sum(p.nbytes for p in fake_pool._POOL) <= 64 << 20
    assert len(fake_pool._POOL) == 2
