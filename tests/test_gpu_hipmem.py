"""The library's memory, copy, zero, stream and event entry points on the
GPU (simwave_amd/hipmem.py over simwave_hip_device_alloc ...
simwave_hip_event_*): host -> device -> host round trips on library
streams, the zero fill, ordering across two streams through a host wait on
an event, and the pinned staging pool."""
import ctypes

import numpy as np
import pytest

from simwave_amd import hipmem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lib():
    return hipmem.default_library()


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int64])
@pytest.mark.parametrize("shape", [(1,), (7, 13), (3, 41, 29)])
def test_round_trip_and_zero(lib, shape, dtype):
    print("hipmem round trip", shape, np.dtype(dtype).name, flush=True)
    rng = np.random.default_rng(1)
    host = (rng.standard_normal(shape) * 1e3).astype(dtype)
    st = hipmem.Stream(lib)
    dev = hipmem.DeviceArray.from_host(lib, host, st)
    assert dev.shape == host.shape and dev.nbytes == host.nbytes
    back = dev.numpy(st)
    assert np.array_equal(back, host)
    dev.zero_(st)
    assert not dev.numpy(st).any()
    # device -> device through the copy entry, then back
    other = hipmem.DeviceArray(lib, shape, dtype, st)
    dev.copy_from_host(host, st)
    hipmem._call(lib, "simwave_hip_copy", ctypes.c_void_p(other.data_ptr()),
                 ctypes.c_void_p(dev.data_ptr()), ctypes.c_ulonglong(dev.nbytes),
                 ctypes.c_void_p(st.handle))
    assert np.array_equal(other.numpy(st), host)


def test_event_orders_work_across_streams(lib):
    """Work on stream a, an event recorded behind it and waited on the host,
    then a copy on stream b sees the result."""
    print("hipmem events", flush=True)
    n = 1 << 20
    host = np.arange(n, dtype=np.float32)
    a, b = hipmem.Stream(lib), hipmem.Stream(lib)
    dev = hipmem.DeviceArray.from_host(lib, host, a)
    dev.zero_(a)
    ev = a.record_event()
    ev.synchronize()
    assert not dev.numpy(b).any()


def test_staging_pool_with_pinned_memory(lib):
    print("hipmem staging", flush=True)
    host = np.arange(3 << 18, dtype=np.float32)
    st = hipmem.Stream(lib)
    dev = hipmem.DeviceArray.from_host(lib, host, st)
    for _ in range(3):
        with hipmem.staging(lib, dev.nbytes) as buf:
            assert buf.nbytes >= dev.nbytes
            hipmem._call(lib, "simwave_hip_copy", ctypes.c_void_p(buf.ctypes.data),
                         ctypes.c_void_p(dev.data_ptr()), ctypes.c_ulonglong(dev.nbytes),
                         ctypes.c_void_p(st.handle))
            st.synchronize()
            assert np.array_equal(buf[:dev.nbytes].view(np.float32), host)
            first = buf.ctypes.data if _ == 0 else first
            assert buf.ctypes.data == first          # the pooled buffer is reused


def test_free_is_stream_ordered(lib):
    """Arrays freed while their stream still has queued work: the work
    completes, and new allocations on the stream see consistent memory."""
    print("hipmem stream-ordered free", flush=True)
    st = hipmem.Stream(lib)
    host = np.full((1 << 20,), 3.0, dtype=np.float32)
    keep = hipmem.DeviceArray.from_host(lib, host, st)
    for _ in range(8):
        tmp = hipmem.DeviceArray.from_host(lib, host * 2, st)
        tmp.zero_(st)
        del tmp                         # freed on st after the queued work
    st.synchronize()
    assert np.array_equal(keep.numpy(st), host)
