"""ServeSGBM.wait keeps a request's buffers alive until the server releases the ticket.

Host logic only: the C library is replaced by a stand-in whose lumen_server_wait first times
out and then completes, as include/lumen_sgbm.h specifies (LUMEN_ETIMEDOUT keeps the ticket).
"""
import threading

import numpy as np
import pytest

import lumenstereo_amd as ls


class _Lib:
    def __init__(self, codes):
        self.codes = list(codes)
        self.calls = 0

    def lumen_server_wait(self, h, ticket, tmo, info):
        self.calls += 1
        return self.codes.pop(0)

    def lumen_server_last_error(self, h):
        return b"request failed"


def _server(codes):
    srv = object.__new__(ls.ServeSGBM)
    srv._lib = _Lib(codes)
    srv._h = None
    srv._live = {}
    srv._waiting = set()
    srv._lock = threading.Lock()
    return srv


def _held():
    return (np.zeros((2, 3), np.uint8), np.zeros((2, 3), np.uint8),
            np.arange(6, dtype=np.int16).reshape(2, 3))


def test_timeout_keeps_buffers_then_wait_returns_output():
    srv = _server([ls.LUMEN_ETIMEDOUT, ls.LUMEN_OK])
    held = _held()
    srv._live[7] = held
    with pytest.raises(TimeoutError):
        srv.wait(7, timeout_s=0)
    assert srv._live[7] is held          # still referenced: the server may still write `out`
    out, info = srv.wait(7)
    assert out is held[2]
    assert 7 not in srv._live and not srv._waiting


def test_failed_request_releases_buffers():
    srv = _server([ls.LUMEN_EDEVICE])
    srv._live[3] = _held()
    with pytest.raises(ls.GpuException):
        srv.wait(3)
    assert 3 not in srv._live


def test_second_waiter_is_refused_and_buffers_stay():
    srv = _server([])
    srv._live[5] = _held()
    srv._waiting.add(5)                   # another thread is waiting on ticket 5
    with pytest.raises(ls.StereoException):
        srv.wait(5)
    assert 5 in srv._live and srv._lib.calls == 0
