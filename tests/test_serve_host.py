"""Serving-loop argument checks that need no GPU."""
import ctypes as C

from tetragpu_amd import serve


def test_bad_config_rejected():
    lib = serve._lib()
    h = C.c_void_p()
    for cfg in (serve.Config(0, 1 << 30, 0, 1), serve.Config(2, 0, 0, 1)):
        assert lib.tg_server_create(C.byref(h), C.byref(cfg)) == -1


def test_no_device_reported():
    import tetragpu_amd as t
    if t.device_count() > 0:
        return
    lib = serve._lib()
    h = C.c_void_p()
    assert lib.tg_server_create(C.byref(h), C.byref(serve.Config(1, 1 << 30, 0, 1))) == -7
