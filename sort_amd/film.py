"""Film output over the C ABI: SORT's EXR file (``encode_exr`` / ``write_exr``, byte-identical to
tinyexr's SaveEXR as texture/texturebase.cpp:27-57 calls it) and the display-server tile protocols
(``display_packet``, ``Display``; core/display_mgr.cpp:86-140). Films are (height, width, 3)
float32 arrays in host memory or torch tensors on the GPU."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib

BLENDER, TEV, TEV_CREATE = 0, 1, 2


def _film(film, width=None, height=None):
    """(pointer, on_device, width, height, keepalive) of a film."""
    try:
        import torch
        if isinstance(film, torch.Tensor):
            assert film.is_cuda and film.dtype == torch.float32 and film.is_contiguous()
            h, w = (film.shape[0], film.shape[1]) if film.dim() == 3 else (height, width)
            return film.data_ptr(), 1, int(w), int(h), film
    except ImportError:
        pass
    a = np.ascontiguousarray(np.asarray(film, np.float32))
    h, w = (a.shape[0], a.shape[1]) if a.ndim == 3 else (height, width)
    return a.ctypes.data, 0, int(w), int(h), a


def _stream(on_device, stream):
    if not on_device or stream is not None:
        return stream
    import torch
    return torch.cuda.current_stream().cuda_stream


def encode_exr(film, width=None, height=None, stream=None) -> bytes:
    ptr, dev, w, h, keep = _film(film, width, height)
    n = lib().sortgpu_exr_size(w, h)
    buf = (ctypes.c_uint8 * n)()
    check(lib().sortgpu_film_encode_exr(ptr, dev, w, h, buf, n, _stream(dev, stream)))
    return bytes(buf)


def write_exr(path, film, width=None, height=None, stream=None) -> None:
    ptr, dev, w, h, keep = _film(film, width, height)
    check(lib().sortgpu_film_write_exr(ptr, dev, w, h, str(path).encode(), _stream(dev, stream)))


def display_packet(mode, film, x=0, y=0, tw=0, th=0, title="sort") -> bytes:
    a = np.ascontiguousarray(np.asarray(film, np.float32))
    h, w = a.shape[0], a.shape[1]
    n = lib().sortgpu_display_packet(mode, title.encode(), a.ctypes.data, w, h, x, y, tw, th, None, 0)
    check(int(n))
    buf = (ctypes.c_uint8 * n)()
    check(int(lib().sortgpu_display_packet(mode, title.encode(), a.ctypes.data, w, h, x, y, tw, th, buf, n)))
    return bytes(buf)


class Display:
    """A connection to a display server (Blender's add-on or tev)."""

    def __init__(self, host: str, port: int):
        self.handle = _lib._P()
        check(lib().sortgpu_display_connect(host.encode(), port, ctypes.byref(self.handle)))

    def send_film(self, film, mode=BLENDER, title="sort", tile=64, stream=None):
        ptr, dev, w, h, keep = _film(film)
        check(lib().sortgpu_display_send_film(self.handle, mode, title.encode(), ptr, dev, w, h, tile,
                                              _stream(dev, stream)))

    def close(self):
        if self.handle:
            lib().sortgpu_display_close(self.handle)
            self.handle = None

    def __del__(self):
        self.close()
