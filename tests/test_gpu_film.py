"""Film output from device memory: the EXR bytes a device film gives (half conversion on the GPU)
equal the host path's and the oracle's, for a rendered frame and for a film holding every special
value; a device film streams to a display server like the host film."""
import socket
import threading

import numpy as np
import pytest
import torch

from oracle import film_oracle as O
from sort_amd import Obvh, film, integrator as I, scenes

pytestmark = pytest.mark.gpu


def test_device_film_exr_equals_host_and_oracle(gpu_ctx):
    sc = scenes.cornell_sphere()
    acc = Obvh(gpu_ctx).build(sc["pos"], sc["idx"], sc["quads"])
    pt = I.PathTracing(acc, sc["materials"], tri_material=sc["tri_material"], normals=sc["normals"],
                       quad_radiance=sc["quad_radiance"], env_radiance=sc["env"], max_recursive_depth=5)
    W, H = 96, 64
    print("[film] render", flush=True)
    img = pt.render(I.camera(**sc["camera"]), W, H, 4, seed=3).reshape(H, W, 3).contiguous()
    host = img.cpu().numpy()
    dev_bytes = film.encode_exr(img)
    assert dev_bytes == film.encode_exr(host) == O.exr_file(host, W, H)
    g = np.random.default_rng(5)
    f = (g.normal(size=(40, 33, 3)) * 10.0 ** g.integers(-9, 6, (40, 33, 3))).astype(np.float32)
    f[0, 0] = [np.inf, -np.nan, 1e-40]
    f[1, 1] = [65520.0, 2.0 ** -15 + 2.0 ** -25, -(1.0 + 2.0 ** -11)]
    d = torch.from_numpy(f).cuda()
    print("[film] specials", flush=True)
    assert film.encode_exr(d) == film.encode_exr(f) == O.exr_file(f, 33, 40)


def test_device_film_display(gpu_ctx):
    g = np.random.default_rng(8)
    f = g.random((70, 130, 3)).astype(np.float32)
    srv = socket.socket(socket.AF_INET, socket.SOCK_STREAM)
    srv.bind(("127.0.0.1", 0))
    srv.listen(1)
    box = []

    def run():
        conn, _ = srv.accept()
        chunks = []
        while True:
            b = conn.recv(1 << 16)
            if not b:
                break
            chunks.append(b)
        box.append(b"".join(chunks))
        conn.close()

    t = threading.Thread(target=run, daemon=True)
    t.start()
    disp = film.Display("127.0.0.1", srv.getsockname()[1])
    disp.send_film(torch.from_numpy(f).cuda(), mode=film.BLENDER)
    disp.close()
    t.join(10)
    srv.close()
    assert np.array_equal(O.read_blender_packets(box[0], 130, 70)[:, :, :3], f)
