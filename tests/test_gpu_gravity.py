"""P2 on the GPU against the oracle: the FFT, the convolution and the field (crates/sr-engine/tests/gpu/fft.rs:103-266
and tests/gpu/grav.rs: rel L2 <= 1e-5 against the twin, Maclaurin disk within 2 %, net self-force)."""
import math

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("shape", [(64, 64), (128, 256), (1024, 2048), (32, 4096), (4096, 32)])
def test_fft2_matches_the_oracle(gpu_lib, shape):
    import torch
    ny, nx = shape
    print(f"fft2 {shape}", flush=True)
    rng = np.random.default_rng(nx + 7 * ny)
    x = (rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape))
    t = torch.tensor(x.astype(np.complex64), device="cuda")
    gpu_lib.fft2(t)
    got = t.cpu().numpy().astype(np.complex128)
    want = np.fft.fft2(x) if max(shape) > 256 else oracle.fft2(x)
    assert rel_l2(got, want) < 1e-5
    gpu_lib.fft2(t, inverse=True)
    assert rel_l2(t.cpu().numpy().astype(np.complex128), x) < 1e-5


@pytest.mark.parametrize("lengths", [32, 64, 128, 256, 512, 1024, 2048, 4096])
def test_every_length_transforms(gpu_lib, lengths):
    import torch
    x = np.random.default_rng(lengths).uniform(-1, 1, (4, lengths)).astype(np.complex64)
    t = torch.tensor(x, device="cuda")
    gpu_lib.fft2(t)
    assert rel_l2(t.cpu().numpy(), np.fft.fft2(x.astype(np.complex128))) < 1e-5


@pytest.mark.parametrize("w,h,dx,G", [(64, 64, 1.0, 1.0), (48, 40, 1.0, 1.0), (200, 96, 0.25, 2.0),
                                      (256, 256, 1.0 / 256, 1.0), (1024, 1024, 1.0 / 1024, 1.0)])
def test_gravity_field_matches_the_oracle(gpu_lib, w, h, dx, G):
    import torch
    print(f"gravity {w}x{h}", flush=True)
    rng = np.random.default_rng(w * 31 + h)
    sigma = rng.uniform(0.0, 1.0, (h, w))
    sigma[rng.uniform(size=(h, w)) < 0.1] = 1e-7  # vacuum cells: no source
    ph = oracle.reference_physics(G=G, dx=dx, sigma_vac=1e-6, max_padded=4096)
    phi, gx, gy, pr = oracle.gravity_field(ph, sigma.astype(np.float32).astype(np.float64), ringed=True)
    g = gpu_lib.Gravity(w, h, G=G, dx=dx, sigma_vac=1e-6)
    s = torch.tensor(sigma.astype(np.float32), device="cuda")
    gphi, ggx, ggy = g.field(s)
    torch.cuda.synchronize()
    # phi: the reference's FFT bar (rel L2 1e-5, tests/gpu/fft.rs); g is a central difference of phi, so f32
    # rounding of phi is amplified by the cancellation: its bar is 1e-4 (the reference's phi-vs-direct-sum bar)
    assert rel_l2(gphi.cpu().numpy(), pr) < 1e-5
    assert rel_l2(ggx.cpu().numpy(), gx) < 1e-4
    assert rel_l2(ggy.cpu().numpy(), gy) < 1e-4


def test_maclaurin_disk_force_on_the_gpu(gpu_lib):
    import torch
    a, side = 40.0, 128
    o = (side - 1) / 2.0
    j, i = np.meshgrid(np.arange(side), np.arange(side), indexing="ij")
    r2 = ((i - o) ** 2 + (j - o) ** 2) / (a * a)
    sigma = np.where(r2 < 1.0, np.sqrt(np.clip(1.0 - r2, 0, None)), 0.0).astype(np.float32)
    g = gpu_lib.Gravity(side, side, sigma_vac=0.0)
    _, gx, gy = g.field(torch.tensor(sigma, device="cuda"))
    gx, gy = gx.cpu().numpy().astype(np.float64), gy.cpu().numpy().astype(np.float64)
    k = math.pi ** 2 / (2 * a)
    dx, dy = i - o, j - o
    r = np.hypot(dx, dy)
    m = (r > 0) & (r <= 0.75 * a)
    err = np.sqrt((gx + k * dx) ** 2 + (gy + k * dy) ** 2)[m] / (k * r[m])
    assert err.max() <= 0.02


def test_net_self_force_is_small(gpu_lib):
    import torch
    rng = np.random.default_rng(11)
    sigma = rng.uniform(0, 1, (96, 80)).astype(np.float32)
    g = gpu_lib.Gravity(80, 96, sigma_vac=1e-6)
    _, gx, gy = g.field(torch.tensor(sigma, device="cuda"))
    gx, gy = gx.cpu().numpy().astype(np.float64), gy.cpu().numpy().astype(np.float64)
    gross = np.sum(sigma * np.hypot(gx, gy))
    assert abs(np.sum(sigma * gx)) <= 1e-5 * gross and abs(np.sum(sigma * gy)) <= 1e-5 * gross


def test_refusals(gpu_lib):
    with pytest.raises(gpu_lib.SrError) as e:
        gpu_lib.Gravity(3000, 64, max_padded=4096)
    assert e.value.kind == "BoxAxis"
