"""sr_fft2_noalloc, the transform that allocates nothing (its twiddle table in a library-owned device array), called
through the C ABI: bitwise equal to sr_fft2 on the same input, forward and inverse, on two calls in a row with the
second one smaller (the shared table rewritten for the smaller length), and sizes past the table refused."""
import ctypes as C

import pytest

pytestmark = pytest.mark.gpu


def run(lib, fn, data, inverse, stream):
    ny, nx = data.shape
    return getattr(lib, fn)(C.c_void_p(data.data_ptr()), nx, ny, int(inverse), stream)


def test_fft2_without_allocating_equals_fft2(gpu_lib):
    import torch
    sr = gpu_lib
    lib = sr.lib()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(11)
    for ny, nx in [(512, 1024), (64, 32)]:  # the second call smaller than the first
        print(f"fft2 without allocating: {ny} x {nx}", flush=True)
        x = torch.complex(torch.randn((ny, nx), device="cuda", generator=gen),
                          torch.randn((ny, nx), device="cuda", generator=gen))
        for inverse in (False, True):
            a, b = x.clone(), x.clone()
            assert run(lib, "sr_fft2", a, inverse, stream) == 0
            assert run(lib, "sr_fft2_noalloc", b, inverse, stream) == 0
            torch.cuda.synchronize()
            assert torch.equal(torch.view_as_real(a), torch.view_as_real(b)), (ny, nx, inverse)
            assert not torch.equal(a, x)
        # the transform itself: forward then inverse returns the input
        c = x.clone()
        assert run(lib, "sr_fft2_noalloc", c, False, stream) == 0
        assert run(lib, "sr_fft2_noalloc", c, True, stream) == 0
        torch.cuda.synchronize()
        assert float((c - x).abs().max()) <= 1e-5 * float(x.abs().max())


def test_fft2_without_allocating_refuses_what_the_table_cannot_hold(gpu_lib):
    import torch
    sr = gpu_lib
    lib = sr.lib()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    x = torch.zeros((1, 16), dtype=torch.complex64, device="cuda")
    # the axis lengths are checked before anything is launched, so the small buffer is never touched
    assert lib.sr_fft2_noalloc(C.c_void_p(x.data_ptr()), 16384, 1, 0, stream) == sr._lib.SR_ERR_LENGTH
    assert b"8192" in lib.sr_last_error()
    assert lib.sr_fft2_noalloc(C.c_void_p(x.data_ptr()), 24, 1, 0, stream) == sr._lib.SR_ERR_NOT_POW2
    assert lib.sr_fft2_noalloc(None, 16, 1, 0, stream) == sr._lib.SR_ERR_ARGUMENT
