"""The streaming kernels that set the benchmark's roofline ceilings
(libtc_stream.so) do the work their rates are quoted for: the copy moves every
byte, the read touches every element, and bad arguments are refused on the
host before any launch."""
import ctypes as C
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _lib():
    lib = C.CDLL(os.path.join(ROOT, "taichi_amd", "libtc_stream.so"))
    lib.tcs_copy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    lib.tcs_read.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    return lib


@pytest.mark.gpu
@pytest.mark.parametrize("nbytes,grid", [(24 << 20, 512), (16 * 1000003, 1024), (16, 7)])
def test_copy_and_read_cover_every_element(dev, nbytes, grid):
    import torch
    lib = _lib()
    s = torch.cuda.current_stream()
    h = C.c_void_p(s.cuda_stream)
    g = torch.Generator(device=dev).manual_seed(nbytes)
    src = torch.randint(-8, 8, (nbytes // 4,), device=dev, generator=g).float()
    dst = torch.full_like(src, float("nan"))
    assert lib.tcs_copy(src.data_ptr(), dst.data_ptr(), nbytes, grid, h) == 0
    part = torch.zeros(grid * 256, device=dev)
    assert lib.tcs_read(src.data_ptr(), nbytes, part.data_ptr(), part.numel(), grid, h) == 0
    torch.cuda.synchronize()
    assert torch.equal(dst, src)
    # small integers: every partial sum and the total are exact in FP32/FP64
    assert float(part.double().sum()) == float(src.double().sum())


@pytest.mark.gpu
def test_bad_arguments_are_refused(dev):
    import torch
    lib = _lib()
    h = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    a = torch.zeros(64, device=dev)
    b = torch.zeros(64, device=dev)
    assert lib.tcs_copy(a.data_ptr(), b.data_ptr(), 20, 1, h) == -1          # not a multiple of 16
    assert lib.tcs_copy(a.data_ptr() + 4, b.data_ptr(), 16, 1, h) == -1      # misaligned
    assert lib.tcs_copy(a.data_ptr(), b.data_ptr(), 16, 0, h) == -1          # no workgroups
    assert lib.tcs_read(a.data_ptr(), 64, b.data_ptr(), 255, 1, h) == -1     # partials too short
    torch.cuda.synchronize()
