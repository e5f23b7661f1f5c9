"""The FP64 pair GEMM of the unfused apply (fqehip_pair_gemm, Y = A B)
against torch's float64 matmul on the same device, through the C ABI.

Shapes cover the LDS-resident kernel (every row-tile count 1..9, k not a
multiple of 4, column counts that leave a partial 32-column panel, k steps
fewer and more than the B prefetch ring) and the shapes it declines (odd
column counts, m above 144, B not 16-byte aligned), which run on the
generic kernel.  Tolerance: normwise relative 1e-14 (both sides are FP64
sums of at most a few hundred products)."""

import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-14

SHAPES = [  # (m, k, ncol)
    (1, 1, 2), (11, 10, 34), (16, 16, 64), (22, 21, 1000),
    (37, 36, 4098), (45, 44, 130), (56, 55, 2), (67, 66, 2050),
    (79, 78, 66), (92, 91, 1026), (106, 105, 100002), (121, 120, 4000),
    (137, 136, 3010), (144, 141, 98),
    # declined by the LDS kernel
    (106, 105, 1001), (150, 149, 512), (394, 392, 300),
]


@pytest.fixture(scope="module")
def lib():
    import torch
    assert torch.cuda.is_available()
    from fqe_amd import _lib
    return _lib


def _run(lib, a, b, y):
    m, k = a.shape
    lib.call("fqehip_pair_gemm", 0, m, k, lib.ptr(a), b.shape[1], lib.ptr(b),
             lib.ptr(y), lib.stream_ptr())


@pytest.mark.parametrize("m,k,ncol", SHAPES)
def test_pair_gemm_matches_matmul(lib, m, k, ncol):
    import torch
    gen = torch.Generator(device="cuda").manual_seed(m * 1000 + k)
    a = torch.randn(m, k, dtype=torch.float64, device="cuda", generator=gen)
    b = torch.randn(k, ncol, dtype=torch.float64, device="cuda", generator=gen)
    y = torch.full((m, ncol), float("nan"), dtype=torch.float64, device="cuda")
    _run(lib, a, b, y)
    ref = a @ b
    torch.cuda.synchronize()
    err = (torch.linalg.norm(y - ref) / torch.linalg.norm(ref)).item()
    print(f"m={m} k={k} ncol={ncol} rel={err:.2e}")
    assert err < TOL


def test_pair_gemm_misaligned_b_and_repeat(lib):
    """B one double past a 16-byte boundary (declined, generic kernel), then
    the aligned call twice in a row into the same output: every call
    overwrites Y."""
    import torch
    gen = torch.Generator(device="cuda").manual_seed(7)
    m, k, ncol = 106, 105, 2050
    a = torch.randn(m, k, dtype=torch.float64, device="cuda", generator=gen)
    buf = torch.randn(k * ncol + 1, dtype=torch.float64, device="cuda",
                      generator=gen)
    b_mis = buf[1:].view(k, ncol)
    b_al = buf[:-1].view(k, ncol)
    for b in (b_mis, b_al, b_al):
        y = torch.full((m, ncol), float("nan"), dtype=torch.float64,
                       device="cuda")
        _run(lib, a, b, y)
        ref = a @ b
        torch.cuda.synchronize()
        err = (torch.linalg.norm(y - ref) / torch.linalg.norm(ref)).item()
        assert err < TOL
