"""The FP64 pair GEMM of the unfused apply (fqehip_pair_gemm, Y = A B)
against torch's float64 matmul on the same device, through the C ABI.

Shapes cover the LDS-resident kernel (every row-tile count 1..9, k not a
multiple of 4, column counts that leave a partial 32-column panel, k steps
fewer and more than the B prefetch ring), the kernel that streams A
through LDS in chunks (m from 145 to 512: complex operators in their real
block form, spin-orbital operators; every row-tile count 1..4 per wave)
and the shapes both decline (odd column counts, m above 512, B not
16-byte aligned), which run on the generic kernel.  Tolerance: normwise
relative 1e-14 (both sides are FP64 sums of at most a few hundred
products)."""

import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-14

SHAPES = [  # (m, k, ncol)
    (1, 1, 2), (11, 10, 34), (16, 16, 64), (22, 21, 1000),
    (37, 36, 4098), (45, 44, 130), (56, 55, 2), (67, 66, 2050),
    (79, 78, 66), (92, 91, 1026), (106, 105, 100002), (121, 120, 4000),
    (137, 136, 3010), (144, 136, 98),
    # A streamed in chunks
    (16, 1300, 66), (144, 141, 98), (150, 149, 512), (200, 198, 2050), (260, 259, 98), (394, 392, 300),
    (394, 392, 20002), (452, 450, 66), (512, 511, 130),
    # declined (generic kernel)
    (106, 105, 1001), (394, 392, 301), (513, 512, 34), (700, 650, 64),
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


# The complex64 engine's pair GEMM: bfloat16 hi/lo products (three of the
# four) accumulated in FP32 on the matrix cores.  Against the float64
# product of the same float32 B and float64 A: normwise relative 2e-5
# (the dropped lo x lo term and the 16-bit splits leave about 2^-16 per
# product; FP32 accumulation over k <= 160 terms).
BX_TOL = 2e-5

BX_SHAPES = [  # (m, k, ncol): A in LDS for m <= 144, k <= 160
    (1, 1, 1), (11, 10, 33), (22, 21, 1000), (45, 44, 131), (79, 78, 2050),
    (106, 105, 100001), (121, 120, 4000), (137, 136, 3010), (144, 160, 98),
    # generic kernel
    (150, 149, 512), (106, 170, 300),
]


@pytest.mark.parametrize("m,k,ncol", BX_SHAPES)
def test_pair_gemm_bf16x3_matches_matmul(lib, m, k, ncol):
    import torch
    gen = torch.Generator(device="cuda").manual_seed(m * 7 + k)
    a = torch.randn(m, k, dtype=torch.float64, device="cuda", generator=gen)
    b = torch.randn(k, ncol, dtype=torch.float32, device="cuda", generator=gen)
    y = torch.full((m, ncol), float("nan"), dtype=torch.float32, device="cuda")
    lib.call("fqehip_pair_gemm", 1, m, k, lib.ptr(a), ncol, lib.ptr(b),
             lib.ptr(y), lib.stream_ptr())
    ref = a @ b.double()
    torch.cuda.synchronize()
    err = (torch.linalg.norm(y.double() - ref) / torch.linalg.norm(ref)).item()
    print(f"bf16x3 m={m} k={k} ncol={ncol} rel={err:.2e}")
    assert err < BX_TOL
