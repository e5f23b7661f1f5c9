"""Engine settings.

``default_dtype``: coefficient dtype of new sectors.  complex128 runs the
FP64 engine; complex64 runs the mixed-precision engine (complex FP32
storage, BF16 hi/lo split matrix-core GEMM with FP32 accumulation, FP64
norms and Taylor accumulation).

``contraction``: arithmetic of the complex128 split apply's alpha-row
contraction.  "fp64" multiplies on the FP64 matrix cores; "ozaki" forms
the same products exactly from 8-bit digits on the int8 matrix cores
(complex128 storage kept), about 1e-11 from the FP64 result.
"""

import numpy

default_dtype = numpy.complex128

# Maximum number of orbitals of the GPU string tables (uint64 strings,
# int32-addressed tables).
max_norb = 32

# Use the fused gather + GEMM + scatter kernels for real operators when the
# sector fits them (else, and for complex operators, the unfused kernels).
fused = True

# complex64 applies of a pair-symmetric real operator: beta-beta block as
# one BF16x3 GEMM and an alpha-only row kernel (no beta scatter).  Off, or
# for other operators, the row kernel also scatters beta excitations.
split = True

# complex128 applies of a pair-symmetric real operator: beta-beta block as
# one FP64 matrix-core GEMM and an FP64 row kernel over the live alpha pair
# rows.  Off, the FP64 row kernel also scatters beta excitations.
split_fp64 = True

# Kernel of the split row pass: 0 = MFMAs over the live pair rows of each
# alpha string only (excitations to another string plus the folded
# diagonal row); 16 = MFMAs over all npair + 1 pair rows.
split_tile = 0

# Kernel of the complex128 split row pass: 0 = column-owned waves (the D'
# values of a wave's columns gathered into its MFMA registers, A' in LDS);
# 32 = the kernel with a shared, double-buffered D tile.
split_tile_fp64 = 0

# Arithmetic of the complex128 split row pass: "fp64" (FP64 matrix cores)
# or "ozaki" (exact int8 digit products, FP64 assembly).
contraction = "fp64"

# With contraction = "ozaki": the beta-beta GEMM as digit products too
# (else FP64 matrix cores for that half).
ozaki_beta = True

# Largest apply workspace (bytes) allocated at once: a complex128 sector
# whose unfused dvec + Y would need more runs the apply in panels of
# orbital pairs within this budget (fqehip_apply_panels).
max_workspace_bytes = 96 << 30

# Single-sector time evolution: record the whole Taylor series once into a
# HIP graph (fqehip_taylor_graph_create) and replay it with one launch per
# time_evolve; the graph and its buffers are cached on the operator for the
# last (sector, time, accuracy, expansion) it ran.
taylor_graph = False
