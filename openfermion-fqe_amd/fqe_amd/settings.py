"""Engine settings.

``default_dtype``: coefficient dtype of new sectors.  complex128 runs the
FP64 engine; complex64 runs the mixed-precision engine (complex FP32
storage, BF16 hi/lo split matrix-core GEMM with FP32 accumulation, FP64
norms and Taylor accumulation).
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
