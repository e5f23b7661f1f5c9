/*
 * fqehip — MI355X (gfx950) engine for the determinant-space action of a
 * spin-restricted 1+2-body Hamiltonian on a CI vector, i.e. the hot path of
 * OpenFermion-FQE's Wavefunction.apply(RestrictedHamiltonian) and
 * Wavefunction.time_evolve.
 *
 * C ABI.  Conventions for every entry point:
 *   - extern "C"; plain pointers and int64 sizes, no torch types;
 *   - d_* arguments are device pointers owned by the caller;
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream);
 *     work is enqueued, nothing synchronises unless stated;
 *   - return 0 (FQEHIP_OK) or an FQEHIP_ERR_* code; fqehip_last_error()
 *     gives the message of the calling thread's last failure.  Nothing
 *     exits the process (the reference's safe_malloc does,
 *     src/fqe/lib/macros.c:64-75).
 *
 * Data layouts (device memory):
 *   strings   uint64[L], address order (lexicographic rank of the ascending
 *             occupation tuple), as FciGraph._astr;
 *   dexc      int32[L][nexc][3] = (source, i*norb+j, sign) rows per target,
 *             ascending i*norb+j, nexc = nele*(norb-nele+1), as
 *             FciGraph._dexca;
 *   maps      int32[sum_ij m_ij][3], map (i,j) at offsets[i*norb+j],
 *             ascending source, as FciGraph._alpha_map[(i,j)];
 *   pair table int32[npair][L]: entry for (pair p, target I) is
 *             sign*(source+1) where <I|E_p|source> = sign, or 0.  Ordered
 *             basis: npair = norb^2, p = i*norb+j, E_p = E_ij.  Compressed
 *             basis: npair = norb(norb+1)/2, p = j + i(i+1)/2 (i >= j),
 *             E_p = E_ij + E_ji (i != j) or E_ii;
 *   coefficients  complex, interleaved (re, im), row-major [lena][lenb]:
 *             complex128 for FQEHIP_C128, complex64 for FQEHIP_C64;
 *   dvec / Y  planar real, [plane][re|im][lena*lenb], double for
 *             FQEHIP_C128, float for FQEHIP_C64.
 */
#ifndef FQEHIP_H_
#define FQEHIP_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FQEHIP_OK 0
#define FQEHIP_ERR_INVALID 1     /* bad size, shape or pointer */
#define FQEHIP_ERR_OOM 2         /* a device allocation failed */
#define FQEHIP_ERR_UNSUPPORTED 3 /* e.g. norb above the supported maximum */
#define FQEHIP_ERR_HIP 4         /* a HIP runtime or launch error */

/* Storage / arithmetic modes. */
#define FQEHIP_C128 0 /* complex FP64 storage, FP64 MFMA GEMM */
#define FQEHIP_C64 1  /* complex FP32 storage, BF16 hi/lo split MFMA GEMM
                         with FP32 accumulation; FP64 norms / Taylor sum */

const char* fqehip_last_error(void);
int fqehip_version(void);

/* ---------------------------------------------------------------------
 * (a) String tables, built on the GPU.
 * Replaces FciGraph._build_strings (src/fqe/fci_graph.py:301-333 ->
 * src/fqe/lib/bitstring.c:29-46, src/fqe/lib/fci_graph.c:27-54,136-147)
 * and map_to_deexc (src/fqe/fci_graph.py:36-63 -> fci_graph.c:57-73).
 * d_strings: uint64[nstrings]; d_dexc: int32[nstrings][nele*(norb-nele+1)][3].
 * ------------------------------------------------------------------- */
int fqehip_string_tables(int norb, int nele, int64_t nstrings,
                         uint64_t* d_strings, int32_t* d_dexc, void* stream);

/* Excitation maps a+_i a_j for all (i,j).  Replaces
 * FciGraph._build_mapping (src/fqe/fci_graph.py:204-244 ->
 * src/fqe/lib/fci_graph.py:60-99 -> fci_graph.c:76-121).
 * d_offsets: int64[norb*norb+1] start row of each map (host-computed
 * prefix sums of C(norb-2,nele-1) / C(norb-1,nele-1)); d_maps: int32[*][3]. */
int fqehip_string_maps(int norb, int nele, int64_t nstrings,
                       const uint64_t* d_strings, const int64_t* d_offsets,
                       int32_t* d_maps, void* stream);

/* Dense pair table (layout above) from dexc; compressed = 0 or 1. */
int fqehip_pair_table(int norb, int nele, int64_t nstrings,
                      const int32_t* d_dexc, int compressed, int32_t* d_ptab,
                      void* stream);

/* ---------------------------------------------------------------------
 * (b) dvec: D[p] = E_p C for all pairs p (pair basis of the tables).
 * Replaces FqeData.calculate_dvec_spatial (src/fqe/fqe_data.py:2149-2160,
 * :2209-2234 -> zdvec_make, src/fqe/lib/fqe_data.c:350-377) and, for the
 * compressed basis, _calculate_dvec_spatial_compressed (:2336-2353).
 * d_dvec: planar [npair][2][lena*lenb].
 * ------------------------------------------------------------------- */
int fqehip_dvec_gather(int precision, int npair, int64_t lena, int64_t lenb,
                       const int32_t* d_ptab_a, const int32_t* d_ptab_b,
                       const void* d_coeff, void* d_dvec, void* stream);

/* ---------------------------------------------------------------------
 * (c) Y = A B on the matrix cores: A real row-major m x k (FP64),
 * B row-major k x ncol and Y row-major m x ncol in the storage precision.
 * Replaces the dvec contraction einsum('ijkl,klmn->ijmn') of
 * src/fqe/fqe_data.py:656 (and :671-672 for the compressed basis); the
 * one-body term is one more row of A.
 * ------------------------------------------------------------------- */
int fqehip_pair_gemm(int precision, int m, int k, const double* d_a,
                     int64_t ncol, const void* d_b, void* d_y, void* stream);

/* ---------------------------------------------------------------------
 * (d) Scatter + epilogue:
 *   v[I] = alpha * (Y[npair][I] + sum_p (E_p Y[p])[I]) + beta * cin[I]
 *   out[I] = v[I];  acc[I] += v[I] (complex128, optional);
 *   norm_partials[block] = partial sums of |v|^2 in FP64 (optional).
 * Y has npair+1 planes (the last holds the one-body term).  Replaces
 * _calculate_coeff_spatial_with_dvec (src/fqe/fqe_data.py:2309-2334 ->
 * zcoeff_make, src/fqe/lib/fqe_data.c:379-406) plus the ax_plus_y of
 * src/fqe/wavefunction.py:396-397 and :564.  If d_done is not NULL and
 * *d_done != 0 the kernel does nothing (device-side early exit of the
 * Taylor loop).  d_cin may alias d_out only when beta == 0.
 * ------------------------------------------------------------------- */
int fqehip_scatter(int precision, int npair, int64_t lena, int64_t lenb,
                   const int32_t* d_ptab_a, const int32_t* d_ptab_b,
                   const void* d_y, double alpha_re, double alpha_im,
                   double beta_re, double beta_im, const void* d_cin,
                   void* d_out, double* d_acc, double* d_norm_partials,
                   const int32_t* d_done, void* stream);

/* Number of partial sums fqehip_scatter / fqehip_apply1 write for a
 * sector of n = lena*lenb determinants. */
int64_t fqehip_norm_partials_count(int64_t n);

/* One-body apply out = alpha * sum_ij h1[i,j] E_ij C + beta * cin.
 * Replaces _apply_array_spatial1 (src/fqe/fqe_data.py:477-531 ->
 * lm_apply_array1, src/fqe/lib/fqe_data.c:272-302).  d_h1: complex128
 * [npair] in the pair basis of the tables. */
int fqehip_apply1(int precision, int npair, int64_t lena, int64_t lenb,
                  const int32_t* d_ptab_a, const int32_t* d_ptab_b,
                  const double* d_h1, const void* d_coeff, double alpha_re,
                  double alpha_im, double beta_re, double beta_im,
                  const void* d_cin, void* d_out, double* d_acc,
                  double* d_norm_partials, const int32_t* d_done,
                  void* stream);

/* Reduce norm partials: norm = sqrt(sum partials) (FP64, fixed order).
 * d_norm_out[0] = norm.  If d_done is not NULL (Taylor mode) and
 * *d_done == 0: when norm < accuracy, sets *d_done = 1 and
 * d_nterms[0] = term; otherwise leaves them. */
int fqehip_norm_finalize(const double* d_partials, int64_t npartials,
                         double* d_norm_out, double accuracy, int term,
                         int32_t* d_done, int32_t* d_nterms, void* stream);

/* Elementwise helpers for the Taylor loop and conversions. */
/* out(complex, storage precision) = in(complex128) */
int fqehip_convert_from_c128(int precision, int64_t n, const double* d_in,
                             void* d_out, void* stream);
/* out(complex128) = scale * in(complex, storage precision) */
int fqehip_convert_to_c128(int precision, int64_t n, const void* d_in,
                           double scale_re, double scale_im, double* d_out,
                           void* stream);

/* ---------------------------------------------------------------------
 * Composed apply and Taylor propagation on one sector.
 * ------------------------------------------------------------------- */
typedef struct fqehip_sector {
  int norb, nalpha, nbeta;
  int npair;               /* pair-basis size of the tables below */
  int64_t lena, lenb;
  const int32_t* ptab_a;   /* [npair][lena] */
  const int32_t* ptab_b;   /* [npair][lenb] */
  /* Tables of the fused path (fqehip_fused_tables); NULL selects the
   * unfused gather / GEMM / scatter kernels. */
  const int32_t* fwd_a;    /* [npair][lena] forward table, alpha */
  const int32_t* fwd_b;    /* [npair][lenb] forward table, beta */
  const int32_t* slot_a;   /* [npair][lena] alpha slot index */
  int nslot;               /* nalpha * (norb - nalpha) */
  /* Beta-beta block of the operator on this sector (fqehip_beta_block),
   * bfloat16 bit patterns [ldh][ldh]; set only for an operator with
   * A[p][q] = A[q][p] (p, q < npair).  When set, complex64 applies take
   * the split path; NULL keeps the beta excitations in the row kernel. */
  const uint16_t* hbb_hi;
  const uint16_t* hbb_lo;
  int64_t ldh;
  /* Tile shape of the complex64 row kernel: with beta scatter (hbb_hi
   * NULL) FQEHIP_TILE_ROWS4 (default, 0 selects it) or FQEHIP_TILE_ROWS8
   * (4 or 8 waves per alpha row); on the split path 0 selects the kernel
   * over the live pair rows of each alpha string (excitations to another
   * string, plus the folded diagonal row), FQEHIP_TILE_SPLIT_ALLROWS the
   * kernel over all npair + 1 rows; on the complex128 split path 0 selects
   * the column-owned FP64 kernel, FQEHIP_TILE_SPLIT_F64_DTILE the one with
   * a shared D tile. */
  int tile;
  /* FP64 beta-beta block of the same operator, column-major [lenb][lenb]
   * (entry [Ib][Ib'] = Hbb[Ib'][Ib], the hcol output of fqehip_beta_block).
   * When set, complex128 applies take the split path (FP64 beta-beta GEMM,
   * FP64 row kernel over the live pair rows); NULL keeps the FP64 row
   * kernel that also scatters beta excitations. */
  const double* hbb_f64;
  /* Digits of the same FP64 beta-beta block for the Ozaki-digit split apply
   * (tile FQEHIP_TILE_OZAKI; fqehip_beta_block_digits): when set, the
   * beta-beta GEMM also runs as digit products; NULL keeps it in FP64. */
  const void* hbb_oz;
} fqehip_sector;

#define FQEHIP_TILE_ROWS4 4
#define FQEHIP_TILE_ROWS8 8
/* split path (hbb set): MFMAs over all npair + 1 pair rows instead of the
 * live rows of each alpha string (the default, tile 0) */
#define FQEHIP_TILE_SPLIT_ALLROWS 16
/* complex128 split path (hbb_f64 set): the row kernel with a shared,
 * double-buffered D tile instead of column-owned waves (the default).
 * Gather depth npair + 1: up to 128 for the column-owned and the digit
 * kernels (norb <= 15 compressed), up to 112 with this tile. */
#define FQEHIP_TILE_SPLIT_F64_DTILE 32
/* complex128 split path (hbb_f64 set): the alpha-row contraction as exact
 * products of 8-bit digits on the int8 matrix cores (Ozaki scheme: A' in
 * five balanced digits per pair row, D' as 48-bit integers per alpha
 * string, 20 digit pairs, int32 accumulation, FP64 assembly); complex128
 * storage and row accumulators, the slot rows in 48 bits per double (12
 * bytes per complex value, relative rounding at most 2^-37) within the
 * FP64 path's slot region, which fqehip_apply_slot_reduce reads back for
 * the same sector.  The workspace grows by the beta gather table and the
 * pair rows' digits, lena doubles (row maxima of |C|) and, with hbb_oz
 * set, by the C digit planes of the beta-beta GEMM (6 * 2 lena * ldk bytes
 * + lena doubles). */
#define FQEHIP_TILE_OZAKI 64

typedef struct fqehip_operator {
  int complex_op;   /* 0: A is (npair+1) x npair; 1: 2(npair+1) x 2npair */
  const double* a;  /* device, row-major */
  const double* h1; /* one-body only apply: complex128 [npair], else NULL */
} fqehip_operator;

/* Tables of the fused apply, from a pair table (layout above):
 *   d_fwd[p][src] = sign*(tgt+1) with <tgt|E_p|src> = sign, else 0;
 *   d_slot[p][src] (optional) = index k of (src, p) among the off-diagonal
 *   de-excitations of tgt, in dexc order, else -1.
 * There is no counterpart table in the reference: its maps
 * (src/fqe/fci_graph.py:204-244) list the same (src, tgt, sign) triples
 * per orbital pair. */
int fqehip_fused_tables(int norb, int nele, int64_t nstrings,
                        const int32_t* d_dexc, int compressed,
                        const int32_t* d_ptab, int32_t* d_fwd,
                        int32_t* d_slot, void* stream);

/* The two kernels of the fused apply, for callers that time or schedule
 * them separately (fqehip_apply runs both).  Preconditions: real operator,
 * fused tables set in the sector, fqehip_fused_supported() true.
 *   fused_rows: slot rows and row accumulators into d_workspace;
 *   slot_reduce: out = alpha * (H C) + beta * cin with the epilogue of
 *   fqehip_scatter.
 * The slot rows' format follows the sector (FQEHIP_TILE_OZAKI with
 * hbb_f64: 48-bit), so both calls take the same sector. */
int fqehip_fused_supported(int precision, const fqehip_sector* s,
                           const fqehip_operator* op);

/* Beta-beta block of a pair-symmetric real operator (A as in
 * fqehip_operator, A[p][q] = A[q][p] for p, q < npair):
 *   Hbb = sum_pq A[p][q] E^b_p E^b_q + sum_q A[npair][q] E^b_q
 * on the lenb beta strings, from the beta forward table d_fwd_b.  Writes
 * d_hcol[Ib][Ib'] = Hbb[Ib'][Ib] (FP64, lenb x lenb: the hbb_f64 input of
 * the complex128 split apply, else scratch) and the bfloat16 hi + lo split
 * of Hbb (the complex64 split apply's input), row-major
 * [ldh][ldh] zero-padded, ldh = fqehip_beta_block_ld(lenb).  It is the
 * same-spin part of the operator that the reference's lm kernels apply
 * per spin (src/fqe/lib/fqe_data.c:668-720); here it is a matrix so that
 * the beta half of every apply is one GEMM.  (lenb + npair) * 8 bytes
 * <= 160 KiB; the same bits on every run. */
int64_t fqehip_beta_block_ld(int64_t lenb);
int fqehip_beta_block(int npair, int64_t lenb, const int32_t* d_fwd_b,
                      const double* d_a, double* d_hcol, uint16_t* d_hhi,
                      uint16_t* d_hlo, int64_t ldh, void* stream);
/* Ozaki digits of an FP64 beta-beta block (d_hcol, the hcol output of
 * fqehip_beta_block) for the digit apply's beta-beta GEMM, written to d_hoz
 * (fqehip_beta_block_digits_bytes(lenb) bytes): int8 [5][lenb][ldk]
 * (ldk = lenb rounded up to 64; per column Ib' of the GEMM, i.e. row Ib'
 * of Hbb, scaled by 2^eh with |Hbb 2^eh| < 2^38 and rounded to integers z,
 * the five exact balanced digits of z, weights 2^32, 2^24, ..., 1), then
 * double [lenb] scales 2^-eh, then double [lenb] offset corrections of
 * the C digits (128 (1 + 2^8 + 2^16 + 2^24 + 2^32) sum of the column's z).
 * Once per operator and beta string count; lenb < 26214. */
int64_t fqehip_beta_block_digits_bytes(int64_t lenb);
int fqehip_beta_block_digits(int64_t lenb, const double* d_hcol, void* d_hoz,
                             void* stream);
/* d_out[Ia][Ib'] = sum_Ib C[Ia][Ib] Hbb[Ib'][Ib] for complex128 C as digit
 * products (the beta-beta half of the Ozaki-digit split apply, for callers
 * that time or schedule it alone): row maxima of |C| into d_work (lena
 * doubles), the C digit planes after them (6 * 2 lena * ldk bytes + lena
 * doubles), then the int8 GEMM.  Replaces, like fqehip_beta_gemm_f64, the
 * same-spin beta part of lm_apply_array12_same_spin_opt
 * (src/fqe/lib/fqe_data.c:668-720). */
int fqehip_beta_gemm_digits(int64_t lena, int64_t lenb, const void* d_hoz,
                            const void* d_coeff, void* d_work, void* d_out,
                            const int32_t* d_done, void* stream);
/* d_out[Ia][Ib'] = sum_Ib d_coeff[Ia][Ib] Hbb[Ib'][Ib] for complex64
 * d_coeff / d_out ([lena][lenb] interleaved), BF16x3 products with FP32
 * accumulation.  The opposite-spin and alpha parts of the split apply
 * run in fqehip_apply_fused_rows after it. */
int fqehip_beta_gemm(int64_t lena, int64_t lenb, const uint16_t* d_hhi,
                     const uint16_t* d_hlo, int64_t ldh, const void* d_coeff,
                     void* d_out, const int32_t* d_done, void* stream);

/* out[Ia][Ib'] = sum_Ib C[Ia][Ib] Hbb[Ib'][Ib] for complex128 C on the FP64
 * matrix cores, from the FP64 columns d_hcol ([Ib][Ib'], the hcol output
 * of fqehip_beta_block): the beta-beta half of the complex128 split apply.
 * Replaces the same-spin beta part of lm_apply_array12_same_spin_opt
 * (src/fqe/lib/fqe_data.c:668-720) as one dense GEMM.  Skipped when
 * d_done and *d_done. */
int fqehip_beta_gemm_f64(int64_t lena, int64_t lenb, const double* d_hcol,
                         const void* d_coeff, void* d_out,
                         const int32_t* d_done, void* stream);
int fqehip_apply_fused_rows(int precision, const fqehip_sector* s,
                            const fqehip_operator* op, const void* d_coeff,
                            void* d_workspace, const int32_t* d_done,
                            void* stream);
int fqehip_apply_slot_reduce(int precision, const fqehip_sector* s,
                             const void* d_workspace, double alpha_re,
                             double alpha_im, double beta_re, double beta_im,
                             const void* d_cin, void* d_out, double* d_acc,
                             double* d_norm_partials, const int32_t* d_done,
                             void* stream);

/* Bytes of workspace for one apply of a sector: the larger of the unfused
 * path (dvec + Y planes) and the fused path (slot rows + row
 * accumulators). */
int64_t fqehip_apply_workspace_bytes(int precision, const fqehip_sector* s);

/* out = alpha * H C + beta * cin (see fqehip_scatter for acc/partials).
 * Replaces FqeData.apply_inplace for (h1, h2) tuples
 * (src/fqe/fqe_data.py:421-475, :582-710) and Wavefunction._apply_array
 * (src/fqe/wavefunction.py:366-399).  d_workspace holds
 * fqehip_apply_workspace_bytes(); d_coeff may alias d_out only if
 * beta == 0. */
int fqehip_apply(int precision, const fqehip_sector* s,
                 const fqehip_operator* op, const void* d_coeff,
                 double alpha_re, double alpha_im, double beta_re,
                 double beta_im, const void* d_cin, void* d_out,
                 double* d_acc, double* d_norm_partials,
                 const int32_t* d_done, void* d_workspace, void* stream);

/* fqehip_apply without a caller workspace and without any device
 * allocation.  No counterpart in the reference (its C kernels allocate
 * scratch with malloc, src/fqe/lib/fqe_data.c:879-996): it is for callers
 * that cannot allocate device memory, such as code under HIP graph capture
 * and tight serving loops.  The workspace is a library-owned device array
 * of FQEHIP_NOALLOC_SCRATCH_BYTES per GPU, held from library load until
 * the process ends; a sector whose apply needs more returns
 * FQEHIP_ERR_UNSUPPORTED (fqehip_apply_workspace_bytes bounds the need;
 * the fused path needs (nslot+1)*2*lena*lenb elements, e.g. norb=12
 * 6a6b complex64 fits, complex128 does not).  Arguments and results are
 * those of fqehip_apply.  All calls on one GPU share its array: calls that
 * overlap on one GPU (other streams, other host threads) race. */
#define FQEHIP_NOALLOC_SCRATCH_BYTES (256ll * 1024 * 1024)
int fqehip_apply_noalloc(int precision, const fqehip_sector* s,
                         const fqehip_operator* op, const void* d_coeff,
                         double alpha_re, double alpha_im, double beta_re,
                         double beta_im, const void* d_cin, void* d_out,
                         double* d_acc, double* d_norm_partials,
                         const int32_t* d_done, void* stream);

/* fqehip_apply on the unfused gather / GEMM / scatter kernels in panels of
 * orbital pairs, for sectors whose dvec and Y (2 npair + 1 planes of
 * 16 lena lenb bytes each) outgrow device memory: per panel of output pairs
 * (the h1' row included), the input pairs are gathered and contracted
 * panel by panel into that panel's Y, which is then scattered into an FP64
 * accumulator; a last pass applies alpha, beta and the epilogue of
 * fqehip_scatter.  Panel sizes follow workspace_bytes (at least 3 planes:
 * the accumulator, one input and one output pair); the whole dvec is
 * gathered once when it fits, else once per output panel.  complex128,
 * two-body operators (real or complex).  Replaces the dvec path of
 * src/fqe/fqe_data.py:646-683 (the reference has no chunked variant: its
 * dvec path allocates the whole dvec, :2209-2234, and its C lm kernels
 * avoid the dvec, src/fqe/lib/fqe_data.c:668-996). */
int fqehip_apply_panels(int precision, const fqehip_sector* s,
                        const fqehip_operator* op, const void* d_coeff,
                        double alpha_re, double alpha_im, double beta_re,
                        double beta_im, const void* d_cin, void* d_out,
                        double* d_acc, double* d_norm_partials,
                        const int32_t* d_done, void* d_workspace,
                        int64_t workspace_bytes, void* stream);

/* Taylor series psi(t) = sum_k (-i t H)^k psi / k! for a real operator
 * (time folded into the scalar), reference
 * src/fqe/wavefunction.py:556-568.  d_term (storage precision) holds psi
 * on entry and is overwritten; d_result (complex128) receives the sum.
 * Stops when ||term_k|| < accuracy or after expansion-1 terms; the
 * number of terms and the converged flag are written to d_status[0..1]
 * (int32: nterms, converged) without a host synchronisation.  Scratch:
 * d_workspace (fqehip_apply_workspace_bytes), d_partials
 * (fqehip_norm_partials_count doubles), d_norm (1 double). */
int fqehip_taylor(int precision, const fqehip_sector* s,
                  const fqehip_operator* op, double time, double accuracy,
                  int expansion, void* d_term, double* d_result,
                  int32_t* d_status, double* d_partials, double* d_norm,
                  void* d_workspace, void* stream);

/* One step of the Chebyshev propagator's three-term recurrence, after the
 * apply that wrote (2/a) H T_(k-1) - T_(k-2) into d_tk (fqehip_apply with
 * alpha = 2/a, beta = -1):
 *   T_k += gamma T_(k-1) (skipped when d_tkm1 is NULL);
 *   result += c T_k (complex128);
 *   d_norm_partials[block] = sum of |T_k|^2 (optional; the blocks of
 *   fqehip_norm_partials_count(n), for fqehip_norm_finalize).
 * Does nothing when d_done and *d_done.  Replaces the per-term axpys of the
 * reference's Chebyshev branch (src/fqe/wavefunction.py:570-602), which
 * make four passes over host wavefunctions per term. */
int fqehip_cheb_update(int precision, int64_t n, void* d_tk,
                       const void* d_tkm1, double gamma_re, double gamma_im,
                       double c_re, double c_im, double* d_result,
                       double* d_norm_partials, const int32_t* d_done,
                       void* stream);

/* The series of fqehip_taylor recorded once into a HIP graph (on a private
 * stream; every apply, norm reduction and device-flag test of the
 * expansion - 1 terms is a node) and replayed with one launch: one host
 * call per time_evolve instead of one per kernel.  Buffers are referenced,
 * not copied: each launch starts from what d_term holds and adds into
 * d_result (fill both before a launch, as for fqehip_taylor), and writes
 * d_status.  The reference loops over its terms in Python
 * (src/fqe/wavefunction.py:556-568). */
typedef struct fqehip_taylor_graph fqehip_taylor_graph;
int fqehip_taylor_graph_create(int precision, const fqehip_sector* s,
                               const fqehip_operator* op, double time,
                               double accuracy, int expansion, void* d_term,
                               double* d_result, int32_t* d_status,
                               double* d_partials, double* d_norm,
                               void* d_workspace, fqehip_taylor_graph** out);
int fqehip_taylor_graph_launch(fqehip_taylor_graph* graph, void* stream);
void fqehip_taylor_graph_destroy(fqehip_taylor_graph* graph);

/* ---------------------------------------------------------------------
 * Davidson eigensolver kernels (the reference's davidsonliu_fqe,
 * src/fqe/algorithm/davidson.py:140-307, builds its preconditioner from
 * one expectation value per determinant, :160-180; here the diagonal is
 * the Slater-Condon closed form).  Vectors: complex128 [n] interleaved; a
 * block of k vectors is k rows of stride ldv elements.
 * ------------------------------------------------------------------- */
/* d_diag[Ia][Ib] = <I|H|I> = sum_i h1[i,i](na_i + nb_i)
 *   + sum_{i!=j} (K_ij - J_ij)(na_i na_j + nb_i nb_j)
 *   - sum_ij J_ij (na_i nb_j + nb_i na_j),
 * J_ij = Re h2[i,j,i,j], K_ij = Re h2[i,j,j,i] (norb x norb, FP64), from
 * the string tables d_astr / d_bstr (fqehip_string_tables). */
int fqehip_diagonal(int norb, int64_t lena, int64_t lenb,
                    const int64_t* d_astr, const int64_t* d_bstr,
                    const double* d_h1diag, const double* d_jmat,
                    const double* d_kmat, double* d_diag, void* stream);
/* d_out[j] = sum_i conj(v_j[i]) x[i], j < k, FP64 with a fixed summation
 * order; d_workspace holds fqehip_block_dot_workspace_bytes(k) bytes. */
int64_t fqehip_block_dot_workspace_bytes(int k);
int fqehip_block_dot(int64_t n, int k, const double* d_v, int64_t ldv,
                     const double* d_x, double* d_out, double* d_workspace,
                     void* stream);
/* x = beta x + sum_j c_j v_j (d_c: k complex128 on the device). */
int fqehip_block_combine(int64_t n, int k, const double* d_v, int64_t ldv,
                         const double* d_c, double beta_re, double beta_im,
                         double* d_x, void* stream);
/* d_out[j][i] = sum_l d_c[l][j] d_v[l][i] for j < m (m <= 16; complex128
 * rows of length n, strides ldv and ldo; d_c complex [k][m] row-major): a
 * thick restart's new basis or images with each row read once instead of
 * once per output (the reference rebuilds its basis in Python,
 * src/fqe/algorithm/davidson.py:211-216). */
int fqehip_block_combine_multi(int64_t n, int k, const double* d_v,
                               int64_t ldv, const double* d_c, int m,
                               double* d_out, int64_t ldo, void* stream);
/* t[i] = -(1/(diag_i - theta) + corr_i) r[i]; |diag_i - theta| is floored
 * at eps; d_corr (FP32 [n], the learned correction) may be NULL.  The
 * reference's update is reciprocal(theta - diag) * residual
 * (src/fqe/algorithm/davidson.py:282-291). */
int fqehip_davidson_precondition(int64_t n, const double* d_r,
                                 const double* d_diag, double theta,
                                 double eps, const float* d_corr, double* d_t,
                                 void* stream);

/* ---------------------------------------------------------------------
 * Learned correction of the Davidson preconditioner (no counterpart in the
 * reference, whose preconditioner is the plain diagonal,
 * src/fqe/algorithm/davidson.py:282-291).  A 3-layer MLP
 * 10 -> 128 -> 128 -> 1 (tanh), FP32, parameters W1 [128][10], b1 [128],
 * W2 [128][128], b2 [128], w3 [128], b3 [1].  Features of determinant I
 * at Ritz value theta (d = <I|H|I>, dmin = min d, reference strings ref_a,
 * ref_b with their highest occupied / lowest empty orbitals):
 *   d - theta, 1/(|d - theta| + 0.05), d - dmin, theta - dmin,
 *   popcount(Ia & ~ref_a), popcount(Ib & ~ref_b),
 *   Ia[homo_a], Ia[lumo_a], Ib[homo_b], Ib[lumo_b].
 * ------------------------------------------------------------------- */
#define FQEHIP_MLP_NFEAT 10
#define FQEHIP_MLP_WIDTH 128
#define FQEHIP_MLP_NPARAMS 18049
/* x[s][f] for the determinants d_idx[s] (flat Ia*lenb+Ib), s < m. */
int fqehip_mlp_features(int norb, int64_t lena, int64_t lenb,
                        const int64_t* d_astr, const int64_t* d_bstr,
                        const double* d_diag, double theta, double dmin,
                        int64_t ref_a, int64_t ref_b, int64_t m,
                        const int64_t* d_idx, float* d_x, void* stream);
/* Inference for every determinant: d_corr[I] (FP32); d_xwork holds the
 * features, lena*lenb*FQEHIP_MLP_NFEAT floats. */
int fqehip_mlp_correction(const float* d_params, int norb, int64_t lena,
                          int64_t lenb, const int64_t* d_astr,
                          const int64_t* d_bstr, const double* d_diag,
                          double theta, double dmin, int64_t ref_a,
                          int64_t ref_b, float* d_xwork, float* d_corr,
                          void* stream);
/* Network output of m feature rows. */
int fqehip_mlp_forward(const float* d_params, int64_t m, const float* d_x,
                       float* d_y, void* stream);
/* Gradient of L = sum_s w_s (y_s - t_s)^2 over m samples:
 * d_grad[0..NPARAMS) = dL/dparams, d_grad[NPARAMS] = L (fixed summation
 * order).  d_workspace: fqehip_mlp_train_workspace_floats(m) floats. */
int64_t fqehip_mlp_train_workspace_floats(int64_t m);
int fqehip_mlp_loss_grad(const float* d_params, int64_t m, const float* d_x,
                         const float* d_target, const float* d_weight,
                         float* d_grad, float* d_workspace, void* stream);
/* One Adam step (bias-corrected, step >= 1) on the parameter vector. */
int fqehip_mlp_adam(float* d_params, const float* d_grad, float* d_m,
                    float* d_v, int step, float lr, float beta1, float beta2,
                    float eps, void* stream);

/* ---------------------------------------------------------------------
 * Batched sectors: one apply and a fixed-length Taylor series per sector,
 * recorded once into a HIP graph over several capture streams and replayed
 * with one launch.  The reference evolves sectors one by one from Python
 * (src/fqe/wavefunction.py:509-611, :961-1054); its Taylor loop stops on
 * convergence instead of after a fixed number of terms (:556-568).
 * ------------------------------------------------------------------- */
typedef struct fqehip_batch_item {
  fqehip_sector sector;      /* copied; its tables must outlive the batch */
  fqehip_operator op;        /* real or complex two-body operator */
  const void* d_coeff0;      /* C0 (storage precision), read only */
  void* d_apply_out;         /* H C0, or NULL to skip the apply */
  void* d_term;              /* Taylor term (storage precision) */
  double* d_result;          /* complex128 sum_{k<=nterms} (-i t H)^k C0 / k! */
  void* d_workspace;         /* fqehip_apply_workspace_bytes(sector) */
} fqehip_batch_item;

typedef struct fqehip_batch fqehip_batch;

/* Records the batch into a graph with up to nstreams parallel branches
 * (1..32), on capture streams of the batch's own (`stream` is not used
 * while recording).  Buffers are referenced, not copied: every launch on
 * a stream reads d_coeff0 and rewrites the outputs.  With group != 0 the
 * sectors that take the fused row kernel without the split (hbb_hi and
 * hbb_f64 NULL, tile not FQEHIP_TILE_OZAKI) run as one group: per Taylor
 * step one row launch and one reduction cover all of them (complex64: the
 * rows4 row body; complex128: the FP64 row body), instead of kernels per
 * sector; the other sectors keep their own kernels on the other branches. */
int fqehip_batch_create(int precision, int nitems,
                        const fqehip_batch_item* items, double time,
                        int nterms, int nstreams, int group, void* stream,
                        fqehip_batch** out);
int fqehip_batch_launch(fqehip_batch* batch, void* stream);
int fqehip_batch_nodes(const fqehip_batch* batch);
/* Number of sectors the batch runs as one group. */
int fqehip_batch_grouped(const fqehip_batch* batch);
void fqehip_batch_destroy(fqehip_batch* batch);

/* ---------------------------------------------------------------------
 * Complex Gram matrix of two blocks of k complex128 vectors of length n:
 * d_out[p][q] = sum_i conj(a[p][i]) b[q][i] (complex128 [k][k]), FP64
 * matrix cores, fixed summation order.  The 2-RDM's dvec overlap
 * (src/fqe/fqe_data.py:1767-1782).  d_workspace:
 * fqehip_gram_workspace_bytes(n, k) bytes; k <= 4096.
 * ------------------------------------------------------------------- */
int64_t fqehip_gram_workspace_bytes(int64_t n, int k);
int fqehip_gram(int64_t n, int k, const double* d_a, const double* d_b,
                double* d_out, double* d_workspace, void* stream);

/* ---------------------------------------------------------------------
 * Spin-orbital apply on a full particle-number set (FqeDataSet): every
 * sector (nalpha, nbeta) with nalpha + nbeta = nele and both <= norb, in
 * ascending nbeta, concatenated into one complex128 vector of
 * fqehip_set_length(norb, nele) elements (each sector row-major
 * lena x lenb).  Spin-orbital index p < norb is alpha orbital p, p >= norb
 * beta orbital p - norb; M = 2 norb; norb <= 16.  Complex arrays are
 * interleaved complex128.
 *
 * fqehip_set_dvec: D[pq] = E_pq C for pq = p*M + q, planar doubles
 * [M*M][2][L] (real parts, then imaginary parts, of each pair).  Replaces
 * FqeDataSet._calculate_dvec_with_coeff (src/fqe/fqe_data_set.py:737-796).
 *
 * fqehip_set_apply with d_v == NULL: out = sum_pq h1[p][q] E_pq C
 * (FqeDataSet._apply1, src/fqe/fqe_data_set.py:153-181); no workspace.
 * With d_v: out = sum_pq h1[p][q] D[pq] + sum_pq E_pq (V . D)[pq], where the
 * caller passes the dvec-form tensors V[i][j][k][l] = -h2[i][k][j][l] and
 * h1' = h1 - sum_k V[:, k, k, :] (FqeDataSet._apply12,
 * src/fqe/fqe_data_set.py:183-209); d_workspace holds
 * fqehip_set_workspace_bytes(norb, nele) bytes.  Spin-flip signs are the
 * reference's (src/fqe/fci_graph_set.py:137-150).  Gathers only: the same
 * bits on every run.  d_out must not alias d_coeff.
 * ------------------------------------------------------------------- */
int64_t fqehip_set_length(int norb, int nele);
int64_t fqehip_set_workspace_bytes(int norb, int nele);
int fqehip_set_dvec(int norb, int nele, const double* d_coeff,
                    double* d_dvec, void* stream);
int fqehip_set_apply(int norb, int nele, const double* d_h1,
                     const double* d_v, const double* d_coeff, double* d_out,
                     double* d_workspace, void* stream);
/* The contraction stage of the two-body apply on its own, on the
 * workspace of fqehip_set_apply: D (planar [M^2][2][L], as fqehip_set_dvec
 * writes it) at its start, Y[pq] = sum_rs V[pq][rs] D[rs] (same layout)
 * after it, as one real FP64 matrix-core GEMM with the block form
 * [[Re V, -Im V], [Im V, Re V]] (the V . D tensordot of
 * src/fqe/fqe_data_set.py:204-205). */
int fqehip_set_contract(int norb, int nele, const double* d_v,
                        double* d_workspace, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FQEHIP_H_ */
