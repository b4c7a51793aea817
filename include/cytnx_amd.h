/*
 * cytnx_amd.h -- C ABI of the MI355X (gfx950) contraction hot path.
 *
 * Plain C types only: raw device pointers, sizes, host arrays of metadata, an opaque stream.
 * Every function returns a status code (CXA_OK on success) and never exits the process; the
 * message of the last failure on the calling thread is returned by cxa_last_error().
 * Device buffers always belong to the caller: no entry point allocates a user-visible output.
 * All device work is enqueued on the given stream (NULL = the null stream) and is asynchronous
 * unless the comment says otherwise.
 *
 * Each entry point names the reference (Cytnx v1.1.1) interface it replaces at the
 * Device.cuda dispatch boundary; paths are relative to the Cytnx source tree.
 */
#ifndef CYTNX_AMD_H
#define CYTNX_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------------ */
#define CXA_OK 0
#define CXA_ERR_INVALID 1     /* argument validation failed (the reference throws cytnx::error) */
#define CXA_ERR_UNSUPPORTED 2 /* dtype / rank / mode this build does not handle                */
#define CXA_ERR_HIP 3         /* a HIP runtime call failed                                     */
#define CXA_ERR_NOMEM 4       /* host or device allocation failed                              */
#define CXA_ERR_NUMERIC 5     /* a numerical routine did not converge                          */

/* ---- dtype ids: the reference's Type_list order (include/Type.hpp:177-179) ---------------- */
#define CXA_VOID 0
#define CXA_C128 1
#define CXA_C64 2
#define CXA_F64 3
#define CXA_F32 4
#define CXA_I64 5
#define CXA_U64 6
#define CXA_I32 7
#define CXA_U32 8
#define CXA_I16 9
#define CXA_U16 10
#define CXA_BOOL 11

/* ---- contraction precision modes ---------------------------------------------------------- */
#define CXA_PREC_FP64 0   /* FP64 MFMA, FP64 accumulation: the drop-in path                   */
#define CXA_PREC_FP32 1   /* operands rounded to FP32, FP32 MFMA, FP32 accumulation            */
#define CXA_PREC_FP32X3 2 /* FP64 operands split hi+lo in FP32; hi*hi + hi*lo + lo*hi on FP32  */
                          /* MFMA with FP32 accumulation                                      */

typedef void *cxa_stream_t; /* a hipStream_t */

/* Message of the last failed call on this thread ("" if none). */
const char *cxa_last_error(void);

/* ABI version of this library (increments on any signature change). */
int cxa_abi_version(void);

/*
 * Physical permute (Tensor::contiguous_ after a lazy permute_).
 * Replaces utils_internal::MoveMemoryGpu<T>
 *   (src/backend/utils_internal_gpu/cuMovemem_gpu.cu:289-304), reached from
 *   StorageImplementation<T>::Move_memory (src/backend/StorageImplementation.cpp:161-177).
 * src is the row-major array of shape old_shape[rank]; dst receives the row-major array whose
 * axis j is old axis mapper[j] (dst = numpy.transpose(src, mapper)).  Any dtype id.
 * dst must be caller-allocated (prod(old_shape) elements) and must not overlap src.
 * rank <= 16.
 */
int cxa_permute(int dtype, int rank, const uint64_t *old_shape, const uint64_t *mapper,
                const void *src, void *dst, cxa_stream_t stream);

/*
 * Block permute: every block of a block (U(1)) tensor permuted by one mapper, the GPU part of
 * BlockUniTensor::permute followed by contiguous() (src/BlockUniTensor.cpp:565-633, which
 * permutes block by block).  Block b is the row-major array of extents shapes[b*rank ..] at
 * element offset src_offsets[b] of src; its permuted copy (axis j = old axis mapper[j]) goes to
 * element offset dst_offsets[b] of dst.  The plan is built once from the metadata (host): blocks
 * of up to 2^15 elements of 4-, 8- or 16-byte dtypes are moved by one batched launch when there
 * are two or more of them, the others by cxa_permute each.  exec enqueues on `stream`; dst must
 * not overlap src.  A plan with batched blocks holds their task list in the memory of the
 * device current at its creation and runs there only.
 */
typedef struct cxa_bp_plan cxa_bp_plan;
int cxa_block_permute_plan(int dtype, int rank, int64_t nblocks, const uint64_t *shapes,
                           const int64_t *src_offsets, const int64_t *dst_offsets,
                           const uint64_t *mapper, cxa_bp_plan **plan);
int cxa_block_permute_exec(const cxa_bp_plan *plan, const void *src, void *dst,
                           cxa_stream_t stream);
int cxa_block_permute_destroy(cxa_bp_plan *plan);

/*
 * Dense row-major matrix product out(Ml x Nr) = inl(Ml x Comm) . inr(Comm x Nr), beta = 0.
 * Replaces cuMatmul_internal_{d,f,cd,cf} (src/backend/linalg_internal_gpu/cuMatmul_internal.cu
 * :25-103; type Matmulfunc_oii, src/backend/linalg_internal_interface.hpp:128-131), called
 * from linalg::Matmul (src/linalg/Matmul.cpp:54-59).  dtype: F64, F32, C128, C64.  Comm may be
 * 0 (out is zero-filled).  FP64 / C128 run on FP64 MFMA, F32 / C64 on FP32 MFMA.
 */
int cxa_matmul(int dtype, void *out, const void *inl, const void *inr, int64_t Ml, int64_t Comm,
               int64_t Nr, cxa_stream_t stream);

/*
 * Batched GEMM in the BLAS column-major convention, one grouped launch for all matrices.
 * Replaces cuGemm_Batch_internal_* (src/backend/linalg_internal_gpu/cuGemm_Batch_internal.cu
 * :10-65; type Gemm_Batchfunc_oii, src/backend/linalg_internal_interface.hpp:100-105), called
 * from linalg::Gemm_Batch (src/linalg/Gemm_Batch.cpp:225-230).  For group g and each of its
 * group_size[g] matrices i (numbered consecutively across groups):
 *   C_i = alpha[g] * op(A_i) * op(B_i) + beta[g] * C_i,  op = 'N' or 'T' (transa/transb[g]),
 *   op(A_i) is m[g] x k[g], op(B_i) is k[g] x n[g], leading dimensions lda/ldb/ldc[g].
 * alpha/beta are the reference's per-group Scalars: group_count doubles for F64 / F32, and
 * 2 * group_count doubles for C128 / C64 ((re, im) of group g at [2g], [2g + 1]).  a/b/c are
 * host arrays of device pointers.  Synchronous with respect to the stream: returns after the
 * launch that reads the uploaded task list has finished.
 */
int cxa_gemm_batch(int dtype, const char *transa, const char *transb, const int64_t *m,
                   const int64_t *n, const int64_t *k, const double *alpha,
                   const void *const *a, const int64_t *lda, const void *const *b,
                   const int64_t *ldb, const double *beta, void *const *c, const int64_t *ldc,
                   int64_t group_count, const int64_t *group_size, cxa_stream_t stream);

/*
 * Dense tensordot: out = sum over the paired axes idxl[i] <-> idxr[i] of L * R, with output
 * axes = free axes of L (in order) then free axes of R (in order).
 * Replaces cuTensordot_internal_* (src/backend/linalg_internal_gpu/cuTensordot_internal.cu
 * :31-188; type Tensordotfunc_oii, src/backend/linalg_internal_interface.hpp:173-175), called
 * from linalg::Tensordot (src/linalg/Tensordot.cpp:158-161).  Operands that remain strided
 * matrices under the implied permutation are read in place; the others are permuted into
 * `workspace` first (size from cxa_tensordot_workspace; may be NULL when that size is 0).
 */
int cxa_tensordot_workspace(int dtype, int rankL, const uint64_t *shapeL, int rankR,
                            const uint64_t *shapeR, int ncomm, const uint64_t *idxl,
                            const uint64_t *idxr, int64_t *bytes);
int cxa_tensordot(int dtype, void *out, const void *L, int rankL, const uint64_t *shapeL,
                  const void *R, int rankR, const uint64_t *shapeR, int ncomm,
                  const uint64_t *idxl, const uint64_t *idxr, void *workspace,
                  cxa_stream_t stream);

/*
 * cxa_tensordot without a workspace argument, allocating nothing: no counterpart in the
 * reference.  It is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  The operand copies that
 * cxa_tensordot permutes into the workspace go to a static device array of the library,
 * CXA_TENSORDOT_SCRATCH_BYTES on every GPU the library runs on (held until the process ends);
 * a call whose cxa_tensordot_workspace size exceeds it returns CXA_ERR_INVALID before any
 * device work.  Results equal cxa_tensordot's bit for bit.  All calls on one GPU share its
 * array: calls that overlap on one GPU (other streams, other threads) race.  The stream must
 * belong to the current device.
 */
#define CXA_TENSORDOT_SCRATCH_BYTES (32ll << 20)
int cxa_tensordot_noalloc(int dtype, void *out, const void *L, int rankL, const uint64_t *shapeL,
                          const void *R, int rankR, const uint64_t *shapeR, int ncomm,
                          const uint64_t *idxl, const uint64_t *idxr, cxa_stream_t stream);

/*
 * Zero-flux block list of a U(1) BlockUniTensor, in the reference's odometer order (last bond
 * fastest).  Replaces the block loop of BlockUniTensor::Init (src/BlockUniTensor.cpp:190-230)
 * with the flux rule of include/UniTensor.hpp:1139-1166 (KET bonds add +q, BRA bonds -q).
 * btype[rank]: -1 KET / +1 BRA; nq[rank]: number of qnums per bond; qnums: all bonds' qnums
 * concatenated.  Writes at most max_blocks tuples (row-major nblocks x rank) to itoi, and the
 * full count to *nblocks (call again with a larger buffer if it exceeds max_blocks).  Host only.
 */
int cxa_u1_enumerate_blocks(int rank, const int *btype, const int64_t *nq, const int64_t *qnums,
                            int64_t max_blocks, uint64_t *itoi, int64_t *nblocks);

/*
 * Block-sparse (U(1) BlockUniTensor) contraction: a plan built once from block metadata, then
 * executed as one grouped MFMA launch over every charge-matched block pair.
 * This is the GPU branch of BlockUniTensor::contract (src/BlockUniTensor.cpp:948-1228), which
 * in the reference has no backend hook and issues one Matmul (+ permute copies and an add pass)
 * per block pair (:1114-1214).  Pairing, target selection and output zero-fill follow it:
 *   for each left block a, each right block b whose qindices on the common legs equal a's
 *   (comm1[i] of L pairs with comm2[i] of R, in the caller's order -- the reference sorts the
 *   common labels, src/utils/vec_intersect.cpp:35-51), out[target] += Lmat_a . Rmat_b, where
 *   target has itoi = (a's free qindices, b's free qindices), Lmat_a = block a with axes
 *   (free ++ common), Rmat_b = block b with axes (common ++ free); output blocks without any
 *   pair are zero.
 * Metadata: itoi[nb*rank] qindex tuples, shape[nb*rank] block extents, off[nb] element offsets
 * of each row-major block within its tensor's buffer.  The output's rank is
 * rankL + rankR - 2*ncomm with axes (L free ++ R free).
 * exec: L, R, C are the buffers the offsets refer to; workspace must hold the bytes reported
 * by cxa_block_contract_info (NULL allowed when 0).  prec: CXA_PREC_* for real dtypes
 * (complex dtypes always use their native precision).  Plans are bound to the device that was
 * current at creation and may be executed concurrently on different streams with different
 * workspaces.  A plan makes the tile list of its dtype's own precision at creation and those of
 * the other real precisions on their first exec, configuration or chain preparation (a
 * synchronous upload): run or configure a precision once before capturing its exec in a graph
 * (an exec inside a capture that would need the upload returns CXA_ERR_INVALID).
 */
typedef struct cxa_bc_plan cxa_bc_plan;
int cxa_block_contract_plan(int dtype, int rankL, int nbL, const uint64_t *itoiL,
                            const uint64_t *shapeL, const int64_t *offL, int rankR, int nbR,
                            const uint64_t *itoiR, const uint64_t *shapeR, const int64_t *offR,
                            int ncomm, const uint64_t *comm1, const uint64_t *comm2, int nbC,
                            const uint64_t *itoiC, const uint64_t *shapeC, const int64_t *offC,
                            cxa_bc_plan **plan);
/* flops: sum over block pairs of 2*M*N*K (x4 for complex); npairs: number of block pairs. */
int cxa_block_contract_info(const cxa_bc_plan *plan, double *flops, int64_t *workspace_bytes,
                            int64_t *npairs);
/* Override the tile shape of every tile (-1 = per-block default: MFMA 16x16..32x32 tiles,
 * row streaming for blocks with N <= 8) and the workgroup split-K mode (-1 automatic, 0 off,
 * 1 on with four waves per tile, 2 on with eight waves per tile for MFMA tiles); used by the
 * tuner.  get_config reports the split mode in the same 0 / 1 / 2 form. */
int cxa_block_contract_set_tile(cxa_bc_plan *plan, int kind);
int cxa_block_contract_set_split(cxa_bc_plan *plan, int mode);
/* The same for one precision mode only (prec = -1: all); get returns the chosen shape, whether
 * split-K is on and the tile count for a precision. */
int cxa_block_contract_set_config(cxa_bc_plan *plan, int prec, int kind, int split);
int cxa_block_contract_get_config(const cxa_bc_plan *plan, int prec, int *kind, int *split,
                                  int *ntiles);
int cxa_block_contract_exec(const cxa_bc_plan *plan, const void *L, const void *R, void *C,
                            void *workspace, int prec, cxa_stream_t stream);
/* exec with operand conjugation for complex plans (ignored for real ones): flags bit 0
 * conjugates L, bit 1 conjugates R as they are read.  Used by the adjoint (reverse-mode)
 * contractions, dL = dC . conj(R)^T and dR = conj(L)^T . dC, which have no counterpart in the
 * reference (Cytnx has no automatic differentiation); callers are differentiable front ends. */
int cxa_block_contract_exec_ex(const cxa_bc_plan *plan, const void *L, const void *R, void *C,
                               void *workspace, int prec, int flags, cxa_stream_t stream);
int cxa_block_contract_destroy(cxa_bc_plan *plan);

/*
 * Strided-batch plan: for b in [0, nbatch), C_b = A_b . B_b (M x N, contraction over K) with
 * every operand a strided matrix view of one buffer (element offsets):
 *   A_b[m, k] at b*a_bstride + m*sam + k*sak,  B_b[k, n] at b*b_bstride + k*sbk + n*sbn,
 *   C_b[m, n] at b*c_bstride + m*ldc + n.
 * No counterpart in the reference, whose dense contractions permute every operand into GEMM
 * order first (src/linalg/Tensordot.cpp:29-110); for contractions whose free or contracted
 * legs are interleaved with a batch leg (a leg carried through unchanged), a batch over that
 * leg reads and writes every tensor in place.  Runs through the block-contract exec / config
 * / destroy entry points and in chains.
 */
int cxa_strided_batch_plan(int dtype, int64_t nbatch, int64_t M, int64_t N, int64_t K,
                           int64_t a_bstride, int64_t sam, int64_t sak, int64_t b_bstride,
                           int64_t sbk, int64_t sbn, int64_t c_bstride, int64_t ldc,
                           cxa_bc_plan **plan);

/*
 * Mix plan: a block contraction followed by a small linear map, in one launch.  Item i (of
 * nitems) has shape M x N, ns slices and no outputs (items[4*i..] = M, N, ns, no); its slices
 * are the next ns rows of `slices` (7 int64 each: A offset, B offset, A strides along m and k,
 * B strides along k and n, K -- element units, as in the block-contract plans), its outputs the
 * next no rows of `outs` (c_off, ldc) with weights coef[CXA_MIX_SLICES*o + s].  exec computes
 *   X_s = A_s . B_s (M x N each, never stored),   C[c_off + m*ldc + n] = sum_s coef[s] X_s[m, n]
 * for every output.  ns <= CXA_MIX_SLICES.  It replaces, for the DMRG effective Hamiltonian,
 * the contractions of L, psi and the two MPO tensors in the matvec network
 * (example/DMRG/dmrg_two_sites_U1.py:18-37, launched per matvec at :20): the slices are the
 * (MPO index, physical indices) slices of L.psi and the weights are the elements of the
 * two-site MPO M1.M2.  FP64 operands only
 * (FP32 copies through a chain); exec / info / set_config / get_config / destroy are the
 * block-contract entry points (tile kinds 16x16, 32x16, 16x32, 32x32; split mode 1 selects
 * four-wave workgroups, two slices per wave, instead of one wave per slice; flops count the
 * products).
 */
#define CXA_MIX_SLICES 8
int cxa_mix_contract_plan(int dtype, int nitems, const int64_t *items, int64_t nslices,
                          const int64_t *slices, int64_t nouts, const int64_t *outs,
                          const double *coef, cxa_bc_plan **plan);

/*
 * The task arrays of cxa_mix_contract_plan for the two-site effective Hamiltonian (host only; no
 * counterpart in the reference, which contracts the network of example/DMRG/
 * dmrg_two_sites_U1.py:18-37 pairwise): L permuted to (l, w, l') (nbL blocks of rank 3), psi
 * (l, s1, s2, r) (nbP blocks of rank 4), the output layout Y with legs (l', s1', s2', w'', r) in
 * any order (nbY blocks of rank 5; r, or l' when it comes after r, contiguous).  pos* give each
 * role's leg index (L: l, w, l'; psi: l, s1, s2, r; Y: l', s1', s2', w'', r); itoi / shape
 * row-major per block, off element offsets.  sec_off: the sector offsets of the bonds w (L),
 * s1, s2 (psi), s1', s2', w'' (Y), concatenated, sec_n their counts.  W12: the dense two-site
 * MPO (w, s1, s1', w'', s2, s2') of extents wdims[6].  One item per (l', r) sector pair of Y;
 * its slices are the L.psi products of every (w, s1, s2) element reaching one of its outputs,
 * its outputs the (s1', s2', w'') sub-matrices with their W12 weights.  read (nbY, may be NULL):
 * the Y blocks the next launch reads; with it the plan addresses the pruned layout (blocks
 * reached by a slice and read, in order, packed; offY unused) and items follow their first
 * block; live (nbY, may be NULL) receives the kept (or, without read, the reached) blocks.
 * Outputs: items (4 each), slices (7), outs (2), coef (CXA_MIX_SLICES each), within the given
 * capacities; counts = (items, slices, outs).  CXA_ERR_UNSUPPORTED if an item needs more than
 * CXA_MIX_SLICES slices.
 */
int cxa_heff2_mix_plan(int nbL, const int64_t *itoiL, const int64_t *shapeL, const int64_t *offL,
                       const int *posL, int nbP, const int64_t *itoiP, const int64_t *shapeP,
                       const int64_t *offP, const int *posP, int nbY, const int64_t *itoiY,
                       const int64_t *shapeY, const int64_t *offY, const int *posY,
                       const int64_t *sec_off, const int64_t *sec_n, const double *W12,
                       const int64_t *wdims, const unsigned char *read, int64_t cap_items,
                       int64_t cap_slices, int64_t cap_outs, int64_t *items, int64_t *slices,
                       int64_t *outs, double *coef, unsigned char *live, int64_t *counts);

/*
 * Vector reductions into device memory; workspace: device buffer of cxa_vec_workspace_bytes()
 * bytes.  Deterministic (fixed reduction order); products are summed in FP64 for every dtype.
 * cxa_dot replaces cuVectordot_internal_{d,f,cd,cf} (src/backend/linalg_internal_gpu/
 * cuVectordot_internal.cu:10-180, type Vectordotfunc_oii, src/backend/linalg_internal_interface
 * .hpp:141-144): result = sum_i x_i y_i, with is_conj sum_i conj(x_i) y_i, in the dtype (two
 * values re, im for C128 / C64).  cxa_nrm2 replaces cuNorm_internal_{d,f,cd,cf}
 * (src/backend/linalg_internal_gpu/cuNorm_internal.cu:14-60, type Normfunc_oii, :87):
 * result[0] = sqrt(sum_i |x_i|^2) in the dtype's real type.  dtype: F64, F32, C128, C64.
 */
int cxa_vec_workspace_bytes(void);
int cxa_dot(int dtype, int64_t n, const void *x, const void *y, int is_conj, void *result,
            void *workspace, cxa_stream_t stream);
int cxa_nrm2(int dtype, int64_t n, const void *x, void *result, void *workspace,
             cxa_stream_t stream);

/*
 * Lowest eigenpair of the symmetric tridiagonal matrix with diagonal a[n] and off-diagonal
 * b[n-1] (host).  Replaces the LAPACKE_dstev call that linalg::Tridiag makes
 * (src/backend/linalg_internal_cpu/Tridiag_internal.cpp:9-42) for the ground-state Lanczos,
 * which only reads eigenpair 0.  z (may be NULL) receives the unit eigenvector (sign arbitrary).
 */
int cxa_tridiag_ground(int n, const double *a, const double *b, double *E, double *z);

/*
 * All eigenpairs of the same symmetric tridiagonal matrix (host): linalg::Tridiag(..., is_V)
 * (src/linalg/Tridiag.cpp:10-104, LAPACKE_dstev in
 * src/backend/linalg_internal_cpu/Tridiag_internal.cpp:9-42).  w[n] ascending; z (may be NULL,
 * n x n) row k = the unit eigenvector of w[k] (sign arbitrary).  Implicit symmetric QR
 * iteration with Wilkinson shifts.
 */
int cxa_tridiag_eigh(int n, const double *a, const double *b, double *w, double *z);

/*
 * Matvec callback used by the Lanczos driver: y = H x for FP64 device vectors, enqueued on
 * `stream`; returns a CXA status.  The LinOp::matvec plug-in point (include/LinOp.hpp:163-167).
 */
typedef int (*cxa_matvec_fn)(void *ctx, const void *x, void *y, cxa_stream_t stream);

/*
 * Matvec chain: a fixed sequence of block-contraction plans applied to a vector.  Step k
 * contracts the running operand (the input vector for k = 0, else step k-1's output) with
 * fixed[k]; fixed_is_left[k] says whether fixed[k] is the plan's L operand.  fixed_elems /
 * out_elems give buffer sizes in elements.  This replaces the per-call Network::Launch of the
 * DMRG effective Hamiltonian (src/RegularNetwork.cpp:901-1007 driven by
 * example/DMRG/dmrg_two_sites_U1.py:12-41) with a planned program whose application allocates
 * nothing.  Plans must be all real FP64 or all C128, and must outlive the chain.  Precision
 * (cxa_chain_prepare) selects FP64 (drop-in), FP32 (FP32 copies of the fixed operands and
 * intermediates; input and output stay FP64) or FP32X3; C128 chains run FP64 only and their
 * vectors are C128 arrays.  cxa_chain_matvec is a cxa_matvec_fn with ctx = the chain.
 * A fixed operand on the plan's right (B) is read at cxa_chain_create into a k-quad packed
 * copy (and at cxa_chain_prepare into its FP32 form); the chain uses those copies, so a fixed
 * operand changed after creation needs a new chain.  cxa_chain_create synchronises the device
 * once when it makes such copies.
 */
typedef struct cxa_chain cxa_chain;
int cxa_chain_create(int nsteps, cxa_bc_plan *const *plans, const void *const *fixed,
                     const int *fixed_is_left, const int64_t *fixed_elems,
                     const int64_t *out_elems, cxa_chain **chain);
int cxa_chain_prepare(cxa_chain *chain, int prec, cxa_stream_t stream);
int cxa_chain_matvec(void *chain, const void *x, void *y, cxa_stream_t stream);
/* one step only (step 0 reads x, the last writes y): per-kernel timing */
int cxa_chain_step(void *chain, int step, const void *x, void *y, cxa_stream_t stream);
/* Autotune: for each step at the chain's current precision, time every candidate tiling
 * (tile shape x workgroup split-K) `reps` times on x -> y and keep the fastest; per-step
 * choices and times (ms) are returned in kinds/splits/ms (arrays of nsteps, may be NULL).
 * In FP32 a step 0 with a packed running vector (cxa_chain_layout bit 3) runs in two forms,
 * reading that copy (FP32 ground-state Lanczos runs) or the FP64 vector (other matvecs): the
 * tiling with the smaller sum of both times is kept and ms reports the packed form's time.
 * Synchronizes the stream. */
int cxa_chain_tune(cxa_chain *chain, const void *x, void *y, int reps, int *kinds, int *splits,
                   double *ms, cxa_stream_t stream);
int cxa_chain_flops(const cxa_chain *chain, double *flops);
/*
 * Operand layouts step `step` runs with (no counterpart in the reference: a diagnostic of this
 * library's packing): bit 0 its fixed B operand is read k4-packed, bit 1 its running A
 * operand (the previous step's output) is read k4-packed, bit 2 its outputs are written
 * k4-packed for the next step, bit 3 (step 0 only, after cxa_chain_prepare for FP32) FP32
 * ground-state runs of cxa_lanczos_run without re-orthogonalisation feed this step a k4-packed
 * FP32 copy of the Lanczos vector, which the residual launch writes.
 */
int cxa_chain_layout(const cxa_chain *chain, int step, int *flags);
int cxa_chain_destroy(cxa_chain *chain);

/*
 * The chain applied `repeats` times to fixed device vectors x -> y, captured once as a graph
 * (current precision) and replayed by cxa_graph_launch on any stream: one host launch per
 * replay instead of one per grouped GEMM.  No counterpart in the reference (its GPU path is
 * synchronous per call).
 */
typedef struct cxa_graph cxa_graph;
int cxa_chain_graph_create(cxa_chain *chain, const void *x, void *y, int repeats,
                           cxa_graph **graph);
int cxa_graph_launch(cxa_graph *graph, cxa_stream_t stream);
int cxa_graph_destroy(cxa_graph *graph);

/*
 * Ground-state Lanczos on FP64 device vectors of length n.  Replaces linalg::Lanczos(...,
 * "Gnd") -> Lanczos_Gnd_Ut (src/linalg/Lanczos.cpp:158-179, src/linalg/Lanczos_Gnd_Ut.cpp
 * :29-199): plain three-term recurrence, stop at the first iteration with
 * |E - E_prev| < CvgCrit or beta == 0, at most Maxiter matvecs, Ritz vector
 * sum_k z_k psi_k.  A context holds the Krylov basis for up to `maxiter` iterations.
 * run: batch = iterations enqueued between host checks (results do not depend on it);
 * flags bit 0 = FP64 full re-orthogonalisation (not in the reference), bit 1 = capture one
 * iteration as a graph and replay it, bit 2 = complex Hermitian operator: the vectors are C128
 * arrays of n / 2 values (n counts doubles and must be even), alpha = Re <v|Hv> as
 * Lanczos_Gnd_Ut takes it (:96-99), and re-orthogonalisation projects with complex
 * coefficients.  Outputs: E_out (host), evec (device, may be NULL), niter (tridiagonal size at
 * the stop), alphas/betas (host, may be NULL, niter entries).
 */
typedef struct cxa_lanczos cxa_lanczos;
int cxa_lanczos_create(int64_t n, int maxiter, cxa_lanczos **ctx);
int cxa_lanczos_run(cxa_lanczos *ctx, cxa_matvec_fn matvec, void *matvec_ctx, const double *Tin,
                    double CvgCrit, int Maxiter, int batch, int flags, double *E_out, double *evec,
                    int *niter, double *alphas, double *betas, cxa_stream_t stream);
int cxa_lanczos_destroy(cxa_lanczos *ctx);
/* No counterpart in the reference (its drivers allocate their Krylov vectors per call): reuse
 * a context's device buffers for vectors of n doubles and Maxiter iterations within what it
 * was created for (n and Maxiter at most the creation values, n * (Maxiter + 1) at most the
 * creation basis), for callers that run many solves of varying size (DMRG bond updates)
 * without freeing and allocating per solve.  Drops captured iteration graphs. */
int cxa_lanczos_reshape(cxa_lanczos *ctx, int64_t n, int maxiter);

/*
 * Lanczos_Exp(Hop, Tin, tau, CvgCrit, Maxiter) with tau = tau_re + i tau_im on device vectors
 * (src/linalg/Lanczos_Exp.cpp:266-390, :404-450): the three-term recurrence from Tin / |Tin|,
 * stopping before v_i at a round-off beta_i (100 eps times the largest |alpha| + beta seen, times
 * sqrt(i)) or after v_i once |exp(tau T)[i, 0]| < CvgCrit or i = min(Maxiter, len) - 1; then
 * out = V exp(tau T) (V^H Tin) (device).  flags bit 2: complex Hermitian operator on C128
 * vectors (n / 2 values, as for cxa_lanczos_run); out holds n doubles.  Real vectors (flags bit
 * 2 clear) with tau_im != 0: out is complex, 2n doubles (n C128 values: a real operator's
 * exp(tau H) Tin for complex tau).  niter = Krylov vectors used.  The context must hold
 * min(Maxiter, len) vectors.
 */
int cxa_lanczos_exp(cxa_lanczos *ctx, cxa_matvec_fn matvec, void *matvec_ctx, const double *Tin,
                    double tau_re, double tau_im, double CvgCrit, int Maxiter, int flags,
                    double *out, int *niter, cxa_stream_t stream);

/*
 * k extreme eigenpairs of a real symmetric (flags bit 2: complex Hermitian, C128 vectors of
 * n / 2 values) operator: linalg::Lanczos(Hop, Tin, which, maxiter, cvg_crit, k, is_V, ncv)
 * (src/linalg/Lanczos.cpp:249-402, :453-583), which the reference runs with ARPACK's implicitly
 * restarted Lanczos (dsaupd/dseupd).  which: 0 = "SA", 1 = "LA", 2 = "LM".  A thick-restart
 * Lanczos with full FP64 re-orthogonalisation: cycles extend the basis to ncv vectors (0: the
 * reference's default min(n, 2k + 10); otherwise k + 2 <= ncv <= n), restart from the
 * k + (ncv - k) / 2 most wanted Ritz vectors, and stop when every wanted Ritz pair has
 * |beta_m y_i[m-1]| <= tol max(eps^(2/3), |theta_i|) (ARPACK's dsconv test; tol <= 0 means
 * machine epsilon) or after Maxiter cycles (the reference's maxiter counts ARPACK's restart
 * iterations).  evals[k] (host) ordered as the reference orders them (SA ascending, LA
 * descending, LM by descending magnitude); evecs (device, may be NULL, k x n) the Ritz vectors
 * as rows; niter = matvecs applied; nconv = pairs within the bound.  The context must hold
 * 2 * ncv vectors (cxa_lanczos_create with maxiter >= 2 * ncv - 1).
 */
int cxa_lanczos_eigsh(cxa_lanczos *ctx, cxa_matvec_fn matvec, void *matvec_ctx, const double *Tin,
                      int which, int k, double tol, int Maxiter, int ncv, int flags,
                      double *evals, double *evecs, int *niter, int *nconv, cxa_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* CYTNX_AMD_H */
