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

#ifdef __cplusplus
}
#endif

#endif /* CYTNX_AMD_H */
