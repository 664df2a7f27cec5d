/*
 * cgfd_fno.h -- C-ABI of the BF16 MFMA kernels of the Fourier-neural-operator
 * surrogate (no reference counterpart: the reference has no learned model;
 * the surrogate maps source location, mechanism and moment to surface
 * peak-ground-velocity maps for a fixed medium, trained on runs of this
 * engine).
 */
#ifndef CGFD_FNO_H
#define CGFD_FNO_H

#ifdef __cplusplus
extern "C" {
#endif

#define CGFD_FNO_EINVAL (-1)
#define CGFD_FNO_EHIP (-3)

/* Batched GEMM on the matrix cores, bf16 inputs, fp32 accumulation:
 *   C[z] = A[z] * B[z]   (accumulate = 0)   or   C[z] += A[z] * B[z]
 * for z < batch.  A is M x K with element (m, k) at A + z*sAz + m*sAm + k*sAk,
 * B is K x N with (k, n) at B + z*sBz + k*sBk + n*sBn (bf16 elements, strides
 * in elements), C is M x N fp32 with (m, n) at C + z*sCz + m*sCm + n.  The
 * caller guarantees that every addressed element lies in its buffer.
 * Asynchronous on `stream` (a hipStream_t; null = the default stream).
 * M/64 and batch must not exceed 65535. */
int cgfd_fno_gemm_bf16(const void *A, long sAz, long sAm, long sAk, const void *B, long sBz, long sBk, long sBn,
                       float *C, long sCz, long sCm, int M, int N, int K, int batch, int accumulate, void *stream);

#ifdef __cplusplus
}
#endif

#endif
