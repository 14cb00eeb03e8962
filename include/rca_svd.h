/*
 * rca_svd.h — C ABI of the MI355X SVD clutter filter for beamformed
 * ensembles (the consumer that follows the beamformer on Doppler data).
 *
 * Replaces Python/clutterfilter/SVDClutterFilter.py:10-23 (SVDClutterFilter)
 * and MATLAB/lib/clutterFilter/SVDClutterFilter.m:1-45.  The reference forms
 * the Casorati matrix Y [nPix x nT] of the ensemble, the Gram matrix Y^H Y,
 * its SVD (the eigen-decomposition of a Hermitian positive semi-definite
 * matrix), and projects Y on a band of the eigenvectors: Y V_k V_k^H.
 *
 * Plain C: device pointers, sizes and scalars.  The Gram matrix is a
 * hand-written FP64 matrix-core kernel; the filter also calls rocSOLVER
 * (ZHEEVD), which keeps a handle and an internal workspace per device.  When
 * the kept or the removed set of eigenvectors has at most 16 members (a
 * clutter filter removing a few tissue vectors), the projection is one
 * hand-written FP64 pass over the complex64 data, Y - (Y U) U^H or (Y U) U^H
 * with U the smaller set; otherwise it is a rocBLAS ZGEMM by the nT x nT
 * projector on an FP64 copy of the data in the caller's workspace.  The
 * call returns when the filter is done (it synchronizes the stream to read
 * the eigensolver's status).
 */
#ifndef RCA_SVD_H
#define RCA_SVD_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Device workspace bytes rca_svd_filter needs for an nPix x nT ensemble. */
size_t rca_svd_workspace_bytes(long long n_pix, int n_t);

/*
 * SVD clutter filter.  No FFI counterpart in the reference (it calls
 * ArrayFire / MATLAB): SVDClutterFilter.py:10-23, SVDClutterFilter.m:17-44.
 *   iq        complex64 ensemble, frame-major: sample (p, t) at iq[t*n_pix + p]
 *             (the beamformer's [nFrame][nPix] output; column-major Casorati
 *             matrix, SVDClutterFilter.m:22)
 *   mask      float [n_pix] noise weights, the data are divided by
 *             sqrt(mask) (SVDClutterFilter.m:21); NULL = no weighting
 *             (SVDClutterFilter.py, which ignores its mask)
 *   keep      host array of n_t bytes: keep[k] != 0 keeps the eigenvector of
 *             the k-th largest eigenvalue of Y^H Y (k = 0 is the largest)
 *   out       complex64 [n_t][n_pix]: Y V_k V_k^H (not re-weighted, as the
 *             MATLAB reference); must not overlap iq
 *   eigval    device double [n_t]: eigenvalues of Y^H Y in descending order
 *             (the reference's singular values s / diag(Eig_val)); may be NULL
 *   workspace device memory of rca_svd_workspace_bytes(n_pix, n_t) bytes
 * Arithmetic: FP64 throughout (Gram, eigen-decomposition, projection).
 */
int rca_svd_filter(const void* iq, const float* mask, long long n_pix, int n_t,
                   const unsigned char* keep, void* out, double* eigval, void* workspace,
                   size_t workspace_bytes, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* RCA_SVD_H */
