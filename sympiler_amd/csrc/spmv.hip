// spmv.hip -- CSR sparse matrix-vector products and the relative residual
// ||L X - B|| / ||B|| reported beside every solve.
//
// Reference: sparse_blas/spmv.cpp:12-21 (y = A x, y zeroed first, entries
// ascending) and spmv.cpp:23-34 (z[i] = a * (z[i] + sum A x) + b * y[i]).
// One lane per row repeats the reference's accumulation order; this file is
// compiled with -ffp-contract=off, so y is bit-identical to the reference.
#include "common.h"

namespace symamd {

namespace {

__global__ __launch_bounds__(256) void k_spmv(int n, const int *__restrict__ Ap,
                                              const int *__restrict__ Ai,
                                              const double *__restrict__ Ax,
                                              const double *__restrict__ x,
                                              double *__restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = 0;
  for (int j = Ap[i]; j < Ap[i + 1]; ++j) acc += Ax[j] * x[Ai[j]];
  y[i] = acc;
}

__global__ __launch_bounds__(256) void k_spmv_axpby(
    int n, const int *__restrict__ Ap, const int *__restrict__ Ai,
    const double *__restrict__ Ax, const double *__restrict__ x,
    const double *__restrict__ y, double *__restrict__ z, double a, double b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = z[i];
  for (int j = Ap[i]; j < Ap[i + 1]; ++j) acc += Ax[j] * x[Ai[j]];
  z[i] = a * acc + b * y[i];
}

// Per column c (blockIdx.y): accumulate sum (L x - b)^2 and sum b^2 over the
// rows of this block into acc[2c], acc[2c+1].
template <typename T>
__global__ __launch_bounds__(256) void k_residual(
    int n, const int *__restrict__ rp, const int *__restrict__ ci,
    const T *__restrict__ val, const int *__restrict__ perm,
    const T *__restrict__ X, const T *__restrict__ B, int64_t rs, int64_t cs,
    double *acc) {
  __shared__ double s_r[4], s_b[4];
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  double r2 = 0, b2 = 0;
  if (q < n) {
    const int i = perm[q];
    const int64_t coff = (int64_t)c * cs;
    double lx = 0;
    for (int j = rp[q]; j < rp[q + 1]; ++j)
      lx += (double)val[j] * (double)X[(int64_t)ci[j] * rs + coff];
    const double bi = (double)B[(int64_t)i * rs + coff];
    r2 = (lx - bi) * (lx - bi);
    b2 = bi * bi;
  }
  for (int off = 32; off > 0; off >>= 1) {
    r2 += __shfl_xor(r2, off);
    b2 += __shfl_xor(b2, off);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { s_r[w] = r2; s_b[w] = b2; }
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(&acc[2 * c], s_r[0] + s_r[1] + s_r[2] + s_r[3]);
    atomicAdd(&acc[2 * c + 1], s_b[0] + s_b[1] + s_b[2] + s_b[3]);
  }
}

}  // namespace

int launch_spmv(symamd_handle h, int n, const int *d_Ap, const int *d_Ai,
                const double *d_Ax, const double *d_x, const double *d_y,
                double *d_z, double a, double b, int mode) {
  if (n == 0) return SYMAMD_OK;
  const int grid = (n + 255) / 256;
  if (mode == 0)
    k_spmv<<<grid, 256, 0, h->stream>>>(n, d_Ap, d_Ai, d_Ax, d_x, d_z);
  else
    k_spmv_axpby<<<grid, 256, 0, h->stream>>>(n, d_Ap, d_Ai, d_Ax, d_x, d_y,
                                              d_z, a, b);
  SYMAMD_HIP(hipGetLastError());
  return SYMAMD_OK;
}

int plan_residual(symamd_handle h, symamd_plan p, const void *d_X,
                  const void *d_B, int k, int64_t rs, int64_t cs,
                  double *h_rel) {
  if (k <= 0) return SYMAMD_OK;
  if (p->n == 0) {
    for (int c = 0; c < k; ++c) h_rel[c] = 0;
    return SYMAMD_OK;
  }
  if (k > 65535) return fail(SYMAMD_ERR_INVALID_ARG, "residual supports at most 65535 columns");
  double *d_acc = nullptr;
  SYMAMD_HIP(hipMallocAsync(&d_acc, sizeof(double) * 2 * k, h->stream));
  SYMAMD_HIP(hipMemsetAsync(d_acc, 0, sizeof(double) * 2 * k, h->stream));
  dim3 grid((p->n + 255) / 256, k);
  if (p->dtype == SYMAMD_F64)
    k_residual<double><<<grid, 256, 0, h->stream>>>(
        p->n, p->d_rp, p->d_ci, (const double *)p->d_val, p->d_perm,
        (const double *)d_X, (const double *)d_B, rs, cs, d_acc);
  else
    k_residual<float><<<grid, 256, 0, h->stream>>>(
        p->n, p->d_rp, p->d_ci, (const float *)p->d_val, p->d_perm,
        (const float *)d_X, (const float *)d_B, rs, cs, d_acc);
  SYMAMD_HIP(hipGetLastError());
  std::vector<double> acc(2 * k);
  SYMAMD_HIP(hipMemcpyAsync(acc.data(), d_acc, sizeof(double) * 2 * k,
                            hipMemcpyDeviceToHost, h->stream));
  SYMAMD_HIP(hipFreeAsync(d_acc, h->stream));
  SYMAMD_HIP(hipStreamSynchronize(h->stream));
  for (int c = 0; c < k; ++c) {
    const double num = std::sqrt(acc[2 * c]), den = std::sqrt(acc[2 * c + 1]);
    h_rel[c] = den > 0 ? num / den : num;
  }
  return SYMAMD_OK;
}

}  // namespace symamd
