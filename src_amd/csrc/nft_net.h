// Learned preconditioner for the GCR pressure solve: column MLP + horizontal graph
// convolutions on gfx950 MFMA (bf16 operands, FP32 accumulation). No counterpart in the
// reference; the FP64 numpy restatement used to check these kernels is oracle/netref.py.
//
// Activations are FP32 row-major [rows][cols] with rows = (sample, column) and cols =
// features; weights are kept FP32 (master, updated by Adam) plus two bf16 copies: Wt
// [N][K] for forward products and W [K][N] for the data gradient. Every reduction has a
// fixed order (no atomics), so training is reproducible.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nft {

typedef short short4v __attribute__((ext_vector_type(4)));
typedef float float16v __attribute__((ext_vector_type(16)));

// float -> bfloat16, round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// C[m][n] = act(sum_k A[m][k] * B[k][n] + bias[n]) for a 32x32 tile per wavefront, with
// v_mfma_f32_32x32x8_bf16: lane l supplies A[row0 + l%32][k0 + 4(l/32) .. +3] and
// B[k0 + 4(l/32) .. +3][col0 + l%32] (Bt holds B transposed, [N][K], so those four are
// contiguous) and receives C rows row0 + 8(v/4) + 4(l/32) + v%4, v = 0..15, of column
// col0 + l%32. K must be a multiple of 8 (callers pad with zeros).
// mask (optional, [M][ldm]): outputs where mask <= 0 are set to 0 (ReLU derivative).
__global__ void __launch_bounds__(64) k_net_gemm(int M, int N, int K, const float *A, int lda,
                                                 const uint16_t *Bt, const float *bias, int relu,
                                                 const float *mask, int ldm, float *C, int ldc) {
  const int lane = threadIdx.x;
  const int row0 = blockIdx.x * 32, col0 = blockIdx.y * 32;
  const int r = row0 + (lane & 31), cc = col0 + (lane & 31);
  const int kq = 4 * (lane >> 5);
  float16v acc = {};
  for (int k0 = 0; k0 < K; k0 += 8) {
    short4v a, b;
    const int k = k0 + kq;
    if (r < M) {
      const float *ap = A + (size_t)r * lda + k;
      a[0] = (short)f2bf(ap[0]); a[1] = (short)f2bf(ap[1]); a[2] = (short)f2bf(ap[2]); a[3] = (short)f2bf(ap[3]);
    } else {
      a[0] = a[1] = a[2] = a[3] = 0;
    }
    if (cc < N) {
      const uint16_t *bp = Bt + (size_t)cc * K + k;
      b[0] = (short)bp[0]; b[1] = (short)bp[1]; b[2] = (short)bp[2]; b[3] = (short)bp[3];
    } else {
      b[0] = b[1] = b[2] = b[3] = 0;
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
  }
  if (cc >= N) return;
  const float bb = bias ? bias[cc] : 0.0f;
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int row = row0 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
    if (row < M) {
      float x = acc[v] + bb;
      if (relu && x < 0.0f) x = 0.0f;
      if (mask && mask[(size_t)row * ldm + cc] <= 0.0f) x = 0.0f;
      C[(size_t)row * ldc + cc] = x;
    }
  }
}

// The same product with a 64x64 C tile per 256-thread block (four wavefronts, one 32x32
// MFMA tile each): A (converted to bf16) and Bt are staged through LDS 32 k at a time with
// coalesced loads, so each A row is read once per 64 output columns instead of once per
// 32. Each tile's MFMA chain runs over k in the same ascending order as k_net_gemm, so the
// two kernels give identical bits. K must be a multiple of 32.
constexpr int NET_TK = 32;        // k per LDS stage
constexpr int NET_LDS_ROW = 36;   // bf16 per staged row (32 + 4 padding)

__global__ void __launch_bounds__(256) k_net_gemm_t(int M, int N, int K, const float *A, int lda,
                                                    const uint16_t *Bt, const float *bias, int relu,
                                                    const float *mask, int ldm, float *C, int ldc) {
  __shared__ uint16_t As[64 * NET_LDS_ROW];
  __shared__ uint16_t Bs[64 * NET_LDS_ROW];
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const int row0 = blockIdx.x * 64, col0 = blockIdx.y * 64;
  const int wr = (w & 1) * 32, wc = (w >> 1) * 32;
  const int lr = t >> 2, lk = (t & 3) * 8;   // staging: 4 threads per row, 8 k each
  float16v acc = {};
  for (int kc = 0; kc < K; kc += NET_TK) {
    {
      const int gr = row0 + lr;
      uint16_t *dst = As + lr * NET_LDS_ROW + lk;
      if (gr < M) {
        const float4 *src = (const float4 *)(A + (size_t)gr * lda + kc + lk);
        const float4 x0 = src[0], x1 = src[1];
        dst[0] = f2bf(x0.x); dst[1] = f2bf(x0.y); dst[2] = f2bf(x0.z); dst[3] = f2bf(x0.w);
        dst[4] = f2bf(x1.x); dst[5] = f2bf(x1.y); dst[6] = f2bf(x1.z); dst[7] = f2bf(x1.w);
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) dst[j] = 0;
      }
      const int gc = col0 + lr;
      uint16_t *bd = Bs + lr * NET_LDS_ROW + lk;
      if (gc < N) {
        const uint16_t *bsrc = Bt + (size_t)gc * K + kc + lk;
#pragma unroll
        for (int j = 0; j < 8; ++j) bd[j] = bsrc[j];
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) bd[j] = 0;
      }
    }
    __syncthreads();
#pragma unroll
    for (int k8 = 0; k8 < NET_TK; k8 += 8) {
      const int kq = k8 + 4 * (lane >> 5);
      const uint16_t *ap = As + (wr + (lane & 31)) * NET_LDS_ROW + kq;
      const uint16_t *bp = Bs + (wc + (lane & 31)) * NET_LDS_ROW + kq;
      short4v a, b;
      a[0] = (short)ap[0]; a[1] = (short)ap[1]; a[2] = (short)ap[2]; a[3] = (short)ap[3];
      b[0] = (short)bp[0]; b[1] = (short)bp[1]; b[2] = (short)bp[2]; b[3] = (short)bp[3];
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  const int cc = col0 + wc + (lane & 31);
  if (cc >= N) return;
  const float bb = bias ? bias[cc] : 0.0f;
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int row = row0 + wr + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
    if (row < M) {
      float x = acc[v] + bb;
      if (relu && x < 0.0f) x = 0.0f;
      if (mask && mask[(size_t)row * ldm + cc] <= 0.0f) x = 0.0f;
      C[(size_t)row * ldc + cc] = x;
    }
  }
}

// Partial weight gradients over a slice of rows: P[s][k][n] = sum_{m in slice s} A[m][k]*D[m][n]
// (bf16 operands, FP32 accumulation), same MFMA tile layout with the roles k <-> m.
__global__ void __launch_bounds__(64) k_net_wgrad(int M, int K, int N, const float *A, int lda,
                                                  const float *D, int ldd, int rows_per_split,
                                                  float *P) {
  const int lane = threadIdx.x;
  const int k0t = blockIdx.x * 32, n0t = blockIdx.y * 32, s = blockIdx.z;
  const int kk = k0t + (lane & 31), nn = n0t + (lane & 31);
  const int mq = 4 * (lane >> 5);
  const int m_begin = s * rows_per_split;
  const int m_end = min(M, m_begin + rows_per_split);
  float16v acc = {};
  for (int m0 = m_begin; m0 < m_end; m0 += 8) {
    short4v a, b;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int m = m0 + mq + j;
      a[j] = (short)((m < m_end && kk < K) ? f2bf(A[(size_t)m * lda + kk]) : 0);
      b[j] = (short)((m < m_end && nn < N) ? f2bf(D[(size_t)m * ldd + nn]) : 0);
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
  }
  if (nn >= N) return;
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int k = k0t + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
    if (k < K) P[((size_t)s * K + k) * N + nn] = acc[v];
  }
}

// out[i] = sum over splits s (in order) of P[s][i]
__global__ void k_net_reduce_splits(int nsplit, int count, const float *P, float *out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float a = 0.0f;
  for (int s = 0; s < nsplit; ++s) a += P[(size_t)s * count + i];
  out[i] = a;
}

// column sums of D over a slice of rows (bias gradient partials): P[s][n]
__global__ void k_net_colsum(int M, int N, const float *D, int ldd, int rows_per_split, float *P) {
  int n = threadIdx.x + blockIdx.x * blockDim.x;
  int s = blockIdx.y;
  if (n >= N) return;
  int m_begin = s * rows_per_split, m_end = min(M, m_begin + rows_per_split);
  float a = 0.0f;
  for (int m = m_begin; m < m_end; ++m) a += D[(size_t)m * ldd + n];
  P[(size_t)s * N + n] = a;
}

// G[row][0:H] = mean over horizontal neighbours of A[row'][0:H] within the row's sample
// (rows = sample*ncol + column); A and G may be column blocks of one wider buffer.
__global__ void k_net_agg(int nrows, int ncol, int H, const int32_t *nptr, const int32_t *nidx,
                          const float *A, int lda, float *G, int ldg) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int row = t / H, h = t % H;
  if (row >= nrows) return;
  int sample = row / ncol, c = row % ncol;
  int b0 = nptr[c], b1 = nptr[c + 1];
  float a = 0.0f;
  for (int q = b0; q < b1; ++q) a += A[((size_t)sample * ncol + nidx[q]) * lda + h];
  G[(size_t)row * ldg + h] = a / (float)(b1 - b0);
}

// dA[row][h] = (dA_direct[row][h] + sum_{c' in N(c)} dG[c'][h] / deg(c')) * (A[row][h] > 0)
// (no mask when A is null: identity activation)
__global__ void k_net_agg_back(int nrows, int ncol, int H, const int32_t *nptr, const int32_t *nidx,
                               const float *dC, int ldc, const float *A, int lda, float *dA, int ldo) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int row = t / H, h = t % H;
  if (row >= nrows) return;
  int sample = row / ncol, c = row % ncol;
  float g = 0.0f;
  for (int q = nptr[c]; q < nptr[c + 1]; ++q) {
    int cn = nidx[q];
    g += dC[((size_t)sample * ncol + cn) * ldc + H + h] / (float)(nptr[cn + 1] - nptr[cn]);
  }
  float v = dC[(size_t)row * ldc + h] + g;
  if (A && A[(size_t)row * lda + h] <= 0.0f) v = 0.0f;
  dA[(size_t)row * ldo + h] = v;
}

// Adam step on one parameter tensor, then refresh its bf16 copies (W [K][N] and Wt [N][K]
// for matrices; bias has neither).
__global__ void k_net_adam(int count, int K, int N, float *w, const float *g, float *m, float *v, float lr,
                           float b1, float b2, float c1, float c2, float eps, uint16_t *wb, uint16_t *wt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float gi = g[i];
  float mi = b1 * m[i] + (1.0f - b1) * gi;
  float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb) {
    int k = i / N, n = i % N;
    uint16_t h = f2bf(wi);
    wb[i] = h;
    wt[(size_t)n * K + k] = h;
  }
}

// bf16 copies of a weight matrix [K][N] (rows >= krows are zero padding)
__global__ void k_net_to_bf16(int K, int N, const float *w, uint16_t *wb, uint16_t *wt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= K * N) return;
  int k = i / N, n = i % N;
  uint16_t h = f2bf(w[i]);
  wb[i] = h;
  wt[(size_t)n * K + k] = h;
}

// network input of one residual: X[c][k] = r[c*L + k] / s (k < L), zero padding to Kp
__global__ void k_net_input(int ncol, int L, int Kp, const double *r, const double *scal, int sidx, float *X) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncol * Kp) return;
  int c = t / Kp, k = t % Kp;
  double s = scal[sidx];
  X[t] = (k < L && s > 0) ? (float)(r[(size_t)c * L + k] / s) : 0.0f;
}

// z[c*L + k] += s * Y[c][k]  (learned correction on top of the line preconditioner)
__global__ void k_net_output(int ncol, int L, int Kp, const float *Y, const double *scal, int sidx, double *z) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncol * L) return;
  int c = t / L, k = t % L;
  double s = scal[sidx];
  z[t] = z[t] + s * (double)Y[(size_t)c * Kp + k];
}

// loss gradient dY = 2 (Y - T) / count on the first L of Kp entries; per-block partial
// sums of squared error (fixed order) for the loss value
__global__ void k_net_loss(int M, int L, int Kp, const float *Y, const float *T, float inv_count, float *dY,
                           float *part) {
  __shared__ float sh[256];
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  float e2 = 0.0f;
  if (t < M * Kp) {
    int k = t % Kp;
    float d = (k < L) ? Y[t] - T[t] : 0.0f;
    dY[t] = 2.0f * d * inv_count;
    e2 = d * d;
  }
  sh[threadIdx.x] = e2;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// network input of a stored residual with its scale given by value
__global__ void k_net_input_s(int ncol, int L, int Kp, const double *r, double s, float *X) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncol * Kp) return;
  int c = t / Kp, k = t % Kp;
  X[t] = (k < L && s > 0) ? (float)(r[(size_t)c * L + k] / s) : 0.0f;
}

// z = q + s * Y (one sample's preconditioned direction)
__global__ void k_net_z(int ncol, int L, int Kp, const double *q, const float *Y, double s, double *z) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncol * L) return;
  int c = t / L, k = t % L;
  z[t] = q[t] + s * (double)Y[(size_t)c * Kp + k];
}

// per-block partials (fixed tree) of <v,v> and <r,v>: part[2*block + {0,1}]
__global__ void k_net_dots(int n, const double *v, const double *r, double *part) {
  __shared__ double sh[2][256];
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  double vv = 0.0, rv = 0.0;
  if (t < n) {
    vv = v[t] * v[t];
    rv = r[t] * v[t];
  }
  sh[0][threadIdx.x] = vv;
  sh[1][threadIdx.x] = rv;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sh[0][0];
    part[2 * blockIdx.x + 1] = sh[1][0];
  }
}

// y = alpha * v + beta * r
__global__ void k_net_axpby(int n, double alpha, const double *v, double beta, const double *r, double *y) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] = alpha * v[t] + beta * r[t];
}

// dY of one sample = s * (L^T y) in the network's [column][Kp] layout
__global__ void k_net_dY(int ncol, int L, int Kp, const double *g, double s, float *dY) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncol * Kp) return;
  int c = t / Kp, k = t % Kp;
  dY[t] = (k < L) ? (float)(s * g[(size_t)c * L + k]) : 0.0f;
}

}  // namespace nft
