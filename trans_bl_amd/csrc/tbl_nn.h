// BF16 U-Net of the learned initial guess: kernel launchers and the network
// object behind the tbl_net handle.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace tbl {
namespace nn {

typedef uint16_t bf16;   // bfloat16 bit pattern

void gemm(int M, int N, int K, const bf16 *A, long sam, long sak, const bf16 *B, long sbk,
          long sbn, float *C, long ldc, float alpha, float beta, hipStream_t st, int splitk = 1);
void im2col(int n, int h, int w, int c, const bf16 *x, bf16 *col, hipStream_t st);
void col2im(int n, int h, int w, int c, const float *dcol, float *dx, int acc, hipStream_t st);
void bias_act(long P, int c, float *z, const float *b, bf16 *y, int relu, hipStream_t st);
void relu_back(long n, const float *dy, const bf16 *y, bf16 *dz, int relu, hipStream_t st);
void bias_grad(long P, int c, const bf16 *dz, float *db, hipStream_t st);
void pool(int n, int h, int w, int c, const bf16 *x, bf16 *y, hipStream_t st);
void pool_back(int n, int h, int w, int c, const float *dy, float *dx, int acc, hipStream_t st);
void up_concat(int n, int h, int w, int c1, int c2, const bf16 *a, const bf16 *b, bf16 *y,
               hipStream_t st);
void up_concat_back(int n, int h, int w, int c1, int c2, const float *dy, float *da, float *db,
                    int acc_b, hipStream_t st);
void to_bf16(long n, const float *a, bf16 *b, hipStream_t st);
void mse(long n, const float *y, const float *t, float *grad, float *loss, hipStream_t st);
void adam(long n, float *w, bf16 *wb, const float *g, float *m, float *v, float lr, float b1,
          float b2, float eps, float c1, float c2, hipStream_t st);
void zero(long n, float *a, hipStream_t st);
// direct 3x3 convolution + bias (+ ReLU) -> bf16 NHWC; weights packed by
// pack_conv3x3 into [tap][cout][cin rounded up to 32]; cout a multiple of 32
int conv3x3_cinp(int cin);
void pack_conv3x3(int cin, int cout, const bf16 *w, bf16 *wp, hipStream_t st);
void conv3x3(int n, int h, int w, int cin, int cout, const bf16 *x, const bf16 *wp,
             const float *bias, bf16 *y, int relu, hipStream_t st);

}  // namespace nn
}  // namespace tbl
