// ll_mlp.hip -- the round-off model: an MLP 3 -> 64 -> 64 -> 1 (FP32, tanh) that predicts
// the log10 of the worst FP64 round-off of a run from (log2 p, log2 N, bits per digit),
// trained and evaluated on the GPU by these kernels.  No reference counterpart: the
// reference sizes its FFT with a fixed bits-per-word cap (lucaslehmer.cpp:244-254).
//
// Parameter layout (floats): W1[64][3], b1[64], W2[64][64], b2[64], w3[64], b3.
// Training is full-batch Adam on the mean squared error:
//   K1 mlp_fwd_bwd   one thread per sample: forward pass, then the per-sample backward
//                    vectors dy, dz2 = dy w3 (1 - h2^2), dz1 = W2^T dz2 (1 - h1^2), stored
//   K2 mlp_grad      one thread per parameter: its gradient summed over the samples in a
//                    fixed order (deterministic), plus the loss
//   K3 mlp_adam      elementwise Adam step
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace llh {

constexpr int kH = 64;
constexpr int kIn = 3;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kH * kIn;
constexpr int kOffW2 = kOffB1 + kH;
constexpr int kOffB2 = kOffW2 + kH * kH;
constexpr int kOffW3 = kOffB2 + kH;
constexpr int kOffB3 = kOffW3 + kH;
constexpr int kNParams = kOffB3 + 1;   // 4481

}  // namespace llh

namespace {

using namespace llh;

// Forward (and, when tgt != nullptr, the per-sample backward vectors) for samples [0, n).
__global__ void __launch_bounds__(128) mlp_fwd_bwd(const float *__restrict__ P, const float *__restrict__ X,
                                                   const float *__restrict__ tgt, int n, float *__restrict__ y_out,
                                                   float *__restrict__ H1, float *__restrict__ H2,
                                                   float *__restrict__ DZ1, float *__restrict__ DZ2,
                                                   float *__restrict__ DY) {
    __shared__ float sp[kNParams];
    for (int i = threadIdx.x; i < kNParams; i += blockDim.x) sp[i] = P[i];
    __syncthreads();
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const float x0 = X[3 * s], x1 = X[3 * s + 1], x2 = X[3 * s + 2];
    float h1[kH], h2[kH];
#pragma unroll 8
    for (int i = 0; i < kH; ++i) {
        const float z = sp[kOffB1 + i] + sp[kOffW1 + 3 * i] * x0 + sp[kOffW1 + 3 * i + 1] * x1 +
                        sp[kOffW1 + 3 * i + 2] * x2;
        h1[i] = tanhf(z);
    }
    float y = sp[kOffB3];
    for (int i = 0; i < kH; ++i) {
        float z = sp[kOffB2 + i];
#pragma unroll 16
        for (int j = 0; j < kH; ++j) z += sp[kOffW2 + kH * i + j] * h1[j];
        h2[i] = tanhf(z);
        y += sp[kOffW3 + i] * h2[i];
    }
    y_out[s] = y;
    if (!tgt) return;
    const float dy = 2.0f * (y - tgt[s]) / (float)n;
    DY[s] = dy;
    float dz2[kH];
#pragma unroll 8
    for (int i = 0; i < kH; ++i) {
        dz2[i] = dy * sp[kOffW3 + i] * (1.0f - h2[i] * h2[i]);
        DZ2[(size_t)s * kH + i] = dz2[i];
        H1[(size_t)s * kH + i] = h1[i];
        H2[(size_t)s * kH + i] = h2[i];
    }
    for (int j = 0; j < kH; ++j) {
        float dh = 0.f;
#pragma unroll 16
        for (int i = 0; i < kH; ++i) dh += sp[kOffW2 + kH * i + j] * dz2[i];
        DZ1[(size_t)s * kH + j] = dh * (1.0f - h1[j] * h1[j]);
    }
}

// Gradient of every parameter, summed over samples in order; thread kNParams writes the loss.
__global__ void __launch_bounds__(256) mlp_grad(const float *__restrict__ X, const float *__restrict__ tgt,
                                                const float *__restrict__ y, int n, const float *__restrict__ H1,
                                                const float *__restrict__ H2, const float *__restrict__ DZ1,
                                                const float *__restrict__ DZ2, const float *__restrict__ DY,
                                                float *__restrict__ G, float *__restrict__ loss) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q > kNParams) return;
    float acc = 0.f;
    if (q == kNParams) {
        for (int s = 0; s < n; ++s) {
            const float d = y[s] - tgt[s];
            acc += d * d;
        }
        *loss = acc / (float)n;
        return;
    }
    if (q < kOffB1) {                       // W1[i][k]
        const int i = q / kIn, k = q % kIn;
        for (int s = 0; s < n; ++s) acc += DZ1[(size_t)s * kH + i] * X[3 * s + k];
    } else if (q < kOffW2) {                // b1[i]
        const int i = q - kOffB1;
        for (int s = 0; s < n; ++s) acc += DZ1[(size_t)s * kH + i];
    } else if (q < kOffB2) {                // W2[i][j]
        const int i = (q - kOffW2) / kH, j = (q - kOffW2) % kH;
        for (int s = 0; s < n; ++s) acc += DZ2[(size_t)s * kH + i] * H1[(size_t)s * kH + j];
    } else if (q < kOffW3) {                // b2[i]
        const int i = q - kOffB2;
        for (int s = 0; s < n; ++s) acc += DZ2[(size_t)s * kH + i];
    } else if (q < kOffB3) {                // w3[i]
        const int i = q - kOffW3;
        for (int s = 0; s < n; ++s) acc += DY[s] * H2[(size_t)s * kH + i];
    } else {                                // b3
        for (int s = 0; s < n; ++s) acc += DY[s];
    }
    G[q] = acc;
}

__global__ void __launch_bounds__(256) mlp_adam(float *__restrict__ P, const float *__restrict__ G,
                                                float *__restrict__ Mm, float *__restrict__ Vv, float lr,
                                                float b1, float b2, float eps, float c1, float c2) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= kNParams) return;
    const float g = G[q];
    const float m = b1 * Mm[q] + (1.f - b1) * g;
    const float v = b2 * Vv[q] + (1.f - b2) * g * g;
    Mm[q] = m;
    Vv[q] = v;
    P[q] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

}  // namespace

namespace llh {

int mlp_n_params() { return kNParams; }

struct MlpBuffers {
    float *P, *X, *T, *Y, *H1, *H2, *DZ1, *DZ2, *DY, *G, *Mm, *Vv, *loss;
};

hipError_t mlp_alloc(MlpBuffers &B, size_t n) {
    hipError_t e;
#define ALLOC(ptr, cnt) \
    if ((e = hipMalloc(&B.ptr, (cnt) * sizeof(float))) != hipSuccess) return e;
    ALLOC(P, kNParams) ALLOC(X, 3 * n + 3) ALLOC(T, n + 1) ALLOC(Y, n + 1) ALLOC(H1, kH * n + kH)
    ALLOC(H2, kH * n + kH) ALLOC(DZ1, kH * n + kH) ALLOC(DZ2, kH * n + kH) ALLOC(DY, n + 1)
    ALLOC(G, kNParams) ALLOC(Mm, kNParams) ALLOC(Vv, kNParams) ALLOC(loss, 1)
#undef ALLOC
    return hipSuccess;
}

void mlp_free(MlpBuffers &B) {
    float *all[] = {B.P, B.X, B.T, B.Y, B.H1, B.H2, B.DZ1, B.DZ2, B.DY, B.G, B.Mm, B.Vv, B.loss};
    for (float *p : all)
        if (p) (void)hipFree(p);
}

hipError_t mlp_forward(const MlpBuffers &B, int n, bool with_backward, hipStream_t st) {
    hipLaunchKernelGGL(mlp_fwd_bwd, dim3((n + 127) / 128), dim3(128), 0, st, B.P, B.X,
                       with_backward ? B.T : nullptr, n, B.Y, B.H1, B.H2, B.DZ1, B.DZ2, B.DY);
    return hipGetLastError();
}

hipError_t mlp_gradient(const MlpBuffers &B, int n, hipStream_t st) {
    hipLaunchKernelGGL(mlp_grad, dim3((kNParams + 1 + 255) / 256), dim3(256), 0, st, B.X, B.T, B.Y, n, B.H1, B.H2,
                       B.DZ1, B.DZ2, B.DY, B.G, B.loss);
    return hipGetLastError();
}

hipError_t mlp_step(const MlpBuffers &B, float lr, int t, hipStream_t st) {
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    hipLaunchKernelGGL(mlp_adam, dim3((kNParams + 255) / 256), dim3(256), 0, st, B.P, B.G, B.Mm, B.Vv, lr, b1, b2,
                       1e-8f, c1, c2);
    return hipGetLastError();
}

}  // namespace llh
