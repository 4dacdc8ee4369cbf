// mlp.hip -- the rank predictor of the learned-sketch rounding: a
// 18 -> 64 -> 64 -> 64 -> 1 perceptron with tanh-GELU activations in FP32,
// trained on the GPU by these kernels (forward, backward, Adam).
//
// Parameter layout (floats, row-major W[in][out]):
//   W1 18x64 | b1 64 | W2 64x64 | b2 64 | W3 64x64 | b3 64 | W4 64 | b4 1
// A workgroup of 256 threads owns 16 samples; weights and the activations
// of its samples live in LDS; each layer is computed as 32 x 64 outputs
// spread over the threads.  Gradients of a workgroup are summed in LDS and
// added to the global gradient with one atomic per parameter.
#include "common.hpp"

#include <cmath>

using namespace ptt;

namespace {

constexpr int NIN = 18, NH = 64, S = 16, NT = 256;  // 16 samples: LDS < 64 KB
constexpr int OW1 = 0, OB1 = OW1 + NIN * NH, OW2 = OB1 + NH, OB2 = OW2 + NH * NH,
              OW3 = OB2 + NH, OB3 = OW3 + NH * NH, OW4 = OB3 + NH, OB4 = OW4 + NH,
              NPARAM = OB4 + 1;

__device__ __forceinline__ float gelu(float x) {
    const float c = 0.7978845608028654f;  // sqrt(2/pi)
    float u = c * (x + 0.044715f * x * x * x);
    return 0.5f * x * (1.0f + tanhf(u));
}

__device__ __forceinline__ float gelu_grad(float x) {
    const float c = 0.7978845608028654f;
    float x2 = x * x;
    float u = c * (x + 0.044715f * x2 * x);
    float t = tanhf(u);
    return 0.5f * (1.0f + t) + 0.5f * x * (1.0f - t * t) * c * (1.0f + 3.0f * 0.044715f * x2);
}

// z[s][j] = b[j] + sum_i a[s][i] W[i][j]
__device__ void dense(const float *a, int nin, const float *W, const float *b, float *z,
                      int ns) {
    for (int e = threadIdx.x; e < ns * NH; e += NT) {
        int s = e / NH, j = e - s * NH;
        float acc = b[j];
        for (int i = 0; i < nin; i++) acc += a[s * nin + i] * W[i * NH + j];
        z[e] = acc;
    }
}

__global__ __launch_bounds__(NT) void k_mlp(int nsamp, const float *__restrict__ X,
                                            const float *__restrict__ Y,
                                            const float *__restrict__ P, float *__restrict__ out,
                                            float *__restrict__ grad, float *__restrict__ loss) {
    __shared__ float w[NPARAM];
    __shared__ float x[S * NIN];
    __shared__ float z1[S * NH], a1[S * NH], z2[S * NH], a2[S * NH], z3[S * NH], a3[S * NH];
    __shared__ float dout[S];
    const int s0 = blockIdx.x * S;
    const int ns = min(S, nsamp - s0);
    for (int e = threadIdx.x; e < NPARAM; e += NT) w[e] = P[e];
    for (int e = threadIdx.x; e < ns * NIN; e += NT) x[e] = X[(long)s0 * NIN + e];
    __syncthreads();
    dense(x, NIN, w + OW1, w + OB1, z1, ns);
    __syncthreads();
    for (int e = threadIdx.x; e < ns * NH; e += NT) a1[e] = gelu(z1[e]);
    __syncthreads();
    dense(a1, NH, w + OW2, w + OB2, z2, ns);
    __syncthreads();
    for (int e = threadIdx.x; e < ns * NH; e += NT) a2[e] = gelu(z2[e]);
    __syncthreads();
    dense(a2, NH, w + OW3, w + OB3, z3, ns);
    __syncthreads();
    for (int e = threadIdx.x; e < ns * NH; e += NT) a3[e] = gelu(z3[e]);
    __syncthreads();
    if (threadIdx.x < ns) {
        int s = threadIdx.x;
        float acc = w[OB4];
        for (int i = 0; i < NH; i++) acc += a3[s * NH + i] * w[OW4 + i];
        if (out) out[s0 + s] = acc;
        if (grad) {
            float diff = acc - Y[s0 + s];
            dout[s] = 2.0f * diff / (float)nsamp;  // d(mean sq err)/d out
            atomicAdd(loss, diff * diff / (float)nsamp);
        }
    }
    if (!grad) return;
    __syncthreads();
    // ---- backward: reuse z/a buffers for deltas, layer by layer ----
    // output layer
    for (int i = threadIdx.x; i < NH; i += NT) {
        float g = 0.0f;
        for (int s = 0; s < ns; s++) g += a3[s * NH + i] * dout[s];
        atomicAdd(&grad[OW4 + i], g);
    }
    if (threadIdx.x == 0) {
        float g = 0.0f;
        for (int s = 0; s < ns; s++) g += dout[s];
        atomicAdd(&grad[OB4], g);
    }
    // delta3 = (W4 dout) * gelu'(z3)  -> stored in z3
    for (int e = threadIdx.x; e < ns * NH; e += NT) {
        int s = e / NH, j = e - s * NH;
        z3[e] = w[OW4 + j] * dout[s] * gelu_grad(z3[e]);
    }
    __syncthreads();
    // layers 3 and 2 (same shape): dW = a_prev^T delta, db = sum delta,
    // delta_prev = (delta W^T) * gelu'(z_prev)
    for (int layer = 3; layer >= 2; layer--) {
        float *delta = (layer == 3) ? z3 : z2;
        float *aprev = (layer == 3) ? a2 : a1;
        float *zprev = (layer == 3) ? z2 : z1;
        const int OW = (layer == 3) ? OW3 : OW2, OB = (layer == 3) ? OB3 : OB2;
        for (int e = threadIdx.x; e < NH * NH; e += NT) {
            int i = e / NH, j = e - i * NH;
            float g = 0.0f;
            for (int s = 0; s < ns; s++) g += aprev[s * NH + i] * delta[s * NH + j];
            atomicAdd(&grad[OW + e], g);
        }
        for (int j = threadIdx.x; j < NH; j += NT) {
            float g = 0.0f;
            for (int s = 0; s < ns; s++) g += delta[s * NH + j];
            atomicAdd(&grad[OB + j], g);
        }
        __syncthreads();  // aprev is overwritten below
        // new delta into aprev's z buffer (zprev), computed from delta and W
        for (int e = threadIdx.x; e < ns * NH; e += NT) {
            int s = e / NH, i = e - s * NH;
            float acc = 0.0f;
            for (int j = 0; j < NH; j++) acc += delta[s * NH + j] * w[OW + i * NH + j];
            // zprev still holds pre-activations; read then overwrite
            aprev[e] = acc * gelu_grad(zprev[e]);
        }
        __syncthreads();
        for (int e = threadIdx.x; e < ns * NH; e += NT) zprev[e] = aprev[e];
        __syncthreads();
    }
    // first layer: delta1 in z1
    for (int e = threadIdx.x; e < NIN * NH; e += NT) {
        int i = e / NH, j = e - i * NH;
        float g = 0.0f;
        for (int s = 0; s < ns; s++) g += x[s * NIN + i] * z1[s * NH + j];
        atomicAdd(&grad[OW1 + e], g);
    }
    for (int j = threadIdx.x; j < NH; j += NT) {
        float g = 0.0f;
        for (int s = 0; s < ns; s++) g += z1[s * NH + j];
        atomicAdd(&grad[OB1 + j], g);
    }
}

__global__ void k_adam(int np, float *__restrict__ P, const float *__restrict__ G,
                       float *__restrict__ m, float *__restrict__ v, float lr, float b1, float b2,
                       float eps, int t) {
    const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < np; i += gridDim.x * blockDim.x) {
        float g = G[i];
        m[i] = b1 * m[i] + (1.0f - b1) * g;
        v[i] = b2 * v[i] + (1.0f - b2) * g * g;
        P[i] -= lr * (m[i] / c1) / (sqrtf(v[i] / c2) + eps);
    }
}

__global__ void k_zero_f(float *p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = 0.0f;
}

}  // namespace

extern "C" {

int ptt_hip_mlp_nparam(void) { return NPARAM; }

void ptt_hip_mlp_forward(int nsamp, const float *X, const float *P, float *out) {
    ptt_hip_init();
    if (nsamp <= 0) return;
    PTT_LAUNCH("mlp_forward", k_mlp, dim3(cdiv(nsamp, S)), dim3(NT), 0, nsamp, X,
               (const float *)nullptr, P, out, (float *)nullptr, (float *)nullptr);
}

void ptt_hip_mlp_grad(int nsamp, const float *X, const float *Y, const float *P, float *grad,
                      float *loss) {
    ptt_hip_init();
    if (ptt_hip_host_visible()) {
        // the gradient is accumulated with device atomics, which host memory
        // reached over the host link does not guarantee
        fprintf(stderr, "ptt_hip_mlp_grad: not available with PARATT_AMD_HOST_VISIBLE\n");
        exit(1);
    }
    PTT_LAUNCH("zero_f", k_zero_f, dim3(cdiv(NPARAM + 1, 256)), dim3(256), 0, grad, NPARAM);
    PTT_LAUNCH("zero_f", k_zero_f, dim3(1), dim3(32), 0, loss, 1);
    if (nsamp <= 0) return;
    PTT_LAUNCH("mlp_backward", k_mlp, dim3(cdiv(nsamp, S)), dim3(NT), 0, nsamp, X, Y, P,
               (float *)nullptr, grad, loss);
}

void ptt_hip_mlp_adam(float *P, const float *grad, float *m, float *v, float lr, int t) {
    ptt_hip_init();
    PTT_LAUNCH("mlp_adam", k_adam, dim3(cdiv(NPARAM, 256)), dim3(256), 0, NPARAM, P, grad, m, v,
               lr, 0.9f, 0.999f, 1e-8f, t);
}

}  // extern "C"
