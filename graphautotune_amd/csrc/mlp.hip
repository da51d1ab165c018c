// Cost model for the autotuner: a 3-layer MLP (in -> H -> H -> 1, ReLU),
// trained and evaluated on the GPU with hand-written kernels.
//
// Every matrix product runs on gfx950 BF16 matrix cores
// (v_mfma_f32_16x16x32_bf16): operands are rounded to bfloat16
// (round-to-nearest-even) as they are loaded, products accumulate in FP32.
// Master weights and Adam state stay FP32. One wavefront owns one 16x16
// output tile. Operand layout of the 16x16x32 instruction:
//   A[i][k]: lane i + 16*(k/8), element k%8
//   B[k][j]: lane j + 16*(k/8), element k%8
//   C[i][j]: lane j + 16*(i/4), element i%4
// The model only ranks kernel configs for the tuner (it replaces the
// reference's XGBoost regressor, tuning/tuning/role/modeling.py:16-162).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

namespace gatm {

__device__ __forceinline__ __bf16 to_bf16(float f) {
    unsigned u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) u = 0x7fc00000u;  // NaN stays NaN
    else u += 0x7fffu + ((u >> 16) & 1u);                 // round to nearest even
    const unsigned short h = (unsigned short)(u >> 16);
    __bf16 r;
    __builtin_memcpy(&r, &h, 2);
    return r;
}

// C[m][n] = sum_k A(m,k) B(k,n) (+ bias[n]) with optional ReLU; element
// (r,c) of X lives at X[r*s0 + c*s1], so transposes are stride swaps.
// Bounds are checked on every load and store: no padding needed.
__global__ void __launch_bounds__(256) gemm_bf16(const float *A, long long sam, long long sak, const float *B,
                                                 long long sbk, long long sbn, float *C, long long scm,
                                                 long long scn, int M, int N, int K, const float *bias, int relu,
                                                 float *Cpre) {
    const int lane = threadIdx.x % 64;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const int tiles_n = (N + 15) / 16;
    const int tm = wave / tiles_n, tn = wave % tiles_n;
    if (tm * 16 >= M) return;  // whole wave exits together
    const int i = tm * 16 + lane % 16;  // A row of this lane
    const int j = tn * 16 + lane % 16;  // B column of this lane
    const int kq = (lane / 16) * 8;
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 32) {
        bf16x8 a, b;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int k = k0 + kq + e;
            const float av = (i < M && k < K) ? A[i * sam + k * sak] : 0.f;
            const float bv = (j < N && k < K) ? B[k * sbk + j * sbn] : 0.f;
            a[e] = to_bf16(av);
            b[e] = to_bf16(bv);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const int r = tm * 16 + (lane / 16) * 4 + e;
        const int c = tn * 16 + lane % 16;
        if (r < M && c < N) {
            float v = acc[e] + (bias ? bias[c] : 0.f);
            if (Cpre) Cpre[r * scm + c * scn] = v;
            if (relu) v = v > 0.f ? v : 0.f;
            C[r * scm + c * scn] = v;
        }
    }
}

__global__ void mask_relu(float *d, const float *pre, long long n) {  // d *= (pre > 0)
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
        if (!(pre[t] > 0.f)) d[t] = 0.f;
}

__global__ void col_sum(const float *d, int rows, int cols, float *out) {  // out[c] = sum_r d[r][c]
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < cols; c += gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int r = 0; r < rows; r++) s += d[(long long)r * cols + c];
        out[c] = s;
    }
}

// dy = 2 (y - t) / B, loss partials
__global__ void mse_grad(const float *y, const float *t, int B, float *dy, float *loss) {
    __shared__ float part[256];
    float s = 0.f;
    for (int r = threadIdx.x; r < B; r += blockDim.x) {
        const float e = y[r] - t[r];
        dy[r] = 2.f * e / B;
        s += e * e;
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = part[0] / B;
}

__global__ void adam(float *p, const float *g, float *m, float *v, long long n, float lr, float b1, float b2,
                     float eps, float c1, float c2) {
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
        const float gt = g[t];
        m[t] = b1 * m[t] + (1.f - b1) * gt;
        v[t] = b2 * v[t] + (1.f - b2) * gt * gt;
        p[t] -= lr * (m[t] / c1) / (sqrtf(v[t] / c2) + eps);
    }
}

}  // namespace gatm

struct gat_mlp {
    int F = 0, H = 0;
    long long P = 0;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    int cap = 0;  // batch capacity of the activation buffers
    float *x = nullptr, *t = nullptr, *z1 = nullptr, *a1 = nullptr, *z2 = nullptr, *a2 = nullptr, *y = nullptr;
    float *dy = nullptr, *d2 = nullptr, *d1 = nullptr, *loss = nullptr;
    long long step = 0;
    hipStream_t s = nullptr;
    // parameter views
    float *W1() { return p; }
    float *b1() { return p + (long long)F * H; }
    float *W2() { return b1() + H; }
    float *b2() { return W2() + (long long)H * H; }
    float *W3() { return b2() + H; }
    float *b3() { return W3() + H; }
    float *gW1() { return g; }
    float *gb1() { return g + (long long)F * H; }
    float *gW2() { return gb1() + H; }
    float *gb2() { return gW2() + (long long)H * H; }
    float *gW3() { return gb2() + H; }
    float *gb3() { return gW3() + H; }
};

using namespace gat;

#define MCHECK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess) GAT_FAIL(GAT_ENODEV, "%s failed: %s", #x, hipGetErrorString(e_));     \
    } while (0)

static int gemm(hipStream_t s, const float *A, long long sam, long long sak, const float *B, long long sbk,
                long long sbn, float *C, long long scm, long long scn, int M, int N, int K, const float *bias,
                int relu, float *Cpre) {
    const long long tiles = (long long)((M + 15) / 16) * ((N + 15) / 16);
    const int blocks = (int)((tiles + 3) / 4);
    hipLaunchKernelGGL(gatm::gemm_bf16, dim3(blocks), dim3(256), 0, s, A, sam, sak, B, sbk, sbn, C, scm, scn, M, N,
                       K, bias, relu, Cpre);
    MCHECK(hipGetLastError());
    return GAT_OK;
}

static int ensure_batch(gat_mlp *n, int B) {
    if (B <= n->cap) return GAT_OK;
    for (float **q : {&n->x, &n->t, &n->z1, &n->a1, &n->z2, &n->a2, &n->y, &n->dy, &n->d2, &n->d1})
        if (*q) (void)hipFree(*q), *q = nullptr;
    const int c = ((B + 63) / 64) * 64;
    MCHECK(hipMalloc(&n->x, sizeof(float) * (size_t)c * n->F));
    MCHECK(hipMalloc(&n->t, sizeof(float) * c));
    for (float **q : {&n->z1, &n->a1, &n->z2, &n->a2, &n->d2, &n->d1})
        MCHECK(hipMalloc(q, sizeof(float) * (size_t)c * n->H));
    MCHECK(hipMalloc(&n->y, sizeof(float) * c));
    MCHECK(hipMalloc(&n->dy, sizeof(float) * c));
    n->cap = c;
    return GAT_OK;
}

static int forward(gat_mlp *n, int B) {
    const int F = n->F, H = n->H;
    int rc = gemm(n->s, n->x, F, 1, n->W1(), H, 1, n->a1, H, 1, B, H, F, n->b1(), 1, n->z1);
    if (!rc) rc = gemm(n->s, n->a1, H, 1, n->W2(), H, 1, n->a2, H, 1, B, H, H, n->b2(), 1, n->z2);
    if (!rc) rc = gemm(n->s, n->a2, H, 1, n->W3(), 1, 1, n->y, 1, 1, B, 1, H, n->b3(), 0, nullptr);
    return rc;
}

static int backward(gat_mlp *n, int B) {
    const int F = n->F, H = n->H;
    hipStream_t s = n->s;
    hipLaunchKernelGGL(gatm::mse_grad, dim3(1), dim3(256), 0, s, n->y, n->t, B, n->dy, n->loss);
    // layer 3: dW3 = A2^T dy, db3 = sum dy, dA2 = dy W3^T
    int rc = gemm(s, n->a2, 1, H, n->dy, 1, 1, n->gW3(), 1, 1, H, 1, B, nullptr, 0, nullptr);
    hipLaunchKernelGGL(gatm::col_sum, dim3(1), dim3(64), 0, s, n->dy, B, 1, n->gb3());
    if (!rc) rc = gemm(s, n->dy, 1, 1, n->W3(), 1, 1, n->d2, H, 1, B, H, 1, nullptr, 0, nullptr);
    hipLaunchKernelGGL(gatm::mask_relu, dim3(256), dim3(256), 0, s, n->d2, n->z2, (long long)B * H);
    // layer 2: dW2 = A1^T dZ2, db2, dA1 = dZ2 W2^T
    if (!rc) rc = gemm(s, n->a1, 1, H, n->d2, H, 1, n->gW2(), H, 1, H, H, B, nullptr, 0, nullptr);
    hipLaunchKernelGGL(gatm::col_sum, dim3((H + 63) / 64), dim3(64), 0, s, n->d2, B, H, n->gb2());
    if (!rc) rc = gemm(s, n->d2, H, 1, n->W2(), 1, H, n->d1, H, 1, B, H, H, nullptr, 0, nullptr);
    hipLaunchKernelGGL(gatm::mask_relu, dim3(256), dim3(256), 0, s, n->d1, n->z1, (long long)B * H);
    // layer 1: dW1 = X^T dZ1, db1
    if (!rc) rc = gemm(s, n->x, 1, F, n->d1, H, 1, n->gW1(), H, 1, F, H, B, nullptr, 0, nullptr);
    hipLaunchKernelGGL(gatm::col_sum, dim3((H + 63) / 64), dim3(64), 0, s, n->d1, B, H, n->gb1());
    MCHECK(hipGetLastError());
    return rc;
}

static int upload(gat_mlp *n, const float *x, const float *t, int B) {
    int rc = ensure_batch(n, B);
    if (rc) return rc;
    MCHECK(hipMemcpyAsync(n->x, x, sizeof(float) * (size_t)B * n->F, hipMemcpyHostToDevice, n->s));
    if (t) MCHECK(hipMemcpyAsync(n->t, t, sizeof(float) * B, hipMemcpyHostToDevice, n->s));
    return GAT_OK;
}

extern "C" {

int gat_mlp_create(int32_t in_dim, int32_t hidden, uint64_t seed, gat_mlp **out) {
    if (!out || in_dim < 1 || hidden < 1 || in_dim > 4096 || hidden > 4096) GAT_FAIL(GAT_EINVAL, "bad MLP shape");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) GAT_FAIL(GAT_ENODEV, "no HIP device visible");
    auto *n = new gat_mlp();
    n->F = in_dim;
    n->H = hidden;
    n->P = (long long)in_dim * hidden + hidden + (long long)hidden * hidden + hidden + hidden + 1;
    // He-uniform init from a seeded stream
    std::vector<float> h(n->P, 0.f);
    Rng r(seed);
    auto fill = [&](long long off, long long cnt, int fan_in) {
        const double lim = std::sqrt(6.0 / fan_in);
        for (long long k = 0; k < cnt; k++) h[off + k] = (float)((r.uniform() * 2 - 1) * lim);
    };
    fill(0, (long long)in_dim * hidden, in_dim);
    fill((long long)in_dim * hidden + hidden, (long long)hidden * hidden, hidden);
    fill((long long)in_dim * hidden + hidden + (long long)hidden * hidden + hidden, hidden, hidden);
    hipError_t e = hipStreamCreateWithFlags(&n->s, hipStreamNonBlocking);
    for (float **q : {&n->p, &n->g, &n->m, &n->v})
        if (e == hipSuccess) e = hipMalloc(q, sizeof(float) * n->P);
    if (e == hipSuccess) e = hipMalloc(&n->loss, sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(n->p, h.data(), sizeof(float) * n->P, hipMemcpyHostToDevice);
    std::vector<float> z(n->P, 0.f);
    if (e == hipSuccess) e = hipMemcpy(n->m, z.data(), sizeof(float) * n->P, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(n->v, z.data(), sizeof(float) * n->P, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        set_error("MLP allocation failed: %s", hipGetErrorString(e));
        delete n;
        return GAT_ENODEV;
    }
    *out = n;
    return GAT_OK;
}

int gat_mlp_param_count(const gat_mlp *n, int64_t *out) {
    if (!n || !out) GAT_FAIL(GAT_EINVAL, "null argument");
    *out = n->P;
    return GAT_OK;
}

int gat_mlp_get_params(gat_mlp *n, float *out, int64_t count) {
    if (!n || !out || count != n->P) GAT_FAIL(GAT_EINVAL, "parameter buffer must hold %lld floats", n ? n->P : 0);
    MCHECK(hipMemcpy(out, n->p, sizeof(float) * n->P, hipMemcpyDeviceToHost));
    return GAT_OK;
}

int gat_mlp_set_params(gat_mlp *n, const float *in, int64_t count) {
    if (!n || !in || count != n->P) GAT_FAIL(GAT_EINVAL, "parameter buffer must hold %lld floats", n ? n->P : 0);
    MCHECK(hipMemcpy(n->p, in, sizeof(float) * n->P, hipMemcpyHostToDevice));
    return GAT_OK;
}

int gat_mlp_forward(gat_mlp *n, const float *x, int32_t batch, float *y_out) {
    if (!n || !x || !y_out || batch < 1) GAT_FAIL(GAT_EINVAL, "bad forward arguments");
    int rc = upload(n, x, nullptr, batch);
    if (!rc) rc = forward(n, batch);
    if (rc) return rc;
    MCHECK(hipMemcpyAsync(y_out, n->y, sizeof(float) * batch, hipMemcpyDeviceToHost, n->s));
    MCHECK(hipStreamSynchronize(n->s));
    return GAT_OK;
}

int gat_mlp_gradients(gat_mlp *n, const float *x, const float *t, int32_t batch, float *grads_out, float *loss_out) {
    if (!n || !x || !t || batch < 1) GAT_FAIL(GAT_EINVAL, "bad gradient arguments");
    int rc = upload(n, x, t, batch);
    if (!rc) rc = forward(n, batch);
    if (!rc) rc = backward(n, batch);
    if (rc) return rc;
    if (grads_out) MCHECK(hipMemcpyAsync(grads_out, n->g, sizeof(float) * n->P, hipMemcpyDeviceToHost, n->s));
    if (loss_out) MCHECK(hipMemcpyAsync(loss_out, n->loss, sizeof(float), hipMemcpyDeviceToHost, n->s));
    MCHECK(hipStreamSynchronize(n->s));
    return GAT_OK;
}

int gat_mlp_train(gat_mlp *n, const float *x, const float *t, int32_t batch, int32_t steps, float lr,
                  float *loss_out) {
    if (!n || !x || !t || batch < 1 || steps < 1 || !(lr > 0)) GAT_FAIL(GAT_EINVAL, "bad training arguments");
    int rc = upload(n, x, t, batch);
    if (rc) return rc;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int k = 0; k < steps; k++) {
        rc = forward(n, batch);
        if (!rc) rc = backward(n, batch);
        if (rc) return rc;
        n->step++;
        const float c1 = 1.f - std::pow(b1, (float)n->step), c2 = 1.f - std::pow(b2, (float)n->step);
        hipLaunchKernelGGL(gatm::adam, dim3(256), dim3(256), 0, n->s, n->p, n->g, n->m, n->v, n->P, lr, b1, b2, eps,
                           c1, c2);
        MCHECK(hipGetLastError());
    }
    if (loss_out) MCHECK(hipMemcpyAsync(loss_out, n->loss, sizeof(float), hipMemcpyDeviceToHost, n->s));
    MCHECK(hipStreamSynchronize(n->s));
    return GAT_OK;
}

void gat_mlp_free(gat_mlp *n) {
    if (!n) return;
    for (float *q : {n->p, n->g, n->m, n->v, n->loss, n->x, n->t, n->z1, n->a1, n->z2, n->a2, n->y, n->dy, n->d2, n->d1})
        if (q) (void)hipFree(q);
    if (n->s) (void)hipStreamDestroy(n->s);
    delete n;
}

}  // extern "C"
