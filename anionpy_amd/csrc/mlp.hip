// mlp.hip -- fp32 kernels of the launch-configuration cost model: a small
// MLP (features -> 3 x 64 hidden ReLU -> per-knob logits) trained on the
// autotuner's own timings.  No counterpart in the reference.
//
// The layers are tiny (64 wide), so each kernel keeps one output element
// per thread with an in-order fp32 FMA-free loop (fp-contract is off):
// results are deterministic and the CPU float64 oracle checks gradients.
#include "planner.h"

namespace ionp {

__global__ void mlp_linear_fwd(const float *x, const float *w, const float *b, float *y, int n, int in, int out,
                               int relu) {
    const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (idx >= (int64_t)n * out) return;
    const int r = (int)(idx / out), c = (int)(idx % out);
    float acc = b ? b[c] : 0.0f;
    const float *xr = x + (int64_t)r * in;
    for (int k = 0; k < in; k++) acc += xr[k] * w[(int64_t)k * out + c];
    y[idx] = (relu && acc < 0.0f) ? 0.0f : acc;
}

// dz = dy masked by the ReLU (y > 0); dx = dz @ w^T
__global__ void mlp_linear_bwd_dx(const float *w, const float *y, const float *dy, float *dx, int n, int in, int out,
                                  int relu) {
    const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (idx >= (int64_t)n * in) return;
    const int r = (int)(idx / in), k = (int)(idx % in);
    float acc = 0.0f;
    for (int c = 0; c < out; c++) {
        const int64_t o = (int64_t)r * out + c;
        const float dz = (relu && !(y[o] > 0.0f)) ? 0.0f : dy[o];
        acc += dz * w[(int64_t)k * out + c];
    }
    dx[idx] = acc;
}

// dw = x^T @ dz, db = sum_r dz
__global__ void mlp_linear_bwd_dw(const float *x, const float *y, const float *dy, float *dw, float *db, int n, int in,
                                  int out, int relu) {
    const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (idx >= (int64_t)(in + 1) * out) return;
    const int k = (int)(idx / out), c = (int)(idx % out);
    float acc = 0.0f;
    for (int r = 0; r < n; r++) {
        const int64_t o = (int64_t)r * out + c;
        const float dz = (relu && !(y[o] > 0.0f)) ? 0.0f : dy[o];
        acc += (k < in ? x[(int64_t)r * in + k] : 1.0f) * dz;
    }
    if (k < in) dw[(int64_t)k * out + c] = acc;
    else if (db) db[c] = acc;
}

// per-sample softmax cross-entropy over each head's logit group; the loss
// is averaged over samples and summed over heads.
__global__ void mlp_softmax_xent(const float *logits, const int *labels, const int *head_off, int heads, int n,
                                 int total, float *dlogits, float *loss_per_sample) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const float *z = logits + (int64_t)r * total;
    float *dz = dlogits + (int64_t)r * total;
    float l = 0.0f;
    for (int h = 0; h < heads; h++) {
        const int a = head_off[h], b = head_off[h + 1];
        float m = z[a];
        for (int i = a + 1; i < b; i++) m = z[i] > m ? z[i] : m;
        float s = 0.0f;
        for (int i = a; i < b; i++) s += expf(z[i] - m);
        const int lab = labels[(int64_t)r * heads + h];
        for (int i = a; i < b; i++) dz[i] = (expf(z[i] - m) / s - (i - a == lab ? 1.0f : 0.0f)) / (float)n;
        l += -(z[a + lab] - m - logf(s));
    }
    loss_per_sample[r] = l;
}

// the same loss against soft targets q (n x total, non-negative; a head's
// targets may sum to any mass m_h, which weights that sample's head):
// loss = -sum_h sum_i q_i log p_i, dz = (m_h p - q) / n.  One-hot q gives
// mlp_softmax_xent.
__global__ void mlp_softmax_xent_soft(const float *logits, const float *q, const int *head_off, int heads, int n,
                                      int total, float *dlogits, float *loss_per_sample) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const float *z = logits + (int64_t)r * total;
    const float *qr = q + (int64_t)r * total;
    float *dz = dlogits + (int64_t)r * total;
    float l = 0.0f;
    for (int h = 0; h < heads; h++) {
        const int a = head_off[h], b = head_off[h + 1];
        float m = z[a];
        for (int i = a + 1; i < b; i++) m = z[i] > m ? z[i] : m;
        float s = 0.0f;
        for (int i = a; i < b; i++) s += expf(z[i] - m);
        const float ls = logf(s);
        float mass = 0.0f;
        for (int i = a; i < b; i++) mass += qr[i];
        for (int i = a; i < b; i++) {
            dz[i] = (mass * (expf(z[i] - m) / s) - qr[i]) / (float)n;
            l -= qr[i] * (z[i] - m - ls);
        }
    }
    loss_per_sample[r] = l;
}

__global__ void sum_reduce_f32(const float *v, int n, float *out) {
    __shared__ float red[256];
    float acc = 0.0f;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += v[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0] / (float)n;
}

__global__ void mlp_adam(float *p, const float *g, float *m, float *v, int64_t count, float lr, float b1, float b2,
                         float eps, float c1, float c2) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

static ionp_status done() {
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? IONP_OK : set_error(IONP_ERR_DEVICE, hipGetErrorString(e));
}

static unsigned blocks(int64_t n) { return (unsigned)((n + 255) / 256 > 0 ? (n + 255) / 256 : 1); }

}  // namespace ionp

using namespace ionp;

extern "C" {

ionp_status ionp_mlp_linear_fwd(const float *x, const float *w, const float *b, float *y, int32_t n, int32_t in,
                                int32_t out, int32_t relu, void *stream) {
    if (n < 0 || in <= 0 || out <= 0) return set_error(IONP_ERR_VALUE, "bad layer shape");
    hipLaunchKernelGGL(mlp_linear_fwd, dim3(blocks((int64_t)n * out)), dim3(256), 0, (hipStream_t)stream, x, w, b, y,
                       n, in, out, relu);
    return done();
}

ionp_status ionp_mlp_linear_bwd(const float *x, const float *w, const float *y, const float *dy, float *dx, float *dw,
                                float *db, int32_t n, int32_t in, int32_t out, int32_t relu, void *stream) {
    if (n <= 0 || in <= 0 || out <= 0) return set_error(IONP_ERR_VALUE, "bad layer shape");
    hipStream_t s = (hipStream_t)stream;
    if (dx) hipLaunchKernelGGL(mlp_linear_bwd_dx, dim3(blocks((int64_t)n * in)), dim3(256), 0, s, w, y, dy, dx, n, in, out, relu);
    hipLaunchKernelGGL(mlp_linear_bwd_dw, dim3(blocks((int64_t)(in + 1) * out)), dim3(256), 0, s, x, y, dy, dw, db, n, in,
                       out, relu);
    return done();
}

ionp_status ionp_mlp_softmax_xent(const float *logits, const int32_t *labels, const int32_t *head_off, int32_t heads,
                                  int32_t n, int32_t total, float *dlogits, float *loss, void *stream) {
    if (n <= 0 || heads <= 0) return set_error(IONP_ERR_VALUE, "bad head layout");
    hipStream_t s = (hipStream_t)stream;
    // per-sample losses live after the scalar slot: loss[1 .. n]
    hipLaunchKernelGGL(mlp_softmax_xent, dim3(blocks(n)), dim3(256), 0, s, logits, labels, head_off, heads, n, total,
                       dlogits, loss + 1);
    hipLaunchKernelGGL(sum_reduce_f32, dim3(1), dim3(256), 0, s, loss + 1, n, loss);
    return done();
}

ionp_status ionp_mlp_softmax_xent_soft(const float *logits, const float *targets, const int32_t *head_off,
                                       int32_t heads, int32_t n, int32_t total, float *dlogits, float *loss,
                                       void *stream) {
    if (n <= 0 || heads <= 0) return set_error(IONP_ERR_VALUE, "bad head layout");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(mlp_softmax_xent_soft, dim3(blocks(n)), dim3(256), 0, s, logits, targets, head_off, heads, n,
                       total, dlogits, loss + 1);
    hipLaunchKernelGGL(sum_reduce_f32, dim3(1), dim3(256), 0, s, loss + 1, n, loss);
    return done();
}

ionp_status ionp_mlp_adam(float *param, const float *grad, float *m, float *v, int64_t count, float lr, float beta1,
                          float beta2, float eps, int32_t step, void *stream) {
    if (count < 0 || step < 1) return set_error(IONP_ERR_VALUE, "bad Adam arguments");
    const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
    hipLaunchKernelGGL(mlp_adam, dim3(blocks(count)), dim3(256), 0, (hipStream_t)stream, param, grad, m, v, count, lr,
                       beta1, beta2, eps, c1, c2);
    return done();
}

}  // extern "C"
