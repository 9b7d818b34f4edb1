// Training and standalone inference of the learned rate-and-state Newton
// guess (network and numerics in eqd_mlp.h), and the C entry points
// eqd_mlp_forward / eqd_mlp_grad / eqd_mlp_train of include/eqdyna_amd.h.
//
// One training step over a batch: mlp_grad_kernel (one wave per 64 samples:
// forward, mean-squared-error loss, backward; per-workgroup partial gradients
// written once, no atomics), mlp_adam_kernel (sums the partials in workgroup
// order, so training is deterministic, then Adam on float32 master weights)
// and mlp_prepare_kernel (bfloat16 copies and transposes for the next
// forward pass).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "../../include/eqdyna_amd.h"

#include "eqd_mlp.h"

namespace eqd {

// Device buffers of one network: float32 master parameters and the bfloat16
// operands derived from them.
struct MlpBuffers {
    float *p = nullptr;       // [kMlpNParams]
    __bf16 *wb = nullptr;     // w1 [64][16], w2, w3, w2t, w3t [64][64]
    float *w4r = nullptr;     // [64]
    MlpNet net() const
    {
        MlpNet M;
        M.w1 = wb;
        M.w2 = wb + 64 * 16;
        M.w3 = M.w2 + 4096;
        M.w2t = M.w3 + 4096;
        M.w3t = M.w2t + 4096;
        M.b1 = p + kOffB1;
        M.b2 = p + kOffB2;
        M.b3 = p + kOffB3;
        M.w4r = w4r;
        M.b4 = p + kOffB4;
        M.norm = p + kOffNorm;
        return M;
    }
};

__global__ void __launch_bounds__(256) mlp_prepare_kernel(const float *p, __bf16 *wb, float *w4r)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    __bf16 *w1 = wb, *w2 = wb + 64 * 16, *w3 = w2 + 4096, *w2t = w3 + 4096, *w3t = w2t + 4096;
    if (j < 64 * 16) {
        const int o = j / 16, k = j % 16;
        w1[j] = to_bf16(k < kMlpIn ? p[kOffW1 + o * kMlpIn + k] : 0.0f);
    }
    if (j < 4096) {
        const int o = j / 64, i = j % 64;
        const __bf16 a = to_bf16(p[kOffW2 + j]), b = to_bf16(p[kOffW3 + j]);
        w2[j] = a;
        w3[j] = b;
        w2t[i * 64 + o] = a;
        w3t[i * 64 + o] = b;
    }
    if (j < 64) w4r[j] = from_bf16(to_bf16(p[kOffW4 + j]));
}

// Standalone inference: one wave per 64 samples, X[n][6] raw inputs, out = r.
__global__ void __launch_bounds__(64) mlp_forward_kernel(MlpNet M, const float *X, int64_t n, float *out)
{
    __shared__ __attribute__((aligned(16))) __bf16 buf0[64 * kLd0];
    __shared__ __attribute__((aligned(16))) __bf16 buf[64 * kLd];
    const int64_t s = (int64_t)blockIdx.x * 64 + threadIdx.x;
    float x[kMlpIn];
    for (int k = 0; k < kMlpIn; k++) x[k] = s < n ? X[s * kMlpIn + k] : M.norm[k];
    const float y = mlp_wave_forward(M, x, buf0, buf);
    if (s < n) out[s] = y * M.norm[2 * kMlpIn + 1] + M.norm[2 * kMlpIn];
}

// Sample of batch position j: an affine permutation of 0..n-1 (mult coprime
// to n; n, mult, add < 2^31 keep the product in 64 bits).
__device__ __forceinline__ int64_t perm(int64_t j, int64_t n, uint64_t mult, uint64_t add)
{
    return (int64_t)(((uint64_t)j * mult + add) % (uint64_t)n);
}

// Forward + backward of one wave over batch positions [b0 + 64*blockIdx.x, +64)
// of a batch of nb samples; writes this workgroup's partial gradient
// (kMlpNTrain floats) and partial squared-error sum.
__global__ void __launch_bounds__(64) mlp_grad_kernel(MlpNet M, const float *X, const float *logv, int64_t n,
                                                      int64_t b0, int64_t nb, uint64_t mult, uint64_t add,
                                                      float *partial, float *partial_loss)
{
    __shared__ __attribute__((aligned(16))) __bf16 a0[64 * kLd0];
    __shared__ __attribute__((aligned(16))) __bf16 a1[64 * kLd], a2[64 * kLd], a3[64 * kLd];
    __shared__ __attribute__((aligned(16))) __bf16 a1t[64 * kLd], a2t[64 * kLd];
    __shared__ __attribute__((aligned(16))) __bf16 dz[64 * kLd], dzt[64 * kLd];
    __shared__ float g1[64 * 64], g2[64 * 64], g3[64 * 64];  // gelu'(z), later dz in float
    __shared__ float dyv[64], red[64];
    const int l = threadIdx.x;
    const int64_t jb = (int64_t)blockIdx.x * 64 + l;  // position within the batch
    const bool valid = jb < nb;
    const int64_t s = valid ? perm(b0 + jb, n, mult, add) : 0;
    float x[kMlpIn];
    for (int k = 0; k < kMlpIn; k++) x[k] = valid ? X[s * kMlpIn + k] : M.norm[k];
    write_input_row(M, x, a0);
    __syncthreads();
    f32x16 acc[2][2];
    // forward, keeping activations (and their transposes) and gelu'(z)
    zero_acc(acc);
    wave_mm(a0, kLd0, M.w1, kMlpK1, kMlpK1, acc);
    epilogue_gelu(acc, M.b1, a1, g1);
    __syncthreads();
    for (int i = 0; i < 64; i++) a1t[i * kLd + l] = a1[l * kLd + i];
    zero_acc(acc);
    wave_mm(a1, kLd, M.w2, 64, 64, acc);
    epilogue_gelu(acc, M.b2, a2, g2);
    __syncthreads();
    for (int i = 0; i < 64; i++) a2t[i * kLd + l] = a2[l * kLd + i];
    zero_acc(acc);
    wave_mm(a2, kLd, M.w3, 64, 64, acc);
    epilogue_gelu(acc, M.b3, a3, g3);
    __syncthreads();
    float y = M.b4[0];
    for (int i = 0; i < 64; i++) y = y + from_bf16(a3[l * kLd + i]) * M.w4r[i];
    const float t = valid ? (logv[s] - M.norm[2 * kMlpIn]) / M.norm[2 * kMlpIn + 1] : y;
    const float r = y - t;
    const float dy = valid ? 2.0f * r / (float)nb : 0.0f;
    dyv[l] = dy;
    red[l] = valid ? r * r : 0.0f;
    __syncthreads();
    float *P = partial + (int64_t)blockIdx.x * kMlpNTrain;
    if (l == 0) {
        float sl = 0.0f, sb = 0.0f;
        for (int m = 0; m < 64; m++) {
            sl = sl + red[m];
            sb = sb + dyv[m];
        }
        partial_loss[blockIdx.x] = sl;
        P[kOffB4] = sb;
    }
    {
        // output layer: lane l owns column l
        float gw = 0.0f;
        for (int m = 0; m < 64; m++) gw = gw + dyv[m] * from_bf16(a3[m * kLd + l]);
        P[kOffW4 + l] = gw;
    }
    // dz3 = dy * w4 * gelu'(z3): row l by lane l
    for (int o = 0; o < 64; o++) {
        const float v = dy * M.w4r[o] * g3[l * 64 + o];
        g3[l * 64 + o] = v;
        const __bf16 b = to_bf16(v);
        dz[l * kLd + o] = b;
        dzt[o * kLd + l] = b;
    }
    __syncthreads();
    {
        float sb = 0.0f;
        for (int m = 0; m < 64; m++) sb = sb + g3[m * 64 + l];
        P[kOffB3 + l] = sb;
    }
    // grad W3 = dz3^T a2
    zero_acc(acc);
    wave_mm(dzt, kLd, a2t, kLd, 64, acc);
    for (int mt = 0; mt < 2; mt++)
        for (int nt = 0; nt < 2; nt++)
            for (int i = 0; i < 16; i++) P[kOffW3 + acc_row(mt, i) * 64 + acc_col(nt)] = acc[mt][nt][i];
    // dz2 = (dz3 W3) * gelu'(z2)
    zero_acc(acc);
    wave_mm(dz, kLd, M.w3t, 64, 64, acc);
    for (int mt = 0; mt < 2; mt++)
        for (int nt = 0; nt < 2; nt++)
            for (int i = 0; i < 16; i++) {
                const int m = acc_row(mt, i), c = acc_col(nt);
                g2[m * 64 + c] = acc[mt][nt][i] * g2[m * 64 + c];
            }
    __syncthreads();
    for (int o = 0; o < 64; o++) {
        const __bf16 b = to_bf16(g2[l * 64 + o]);
        dz[l * kLd + o] = b;
        dzt[o * kLd + l] = b;
    }
    {
        float sb = 0.0f;
        for (int m = 0; m < 64; m++) sb = sb + g2[m * 64 + l];
        P[kOffB2 + l] = sb;
    }
    __syncthreads();
    // grad W2 = dz2^T a1
    zero_acc(acc);
    wave_mm(dzt, kLd, a1t, kLd, 64, acc);
    for (int mt = 0; mt < 2; mt++)
        for (int nt = 0; nt < 2; nt++)
            for (int i = 0; i < 16; i++) P[kOffW2 + acc_row(mt, i) * 64 + acc_col(nt)] = acc[mt][nt][i];
    // dz1 = (dz2 W2) * gelu'(z1)
    zero_acc(acc);
    wave_mm(dz, kLd, M.w2t, 64, 64, acc);
    for (int mt = 0; mt < 2; mt++)
        for (int nt = 0; nt < 2; nt++)
            for (int i = 0; i < 16; i++) {
                const int m = acc_row(mt, i), c = acc_col(nt);
                g1[m * 64 + c] = acc[mt][nt][i] * g1[m * 64 + c];
            }
    __syncthreads();
    {
        // grad b1 and grad W1 = dz1^T a0 (K = 6): lane l owns output l
        float sb = 0.0f, gw[kMlpIn] = {0, 0, 0, 0, 0, 0};
        for (int m = 0; m < 64; m++) {
            const float d = g1[m * 64 + l];
            sb = sb + d;
            for (int k = 0; k < kMlpIn; k++) gw[k] = gw[k] + d * from_bf16(a0[m * kLd0 + k]);
        }
        P[kOffB1 + l] = sb;
        for (int k = 0; k < kMlpIn; k++) P[kOffW1 + l * kMlpIn + k] = gw[k];
    }
}

// Sum of the partial gradients in workgroup order; with lr > 0 an Adam step
// (step number t >= 1) on the master parameters, else the sum goes to grad.
__global__ void __launch_bounds__(256) mlp_adam_kernel(float *p, float *mom, float *vel, const float *partial, int nwg,
                                                       const float *partial_loss, float *loss_acc, float lr,
                                                       float beta1, float beta2, float eps, int t, float *grad)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j == 0 && loss_acc) {
        float s = 0.0f;
        for (int w = 0; w < nwg; w++) s = s + partial_loss[w];
        loss_acc[0] = loss_acc[0] + s;
    }
    if (j >= kMlpNTrain) return;
    float g = 0.0f;
    for (int w = 0; w < nwg; w++) g = g + partial[(int64_t)w * kMlpNTrain + j];
    if (grad) {
        grad[j] = g;
        return;
    }
    const float m = beta1 * mom[j] + (1.0f - beta1) * g;
    const float v = beta2 * vel[j] + (1.0f - beta2) * g * g;
    mom[j] = m;
    vel[j] = v;
    const float mh = m / (1.0f - powf(beta1, (float)t));
    const float vh = v / (1.0f - powf(beta2, (float)t));
    p[j] = p[j] - lr * mh / (sqrtf(vh) + eps);
}

hipError_t launch_mlp_prepare(const float *p, __bf16 *wb, float *w4r, hipStream_t s)
{
    mlp_prepare_kernel<<<16, 256, 0, s>>>(p, wb, w4r);
    return hipGetLastError();
}

}  // namespace eqd

using namespace eqd;

namespace {

struct DevMem {
    std::vector<void *> ptrs;
    ~DevMem()
    {
        for (void *q : ptrs) (void)hipFree(q);
    }
    template <typename T>
    T *alloc(size_t count)
    {
        void *q = nullptr;
        if (hipMalloc(&q, count * sizeof(T) + 16) != hipSuccess) return nullptr;
        ptrs.push_back(q);
        return (T *)q;
    }
};

int prepare(MlpBuffers &B, hipStream_t s)
{
    mlp_prepare_kernel<<<16, 256, 0, s>>>(B.p, B.wb, B.w4r);
    return hipGetLastError() == hipSuccess ? EQD_OK : EQD_EHIP;
}

bool alloc_net(DevMem &D, MlpBuffers &B)
{
    B.p = D.alloc<float>(kMlpNParams);
    B.wb = D.alloc<__bf16>(64 * 16 + 4 * 4096);
    B.w4r = D.alloc<float>(64);
    return B.p && B.wb && B.w4r;
}

// ---------------- scratch of the allocation-free entry point
// The network's prepared operands (bf16 weights w1, w2, w3 and their
// transposes, and w4 as float) of eqd_mlp_forward_noalloc live in this
// file-scope device array: kNoAllocScratch bytes per device, held until the
// process ends.  Host code never uses its name as an address: the device
// address comes from hipGetSymbolAddress, looked up once per device.
constexpr size_t kWbCount = 64 * 16 + 4 * 4096;                       // bf16 values
constexpr size_t kW4rOffset = (kWbCount * sizeof(__bf16) + 255) / 256 * 256;  // bytes
constexpr size_t kNoAllocScratch = kW4rOffset + 64 * sizeof(float);
static_assert(kNoAllocScratch == EQD_MLP_NOALLOC_SCRATCH_BYTES, "scratch size and header disagree");
static __device__ __align__(256) unsigned char scratch[kNoAllocScratch];

constexpr int kMaxDevices = 64;
void *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

// device address of this device's scratch array (device current), or null
void *scratch_on(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        g_scratch_ptr[device] = ptr;
    }
    return g_scratch_ptr[device];
}

int check_params(const float *params)
{
    const float *nrm = params + kOffNorm;
    for (int k = 0; k < kMlpNParams; k++)
        if (!(params[k] == params[k])) return EQD_EINVAL;
    if (nrm[2 * kMlpIn + 1] == 0.0f) return EQD_EINVAL;
    return EQD_OK;
}

}  // namespace

extern "C" {

int eqd_mlp_forward(const float *params, const float *X, int64_t n, float *r_out, int32_t device)
{
    if (!params || (n > 0 && (!X || !r_out)) || n < 0) return EQD_EINVAL;
    if (check_params(params) != EQD_OK) return EQD_EINVAL;
    if (n == 0) return EQD_OK;
    if (hipSetDevice(device) != hipSuccess) return EQD_EHIP;
    DevMem D;
    MlpBuffers B;
    if (!alloc_net(D, B)) return EQD_EOOM;
    float *dX = D.alloc<float>((size_t)n * kMlpIn), *dy = D.alloc<float>((size_t)n);
    if (!dX || !dy) return EQD_EOOM;
    if (hipMemcpy(B.p, params, kMlpNParams * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dX, X, (size_t)n * kMlpIn * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return EQD_EHIP;
    if (prepare(B, 0)) return EQD_EHIP;
    mlp_forward_kernel<<<(unsigned)((n + 63) / 64), 64>>>(B.net(), dX, n, dy);
    if (hipGetLastError() != hipSuccess) return EQD_EHIP;
    if (hipMemcpy(r_out, dy, (size_t)n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return EQD_EHIP;
    return EQD_OK;
}

int eqd_mlp_forward_noalloc(const float *params, const float *X, int64_t n, float *r_out, int32_t device,
                            void *stream)
{
    if (!params || (n > 0 && (!X || !r_out)) || n < 0) return EQD_EINVAL;
    if (n > (int64_t)0xFFFFFFFF * 64) return EQD_EINVAL;  // one launch covers at most 2^32 - 1 waves
    if (n == 0) return EQD_OK;
    if (hipSetDevice(device) != hipSuccess) return EQD_EHIP;
    unsigned char *sc = (unsigned char *)scratch_on(device);
    if (!sc) return EQD_EHIP;
    MlpBuffers B;
    B.p = const_cast<float *>(params);
    B.wb = (__bf16 *)sc;
    B.w4r = (float *)(sc + kW4rOffset);
    const hipStream_t st = (hipStream_t)stream;
    mlp_prepare_kernel<<<16, 256, 0, st>>>(B.p, B.wb, B.w4r);
    mlp_forward_kernel<<<(unsigned)((n + 63) / 64), 64, 0, st>>>(B.net(), X, n, r_out);
    return hipGetLastError() == hipSuccess ? EQD_OK : EQD_EHIP;
}

int eqd_mlp_grad(const float *params, const float *X, const float *logv, int64_t n, float *grad_out,
                 float *loss_out, int32_t device)
{
    if (!params || !X || !logv || !grad_out || n <= 0 || n > (int64_t)1 << 24) return EQD_EINVAL;
    if (check_params(params) != EQD_OK) return EQD_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return EQD_EHIP;
    DevMem D;
    MlpBuffers B;
    if (!alloc_net(D, B)) return EQD_EOOM;
    const int nwg = (int)((n + 63) / 64);
    float *dX = D.alloc<float>((size_t)n * kMlpIn), *dl = D.alloc<float>((size_t)n);
    float *part = D.alloc<float>((size_t)nwg * kMlpNTrain), *pl = D.alloc<float>(nwg), *la = D.alloc<float>(1);
    float *g = D.alloc<float>(kMlpNTrain);
    if (!dX || !dl || !part || !pl || !la || !g) return EQD_EOOM;
    if (hipMemcpy(B.p, params, kMlpNParams * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dX, X, (size_t)n * kMlpIn * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dl, logv, (size_t)n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(la, 0, sizeof(float)) != hipSuccess)
        return EQD_EHIP;
    if (prepare(B, 0)) return EQD_EHIP;
    mlp_grad_kernel<<<nwg, 64>>>(B.net(), dX, dl, n, 0, n, 1, 0, part, pl);
    mlp_adam_kernel<<<(kMlpNTrain + 255) / 256, 256>>>(B.p, nullptr, nullptr, part, nwg, pl, la, 0.0f, 0.0f, 0.0f,
                                                       0.0f, 1, g);
    if (hipGetLastError() != hipSuccess) return EQD_EHIP;
    float ls = 0.0f;
    if (hipMemcpy(grad_out, g, kMlpNTrain * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(&ls, la, sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        return EQD_EHIP;
    if (loss_out) *loss_out = ls / (float)n;
    return EQD_OK;
}

static uint64_t gcd64(uint64_t a, uint64_t b)
{
    while (b) {
        const uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

static uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int eqd_mlp_train(const float *X, const float *logv, int64_t n, const eqd_mlp_train_opts *o, float *params,
                  float *epoch_loss)
{
    if (!X || !logv || !o || !params || n <= 0 || n >= ((int64_t)1 << 31) || o->epochs <= 0 || o->batch <= 0 ||
        o->batch > (1 << 20) || !(o->lr > 0.0f))
        return EQD_EINVAL;
    if (check_params(params) != EQD_OK) return EQD_EINVAL;
    if (hipSetDevice(o->device) != hipSuccess) return EQD_EHIP;
    DevMem D;
    MlpBuffers B;
    if (!alloc_net(D, B)) return EQD_EOOM;
    const int64_t bs = o->batch < n ? o->batch : n;
    const int nwg = (int)((bs + 63) / 64);
    float *dX = D.alloc<float>((size_t)n * kMlpIn), *dl = D.alloc<float>((size_t)n);
    float *part = D.alloc<float>((size_t)nwg * kMlpNTrain), *pl = D.alloc<float>(nwg);
    float *mom = D.alloc<float>(kMlpNTrain), *vel = D.alloc<float>(kMlpNTrain), *la = D.alloc<float>(o->epochs);
    if (!dX || !dl || !part || !pl || !mom || !vel || !la) return EQD_EOOM;
    if (hipMemcpy(B.p, params, kMlpNParams * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dX, X, (size_t)n * kMlpIn * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dl, logv, (size_t)n * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(mom, 0, kMlpNTrain * sizeof(float)) != hipSuccess ||
        hipMemset(vel, 0, kMlpNTrain * sizeof(float)) != hipSuccess ||
        hipMemset(la, 0, o->epochs * sizeof(float)) != hipSuccess)
        return EQD_EHIP;
    if (prepare(B, 0)) return EQD_EHIP;
    uint64_t rs = o->seed;
    int t = 0;
    const int64_t nbatch = n / bs;  // full batches per epoch
    for (int e = 0; e < o->epochs; e++) {
        uint64_t mult;
        do {
            mult = splitmix(rs) % (uint64_t)n;
        } while (mult == 0 || gcd64(mult, (uint64_t)n) != 1);
        const uint64_t add = splitmix(rs) % (uint64_t)n;
        for (int64_t b = 0; b < nbatch; b++) {
            t++;
            mlp_grad_kernel<<<nwg, 64>>>(B.net(), dX, dl, n, b * bs, bs, mult, add, part, pl);
            mlp_adam_kernel<<<(kMlpNTrain + 255) / 256, 256>>>(B.p, mom, vel, part, nwg, pl, la + e, o->lr,
                                                               o->beta1, o->beta2, o->eps, t, nullptr);
            mlp_prepare_kernel<<<16, 256>>>(B.p, B.wb, B.w4r);
        }
        if (hipGetLastError() != hipSuccess) return EQD_EHIP;
    }
    if (hipDeviceSynchronize() != hipSuccess) return EQD_EHIP;
    if (hipMemcpy(params, B.p, kMlpNTrain * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return EQD_EHIP;
    if (epoch_loss) {
        std::vector<float> h(o->epochs);
        if (hipMemcpy(h.data(), la, o->epochs * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) return EQD_EHIP;
        for (int e = 0; e < o->epochs; e++) epoch_loss[e] = h[e] / (float)(nbatch * bs);
    }
    return EQD_OK;
}

}  // extern "C"
