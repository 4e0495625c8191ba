// wenofd_mlp.cpp — C ABI of the switch MLP trainer (wenofd_mlp_*): a
// device-resident sample set, the project's own FP32 forward/backward/Adam
// kernels (wenofd_switch.hip), and bf16 MFMA inference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "wenofd_amd.h"
#include "wenofd_host.hpp"
#include "wenofd_kernels.hpp"
#include "wenofd_mlp.hpp"
#include "wenofd_switch.hpp"

using namespace wenofd;

#define MHIP_TRY(expr)                                                                           \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return set_last_error(WENOFD_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {

int fail(int code, const std::string& msg) { return set_last_error(code, msg); }

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double u01(uint64_t& s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

// PyTorch-style Linear initialisation: weights and biases ~ U(-1/sqrt(fan_in), 1/sqrt(fan_in)).
void default_params(uint64_t seed, std::vector<float>& P) {
    P.assign(kNParams, 0.0f);
    uint64_t s = seed;
    auto fill = [&](int off, int n, int fan_in) {
        const double b = 1.0 / std::sqrt((double)fan_in);
        for (int k = 0; k < n; ++k) P[off + k] = (float)((2.0 * u01(s) - 1.0) * b);
    };
    fill(kOffW1, kHid * kFeat, kFeat);
    fill(kOffB1, kHid, kFeat);
    fill(kOffW2, kHid * kHid, kHid);
    fill(kOffB2, kHid, kHid);
    fill(kOffW3, kHid, kHid);
    fill(kOffB3, 1, kHid);
}

int ensure_idx(wenofd_mlp* m, long long n) {
    if (n <= m->idx_cap) return 0;
    (void)hipFree(m->idx);
    m->idx = nullptr;
    MHIP_TRY(hipMalloc(&m->idx, n * sizeof(int)));
    m->idx_cap = n;
    return 0;
}

int ensure_partial(wenofd_mlp* m, int n) {
    const int nb = mlp_grad_blocks(n);
    if (nb <= m->partial_blocks) return 0;
    (void)hipFree(m->partial);
    (void)hipFree(m->lpart);
    m->partial = nullptr;
    m->lpart = nullptr;
    MHIP_TRY(hipMalloc(&m->partial, (size_t)nb * kNParams * sizeof(float)));
    MHIP_TRY(hipMalloc(&m->lpart, (size_t)nb * sizeof(float)));
    m->partial_blocks = nb;
    return 0;
}

int ensure_losses(wenofd_mlp* m, long long n) {
    if (n <= m->losses_cap) return 0;
    (void)hipFree(m->losses);
    m->losses = nullptr;
    MHIP_TRY(hipMalloc(&m->losses, n * sizeof(float)));
    m->losses_cap = n;
    return 0;
}

int reset_adam(wenofd_mlp* m) {
    MHIP_TRY(hipMemsetAsync(m->m, 0, kNParams * sizeof(float), m->stream));
    MHIP_TRY(hipMemsetAsync(m->v, 0, kNParams * sizeof(float), m->stream));
    m->adam_t = 0;
    return 0;
}

}  // namespace

namespace wenofd {

int mlp_reserve(wenofd_mlp* m, long long need) {
    if (need <= m->cap) return 0;
    const long long cap = std::max(need, std::max(2 * m->cap, (long long)1 << 20));
    float* X = nullptr;
    unsigned char* Y = nullptr;
    MHIP_TRY(hipMalloc(&X, cap * kFeat * sizeof(float)));
    hipError_t e = hipMalloc(&Y, cap);
    if (e != hipSuccess) {
        (void)hipFree(X);
        return fail(WENOFD_E_HIP, std::string("sample set allocation: ") + hipGetErrorString(e));
    }
    if (m->n > 0) {
        MHIP_TRY(hipMemcpyAsync(X, m->X, m->n * kFeat * sizeof(float), hipMemcpyDeviceToDevice, m->stream));
        MHIP_TRY(hipMemcpyAsync(Y, m->Y, m->n, hipMemcpyDeviceToDevice, m->stream));
        MHIP_TRY(hipStreamSynchronize(m->stream));
    }
    (void)hipFree(m->X);
    (void)hipFree(m->Y);
    m->X = X;
    m->Y = Y;
    m->cap = cap;
    return 0;
}

int mlp_reserve_dense(wenofd_mlp* m, long long faces) {
    if (faces <= m->dense_cap) return 0;
    (void)hipFree(m->dense_X);
    (void)hipFree(m->dense_Y);
    m->dense_X = nullptr;
    m->dense_Y = nullptr;
    MHIP_TRY(hipMalloc(&m->dense_X, faces * 4 * kFeat * sizeof(float)));
    MHIP_TRY(hipMalloc(&m->dense_Y, faces * 4));
    m->dense_cap = faces;
    const long long nblk = (faces * 4 + 255) / 256;
    if (nblk > m->counts_cap) {
        (void)hipFree(m->counts);
        (void)hipFree(m->offsets);
        m->counts = nullptr;
        m->offsets = nullptr;
        MHIP_TRY(hipMalloc(&m->counts, nblk * sizeof(int)));
        MHIP_TRY(hipMalloc(&m->offsets, nblk * sizeof(long long)));
        m->counts_cap = nblk;
    }
    return 0;
}

int mlp_append_dense(wenofd_mlp* m, long long nsamp, unsigned long long seed, double p_pos, double p_neg,
                     long long* added) {
    const long long nblk = (nsamp + 255) / 256;
    MHIP_TRY(launch_keep_count(m->dense_Y, nsamp, seed, p_pos, p_neg, m->counts, m->stream));
    std::vector<int> counts(nblk);
    MHIP_TRY(hipMemcpyAsync(counts.data(), m->counts, nblk * sizeof(int), hipMemcpyDeviceToHost, m->stream));
    MHIP_TRY(hipStreamSynchronize(m->stream));
    std::vector<long long> offs(nblk);
    long long tot = 0;
    for (long long b = 0; b < nblk; ++b) {
        offs[b] = m->n + tot;
        tot += counts[b];
    }
    int rc = mlp_reserve(m, m->n + tot);
    if (rc) return rc;
    MHIP_TRY(hipMemcpyAsync(m->offsets, offs.data(), nblk * sizeof(long long), hipMemcpyHostToDevice, m->stream));
    MHIP_TRY(launch_keep_write(m->dense_X, m->dense_Y, nsamp, seed, p_pos, p_neg, m->offsets, m->X, m->Y,
                               m->stream));
    std::vector<unsigned char> ynew(tot);
    if (tot > 0)
        MHIP_TRY(hipMemcpyAsync(ynew.data(), m->Y + m->n, tot, hipMemcpyDeviceToHost, m->stream));
    MHIP_TRY(hipStreamSynchronize(m->stream));
    for (unsigned char y : ynew) m->positives += y ? 1 : 0;
    m->n += tot;
    if (added) *added = tot;
    return 0;
}

}  // namespace wenofd

extern "C" {

int wenofd_mlp_create(int device, unsigned long long seed, wenofd_mlp** out) {
    if (!out) return fail(WENOFD_E_ARG, "null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(WENOFD_E_HIP, "no HIP device available");
    if (device < 0 && hipGetDevice(&device) != hipSuccess) return fail(WENOFD_E_HIP, "hipGetDevice failed");
    if (device >= ndev) return fail(WENOFD_E_ARG, "device index out of range");
    auto* m = new wenofd_mlp();
    m->device = device;
    auto bail = [&](int rc) {
        wenofd_mlp_destroy(m);
        return rc;
    };
    if (hipSetDevice(device) != hipSuccess) return bail(fail(WENOFD_E_HIP, "hipSetDevice failed"));
    hipError_t e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&m->P, kNParams * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&m->m, kNParams * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&m->v, kNParams * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&m->grad, kNParams * sizeof(float));
    if (e != hipSuccess) return bail(fail(WENOFD_E_HIP, std::string("allocation: ") + hipGetErrorString(e)));
    std::vector<float> P;
    default_params(seed, P);
    int rc = wenofd_mlp_set_params(m, P.data());
    if (rc) return bail(rc);
    *out = m;
    return 0;
}

int wenofd_mlp_destroy(wenofd_mlp* m) {
    if (!m) return 0;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (void* p : {(void*)m->P, (void*)m->m, (void*)m->v, (void*)m->X, (void*)m->Y, (void*)m->grad,
                    (void*)m->partial, (void*)m->lpart, (void*)m->idx, (void*)m->losses, (void*)m->dense_X,
                    (void*)m->dense_Y, (void*)m->counts, (void*)m->offsets})
        (void)hipFree(p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return 0;
}

int wenofd_mlp_set_params(wenofd_mlp* m, const float* params) {
    if (!m || !params) return fail(WENOFD_E_ARG, "null argument");
    MHIP_TRY(hipSetDevice(m->device));
    MHIP_TRY(hipMemcpyAsync(m->P, params, kNParams * sizeof(float), hipMemcpyHostToDevice, m->stream));
    int rc = reset_adam(m);
    if (rc) return rc;
    MHIP_TRY(hipStreamSynchronize(m->stream));
    return 0;
}

int wenofd_mlp_get_params(wenofd_mlp* m, float* params) {
    if (!m || !params) return fail(WENOFD_E_ARG, "null argument");
    MHIP_TRY(hipSetDevice(m->device));
    MHIP_TRY(hipMemcpyAsync(params, m->P, kNParams * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MHIP_TRY(hipStreamSynchronize(m->stream));
    return 0;
}

int wenofd_mlp_dataset_add(wenofd_mlp* m, const float* X, const unsigned char* Y, long long n) {
    if (!m || (n > 0 && (!X || !Y)) || n < 0) return fail(WENOFD_E_ARG, "bad argument");
    MHIP_TRY(hipSetDevice(m->device));
    int rc = mlp_reserve(m, m->n + n);
    if (rc) return rc;
    MHIP_TRY(hipMemcpyAsync(m->X + m->n * kFeat, X, n * kFeat * sizeof(float), hipMemcpyHostToDevice, m->stream));
    MHIP_TRY(hipMemcpyAsync(m->Y + m->n, Y, n, hipMemcpyHostToDevice, m->stream));
    MHIP_TRY(hipStreamSynchronize(m->stream));
    for (long long k = 0; k < n; ++k) m->positives += Y[k] ? 1 : 0;
    m->n += n;
    return 0;
}

int wenofd_mlp_dataset_clear(wenofd_mlp* m) {
    if (!m) return fail(WENOFD_E_ARG, "null handle");
    m->n = 0;
    m->positives = 0;
    return 0;
}

int wenofd_mlp_dataset_size(wenofd_mlp* m, long long* n, long long* positives) {
    if (!m) return fail(WENOFD_E_ARG, "null handle");
    if (n) *n = m->n;
    if (positives) *positives = m->positives;
    return 0;
}

int wenofd_mlp_dataset_get(wenofd_mlp* m, long long start, long long n, float* X, unsigned char* Y) {
    if (!m || start < 0 || n < 0 || start + n > m->n) return fail(WENOFD_E_ARG, "range outside the sample set");
    MHIP_TRY(hipSetDevice(m->device));
    if (X) MHIP_TRY(hipMemcpyAsync(X, m->X + start * kFeat, n * kFeat * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    if (Y) MHIP_TRY(hipMemcpyAsync(Y, m->Y + start, n, hipMemcpyDeviceToHost, m->stream));
    MHIP_TRY(hipStreamSynchronize(m->stream));
    return 0;
}

int wenofd_mlp_grad(wenofd_mlp* m, const int* idx, int n, float pos_weight, float* grad, float* loss) {
    if (!m || !idx || n <= 0) return fail(WENOFD_E_ARG, "bad argument");
    for (int k = 0; k < n; ++k)
        if (idx[k] < 0 || idx[k] >= m->n) return fail(WENOFD_E_ARG, "sample index outside the sample set");
    MHIP_TRY(hipSetDevice(m->device));
    int rc;
    if ((rc = ensure_idx(m, n)) || (rc = ensure_partial(m, n)) || (rc = ensure_losses(m, 1))) return rc;
    MHIP_TRY(hipMemcpyAsync(m->idx, idx, n * sizeof(int), hipMemcpyHostToDevice, m->stream));
    MHIP_TRY(launch_mlp_grad(m->P, m->X, m->Y, m->idx, n, pos_weight, m->partial, m->lpart, m->grad, m->losses,
                             m->stream));
    if (grad) MHIP_TRY(hipMemcpyAsync(grad, m->grad, kNParams * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    if (loss) MHIP_TRY(hipMemcpyAsync(loss, m->losses, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MHIP_TRY(hipStreamSynchronize(m->stream));
    return 0;
}

int wenofd_mlp_train(wenofd_mlp* m, int epochs, int batch, float lr, float pos_weight, unsigned long long seed,
                     float* epoch_losses) {
    if (!m || epochs <= 0 || batch <= 0 || batch > (1 << 22)) return fail(WENOFD_E_ARG, "bad argument");
    if (m->n == 0) return fail(WENOFD_E_ARG, "empty sample set");
    if (m->n > 0x7fffffffLL) return fail(WENOFD_E_ARG, "sample set too large");
    MHIP_TRY(hipSetDevice(m->device));
    const int N = (int)m->n;
    const int nb = (N + batch - 1) / batch;
    int rc;
    if ((rc = ensure_idx(m, N)) || (rc = ensure_partial(m, std::min(batch, N))) || (rc = ensure_losses(m, nb)))
        return rc;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    std::vector<int> perm(N);
    std::vector<float> losses(nb);
    uint64_t s = seed;
    for (int e = 0; e < epochs; ++e) {
        for (int k = 0; k < N; ++k) perm[k] = k;
        for (int k = N - 1; k > 0; --k) std::swap(perm[k], perm[(int)(splitmix(s) % (uint64_t)(k + 1))]);
        MHIP_TRY(hipMemcpyAsync(m->idx, perm.data(), (size_t)N * sizeof(int), hipMemcpyHostToDevice, m->stream));
        for (int b = 0; b < nb; ++b) {
            const int n = std::min(batch, N - b * batch);
            MHIP_TRY(launch_mlp_grad(m->P, m->X, m->Y, m->idx + (size_t)b * batch, n, pos_weight, m->partial,
                                     m->lpart, m->grad, m->losses + b, m->stream));
            m->adam_t += 1;
            const float c1 = 1.0f - std::pow(b1, (float)m->adam_t);
            const float c2 = 1.0f - std::pow(b2, (float)m->adam_t);
            MHIP_TRY(launch_adam(m->P, m->grad, m->m, m->v, lr, b1, b2, eps, c1, c2, m->stream));
        }
        MHIP_TRY(hipMemcpyAsync(losses.data(), m->losses, nb * sizeof(float), hipMemcpyDeviceToHost, m->stream));
        MHIP_TRY(hipStreamSynchronize(m->stream));  // also keeps perm alive until the copy is done
        double acc = 0;
        for (int b = 0; b < nb; ++b) acc += (double)losses[b] * std::min(batch, N - b * batch);
        if (epoch_losses) epoch_losses[e] = (float)(acc / N);
    }
    return 0;
}

int wenofd_mlp_infer(wenofd_mlp* m, const float* X, long long n, float* logits) {
    if (!m || !X || !logits || n < 0) return fail(WENOFD_E_ARG, "bad argument");
    if (n == 0) return 0;
    MHIP_TRY(hipSetDevice(m->device));
    float *dX = nullptr, *dz = nullptr;
    MHIP_TRY(hipMalloc(&dX, n * kFeat * sizeof(float)));
    hipError_t e = hipMalloc(&dz, n * sizeof(float));
    if (e == hipSuccess) e = hipMemcpyAsync(dX, X, n * kFeat * sizeof(float), hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) e = launch_mlp_infer(m->P, dX, n, dz, m->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(logits, dz, n * sizeof(float), hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(dX);
    (void)hipFree(dz);
    if (e != hipSuccess) return fail(WENOFD_E_HIP, std::string("mlp_infer: ") + hipGetErrorString(e));
    return 0;
}

// Confusion counts {TP, FP, FN, TN} of the deployed (bf16 MFMA) network over
// samples [start, start + n) at threshold thr.
int wenofd_mlp_evaluate(wenofd_mlp* m, long long start, long long n, float thr, long long* counts4) {
    if (!m || !counts4 || start < 0 || n <= 0 || start + n > m->n) return fail(WENOFD_E_ARG, "bad argument");
    MHIP_TRY(hipSetDevice(m->device));
    float* dz = nullptr;
    MHIP_TRY(hipMalloc(&dz, n * sizeof(float)));
    std::vector<float> z(n);
    std::vector<unsigned char> y(n);
    hipError_t e = launch_mlp_infer(m->P, m->X + start * kFeat, n, dz, m->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(z.data(), dz, n * sizeof(float), hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(y.data(), m->Y + start, n, hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    (void)hipFree(dz);
    if (e != hipSuccess) return fail(WENOFD_E_HIP, std::string("mlp_evaluate: ") + hipGetErrorString(e));
    long long tp = 0, fp = 0, fn = 0, tn = 0;
    for (long long k = 0; k < n; ++k) {
        const bool p = z[k] > thr, t = y[k] != 0;
        tp += p && t;
        fp += p && !t;
        fn += !p && t;
        tn += !p && !t;
    }
    counts4[0] = tp;
    counts4[1] = fp;
    counts4[2] = fn;
    counts4[3] = tn;
    return 0;
}

}  // extern "C"
