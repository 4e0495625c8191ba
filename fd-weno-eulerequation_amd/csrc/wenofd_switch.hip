// wenofd_switch.hip — kernels of the learned troubled-stencil switch for
// gfx950: sample harvest, bf16 MFMA inference of the switch MLP (flags per
// face), flag dilation, dataset compaction, and the MLP's own training
// kernels (forward + backward + Adam, FP32).
//
// The arithmetic of features, labels and fluxes is in wenofd_switch.hpp; the
// fluxes that consume the flags are the hybrid variants of the sweep tiles in
// wenofd_kernels.hip.
#include <hip/hip_runtime.h>

#include "wenofd_kernels.hpp"
#include "wenofd_math.hpp"
#include "wenofd_switch.hpp"

namespace wenofd {

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef short i16x2 __attribute__((ext_vector_type(2)));

// ---------------------------------------------------------------------------
// MLP inference on the matrix cores, transposed so that one layer's output
// registers are the next layer's B operand without any data exchange:
//   H1^T (32 x 16 samples) = W1 (32 x 16 features, 10 used) . X^T
//   H2^T (32 x 16)         = W2 (32 x 32) . relu(H1^T + b1)
//   z (1 x 16)             = w3 . relu(H2^T + b2) + b3   (VALU FMAs + 2 shuffles)
// v_mfma_f32_16x16x16_bf16 lane layouts (lane l, j = 0..3):
//   A[m = l%16][k = 4(l/16)+j], B[k = 4(l/16)+j][n = l%16], D[m = 4(l/16)+j][n = l%16].
// One call handles 16 samples: lane l supplies features 4(l/16)..+3 of sample
// l%16 and receives that sample's logit.
// ---------------------------------------------------------------------------
struct MlpRegs {
    bf16x4 w1[2];     // [m]: W1[16m + l%16][4(l/16)+j]
    bf16x4 w2[2][2];  // [m][kt]: W2[16m + l%16][16kt + 4(l/16)+j]
    float b1[2][4], b2[2][4], w3[2][4];
    float b3;
};

__device__ __forceinline__ void mlp_load(const float* __restrict__ P, MlpRegs& R) {
    const int lane = threadIdx.x % 64, col = lane % 16, grp = lane / 16;
#pragma unroll
    for (int m = 0; m < 2; ++m) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = 4 * grp + j;
            R.w1[m][j] = (__bf16)(k < kFeat ? P[kOffW1 + (16 * m + col) * kFeat + k] : 0.0f);
            R.w2[m][0][j] = (__bf16)P[kOffW2 + (16 * m + col) * kHid + k];
            R.w2[m][1][j] = (__bf16)P[kOffW2 + (16 * m + col) * kHid + 16 + k];
            R.b1[m][j] = P[kOffB1 + 16 * m + k];
            R.b2[m][j] = P[kOffB2 + 16 * m + k];
            R.w3[m][j] = P[kOffW3 + 16 * m + k];
        }
    }
    R.b3 = P[kOffB3];
}

// The biases enter as the accumulators' initial values (the lane's C rows are
// exactly the units whose biases it holds), so no separate bias additions.
__device__ __forceinline__ float mlp_group(const MlpRegs& R, bf16x4 x) {
    bf16x4 h[2];
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        const f32x4 b = {R.b1[m][0], R.b1[m][1], R.b1[m][2], R.b1[m][3]};
        const f32x4 c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(R.w1[m], x, b, 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 4; j += 2) {
            // ReLU after the rounding to bf16, on the pair's sign bits: a
            // bf16 read as a 16-bit integer is negative exactly when its sign
            // bit is set, so a packed integer max with 0 zeroes the negative
            // values (and -0, and NaNs with the sign bit)
            const bf16x2 p = __builtin_convertvector((f32x2){c[j], c[j + 1]}, bf16x2);
            const i16x2 r = __builtin_elementwise_max(__builtin_bit_cast(i16x2, p), (i16x2){0, 0});
            const bf16x2 q = __builtin_bit_cast(bf16x2, r);
            h[m][j] = q[0];
            h[m][j + 1] = q[1];
        }
    }
    float z = 0.0f;
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        const f32x4 b = {R.b2[m][0], R.b2[m][1], R.b2[m][2], R.b2[m][3]};
        f32x4 d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(R.w2[m][0], h[0], b, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(R.w2[m][1], h[1], d, 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 4; ++j) z = __builtin_fmaf(R.w3[m][j], relu_f(d[j]), z);
    }
    z += __shfl_xor(z, 16, 64);
    z += __shfl_xor(z, 32, 64);
    return z + R.b3;
}

// bf16 features of one sample as 4 operand quads: quad q holds features 4q..4q+3.
struct SampleBf16 {
    bf16x4 q[4];
};

__device__ __forceinline__ SampleBf16 to_bf16(const float* x) {
    SampleBf16 s;
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int j = 0; j < 4; ++j) s.q[q][j] = (__bf16)(4 * q + j < kFeat ? x[4 * q + j] : 0.0f);
    return s;
}

// Logits of the wave's 64 x 4 samples: sX[field][lane] holds the features of
// the lane's face; sZ[field][lane] receives the logits. Wave-local LDS.
__device__ __forceinline__ void mlp_wave(const MlpRegs& R, const SampleBf16 (*sX)[64], float (*sZ)[64]) {
    const int lane = threadIdx.x % 64;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const bf16x4 x = sX[k][16 * q + lane % 16].q[lane / 16];
            const float z = mlp_group(R, x);
            if (lane < 16) sZ[k][16 * q + lane] = z;
        }
    }
}

// ---------------------------------------------------------------------------
// Per-face work over one direction: MODE 0 classifies (flags, optional
// logits), MODE 1 harvests (features + labels, dense per face).
// x: block = 256 consecutive faces of one interior row (split values of the
// 261 points in LDS). y: block = 64 columns x 4 face rows (split values of 9
// point rows in LDS). In both, wave w's lane l owns one face.
// Face index: x faces q = row (nx+1) + (f-4); y faces q = (f-4) nx + (i-4).
// ---------------------------------------------------------------------------
constexpr int kSwT = 256;

struct SwitchOut {
    const float* params;  // MODE 0: parameters (float, device)
    float thr;
    unsigned char* flags;  // MODE 0: raw flag per face
    float* logits;         // MODE 0, optional: [faces][4]
    float* X;              // MODE 1: [faces][4][10]
    unsigned char* Y;      // MODE 1: [faces][4]
};

template <int D, int MODE>
__global__ __launch_bounds__(kSwT) void k_switch_faces(const double* __restrict__ U, Geom g, SwitchOut o) {
    // split values (x: [4][264] x3 doubles; y: [4][9][64] x2 doubles), reused
    // as the bf16 feature tile once the features are in registers
    constexpr int kSplit = D == 0 ? 3 * 4 * 264 : 2 * 4 * 9 * 64;
    constexpr int kTile = 4 * 4 * 64 * (int)sizeof(SampleBf16) / 8;
    __shared__ __align__(16) double smem[kSplit > kTile ? kSplit : kTile];
    __shared__ float sZ[4][4][64];
    const int t = threadIdx.x, w = t / 64, lane = t % 64;
    int face_ok;
    size_t q;
    double vp[5][4], vm[5][4], Rm[16];
    if constexpr (D == 0) {
        double (*sP)[264] = reinterpret_cast<double (*)[264]>(smem);
        double (*sM)[264] = reinterpret_cast<double (*)[264]>(smem + 4 * 264);
        double (*sU)[264] = reinterpret_cast<double (*)[264]>(smem + 8 * 264);
        const int row = blockIdx.y, j = row + 4;
        const int f0 = 4 + blockIdx.x * kSwT;
        const double* rowp = U + (size_t)j * g.pitch;
        for (int p = t; p < kSwT + 5; p += kSwT) {
            const int pt = f0 - 3 + p;
            if (pt < g.nx + 8) {
                double u[4], fm[4], fp[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) u[r] = rowp[r * g.plane + pt];
                flux_split<0>(u, g.gamma, fm, fp);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    sP[r][p] = fp[r];
                    sM[r][p] = fm[r];
                    sU[r][p] = u[r];
                }
            }
        }
        __syncthreads();
        const int f = f0 + t;
        face_ok = f <= g.nx + 4;
        q = (size_t)row * (g.nx + 1) + (f - 4);
        if (face_ok) {
            double Ul[4], Ur[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                Ul[r] = sU[r][t + 2];
                Ur[r] = sU[r][t + 3];
            }
            face_char_stencils<0>(
                Ul, Ur, [&](int s, int r) { return sP[r][t + s]; }, [&](int s, int r) { return sM[r][t + s]; },
                g.gamma, vp, vm, Rm);
        }
    } else {
        double (*sP)[9][64] = reinterpret_cast<double (*)[9][64]>(smem);
        double (*sM)[9][64] = reinterpret_cast<double (*)[9][64]>(smem + 4 * 9 * 64);
        const int tx = lane, ty = w;
        const int i = 4 + blockIdx.x * 64 + tx;
        const bool col_ok = i < g.nx + 4;
        const int F0 = 4 + blockIdx.y * 4;
        for (int rr = ty; rr < 9; rr += 4) {
            const int prow = F0 - 3 + rr;
            if (col_ok && prow < g.ny + 8) {
                double u[4], fm[4], fp[4];
                const size_t o = (size_t)prow * g.pitch + i;
#pragma unroll
                for (int r = 0; r < 4; ++r) u[r] = U[r * g.plane + o];
                flux_split<1>(u, g.gamma, fm, fp);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    sP[r][rr][tx] = fp[r];
                    sM[r][rr][tx] = fm[r];
                }
            }
        }
        __syncthreads();
        const int fr = F0 + ty;
        face_ok = col_ok && fr <= g.ny + 4;
        q = (size_t)(fr - 4) * g.nx + (i - 4);
        if (face_ok) {
            double Ul[4], Ur[4];
            const size_t ol = (size_t)(fr - 1) * g.pitch + i;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                Ul[r] = U[r * g.plane + ol];
                Ur[r] = U[r * g.plane + ol + g.pitch];
            }
            face_char_stencils<1>(
                Ul, Ur, [&](int s, int r) { return sP[r][ty + s][tx]; },
                [&](int s, int r) { return sM[r][ty + s][tx]; }, g.gamma, vp, vm, Rm);
        }
    }
    float x[4][kFeat];
    int lab[4];
    if (face_ok) face_samples(vp, vm, x, lab);
    if constexpr (MODE == 1) {
        if (face_ok) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
#pragma unroll
                for (int c = 0; c < kFeat; ++c) o.X[(q * 4 + k) * kFeat + c] = x[k][c];
                o.Y[q * 4 + k] = (unsigned char)lab[k];
            }
        }
    } else {
        MlpRegs R;
        mlp_load(o.params, R);
        __syncthreads();  // every wave is done with the split values
        SampleBf16 (*sX)[4][64] = reinterpret_cast<SampleBf16 (*)[4][64]>(smem);  // [wave][field][lane]
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (!face_ok) {
#pragma unroll
                for (int c = 0; c < kFeat; ++c) x[k][c] = 0.0f;
            }
            sX[w][k][lane] = to_bf16(x[k]);
        }
        __syncthreads();
        mlp_wave(R, sX[w], sZ[w]);
        __syncthreads();
        if (face_ok) {
            unsigned char mask = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const float z = sZ[w][k][lane];
                if (z > o.thr) mask |= (unsigned char)(1u << k);
                if (o.logits) o.logits[q * 4 + k] = z;
            }
            o.flags[q] = mask;
        }
    }
}

// out[r][c] = bitwise OR of the field masks of in over the (2 rad + 1)^2
// neighbourhood (rows x cols grid of faces of one direction): each field's
// flags dilated on their own.
__global__ __launch_bounds__(256) void k_dilate(const unsigned char* __restrict__ in, unsigned char* __restrict__ out,
                                                int rows, int cols, int rad) {
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long)rows * cols) return;
    const int r = (int)(t / cols), c = (int)(t % cols);
    unsigned v = 0;
    for (int dr = -rad; dr <= rad; ++dr) {
        const int rr = r + dr;
        if (rr < 0 || rr >= rows) continue;
        for (int dc = -rad; dc <= rad; ++dc) {
            const int cc = c + dc;
            if (cc >= 0 && cc < cols) v |= in[(size_t)rr * cols + cc];
        }
    }
    out[t] = (unsigned char)(v & 15u);
}

// ---------------------------------------------------------------------------
// The list of one direction's faces with flagged fields, grouped by their
// number of flagged fields (1 .. 4; within a group in face order), so that a
// wave of the correction kernel runs one field count: k_list_count counts
// each group in each run of kListFaces faces, k_list_scan turns the counts
// into offsets (one block), k_list_write writes the faces. Deterministic:
// every position depends on the masks alone.
// ---------------------------------------------------------------------------
constexpr int kListFaces = 1024;  // faces per counting block (4 per thread)

// Counts of the four groups packed 16 bits each (a block holds <= 1024 faces).
__device__ __forceinline__ unsigned long long group_counts(unsigned m) {
    return m ? 1ull << (16 * (__popc(m) - 1)) : 0ull;
}

__global__ __launch_bounds__(256) void k_list_count(const unsigned char* __restrict__ m, long F,
                                                    int* __restrict__ cnt) {
    __shared__ unsigned long long s4[4];
    const long f0 = (long)blockIdx.x * kListFaces + 4 * threadIdx.x;
    unsigned long long c = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (f0 + q < F) c += group_counts(m[f0 + q] & 15u);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    if (threadIdx.x % 64 == 0) s4[threadIdx.x / 64] = c;
    __syncthreads();
    if (threadIdx.x < 4) {
        const unsigned long long t = s4[0] + s4[1] + s4[2] + s4[3];
        cnt[4 * blockIdx.x + threadIdx.x] = (int)((t >> (16 * threadIdx.x)) & 0xFFFFu);
    }
}

// off[4 b + g] = position of block b's first face of group g: the faces of
// groups below g, then those of group g in the blocks before b. *n = the
// total. One block of 1024 threads; thread t takes a contiguous run of blocks.
__global__ __launch_bounds__(1024) void k_list_scan(const int* __restrict__ cnt, int nb, int* __restrict__ off,
                                                    int* __restrict__ n) {
    __shared__ int s[1024];
    __shared__ int gbase[5];
    const int t = threadIdx.x;
    const int per = (nb + 1023) / 1024;
    const int b0 = t * per, b1 = b0 + per < nb ? b0 + per : nb;
    if (t == 0) gbase[0] = 0;
    for (int grp = 0; grp < 4; ++grp) {
        int sum = 0;
        for (int b = b0; b < b1; ++b) sum += cnt[4 * b + grp];
        s[t] = sum;
        __syncthreads();
        for (int d = 1; d < 1024; d <<= 1) {  // inclusive Hillis-Steele scan of the runs' sums
            const int v = t >= d ? s[t - d] : 0;
            __syncthreads();
            s[t] += v;
            __syncthreads();
        }
        if (t == 1023) gbase[grp + 1] = gbase[grp] + s[1023];
        __syncthreads();
        int run = gbase[grp] + s[t] - sum;
        for (int b = b0; b < b1; ++b) {
            off[4 * b + grp] = run;
            run += cnt[4 * b + grp];
        }
        __syncthreads();
    }
    if (t == 0) *n = gbase[4];
}

__global__ __launch_bounds__(256) void k_list_write(const unsigned char* __restrict__ m, long F,
                                                    const int* __restrict__ off, unsigned* __restrict__ list) {
    __shared__ unsigned long long s[256];
    const int t = threadIdx.x;
    const long f0 = (long)blockIdx.x * kListFaces + 4 * t;
    unsigned mk[4];
    unsigned long long c = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        mk[q] = f0 + q < F ? (m[f0 + q] & 15u) : 0u;
        c += group_counts(mk[q]);
    }
    s[t] = c;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {
        const unsigned long long v = t >= d ? s[t - d] : 0ull;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    unsigned long long pos = s[t] - c;  // this thread's first position in each group, within the block
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (f0 + q >= F) break;
        if (!mk[q]) continue;
        const int grp = __popc(mk[q]) - 1;
        const unsigned e = (unsigned)off[4 * blockIdx.x + grp] + (unsigned)((pos >> (16 * grp)) & 0xFFFFull);
        pos += 1ull << (16 * grp);
        list[e] = (unsigned)(f0 + q);
    }
}

__global__ __launch_bounds__(256) void k_fill_u8(unsigned char* __restrict__ p, long n, unsigned char v) {
    const long k = (long)blockIdx.x * 256 + threadIdx.x;
    if (k < n) p[k] = v;
}

// MLP inference over given features (the classifier's arithmetic), n samples.
__global__ __launch_bounds__(256) void k_mlp_infer(const float* __restrict__ P, const float* __restrict__ X, long n,
                                                   float* __restrict__ z) {
    __shared__ SampleBf16 sX[4][4][64];
    __shared__ float sZ[4][4][64];
    const int t = threadIdx.x, w = t / 64, lane = t % 64;
    MlpRegs R;
    mlp_load(P, R);
    const long s0 = ((long)blockIdx.x * 4 + w) * 256;  // the wave's 256 samples: s0 + 64 k + lane
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const long s = s0 + 64 * k + lane;
        float x[kFeat];
#pragma unroll
        for (int c = 0; c < kFeat; ++c) x[c] = s < n ? X[s * kFeat + c] : 0.0f;
        sX[w][k][lane] = to_bf16(x);
    }
    __syncthreads();
    mlp_wave(R, sX[w], sZ[w]);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const long s = s0 + 64 * k + lane;
        if (s < n) z[s] = sZ[w][k][lane];
    }
}

// ---------------------------------------------------------------------------
// Dataset compaction: sample s = face * 4 + field is kept when
// u(seed, s) < (label ? p_pos : p_neg), u uniform in [0, 1) from splitmix64.
// Two passes (count per block, then write at host-scanned offsets) keep the
// order deterministic.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool keep_sample(unsigned long long seed, long s, int label, double p_pos, double p_neg) {
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(s + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
    return u < (label ? p_pos : p_neg);
}

__global__ __launch_bounds__(256) void k_keep_count(const unsigned char* __restrict__ Y, long n,
                                                    unsigned long long seed, double p_pos, double p_neg,
                                                    int* __restrict__ counts) {
    __shared__ int sc[4];
    const long s = (long)blockIdx.x * 256 + threadIdx.x;
    const bool keep = s < n && keep_sample(seed, s, Y[s], p_pos, p_neg);
    const int c = __popcll(__ballot(keep));
    if (threadIdx.x % 64 == 0) sc[threadIdx.x / 64] = c;
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = sc[0] + sc[1] + sc[2] + sc[3];
}

__global__ __launch_bounds__(256) void k_keep_write(const float* __restrict__ X, const unsigned char* __restrict__ Y,
                                                    long n, unsigned long long seed, double p_pos, double p_neg,
                                                    const long long* __restrict__ offsets, float* __restrict__ Xo,
                                                    unsigned char* __restrict__ Yo) {
    __shared__ int sCnt[4];
    const int t = threadIdx.x, lane = t % 64, w = t / 64;
    const long s = (long)blockIdx.x * 256 + t;
    const bool keep = s < n && keep_sample(seed, s, Y[s], p_pos, p_neg);
    const unsigned long long m = __ballot(keep);
    if (lane == 0) sCnt[w] = __popcll(m);
    __syncthreads();
    int base = 0;
    for (int q = 0; q < w; ++q) base += sCnt[q];
    if (keep) {
        const long long o = offsets[blockIdx.x] + base + __popcll(m & ((1ull << lane) - 1ull));
#pragma unroll
        for (int c = 0; c < kFeat; ++c) Xo[o * kFeat + c] = X[s * kFeat + c];
        Yo[o] = Y[s];
    }
}

// ---------------------------------------------------------------------------
// Training, FP32. k_mlp_grad: each block takes 64 samples of the batch
// (4 threads per sample split the 32 hidden units), runs the forward and the
// backward pass with activations in LDS, then reduces the per-sample
// gradients parameter by parameter in sample order into its partial row.
// k_mlp_reduce sums the partial rows in block order; k_adam updates.
// Loss: weighted binary cross-entropy with logits,
//   pw y softplus(-z) + (1 - y) softplus(z), mean over the batch.
// ---------------------------------------------------------------------------
constexpr int kGradS = 64;

__global__ __launch_bounds__(256) void k_mlp_grad(const float* __restrict__ P, const float* __restrict__ X,
                                                  const unsigned char* __restrict__ Y, const int* __restrict__ idx,
                                                  int n, float pw, float* __restrict__ partial,
                                                  float* __restrict__ lpart) {
    __shared__ float sX[kGradS][kFeat];
    __shared__ float sA1[kGradS][kHid], sH1[kGradS][kHid], sA2[kGradS][kHid], sH2[kGradS][kHid];
    __shared__ float sD1[kGradS][kHid], sD2[kGradS][kHid];
    __shared__ float sDz[kGradS], sLoss[kGradS];
    const int t = threadIdx.x, sl = t / 4, qq = t % 4;
    const int sg = blockIdx.x * kGradS + sl;
    const bool valid = sg < n;
    const int sidx = valid ? idx[sg] : 0;
    if (qq == 0) {
#pragma unroll
        for (int c = 0; c < kFeat; ++c) sX[sl][c] = valid ? X[(long)sidx * kFeat + c] : 0.0f;
    }
    __syncthreads();
    for (int u = 8 * qq; u < 8 * qq + 8; ++u) {
        float a = P[kOffB1 + u];
        for (int c = 0; c < kFeat; ++c) a += P[kOffW1 + u * kFeat + c] * sX[sl][c];
        sA1[sl][u] = a;
        sH1[sl][u] = relu_f(a);
    }
    __syncthreads();
    float zp = 0.0f;
    for (int u = 8 * qq; u < 8 * qq + 8; ++u) {
        float a = P[kOffB2 + u];
        for (int c = 0; c < kHid; ++c) a += P[kOffW2 + u * kHid + c] * sH1[sl][c];
        sA2[sl][u] = a;
        const float h = relu_f(a);
        sH2[sl][u] = h;
        zp += P[kOffW3 + u] * h;
    }
    zp += __shfl_xor(zp, 1, 64);
    zp += __shfl_xor(zp, 2, 64);
    const float z = zp + P[kOffB3];
    const float y = (valid && Y[sidx]) ? 1.0f : 0.0f;
    const float sgm = 1.0f / (1.0f + expf(-z));
    const float dz = valid ? ((1.0f - y) * sgm - pw * y * (1.0f - sgm)) : 0.0f;
    if (qq == 0) {
        const float spz = fmaxf(z, 0.0f) + log1pf(expf(-fabsf(z)));  // softplus(z)
        const float spm = spz - z;                                    // softplus(-z)
        sLoss[sl] = valid ? (pw * y * spm + (1.0f - y) * spz) : 0.0f;
        sDz[sl] = dz;
    }
    for (int u = 8 * qq; u < 8 * qq + 8; ++u) sD2[sl][u] = dz * P[kOffW3 + u] * relu_grad_f(sA2[sl][u]);
    __syncthreads();
    for (int u = 8 * qq; u < 8 * qq + 8; ++u) {
        float d = 0.0f;
        for (int c = 0; c < kHid; ++c) d += P[kOffW2 + c * kHid + u] * sD2[sl][c];
        sD1[sl][u] = d * relu_grad_f(sA1[sl][u]);
    }
    __syncthreads();
    float* out = partial + (size_t)blockIdx.x * kNParams;
    for (int p = t; p < kNParams; p += 256) {
        float acc = 0.0f;
        if (p < kOffB1) {
            const int u = p / kFeat, c = p % kFeat;
            for (int s = 0; s < kGradS; ++s) acc += sD1[s][u] * sX[s][c];
        } else if (p < kOffW2) {
            const int u = p - kOffB1;
            for (int s = 0; s < kGradS; ++s) acc += sD1[s][u];
        } else if (p < kOffB2) {
            const int u = (p - kOffW2) / kHid, c = (p - kOffW2) % kHid;
            for (int s = 0; s < kGradS; ++s) acc += sD2[s][u] * sH1[s][c];
        } else if (p < kOffW3) {
            const int u = p - kOffB2;
            for (int s = 0; s < kGradS; ++s) acc += sD2[s][u];
        } else if (p < kOffB3) {
            const int c = p - kOffW3;
            for (int s = 0; s < kGradS; ++s) acc += sDz[s] * sH2[s][c];
        } else {
            for (int s = 0; s < kGradS; ++s) acc += sDz[s];
        }
        out[p] = acc;
    }
    if (t == 0) {
        float l = 0.0f;
        for (int s = 0; s < kGradS; ++s) l += sLoss[s];
        lpart[blockIdx.x] = l;
    }
}

__global__ __launch_bounds__(256) void k_mlp_reduce(const float* __restrict__ partial, const float* __restrict__ lpart,
                                                    int nblocks, int n, float* __restrict__ grad,
                                                    float* __restrict__ loss) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p < kNParams) {
        float acc = 0.0f;
        for (int b = 0; b < nblocks; ++b) acc += partial[(size_t)b * kNParams + p];
        grad[p] = acc / (float)n;
    }
    if (p == kNParams) {
        float l = 0.0f;
        for (int b = 0; b < nblocks; ++b) l += lpart[b];
        *loss = l / (float)n;
    }
}

__global__ __launch_bounds__(256) void k_adam(float* __restrict__ P, const float* __restrict__ grad,
                                              float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                                              float b2, float eps, float c1, float c2) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= kNParams) return;
    const float g = grad[p];
    const float mp = b1 * m[p] + (1.0f - b1) * g;
    const float vp = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mp;
    v[p] = vp;
    P[p] -= lr * (mp / c1) / (sqrtf(vp / c2) + eps);  // c1 = 1 - b1^t, c2 = 1 - b2^t
}

// ---------------------------------------------------------------------------
// Launch wrappers.
// ---------------------------------------------------------------------------
static dim3 switch_grid(int dir, const Geom& g) {
    return dir == 0 ? dim3((g.nx + 1 + kSwT - 1) / kSwT, g.ny) : dim3((g.nx + 63) / 64, (g.ny + 1 + 3) / 4);
}

hipError_t launch_classify(int dir, const double* U, const Geom& g, const float* params, float thr,
                           unsigned char* flags, float* logits, hipStream_t s) {
    SwitchOut o{params, thr, flags, logits, nullptr, nullptr};
    if (dir == 0)
        hipLaunchKernelGGL((k_switch_faces<0, 0>), switch_grid(0, g), dim3(kSwT), 0, s, U, g, o);
    else
        hipLaunchKernelGGL((k_switch_faces<1, 0>), switch_grid(1, g), dim3(kSwT), 0, s, U, g, o);
    return hipGetLastError();
}

hipError_t launch_harvest(int dir, const double* U, const Geom& g, float* X, unsigned char* Y, hipStream_t s) {
    SwitchOut o{nullptr, 0.0f, nullptr, nullptr, X, Y};
    if (dir == 0)
        hipLaunchKernelGGL((k_switch_faces<0, 1>), switch_grid(0, g), dim3(kSwT), 0, s, U, g, o);
    else
        hipLaunchKernelGGL((k_switch_faces<1, 1>), switch_grid(1, g), dim3(kSwT), 0, s, U, g, o);
    return hipGetLastError();
}

hipError_t launch_dilate(const unsigned char* in, unsigned char* out, int rows, int cols, int rad, hipStream_t s) {
    const long n = (long)rows * cols;
    hipLaunchKernelGGL(k_dilate, dim3((n + 255) / 256), dim3(256), 0, s, in, out, rows, cols, rad);
    return hipGetLastError();
}

long list_blocks(long F) { return (F + kListFaces - 1) / kListFaces; }

hipError_t launch_build_list(const unsigned char* masks, long F, int* cnt, int* off, unsigned* list, int* n,
                             hipStream_t s) {
    const long nb = list_blocks(F);
    hipLaunchKernelGGL(k_list_count, dim3(nb), dim3(256), 0, s, masks, F, cnt);
    hipLaunchKernelGGL(k_list_scan, dim3(1), dim3(1024), 0, s, cnt, (int)nb, off, n);
    hipLaunchKernelGGL(k_list_write, dim3(nb), dim3(256), 0, s, masks, F, off, list);
    return hipGetLastError();
}

hipError_t launch_fill_u8(unsigned char* p, long n, unsigned char v, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_u8, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_mlp_infer(const float* P, const float* X, long n, float* z, hipStream_t s) {
    hipLaunchKernelGGL(k_mlp_infer, dim3((n + 1023) / 1024), dim3(256), 0, s, P, X, n, z);
    return hipGetLastError();
}

hipError_t launch_keep_count(const unsigned char* Y, long n, unsigned long long seed, double p_pos, double p_neg,
                             int* counts, hipStream_t s) {
    hipLaunchKernelGGL(k_keep_count, dim3((n + 255) / 256), dim3(256), 0, s, Y, n, seed, p_pos, p_neg, counts);
    return hipGetLastError();
}

hipError_t launch_keep_write(const float* X, const unsigned char* Y, long n, unsigned long long seed, double p_pos,
                             double p_neg, const long long* offsets, float* Xo, unsigned char* Yo, hipStream_t s) {
    hipLaunchKernelGGL(k_keep_write, dim3((n + 255) / 256), dim3(256), 0, s, X, Y, n, seed, p_pos, p_neg, offsets,
                       Xo, Yo);
    return hipGetLastError();
}

int mlp_grad_blocks(int n) { return (n + kGradS - 1) / kGradS; }

hipError_t launch_mlp_grad(const float* P, const float* X, const unsigned char* Y, const int* idx, int n, float pw,
                           float* partial, float* lpart, float* grad, float* loss, hipStream_t s) {
    const int nb = mlp_grad_blocks(n);
    hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(256), 0, s, P, X, Y, idx, n, pw, partial, lpart);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((kNParams + 1 + 255) / 256), dim3(256), 0, s, partial, lpart, nb, n, grad,
                       loss);
    return hipGetLastError();
}

hipError_t launch_adam(float* P, const float* grad, float* m, float* v, float lr, float b1, float b2, float eps,
                       float c1, float c2, hipStream_t s) {
    hipLaunchKernelGGL(k_adam, dim3((kNParams + 255) / 256), dim3(256), 0, s, P, grad, m, v, lr, b1, b2, eps, c1, c2);
    return hipGetLastError();
}

}  // namespace wenofd
