// nq_mlp.hip -- the learned cost model: per-constellation features and a
// 3-layer FP32 MLP (24 -> 64 -> 64 -> 1, ReLU) with its forward, backward
// and Adam kernels.  The model predicts log(1 + DFS nodes) of a
// constellation; it only orders work, never changes a count.
//
// Parameter layout (NQ_MLP_NPARAM floats):
//   W1[64][24] | b1[64] | W2[64][64] | b2[64] | W3[64] | b3
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>
#include <math.h>
#include <algorithm>
#include "nq_task.h"
#include "nq_kernels.h"
#include "nq_dfs.h"

namespace nqk {

constexpr int kIn = 24, kH = 64;
constexpr int kW1 = 0, kB1 = kW1 + kH * kIn, kW2 = kB1 + kH, kB2 = kW2 + kH * kH, kW3 = kB2 + kH,
              kB3 = kW3 + kH, kNP = kB3 + 1;

// ---------------------------------------------------------------------------
// features: 24 floats per record, from the record and its decoded descriptor
// ---------------------------------------------------------------------------
__global__ void k_features(const uint4 *__restrict__ rec, const uint4 *__restrict__ root,
                           const uint4 *__restrict__ sched, uint32_t m, float *__restrict__ feat)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m) return;
    const uint4 r = rec[c];
    const int N = (int)((sched[kSchedU4 * c + 1].x >> 16) & 31u);
    nqt::Desc D;
    nqt::decode(N, r.x, r.y, r.z, r.w, D);
    const uint4 fr = root[c];
    const uint32_t mask = (1u << N) - 1u;
    const float inv = 1.0f / (float)N;
    float *f = feat + (size_t)c * kIn;
    f[0] = D.i * inv;
    f[1] = D.j * inv;
    f[2] = D.k * inv;
    f[3] = D.l * inv;
    f[4] = D.start * inv;
    f[5] = D.fid * (1.0f / 27.0f);
    f[6] = __builtin_popcount(D.free0) * inv;
    f[7] = (D.endmark - D.start) * inv;
    f[8] = (D.mark1 > D.start ? D.mark1 - D.start : 0) * inv;
    f[9] = (D.mark2 > D.start ? D.mark2 - D.start : 0) * inv;
    f[10] = (D.jmark > D.start ? D.jmark - D.start : 0) * inv;
    f[11] = (float)D.wshift * (1.0f / 3.0f);
    f[12] = __builtin_popcount(fr.x & mask) * inv;
    f[13] = __builtin_popcount(fr.y & mask) * inv;
    f[14] = __builtin_popcount(fr.z & mask) * inv;
    // free squares of the next five rows under the current diagonals/columns
    float lsum = 0.f;
    for (int t = 1; t <= 5; t++) {
        const uint32_t nf = mask & ~((fr.x << t) | (fr.y >> t) | fr.z);
        const int pc = __builtin_popcount(nf);
        f[14 + t] = pc * inv;
        lsum += log2f(1.0f + (float)pc);
    }
    f[20] = (float)D.T * inv;
    f[21] = (float)N * (1.0f / 21.0f);
    int marks = 0;
    for (int d = 0; d < (int)D.T; d++) marks += ((nqt::param_byte(D.tab, d) & 3u) > 1u);
    f[22] = (float)marks * 0.5f;
    f[23] = lsum * 0.1f;
}

__device__ __forceinline__ float wave_sum_f32(float v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ float lane_f32(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// ---------------------------------------------------------------------------
// inference: y[b] for rows b = 0..m-1.  Lane o of a wave holds hidden unit
// o's weights in registers (row o of W1 and of W2, W3[o]) for the whole
// kernel; the wave walks rows (wave w of the grid takes w, w + waves, ...),
// broadcasting each row's features and first-layer activations by readlane.
// Each hidden unit sums its inputs in index order; the output is a wave sum.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_mlp_infer(const float *__restrict__ params, const float *__restrict__ x,
                                                   uint32_t m, float *__restrict__ y)
{
    const int o = threadIdx.x & 63;
    float w1[kIn], w2[kH];
#pragma unroll
    for (int i = 0; i < kIn; i++) w1[i] = params[kW1 + o * kIn + i];
#pragma unroll
    for (int i = 0; i < kH; i++) w2[i] = params[kW2 + o * kH + i];
    const float b1 = params[kB1 + o], b2 = params[kB2 + o], w3 = params[kW3 + o], b3 = params[kB3];
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    for (uint32_t b = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); b < m; b += waves) {
        const float xv = o < kIn ? x[(size_t)b * kIn + o] : 0.f;
        float z = b1;
#pragma unroll
        for (int i = 0; i < kIn; i++) z = fmaf(w1[i], lane_f32(xv, i), z);
        const float h1 = fmaxf(z, 0.f);
        z = b2;
#pragma unroll
        for (int i = 0; i < kH; i++) z = fmaf(w2[i], lane_f32(h1, i), z);
        const float out = wave_sum_f32(w3 * fmaxf(z, 0.f)) + b3;
        if (o == 0) y[b] = out;
    }
}

// ---------------------------------------------------------------------------
// training step: rows kernel (forward + per-row backward), weight-gradient
// partials over row chunks, fixed-order sum of the partials, Adam
// ---------------------------------------------------------------------------
constexpr int kRowsPerBlock = 4;      // one wave per row, one lane per hidden unit
constexpr int kP1 = kIn + 1, kP2 = kH + 1;   // padded LDS row strides: a lane per row, distinct banks
constexpr int kChunk = 64;            // rows per weight-gradient chunk
constexpr int kMaxParts = 64;         // weight-gradient blocks (partials) per parameter group

// Forward and per-row backward of the loss mean((y - t)^2) over rows
// b = 0..m-1 (row b is x[rows[b]], or x[b] without rows): stores the
// activations h1, h2 and the pre-ReLU gradients dz1, dz2, dz3 of each row for
// the weight gradient, the prediction, and adds the rows' loss * inv_m to
// *loss.  Wave r of the block takes row blockIdx.x * 4 + r; lane o computes
// hidden unit o of both layers (weights staged in LDS with padded rows, so a
// wave reads 64 rows of W at one column without bank conflicts).  Waves past
// the last row keep to the barriers and store nothing.
__global__ void __launch_bounds__(256) k_mlp_rows(const float *__restrict__ params, const float *__restrict__ x,
                                                  const float *__restrict__ target,
                                                  const uint32_t *__restrict__ rows, uint32_t m, float inv_m,
                                                  float *__restrict__ pred, float *__restrict__ h1s,
                                                  float *__restrict__ h2s, float *__restrict__ dz1s,
                                                  float *__restrict__ dz2s, float *__restrict__ dz3s,
                                                  float *__restrict__ loss)
{
    __shared__ float w1[kH * kP1];
    __shared__ float w2[kH * kP2];
    __shared__ float xs[kRowsPerBlock][kIn];
    __shared__ float hs[kRowsPerBlock][kH];
    __shared__ float ds[kRowsPerBlock][kH];
    __shared__ float ls[kRowsPerBlock];
    for (int a = threadIdx.x; a < kH * kIn; a += blockDim.x) w1[(a / kIn) * kP1 + a % kIn] = params[kW1 + a];
    for (int a = threadIdx.x; a < kH * kH; a += blockDim.x) w2[(a / kH) * kP2 + a % kH] = params[kW2 + a];
    const int r = threadIdx.x >> 6, o = threadIdx.x & 63;
    const uint32_t b = blockIdx.x * kRowsPerBlock + r;
    const bool live = b < m;
    const uint32_t src = live ? (rows ? rows[b] : b) : 0u;
    if (o < kIn) xs[r][o] = live ? x[(size_t)src * kIn + o] : 0.f;
    __syncthreads();
    float z = params[kB1 + o];
#pragma unroll
    for (int i = 0; i < kIn; i++) z = fmaf(w1[o * kP1 + i], xs[r][i], z);
    const float h1 = fmaxf(z, 0.f);
    hs[r][o] = h1;
    __syncthreads();
    z = params[kB2 + o];
#pragma unroll 16
    for (int i = 0; i < kH; i++) z = fmaf(w2[o * kP2 + i], hs[r][i], z);
    const float h2 = fmaxf(z, 0.f);
    const float w3 = params[kW3 + o];
    const float y = wave_sum_f32(w3 * h2) + params[kB3];
    const float e = y - (live ? target[src] : y);
    const float dz3 = 2.f * e * inv_m;
    const float dz2 = h2 > 0.f ? dz3 * w3 : 0.f;
    ds[r][o] = dz2;
    if (o == 0) ls[r] = e * e;
    __syncthreads();
    float g = 0.f;
    if (h1 > 0.f) {
#pragma unroll 16
        for (int q = 0; q < kH; q++) g = fmaf(w2[q * kP2 + o], ds[r][q], g);
    }
    if (live) {
        const size_t row = (size_t)b * kH + o;
        h1s[row] = h1;
        h2s[row] = h2;
        dz2s[row] = dz2;
        dz1s[row] = g;
        if (o == 0) {
            pred[b] = y;
            dz3s[b] = dz3;
        }
    }
    if (threadIdx.x == 0) atomicAdd(loss, (ls[0] + ls[1] + ls[2] + ls[3]) * inv_m);
}

// Weight-gradient partials.  blockIdx.y picks the parameter group (0: W1,
// b1; 1: W2, b2; 2: W3, b3); block x of gridDim.x takes the row chunks
// x, x + gridDim.x, ... (64 rows each, staged in LDS, rows past m as zeros)
// and writes its sums to part[x][parameter].  Every output is summed in a
// fixed order, so the gradient does not depend on scheduling.
__global__ void __launch_bounds__(256) k_mlp_wgrad_part(const float *__restrict__ x,
                                                        const uint32_t *__restrict__ rows, uint32_t m,
                                                        const float *__restrict__ h1s,
                                                        const float *__restrict__ h2s,
                                                        const float *__restrict__ dz1s,
                                                        const float *__restrict__ dz2s,
                                                        const float *__restrict__ dz3s, float *__restrict__ part)
{
    __shared__ float A[kChunk][kH];       // dz1 (group 0), dz2 (group 1), h2 (group 2)
    __shared__ float B[kChunk][kH + 1];   // x (group 0), h1 (group 1), dz3 in column 0 (group 2)
    constexpr int kOutMax = (kH * kH + kH + 255) / 256;   // outputs per thread, group 1
    const int grp = blockIdx.y;
    const int nout = grp == 0 ? kH * kIn + kH : grp == 1 ? kH * kH + kH : kH + 1;
    float acc[kOutMax];
#pragma unroll
    for (int k = 0; k < kOutMax; k++) acc[k] = 0.f;
    const uint32_t nchunk = (m + kChunk - 1) / kChunk;
    for (uint32_t c = blockIdx.x; c < nchunk; c += gridDim.x) {
        const uint32_t b0 = c * kChunk;
        for (int a = threadIdx.x; a < kChunk * kH; a += blockDim.x) {
            const int j = a / kH, k = a % kH;
            const uint32_t b = b0 + j;
            const bool ok = b < m;
            const size_t row = (size_t)b * kH + k;
            A[j][k] = ok ? (grp == 0 ? dz1s[row] : grp == 1 ? dz2s[row] : h2s[row]) : 0.f;
            if (grp == 1) B[j][k] = ok ? h1s[row] : 0.f;
        }
        if (grp == 0) {
            for (int a = threadIdx.x; a < kChunk * kIn; a += blockDim.x) {
                const int j = a / kIn, i = a % kIn;
                const uint32_t b = b0 + j;
                B[j][i] = b < m ? x[(size_t)(rows ? rows[b] : b) * kIn + i] : 0.f;
            }
        } else if (grp == 2) {
            for (int j = threadIdx.x; j < kChunk; j += blockDim.x) B[j][0] = b0 + j < m ? dz3s[b0 + j] : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kOutMax; k++) {
            const int q = threadIdx.x + 256 * k;
            if (q >= nout) continue;
            float g = acc[k];
            if (grp == 0) {
                if (q < kH * kIn) {
                    const int o = q / kIn, i = q % kIn;
                    for (int j = 0; j < kChunk; j++) g = fmaf(A[j][o], B[j][i], g);
                } else {
                    for (int j = 0; j < kChunk; j++) g += A[j][q - kH * kIn];
                }
            } else if (grp == 1) {
                if (q < kH * kH) {
                    const int o = q / kH, i = q % kH;
                    for (int j = 0; j < kChunk; j++) g = fmaf(A[j][o], B[j][i], g);
                } else {
                    for (int j = 0; j < kChunk; j++) g += A[j][q - kH * kH];
                }
            } else {
                if (q < kH) {
                    for (int j = 0; j < kChunk; j++) g = fmaf(B[j][0], A[j][q], g);
                } else {
                    for (int j = 0; j < kChunk; j++) g += B[j][0];
                }
            }
            acc[k] = g;
        }
        __syncthreads();
    }
    const int base = grp == 0 ? kW1 : grp == 1 ? kW2 : kW3;   // each group's parameters are contiguous
    float *out = part + (size_t)blockIdx.x * kNP + base;
#pragma unroll
    for (int k = 0; k < kOutMax; k++) {
        const int q = threadIdx.x + 256 * k;
        if (q < nout) out[q] = acc[k];
    }
}

// grad[p] = sum of the nparts partials of parameter p, in block order
__global__ void k_mlp_grad_sum(const float *__restrict__ part, uint32_t nparts, float *__restrict__ grad)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (uint32_t)kNP) return;
    float g = 0.f;
    for (uint32_t c = 0; c < nparts; c++) g += part[(size_t)c * kNP + p];
    grad[p] = g;
}

__global__ void k_adam(float *__restrict__ params, const float *__restrict__ grad, float *__restrict__ m1,
                       float *__restrict__ m2, uint32_t n, float lr, float b1, float b2, float eps, float bc1,
                       float bc2, float scale)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float g = grad[p] * scale;
    const float a = b1 * m1[p] + (1.f - b1) * g;
    const float v = b2 * m2[p] + (1.f - b2) * g * g;
    m1[p] = a;
    m2[p] = v;
    params[p] -= lr * (a / bc1) / (sqrtf(v / bc2) + eps);
}

hipError_t launch_features(hipStream_t s, const uint4 *rec, const uint4 *root, const uint4 *sched,
                           uint32_t m, float *feat)
{
    if (m == 0) return hipSuccess;
    k_features<<<(m + 255) / 256, 256, 0, s>>>(rec, root, sched, m, feat);
    return hipGetLastError();
}

hipError_t launch_mlp_forward(hipStream_t s, const float *params, const float *x, uint32_t m, float *y)
{
    if (m == 0) return hipSuccess;
    const uint32_t blocks = std::min<uint32_t>((m + 3) / 4, 2048u);   // 4 waves per block, rows strided
    k_mlp_infer<<<blocks, 256, 0, s>>>(params, x, m, y);
    return hipGetLastError();
}

__global__ void k_iota(uint32_t *__restrict__ p, uint32_t m)
{
    const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < m) p[a] = a;
}

}  // namespace nqk

// ---------------------------------------------------------------------------
// host side of the cost-model entry points
// ---------------------------------------------------------------------------
#include "nq_ctx.h"

namespace {

struct MlpBufs {
    nqk_buf::Buf<float> params, grad, m1, m2, x, y, pred, h1, h2, dz1, dz2, dz3, loss, part;
    nqk_buf::Buf<uint32_t> rows;
};

}  // namespace

extern "C" {

int nq_ctx_features(nq_ctx *c, float *feat_out)
{
    NQ_CTX_CHECK(c);
    if (!c->loaded) return nq_fail(NQ_ESTATE, "no task set loaded");
    if (!feat_out) return nq_fail(NQ_EINVAL, "null output");
    const uint64_t m = c->m;
    if (m == 0) return 0;
    nqk_buf::Buf<float> f;
    NQ_HIPCHK(f.reserve(m * nqk::kIn));
    NQ_HIPCHK(nqk::launch_features(c->stream, c->rec.p, c->root.p, c->sched.p,
                                   (uint32_t)m, f.p));
    NQ_HIPCHK(hipMemcpyAsync(feat_out, f.p, m * nqk::kIn * sizeof(float), hipMemcpyDeviceToHost,
                             c->stream));
    NQ_HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int nq_ctx_order_by_cost(nq_ctx *c, const float *params, uint32_t *order_out)
{
    NQ_CTX_CHECK(c);
    if (!c->loaded) return nq_fail(NQ_ESTATE, "no task set loaded");
    if (!params) return nq_fail(NQ_EINVAL, "null parameters");
    const uint32_t m = (uint32_t)c->m;
    if (m == 0) return 0;
    hipStream_t s = c->stream;
    NQ_HIPCHK(c->cm_params.reserve(nqk::kNP));
    NQ_HIPCHK(c->cm_x.reserve((size_t)m * nqk::kIn));
    NQ_HIPCHK(c->cm_y.reserve(m));
    NQ_HIPCHK(c->cm_ys.reserve(m));
    NQ_HIPCHK(c->cm_idx.reserve(m));
    NQ_HIPCHK(hipMemcpyAsync(c->cm_params.p, params, nqk::kNP * sizeof(float), hipMemcpyHostToDevice, s));
    NQ_HIPCHK(nqk::launch_features(s, c->rec.p, c->root.p, c->sched.p, m, c->cm_x.p));
    NQ_HIPCHK(nqk::launch_mlp_forward(s, c->cm_params.p, c->cm_x.p, m, c->cm_y.p));
    nqk::k_iota<<<(m + 255) / 256, 256, 0, s>>>(c->cm_idx.p, m);
    NQ_HIPCHK(hipGetLastError());
    // predicted heaviest first; ties keep record order (the radix sort is stable)
    size_t tmp = 0;
    NQ_HIPCHK(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tmp, c->cm_y.p, c->cm_ys.p, c->cm_idx.p,
                                                           c->order.p, (int)m, 0, 32, s));
    NQ_HIPCHK(c->cm_tmp.reserve(tmp ? tmp : 1));
    NQ_HIPCHK(hipcub::DeviceRadixSort::SortPairsDescending(c->cm_tmp.p, tmp, c->cm_y.p, c->cm_ys.p, c->cm_idx.p,
                                                           c->order.p, (int)m, 0, 32, s));
    if (order_out) NQ_HIPCHK(hipMemcpyAsync(order_out, c->order.p, (size_t)m * 4, hipMemcpyDeviceToHost, s));
    NQ_HIPCHK(hipStreamSynchronize(s));
    return 0;
}

int nq_mlp_forward(nq_ctx *c, const float *params, const float *x, uint64_t m, float *y_out)
{
    NQ_CTX_CHECK(c);
    if (!params || (m && (!x || !y_out))) return nq_fail(NQ_EINVAL, "null pointer");
    if (m >= (1ull << 31)) return nq_fail(NQ_EINVAL, "too many rows");
    if (m == 0) return 0;
    hipStream_t s = c->stream;
    nqk_buf::Buf<float> P, X, Y;
    NQ_HIPCHK(P.reserve(nqk::kNP));
    NQ_HIPCHK(X.reserve(m * nqk::kIn));
    NQ_HIPCHK(Y.reserve(m));
    NQ_HIPCHK(hipMemcpyAsync(P.p, params, nqk::kNP * sizeof(float), hipMemcpyHostToDevice, s));
    NQ_HIPCHK(hipMemcpyAsync(X.p, x, m * nqk::kIn * sizeof(float), hipMemcpyHostToDevice, s));
    NQ_HIPCHK(nqk::launch_mlp_forward(s, P.p, X.p, (uint32_t)m, Y.p));
    NQ_HIPCHK(hipMemcpyAsync(y_out, Y.p, m * sizeof(float), hipMemcpyDeviceToHost, s));
    NQ_HIPCHK(hipStreamSynchronize(s));
    return 0;
}

namespace {

// one minibatch: forward and per-row backward, weight-gradient partials, their
// fixed-order sum (into B.grad); the loss is added to B.loss
int batch_grad(hipStream_t s, MlpBufs &B, const uint32_t *rows, uint32_t nb, float inv_m)
{
    using namespace nqk;
    NQ_HIPCHK(B.part.reserve((size_t)kMaxParts * kNP));
    k_mlp_rows<<<(nb + kRowsPerBlock - 1) / kRowsPerBlock, 64 * kRowsPerBlock, 0, s>>>(
        B.params.p, B.x.p, B.y.p, rows, nb, inv_m, B.pred.p, B.h1.p, B.h2.p, B.dz1.p, B.dz2.p, B.dz3.p, B.loss.p);
    NQ_HIPCHK(hipGetLastError());
    const uint32_t nparts = std::min<uint32_t>((nb + kChunk - 1) / kChunk, kMaxParts);
    k_mlp_wgrad_part<<<dim3(nparts, 3), 256, 0, s>>>(B.x.p, rows, nb, B.h1.p, B.h2.p, B.dz1.p, B.dz2.p, B.dz3.p,
                                                     B.part.p);
    NQ_HIPCHK(hipGetLastError());
    k_mlp_grad_sum<<<(kNP + 255) / 256, 256, 0, s>>>(B.part.p, nparts, B.grad.p);
    NQ_HIPCHK(hipGetLastError());
    return 0;
}

int stage_xy(hipStream_t s, MlpBufs &B, const float *params, const float *x, const float *y, uint64_t m,
             uint32_t rows_cap)
{
    NQ_HIPCHK(B.params.reserve(nqk::kNP));
    NQ_HIPCHK(B.grad.reserve(nqk::kNP));
    NQ_HIPCHK(B.x.reserve(m * nqk::kIn));
    NQ_HIPCHK(B.y.reserve(m));
    NQ_HIPCHK(B.pred.reserve(rows_cap));
    NQ_HIPCHK(B.h1.reserve((size_t)rows_cap * nqk::kH));
    NQ_HIPCHK(B.h2.reserve((size_t)rows_cap * nqk::kH));
    NQ_HIPCHK(B.dz1.reserve((size_t)rows_cap * nqk::kH));
    NQ_HIPCHK(B.dz2.reserve((size_t)rows_cap * nqk::kH));
    NQ_HIPCHK(B.dz3.reserve(rows_cap));
    NQ_HIPCHK(B.loss.reserve(1));
    NQ_HIPCHK(hipMemcpyAsync(B.params.p, params, nqk::kNP * sizeof(float), hipMemcpyHostToDevice, s));
    NQ_HIPCHK(hipMemcpyAsync(B.x.p, x, m * nqk::kIn * sizeof(float), hipMemcpyHostToDevice, s));
    NQ_HIPCHK(hipMemcpyAsync(B.y.p, y, m * sizeof(float), hipMemcpyHostToDevice, s));
    return 0;
}

}  // namespace

int nq_mlp_loss_grad(nq_ctx *c, const float *params, const float *x, const float *y, uint64_t m, float *loss_out,
                     float *grad_out)
{
    NQ_CTX_CHECK(c);
    if (!params || !loss_out || !grad_out || (m && (!x || !y))) return nq_fail(NQ_EINVAL, "null pointer");
    if (m == 0 || m >= (1ull << 24)) return nq_fail(NQ_EINVAL, "rows must be in 1..2^24");
    hipStream_t s = c->stream;
    MlpBufs B;
    if (int r = stage_xy(s, B, params, x, y, m, (uint32_t)m)) return r;
    NQ_HIPCHK(hipMemsetAsync(B.loss.p, 0, sizeof(float), s));
    if (int r = batch_grad(s, B, nullptr, (uint32_t)m, 1.0f / (float)m)) return r;
    NQ_HIPCHK(hipMemcpyAsync(grad_out, B.grad.p, nqk::kNP * sizeof(float), hipMemcpyDeviceToHost, s));
    NQ_HIPCHK(hipMemcpyAsync(loss_out, B.loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
    NQ_HIPCHK(hipStreamSynchronize(s));
    return 0;
}

int nq_mlp_train(nq_ctx *c, float *params, const float *x, const float *y, uint64_t m, const uint32_t *perm,
                 int epochs, int batch, float lr, float *loss_out)
{
    NQ_CTX_CHECK(c);
    if (!params || !x || !y || !perm || !loss_out) return nq_fail(NQ_EINVAL, "null pointer");
    if (m == 0 || m >= (1ull << 24)) return nq_fail(NQ_EINVAL, "rows must be in 1..2^24");
    if (epochs < 1 || batch < 1 || batch > 65536) return nq_fail(NQ_EINVAL, "bad epochs/batch");
    for (uint64_t a = 0; a < m * (uint64_t)epochs; a++)
        if (perm[a] >= m) return nq_fail(NQ_EINVAL, "perm entry out of range");
    hipStream_t s = c->stream;
    MlpBufs B;
    if (int r = stage_xy(s, B, params, x, y, m, (uint32_t)batch)) return r;
    NQ_HIPCHK(B.m1.reserve(nqk::kNP));
    NQ_HIPCHK(B.m2.reserve(nqk::kNP));
    NQ_HIPCHK(B.rows.reserve(m * (uint64_t)epochs));
    NQ_HIPCHK(hipMemsetAsync(B.m1.p, 0, nqk::kNP * sizeof(float), s));
    NQ_HIPCHK(hipMemsetAsync(B.m2.p, 0, nqk::kNP * sizeof(float), s));
    NQ_HIPCHK(hipMemcpyAsync(B.rows.p, perm, m * (uint64_t)epochs * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    double p1 = 1.0, p2 = 1.0;
    for (int e = 0; e < epochs; e++) {
        NQ_HIPCHK(hipMemsetAsync(B.loss.p, 0, sizeof(float), s));
        const uint32_t *rows = B.rows.p + (uint64_t)e * m;
        uint32_t nbatch = 0;
        for (uint64_t off = 0; off < m; off += (uint64_t)batch) {
            const uint32_t nb = (uint32_t)std::min<uint64_t>((uint64_t)batch, m - off);
            if (int r = batch_grad(s, B, rows + off, nb, 1.0f / (float)nb)) return r;
            p1 *= b1;
            p2 *= b2;
            nqk::k_adam<<<(nqk::kNP + 255) / 256, 256, 0, s>>>(B.params.p, B.grad.p, B.m1.p, B.m2.p, nqk::kNP, lr,
                                                              b1, b2, eps, (float)(1.0 - p1), (float)(1.0 - p2),
                                                              1.0f);
            NQ_HIPCHK(hipGetLastError());
            nbatch++;
        }
        float l = 0.f;
        NQ_HIPCHK(hipMemcpyAsync(&l, B.loss.p, sizeof(float), hipMemcpyDeviceToHost, s));
        NQ_HIPCHK(hipStreamSynchronize(s));
        loss_out[e] = l / (float)nbatch;
    }
    NQ_HIPCHK(hipMemcpyAsync(params, B.params.p, nqk::kNP * sizeof(float), hipMemcpyDeviceToHost, s));
    NQ_HIPCHK(hipStreamSynchronize(s));
    return 0;
}

}  // extern "C"
