// nq_probe.hip -- nq_issue_probe: the VALU issue rate of this device on a
// stream of independent 32-bit v_or_b32 / v_and_b32 ops, measured at a
// chosen number of waves per SIMD.  The published INT32 vector peak counts
// two lane-ops per lane per clock; this is what plain bitwise code of the
// DFS step's kind actually issues, the ceiling its kernels are compared with.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "nq_ctx.h"

namespace {

constexpr uint32_t kProbeOpsPerIter = 32;   // per lane: 8 chains x (or, and, or, and)
constexpr uint32_t kProbeOr = 0x01234567u, kProbeAnd = 0xFEDCBA98u;

// One wave per workgroup.  Lane g runs 8 chains a_i = ((a_i | k_i) & m_i)
// twice per iteration, a_i starting at g + i, k_i = kProbeOr << i,
// m_i = kProbeAnd >> i; the 32 ops of an iteration are one asm block, so the
// compiler neither folds the chains nor schedules anything between them, and
// consecutive ops belong to different chains.  The lane stores the XOR of its
// chains (the host checks it).
__global__ void __launch_bounds__(64) k_issue_probe(uint32_t iters, uint32_t *__restrict__ out)
{
    const uint32_t g = blockIdx.x * 64u + threadIdx.x;
    uint32_t a0 = g, a1 = g + 1u, a2 = g + 2u, a3 = g + 3u, a4 = g + 4u, a5 = g + 5u, a6 = g + 6u, a7 = g + 7u;
    const uint32_t k0 = kProbeOr, k1 = kProbeOr << 1, k2 = kProbeOr << 2, k3 = kProbeOr << 3, k4 = kProbeOr << 4,
                   k5 = kProbeOr << 5, k6 = kProbeOr << 6, k7 = kProbeOr << 7;
    const uint32_t m0 = kProbeAnd, m1 = kProbeAnd >> 1, m2 = kProbeAnd >> 2, m3 = kProbeAnd >> 3,
                   m4 = kProbeAnd >> 4, m5 = kProbeAnd >> 5, m6 = kProbeAnd >> 6, m7 = kProbeAnd >> 7;
    for (uint32_t i = 0; i < iters; i++) {
        asm volatile(
            "v_or_b32 %0, %0, %8\n\tv_or_b32 %1, %1, %9\n\tv_or_b32 %2, %2, %10\n\tv_or_b32 %3, %3, %11\n\t"
            "v_or_b32 %4, %4, %12\n\tv_or_b32 %5, %5, %13\n\tv_or_b32 %6, %6, %14\n\tv_or_b32 %7, %7, %15\n\t"
            "v_and_b32 %0, %0, %16\n\tv_and_b32 %1, %1, %17\n\tv_and_b32 %2, %2, %18\n\tv_and_b32 %3, %3, %19\n\t"
            "v_and_b32 %4, %4, %20\n\tv_and_b32 %5, %5, %21\n\tv_and_b32 %6, %6, %22\n\tv_and_b32 %7, %7, %23\n\t"
            "v_or_b32 %0, %0, %8\n\tv_or_b32 %1, %1, %9\n\tv_or_b32 %2, %2, %10\n\tv_or_b32 %3, %3, %11\n\t"
            "v_or_b32 %4, %4, %12\n\tv_or_b32 %5, %5, %13\n\tv_or_b32 %6, %6, %14\n\tv_or_b32 %7, %7, %15\n\t"
            "v_and_b32 %0, %0, %16\n\tv_and_b32 %1, %1, %17\n\tv_and_b32 %2, %2, %18\n\tv_and_b32 %3, %3, %19\n\t"
            "v_and_b32 %4, %4, %20\n\tv_and_b32 %5, %5, %21\n\tv_and_b32 %6, %6, %22\n\tv_and_b32 %7, %7, %23"
            : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7)
            : "v"(k0), "v"(k1), "v"(k2), "v"(k3), "v"(k4), "v"(k5), "v"(k6), "v"(k7), "v"(m0), "v"(m1), "v"(m2),
              "v"(m3), "v"(m4), "v"(m5), "v"(m6), "v"(m7));
    }
    out[g] = a0 ^ a1 ^ a2 ^ a3 ^ a4 ^ a5 ^ a6 ^ a7;
}

}  // namespace

extern "C" int nq_issue_probe(nq_ctx *c, int waves_per_simd, uint32_t iters, double *ms_out, double *clk_out,
                              uint32_t *xor_out)
{
    NQ_CTX_CHECK(c);
    if (waves_per_simd < 1 || waves_per_simd > 8)
        return nq_fail(NQ_EINVAL, "waves per SIMD %d outside 1..8", waves_per_simd);
    if (iters < 1u || iters > (1u << 24)) return nq_fail(NQ_EINVAL, "iterations %u outside 1..2^24", iters);
    // four SIMDs per CU; one wave per workgroup, all resident at once
    const uint32_t waves = (uint32_t)c->cu * 4u * (uint32_t)waves_per_simd;
    nqk_buf::Buf<uint32_t> out;
    NQ_HIPCHK(out.reserve((size_t)waves * 64u));
    hipStream_t st = c->stream;
    k_issue_probe<<<waves, 64, 0, st>>>(1u, out.p);   // warm-up: code and wave launch paths
    NQ_HIPCHK(hipGetLastError());
    NQ_HIPCHK(hipEventRecord(c->ev[0], st));
    k_issue_probe<<<waves, 64, 0, st>>>(iters, out.p);
    NQ_HIPCHK(hipGetLastError());
    NQ_HIPCHK(hipEventRecord(c->ev[1], st));
    std::vector<uint32_t> h((size_t)waves * 64u);
    NQ_HIPCHK(hipMemcpyAsync(h.data(), out.p, h.size() * 4u, hipMemcpyDeviceToHost, st));
    NQ_HIPCHK(hipStreamSynchronize(st));
    float ms = 0.f;
    NQ_HIPCHK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    uint32_t x = 0;
    for (uint32_t v : h) x ^= v;
    // wave instructions per SIMD, and clocks per wave instruction per SIMD
    const double per_simd = (double)waves_per_simd * iters * kProbeOpsPerIter;
    if (ms_out) *ms_out = ms;
    if (clk_out) *clk_out = (double)ms * 1e-3 * (double)c->clock_khz * 1e3 / per_simd;
    if (xor_out) *xor_out = x;
    return 0;
}
