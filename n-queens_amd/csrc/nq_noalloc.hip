// nq_noalloc.hip -- nq_count_noalloc: the one-shot count of nq_count with
// every device buffer carved out of one file-scope static device array, so
// that the call itself allocates nothing (for callers that cannot allocate:
// code under graph capture, tight serving loops, harnesses that forbid it).
//
// The array (`scratch`, kScratchBytes) is part of the library's code object:
// each device the library runs on holds its own copy, at its own address,
// until the process ends.  Host code never uses the array's name as an
// address (on the host it is only a stand-in for the device symbol): the
// address is looked up once per device with hipGetSymbolAddress when a
// context is created (nq_scratch_attach), and that pointer is what the
// kernels and copies receive.  Every call on a device shares the array and
// leaves its contents behind; each call clears what must start at zero, and
// calls that overlap on one device race.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <mutex>

#include "nq_task.h"
#include "nq_kernels.h"
#include "nq_ctx.h"

namespace {

// 128 MiB: descriptors of up to kNoAllocMaxRecords records, the rest work items.
constexpr size_t kScratchBytes = size_t(128) << 20;
constexpr uint64_t kNoAllocMaxRecords = 1u << 18;
constexpr int kMaxDevices = 64;

__device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_attach_lock;
void *g_scratch_ptr[kMaxDevices] = {};

// Lowest and highest terminal depth of the decoded set: tr[0] = min, tr[1] = max
// (tr[0] starts at ~0, tr[1] at 0).
__global__ void k_trange(const uint4 *__restrict__ sched, uint32_t m, uint32_t *__restrict__ tr)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= m) return;
    const uint32_t T = sched[nqk::kSchedU4 * c + 1].x & 0xFFu;
    atomicMin(tr, T);
    atomicMax(tr + 1, T);
}

// Consecutive 256-byte-aligned views into the scratch array (device addresses).
struct Carve {
    uintptr_t base, off = 0, end;
    template <class T>
    T *take(size_t n)
    {
        off = (off + 255u) & ~uintptr_t(255);
        T *p = (T *)(base + off);
        off += n * sizeof(T);
        return p;
    }
    bool fits() const { return off <= end; }
};

}  // namespace

int nq_scratch_attach(int device, void **ptr)
{
    *ptr = nullptr;
    if (device < 0 || device >= kMaxDevices)
        return nq_fail(NQ_EUNSUPPORTED, "device %d beyond the %d devices the static scratch table holds", device,
                       kMaxDevices);
    std::lock_guard<std::mutex> g(g_attach_lock);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        NQ_HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    return 0;
}

extern "C" int nq_count_noalloc(nq_ctx *c, int N, const uint32_t *rec4, uint64_t m, const nq_sched *sched,
                                uint64_t *weighted_out, uint64_t *nodes_out, uint64_t *total_out)
{
    NQ_CTX_CHECK(c);
    if (!c->scratch) return nq_fail(NQ_ESTATE, "context has no static scratch array");
    if (N < nqt::kMinN || N > nqt::kMaxN)
        return nq_fail(NQ_EINVAL, "N=%d outside the u32 path (%d..%d)", N, nqt::kMinN, nqt::kMaxN);
    if (m && !rec4) return nq_fail(NQ_EINVAL, "null record pointer");
    if (m > kNoAllocMaxRecords)
        return nq_fail(NQ_ECAPACITY, "%llu records: the static scratch array holds at most %llu",
                       (unsigned long long)m, (unsigned long long)kNoAllocMaxRecords);
    if (sched && sched->split_depth == -1)
        return nq_fail(NQ_EINVAL, "per-record split depths are not available without a loaded set");
    if (m == 0) {
        if (total_out) *total_out = 0;
        return 0;
    }
    hipStream_t st = c->stream;
    const uint32_t mu = (uint32_t)m;
    Carve cv{(uintptr_t)c->scratch, 0, kScratchBytes};
    unsigned long long *total = cv.take<unsigned long long>(1);
    unsigned long long *item_count = cv.take<unsigned long long>(1);
    uint32_t *small = cv.take<uint32_t>(4);   // decode status, queue head, T min, T max
    uint4 *rec = cv.take<uint4>(m);
    nq_run r;
    r.m = mu;
    uint4 *root = cv.take<uint4>(m), *sch = cv.take<uint4>(nqk::kSchedU4 * m);
    r.root = root;
    r.sched = sch;
    r.fan = 1u;   // one expansion lane per record: per-record count arrays leave the array's room to work items
    r.nitems = cv.take<uint32_t>(m * r.fan);
    r.off = cv.take<uint32_t>(m * r.fan);
    r.scan_tmp = cv.take<unsigned long long>(m * r.fan / 1024 + 1);
    r.raw = cv.take<unsigned long long>(m);
    r.weighted = cv.take<unsigned long long>(m);
    r.nodes = cv.take<unsigned long long>(m);
    r.total = total;
    r.item_count = item_count;
    r.head = small + 1;
    cv.off = (cv.off + 255u) & ~uintptr_t(255);
    const uint64_t max_items = cv.off < cv.end ? (cv.end - cv.off) / (16u * nqk::kItemU4) : 0u;
    r.items = (uint4 *)(cv.base + cv.off);
    if (!cv.fits()) return nq_fail(NQ_ECAPACITY, "record views exceed the static scratch array");

    NQ_HIPCHK(hipMemcpyAsync(rec, rec4, m * 16, hipMemcpyHostToDevice, st));
    NQ_HIPCHK(nqk::launch_fill_u32(st, small, 1, 0));
    NQ_HIPCHK(nqk::launch_fill_u32(st, small + 2, 1, ~0u));
    NQ_HIPCHK(nqk::launch_fill_u32(st, small + 3, 1, 0));
    NQ_HIPCHK(nqk::launch_decode(st, N, nullptr, rec, mu, root, sch, small));
    k_trange<<<(mu + 255) / 256, 256, 0, st>>>(sch, mu, small + 2);
    NQ_HIPCHK(hipGetLastError());
    uint32_t hs[4] = {0, 0, 0, 0};
    NQ_HIPCHK(hipMemcpyAsync(hs, small, sizeof hs, hipMemcpyDeviceToHost, st));
    NQ_HIPCHK(hipStreamSynchronize(st));
    if (hs[0]) return nq_fail(NQ_EUNSUPPORTED, "record decode failed (status mask 0x%x: 2=P5 record past its switch row, 4=shift, "
                                               "8=no terminal row, 16=depth)", hs[0]);
    r.Tmin = hs[2];
    r.Tmax = hs[3];
    if (r.Tmax > 16u) return nq_fail(NQ_EUNSUPPORTED, "terminal depth %u > 16 (preset too small for N)", r.Tmax);
    const int nodes = nodes_out != nullptr;
    if (!sched) {
        const nqk::WalkLds lay = nqk::walk_lds(r.Tmax, 0u);
        if (lay.bytes(64) > (size_t)c->lds_max) return nq_fail(NQ_ECAPACITY, "LDS stack too large");
        NQ_HIPCHK(nqk::launch_fill_u64(st, total, 1, 0));
        if (!nodes) NQ_HIPCHK(nqk::launch_fill_u64(st, r.nodes, m, 0));
        NQ_HIPCHK(nqk::launch_plain(st, nodes != 0, 64, lay, root, sch, mu, r.weighted, r.nodes, total));
    } else {
        nq_sched s = *sched;
        s.use_graph = 0;   // eager launches: instantiating a graph would allocate
        if (int e = nq_prepare_sched(r, c->lds_max, &s)) return e;
        uint64_t n_items = 0;
        if (int e = nq_count_items(st, r, s.split_depth, s.block, &n_items)) return e;
        if (n_items > max_items)
            return nq_fail(NQ_ECAPACITY, "split depth %d gives %llu work items: the static scratch array holds %llu "
                           "(use a shallower split)", s.split_depth, (unsigned long long)n_items,
                           (unsigned long long)max_items);
        const size_t lds_bytes = s.lds_stack ? nq_batched_lds(r, s.split_depth).bytes(s.block) : 0u;
        int grid = 0;
        if (int e = nq_batched_grid(c->cu, s, lds_bytes, nodes, n_items, &grid)) return e;
        if (int e = nq_enqueue_batched(st, r, s, nodes, grid, nullptr)) return e;
    }
    unsigned long long tot = 0;
    NQ_HIPCHK(hipMemcpyAsync(&tot, total, 8, hipMemcpyDeviceToHost, st));
    if (weighted_out) NQ_HIPCHK(hipMemcpyAsync(weighted_out, r.weighted, m * 8, hipMemcpyDeviceToHost, st));
    if (nodes_out) NQ_HIPCHK(hipMemcpyAsync(nodes_out, r.nodes, m * 8, hipMemcpyDeviceToHost, st));
    NQ_HIPCHK(hipStreamSynchronize(st));
    if (total_out) *total_out = tot;
    return 0;
}
