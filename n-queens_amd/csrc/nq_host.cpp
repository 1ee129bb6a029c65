// nq_host.cpp -- the C ABI of libnq (declared in include/nq.h): contexts,
// resident task sets, the plain and batched count runs, Unique support and
// the host-side decode used by the CPU tests.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdarg.h>
#include <string.h>
#include <string>
#include <vector>
#include <algorithm>

#define NQ_HD __host__ __device__
#include "nq_task.h"
#include "nq_kernels.h"

#include "nq_ctx.h"

namespace {
thread_local std::string g_err;
}  // namespace

int nq_fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int nq_check_ctx(nq_ctx *c)
{
    if (!c) return nq_fail(NQ_EINVAL, "null context");
    NQ_HIPCHK(hipSetDevice(c->device));
    return 0;
}

namespace {

int check_ctx(nq_ctx *c) { return nq_check_ctx(c); }

void drop_graph(nq_ctx *c)
{
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->gexec = nullptr;
    c->graph = nullptr;
    c->gkey = nq_graph_key{};
}

// The resident set of a context as the runs read it.
nq_run ctx_run(nq_ctx *c)
{
    nq_run r;
    r.m = (uint32_t)c->m;
    r.fan = c->fan;
    r.Tmax = c->Tmax;
    r.Tmin = c->Tmin;
    r.split_min = c->split_min;
    r.root = c->root.p;
    r.sched = c->sched.p;
    r.order = c->order.p;
    r.splits = c->has_splits ? c->splits.p : nullptr;
    r.nitems = c->nitems.p;
    r.off = c->off.p;
    r.head = c->head.p;
    r.items = c->items.p;
    r.raw = c->raw.p;
    r.nodes = c->nodes.p;
    r.weighted = c->weighted.p;
    r.total = c->total.p;
    r.item_count = c->item_count.p;
    r.scan_tmp = c->scan_tmp.p;
    return r;
}

// Stack slots per lane of the expansion DFS (frames of depths 0..T-2).
uint32_t expand_slots(const nq_run &r)
{
    return r.Tmax > 1u ? r.Tmax : 1u;
}

int elapsed(nq_ctx *c, int a, int b, double *ms)
{
    float f = 0.f;
    NQ_HIPCHK(hipEventElapsedTime(&f, c->ev[a], c->ev[b]));
    *ms = f;
    return 0;
}

// Item count of the split at `split`; sizes the context's item buffer.
int size_items(nq_ctx *c, int split, int block)
{
    uint64_t cnt = 0;
    if (int r = nq_count_items(c->stream, ctx_run(c), split, block, &cnt)) return r;
    // item indices and offsets are u32 (the queue head's claims past the end
    // are checked against the launch in nq_ctx_run_batched)
    if (cnt >= (1ull << 32)) return nq_fail(NQ_ECAPACITY, "split depth %d gives %llu work items (limit 2^32)", split, cnt);
    NQ_HIPCHK(c->items.reserve(nqk::kItemU4 * cnt));
    c->item_total = cnt;
    c->sized_split = split;
    return 0;
}

}  // namespace

// LDS layout of the batched walk: a walk from depth d0 pushes at most
// T-1-d0 frames, so the shallowest item start depth over the set decides the
// stack size (with the deepest T).  For per-record depths that minimum is
// computed per record by nq_ctx_set_splits.
nqk::WalkLds nq_batched_lds(const nq_run &r, int split)
{
    const uint32_t d0 = split < 0 ? r.split_min : nqk::item_depth((uint32_t)split, r.Tmin);
    return nqk::walk_lds(r.Tmax, d0);
}

int nq_prepare_sched(const nq_run &r, int lds_max, nq_sched *s)
{
    if (s->block != 64 && s->block != 128 && s->block != 256) return nq_fail(NQ_EINVAL, "block must be 64, 128 or 256");
    if (s->chunk < 1) s->chunk = 64;
    if (s->split_depth < -1) return nq_fail(NQ_EINVAL, "split depth must be >= 0, or -1 for per-record depths");
    if (s->split_depth == -1 && !r.splits) return nq_fail(NQ_ESTATE, "no per-record split depths set");
    // items live at depth <= T-2 (a deeper split falls back to the root item)
    const int max_split = r.Tmax >= 2 ? (int)r.Tmax - 2 : 0;
    if (s->split_depth > max_split) s->split_depth = max_split;
    const nqk::WalkLds lay = nq_batched_lds(r, s->split_depth);
    const size_t lds_bytes = s->lds_stack ? lay.bytes(s->block) : 0u;
    if (lds_bytes > (size_t)lds_max) return nq_fail(NQ_ECAPACITY, "LDS stack %zu B > %d B", lds_bytes, lds_max);
    return 0;
}

int nq_batched_grid(int cu, const nq_sched &s, size_t lds_bytes, int nodes, uint64_t items, int *grid)
{
    int per_cu = 0;
    NQ_HIPCHK(nqk::batched_occupancy(nodes != 0, s.lds_stack != 0, s.block, lds_bytes, &per_cu));
    if (per_cu < 1) per_cu = 1;
    if (s.waves_per_eu > 0) {
        int want = (s.waves_per_eu * 4 * 64) / s.block;   // 4 SIMDs per CU
        if (want < 1) want = 1;
        per_cu = std::min(per_cu, want);
    }
    *grid = cu * per_cu;
    // the u32 queue head ends at most one chunk per wave past the item count
    if (items + (uint64_t)*grid * (s.block / 64) * (uint64_t)s.chunk >= (1ull << 32))
        return nq_fail(NQ_ECAPACITY, "%llu work items and %d claims of %d past the end overflow the u32 queue head",
                       (unsigned long long)items, *grid * (s.block / 64), s.chunk);
    return 0;
}

int nq_count_items(hipStream_t st, const nq_run &r, int split, int block, uint64_t *count)
{
    NQ_HIPCHK(nqk::launch_expand(st, false, false, block, expand_slots(r), r.root, r.sched, r.order,
                                 split < 0 ? r.splits : nullptr, r.m, (uint32_t)(split < 0 ? 0 : split), r.fan,
                                 r.nitems, nullptr, nullptr, nullptr, nullptr));
    NQ_HIPCHK(nqk::launch_scan(st, r.nitems, r.m * r.fan, r.off, r.item_count, r.scan_tmp, r.order, r.fan));
    unsigned long long cnt = 0;
    NQ_HIPCHK(hipMemcpyAsync(&cnt, r.item_count, sizeof cnt, hipMemcpyDeviceToHost, st));
    NQ_HIPCHK(hipStreamSynchronize(st));
    *count = cnt;
    return 0;
}

// Every run writes the whole expansion from the resident records.  The item
// counts per (record, root candidate) are those of the sizing pass
// (nq_count_items), kept by record with the resident set for its split depth;
// the run scans them in its own processing order.
int nq_enqueue_batched(hipStream_t st, const nq_run &r, const nq_sched &s, int nodes, int grid, hipEvent_t *ev)
{
    const uint32_t m = r.m, sd = expand_slots(r);
    const nqk::WalkLds lay = nq_batched_lds(r, s.split_depth);
    if (ev) NQ_HIPCHK(hipEventRecord(ev[0], st));
    NQ_HIPCHK(nqk::launch_fill_u64(st, r.raw, m, 0));
    NQ_HIPCHK(nqk::launch_fill_u64(st, r.total, 1, 0));
    NQ_HIPCHK(nqk::launch_fill_u32(st, r.head, 1, 0));
    NQ_HIPCHK(nqk::launch_fill_u64(st, r.nodes, m, 0));
    const uint8_t *sp8 = s.split_depth < 0 ? r.splits : nullptr;
    const uint32_t su = (uint32_t)(s.split_depth < 0 ? 0 : s.split_depth);
    NQ_HIPCHK(nqk::launch_scan(st, r.nitems, m * r.fan, r.off, r.item_count, r.scan_tmp, r.order, r.fan));
    NQ_HIPCHK(nqk::launch_expand(st, nodes != 0, true, s.block, sd, r.root, r.sched, r.order, sp8, m,
                                 su, r.fan, r.nitems, r.off, r.items, r.raw, r.nodes));
    if (ev) NQ_HIPCHK(hipEventRecord(ev[1], st));
    NQ_HIPCHK(nqk::launch_batched(st, nodes != 0, s.lds_stack != 0, grid, s.block, lay, r.items,
                                  r.item_count, r.head, (uint32_t)s.chunk, r.raw, r.nodes));
    if (ev) NQ_HIPCHK(hipEventRecord(ev[2], st));
    NQ_HIPCHK(nqk::launch_finalize(st, r.sched, m, r.raw, r.weighted, r.total));
    if (ev) NQ_HIPCHK(hipEventRecord(ev[3], st));
    return 0;
}

extern "C" {

int nq_abi_version(void) { return NQ_ABI_VERSION; }

const char *nq_last_error(void) { return g_err.c_str(); }

namespace {

// Device properties, stream, events and the small fixed buffers of a new
// context (the device is current).
int init_ctx(nq_ctx *c)
{
    hipDeviceProp_t prop;
    NQ_HIPCHK(hipGetDeviceProperties(&prop, c->device));
    c->cu = prop.multiProcessorCount;
    c->lds_max = (int)prop.sharedMemPerBlock;
    c->wave = prop.warpSize;
    c->clock_khz = prop.clockRate;
    if (c->wave != 64) return nq_fail(NQ_EUNSUPPORTED, "wavefront size %d, kernels assume 64", prop.warpSize);
    NQ_HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto &e : c->ev) NQ_HIPCHK(hipEventCreate(&e));
    NQ_HIPCHK(c->total.reserve(1));
    NQ_HIPCHK(c->err.reserve(1));
    NQ_HIPCHK(c->head.reserve(1));
    NQ_HIPCHK(c->item_count.reserve(1));
    NQ_HIPCHK(c->sym.reserve(2));
    // the device address of this device's static scratch array (nq_count_noalloc)
    if (int r = nq_scratch_attach(c->device, &c->scratch)) return r;
    return 0;
}

}  // namespace

int nq_ctx_create(int device, nq_ctx **out)
{
    if (!out) return nq_fail(NQ_EINVAL, "null out pointer");
    *out = nullptr;
    int ndev = 0;
    NQ_HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return nq_fail(NQ_EINVAL, "device %d out of range (%d devices)", device, ndev);
    NQ_HIPCHK(hipSetDevice(device));
    nq_ctx *c = new nq_ctx;
    c->device = device;
    if (int r = init_ctx(c)) {
        nq_ctx_destroy(c);   // releases whatever was created; keeps the message of the failure
        return r;
    }
    *out = c;
    return 0;
}

int nq_ctx_destroy(nq_ctx *c)
{
    if (!c) return 0;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graph(c);
    for (auto &e : c->ev) if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

int nq_device_info(nq_ctx *c, int32_t *out4)
{
    if (int r = check_ctx(c)) return r;
    if (!out4) return nq_fail(NQ_EINVAL, "null out pointer");
    out4[0] = c->cu; out4[1] = c->lds_max; out4[2] = c->wave; out4[3] = c->clock_khz;
    return 0;
}

namespace {

// Expands the m records at dev[0] by `levels` preset queens on the device
// (two call-owned buffers, ping-pong); the result ends in dev[*cur], its
// count in *m_io.
int expand_device(nq_ctx *c, int N, nqk_buf::Buf<uint4> dev[2], int *cur, uint64_t *m_io, int levels)
{
    nqk_buf::Buf<uint32_t> cnt, off;
    nqk_buf::Buf<unsigned long long> total, tsum;
    NQ_HIPCHK(total.reserve(1));
    uint64_t m = *m_io;
    int a = 0;
    for (int lv = 0; lv < levels && m; lv++) {
        NQ_HIPCHK(cnt.reserve(m));
        NQ_HIPCHK(off.reserve(m));
        NQ_HIPCHK(tsum.reserve(m / 1024 + 1));
        NQ_HIPCHK(nqk::launch_fill_u32(c->stream, c->err.p, 1, 0));
        NQ_HIPCHK(nqk::launch_preset(c->stream, false, N, dev[a].p, (uint32_t)m, cnt.p, nullptr, nullptr, c->err.p));
        NQ_HIPCHK(nqk::launch_scan(c->stream, cnt.p, (uint32_t)m, off.p, total.p, tsum.p));
        unsigned long long n = 0;
        uint32_t err = 0;
        NQ_HIPCHK(hipMemcpyAsync(&n, total.p, 8, hipMemcpyDeviceToHost, c->stream));
        NQ_HIPCHK(hipMemcpyAsync(&err, c->err.p, 4, hipMemcpyDeviceToHost, c->stream));
        NQ_HIPCHK(hipStreamSynchronize(c->stream));
        if (err) return nq_fail(NQ_EUNSUPPORTED, "preset expansion failed at level %d (status mask 0x%x: "
                                "4=start row pre-placed or below the board, 64=field beyond 32 bits)", lv + 1, err);
        if (n >= (1ull << 31)) return nq_fail(NQ_ECAPACITY, "preset expansion gives %llu records (limit 2^31)", n);
        NQ_HIPCHK(dev[a ^ 1].reserve(n ? n : 1));
        NQ_HIPCHK(nqk::launch_preset(c->stream, true, N, dev[a].p, (uint32_t)m, cnt.p, off.p, dev[a ^ 1].p,
                                     c->err.p));
        NQ_HIPCHK(hipStreamSynchronize(c->stream));
        a ^= 1;
        m = n;
    }
    *cur = a;
    *m_io = m;
    return 0;
}

// Shared body of nq_ctx_load (one N), nq_ctx_load_mixed (N per record) and
// nq_ctx_load_expanded (`levels` preset queens added on the device first).
int load_impl(nq_ctx *c, int N, const int32_t *Ns, const uint32_t *rec4, uint64_t m, const uint32_t *order,
              int levels = 0, uint64_t *m_out = nullptr)
{
    if (int r = check_ctx(c)) return r;
    if (!Ns && (N < nqt::kMinN || N > nqt::kMaxN)) return nq_fail(NQ_EINVAL, "N=%d outside the u32 path (%d..%d)", N, nqt::kMinN, nqt::kMaxN);
    std::vector<uint8_t> n8;
    if (Ns) {
        n8.resize(m);
        for (uint64_t a = 0; a < m; a++) {
            if (Ns[a] < nqt::kMinN || Ns[a] > nqt::kMaxN)
                return nq_fail(NQ_EINVAL, "record %llu: N=%d outside the u32 path (%d..%d)", (unsigned long long)a, Ns[a],
                               nqt::kMinN, nqt::kMaxN);
            n8[a] = (uint8_t)Ns[a];
        }
    }
    if (m && !rec4) return nq_fail(NQ_EINVAL, "null record pointer");
    if (m >= (1ull << 31)) return nq_fail(NQ_EINVAL, "too many records (%llu)", (unsigned long long)m);
    // records expanded on the device (nq_ctx_load_expanded): the set to decode
    nqk_buf::Buf<uint4> xdev[2];
    int xcur = 0;
    if (levels > 0) {
        NQ_HIPCHK(c->err.reserve(1));
        NQ_HIPCHK(xdev[0].reserve(m ? m : 1));
        if (m) NQ_HIPCHK(hipMemcpyAsync(xdev[0].p, rec4, m * 16, hipMemcpyHostToDevice, c->stream));
        if (int r = expand_device(c, N, xdev, &xcur, &m, levels)) return r;
    }
    if (m_out) *m_out = m;
    if (order) {
        std::vector<uint8_t> seen(m, 0);
        for (uint64_t a = 0; a < m; a++) {
            if (order[a] >= m || seen[order[a]]) return nq_fail(NQ_EINVAL, "order is not a permutation of 0..m-1");
            seen[order[a]] = 1;
        }
    }
    // expansion lanes per record: one per root candidate, at most N of them
    uint32_t fan = (uint32_t)N;
    if (Ns) {
        fan = 1;
        for (uint8_t n : n8) fan = std::max<uint32_t>(fan, n);
    }
    if ((uint64_t)m * fan >= (1ull << 32))
        return nq_fail(NQ_ECAPACITY, "%llu records of up to N=%u: the expansion takes fewer than 2^32 / N (load in chunks)",
                       (unsigned long long)m, fan);
    drop_graph(c);
    c->loaded = false;
    c->sized_split = -2;
    c->has_splits = false;
    c->fan = fan;
    const size_t mm = m ? m : 1;
    NQ_HIPCHK(c->rec.reserve(mm));
    NQ_HIPCHK(c->root.reserve(mm));
    NQ_HIPCHK(c->sched.reserve(nqk::kSchedU4 * mm));
    NQ_HIPCHK(c->order.reserve(mm));
    NQ_HIPCHK(c->nitems.reserve(mm * fan));
    NQ_HIPCHK(c->off.reserve(mm * fan));
    NQ_HIPCHK(c->scan_tmp.reserve(mm * fan / 1024 + 1));
    NQ_HIPCHK(c->weighted.reserve(mm));
    NQ_HIPCHK(c->nodes.reserve(mm));
    NQ_HIPCHK(c->raw.reserve(mm));
    if (Ns) NQ_HIPCHK(c->nsz.reserve(mm));
    c->N = Ns ? 0 : N;
    c->m = m;
    std::vector<uint32_t> ord;
    if (!order) {
        ord.resize(mm);
        for (uint64_t a = 0; a < mm; a++) ord[a] = (uint32_t)a;
        order = ord.data();
    }
    if (m) {
        if (levels > 0)
            NQ_HIPCHK(hipMemcpyAsync(c->rec.p, xdev[xcur].p, m * 16, hipMemcpyDeviceToDevice, c->stream));
        else
            NQ_HIPCHK(hipMemcpyAsync(c->rec.p, rec4, m * 16, hipMemcpyHostToDevice, c->stream));
        NQ_HIPCHK(hipMemcpyAsync(c->order.p, order, m * 4, hipMemcpyHostToDevice, c->stream));
        if (Ns) NQ_HIPCHK(hipMemcpyAsync(c->nsz.p, n8.data(), m, hipMemcpyHostToDevice, c->stream));
    }
    NQ_HIPCHK(nqk::launch_fill_u32(c->stream, c->err.p, 1, 0));
    NQ_HIPCHK(nqk::launch_decode(c->stream, N, Ns ? c->nsz.p : nullptr, c->rec.p, (uint32_t)m, c->root.p,
                                 c->sched.p, c->err.p));
    NQ_HIPCHK(nqk::launch_fill_u64(c->stream, c->weighted.p, mm, 0));
    NQ_HIPCHK(nqk::launch_fill_u64(c->stream, c->nodes.p, mm, 0));
    uint32_t err = 0;
    NQ_HIPCHK(hipMemcpyAsync(&err, c->err.p, 4, hipMemcpyDeviceToHost, c->stream));
    NQ_HIPCHK(hipStreamSynchronize(c->stream));
    if (err) return nq_fail(NQ_EUNSUPPORTED, "record decode failed (status mask 0x%x: 2=P5 record past its switch row, 4=shift, "
                                          "8=no terminal row, 16=depth)", err);
    // the terminal depth is N-1-preset for every record of a set; take the max
    c->Tmax = 0;
    c->Tmin = 0;
    c->Trec.assign(m, 0);
    if (m) {
        std::vector<uint4> sc(nqk::kSchedU4 * m);
        NQ_HIPCHK(hipMemcpy(sc.data(), c->sched.p, nqk::kSchedU4 * m * 16, hipMemcpyDeviceToHost));
        c->Tmin = 0xFFu;
        for (uint64_t a = 0; a < m; a++) {
            const uint32_t T = sc[nqk::kSchedU4 * a + 1].x & 0xFFu;
            c->Trec[a] = (uint8_t)T;
            c->Tmax = std::max(c->Tmax, T);
            c->Tmin = std::min(c->Tmin, T);
        }
    }
    // the walk keeps 16 parameter bytes per record (depths 0..15), so the
    // terminal depth T = N-1-preset must not exceed 16
    if (c->Tmax > 16u) return nq_fail(NQ_EUNSUPPORTED, "terminal depth %u > 16 (preset too small for N)", c->Tmax);
    c->loaded = true;
    return 0;
}

}  // namespace

int nq_ctx_load(nq_ctx *c, int N, const uint32_t *rec4, uint64_t m, const uint32_t *order)
{
    return load_impl(c, N, nullptr, rec4, m, order);
}

int nq_ctx_load_mixed(nq_ctx *c, const int32_t *Ns, const uint32_t *rec4, uint64_t m, const uint32_t *order)
{
    if (m && !Ns) return nq_fail(NQ_EINVAL, "null N array");
    return load_impl(c, 0, Ns, rec4, m, order);
}

int nq_ctx_load_expanded(nq_ctx *c, int N, const uint32_t *rec4, uint64_t m, int levels, uint64_t *m_out)
{
    if (levels < 1 || levels > 8) return nq_fail(NQ_EINVAL, "levels must be 1..8, got %d", levels);
    return load_impl(c, N, nullptr, rec4, m, nullptr, levels, m_out);
}

int nq_expand_records(nq_ctx *c, int N, const uint32_t *rec4, uint64_t m, int levels, uint32_t *out4,
                      uint64_t cap, uint64_t *m_out)
{
    if (int r = check_ctx(c)) return r;
    if (N < nqt::kMinN || N > nqt::kMaxN) return nq_fail(NQ_EINVAL, "N=%d outside %d..%d", N, nqt::kMinN, nqt::kMaxN);
    if (levels < 1 || levels > 8) return nq_fail(NQ_EINVAL, "levels must be 1..8, got %d", levels);
    if (!m_out) return nq_fail(NQ_EINVAL, "null count pointer");
    if (m && !rec4) return nq_fail(NQ_EINVAL, "null record pointer");
    if (m >= (1ull << 31)) return nq_fail(NQ_EINVAL, "too many records (%llu)", (unsigned long long)m);
    nqk_buf::Buf<uint4> dev[2];
    int cur = 0;
    NQ_HIPCHK(c->err.reserve(1));
    NQ_HIPCHK(dev[0].reserve(m ? m : 1));
    if (m) NQ_HIPCHK(hipMemcpyAsync(dev[0].p, rec4, m * 16, hipMemcpyHostToDevice, c->stream));
    uint64_t n = m;
    if (int r = expand_device(c, N, dev, &cur, &n, levels)) return r;
    *m_out = n;
    if (!out4) return 0;
    if (cap < n) return nq_fail(NQ_ECAPACITY, "output holds %llu records, expansion gives %llu",
                                (unsigned long long)cap, (unsigned long long)n);
    if (n) NQ_HIPCHK(hipMemcpyAsync(out4, dev[cur].p, n * 16, hipMemcpyDeviceToHost, c->stream));
    NQ_HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int nq_ctx_set_splits(nq_ctx *c, const uint8_t *depth, uint64_t m)
{
    if (int r = check_ctx(c)) return r;
    if (!c->loaded) return nq_fail(NQ_ESTATE, "no task set loaded");
    if (m != c->m) return nq_fail(NQ_EINVAL, "expected %llu depths, got %llu", (unsigned long long)c->m,
                                  (unsigned long long)m);
    if (m && !depth) return nq_fail(NQ_EINVAL, "null depth array");
    uint32_t mn = 255;
    for (uint64_t a = 0; a < m; a++) {
        if (depth[a] > 15) return nq_fail(NQ_EINVAL, "split depth %u > 15 at record %llu", depth[a],
                                          (unsigned long long)a);
        // the depth the record's items really start at (its root when the
        // depth is 0 or beyond T-2): the stack is sized for the shallowest
        mn = std::min<uint32_t>(mn, nqk::item_depth(depth[a], c->Trec[a]));
    }
    NQ_HIPCHK(c->splits.reserve(m ? m : 1));
    if (m) NQ_HIPCHK(hipMemcpyAsync(c->splits.p, depth, m, hipMemcpyHostToDevice, c->stream));
    NQ_HIPCHK(hipStreamSynchronize(c->stream));
    drop_graph(c);
    c->has_splits = true;
    c->split_min = m ? mn : 0;
    c->sized_split = -2;
    return 0;
}

int nq_ctx_run_plain(nq_ctx *c, int count_nodes, uint64_t *total_out, nq_timing *t)
{
    if (int r = check_ctx(c)) return r;
    if (!c->loaded) return nq_fail(NQ_ESTATE, "no task set loaded");
    const int block = 64;
    const nqk::WalkLds lay = nqk::walk_lds(c->Tmax, 0u);
    if (lay.bytes(block) > (size_t)c->lds_max) return nq_fail(NQ_ECAPACITY, "LDS stack too large");
    hipStream_t st = c->stream;
    NQ_HIPCHK(nqk::launch_fill_u64(st, c->total.p, 1, 0));
    if (!count_nodes) NQ_HIPCHK(nqk::launch_fill_u64(st, c->nodes.p, c->m ? c->m : 1, 0));
    NQ_HIPCHK(hipEventRecord(c->ev[0], st));
    NQ_HIPCHK(hipEventRecord(c->ev[1], st));
    NQ_HIPCHK(nqk::launch_plain(st, count_nodes != 0, block, lay, c->root.p, c->sched.p, (uint32_t)c->m,
                             c->weighted.p, c->nodes.p, c->total.p));
    NQ_HIPCHK(hipEventRecord(c->ev[2], st));
    NQ_HIPCHK(hipEventRecord(c->ev[3], st));
    unsigned long long tot = 0;
    NQ_HIPCHK(hipMemcpyAsync(&tot, c->total.p, 8, hipMemcpyDeviceToHost, st));
    NQ_HIPCHK(hipStreamSynchronize(st));
    if (total_out) *total_out = tot;
    if (t) {
        if (int r = elapsed(c, 1, 2, &t->kernel_ms)) return r;
        if (int r = elapsed(c, 0, 3, &t->total_ms)) return r;
        t->items = c->m;
        t->launches = 1;
    }
    return 0;
}

int nq_ctx_run_batched(nq_ctx *c, const nq_sched *sp, int count_nodes, uint64_t *total_out, nq_timing *t)
{
    if (int r = check_ctx(c)) return r;
    if (!c->loaded) return nq_fail(NQ_ESTATE, "no task set loaded");
    if (!sp) return nq_fail(NQ_EINVAL, "null schedule");
    nq_sched s = *sp;
    nq_run run = ctx_run(c);
    if (s.split_depth == -1 && !c->has_splits) return nq_fail(NQ_ESTATE, "no per-record split depths set");
    if (int r = nq_prepare_sched(run, c->lds_max, &s)) return r;
    if (c->m == 0) {
        if (total_out) *total_out = 0;
        if (t) memset(t, 0, sizeof *t);
        return 0;
    }
    if (c->sized_split != s.split_depth) {
        drop_graph(c);
        if (int r = size_items(c, s.split_depth, s.block)) return r;
        run.items = c->items.p;
    }
    const size_t lds_bytes = s.lds_stack ? nq_batched_lds(run, s.split_depth).bytes(s.block) : 0u;
    int grid = 0;
    if (int r = nq_batched_grid(c->cu, s, lds_bytes, count_nodes, c->item_total, &grid)) return r;
    if (s.use_graph) {
        nq_graph_key k;
        k.split = s.split_depth; k.block = s.block; k.lds = s.lds_stack; k.chunk = s.chunk; k.grid = grid;
        k.nodes = count_nodes;
        if (!c->gexec || !(k == c->gkey)) {
            drop_graph(c);
            NQ_HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            int r = nq_enqueue_batched(c->stream, run, s, count_nodes, grid, nullptr);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(c->stream, &g);
            if (r) { if (g) (void)hipGraphDestroy(g); return r; }
            NQ_HIPCHK(e);
            c->graph = g;
            NQ_HIPCHK(hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0));
            c->gkey = k;
        }
        // events bracket the whole replay; the DFS kernel alone is not timed in graph mode
        NQ_HIPCHK(hipEventRecord(c->ev[0], c->stream));
        NQ_HIPCHK(hipGraphLaunch(c->gexec, c->stream));
        NQ_HIPCHK(hipEventRecord(c->ev[3], c->stream));
    } else {
        if (int r = nq_enqueue_batched(c->stream, run, s, count_nodes, grid, c->ev)) return r;
    }
    unsigned long long tot = 0, items = 0;
    NQ_HIPCHK(hipMemcpyAsync(&tot, c->total.p, 8, hipMemcpyDeviceToHost, c->stream));
    NQ_HIPCHK(hipMemcpyAsync(&items, c->item_count.p, 8, hipMemcpyDeviceToHost, c->stream));
    NQ_HIPCHK(hipStreamSynchronize(c->stream));
    if (items != c->item_total) return nq_fail(NQ_ECAPACITY, "work items changed between sizing and run");
    if (total_out) *total_out = tot;
    if (t) {
        if (s.use_graph) t->kernel_ms = -1.0;
        else if (int r = elapsed(c, 1, 2, &t->kernel_ms)) return r;
        if (int r = elapsed(c, 0, 3, &t->total_ms)) return r;
        t->items = items;
        t->launches = 1;
    }
    return 0;
}

int nq_ctx_results(nq_ctx *c, uint64_t *weighted_out, uint64_t *nodes_out)
{
    if (int r = check_ctx(c)) return r;
    if (!c->loaded) return nq_fail(NQ_ESTATE, "no task set loaded");
    if (c->m) {
        if (weighted_out) NQ_HIPCHK(hipMemcpyAsync(weighted_out, c->weighted.p, c->m * 8, hipMemcpyDeviceToHost, c->stream));
        if (nodes_out) NQ_HIPCHK(hipMemcpyAsync(nodes_out, c->nodes.p, c->m * 8, hipMemcpyDeviceToHost, c->stream));
    }
    NQ_HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int nq_count(nq_ctx *c, int N, const uint32_t *rec4, uint64_t m, const nq_sched *sched, uint64_t *weighted_out,
             uint64_t *nodes_out, uint64_t *total_out)
{
    if (int r = nq_ctx_load(c, N, rec4, m, nullptr)) return r;
    const int nodes = nodes_out != nullptr;
    int r = sched ? nq_ctx_run_batched(c, sched, nodes, total_out, nullptr)
                  : nq_ctx_run_plain(c, nodes, total_out, nullptr);
    if (r) return r;
    return nq_ctx_results(c, weighted_out, nodes_out);
}

int nq_symmetric_counts(nq_ctx *c, int N, uint64_t *f90_out, uint64_t *f180_out)
{
    if (int r = check_ctx(c)) return r;
    if (N < 2 || N > 27) return nq_fail(NQ_EINVAL, "N=%d outside 2..27", N);
    NQ_HIPCHK(nqk::launch_fill_u64(c->stream, c->sym.p, 2, 0));
    NQ_HIPCHK(nqk::launch_symmetric(c->stream, N, c->sym.p));
    unsigned long long v[2] = {0, 0};
    NQ_HIPCHK(hipMemcpyAsync(v, c->sym.p, 16, hipMemcpyDeviceToHost, c->stream));
    NQ_HIPCHK(hipStreamSynchronize(c->stream));
    if (f90_out) *f90_out = v[0];
    if (f180_out) *f180_out = v[1];
    return 0;
}

int nq_small_total(nq_ctx *c, int N, uint64_t *total_out)
{
    if (int r = check_ctx(c)) return r;
    if (N < 1 || N > 12) return nq_fail(NQ_EINVAL, "N=%d outside 1..12", N);
    if (!total_out) return nq_fail(NQ_EINVAL, "null out pointer");
    NQ_HIPCHK(nqk::launch_fill_u64(c->stream, c->total.p, 1, 0));
    NQ_HIPCHK(nqk::launch_small_total(c->stream, N, c->total.p));
    unsigned long long v = 0;
    NQ_HIPCHK(hipMemcpyAsync(&v, c->total.p, 8, hipMemcpyDeviceToHost, c->stream));
    NQ_HIPCHK(hipStreamSynchronize(c->stream));
    *total_out = v;
    return 0;
}

int nq_decode_host(int N, const uint32_t *rec4, uint64_t m, int32_t *out10)
{
    if (N < nqt::kMinN || N > nqt::kMaxN) return nq_fail(NQ_EINVAL, "N=%d outside %d..%d", N, nqt::kMinN, nqt::kMaxN);
    if (m && (!rec4 || !out10)) return nq_fail(NQ_EINVAL, "null pointer");
    for (uint64_t a = 0; a < m; a++) {
        nqt::Desc D;
        const uint32_t *r = rec4 + 4 * a;
        nqt::decode(N, r[0], r[1], r[2], r[3], D);
        int32_t *o = out10 + 10 * a;
        o[0] = D.fid; o[1] = D.start; o[2] = D.jmark; o[3] = D.endmark; o[4] = D.mark1; o[5] = D.mark2;
        o[6] = 1 << D.wshift;
        o[7] = (int32_t)D.free0;
        o[8] = (int32_t)D.T;
        o[9] = D.status;
    }
    return 0;
}

}  // extern "C"
