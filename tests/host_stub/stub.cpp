// Stand-in HIP runtime for the host-logic check (tests/test_host_logic.py):
// allocations are host memory, streams and events are tokens, graph capture
// records the launches and a graph launch replays them, and the kernel
// launchers touch every buffer range a launch would read or write, so a
// wrong size or offset in the host library shows under AddressSanitizer.
#include <hip/hip_runtime.h>
#include "../../mpfft_amd/csrc/kernels.h"
#include <cstdlib>
#include <cstring>
#include <cstdio>
#include <atomic>
#include <mutex>
#include <vector>
#include <functional>
#include <set>
struct ihipStream_t { int id; bool capturing = false; std::vector<std::function<void()>> *cap = nullptr; };
struct ihipEvent_t { int x; };
struct ihipGraph { std::vector<std::function<void()>> ops; };
struct hipGraphExec { std::vector<std::function<void()>> ops; };
static thread_local hipStream_t g_capture = nullptr;
static std::vector<std::function<void()>> *cur_cap() { return g_capture ? g_capture->cap : nullptr; }
static void run_or_record(std::function<void()> f) { if (auto c = cur_cap()) c->push_back(f); else f(); }
// failure injection: the k-th hipMalloc from now fails (k = 0: the next one)
static std::atomic<int> g_malloc_fail_in{-1};
void stub_fail_malloc_in(int k) { g_malloc_fail_in = k; }
hipError_t hipMalloc(void **p, size_t b) {
    int k = g_malloc_fail_in.load();
    while (k >= 0 && !g_malloc_fail_in.compare_exchange_weak(k, k - 1)) {}
    if (k == 0) { *p = nullptr; return hipErrorUnknown; }
    *p = malloc(b ? b : 1); return hipSuccess; }
hipError_t hipFree(void *p) { free(p); return hipSuccess; }
hipError_t hipGetSymbolAddress(void **p, const void *symbol) { *p = const_cast<void *>(symbol); return hipSuccess; }
hipError_t hipHostMalloc(void **p, size_t b, unsigned) { *p = malloc(b ? b : 1); return hipSuccess; }
hipError_t hipHostFree(void *p) { free(p); return hipSuccess; }
hipError_t hipMemcpy(void *d, const void *s, size_t b, hipMemcpyKind) { memcpy(d, s, b); return hipSuccess; }
hipError_t hipMemcpyAsync(void *d, const void *s, size_t b, hipMemcpyKind, hipStream_t) { run_or_record([=] { memcpy(d, s, b); }); return hipSuccess; }
hipError_t hipMemcpy2DAsync(void *d, size_t dp, const void *s, size_t sp, size_t w, size_t h, hipMemcpyKind, hipStream_t) {
    for (size_t i = 0; i < h; i++) memcpy((char *)d + i * dp, (const char *)s + i * sp, w); return hipSuccess; }
hipError_t hipStreamCreate(hipStream_t *s) { *s = new ihipStream_t{1}; return hipSuccess; }
hipError_t hipStreamDestroy(hipStream_t s) { delete s; return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t s) { (void)s->id; return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned) { (void)s->id; (void)e->x; return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t *e) { *e = new ihipEvent_t{0}; return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { *e = new ihipEvent_t{0}; return hipSuccess; }
hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) { (void)e->x; if (s) (void)s->id; return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t e) { (void)e->x; return hipSuccess; }
// the stand-in runs work at enqueue: an event is complete once recorded
hipError_t hipEventQuery(hipEvent_t e) { (void)e->x; return hipSuccess; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) { (void)a->x; (void)b->x; *ms = (float)(rand() % 100 + 1) * 0.01f; return hipSuccess; }
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
hipError_t hipGetDeviceCount(int *n) { *n = 1; return hipSuccess; }
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t, int) { *v = 256; return hipSuccess; }
hipError_t hipGetLastError() { return hipSuccess; }
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) { s->cap = new std::vector<std::function<void()>>; g_capture = s; return hipSuccess; }
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g) { *g = new ihipGraph{*s->cap}; delete s->cap; s->cap = nullptr; g_capture = nullptr; return hipSuccess; }
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *, char *, size_t) { *e = new hipGraphExec{g->ops}; return hipSuccess; }
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s) { (void)s->id; for (auto &f : e->ops) f(); return hipSuccess; }
hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete e; return hipSuccess; }
hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }

namespace mpfft {
static const int kL[] = {1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 64};   // as mp_fft.hip
int limbs_for_precision(long P) { long need = (P + 31) / 32; for (int L : kL) if (L >= need) return L; return -1; }
static thread_local volatile uint32_t sink;   // keeps the read-touches (one per thread)
static void touch_r(const uint32_t *p, size_t n) { uint32_t s = 0; for (size_t i = 0; i < n; i++) s += p[i]; sink = s; }
static void touch_w(uint32_t *p, size_t n) { for (size_t i = 0; i < n; i++) p[i] = 0x5a5a5a5au; }
// The subtree kernel's tables against the recursion itself
// (serial_mp_fft3.c:180-182, restated here): the subtrees tile the N
// positions, each is a node of the tree (or a size-4 leaf), the nodes of
// every level up to the subtree size are exactly the subtrees' offsets plus
// the relative offsets, the level launches start at the next size, and the
// workgroup fits the LDS.
std::atomic<int> stub_sub_checked{0}, stub_sub_errors{0};
static void sub_walk(uint32_t n, uint32_t off, std::vector<std::set<uint32_t>> &lv,
                     std::set<uint32_t> &leaf4)
{
    if (n <= 4) { if (n == 4) leaf4.insert(off); return; }
    uint32_t l = 0; while ((1u << l) < n) l++;
    lv[l].insert(off);
    sub_walk(n / 2, off, lv, leaf4);
    sub_walk(n / 4, off + n / 2, lv, leaf4);
    sub_walk(n / 4, off + 3 * n / 4, lv, leaf4);
}
static void sub_check(int L, uint32_t N, const SubArgs &sa, const LevelArgs *lv, int nlev)
{
    int bad = 0;
    std::vector<std::set<uint32_t>> glv(40);
    std::set<uint32_t> leaf4;
    sub_walk(N, 0, glv, leaf4);
    std::vector<int> cover(N, 0);
    std::vector<std::set<uint32_t>> got(40);
    for (uint32_t i = 0; i < sa.nsub; i++) {
        const uint32_t kind = sa.sub[i] & 1u, o = sa.sub[i] >> 1, lg = sa.lgs[kind], sz = 1u << lg;
        if (o + sz > N) { bad++; continue; }
        for (uint32_t j = 0; j < sz; j++) cover[o + j]++;
        if (sz == N ? o != 0 : sz >= 8 ? !glv[lg].count(o) : !(sz == 4 && leaf4.count(o))) bad++;
        for (uint32_t l = 3; l <= lg; l++)
            for (uint32_t r = 0; r < sa.rel_cnt[kind][l]; r++)
                got[l].insert(o + sa.rel[sa.rel_first[kind][l] + r]);
    }
    for (uint32_t j = 0; j < N; j++) bad += cover[j] != 1;
    for (uint32_t l = 3; l <= sa.lgs[0]; l++) bad += got[l] != glv[l];
    if (nlev > 0 && lv[0].n != (2u << sa.lgs[0])) bad++;
    if (sizeof(uint32_t) * (L + 1) * (2 * (size_t)sa.nthr + (2u << sa.lgs[0])) > 160 * 1024) bad++;
    if (sa.nthr % 64 || sa.nthr > 256) bad++;
    stub_sub_checked++;
    if (bad) { stub_sub_errors++; printf("sub tables: N=%u L=%d S=%u: %d mismatches\n", N, L, 1u << sa.lgs[0], bad); }
}
// identity mode: the plain transform's launch copies its input to its
// output, so a caller's data movement around it can be checked end to end
std::atomic<int> stub_identity_mp{0};
hipError_t mp_chunk_relayout_launch(const uint32_t *src, uint32_t *dst, uint32_t N, uint32_t R,
                                    uint32_t cl, bool to_soa, hipStream_t) {
    if (!N || !R || !cl || N % cl) return hipErrorInvalidValue;
    run_or_record([=] {
        for (uint32_t j = 0; j < R; j++)
            for (uint32_t e = 0; e < 2 * N; e++) {
                const uint32_t im = e >= N, k = e - im * N, c = k / cl, loc = k - c * cl + im * cl;
                const size_t chunk = (size_t)c * R * 2 * cl + (size_t)j * 2 * cl + loc;
                const size_t soa = (size_t)j * 2 * N + e;
                if (to_soa) dst[soa] = src[chunk];
                else dst[chunk] = src[soa];
            }
    });
    return hipSuccess;
}
hipError_t mp_fft_launch(int L, int, const LeafArgs &la, const LevelArgs *lv, int nlev, uint32_t batch, hipStream_t, const SubArgs *sub) {
    if (stub_identity_mp.load()) {
        const size_t w = 2 * (size_t)la.N * (L + 1);
        LeafArgs a = la;
        run_or_record([=] { for (uint32_t b = 0; b < batch; b++) memcpy(a.data + a.stride * b, a.in + a.in_stride * b, 4 * w); });
        return hipSuccess;
    }
    if (sub && la.npos) sub_check(L, la.N, *sub, lv, nlev);
    const size_t w = 2 * (size_t)la.N * (L + 1);
    LeafArgs a = la; std::vector<LevelArgs> v(lv, lv + nlev);
    const bool has_sub = sub != nullptr;
    SubArgs sa = has_sub ? *sub : SubArgs{};
    run_or_record([=] {
        // the subtree kernel reads its subtree list and every relative node
        // offset of its levels
        if (a.npos && has_sub) {
            touch_r(sa.sub, sa.nsub);
            for (int k = 0; k < 2; k++)
                for (uint32_t l = 3; l <= sa.lgs[k] && l < (uint32_t)kSubMaxLevel; l++)
                    if (sa.rel_cnt[k][l]) touch_r(sa.rel + sa.rel_first[k][l], sa.rel_cnt[k][l]);
        }
        for (uint32_t b = 0; b < batch; b++) {
            if (a.npos) { touch_r(a.in + a.in_stride * b, w); touch_w(a.data + a.stride * b, w); touch_w(a.aos + a.aos_stride * b, (size_t)2 * a.N * a.aos_words); }
            for (auto &x : v) { touch_w(x.data + x.stride * b, w); touch_r(x.lut, 1); }
        }
    });
    return hipSuccess;
}
hipError_t mp_unswap_scale_launch(int L, uint32_t *d, uint64_t stride, uint32_t N, int32_t, uint32_t batch, hipStream_t) {
    run_or_record([=] { for (uint32_t b = 0; b < batch; b++) touch_w(d + stride * b, 2 * (size_t)N * (L + 1)); }); return hipSuccess; }
// writes every partial of every block (as the kernel does)
hipError_t mp_err_launch(int, int, const uint32_t *, const uint32_t *, uint64_t, uint32_t N, double *partial, uint32_t batch, hipStream_t) {
    const size_t nd = (size_t)((2 * N + kErrBlock - 1) / kErrBlock) * batch * kErrPartial;
    run_or_record([=] { for (size_t i = 0; i < nd; i++) partial[i] = 1.0; });
    return hipSuccess; }
hipError_t mp_binop_launch(int, int, int, uint32_t, const uint32_t *, const uint32_t *, uint32_t *, hipStream_t) { return hipSuccess; }
static void fx_touch(int L, const FxArgs &a, bool panel) {
    const size_t w = 2 * (size_t)a.N * (L + 1);
    for (uint32_t b = 0; b < a.batch; b++) { touch_r(a.in + a.in_stride * b, w); touch_w(a.out + a.out_stride * b, w); }
    const size_t fw = 2 * (size_t)a.N * (L + 2) * (panel ? fxp_groups(a.batch) * kPanel : a.batch);
    touch_w(a.buf0, fw); touch_w(a.buf1, fw); touch_r(a.tw, 1); touch_w(a.flag, 1);
}
hipError_t fx_eoff_launch(const FxArgs &a, hipStream_t) {
    if (a.eoff_nblk != fx_eoff_blocks(a.N)) return hipErrorInvalidValue;
    FxArgs c = a;
    run_or_record([=] {
        for (uint32_t b = 0; b < c.batch; b++) touch_r(c.in + c.in_stride * b, 2 * (size_t)c.N);
        touch_w((uint32_t *)c.eoff_part, (size_t)c.batch * c.eoff_nblk);
        touch_w((uint32_t *)c.eoff, c.batch); touch_w(c.flag, 1);
    });
    return hipSuccess;
}
hipError_t fx_fft_launch(int L, int, const FxArgs &a, int, int, hipStream_t) { FxArgs c = a; run_or_record([=] { fx_touch(L, c, false); }); return hipSuccess; }
hipError_t fxp_fft_launch(int L, int, const FxArgs &a, int, int, hipStream_t) {
    if (!fxp_eligible(a.N, a.batch)) return hipErrorInvalidValue;
    FxArgs c = a; run_or_record([=] { fx_touch(L, c, true); }); return hipSuccess; }
hipError_t fxs_fft_launch(int L, int, const FxArgs &a, const FxsPass *, int np, uint32_t, int,
                          int, hipStream_t) {
    if (!fxs_eligible(a.N, L + 2) || np < 2 || np > kFxsMaxPasses) return hipErrorInvalidValue;
    FxArgs c = a; run_or_record([=] { fx_touch(L, c, true); }); return hipSuccess; }
hipError_t dd_fft_launch(int K, int L, int, const FxArgs &a, int, int, hipStream_t) {
    FxArgs c = a;
    run_or_record([=] {
        const size_t w = 2 * (size_t)c.N * (L + 1);
        for (uint32_t b = 0; b < c.batch; b++) { touch_r(c.in + c.in_stride * b, w); touch_w(c.out + c.out_stride * b, w); }
        const size_t fw = (size_t)K * 2 * c.N * c.batch * 2;      // doubles -> 32-bit words
        touch_w(c.buf0, fw); touch_w(c.buf1, fw); touch_r(c.tw, 1); touch_w(c.flag, 1);
    });
    return hipSuccess;
}
hipError_t d_fft_launch(const DTables &, const double *, double *, uint64_t, uint32_t, hipStream_t) { return hipSuccess; }
}
