// Batched form: many transforms of uneven size and precision in one HIP
// graph launch (include/mpfft.h, mpfft_batch_*).
//
// The reference runs one transform per call (mpFFT.h:38); a workload of
// thousands of small transforms would pay a launch sequence per transform,
// and a single transform of n <= 2^14 fills only a sliver of the device.
// Here transforms of equal (n, prec) are grouped and each group runs as one
// plan of the group's batch (its kernels cover every transform of the group
// in each launch).  The groups are dealt over up to kMaxStreams streams,
// captured into one graph (fork from the batch's stream, join back) and
// replayed: one launch per execute, independent groups overlapping on the
// device.  A fixed-form group takes the panel or the per-lane kernels by the
// rule MPFFT_FORM_FIXED plans use (mpfft::fxp_preferred, from measured
// crossovers), a class of more than one panel whose last panel that rule
// rejects as mostly padding being split into its full panels (panel kernels)
// and the rest (per-lane kernels); the autotuner times every group on the caller's
// buffers, then the graph with the groups dealt over 1, 2 and 4 streams, and
// keeps the fastest.  Built only on the public plan API.
#include "../../include/mpfft.h"
#include "kernels.h"

#include <algorithm>
#include <map>
#include <mutex>
#include <thread>
#include <utility>
#include <vector>

namespace {

#define HIP_TRY(expr)                                          \
    do {                                                       \
        if ((expr) != hipSuccess) return MPFFT_ERR_HIP;        \
    } while (0)

constexpr int kMaxStreams = 4;
constexpr int kTimedReps = 3;

struct Group {
    unsigned long N = 0;
    long P = 0, L = 0;
    size_t words = 0;                 // words per transform
    size_t off = 0;                   // first word of the group in the packed buffers
    std::vector<unsigned> idx;        // its transforms, in input order
    int form = 0;                     // form of its plan
    mpfft_plan plan = nullptr;
    double ms = 0;                    // measured time of the group alone
    int stream = 0;                   // stream of the graph that runs it
};

bool is_pow2(unsigned long n) { return n && !(n & (n - 1)); }

}  // namespace

struct mpfft_batch_s {
    int device = 0, form = 0;
    unsigned count = 0;
    std::vector<Group> groups;
    std::vector<unsigned> grp_of, rank_of;   // per transform: group, position in it
    size_t words = 0;
    hipStream_t stream = nullptr, last = nullptr;
    hipStream_t side[kMaxStreams] = {};
    hipEvent_t fork = nullptr, join[kMaxStreams] = {}, t0 = nullptr, t1 = nullptr;
    int nstreams = 1;
    double graph_ms = 0;
    bool tuned = false;
    hipGraphExec_t exec = nullptr;
    const void *g_in = nullptr;
    const void *g_out = nullptr;
    std::mutex mu;

    ~mpfft_batch_s()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        for (auto &g : groups)
            if (g.plan) mpfft_plan_destroy(g.plan);
        for (int s = 0; s < kMaxStreams; s++) {
            if (join[s]) (void)hipEventDestroy(join[s]);
            if (side[s]) (void)hipStreamDestroy(side[s]);
        }
        for (auto e : {fork, t0, t1})
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

int run_group(const Group &g, const uint32_t *in, uint32_t *out, hipStream_t st)
{
    return mpfft_execute_limbs(g.plan, in + g.off, out + g.off, st);
}

// Capture every group into one graph over S streams.
int capture(mpfft_batch_s *b, int S, const uint32_t *in, uint32_t *out, hipGraphExec_t *exec)
{
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = MPFFT_OK;
    if (S > 1) {
        if (hipEventRecord(b->fork, b->stream) != hipSuccess) rc = MPFFT_ERR_HIP;
        for (int s = 0; s < S && !rc; s++)
            if (hipStreamWaitEvent(b->side[s], b->fork, 0) != hipSuccess) rc = MPFFT_ERR_HIP;
    }
    for (size_t i = 0; i < b->groups.size() && !rc; i++) {
        const Group &g = b->groups[i];
        rc = run_group(g, in, out, S > 1 ? b->side[g.stream] : b->stream);
    }
    if (S > 1)
        for (int s = 0; s < S && !rc; s++)
            if (hipEventRecord(b->join[s], b->side[s]) != hipSuccess ||
                hipStreamWaitEvent(b->stream, b->join[s], 0) != hipSuccess)
                rc = MPFFT_ERR_HIP;
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &graph);   // ends it in every case
    if (!rc && e != hipSuccess) rc = MPFFT_ERR_HIP;
    if (!rc && hipGraphInstantiate(exec, graph, nullptr, nullptr, 0) != hipSuccess)
        rc = MPFFT_ERR_HIP;
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
}

// Deal the groups over S streams, longest measured time first, each to the
// stream with the least time so far.
void assign_streams(mpfft_batch_s *b, int S)
{
    std::vector<size_t> ord(b->groups.size());
    for (size_t i = 0; i < ord.size(); i++) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](size_t x, size_t y) {
        return b->groups[x].ms > b->groups[y].ms;
    });
    double load[kMaxStreams] = {0, 0, 0, 0};
    for (size_t i : ord) {
        int best = 0;
        for (int s = 1; s < S; s++)
            if (load[s] < load[best]) best = s;
        Group &g = b->groups[i];
        g.stream = best;
        load[best] += g.ms;
    }
}

// time `reps` launches of an instantiated graph (ms per launch)
int time_graph(mpfft_batch_s *b, hipGraphExec_t exec, double *ms)
{
    HIP_TRY(hipGraphLaunch(exec, b->stream));                // warm-up
    HIP_TRY(hipEventRecord(b->t0, b->stream));
    for (int r = 0; r < kTimedReps; r++) HIP_TRY(hipGraphLaunch(exec, b->stream));
    HIP_TRY(hipEventRecord(b->t1, b->stream));
    HIP_TRY(hipEventSynchronize(b->t1));
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, b->t0, b->t1));
    *ms = t / kTimedReps;
    return MPFFT_OK;
}

int tune(mpfft_batch_s *b, const uint32_t *in, uint32_t *out)
{
    // a graph launched before on a caller's stream finishes first: the
    // timing runs below write the same buffers, and its exec is replaced
    if (b->last) HIP_TRY(hipStreamSynchronize(b->last));
    // 1. every group alone on the batch's stream (its weight in the dealing)
    for (auto &g : b->groups) {
        int rc = run_group(g, in, out, b->stream);                    // warm-up
        if (rc) return rc;
        HIP_TRY(hipEventRecord(b->t0, b->stream));
        for (int r = 0; r < kTimedReps && !rc; r++) rc = run_group(g, in, out, b->stream);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(b->t1, b->stream));
        HIP_TRY(hipEventSynchronize(b->t1));
        float t = 0;
        HIP_TRY(hipEventElapsedTime(&t, b->t0, b->t1));
        g.ms = t / kTimedReps;
    }
    // 2. the graph over 1, 2, 4 streams
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    double best = 0;
    for (int S = 1; S <= kMaxStreams; S *= 2) {
        if (S > 1 && (size_t)S > b->groups.size()) break;
        assign_streams(b, S);
        hipGraphExec_t ex = nullptr;
        int rc = capture(b, S, in, out, &ex);
        double ms = 0;
        if (!rc) rc = time_graph(b, ex, &ms);
        if (rc) {
            if (ex) (void)hipGraphExecDestroy(ex);
            return rc;
        }
        if (!b->exec || ms < best) {
            if (b->exec) (void)hipGraphExecDestroy(b->exec);
            b->exec = ex;
            best = ms;
            b->nstreams = S;
        } else {
            (void)hipGraphExecDestroy(ex);
        }
    }
    assign_streams(b, b->nstreams);
    b->graph_ms = best;
    b->g_in = in;
    b->g_out = out;
    b->tuned = true;
    return MPFFT_OK;
}

}  // namespace

extern "C" {

int mpfft_batch_create(mpfft_batch *batch, unsigned int count, const unsigned long *n,
                       const long *prec, int form)
{
    if (!batch || !n || !prec || count == 0) return MPFFT_ERR_ARG;
    *batch = nullptr;
    const bool fixed = form == MPFFT_FORM_FIXED || form == MPFFT_FORM_FIXED_LANE ||
                       form == MPFFT_FORM_FIXED_PANEL;
    const bool words = form == MPFFT_FORM_DD || form == MPFFT_FORM_QD;
    if (!fixed && !words && form != MPFFT_FORM_AUTO && form != MPFFT_FORM_MP &&
        form != MPFFT_FORM_MP_C1 && form != MPFFT_FORM_MP_C2)
        return MPFFT_ERR_FORM;
    std::unique_ptr<mpfft_batch_s> b(new mpfft_batch_s);
    b->form = form;
    b->count = count;
    // groups of equal (n, prec), ordered by (n, prec); that order is the
    // packed layout
    std::map<std::pair<unsigned long, long>, std::vector<unsigned>> key;
    for (unsigned i = 0; i < count; i++) {
        if (!is_pow2(n[i])) return MPFFT_ERR_SIZE;
        if (mpfft_limbs_for_precision(prec[i]) < 1) return MPFFT_ERR_PREC;
        key[{n[i], prec[i]}].push_back(i);
    }
    b->grp_of.resize(count);
    b->rank_of.resize(count);
    size_t off = 0;
    // one group: transforms [r0, r1) of the (n, prec) class kv, its plan form
    auto add_group = [&](const std::pair<const std::pair<unsigned long, long>,
                                         std::vector<unsigned>> &kv,
                         size_t r0, size_t r1, int gform) {
        Group g;
        g.N = kv.first.first;
        g.P = kv.first.second;
        g.L = mpfft_limbs_for_precision(g.P);
        g.words = 2 * g.N * (size_t)(g.L + 1);
        g.off = off;
        g.idx.assign(kv.second.begin() + (long)r0, kv.second.begin() + (long)r1);
        for (size_t r = 0; r < g.idx.size(); r++) {
            b->grp_of[g.idx[r]] = (unsigned)b->groups.size();
            b->rank_of[g.idx[r]] = (unsigned)r;
        }
        off += g.words * g.idx.size();
        g.form = gform;
        b->groups.push_back(std::move(g));
    };
    for (auto &kv : key) {
        const uint32_t cnt = (uint32_t)kv.second.size();
        const uint32_t N = (uint32_t)kv.first.first;
        const int W = (int)mpfft_limbs_for_precision(kv.first.second) + 2;
        // MPFFT_FORM_AUTO goes to the plan as it is (mpfft_plan_create
        // resolves it, with the group's count as the batch)
        if (form == MPFFT_FORM_FIXED) {
            if (mpfft::fxp_preferred(N, cnt, W)) {
                add_group(kv, 0, cnt, MPFFT_FORM_FIXED_PANEL);
            } else if (cnt > mpfft::kPanel && mpfft::fxp_eligible(N, cnt) &&
                       !mpfft::fxl_preferred(N, W)) {
                // a class whose last panel would be mostly padding: its full
                // panels on the panel kernels, the rest on the per-lane ones
                const uint32_t full = cnt / mpfft::kPanel * mpfft::kPanel;
                add_group(kv, 0, full, MPFFT_FORM_FIXED_PANEL);
                add_group(kv, full, cnt, MPFFT_FORM_FIXED_LANE);
            } else {
                add_group(kv, 0, cnt, MPFFT_FORM_FIXED_LANE);
            }
        } else if (form == MPFFT_FORM_FIXED_PANEL && !mpfft::fxp_eligible(N, cnt)) {
            add_group(kv, 0, cnt, MPFFT_FORM_FIXED_LANE);
        } else {
            add_group(kv, 0, cnt, form);
        }
    }
    b->words = off;
    HIP_TRY(hipGetDevice(&b->device));
    HIP_TRY(hipStreamCreate(&b->stream));
    for (int s = 0; s < kMaxStreams; s++) {
        HIP_TRY(hipStreamCreate(&b->side[s]));
        HIP_TRY(hipEventCreateWithFlags(&b->join[s], hipEventDisableTiming));
    }
    HIP_TRY(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    HIP_TRY(hipEventCreate(&b->t0));
    HIP_TRY(hipEventCreate(&b->t1));
    // the groups' plans, built on a few host threads (each builds its
    // twiddle table with MPFR)
    const size_t njobs = b->groups.size();
    std::vector<int> rcs(njobs, MPFFT_OK);
    const unsigned nt = std::max(1u, std::min(8u, (unsigned)njobs));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++)
        th.emplace_back([&, t] {
            if (hipSetDevice(b->device) != hipSuccess) {
                for (size_t j = t; j < njobs; j += nt) rcs[j] = MPFFT_ERR_HIP;
                return;
            }
            for (size_t j = t; j < njobs; j += nt) {
                Group &g = b->groups[j];
                rcs[j] = mpfft_plan_create(&g.plan, g.N, g.P, g.form, (unsigned)g.idx.size());
            }
        });
    for (auto &x : th) x.join();
    for (int rc : rcs)
        if (rc) return rc;                   // the destructor frees what was built
    for (auto &g : b->groups) g.form = mpfft_plan_form(g.plan);   // what each plan runs
    *batch = b.release();
    return MPFFT_OK;
}

int mpfft_batch_destroy(mpfft_batch b)
{
    if (!b) return MPFFT_ERR_ARG;
    (void)hipSetDevice(b->device);
    (void)hipDeviceSynchronize();
    delete b;
    return MPFFT_OK;
}

int mpfft_batch_words(mpfft_batch b, size_t *words)
{
    if (!b || !words) return MPFFT_ERR_ARG;
    *words = b->words;
    return MPFFT_OK;
}

int mpfft_batch_info(mpfft_batch b, unsigned int i, size_t *offset, long *limbs, int *form)
{
    if (!b || i >= b->count) return MPFFT_ERR_ARG;
    std::lock_guard<std::mutex> lk(b->mu);
    const Group &g = b->groups[b->grp_of[i]];
    if (offset) *offset = g.off + g.words * b->rank_of[i];
    if (limbs) *limbs = g.L;
    if (form) *form = g.form;
    return MPFFT_OK;
}

int mpfft_batch_tune(mpfft_batch b, const uint32_t *d_in, uint32_t *d_out)
{
    if (!b || !d_in || !d_out) return MPFFT_ERR_ARG;
    std::lock_guard<std::mutex> lk(b->mu);
    HIP_TRY(hipSetDevice(b->device));
    const int rc = tune(b, d_in, d_out);
    b->last = b->stream;
    return rc;
}

int mpfft_batch_schedule(mpfft_batch b, int *streams, double *ms)
{
    if (!b) return MPFFT_ERR_ARG;
    std::lock_guard<std::mutex> lk(b->mu);
    if (streams) *streams = b->nstreams;
    if (ms) *ms = b->graph_ms;
    return MPFFT_OK;
}

int mpfft_batch_execute_limbs(mpfft_batch b, const uint32_t *d_in, uint32_t *d_out, void *stream)
{
    if (!b || !d_in || !d_out) return MPFFT_ERR_ARG;
    std::lock_guard<std::mutex> lk(b->mu);
    HIP_TRY(hipSetDevice(b->device));
    int rc;
    if (!b->tuned && (rc = tune(b, d_in, d_out))) return rc;
    if (b->exec && (b->g_in != d_in || b->g_out != d_out)) {
        if (b->last) HIP_TRY(hipStreamSynchronize(b->last));     // the old graph is done
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    if (!b->exec) {
        if ((rc = capture(b, b->nstreams, d_in, d_out, &b->exec))) return rc;
        b->g_in = d_in;
        b->g_out = d_out;
    }
    hipStream_t st = stream ? (hipStream_t)stream : b->stream;
    HIP_TRY(hipGraphLaunch(b->exec, st));
    b->last = st;
    return MPFFT_OK;
}

int mpfft_batch_check(mpfft_batch b)
{
    if (!b) return MPFFT_ERR_ARG;
    std::lock_guard<std::mutex> lk(b->mu);
    HIP_TRY(hipSetDevice(b->device));
    if (b->last) HIP_TRY(hipStreamSynchronize(b->last));
    int out = MPFFT_OK;
    for (auto &g : b->groups) {
        const int rc = mpfft_plan_check(g.plan);
        if (rc && out == MPFFT_OK) out = rc;
    }
    return out;
}

}  // extern "C"
