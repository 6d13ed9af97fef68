// libmpfft host side: plans, twiddle tables, mpfr_t marshalling and the
// reference entry points (include/mpfft.h).
//
// The reference computes one transform per call with depth-first recursion
// over mpfr_t arrays (lab/serial/mp/serial_mp_fft3.c:121-253).  Here a plan
// flattens that recursion once into leaf tables and per-size node lists,
// keeps the twiddles (the reference's lut() sequence, evaluated by MPFR at
// the plan's precision) resident in HBM, and drives the HIP kernels.
#include "../../include/mpfft.h"
#include "kernels.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include <unistd.h>

static_assert(sizeof(mp_limb_t) == 8, "64-bit GMP limbs expected");

using namespace mpfft;

// Workspace of mpfft_roundtrip_stats_noalloc: the error kernel's per-block
// partial sums (kErrPartial doubles per block of kErrBlock reals), 4 MiB per
// device, in the code object instead of a runtime allocation.  Host code
// never passes this name to a kernel or copy: each device's copy has its
// own address, which scratch_on() looks up once per device.
constexpr size_t kScratchDoubles = size_t(1) << 19;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

thread_local int g_last_error = MPFFT_OK;

// per device: the address of `scratch` there (hipGetSymbolAddress), and a
// static host buffer for the partial sums read back; both guarded by
// g_scratch_mu, which a no-allocation call holds from its launch to its read
// back, so such calls serialize within a process
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices];
double g_scratch_host[kScratchDoubles];

// the device address of `scratch` on device dev (which must be current);
// g_scratch_mu held
int scratch_on(int dev, double **ptr)
{
    if (dev < 0 || dev >= kMaxDevices) return MPFFT_ERR_NODEV;
    if (!g_scratch_ptr[dev]) {
        void *q = nullptr;
        if (hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)) != hipSuccess || !q)
            return MPFFT_ERR_HIP;
        g_scratch_ptr[dev] = static_cast<double *>(q);
    }
    *ptr = g_scratch_ptr[dev];
    return MPFFT_OK;
}

constexpr long kMaxPrec = 2048;          // largest compiled limb count * 32 (panels)
constexpr long kMaxLimbsAll = 32;        // limb counts every kernel family compiles
constexpr long kMaxLimbsMp = 64;         // the plain op orders' kernels
constexpr long kExpLimit = 1L << 29;     // accepted input exponent range
constexpr unsigned long kMaxN = 1UL << 24;
constexpr long kDdMinPrec = 53;          // precisions of the double- and quad-word forms
constexpr long kDdMaxPrec = 106, kQdMaxPrec = 212;
constexpr int kDdExpLimit = 900;         // |exponent| of their inputs (the lowest word
constexpr int kQdExpLimit = 800;         // stays a normal binary64)

#define HIP_TRY(expr)                                          \
    do {                                                       \
        if ((expr) != hipSuccess) return MPFFT_ERR_HIP;        \
    } while (0)

bool is_pow2(unsigned long n) { return n && !(n & (n - 1)); }

bool is_mp(int form)
{
    return form == MPFFT_FORM_MP || form == MPFFT_FORM_MP_C1 || form == MPFFT_FORM_MP_C2;
}

bool is_fixed(int form)
{
    return form == MPFFT_FORM_FIXED || form == MPFFT_FORM_FIXED_LANE ||
           form == MPFFT_FORM_FIXED_PANEL;
}

// the FP64 word forms: binary64 words per real (0: not one of them)
int word_count(int form)
{
    return form == MPFFT_FORM_DD ? 2 : form == MPFFT_FORM_QD ? 4 : 0;
}

// forms that read their input through a range check (MPFFT_ERR_VALUE)
bool has_range_flag(int form) { return is_fixed(form) || word_count(form) > 0; }

// forms whose device data use the SoA-limb layout of include/mpfft.h
bool is_limb_form(int form) { return is_mp(form) || has_range_flag(form); }

// the reference's operation order of a multiprecision form: CODE 3
// (serial_mp_fft3.c), CODE 1 (serial_mp_fft1.c) or CODE 2 (serial_mp_fft2.c)
int mp_code(int form)
{
    return form == MPFFT_FORM_MP_C1 ? 1 : form == MPFFT_FORM_MP_C2 ? 2 : 3;
}

// ---- recursion tree (serial_mp_fft3.c:180-182) ----------------------------

struct Tree {
    struct Leaf { int32_t sn; uint32_t off, n; };
    std::vector<Leaf> leaves;
    std::vector<std::vector<uint32_t>> level;   // level[l]: offsets of size-2^l nodes
    uint32_t lgN = 0;

    void walk(uint32_t n, uint32_t s, int32_t sn, uint32_t off, uint32_t leaf_max)
    {
        if (n <= leaf_max) { leaves.push_back({sn, off, n}); return; }
        uint32_t l = 0; while ((1u << l) < n) l++;
        level[l].push_back(off);
        walk(n / 2, 2 * s, sn, off, leaf_max);
        walk(n / 4, 4 * s, sn + (int32_t)s, off + n / 2, leaf_max);
        walk(n / 4, 4 * s, sn - (int32_t)s, off + 3 * n / 4, leaf_max);
    }
    Tree(uint32_t N, uint32_t leaf_max)
    {
        while ((1u << lgN) < N) lgN++;
        level.resize(lgN + 2);
        walk(N, 1, 0, 0, leaf_max);
    }
};

// ---- device buffer helper -------------------------------------------------

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { if (p) (void)hipFree(p); }
    // bytes is the size of what p holds: 0 after a failed allocation, so a
    // size comparison never takes a failed buffer for a usable one
    int alloc(size_t b)
    {
        if (p) { (void)hipFree(p); p = nullptr; }
        bytes = 0;
        if (!b) return MPFFT_OK;
        if (hipMalloc(&p, b) != hipSuccess) { p = nullptr; return MPFFT_ERR_HIP; }
        bytes = b;
        return MPFFT_OK;
    }
    bool holds(size_t b) const { return p && bytes == b; }
    template <class T> int upload(const std::vector<T> &v)
    {
        int rc = alloc(sizeof(T) * (v.empty() ? 1 : v.size()));
        if (rc) return rc;
        if (!v.empty())
            HIP_TRY(hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
        return MPFFT_OK;
    }
    template <class T> T *as() const { return (T *)p; }
};

// page-locked host memory for the staging of mpfr_t data (transfers at full
// PCIe rate, asynchronous with respect to the host)
struct PinnedBuf {
    void *p = nullptr;
    size_t bytes = 0;
    ~PinnedBuf() { if (p) (void)hipHostFree(p); }
    int alloc(size_t b)
    {
        if (b == bytes && p) return MPFFT_OK;
        if (p) { (void)hipHostFree(p); p = nullptr; }
        bytes = 0;
        if (hipHostMalloc(&p, b, hipHostMallocDefault) != hipSuccess) {
            p = nullptr;
            return MPFFT_ERR_HIP;
        }
        bytes = b;
        return MPFFT_OK;
    }
    template <class T> T *as() const { return (T *)p; }
};

// ---- host worker pool -------------------------------------------------------

// The threads that marshal mpfr_t arrays, started on first use and kept for
// the process (a thread spawned per call costs tens of microseconds, the
// order of a whole small call).  One job at a time: start(n, f) runs f(w)
// for w < n on n of the pool's threads and returns; wait() returns when all
// n have finished.  Callers hold user() across start .. wait.
class HostPool {
public:
    static HostPool &get()
    {
        // never destroyed: its threads end with the process.  A child made
        // by fork() inherits the pool but not its threads: it starts its own
        // (the parent's object is left alone)
        static std::mutex mk;
        static HostPool *pool = nullptr;
        std::lock_guard<std::mutex> lk(mk);
        if (!pool || pool->pid_ != getpid()) pool = new HostPool();
        return *pool;
    }
    unsigned size() const { return (unsigned)th_.size(); }
    std::mutex &user() { return user_; }
    void start(unsigned n, std::function<void(unsigned)> f)
    {
        std::lock_guard<std::mutex> lk(mu_);
        job_ = std::move(f);
        njob_ = std::min(n, size());
        left_ = njob_;
        gen_++;
        cv_.notify_all();
    }
    void wait()
    {
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&] { return left_ == 0; });
        job_ = nullptr;
    }

private:
    HostPool() : pid_(getpid())
    {
        unsigned n = std::thread::hardware_concurrency();
        n = std::max(1u, std::min(n, 16u));
        for (unsigned i = 0; i < n; i++) th_.emplace_back([this, i] { loop(i); });
    }
    void loop(unsigned i)
    {
        unsigned seen = 0;
        for (;;) {
            std::function<void(unsigned)> f;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
                if (i >= njob_) continue;
                f = job_;
            }
            f(i);
            std::lock_guard<std::mutex> lk(mu_);
            if (--left_ == 0) done_.notify_all();
        }
    }
    pid_t pid_;
    std::vector<std::thread> th_;
    std::mutex mu_, user_;
    std::condition_variable cv_, done_;
    std::function<void(unsigned)> job_;
    unsigned njob_ = 0, left_ = 0, gen_ = 0;
};

// ---- mpfr_t <-> SoA limbs --------------------------------------------------

int mpfr_put(mpfr_srcptr x, uint32_t *hdr, uint32_t *limb, size_t M, size_t i, long L)
{
    if (mpfr_nan_p(x) || mpfr_inf_p(x)) return MPFFT_ERR_VALUE;
    const uint32_t neg = mpfr_signbit(x) ? 1u : 0u;
    if (mpfr_zero_p(x)) {
        hdr[i] = 0x80000000u | neg;
        for (long j = 0; j < L; j++) limb[j * M + i] = 0;
        return MPFFT_OK;
    }
    const long e = mpfr_get_exp(x);
    if (e >= kExpLimit || e <= -kExpLimit) return MPFFT_ERR_VALUE;
    hdr[i] = ((uint32_t)e << 1) | neg;
    const mp_limb_t *d = (const mp_limb_t *)mpfr_custom_get_significand(x);
    const long n64 = (mpfr_get_prec(x) + 63) / 64;
    for (long j = 0; j < L; j++) {
        const long t = j >> 1;               // 64-bit limb index from the top
        uint32_t w = 0;
        if (t < n64) {
            const mp_limb_t v = d[n64 - 1 - t];
            w = (j & 1) ? (uint32_t)v : (uint32_t)(v >> 32);
        }
        limb[j * M + i] = w;
    }
    return MPFFT_OK;
}

// mpfr_put for the points k0 .. k1-1 of one array (SoA positions base + k),
// in blocks of 32 points: the headers and significand pointers first, then
// the limb rows one at a time, so the stores run along each row
int mpfr_put_range(mpfr_t *x, unsigned long k0, unsigned long k1, size_t base, uint32_t *hdr,
                   uint32_t *limb, size_t M, long L)
{
    constexpr unsigned kBlk = 32;
    const mp_limb_t *d[kBlk];
    long n64[kBlk];
    for (unsigned long b0 = k0; b0 < k1; b0 += kBlk) {
        const unsigned nb = (unsigned)std::min<unsigned long>(kBlk, k1 - b0);
        for (unsigned u = 0; u < nb; u++) {
            mpfr_srcptr v = x[b0 + u];
            const size_t i = base + b0 + u;
            d[u] = nullptr;
            n64[u] = 0;
            if (mpfr_nan_p(v) || mpfr_inf_p(v)) return MPFFT_ERR_VALUE;
            const uint32_t neg = mpfr_signbit(v) ? 1u : 0u;
            if (mpfr_zero_p(v)) { hdr[i] = 0x80000000u | neg; continue; }
            const long e = mpfr_get_exp(v);
            if (e >= kExpLimit || e <= -kExpLimit) return MPFFT_ERR_VALUE;
            hdr[i] = ((uint32_t)e << 1) | neg;
            d[u] = (const mp_limb_t *)mpfr_custom_get_significand(v);
            n64[u] = (mpfr_get_prec(v) + 63) / 64;
        }
        for (long j = 0; j < L; j++) {
            const long t = j >> 1;           // 64-bit limb index from the top
            uint32_t *row = limb + j * M + base + b0;
            for (unsigned u = 0; u < nb; u++) {
                uint32_t w = 0;
                if (t < n64[u]) {
                    const mp_limb_t v = d[u][n64[u] - 1 - t];
                    w = (j & 1) ? (uint32_t)v : (uint32_t)(v >> 32);
                }
                row[u] = w;
            }
        }
    }
    return MPFFT_OK;
}

void mpfr_take(mpfr_ptr x, const uint32_t *hdr, const uint32_t *limb, size_t M, size_t i, long L)
{
    const uint32_t h = hdr[i], neg = h & 1u;
    if ((h & ~1u) == 0x80000000u) { mpfr_set_zero(x, neg ? -1 : 1); return; }
    const mpfr_prec_t prec = mpfr_get_prec(x);
    const long n64 = (prec + 63) / 64;
    mp_limb_t *d = (mp_limb_t *)mpfr_custom_get_significand(x);
    for (long t = 0; t < n64; t++) {
        const long j = 2 * t;
        const uint64_t hi = j < L ? limb[j * M + i] : 0;
        const uint64_t lo = j + 1 < L ? limb[(j + 1) * M + i] : 0;
        d[n64 - 1 - t] = (mp_limb_t)((hi << 32) | lo);
    }
    const mpfr_exp_t e = (mpfr_exp_t)((int32_t)h >> 1);
    mpfr_custom_init_set(x, neg ? -MPFR_REGULAR_KIND : MPFR_REGULAR_KIND, e, prec, d);
}

// mpfr_take for the points k0 .. k1-1 of one array (SoA positions base + k)
// in blocks of 32 points: the limb rows read one at a time along the row
// (32 words each), then each point's significand assembled
void mpfr_take_range(mpfr_t *x, unsigned long k0, unsigned long k1, size_t base,
                     const uint32_t *hdr, const uint32_t *limb, size_t M, long L)
{
    constexpr unsigned kBlk = 32;
    std::vector<uint32_t> tmp((size_t)L * kBlk);
    for (unsigned long b0 = k0; b0 < k1; b0 += kBlk) {
        const unsigned nb = (unsigned)std::min<unsigned long>(kBlk, k1 - b0);
        for (long j = 0; j < L; j++) {
            const uint32_t *row = limb + j * M + base + b0;
            for (unsigned u = 0; u < nb; u++) tmp[(size_t)j * kBlk + u] = row[u];
        }
        for (unsigned u = 0; u < nb; u++) {
            mpfr_ptr v = x[b0 + u];
            const uint32_t h = hdr[base + b0 + u], neg = h & 1u;
            if ((h & ~1u) == 0x80000000u) { mpfr_set_zero(v, neg ? -1 : 1); continue; }
            const mpfr_prec_t prec = mpfr_get_prec(v);
            const long n64 = (prec + 63) / 64;
            mp_limb_t *d = (mp_limb_t *)mpfr_custom_get_significand(v);
            for (long t = 0; t < n64; t++) {
                const long j = 2 * t;
                const uint64_t hi = j < L ? tmp[(size_t)j * kBlk + u] : 0;
                const uint64_t lo = j + 1 < L ? tmp[(size_t)(j + 1) * kBlk + u] : 0;
                d[n64 - 1 - t] = (mp_limb_t)((hi << 32) | lo);
            }
            const mpfr_exp_t e = (mpfr_exp_t)((int32_t)h >> 1);
            mpfr_custom_init_set(v, neg ? -MPFR_REGULAR_KIND : MPFR_REGULAR_KIND, e, prec, d);
        }
    }
}

// ---- twiddles: the reference's lut() (serial_mp_fft3.c:71-82) ---------------

// r, s, d (code 3) or r, i (codes 1, 2) of k = 0..N/4-1 at precision P, as
// SoA (M = 3N/4 or N/2)
void build_mp_lut(unsigned long N, long P, long L, int code, std::vector<uint32_t> &buf)
{
    const size_t Q = N / 4, M = (code == 3 ? 3 : 2) * Q;
    buf.assign(M * (L + 1), 0);
    uint32_t *hdr = buf.data(), *limb = hdr + M;
    unsigned nt = std::thread::hardware_concurrency();
    if (nt < 1) nt = 1;
    if (nt > 16) nt = 16;
    if (Q < 256) nt = 1;
    auto work = [&](size_t k0, size_t k1) {
        mpfr_t step, ang, c, s, t;
        mpfr_inits2(P, step, ang, c, s, t, (mpfr_ptr)0);
        // RN(RN(pi) / N): exact division by a power of two
        mpfr_const_pi(step, MPFR_RNDN);
        mpfr_div_ui(step, step, N, MPFR_RNDN);
        for (size_t k = k0; k < k1; k++) {
            // the angle is the rounded product RN(step * 2k), as in the reference
            mpfr_mul_ui(ang, step, 2 * k, MPFR_RNDN);
            mpfr_cos(c, ang, MPFR_RNDN);
            mpfr_sin(s, ang, MPFR_RNDN);
            mpfr_put(c, hdr, limb, M, k, L);
            if (code != 3) {
                mpfr_put(s, hdr, limb, M, Q + k, L);
                continue;
            }
            mpfr_add(t, c, s, MPFR_RNDN);
            mpfr_put(t, hdr, limb, M, Q + k, L);
            mpfr_sub(t, c, s, MPFR_RNDN);
            mpfr_put(t, hdr, limb, M, 2 * Q + k, L);
        }
        mpfr_clears(step, ang, c, s, t, (mpfr_ptr)0);
        mpfr_free_cache();
    };
    std::vector<std::thread> th;
    for (unsigned i = 0; i < nt; i++)
        th.emplace_back(work, Q * i / nt, Q * (i + 1) / nt);
    for (auto &x : th) x.join();
}

}  // namespace

// ---- plan ------------------------------------------------------------------

struct mpfft_plan_s {
    unsigned long N = 0;
    long P = 0;
    int form = 0;
    unsigned batch = 1;
    long L = 0;              // device limbs per real (mp form)
    size_t words = 0;        // uint32 words (mp) or doubles (double forms) per transform
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;

    // mp form
    DevBuf lut, pos_sn, pos_nr, nodes, aos;
    LeafArgs leaf{};
    std::vector<LevelArgs> levels;
    DevBuf sub_list, sub_rel;            // subtrees finished in LDS
    SubArgs sub{};
    bool use_sub = false;
    // fixed form
    DevBuf fx_tw, fx_buf0, fx_buf1, fx_flag;
    DevBuf fx_eoff;          // block exponents: batch, then batch x fx_eoff_blocks(N) partials
    int fx_lim = 0;
    uint32_t fx_ntw = 0;
    bool fx_panel = false;   // panel kernels (matrix-core twiddle products)
    int fx_prb = 2;          // panel kernels: log2 of the main radix (fxp_plan)
    int fx_npass = 0;        // passes of the panel or lane plan
    bool fx_fs = false;      // panel kernels: the four-step plan (fxs_plan)
    FxsPass fxs[kFxsMaxPasses] = {};
    uint32_t fxs_c16e = 0;   // its table entry of w^0 = 1
    uint32_t ncu = 0;        // compute units of the plan's device
    // double forms
    DevBuf tw, dleaf_sn, dleaf_off, dleaf_n, dlvl_off, dlvl_begin;
    DTables dt{};
    // staging for host / mpfr execution
    DevBuf d_in, d_out, d_aux;
    DevBuf errbuf;                       // per-block partial sums of the error statistics
    PinnedBuf h_stage_in, h_stage_out;
    double timing[4] = {0, 0, 0, 0};     // last mpfr_t call: marshal in, compute, marshal out, total (ms)
    hipEvent_t time_ev[3] = {};          // mpfr_t calls: upload start, upload end, compute end

    ~mpfft_plan_s()
    {
        for (auto e : time_ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

int build_mp_plan(mpfft_plan_s *p)
{
    const unsigned long N = p->N;
    Tree t((uint32_t)N, 4);
    // per output position: its leaf's sn, size and the position inside it
    std::vector<int32_t> psn(N);
    std::vector<uint32_t> pnr(N);
    for (const auto &lf : t.leaves)
        for (uint32_t r = 0; r < lf.n; r++) {
            psn[lf.off + r] = lf.sn;
            pnr[lf.off + r] = lf.n | (r << 8);
        }
    int rc;
    if ((rc = p->pos_sn.upload(psn)) || (rc = p->pos_nr.upload(pnr))) return rc;
    std::vector<uint32_t> all;
    std::vector<size_t> first(t.level.size(), 0);
    for (size_t l = 0; l < t.level.size(); l++) {
        first[l] = all.size();
        all.insert(all.end(), t.level[l].begin(), t.level[l].end());
    }
    if ((rc = p->nodes.upload(all))) return rc;
    // per level n >= 8: entries k < n/4 gathered from the base table
    // (entry k of level n = base entry k N/n), so a wave's lanes read
    // consecutive entries
    const int rows = mp_code(p->form) == 3 ? 3 : 2;
    size_t E = 0;
    std::vector<uint32_t> lvl_off(t.lgN + 2, 0);
    for (uint32_t l = 3; l <= t.lgN; l++) { lvl_off[l] = (uint32_t)E; E += (1UL << l) / 4; }
    if (N >= 8) {
        std::vector<uint32_t> base;
        build_mp_lut(N, p->P, p->L, mp_code(p->form), base);
        const size_t Q = N / 4, MB = rows * Q, ML = rows * E;
        std::vector<uint32_t> lut(ML * (p->L + 1));
        for (uint32_t l = 3; l <= t.lgN; l++) {
            const size_t n = 1UL << l, step = N / n;
            for (size_t k = 0; k < n / 4; k++)
                for (int u = 0; u < rows; u++) {
                    const size_t src = u * Q + k * step, dst = u * E + lvl_off[l] + k;
                    for (long j = 0; j <= p->L; j++) lut[j * ML + dst] = base[j * MB + src];
                }
        }
        if ((rc = p->lut.upload(lut))) return rc;
    }
    // element-major staging copy of the input for the leaf gathers
    const uint32_t aw = (uint32_t)((p->L + 1 + 3) & ~3L);
    if ((rc = p->aos.alloc(sizeof(uint32_t) * (size_t)aw * 2 * N * p->batch))) return rc;
    LeafArgs &la = p->leaf;
    la.aos = p->aos.as<uint32_t>();
    la.aos_words = aw;
    la.aos_stride = (uint64_t)aw * 2 * N;
    la.pos_sn = p->pos_sn.as<int32_t>(); la.pos_nr = p->pos_nr.as<uint32_t>();
    la.npos = N >= 2 ? (uint32_t)N : 0u;
    la.N = (uint32_t)N; la.P = (int)p->P; la.in_swap = 0;
    la.in_stride = la.stride = p->words;
    // subtrees finished in LDS: the levels up to node size S run inside
    // them (MPFFT_SUB_SIZE overrides S; 0 keeps every level on HBM)
    uint32_t S = N >= 8 ? mp_sub_size((int)p->L) : 0u;
    if (const char *e = getenv("MPFFT_SUB_SIZE")) S = (uint32_t)strtoul(e, nullptr, 10);
    if (S && (!is_pow2(S) || S < 8)) S = 0;
    if (S > N) S = (uint32_t)N;
    const uint32_t nthr = S ? std::min(256u, std::max(64u, S / 2)) : 0u;
    if (S && sizeof(uint32_t) * (p->L + 1) * (2 * (size_t)nthr + 2 * (size_t)S) > 160 * 1024)
        S = 0;                                    // does not fit a workgroup's LDS
    p->use_sub = S != 0;
    uint32_t lgS = 0;
    while ((1u << lgS) < S) lgS++;
    if (p->use_sub) {
        // maximal subtrees of size <= S: sizes S and S/2 when N > S
        std::vector<uint32_t> list;
        std::function<void(uint32_t, uint32_t)> front = [&](uint32_t n, uint32_t off) {
            if (n <= S) { list.push_back(off << 1 | (n == S ? 0u : 1u)); return; }
            front(n / 2, off);
            front(n / 4, off + n / 2);
            front(n / 4, off + 3 * n / 4);
        };
        front((uint32_t)N, 0);
        std::vector<uint32_t> rel;
        SubArgs &sa = p->sub;
        sa = SubArgs{};
        for (int kind = 0; kind < 2; kind++) {
            const uint32_t s = kind ? S / 2 : S;
            sa.lgs[kind] = kind ? lgS - 1 : lgS;
            if (s < 8) continue;
            Tree ts(s, 4);
            for (uint32_t l = 3; l <= ts.lgN && l < (uint32_t)kSubMaxLevel; l++) {
                sa.rel_first[kind][l] = (uint32_t)rel.size();
                sa.rel_cnt[kind][l] = (uint32_t)ts.level[l].size();
                rel.insert(rel.end(), ts.level[l].begin(), ts.level[l].end());
            }
        }
        if ((rc = p->sub_list.upload(list)) || (rc = p->sub_rel.upload(rel))) return rc;
        sa.lut = p->lut.as<uint32_t>();
        sa.lut_e = (uint32_t)E;
        sa.lut_m = (uint32_t)(rows * E);
        sa.code = mp_code(p->form);
        sa.sub = p->sub_list.as<uint32_t>();
        sa.nsub = (uint32_t)list.size();
        sa.rel = p->sub_rel.as<uint32_t>();
        for (uint32_t l = 3; l <= lgS && l < (uint32_t)kSubMaxLevel; l++) sa.lut_off[l] = lvl_off[l];
        sa.nthr = nthr;
    }
    p->levels.clear();
    for (uint32_t l = p->use_sub ? lgS + 1 : 3; l <= t.lgN; l++) {
        LevelArgs a{};
        const uint32_t n = 1u << l, cnt = (uint32_t)t.level[l].size();
        a.lut = p->lut.as<uint32_t>();
        a.code = mp_code(p->form);
        a.lut_e = (uint32_t)E;
        a.lut_m = (uint32_t)(rows * E);
        a.lut_off = lvl_off[l];
        a.node_off = p->nodes.as<uint32_t>() + first[l];
        a.stride = p->words;
        a.N = (uint32_t)N; a.n = n; a.cnt = cnt;
        a.P = (int)p->P;
        p->levels.push_back(a);
    }
    return MPFFT_OK;
}

// Fixed form: for every radix-4 pass (Ns = 1 or 2, then times 4 while
// Ns < N), entries 3k + r - 1 (k < Ns, r = 1..3) hold w^t = c + i d =
// exp(-2 pi i t / N), t = r k N / (4 Ns), rounded to F = 32 (L + 1)
// fraction bits (MPFR at F + 64 bits, then one rounding to nearest).
//   per-lane kernels: c, d - c and c + d, each as DT B-bit digits of its
//     magnitude (least significant first) and a sign word; SoA over the
//     3 T3 reals;
//   panel kernels: c, d and -d per entry (entry-major), each as the byte
//     table of its balanced base-256 digits that the matrix-core product
//     reads (byte u = digit U - u, zero outside the digits; mf_geom).
int build_fx_plan(mpfft_plan_s *p)
{
    const unsigned long N = p->N;
    const long W = p->L + 2, F = 32 * (W - 1);
    const long B = fx_lane_digit_bits((int)W);            // 29, or 28 at W = 66
    const long DT = (F + 1 + B - 1) / B;
    int lg0 = 0;
    while ((1UL << lg0) < N) lg0++;
    if (p->form == MPFFT_FORM_FIXED_PANEL && !fxp_eligible((uint32_t)N, p->batch))
        return MPFFT_ERR_SIZE;
    p->fx_panel = p->form == MPFFT_FORM_FIXED_PANEL ||
                  (p->form == MPFFT_FORM_FIXED && fxp_preferred((uint32_t)N, p->batch, (int)W));
    int ncu = 0;
    HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, p->device));
    p->ncu = ncu > 0 ? (uint32_t)ncu : 1u;
    const MfGeom G = mf_geom((int)W);
    std::vector<unsigned long> ent;               // t of every entry, in table order
    if (p->fx_panel) {
        // the panel passes (fxp_plan), then the butterflies' constants.  The
        // main radix: the width's default, or MPFFT_PANEL_RADIX (4, 8 or 16)
        // from the environment where the width allows it (for timing the
        // alternatives)
        p->fx_prb = fxp_radix_default((int)W);
        bool radix_set = false;
        if (const char *e = getenv("MPFFT_PANEL_RADIX")) {
            const int r = atoi(e), rb = r == 16 ? 4 : r == 8 ? 3 : r == 4 ? 2 : 0;
            if (rb && fxp_radix_ok((int)W, rb)) { p->fx_prb = rb; radix_set = true; }
        }
        // the four-step plan (its first and last passes convert the input
        // and output) where N = 16^m, m >= 3, when MPFFT_PANEL_PLAN=fs asks
        // for it: measured on MI355X at N = 2^16, batch 64, it ties the plan
        // with conversion passes (1024 bits 14.4 k against 14.4-14.9 k FFT/s,
        // 256 bits 49.4 k against 50.0 k): its converting passes save what
        // the conversions cost, and the transposition between its steps
        // costs about as much again
        const char *pe = getenv("MPFFT_PANEL_PLAN");
        p->fx_fs = fxs_eligible((uint32_t)N, (int)W) && !radix_set &&
                   (pe && strcmp(pe, "fs") == 0);
        uint32_t nent = 0;
        if (p->fx_fs) {
            const int np = fxs_plan((uint32_t)N, p->batch, p->fxs, &nent);
            for (int i = 0; i < np; i++) {
                const FxsPass &q = p->fxs[i];
                if (q.mode == kFxsTwist) {
                    // w^(k1 n2), n2 = j + 16 r, in the order (k1, j, r)
                    for (unsigned long k1 = 0; k1 < N / 256; k1++)
                        for (unsigned long j = 0; j < 16; j++)
                            for (unsigned long r = 0; r < 16; r++)
                                ent.push_back((k1 * (j + 16 * r)) % N);
                } else if (q.mode != kFxsFirst) {
                    for (unsigned long k = 0; k < q.Ns; k++)
                        for (unsigned long r = 1; r < 16; r++)
                            ent.push_back(r * k * (N / (16UL * q.Ns)));
                }
            }
            p->fx_npass = np;
            p->fxs_c16e = nent + 1;
        } else {
            FxpPass ps[32];
            const int np = fxp_plan((uint32_t)N, (int)W, p->fx_prb, ps, &nent);
            for (int i = 0; i < np; i++)
                for (unsigned long k = 0; k < ps[i].Ns; k++)
                    for (unsigned long r = 1; r < ps[i].R; r++)
                        ent.push_back(r * k * (N / ((unsigned long)ps[i].R * ps[i].Ns)));
            p->fx_npass = np;
        }
        if (ent.size() != nent) return MPFFT_ERR_ARG;
        // the butterflies' constants (kPanelConstEntries of them, fxp_plan)
        for (unsigned long t : {N / 8, 0UL, N / 16, 3 * N / 16, 9 * N / 16, N / 8, 3 * N / 8})
            ent.push_back(t);
    } else {
        // the lane passes (fxl_plan): radix 16 r-major, radix 4 and 2 k-major,
        // then w16^m, m = 0 .. 9, where a radix-16 pass needs them
        MwPass ps[32];
        uint32_t nent = 0;
        const int np = N > 1 ? fxl_plan((uint32_t)N, (int)W, ps, &nent) : 0;
        bool r16 = false;
        for (int i = 0; i < np; i++) {
            const unsigned long R = ps[i].R, Ns = ps[i].Ns;
            if (R == 16) {
                r16 = true;
                for (unsigned long r = 1; r < R; r++)
                    for (unsigned long k = 0; k < Ns; k++) ent.push_back(r * k * (N / (R * Ns)));
            } else {
                for (unsigned long k = 0; k < Ns; k++)
                    for (unsigned long r = 1; r < R; r++) ent.push_back(r * k * (N / (R * Ns)));
            }
        }
        if (ent.size() != nent) return MPFFT_ERR_ARG;
        if (r16)
            for (unsigned long m = 0; m < 10; m++) ent.push_back(m * (N / 16));
        p->fx_npass = np;
    }
    const size_t T3 = ent.size(), MT = 3 * T3;
    const size_t tabn = p->fx_panel ? T3 * 3 * (size_t)G.RW : MT * (DT + 1);
    std::vector<uint32_t> tab(tabn > 0 ? tabn : 1, 0);
    unsigned nt = std::thread::hardware_concurrency();
    nt = std::max(1u, std::min(nt, 16u));
    if (T3 < 256) nt = 1;
    auto work = [&](size_t t0, size_t t1) {
        mpfr_t th, c, d, v;
        mpfr_inits2(F + 64, th, c, d, v, (mpfr_ptr)0);
        mpz_t z;
        mpz_init(z);
        std::vector<uint32_t> words((size_t)(F / 32 + 4));
        std::vector<uint8_t> bytes((size_t)G.NTB + 8), R((size_t)G.RW * 4);
        for (size_t e = t0; e < t1; e++) {
            const size_t t = ent[e];
            mpfr_const_pi(th, MPFR_RNDN);
            mpfr_mul_ui(th, th, 2 * t, MPFR_RNDN);
            mpfr_div_ui(th, th, N, MPFR_RNDN);
            mpfr_sin_cos(d, c, th, MPFR_RNDN);
            mpfr_neg(d, d, MPFR_RNDN);                    // w = cos - i sin
            for (int u = 0; u < 3; u++) {
                if (p->fx_panel) {
                    if (u == 0) mpfr_set(v, c, MPFR_RNDN);
                    else if (u == 1) mpfr_set(v, d, MPFR_RNDN);
                    else mpfr_neg(v, d, MPFR_RNDN);
                } else {
                    if (u == 0) mpfr_set(v, c, MPFR_RNDN);
                    else if (u == 1) mpfr_sub(v, d, c, MPFR_RNDN);
                    else mpfr_add(v, c, d, MPFR_RNDN);
                }
                mpfr_mul_2si(v, v, F, MPFR_RNDN);
                mpfr_get_z(z, v, MPFR_RNDN);
                if (p->fx_panel) {
                    // two's complement over NTB bytes, then balanced digits:
                    // bytes of (T + 0x80..80) minus 128 (|T| <= 2^F, so the
                    // sum stays in [0, 256^NTB))
                    const bool neg = mpz_sgn(z) < 0;
                    mpz_abs(z, z);
                    std::fill(bytes.begin(), bytes.end(), 0);
                    size_t cnt = 0;
                    mpz_export(bytes.data(), &cnt, -1, 1, 0, 0, z);
                    if (neg) {
                        unsigned cy = 1;
                        for (int i = 0; i < G.NTB; i++) {
                            const unsigned x = (unsigned)(uint8_t)~bytes[i] + cy;
                            bytes[i] = (uint8_t)x;
                            cy = x >> 8;
                        }
                    }
                    unsigned cy = 0;
                    for (int i = 0; i < G.NTB; i++) {
                        const unsigned x = (unsigned)bytes[i] + 0x80u + cy;
                        bytes[i] = (uint8_t)((x & 0xffu) ^ 0x80u);   // digit as int8
                        cy = x >> 8;
                    }
                    std::fill(R.begin(), R.end(), 0);
                    for (int b = 0; b < 4 * G.RW; b++) {
                        const int idx = G.U - b;
                        if (idx >= 0 && idx < G.NTB) R[b] = bytes[idx];
                    }
                    uint32_t *dst = tab.data() + (e * 3 + u) * (size_t)G.RW;
                    for (int w = 0; w < G.RW; w++)
                        dst[w] = (uint32_t)R[4 * w] | (uint32_t)R[4 * w + 1] << 8 |
                                 (uint32_t)R[4 * w + 2] << 16 | (uint32_t)R[4 * w + 3] << 24;
                    continue;
                }
                const size_t q = u * T3 + e;
                tab[(size_t)DT * MT + q] = mpz_sgn(z) < 0 ? 1u : 0u;
                mpz_abs(z, z);
                std::fill(words.begin(), words.end(), 0u);
                size_t cnt = 0;
                mpz_export(words.data(), &cnt, -1, sizeof(uint32_t), 0, 0, z);
                for (long k = 0; k < DT; k++) {
                    const long bit = B * k, wi = bit >> 5, sh = bit & 31;
                    const uint64_t lo = words[wi], hi = words[wi + 1];
                    tab[(size_t)k * MT + q] =
                        (uint32_t)(((hi << 32 | lo) >> sh) & ((1ull << B) - 1));
                }
            }
        }
        mpz_clear(z);
        mpfr_clears(th, c, d, v, (mpfr_ptr)0);
        mpfr_free_cache();
    };
    std::vector<std::thread> th;
    for (unsigned i = 0; i < nt; i++) th.emplace_back(work, T3 * i / nt, T3 * (i + 1) / nt);
    for (auto &x : th) x.join();
    int rc;
    if ((rc = p->fx_tw.upload(tab))) return rc;
    const size_t fxw = 2 * N * (size_t)W * (p->fx_panel ? fxp_groups(p->batch) * kPanel : p->batch);
    if ((rc = p->fx_buf0.alloc(sizeof(uint32_t) * fxw)) ||
        (rc = p->fx_buf1.alloc(sizeof(uint32_t) * fxw)) ||
        (rc = p->fx_flag.upload(std::vector<uint32_t>{0u})) ||
        (rc = p->fx_eoff.alloc(sizeof(int32_t) * (size_t)p->batch *
                               (1 + fx_eoff_blocks((uint32_t)N)))))
        return rc;
    p->fx_ntw = (uint32_t)T3;
    p->fx_lim = 29 - lg0;
    return MPFFT_OK;
}

// Double- and quad-word forms: twiddles of the passes (pass after pass,
// entry (r - 1) Ns + k = w^(r k N / (R Ns)), k fastest) as K words of the
// real part, then K of the imaginary part, each word the binary64 nearest to
// what the words before it leave of the value (MPFR at 512 bits), and the
// word ping-pong buffers.
int build_dd_plan(mpfft_plan_s *p)
{
    const int K = word_count(p->form);
    const unsigned long N = p->N;
    int lg0 = 0;
    while ((1UL << lg0) < N) lg0++;
    (void)lg0;
    MwPass ps[32];
    uint32_t nent = 0;
    const int np = N > 1 ? mw_plan((uint32_t)N, K, ps, &nent) : 0;
    std::vector<unsigned long> ent;
    bool r16 = false;
    for (int i = 0; i < np; i++) {
        for (unsigned long r = 1; r < ps[i].R; r++)
            for (unsigned long k = 0; k < ps[i].Ns; k++)
                ent.push_back(r * k * (N / ((unsigned long)ps[i].R * ps[i].Ns)));
        r16 |= ps[i].R == 16;
    }
    if (ent.size() != nent) return MPFFT_ERR_ARG;
    if (r16)
        for (unsigned long m = 0; m < 10; m++) ent.push_back(m * (N / 16));   // w16^m
    p->fx_npass = np;
    std::vector<double> tab(2 * K * (ent.empty() ? 1 : ent.size()), 0.0);
    unsigned nt = std::max(1u, std::min(std::thread::hardware_concurrency(), 16u));
    if (ent.size() < 1024) nt = 1;
    auto work = [&](size_t e0, size_t e1) {
        mpfr_t th, c, d, t;
        mpfr_inits2(512, th, c, d, t, (mpfr_ptr)0);
        for (size_t e = e0; e < e1; e++) {
            mpfr_const_pi(th, MPFR_RNDN);
            mpfr_mul_ui(th, th, 2 * ent[e], MPFR_RNDN);
            mpfr_div_ui(th, th, N, MPFR_RNDN);
            mpfr_sin_cos(d, c, th, MPFR_RNDN);
            mpfr_neg(d, d, MPFR_RNDN);                    // w = cos - i sin
            for (int u = 0; u < 2; u++) {
                mpfr_set(t, u ? d : c, MPFR_RNDN);
                for (int w = 0; w < K; w++) {
                    const double v = mpfr_get_d(t, MPFR_RNDN);
                    tab[2 * K * e + K * u + w] = v;
                    mpfr_sub_d(t, t, v, MPFR_RNDN);       // exact at 512 bits
                }
            }
        }
        mpfr_clears(th, c, d, t, (mpfr_ptr)0);
        mpfr_free_cache();
    };
    std::vector<std::thread> th;
    for (unsigned i = 0; i < nt; i++)
        th.emplace_back(work, ent.size() * i / nt, ent.size() * (i + 1) / nt);
    for (auto &x : th) x.join();
    int rc;
    if ((rc = p->fx_tw.upload(tab))) return rc;
    const size_t bytes = K * sizeof(double) * 2 * N * (size_t)p->batch;
    if ((rc = p->fx_buf0.alloc(bytes)) || (rc = p->fx_buf1.alloc(bytes)) ||
        (rc = p->fx_flag.upload(std::vector<uint32_t>{0u})) ||
        (rc = p->fx_eoff.alloc(sizeof(int32_t) * (size_t)p->batch *
                               (1 + fx_eoff_blocks((uint32_t)N)))))
        return rc;
    p->fx_ntw = (uint32_t)ent.size();
    p->fx_lim = K == 2 ? kDdExpLimit : kQdExpLimit;
    return MPFFT_OK;
}

int run_fx(mpfft_plan_s *p, const uint32_t *d_in, uint32_t *d_out, hipStream_t st,
           int s0 = 0, int s1 = 1 << 30)
{
    if (s0 < 0 || s0 > s1) return MPFFT_ERR_ARG;
    FxArgs a{};
    a.in = d_in; a.in_stride = p->words;
    a.out = d_out; a.out_stride = p->words;
    a.buf0 = p->fx_buf0.as<uint32_t>(); a.buf1 = p->fx_buf1.as<uint32_t>();
    a.stride = 2 * p->N * (uint64_t)(p->L + 2);
    a.tw = p->fx_tw.as<uint32_t>();
    a.ntw = p->fx_ntw;
    a.flag = p->fx_flag.as<uint32_t>();
    a.N = (uint32_t)p->N; a.batch = p->batch;
    a.lim = p->fx_lim; a.P = (int)p->P;
    a.ncu = p->ncu;
    a.prb = p->fx_prb;
    a.eoff = p->fx_eoff.as<int32_t>();
    a.eoff_part = a.eoff + p->batch;
    a.eoff_nblk = fx_eoff_blocks((uint32_t)p->N);
    const int K = word_count(p->form);
    if (K) a.stride = (uint64_t)p->N * K;                      // K/2 double2 per real
    // stage 0 reads the input: the block exponents come first
    if (s0 == 0 && s1 > 0 && fx_eoff_launch(a, st) != hipSuccess) return MPFFT_ERR_HIP;
    hipError_t e = K            ? dd_fft_launch(K, (int)p->L, (int)p->P, a, s0, s1, st)
                   : p->fx_fs    ? fxs_fft_launch((int)p->L, (int)p->P, a, p->fxs, p->fx_npass,
                                                  p->fxs_c16e, s0, s1, st)
                   : p->fx_panel ? fxp_fft_launch((int)p->L, (int)p->P, a, s0, s1, st)
                                 : fx_fft_launch((int)p->L, (int)p->P, a, s0, s1, st);
    return e == hipSuccess ? MPFFT_OK : MPFFT_ERR_HIP;
}

// read and clear the fixed form's range flag (waits for the plan's stream)
int fx_check(mpfft_plan_s *p)
{
    if (!has_range_flag(p->form)) return MPFFT_OK;
    uint32_t f = 0;
    HIP_TRY(hipStreamSynchronize(p->stream));
    HIP_TRY(hipMemcpy(&f, p->fx_flag.p, sizeof f, hipMemcpyDeviceToHost));
    if (f) {
        const uint32_t zero = 0;
        HIP_TRY(hipMemcpy(p->fx_flag.p, &zero, sizeof zero, hipMemcpyHostToDevice));
        return MPFFT_ERR_VALUE;
    }
    return MPFFT_OK;
}

int build_d_plan(mpfft_plan_s *p)
{
    const unsigned long N = p->N;
    const int code = p->form == MPFFT_FORM_D5 ? 5 : 6;
    const Tree t((uint32_t)N, code == 5 ? 4 : 16);
    std::vector<int32_t> lsn;
    std::vector<uint32_t> loff, ln;
    for (const auto &lf : t.leaves) { lsn.push_back(lf.sn); loff.push_back(lf.off); ln.push_back(lf.n); }
    std::vector<uint32_t> all, begin(t.lgN + 2, 0);
    for (uint32_t l = 0; l <= t.lgN; l++) {
        begin[l] = (uint32_t)all.size();
        all.insert(all.end(), t.level[l].begin(), t.level[l].end());
    }
    begin[t.lgN + 1] = (uint32_t)all.size();
    // twiddles and constants as serial_d_fft5.c:52-62 / serial_d_fft6.c:52-62.
    // The reference's documented build (gcc -O3) fuses its cos(x) and sin(x)
    // of one argument into a single glibc sincos(x) call, whose sine differs
    // from glibc's sin(x) by one ulp for some angles; calling sincos here
    // reproduces that build's table bit for bit whatever this compiler fuses.
    std::vector<double> tw(2 * (N / 4) + 1, 0.0);
    for (unsigned long k = 0; k < N / 4; k++)
        sincos(2. * M_PI * k / N, &tw[N / 4 + k], &tw[k]);
    int rc;
    if ((rc = p->dleaf_sn.upload(lsn)) || (rc = p->dleaf_off.upload(loff)) ||
        (rc = p->dleaf_n.upload(ln)) || (rc = p->dlvl_off.upload(all)) ||
        (rc = p->dlvl_begin.upload(begin)) || (rc = p->tw.upload(tw)))
        return rc;
    DTables &d = p->dt;
    d.leaf_sn = p->dleaf_sn.as<int32_t>(); d.leaf_off = p->dleaf_off.as<uint32_t>();
    d.leaf_n = p->dleaf_n.as<uint32_t>(); d.nleaf = (uint32_t)lsn.size();
    d.lvl_off = p->dlvl_off.as<uint32_t>(); d.lvl_begin = p->dlvl_begin.as<uint32_t>();
    d.wr = p->tw.as<double>(); d.wi = p->tw.as<double>() + N / 4;
    d.c[0] = cos(2. * M_PI * 1 / 8);
    d.c[1] = cos(2. * M_PI * 1 / 16);
    d.c[2] = cos(2. * M_PI * 1 / 16) - cos(2. * M_PI * 3 / 16);
    d.c[3] = cos(2. * M_PI * 3 / 16);
    d.c[4] = cos(2. * M_PI * 1 / 16) + cos(2. * M_PI * 3 / 16);
    d.N = (uint32_t)N; d.lgN = t.lgN; d.lmin = code == 5 ? 3 : 5; d.code = code;
    return MPFFT_OK;
}

// stages [s0, s1): 0 = leaves (and the levels finished in LDS), i >= 1 =
// the combine levels on HBM, smallest node size first
// Stages [s0, s1) of the forward transform; with `inverse` (whole range
// only) the inverse by the reference's swap trick (serial_mp_fft3.c:367-385):
// out = swap(FFT(swap(in))) / N, swap exchanging real and imaginary parts.
int run_mp(mpfft_plan_s *p, const uint32_t *d_in, uint32_t *d_out, hipStream_t st,
           int s0 = 0, int s1 = 1 << 30, bool inverse = false)
{
    const int nst = 1 + (int)p->levels.size();
    if (s1 > nst) s1 = nst;
    if (s0 < 0 || s0 > s1) return MPFFT_ERR_ARG;
    if (inverse && (s0 != 0 || s1 != nst)) return MPFFT_ERR_ARG;
    if (p->N == 1) {                              // both directions: the identity
        if (s0 == 0 && s1 >= 1)
            HIP_TRY(hipMemcpyAsync(d_out, d_in, sizeof(uint32_t) * p->words * p->batch,
                                   hipMemcpyDeviceToDevice, st));
        return MPFFT_OK;
    }
    LeafArgs la = p->leaf;
    la.in = d_in; la.data = d_out;
    la.in_swap = inverse ? 1u : 0u;
    if (s0 > 0) la.npos = 0;                      // no leaf launch
    const int l0 = s0 > 0 ? s0 - 1 : 0, l1 = s1 - 1;
    std::vector<LevelArgs> lv(p->levels.begin() + l0, p->levels.begin() + (l1 > l0 ? l1 : l0));
    for (auto &a : lv) a.data = d_out;
    hipError_t e = mp_fft_launch((int)p->L, (int)p->P, la, lv.data(), (int)lv.size(),
                                 p->batch, st, p->use_sub ? &p->sub : nullptr);
    if (e == hipSuccess && inverse) {
        int32_t lg = 0;
        while ((1UL << lg) < p->N) lg++;
        e = mp_unswap_scale_launch((int)p->L, d_out, p->words, (uint32_t)p->N, lg, p->batch, st);
    }
    return e == hipSuccess ? MPFFT_OK : MPFFT_ERR_HIP;
}

// one whole transform of a limb-layout form
int run_limbs(mpfft_plan_s *p, const uint32_t *d_in, uint32_t *d_out, hipStream_t st)
{
    return has_range_flag(p->form) ? run_fx(p, d_in, d_out, st) : run_mp(p, d_in, d_out, st);
}

// Combine the per-block partials of mp_err_kernel into, per transform,
// {rRMSE re, rRMSE im, max|aE| re, max|aE| im} and their log2.
void combine_err(const double *part, uint32_t nblk, uint32_t batch, double *out)
{
    for (uint32_t b = 0; b < batch; b++) {
        for (int h = 0; h < 2; h++) {
            double S[2] = {0, 0}, E[2] = {0, 0}, lg = -INFINITY;
            for (int q = 0; q < 2; q++) {
                double emax = -INFINITY;
                for (uint32_t k = 0; k < nblk; k++) {
                    const double *o = part + ((size_t)b * nblk + k) * kErrPartial + 5 * h;
                    if (o[2 * q] > 0) emax = fmax(emax, o[2 * q + 1]);
                }
                double sum = 0;
                if (emax > -INFINITY)
                    for (uint32_t k = 0; k < nblk; k++) {
                        const double *o = part + ((size_t)b * nblk + k) * kErrPartial + 5 * h;
                        if (o[2 * q] > 0) sum += ldexp(o[2 * q], (int)(o[2 * q + 1] - emax));
                    }
                S[q] = sum; E[q] = emax > -INFINITY ? emax : 0;
            }
            for (uint32_t k = 0; k < nblk; k++)
                lg = fmax(lg, part[((size_t)b * nblk + k) * kErrPartial + 5 * h + 4]);
            double rr, lrr;
            if (S[0] == 0) { rr = NAN; lrr = NAN; }
            else if (S[1] == 0) { rr = 0; lrr = -INFINITY; }
            else {
                // sqrt(S1 2^E1 / (S0 2^E0)), with the exponent halved exactly
                const long de = (long)(E[1] - E[0]);
                const long half = de >= 0 ? de / 2 : -((-de + 1) / 2);
                const double m = sqrt(S[1] / S[0] * ldexp(1.0, (int)(de - 2 * half)));
                rr = ldexp(m, (int)half);
                lrr = log2(m) + (double)half;
            }
            double *o = out + 8 * (size_t)b;
            o[h] = rr;
            o[2 + h] = lg > -INFINITY ? exp2(lg) : 0.0;
            o[4 + h] = lrr;
            o[6 + h] = lg;
        }
    }
}

int ensure_stage(mpfft_plan_s *p)
{
    const size_t bytes = (is_limb_form(p->form) ? sizeof(uint32_t) : sizeof(double)) *
                         p->words * p->batch;
    if (!p->d_in.holds(bytes) || !p->d_out.holds(bytes)) {
        int rc;
        if ((rc = p->d_in.alloc(bytes)) || (rc = p->d_out.alloc(bytes))) return rc;
    }
    return MPFFT_OK;
}

// cache of plans for the reference entry points, keyed by (N, P, form)
std::mutex g_cache_mu;
std::map<std::tuple<unsigned long, long, int>, std::unique_ptr<mpfft_plan_s>> g_cache;

mpfft_plan_s *cached_plan(unsigned long N, long P, int form, int *rc)
{
    std::lock_guard<std::mutex> g(g_cache_mu);
    auto key = std::make_tuple(N, P, form);
    auto it = g_cache.find(key);
    if (it != g_cache.end()) { *rc = MPFFT_OK; return it->second.get(); }
    mpfft_plan pl = nullptr;
    *rc = mpfft_plan_create(&pl, N, P, form, 1);
    if (*rc) return nullptr;
    g_cache[key].reset(pl);
    return pl;
}

int legacy_mp(unsigned long n, unsigned long s, long sn, mpfr_t *inr, mpfr_t *ini,
              mpfr_t *outr, mpfr_t *outi)
{
    if (!inr || !ini || !outr || !outi) return MPFFT_ERR_ARG;
    if (s != 1 || sn != 0) return MPFFT_ERR_SUBTREE;
    if (!is_pow2(n) || n > kMaxN) return MPFFT_ERR_SIZE;
    const long P = (long)mpfr_get_prec(outr[0]);
    int rc;
    mpfft_plan_s *p = cached_plan(n, P, MPFFT_FORM_MP, &rc);
    if (!p) return rc;
    return mpfft_execute_mpfr(p, inr, ini, outr, outi);
}

int legacy_d(unsigned long n, unsigned long s, long sn, const double *inr, const double *ini,
             double *outr, double *outi)
{
    if (!inr || !ini || !outr || !outi) return MPFFT_ERR_ARG;
    if (s != 1 || sn != 0) return MPFFT_ERR_SUBTREE;
    if (!is_pow2(n) || n > kMaxN) return MPFFT_ERR_SIZE;
    int rc;
    mpfft_plan_s *p = cached_plan(n, 0, MPFFT_FORM_D6, &rc);
    if (!p) return rc;
    std::vector<double> hin(2 * n), hout(2 * n);
    memcpy(hin.data(), inr, sizeof(double) * n);
    memcpy(hin.data() + n, ini, sizeof(double) * n);
    if ((rc = mpfft_execute_f64_host(p, hin.data(), hout.data()))) return rc;
    memcpy(outr, hout.data(), sizeof(double) * n);
    memcpy(outi, hout.data() + n, sizeof(double) * n);
    return MPFFT_OK;
}

// The reference entry points return void and their callers read no status
// (mpFFT.c:84-91), so a failure would leave the outputs unwritten without a
// sign: every failure is recorded (mpfft_last_error) and printed to stderr,
// and a fatal one (a HIP error or no usable GPU) aborts the process.
// MPFFT_NO_ABORT=1 in the environment turns the abort into the record and
// the message only, for callers that check mpfft_last_error themselves.
void legacy_report(const char *name, int rc)
{
    g_last_error = rc;
    if (rc == MPFFT_OK) return;
    fprintf(stderr, "mpfft: %s: %s (%d); outputs not written\n", name, mpfft_strerror(rc), rc);
    if (rc == MPFFT_ERR_HIP || rc == MPFFT_ERR_NODEV) {
        const char *e = getenv("MPFFT_NO_ABORT");
        if (!(e && e[0] == '1')) abort();
    }
}

}  // namespace

// ---- C ABI -----------------------------------------------------------------

extern "C" {

int mpfft_last_error(void) { return g_last_error; }

const char *mpfft_strerror(int code)
{
    switch (code) {
    case MPFFT_OK: return "ok";
    case MPFFT_ERR_SIZE: return "transform size is not a supported power of two";
    case MPFFT_ERR_PREC: return "precision unsupported or inconsistent across arrays";
    case MPFFT_ERR_VALUE: return "input holds NaN, Inf or an exponent beyond +-2^29";
    case MPFFT_ERR_SUBTREE: return "only the top-level call (s = 1, sn = 0) is supported";
    case MPFFT_ERR_HIP: return "HIP runtime error";
    case MPFFT_ERR_ARG: return "bad argument";
    case MPFFT_ERR_NODEV: return "no usable GPU";
    case MPFFT_ERR_FORM: return "form not valid for this call";
    case MPFFT_ERR_SCRATCH: return "beyond the fixed workspace of a no-allocation entry point";
    default: return "unknown error";
    }
}

long mpfft_limbs_for_precision(long prec)
{
    if (prec < 2 || prec > kMaxPrec) return -1;
    return limbs_for_precision(prec);
}

int mpfft_plan_create(mpfft_plan *plan, unsigned long n, long prec, int form, unsigned int batch)
{
    if (!plan || batch == 0) return MPFFT_ERR_ARG;
    *plan = nullptr;
    if (!is_pow2(n) || n > kMaxN) return MPFFT_ERR_SIZE;
    // the chosen form, by the measured rates at matched error: the
    // double-word form where it applies (at 106 bits 4.5x the plain form
    // and 1.7x the fixed form at 128 bits, N = 2^16, batch 64), the fixed
    // form elsewhere (the quad-word form at 212 bits runs at 0.83x the fixed
    // form at 256 bits, with a looser bound)
    if (form == MPFFT_FORM_AUTO)
        form = (prec >= kDdMinPrec && prec <= kDdMaxPrec) ? MPFFT_FORM_DD : MPFFT_FORM_FIXED;
    if (!is_limb_form(form) && form != MPFFT_FORM_D5 && form != MPFFT_FORM_D6)
        return MPFFT_ERR_FORM;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MPFFT_ERR_NODEV;
    std::unique_ptr<mpfft_plan_s> p(new mpfft_plan_s);
    p->N = n; p->form = form; p->batch = batch;
    HIP_TRY(hipGetDevice(&p->device));
    HIP_TRY(hipStreamCreate(&p->stream));
    int rc;
    if (is_limb_form(form)) {
        if (prec < 2 || prec > kMaxPrec) return MPFFT_ERR_PREC;
        if (form == MPFFT_FORM_DD && (prec < kDdMinPrec || prec > kDdMaxPrec))
            return MPFFT_ERR_PREC;
        if (form == MPFFT_FORM_QD && (prec < kDdMinPrec || prec > kQdMaxPrec))
            return MPFFT_ERR_PREC;
        p->P = prec;
        p->L = limbs_for_precision(prec);
        p->words = 2 * n * (size_t)(p->L + 1);
        // beyond 32 limbs: the plain op orders and the fixed form (panels
        // and per-lane kernels at 64 limbs)
        if (p->L > kMaxLimbsAll && !is_mp(form) && !is_fixed(form))
            return MPFFT_ERR_PREC;
        rc = word_count(form)       ? build_dd_plan(p.get())
             : is_fixed(form)      ? build_fx_plan(p.get())
                                   : build_mp_plan(p.get());
    } else {
        p->P = 53;
        p->words = 2 * n;
        rc = build_d_plan(p.get());
    }
    if (rc) return rc;
    *plan = p.release();
    return MPFFT_OK;
}

int mpfft_plan_destroy(mpfft_plan plan)
{
    delete plan;
    return MPFFT_OK;
}

int mpfft_plan_info(mpfft_plan p, long *limbs, size_t *words)
{
    if (!p) return MPFFT_ERR_ARG;
    if (limbs) *limbs = p->L;
    if (words) *words = p->words;
    return MPFFT_OK;
}

int mpfft_plan_sync(mpfft_plan p)
{
    if (!p) return MPFFT_ERR_ARG;
    HIP_TRY(hipSetDevice(p->device));
    HIP_TRY(hipStreamSynchronize(p->stream));
    return MPFFT_OK;
}

int mpfft_execute_limbs(mpfft_plan p, const uint32_t *d_in, uint32_t *d_out, void *stream)
{
    if (!p || !d_in || !d_out) return MPFFT_ERR_ARG;
    if (!is_limb_form(p->form)) return MPFFT_ERR_FORM;
    HIP_TRY(hipSetDevice(p->device));
    return run_limbs(p, d_in, d_out, stream ? (hipStream_t)stream : p->stream);
}

int mpfft_plan_check(mpfft_plan p)
{
    if (!p) return MPFFT_ERR_ARG;
    HIP_TRY(hipSetDevice(p->device));
    HIP_TRY(hipStreamSynchronize(p->stream));
    return fx_check(p);
}

int mpfft_execute_inverse_limbs(mpfft_plan p, const uint32_t *d_in, uint32_t *d_out,
                                void *stream)
{
    if (!p || !d_in || !d_out) return MPFFT_ERR_ARG;
    if (!is_mp(p->form)) return MPFFT_ERR_FORM;
    HIP_TRY(hipSetDevice(p->device));
    return run_mp(p, d_in, d_out, stream ? (hipStream_t)stream : p->stream, 0, 1 << 30, true);
}

int mpfft_roundtrip_stats(mpfft_plan p, const uint32_t *d_x, const uint32_t *d_y,
                          void *stream, double *out)
{
    if (!p || !d_x || !d_y || !out) return MPFFT_ERR_ARG;
    if (!is_mp(p->form)) return MPFFT_ERR_FORM;
    std::lock_guard<std::mutex> g(p->mu);
    HIP_TRY(hipSetDevice(p->device));
    hipStream_t st = stream ? (hipStream_t)stream : p->stream;
    const uint32_t nblk = (uint32_t)((2 * p->N + kErrBlock - 1) / kErrBlock);
    const size_t nd = (size_t)nblk * p->batch * kErrPartial;
    if (!p->errbuf.holds(nd * sizeof(double))) {
        int rc = p->errbuf.alloc(nd * sizeof(double));
        if (rc) return rc;
    }
    HIP_TRY(mp_err_launch((int)p->L, (int)p->P, d_x, d_y, p->words, (uint32_t)p->N,
                          p->errbuf.as<double>(), p->batch, st));
    std::vector<double> part(nd);
    HIP_TRY(hipMemcpyAsync(part.data(), p->errbuf.p, nd * sizeof(double),
                           hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    combine_err(part.data(), nblk, p->batch, out);
    return MPFFT_OK;
}

int mpfft_roundtrip_stats_noalloc(mpfft_plan p, const uint32_t *d_x, const uint32_t *d_y,
                                  void *stream, double *out)
{
    if (!p || !d_x || !d_y || !out) return MPFFT_ERR_ARG;
    if (!is_mp(p->form)) return MPFFT_ERR_FORM;
    const uint32_t nblk = (uint32_t)((2 * p->N + kErrBlock - 1) / kErrBlock);
    const size_t nd = (size_t)nblk * p->batch * kErrPartial;
    if (nd > kScratchDoubles) return MPFFT_ERR_SCRATCH;
    std::lock_guard<std::mutex> g(p->mu);
    HIP_TRY(hipSetDevice(p->device));
    std::lock_guard<std::mutex> gs(g_scratch_mu);
    double *ws = nullptr;
    int rc = scratch_on(p->device, &ws);
    if (rc) return rc;
    hipStream_t st = stream ? (hipStream_t)stream : p->stream;
    // every block writes all kErrPartial of its doubles: nothing to clear
    HIP_TRY(mp_err_launch((int)p->L, (int)p->P, d_x, d_y, p->words, (uint32_t)p->N, ws,
                          p->batch, st));
    HIP_TRY(hipMemcpyAsync(g_scratch_host, ws, nd * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    combine_err(g_scratch_host, nblk, p->batch, out);
    return MPFFT_OK;
}

int mpfft_roundtrip_host(mpfft_plan p, const uint32_t *h_x, double *out)
{
    if (!p || !h_x || !out) return MPFFT_ERR_ARG;
    if (!is_mp(p->form)) return MPFFT_ERR_FORM;
    int rc;
    {
        std::lock_guard<std::mutex> g(p->mu);
        HIP_TRY(hipSetDevice(p->device));
        if ((rc = ensure_stage(p))) return rc;
        if (!p->d_aux.holds(p->d_in.bytes) && (rc = p->d_aux.alloc(p->d_in.bytes))) return rc;
        HIP_TRY(hipMemcpyAsync(p->d_in.p, h_x, p->d_in.bytes, hipMemcpyHostToDevice, p->stream));
        if ((rc = run_mp(p, p->d_in.as<uint32_t>(), p->d_out.as<uint32_t>(), p->stream)))
            return rc;
        if ((rc = run_mp(p, p->d_out.as<uint32_t>(), p->d_aux.as<uint32_t>(), p->stream,
                         0, 1 << 30, true)))
            return rc;
    }
    return mpfft_roundtrip_stats(p, p->d_in.as<uint32_t>(), p->d_aux.as<uint32_t>(),
                                 p->stream, out);
}

int mpfft_execute_limbs_range(mpfft_plan p, const uint32_t *d_in, uint32_t *d_out,
                              void *stream, int stage_begin, int stage_end)
{
    if (!p || !d_in || !d_out) return MPFFT_ERR_ARG;
    if (!is_limb_form(p->form)) return MPFFT_ERR_FORM;
    HIP_TRY(hipSetDevice(p->device));
    hipStream_t st = stream ? (hipStream_t)stream : p->stream;
    if (has_range_flag(p->form)) return run_fx(p, d_in, d_out, st, stage_begin, stage_end);
    return run_mp(p, d_in, d_out, st, stage_begin, stage_end);
}

int mpfft_plan_form(mpfft_plan p)
{
    if (!p) return MPFFT_ERR_ARG;
    if (p->form == MPFFT_FORM_FIXED)
        return p->fx_panel ? MPFFT_FORM_FIXED_PANEL : MPFFT_FORM_FIXED_LANE;
    return p->form;
}

int mpfft_plan_panel_plan(mpfft_plan p)
{
    if (!p) return MPFFT_ERR_ARG;
    if (!is_fixed(p->form) || !p->fx_panel) return 0;
    return p->fx_fs ? 2 : 1;
}

int mpfft_plan_stages(mpfft_plan p)
{
    if (!p) return MPFFT_ERR_ARG;
    if (word_count(p->form)) return p->N == 1 ? 1 : p->fx_npass;
    if (is_fixed(p->form))
        return p->fx_fs ? p->fx_npass : p->fx_panel ? p->fx_npass + 2
                                      : (p->fx_npass > 0 ? p->fx_npass : 2);
    return is_mp(p->form) ? 1 + (int)p->levels.size() : MPFFT_ERR_FORM;
}

int mpfft_execute_limbs_host(mpfft_plan p, const uint32_t *h_in, uint32_t *h_out)
{
    if (!p || !h_in || !h_out) return MPFFT_ERR_ARG;
    if (!is_limb_form(p->form)) return MPFFT_ERR_FORM;
    std::lock_guard<std::mutex> g(p->mu);
    HIP_TRY(hipSetDevice(p->device));
    int rc = ensure_stage(p);
    if (rc) return rc;
    const size_t bytes = sizeof(uint32_t) * p->words * p->batch;
    HIP_TRY(hipMemcpyAsync(p->d_in.p, h_in, bytes, hipMemcpyHostToDevice, p->stream));
    if ((rc = run_limbs(p, p->d_in.as<uint32_t>(), p->d_out.as<uint32_t>(), p->stream)))
        return rc;
    HIP_TRY(hipMemcpyAsync(h_out, p->d_out.p, bytes, hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    return fx_check(p);
}

int mpfft_mpfr_to_limbs(unsigned long n, long limbs, mpfr_t *re, mpfr_t *im, uint32_t *h_out)
{
    if (!re || !im || !h_out || limbs < 1) return MPFFT_ERR_ARG;
    const size_t M = 2 * n;
    uint32_t *hdr = h_out, *limb = h_out + M;
    const int rc = mpfr_put_range(re, 0, n, 0, hdr, limb, M, limbs);
    return rc ? rc : mpfr_put_range(im, 0, n, n, hdr, limb, M, limbs);
}

int mpfft_limbs_to_mpfr(unsigned long n, long limbs, const uint32_t *h_in, mpfr_t *re, mpfr_t *im)
{
    if (!re || !im || !h_in || limbs < 1) return MPFFT_ERR_ARG;
    const size_t M = 2 * n;
    const uint32_t *hdr = h_in, *limb = h_in + M;
    for (unsigned long k = 0; k < n; k++) {
        mpfr_take(re[k], hdr, limb, M, k, limbs);
        mpfr_take(im[k], hdr, limb, M, n + k, limbs);
    }
    return MPFFT_OK;
}

int mpfft_execute_mpfr(mpfft_plan p, mpfr_t *inr, mpfr_t *ini, mpfr_t *outr, mpfr_t *outi)
{
    if (!p || !inr || !ini || !outr || !outi) return MPFFT_ERR_ARG;
    if (!is_limb_form(p->form)) return MPFFT_ERR_FORM;
    if (p->batch != 1) return MPFFT_ERR_ARG;
    const unsigned long n = p->N;
    for (unsigned long k = 0; k < n; k++) {
        if (mpfr_get_prec(inr[k]) != p->P || mpfr_get_prec(ini[k]) != p->P ||
            mpfr_get_prec(outr[k]) != p->P || mpfr_get_prec(outi[k]) != p->P)
            return MPFFT_ERR_PREC;
    }
    std::lock_guard<std::mutex> g(p->mu);
    HIP_TRY(hipSetDevice(p->device));
    int rc;
    if ((rc = ensure_stage(p))) return rc;
    const size_t bytes = sizeof(uint32_t) * p->words;
    if ((rc = p->h_stage_in.alloc(bytes)) || (rc = p->h_stage_out.alloc(bytes))) return rc;
    if (!p->d_aux.holds(bytes) && (rc = p->d_aux.alloc(bytes))) return rc;

    // The input is converted in chunks of cl points (8 chunks from 8192
    // points on, else one): the pool's threads convert every chunk in turn,
    // each its share of the chunk's points, into page-locked chunk-major
    // staging (each chunk L + 1 rows of 2 cl words, real parts then
    // imaginary parts), while this thread uploads every chunk as soon as all
    // its shares are in; on the device the chunks are laid out as SoA rows
    // and the transform runs; the result comes back in one copy and the pool
    // converts it.  (Chunked downloads overlapped with the conversion out
    // measured slower: waiting threads, polling or blocked on events, slowed
    // the runtime's own copy handling on the host.)
    const long L = p->L;
    const uint32_t R = (uint32_t)L + 1;
    const unsigned long cl = n >= 8192 ? n / 8 : n;
    const unsigned C = (unsigned)(n / cl);
    const size_t cw = (size_t)R * 2 * cl;                        // words per chunk
    const size_t M = 2 * n;
    uint32_t *hi = p->h_stage_in.as<uint32_t>(), *ho = p->h_stage_out.as<uint32_t>();
    uint32_t *di = p->d_in.as<uint32_t>(), *dout = p->d_out.as<uint32_t>();
    uint32_t *ds = p->d_aux.as<uint32_t>();
    for (auto &e : p->time_ev)
        if (!e) HIP_TRY(hipEventCreate(&e));
    HostPool &pool = HostPool::get();
    std::lock_guard<std::mutex> pg(pool.user());
    const unsigned nthr =
        (unsigned)std::max(1UL, std::min((unsigned long)pool.size(), (cl + 255) / 256));
    auto t0 = std::chrono::steady_clock::now();

    // in: conversions on the pool, uploads here in chunk order
    std::vector<int> parts(C, 0);                                // shares converted
    std::mutex pm;
    std::condition_variable pcv;
    std::atomic<int> err{0};
    pool.start(nthr, [&](unsigned w) {
        // share w of a chunk: its points [w cl / nthr, (w + 1) cl / nthr)
        const unsigned long s0 = (unsigned long)w * cl / nthr;
        const unsigned long s1 = (unsigned long)(w + 1) * cl / nthr;
        for (unsigned c = 0; c < C; c++) {
            const unsigned long k0 = c * cl;
            uint32_t *blk = hi + c * cw;
            // position of point k in the chunk: k - k0 (real), cl + k - k0
            // (imaginary), as offsets modulo 2^64
            int r = err.load() ? MPFFT_OK
                               : mpfr_put_range(inr, k0 + s0, k0 + s1, (size_t)0 - k0, blk,
                                                blk + 2 * cl, 2 * cl, L);
            if (!r && !err.load())
                r = mpfr_put_range(ini, k0 + s0, k0 + s1, (size_t)cl - k0, blk, blk + 2 * cl,
                                   2 * cl, L);
            if (r) {
                int z = 0;
                err.compare_exchange_strong(z, r);
            }
            std::lock_guard<std::mutex> lk(pm);
            if (++parts[c] == (int)nthr) pcv.notify_one();
        }
    });
    hipError_t he = hipEventRecord(p->time_ev[0], p->stream);
    for (unsigned c = 0; c < C && he == hipSuccess; c++) {
        {
            std::unique_lock<std::mutex> lk(pm);
            pcv.wait(lk, [&] { return parts[c] == (int)nthr; });
        }
        if (err.load()) break;
        he = hipMemcpyAsync(ds + c * cw, hi + c * cw, sizeof(uint32_t) * cw, hipMemcpyHostToDevice,
                            p->stream);
    }
    pool.wait();
    rc = err.load();
    if (!rc && he != hipSuccess) rc = MPFFT_ERR_HIP;
    if (rc) {
        (void)hipStreamSynchronize(p->stream);
        return rc;
    }
    // transform, download
    if (hipEventRecord(p->time_ev[1], p->stream) != hipSuccess ||
        mp_chunk_relayout_launch(ds, di, (uint32_t)n, R, (uint32_t)cl, true, p->stream) != hipSuccess)
        rc = MPFFT_ERR_HIP;
    if (!rc) rc = run_limbs(p, di, dout, p->stream);
    if (!rc && (hipEventRecord(p->time_ev[2], p->stream) != hipSuccess ||
                hipMemcpyAsync(ho, dout, sizeof(uint32_t) * M * R, hipMemcpyDeviceToHost,
                               p->stream) != hipSuccess))
        rc = MPFFT_ERR_HIP;
    if (hipStreamSynchronize(p->stream) != hipSuccess && !rc) rc = MPFFT_ERR_HIP;
    if (rc) return rc;
    // out: the pool converts its shares of the points
    pool.start(nthr, [&](unsigned w) {
        const unsigned long k0 = (unsigned long)w * n / nthr, k1 = (unsigned long)(w + 1) * n / nthr;
        mpfr_take_range(outr, k0, k1, 0, ho, ho + M, M, L);
        mpfr_take_range(outi, k0, k1, n, ho, ho + M, M, L);
    });
    pool.wait();
    auto t2 = std::chrono::steady_clock::now();
    // in: from the call's start to the last upload's end (the upload-start
    // event completes at once on the idle stream); compute: the device's
    // layout change and the transform; out: the rest
    float tin = 0, tgpu = 0;
    (void)hipEventElapsedTime(&tin, p->time_ev[0], p->time_ev[1]);
    (void)hipEventElapsedTime(&tgpu, p->time_ev[1], p->time_ev[2]);
    const double ttot = std::chrono::duration<double, std::milli>(t2 - t0).count();
    p->timing[0] = tin;
    p->timing[1] = tgpu;
    p->timing[2] = ttot - tin - tgpu;
    p->timing[3] = ttot;
    return fx_check(p);
}

int mpfft_plan_last_timing(mpfft_plan p, double *ms)
{
    if (!p || !ms) return MPFFT_ERR_ARG;
    for (int i = 0; i < 4; i++) ms[i] = p->timing[i];
    return MPFFT_OK;
}

int mpfft_execute_f64(mpfft_plan p, const double *d_in, double *d_out, void *stream)
{
    if (!p || !d_in || !d_out) return MPFFT_ERR_ARG;
    if (is_limb_form(p->form)) return MPFFT_ERR_FORM;
    HIP_TRY(hipSetDevice(p->device));
    hipError_t e = d_fft_launch(p->dt, d_in, d_out, p->words, p->batch,
                                stream ? (hipStream_t)stream : p->stream);
    return e == hipSuccess ? MPFFT_OK : MPFFT_ERR_HIP;
}

int mpfft_execute_f64_host(mpfft_plan p, const double *h_in, double *h_out)
{
    if (!p || !h_in || !h_out) return MPFFT_ERR_ARG;
    if (is_limb_form(p->form)) return MPFFT_ERR_FORM;
    std::lock_guard<std::mutex> g(p->mu);
    HIP_TRY(hipSetDevice(p->device));
    int rc = ensure_stage(p);
    if (rc) return rc;
    const size_t bytes = sizeof(double) * p->words * p->batch;
    HIP_TRY(hipMemcpyAsync(p->d_in.p, h_in, bytes, hipMemcpyHostToDevice, p->stream));
    if ((rc = mpfft_execute_f64(p, p->d_in.as<double>(), p->d_out.as<double>(), p->stream)))
        return rc;
    HIP_TRY(hipMemcpyAsync(h_out, p->d_out.p, bytes, hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    return MPFFT_OK;
}

int mpfft_mp_binop_host(int op, unsigned long count, long prec, const uint32_t *x,
                        const uint32_t *y, uint32_t *z)
{
    if (!x || !y || !z || op < 0 || op > 2 || count == 0 || count > (1UL << 26))
        return MPFFT_ERR_ARG;
    const long L = mpfft_limbs_for_precision(prec);
    if (L < 1 || L > kMaxLimbsMp) return MPFFT_ERR_PREC;
    const size_t bytes = sizeof(uint32_t) * count * (L + 1);
    DevBuf dx, dy, dz;
    int rc;
    if ((rc = dx.alloc(bytes)) || (rc = dy.alloc(bytes)) || (rc = dz.alloc(bytes))) return rc;
    HIP_TRY(hipMemcpy(dx.p, x, bytes, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dy.p, y, bytes, hipMemcpyHostToDevice));
    if (mp_binop_launch((int)L, (int)prec, op, (uint32_t)count, dx.as<uint32_t>(),
                        dy.as<uint32_t>(), dz.as<uint32_t>(), nullptr) != hipSuccess)
        return MPFFT_ERR_HIP;
    HIP_TRY(hipMemcpy(z, dz.p, bytes, hipMemcpyDeviceToHost));
    return MPFFT_OK;
}

void mp_sr24_cp_fft(unsigned long n, unsigned long s, long int sn, mpfr_t *inr, mpfr_t *ini,
                    mpfr_t *outr, mpfr_t *outi)
{
    legacy_report("mp_sr24_cp_fft", legacy_mp(n, s, sn, inr, ini, outr, outi));
}

void mp_sr24_cp_fft_v2(unsigned long n, unsigned long log2n, unsigned long s, long int sn,
                       mpfr_t *inr, mpfr_t *ini, mpfr_t *outr, mpfr_t *outi)
{
    legacy_report("mp_sr24_cp_fft_v2", log2n != 0 ? MPFFT_ERR_SUBTREE
                                                  : legacy_mp(n, s, sn, inr, ini, outr, outi));
}

void d_sr24_cp_fft(unsigned long n, unsigned long s, long int sn, double *inr, double *ini,
                   double *outr, double *outi)
{
    legacy_report("d_sr24_cp_fft", legacy_d(n, s, sn, inr, ini, outr, outi));
}

void d_sr24_cp_fft_v2(unsigned long n, unsigned long log2n, unsigned long s, long int sn,
                      double *inr, double *ini, double *outr, double *outi)
{
    legacy_report("d_sr24_cp_fft_v2", log2n != 0 ? MPFFT_ERR_SUBTREE
                                                 : legacy_d(n, s, sn, inr, ini, outr, outi));
}

}  // extern "C"
