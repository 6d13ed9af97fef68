// Host-logic check of libmpfft without a GPU: the host library (plans,
// twiddle tables, batched form) built against a stand-in HIP runtime whose
// kernel launchers touch every buffer range a launch would read or write
// (stub.cpp), run under AddressSanitizer; plus the pass-plan invariants the
// kernels rely on, for every size and compiled width.
#include "../../include/mpfft.h"
#include "../../mpfft_amd/csrc/kernels.h"

#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include <sys/wait.h>
#include <unistd.h>

void stub_fail_malloc_in(int k);      // stub.cpp
#include <atomic>
namespace mpfft { extern std::atomic<int> stub_sub_checked, stub_sub_errors, stub_identity_mp; }   // stub.cpp

static int fails = 0;
#define CHECK(c, ...)                                   \
    do {                                                \
        if (!(c)) {                                     \
            fails++;                                    \
            printf("FAIL %s:%d ", __FILE__, __LINE__);  \
            printf(__VA_ARGS__);                        \
            printf("\n");                               \
        }                                               \
    } while (0)

static void plans()
{
    using namespace mpfft;
    const int widths[] = {3, 4, 5, 6, 8, 10, 14, 18, 26, 34, 66};
    for (int W : widths) {
        // the twiddle tables fit at the slot end, or get their own area
        const MfGeom G = mf_geom(W);
        CHECK(mf_tail_fits(W) || mf_table_words(W) == 3 * G.RW, "W=%d table area", W);
        CHECK(G.KT * 32 >= G.NXB && G.MT * 32 >= G.NCOL, "W=%d tiles", W);
        for (int rb = 2; rb <= 4; rb++)
        for (int lg = 4; lg <= 24; lg++) {
            const uint32_t N = 1u << lg;
            FxpPass p[32];
            uint32_t nent = 0, prod = 1, ent = 0;
            const int np = fxp_plan(N, W, rb, p, &nent);
            const int mb = fxp_radix_ok(W, rb) ? rb : 2;
            for (int i = 0; i < np; i++) {
                CHECK(p[i].Ns == prod && p[i].toff == ent, "fxp W=%d lg=%d pass %d", W, lg, i);
                CHECK(p[i].R != 8 || W <= kPanelR8MaxW, "fxp radix 8 at W=%d", W);
                CHECK(p[i].R != 16 || W <= kPanelR16MaxW, "fxp radix 16 at W=%d", W);
                // a radix-16 pass needs N >= 16 (one item covers 16 positions)
                CHECK(p[i].R != 16 || N >= 16, "fxp radix 16 at N=%u", N);
                // every pass but a first one (or a first two: radix 2 then 4)
                // has the main radix
                CHECK(i == 0 || p[i].R == (1u << mb) || (i == 1 && p[0].R == 2 && p[i].R == 4),
                      "fxp W=%d rb=%d lg=%d pass %d radix %u", W, rb, lg, i, p[i].R);
                ent += (p[i].R - 1) * p[i].Ns;
                prod *= p[i].R;
            }
            CHECK(prod == N && ent == nent, "fxp W=%d lg=%d product", W, lg);
        }
    }
    for (int K : {2, 4})
        for (int lg = 1; lg <= 24; lg++) {
            const uint32_t N = 1u << lg;
            MwPass p[32];
            uint32_t nent = 0, prod = 1, ent = 0;
            const int np = mw_plan(N, K, p, &nent);
            for (int i = 0; i < np; i++) {
                CHECK(p[i].Ns == prod && p[i].toff == ent, "mw K=%d lg=%d pass %d", K, lg, i);
                CHECK(p[i].R != 16 || N >= 16, "mw K=%d lg=%d r16", K, lg);
                // the radix-2 kernel has no twiddles: first pass only
                CHECK(p[i].R != 2 || (i == 0 && p[i].Ns == 1), "mw K=%d lg=%d r2 at %d", K, lg, i);
                // radix 4 only right after the first radix 16 (from N = 16 on)
                CHECK(lg < 4 || p[i].R != 4 ||
                          (i > 0 && p[i - 1].R == 16 && p[i - 1].Ns == (p[0].R == 2 ? 2u : 1u)),
                      "mw K=%d lg=%d r4 at %d", K, lg, i);
                ent += (p[i].R - 1) * p[i].Ns;
                prod *= p[i].R;
            }
            CHECK(prod == N && ent == nent, "mw K=%d lg=%d product", K, lg);
            // one pass per radix-16 factor, plus one for each of bits 0 and 1
            // of log2 N
            CHECK(lg < 4 || np == lg / 4 + (lg & 1) + ((lg >> 1) & 1),
                  "mw K=%d lg=%d passes %d", K, lg, np);
        }
    // the fixed form on lanes: radix 16 up to kFxl16MaxW words (the word
    // forms' plan), radix 2 (first only) and 4 beyond
    for (int W : {3, 6, 10, 12, 34, 66})
        for (int lg = 1; lg <= 24; lg++) {
            const uint32_t N = 1u << lg;
            MwPass p[32];
            uint32_t nent = 0, prod = 1, ent = 0;
            const int np = fxl_plan(N, W, p, &nent);
            for (int i = 0; i < np; i++) {
                CHECK(p[i].Ns == prod && p[i].toff == ent, "fxl W=%d lg=%d pass %d", W, lg, i);
                CHECK(p[i].R != 16 || (W <= kFxl16MaxW && N >= 16), "fxl W=%d lg=%d r16", W, lg);
                CHECK(p[i].R != 2 || (i == 0 && p[i].Ns == 1), "fxl W=%d lg=%d r2 at %d", W, lg, i);
                ent += (p[i].R - 1) * p[i].Ns;
                prod *= p[i].R;
            }
            CHECK(prod == N && ent == nent, "fxl W=%d lg=%d product", W, lg);
            CHECK(W > kFxl16MaxW || lg < 4 || np == lg / 4 + (lg & 1) + ((lg >> 1) & 1),
                  "fxl W=%d lg=%d passes %d", W, lg, np);
        }
    // panels: a partial last panel is padded; the preferred rule
    CHECK(fxp_groups(1) == 1 && fxp_groups(64) == 1 && fxp_groups(65) == 2, "groups");
    CHECK(fxp_preferred(1024, 64, 10) && fxp_preferred(1024, 120, 10) &&
              !fxp_preferred(1 << 16, 80, 10) && fxp_preferred(256, 3, 34) &&
              !fxp_preferred(8, 64, 10) && !fxp_preferred(4096, 64, 10) &&
              !fxp_preferred(1024, 64, 4) && fxp_preferred(4096, 64, 12) &&
              fxl_preferred(1 << 16, 10) && !fxl_preferred(1 << 16, 12),
          "preferred rule");
}

static void batch(int form, std::vector<unsigned long> n, std::vector<long> P, int expect_check)
{
    mpfft_batch b;
    int rc = mpfft_batch_create(&b, (unsigned)n.size(), n.data(), P.data(), form);
    CHECK(rc == 0, "create form %d rc %d", form, rc);
    if (rc) return;
    size_t words = 0;
    mpfft_batch_words(b, &words);
    // every transform owns a disjoint, tight range of the packed buffers
    std::vector<char> used(words, 0);
    for (unsigned i = 0; i < n.size(); i++) {
        size_t off;
        long L;
        int f;
        mpfft_batch_info(b, i, &off, &L, &f);
        CHECK(L == mpfft_limbs_for_precision(P[i]), "limbs");
        const size_t w = 2 * n[i] * (L + 1);
        for (size_t k = off; k < off + w; k++) {
            CHECK(k < words && !used[k], "overlap at transform %u", i);
            if (k < words) used[k] = 1;
        }
    }
    for (size_t k = 0; k < words; k++) CHECK(used[k], "gap at word %zu", k);
    std::vector<uint32_t> in(words, 0), out(words, 0), in2(words, 0), out2(words, 0);
    CHECK(mpfft_batch_tune(b, in.data(), out.data()) == 0, "tune");
    CHECK(mpfft_batch_execute_limbs(b, in.data(), out.data(), nullptr) == 0, "execute");
    CHECK(mpfft_batch_execute_limbs(b, in2.data(), out2.data(), nullptr) == 0, "execute again");
    // the stand-in launchers set the range flag of the fixed and word forms
    CHECK(mpfft_batch_check(b) == expect_check, "check");
    int S = 0;
    double ms = 0;
    mpfft_batch_schedule(b, &S, &ms);
    CHECK(S == 1 || S == 2 || S == 4, "streams %d", S);
    mpfft_batch_destroy(b);
}

// A staging allocation that fails (d_in allocated, d_out not) and a retry:
// the retry allocates again instead of taking the failed buffer for a
// usable one (the stand-in kernels would write through a null pointer).
static void alloc_failure_then_retry()
{
    mpfft_plan p;
    CHECK(mpfft_plan_create(&p, 64, 256, MPFFT_FORM_MP, 2) == 0, "plan");
    long L;
    size_t w;
    mpfft_plan_info(p, &L, &w);
    std::vector<uint32_t> x(2 * w, 0), y(2 * w, 0);
    stub_fail_malloc_in(1);                       // d_in succeeds, d_out fails
    CHECK(mpfft_execute_limbs_host(p, x.data(), y.data()) == MPFFT_ERR_HIP, "failed alloc");
    CHECK(mpfft_execute_limbs_host(p, x.data(), y.data()) == MPFFT_OK, "retry");
    stub_fail_malloc_in(0);                       // a plan whose first buffer fails
    mpfft_plan q = nullptr;
    CHECK(mpfft_plan_create(&q, 64, 256, MPFFT_FORM_MP, 3) == MPFFT_ERR_HIP, "plan alloc");
    stub_fail_malloc_in(-1);
    CHECK(mpfft_execute_limbs_host(p, x.data(), y.data()) == MPFFT_OK, "after");
    mpfft_plan_destroy(p);
}

// The no-allocation statistics: no allocation on any call (the stand-in's
// next hipMalloc would fail), the same results as the allocating entry
// point, and the size limit.
static void noalloc_stats()
{
    mpfft_plan p;
    CHECK(mpfft_plan_create(&p, 4096, 256, MPFFT_FORM_MP, 3) == 0, "plan");
    long L;
    size_t w;
    mpfft_plan_info(p, &L, &w);
    std::vector<uint32_t> x(3 * w, 0);
    std::vector<double> a(24), b(24);
    stub_fail_malloc_in(0);
    CHECK(mpfft_roundtrip_stats_noalloc(p, x.data(), x.data(), nullptr, a.data()) == 0, "noalloc");
    CHECK(mpfft_roundtrip_stats_noalloc(p, x.data(), x.data(), nullptr, a.data()) == 0, "again");
    stub_fail_malloc_in(-1);
    CHECK(mpfft_roundtrip_stats(p, x.data(), x.data(), nullptr, b.data()) == 0, "alloc");
    for (int i = 0; i < 24; i++) CHECK(a[i] == b[i] || (a[i] != a[i] && b[i] != b[i]), "same %d", i);
    mpfft_plan_destroy(p);
    mpfft_plan q;
    CHECK(mpfft_plan_create(&q, 1UL << 16, 64, MPFFT_FORM_MP, 103) == 0, "plan big");
    mpfft_plan_info(q, &L, &w);
    CHECK(mpfft_roundtrip_stats_noalloc(q, x.data(), x.data(), nullptr, a.data()) == MPFFT_ERR_SCRATCH,
          "limit");
    mpfft_plan_destroy(q);
}

// The reference entry points on failure (argv[1] selects the case; the test
// runs each in its own process): a precision mismatch is recorded and printed,
// a HIP error aborts unless MPFFT_NO_ABORT=1.
static int legacy(const char *what)
{
    const unsigned long n = 16;
    mpfr_t a[4][16];
    for (int k = 0; k < 4; k++)
        for (unsigned long i = 0; i < n; i++) {
            mpfr_init2(a[k][i], (k == 3 && what[0] == 'p') ? 128 : 256);
            mpfr_set_ui(a[k][i], (unsigned long)i, MPFR_RNDN);
        }
    mpfr_t *r = a[0], *im = a[1], *o = a[2], *oi = a[3];
    if (what[0] == 'h') stub_fail_malloc_in(0);   // the first device allocation fails
    mp_sr24_cp_fft(n, 1, 0, r, im, o, oi);
    printf("last_error %d\n", mpfft_last_error());
    return 0;
}

// The batched form's host logic from several host threads at once: plans
// built on threads inside each mpfft_batch_create, tuning (timing, dealing,
// capture), replays and re-capture on other buffers, and destruction, for
// two batches used concurrently (run under ThreadSanitizer).
static void batches_concurrently()
{
    auto work = [](int seed) {
        srand(seed);
        std::vector<unsigned long> n;
        std::vector<long> P;
        const long ps[] = {128, 256, 512, 1024};
        for (int i = 0; i < 40; i++) {
            n.push_back(1UL << (4 + rand() % 6));
            P.push_back(ps[rand() % 4]);
        }
        for (int form : {MPFFT_FORM_FIXED, MPFFT_FORM_MP}) {
            mpfft_batch b;
            CHECK(mpfft_batch_create(&b, (unsigned)n.size(), n.data(), P.data(), form) == 0, "create");
            size_t words = 0;
            mpfft_batch_words(b, &words);
            std::vector<uint32_t> x(words, 0), y(words, 0), y2(words, 0);
            CHECK(mpfft_batch_tune(b, x.data(), y.data()) == 0, "tune");
            for (int r = 0; r < 3; r++)
                CHECK(mpfft_batch_execute_limbs(b, x.data(), y.data(), nullptr) == 0, "exec");
            CHECK(mpfft_batch_execute_limbs(b, x.data(), y2.data(), nullptr) == 0, "recapture");
            CHECK(mpfft_batch_tune(b, x.data(), y.data()) == 0, "retune");
            mpfft_batch_check(b);
            mpfft_batch_destroy(b);
        }
    };
    std::thread a(work, 3), c(work, 4);
    a.join();
    c.join();
}

// The drop-in mpfr_t call's pipeline (chunked conversion on the host pool,
// uploads, the device's chunk <-> SoA layout changes, chunked downloads and
// conversions out) with the plain transform as the identity: the outputs
// must equal the inputs bit for bit, at one chunk (n = 16) and at 8 and 64
// chunks, precisions of one and several 64-bit limbs, signed zeros.
static void mpfr_pipeline_one(unsigned long n, long P, int seed)
{
    mpfft_plan p;
    CHECK(mpfft_plan_create(&p, n, P, MPFFT_FORM_MP, 1) == 0, "mpfr plan n=%lu P=%ld", n, P);
    std::unique_ptr<mpfr_t[]> a(new mpfr_t[4 * n]);
    gmp_randstate_t st;
    gmp_randinit_default(st);
    gmp_randseed_ui(st, (unsigned long)seed);
    for (unsigned long i = 0; i < 4 * n; i++) {
        mpfr_init2(a[i], P);
        mpfr_urandomb(a[i], st);
        if (i % 7 == 3) mpfr_neg(a[i], a[i], MPFR_RNDN);
        if (i % 97 == 5) mpfr_set_zero(a[i], (i & 1) ? -1 : 1);
        mpfr_mul_2si(a[i], a[i], (long)(i % 13) - 6, MPFR_RNDN);
    }
    mpfr_t *inr = a.get(), *ini = a.get() + n, *outr = a.get() + 2 * n, *outi = a.get() + 3 * n;
    CHECK(mpfft_execute_mpfr(p, inr, ini, outr, outi) == 0, "mpfr exec n=%lu", n);
    int bad = 0;
    for (unsigned long k = 0; k < n; k++) {
        bad += mpfr_cmp(outr[k], inr[k]) != 0 || mpfr_signbit(outr[k]) != mpfr_signbit(inr[k]);
        bad += mpfr_cmp(outi[k], ini[k]) != 0 || mpfr_signbit(outi[k]) != mpfr_signbit(ini[k]);
    }
    CHECK(bad == 0, "mpfr pipeline n=%lu P=%ld: %d mismatches", n, P, bad);
    double ms[4];
    mpfft_plan_last_timing(p, ms);
    CHECK(ms[3] >= 0, "timing");
    for (unsigned long i = 0; i < 4 * n; i++) mpfr_clear(a[i]);
    gmp_randclear(st);
    mpfft_plan_destroy(p);
}

static void mpfr_pipeline(bool threads)
{
    mpfft::stub_identity_mp.store(1);
    if (threads) {
        std::thread a([] { mpfr_pipeline_one(4096, 256, 1); mpfr_pipeline_one(16, 64, 2); });
        std::thread b([] { mpfr_pipeline_one(1UL << 15, 100, 3); });
        a.join();
        b.join();
    } else {
        mpfr_pipeline_one(16, 53, 4);
        mpfr_pipeline_one(4096, 256, 5);
        mpfr_pipeline_one(1UL << 15, 1100, 6);
    }
    mpfft::stub_identity_mp.store(0);
}

// the mpfr_t call in a child made by fork() after the parent's host pool has
// started: the child has none of the pool's threads and must start its own
static void mpfr_pipeline_forked()
{
    fflush(stdout);
    const pid_t pid = fork();
    if (pid == 0) {
        mpfft::stub_identity_mp.store(1);
        mpfr_pipeline_one(4096, 256, 7);
        _exit(fails ? 1 : 0);
    }
    int status = 0;
    CHECK(pid > 0 && waitpid(pid, &status, 0) == pid, "fork");
    CHECK(WIFEXITED(status) && WEXITSTATUS(status) == 0, "mpfr call in a forked child: status %d", status);
}

// the plain form's subtrees finished in LDS (MPFFT_SUB_SIZE): their tables,
// checked by the stand-in launcher against the recursion, at every
// transform size and subtree size, at 2, 8, 16 and 32 limbs (a size whose
// workgroup would not fit the LDS falls back to the level kernels)
static void subtrees()
{
    using mpfft::stub_sub_checked;
    using mpfft::stub_sub_errors;
    const int before = stub_sub_checked;
    for (const char *S : {(const char *)nullptr, "8", "16", "64", "256", "512", "2048"}) {
        if (S) setenv("MPFFT_SUB_SIZE", S, 1); else unsetenv("MPFFT_SUB_SIZE");
        for (long prec : {64L, 256L, 512L, 1024L})
            for (int lg = 3; lg <= (prec >= 512 ? 12 : 16); lg++) {
                mpfft_plan p;
                const unsigned long N = 1UL << lg;
                if (mpfft_plan_create(&p, N, prec, lg % 3 == 0 ? MPFFT_FORM_MP_C2 : MPFFT_FORM_MP, 2)) {
                    CHECK(false, "plan N=%lu P=%ld", N, prec);
                    continue;
                }
                std::vector<uint32_t> buf(2 * 2 * N * ((prec + 31) / 32 + 1), 0);
                CHECK(mpfft_execute_limbs_host(p, buf.data(), buf.data()) == 0, "run N=%lu", N);
                mpfft_plan_destroy(p);
            }
    }
    unsetenv("MPFFT_SUB_SIZE");
    CHECK(stub_sub_errors == 0, "%d subtree table errors", (int)stub_sub_errors);
    CHECK(stub_sub_checked > before + 100, "subtree launches checked: %d", stub_sub_checked - before);
}

int main(int argc, char **argv)
{
    if (argc > 1 && std::string(argv[1]) == "threads") {
        batches_concurrently();
        mpfr_pipeline(true);
        printf(fails ? "FAIL\n" : "PASS\n");
        return fails ? 1 : 0;
    }
    if (argc > 1) return legacy(argv[1]);
    plans();
    mpfr_pipeline(false);
    mpfr_pipeline_forked();
    subtrees();
    alloc_failure_then_retry();
    noalloc_stats();
    batch(MPFFT_FORM_MP, {1, 2, 4, 8, 16, 64, 512, 512, 64, 32, 256, 128, 8, 512},
          {64, 100, 128, 256, 53, 128, 256, 256, 1024, 200, 64, 512, 128, 100}, 0);
    srand(11);
    std::vector<unsigned long> n;
    std::vector<long> P;
    const long ps[] = {128, 256, 512, 1024, 2048};
    for (int i = 0; i < 160; i++) {
        n.push_back(1UL << (4 + rand() % 7));
        P.push_back(ps[rand() % 5]);
    }
    batch(MPFFT_FORM_FIXED, n, P, MPFFT_ERR_VALUE);
    batch(MPFFT_FORM_FIXED_PANEL, {256, 256, 256, 16}, {256, 256, 256, 2048}, MPFFT_ERR_VALUE);
    // a plan of every form, created and run through the host path
    const int forms[] = {MPFFT_FORM_MP, MPFFT_FORM_MP_C1, MPFFT_FORM_MP_C2, MPFFT_FORM_FIXED,
                         MPFFT_FORM_FIXED_LANE, MPFFT_FORM_DD, MPFFT_FORM_QD};
    for (int f : forms)
        for (unsigned long N : {1UL, 2UL, 16UL, 128UL, 2048UL}) {
            const long prec = f == MPFFT_FORM_DD ? 106 : f == MPFFT_FORM_QD ? 212 : 256;
            mpfft_plan p;
            int rc = mpfft_plan_create(&p, N, prec, f, 3);
            CHECK(rc == 0, "plan form %d N %lu rc %d", f, N, rc);
            if (rc) continue;
            long L;
            size_t w;
            mpfft_plan_info(p, &L, &w);
            std::vector<uint32_t> x(3 * w, 0), y(3 * w, 0);
            mpfft_execute_limbs_host(p, x.data(), y.data());
            mpfft_plan_destroy(p);
        }
    printf(fails ? "FAIL\n" : "PASS\n");
    return fails ? 1 : 0;
}
