// fk_engine.hip -- host runtime and C ABI (include/fk_engine.h).
//
// An engine owns one family's resident tables on one device: the wheel
// residues, the LDS-stage prime tables (pattern words and marking pairs)
// and its place in the device-wide prime table that the per-candidate
// rounds read.  A batch owns one HIP stream and the grow-only queues a
// launch sequence needs; a sweep is a batch of one window.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fk_engine.h"
#include "fk_device.h"
#include "fk_kill.h"
#include "fk_launch.h"
#include "fk_plan.h"

using namespace fk;

// ------------------------------------------------------------- errors

static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(call)                                                       \
    do {                                                                   \
        hipError_t e_ = (call);                                            \
        if (e_ != hipSuccess)                                              \
            return fail(e_ == hipErrorOutOfMemory ? FK_E_OOM : FK_E_DEVICE, \
                        "%s failed: %s", #call, hipGetErrorString(e_));    \
    } while (0)

// ----------------------------------------------- device-wide prime table

namespace {

struct DeviceCtx {
    int device = -1;
    int cus = 0;
    uint32_t limit = 0;             // primes below this are in the table
    std::vector<uint32_t> primes;   // host copy
    PrimeK* d_primes = nullptr;     // Barrett form (table builds)
    PrimeD* d_primed = nullptr;     // double-precision form (rounds)
};

std::mutex g_ctx_mu;
std::vector<std::unique_ptr<DeviceCtx>> g_ctx;

int ctx_for(int device, uint32_t need_limit, DeviceCtx** out)
{
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    DeviceCtx* c = nullptr;
    for (auto& p : g_ctx)
        if (p->device == device) c = p.get();
    if (!c) {
        g_ctx.emplace_back(new DeviceCtx());
        c = g_ctx.back().get();
        c->device = device;
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, device));
        c->cus = prop.multiProcessorCount;
    }
    if (need_limit > c->limit) {
        uint32_t lim = std::max(need_limit, 1u << 16);
        std::vector<uint32_t> ps = primes_below(lim);
        std::vector<PrimeK> pk(ps.size());
        std::vector<PrimeD> pd(ps.size());
        for (size_t i = 0; i < ps.size(); i++) {
            pk[i] = make_primek(ps[i]);
            pd[i] = make_primed(ps[i]);
        }
        PrimeK* d = nullptr;
        PrimeD* dd = nullptr;
        HIPCHK(hipSetDevice(device));
        HIPCHK(hipMalloc(&d, pk.size() * sizeof(PrimeK)));
        HIPCHK(hipMemcpy(d, pk.data(), pk.size() * sizeof(PrimeK),
                         hipMemcpyHostToDevice));
        HIPCHK(hipMalloc(&dd, pd.size() * sizeof(PrimeD)));
        HIPCHK(hipMemcpy(dd, pd.data(), pd.size() * sizeof(PrimeD),
                         hipMemcpyHostToDevice));
        // engines still hold indices only, which stay valid: the table is
        // the same ascending prime list, extended
        if (c->d_primes) (void)hipFree(c->d_primes);
        if (c->d_primed) (void)hipFree(c->d_primed);
        c->d_primes = d;
        c->d_primed = dd;
        c->primes.swap(ps);
        c->limit = lim;
    }
    *out = c;
    return FK_OK;
}

uint32_t prime_index(const std::vector<uint32_t>& ps, uint32_t q)
{
    return (uint32_t)(std::lower_bound(ps.begin(), ps.end(), q) - ps.begin());
}

template <class T>
int dev_upload(const std::vector<T>& v, T** out)
{
    *out = nullptr;
    size_t bytes = std::max<size_t>(v.size(), 1) * sizeof(T);
    HIPCHK(hipMalloc(out, bytes));
    if (!v.empty())
        HIPCHK(hipMemcpy(*out, v.data(), v.size() * sizeof(T),
                         hipMemcpyHostToDevice));
    return FK_OK;
}

inline u128 mk128(const uint64_t v[2]) { return ((u128)v[1] << 64) | v[0]; }

}  // namespace

// --------------------------------------------------------------- engine

struct fk_batch;

struct fk_engine {
    fk_engine_desc desc;
    int device = 0;
    int family = 0, n = 0;
    uint32_t q2 = 0;
    std::vector<uint32_t> wheel;
    uint64_t D = 1;
    std::vector<uint64_t> offs;      // the wheel's residues (host-built wheels)
    uint64_t nlines = 0;             // the wheel's residue count
    bool factored = false;           // residues expanded on the device
    uint64_t W = 1;
    u128 dval = 0;
    uint32_t q_pat = 0, q_med = 0, threads = 256, wpt = 8;
    uint32_t round_lane_prime = 8192, growth = 2;
    uint32_t short_words = 1024;     // 0: no short-line sieve
    uint32_t segment_cap = 0;        // 0: queue segment from the expectation
    uint32_t q_mark = 0;             // [q_pat, q_mark) stride-marked
    uint32_t q_bits = 0;             // [q_mark, q_bits) ORed in from bitmaps
    uint32_t npat = 0;
    std::vector<uint32_t> lds_q;
    std::vector<uint32_t> kill_w;    // |K_q| per LDS prime
    uint32_t pat_words = 0;
    uint32_t round_first = 0, round_end = 0;
    uint32_t nrounds = 0;
    double surv_lds = 1.0;
    double mark_density = 0.0;       // sum |K_q| / q over marked primes
    double line_density = 1.0;       // nlines / D
    DevEngine host{};                // with device pointers
    DevEngine* d_engine = nullptr;   // device copy of `host`
    std::vector<void*> allocs;
    uint64_t device_bytes = 0;
    fk_batch* own = nullptr;
    // Prefix wheels for short windows (the wheel without its last 1, 2, ...
    // primes, which their LDS stage then sieves): sieve-side tables only,
    // the round table is the parent's.  A window whose lines hold few
    // positions on the full wheel is sieved on the one with the least work.
    std::vector<fk_engine*> alts;
    const fk_engine* parent = nullptr;
};

static void engine_free(fk_engine* e)
{
    if (!e) return;
    for (fk_engine* a : e->alts) {
        engine_free(a);
        delete a;
    }
    e->alts.clear();
    (void)hipSetDevice(e->device);
    for (void* p : e->allocs) (void)hipFree(p);
    e->allocs.clear();
}

template <class T>
static int eng_upload(fk_engine* e, const std::vector<T>& v, const T** out)
{
    T* d = nullptr;
    int rc = dev_upload(v, &d);
    if (rc) return rc;
    e->allocs.push_back(d);
    e->device_bytes += std::max<size_t>(v.size(), 1) * sizeof(T);
    *out = d;
    return FK_OK;
}

// In-tile test rounds are cut where the expected survival over the round
// reaches this fraction (fewer rounds: fewer barriers, more tests on
// positions an earlier cut would have dropped).
static constexpr double kTileRoundSurvival = 0.25;

// Largest survivor queue segment of one engine in one batch, in records
// (24 B each, two queues): the sieve's per-engine counters are 32-bit, and
// a window expected to need more is refused (split it) rather than risk a
// counter that wraps.
static constexpr uint32_t kSegmentMax = 1u << 30;

// Wheels of more residues than this are built factored (the last prime's
// lifts expanded on the device): the Euler 37# wheel at n = 17 has
// 598,752,000 (4.8 GB), the 31# one 29,937,600.
static constexpr uint64_t kHostWheelResidues = 1ull << 25;

static int build_engine(fk_engine* e)
{
    const fk_engine_desc& d = e->desc;
    e->family = d.family;
    e->n = d.n;
    e->q2 = d.q2;
    e->device = d.device;
    if (d.family < 1 || d.family > 3)
        return fail(FK_E_ARG, "unknown family %d", d.family);
    if (d.n < 1 || d.n > 64)
        return fail(FK_E_ARG, "n = %d outside [1, 64]", d.n);
    if (d.family == FK_FAMILY_AP && d.n > 26)
        return fail(FK_E_ARG, "AP n = %d: P(n) must fit 128 bits (n <= 26)",
                    d.n);
    if (d.q2 < 3 || d.q2 > (1u << 30))
        return fail(FK_E_ARG, "q2 = %u outside [3, 2^30]", d.q2);

    const uint32_t first = first_kill_prime(d.family, d.n);
    std::vector<uint32_t> kp;
    for (uint32_t q : primes_below(d.q2))
        if (q >= first) kp.push_back(q);

    // ---- the wheel
    std::vector<uint32_t> wheel;
    if (d.wheel_primes[0]) {
        for (int i = 0; i < 12 && d.wheel_primes[i]; i++)
            wheel.push_back(d.wheel_primes[i]);
        for (size_t i = 0; i < wheel.size(); i++)
            if (i >= kp.size() || wheel[i] != kp[i])
                return fail(FK_E_ARG, "wheel primes must be the first kill "
                            "primes of the family, in order");
    } else {
        int nw = d.nwheel;
        if (nw <= 0) {              // widest wheel with period <= 2^17
            uint64_t m = 1;
            nw = 0;
            while (nw < (int)kp.size() && m * kp[nw] <= (1u << 17)) {
                m *= kp[nw];
                nw++;
            }
            if (nw == 0 && !kp.empty()) nw = 1;
        }
        for (int i = 0; i < nw && i < (int)kp.size(); i++)
            wheel.push_back(kp[i]);
    }
    if (wheel.empty()) return fail(FK_E_ARG, "no kill prime below q2");
    bool ok = false;
    e->nlines = wheel_count(d.family, d.n, wheel, &ok);
    if (!ok) return fail(FK_E_ARG, "wheel period exceeds 2^63");
    if (e->nlines == 0)
        return fail(FK_E_ARG, "the wheel admits no residue at n = %d", d.n);
    if (e->nlines > (1u << 30))
        return fail(FK_E_ARG, "wheel has %llu residues (> 2^30)",
                    (unsigned long long)e->nlines);
    // a wheel of many residues is built factored: the residues of all but
    // its last prime on the host, lifted over the last prime on the device
    e->factored = e->nlines > kHostWheelResidues && wheel.size() > 1;
    std::vector<uint64_t> fbase;
    std::vector<uint32_t> flifts;
    uint32_t fnj = 0;
    uint64_t fMb = 0;
    if (e->factored) {
        std::vector<uint32_t> head(wheel.begin(), wheel.end() - 1);
        fbase = wheel_offsets(d.family, d.n, head, &fMb, &ok);
        if (!ok) return fail(FK_E_ARG, "wheel period exceeds 2^63");
        flifts = wheel_lifts(d.family, d.n, wheel.back(), fMb, &fnj);
        e->D = fMb * wheel.back();
        if ((uint64_t)fbase.size() * fnj != e->nlines)
            return fail(FK_E_ARG, "factored wheel count disagrees");
    } else {
        e->offs = wheel_offsets(d.family, d.n, wheel, &e->D, &ok);
        if (!ok) return fail(FK_E_ARG, "wheel period exceeds 2^63");
    }
    e->wheel = wheel;
    e->line_density = (double)e->nlines / (double)e->D;

    // ---- tuning
    e->threads = d.threads ? d.threads : 256;
    e->wpt = d.words_per_thread ? d.words_per_thread : 8;
    if (!sieve_shape_ok((int)e->threads, (int)e->wpt))
        return fail(FK_E_ARG, "unsupported sieve shape threads=%u wpt=%u",
                    e->threads, e->wpt);
    e->q_pat = d.q_pat ? d.q_pat : 128;
    e->q_med = d.q_med ? d.q_med : 2048;
    if (e->q_med > 65536) e->q_med = 65536;     // pairs pack q in 16 bits
    if (e->q_med > d.q2) e->q_med = d.q2;
    if (e->q_pat > e->q_med) e->q_pat = e->q_med;
    e->round_lane_prime = d.round_lane_prime ? d.round_lane_prime : 8192;
    if (d.short_words > 4096)
        return fail(FK_E_ARG, "short_words = %u > 4096", d.short_words);
    e->short_words = d.short_words ? d.short_words : 1024;
    if (d.segment_cap > kSegmentMax)
        return fail(FK_E_ARG, "segment_cap = %u > 2^30", d.segment_cap);
    e->segment_cap = d.segment_cap;
    e->q_mark = d.q_mark ? d.q_mark : e->q_pat;
    if (e->q_mark < e->q_pat) e->q_mark = e->q_pat;
    if (e->q_mark > e->q_med) e->q_mark = e->q_med;
    e->q_bits = d.q_bits ? d.q_bits : e->q_mark;
    if (e->q_bits < e->q_mark) e->q_bits = e->q_mark;
    if (e->q_bits > e->q_med) e->q_bits = e->q_med;
    // round growth: an integer factor (2..99), or hundredths from 101 up
    // (141: each round's prime bound 1.41 times the last)
    if (d.round_growth == 1 || d.round_growth == 100 || d.round_growth > 9999)
        return fail(FK_E_ARG, "round_growth = %u: use 2..99, or hundredths "
                    "101..9999", d.round_growth);
    e->growth = d.round_growth ? d.round_growth : 2;

    // ---- LDS-stage primes: past the wheel, below q_med; in order the
    // pattern primes, the stride-marked primes, the bitmap primes (ORed in
    // as 64-bit windows below q_bits, tested per surviving position above)
    std::vector<uint32_t> lq, lm32, ldinv, pat_off, pairs, bm_off, bm;
    std::vector<uint32_t> sbm_off, sbm;     // every prime's bitmap
    std::vector<uint64_t> lm64, pat;
    uint32_t nmark = 0, nbits = 0;
    double surv = 1.0;
    for (size_t i = wheel.size(); i < kp.size() && kp[i] < e->q_med; i++) {
        const uint32_t q = kp[i];
        const uint32_t dinv = inv_mod((uint32_t)(e->D % q), q);
        std::vector<uint32_t> ks = kill_set(d.family, d.n, q);
        std::vector<uint32_t> kl;                 // line-domain residues
        for (uint32_t f : ks) kl.push_back((uint32_t)((uint64_t)f * dinv % q));
        std::sort(kl.begin(), kl.end());
        const uint32_t idx = (uint32_t)lq.size();
        lq.push_back(q);
        lm32.push_back((uint32_t)(((uint64_t)1 << 32) / q));
        lm64.push_back(make_primek(q).m64);
        ldinv.push_back(dinv);
        e->kill_w.push_back((uint32_t)kl.size());
        surv *= 1.0 - (double)kl.size() / q;
        append_kill_bitmap(kl, q, kShortTailBits, sbm_off, sbm);
        if (q < e->q_pat) {
            pat_off.push_back((uint32_t)pat.size());
            std::vector<uint8_t> dead(q, 0);
            for (uint32_t r : kl) dead[r] = 1;
            for (uint32_t s = 0; s < q; s++) {
                uint64_t w = 0;
                for (int b = 0; b < 64; b++)
                    if (dead[(s + b) % q]) w |= 1ull << b;
                pat.push_back(w);
            }
            e->npat = idx + 1;
        } else if (q < e->q_mark) {
            for (uint32_t r : kl) pairs.push_back((idx << 16) | r);
            e->mark_density += (double)kl.size() / q;
            nmark = idx + 1;
        } else {
            append_kill_bitmap(kl, q, 64, bm_off, bm);
            if (q < e->q_bits) nbits = idx + 1;
        }
    }
    if (nmark < e->npat) nmark = e->npat;
    if (nbits < nmark) nbits = nmark;
    // in-tile rounds: cut the tested primes where the expected survival of
    // the round reaches kTileRoundSurvival
    std::vector<uint32_t> tr_end;
    {
        double lsv = 0.0;
        for (uint32_t i = nbits; i < (uint32_t)lq.size(); i++) {
            lsv += std::log(1.0 - (double)e->kill_w[i] / lq[i]);
            if (lsv <= std::log(kTileRoundSurvival)) {
                tr_end.push_back(i + 1);
                lsv = 0.0;
            }
        }
        if (tr_end.empty() || tr_end.back() != (uint32_t)lq.size())
            if ((uint32_t)lq.size() > nbits) tr_end.push_back((uint32_t)lq.size());
    }
    // the tile's LDS: bits, s0, per-prime info, patterns, kill bitmaps
    {
        const size_t S = (size_t)64 * e->threads * e->wpt;
        const size_t need = S / 8 + 16 * ((lq.size() + 3) / 4) +
                            16 * lq.size() + 8 * pat.size() +
                            4 * ((bm.size() + 3) & ~(size_t)3) +
                            2 * std::min<size_t>(S / 16, 4096) *
                                (S <= 65536 ? 2 : 4);
        if (need > 160 * 1024)
            return fail(FK_E_ARG, "tile needs %zu bytes of LDS (> 160 KiB): "
                        "lower q_med, q_pat or the tile", need);
    }
    // the short-line sieve's LDS: per-prime arrays and every kill bitmap
    // (the kernel's own survivor buffer takes 4 KB more)
    if (short_lds(lq.size(), sbm.size()) > 152 * 1024) e->short_words = 0;
    e->lds_q = lq;
    e->pat_words = (uint32_t)pat.size();
    e->surv_lds = surv;

    // ---- round primes: [q_med, q2) in the device-wide table
    DeviceCtx* ctx = nullptr;
    int rc = ctx_for(d.device, d.q2, &ctx);
    if (rc) return rc;
    e->round_first = prime_index(ctx->primes, std::max(e->q_med, first));
    e->round_first = std::max(e->round_first,
                              prime_index(ctx->primes, wheel.back() + 1));
    e->round_end = prime_index(ctx->primes, d.q2);
    if (e->round_end < e->round_first) e->round_end = e->round_first;
    // round boundaries: prime values q_med * growth^r, as table indices
    std::vector<uint32_t> rb(1, e->round_first);
    if (e->round_end > e->round_first) {
        const double g = e->growth >= 100 ? e->growth / 100.0 : (double)e->growth;
        for (double v = (double)e->q_med * g;; v *= g) {
            const uint32_t idx = std::min(
                e->round_end,
                prime_index(ctx->primes, (uint32_t)std::min<double>(v, (double)d.q2)));
            if (idx > rb.back()) rb.push_back(idx);
            if (v >= (double)d.q2) break;
        }
        if (rb.back() < e->round_end) rb.push_back(e->round_end);
    }
    if (rb.size() > 33)
        return fail(FK_E_ARG, "more than 32 compaction rounds (raise "
                    "round_growth)");
    e->nrounds = (uint32_t)rb.size() - 1;

    // ---- family constants
    e->W = (d.family == FK_FAMILY_DICKSON) ? dickson_wheel(d.n) : 1;
    e->dval = (d.family == FK_FAMILY_AP) ? ap_difference(d.n) : 0;
    if (e->W > 0xffffffffull) return fail(FK_E_ARG, "W(n) exceeds 32 bits");

    HIPCHK(hipSetDevice(d.device));
    DevEngine& h = e->host;
    h.family = d.family;
    h.n = d.n;
    h.nlines = (uint32_t)e->nlines;
    h.nlds = (uint32_t)lq.size();
    h.npat = e->npat;
    h.nmark = nmark;
    h.nbits = nbits;
    h.npairs = (uint32_t)pairs.size();
    h.pat_words = e->pat_words;
    h.bm_words = (uint32_t)bm.size();
    h.sbm_words = (uint32_t)sbm.size();
    h.ntr = (uint32_t)tr_end.size();
    h.round_first = e->round_first;
    h.round_end = e->round_end;
    // AP rounds test one index per prime when no round prime divides d:
    // the primes dividing d are the first n, so the first round prime
    // decides
    h.kidx = 0;
    if (kill_indexable(d.family))
        h.kidx = (e->round_first >= e->round_end ||
                  e->dval % ctx->primes[e->round_first] != 0) ? 1u : 0u;
    h.kstride = h.kidx ? kIndexStride : round_row_stride(d.family, d.n);
    h.pronic_first = e->round_end;
    if (d.family == FK_FAMILY_EULER)
        h.pronic_first = std::min(
            e->round_end,
            std::max(e->round_first,
                     prime_index(ctx->primes, (uint32_t)(d.n - 1) * (uint32_t)d.n + 1u)));
    h.nround = e->nrounds;
    for (size_t i = 0; i < 33; i++)
        h.rb[i] = i < rb.size() ? rb[i] : e->round_end;
    h.D = e->D;
    h.W = (uint32_t)e->W;
    for (int i = 0; i < 4; i++) {
        h.d[i] = (uint32_t)(e->dval >> (32 * i));
        h.d[4 + i] = 0;
    }
    if (!e->factored) {
        if ((rc = eng_upload(e, e->offs, &h.offs))) return rc;
    } else {
        // expand the factored wheel into the engine's residue array
        uint64_t* offs = nullptr;
        HIPCHK(hipMalloc(&offs, e->nlines * sizeof(uint64_t)));
        e->allocs.push_back(offs);
        e->device_bytes += e->nlines * sizeof(uint64_t);
        uint64_t* db = nullptr;
        uint32_t* dl = nullptr;
        hipError_t err = hipMalloc(&db, fbase.size() * sizeof(uint64_t));
        if (err == hipSuccess) err = hipMalloc(&dl, flifts.size() * sizeof(uint32_t));
        if (err == hipSuccess)
            err = hipMemcpy(db, fbase.data(), fbase.size() * sizeof(uint64_t),
                            hipMemcpyHostToDevice);
        if (err == hipSuccess)
            err = hipMemcpy(dl, flifts.data(), flifts.size() * sizeof(uint32_t),
                            hipMemcpyHostToDevice);
        if (err == hipSuccess)
            err = launch_wheel_expand(db, fbase.size(), dl, fnj, wheel.back(),
                                      fMb, offs, nullptr);
        if (err == hipSuccess) err = hipDeviceSynchronize();
        if (db) (void)hipFree(db);
        if (dl) (void)hipFree(dl);
        if (err != hipSuccess)
            return fail(err == hipErrorOutOfMemory ? FK_E_OOM : FK_E_DEVICE,
                        "factored wheel: %s", hipGetErrorString(err));
        h.offs = offs;
    }
    if ((rc = eng_upload(e, lq, &h.lds_q))) return rc;
    if ((rc = eng_upload(e, lm32, &h.lds_m32))) return rc;
    if ((rc = eng_upload(e, lm64, &h.lds_m64))) return rc;
    if ((rc = eng_upload(e, ldinv, &h.lds_dinv))) return rc;
    if ((rc = eng_upload(e, pat_off, &h.pat_off))) return rc;
    if ((rc = eng_upload(e, pat, &h.pat))) return rc;
    if ((rc = eng_upload(e, pairs, &h.pairs))) return rc;
    if ((rc = eng_upload(e, bm_off, &h.bm_off))) return rc;
    if ((rc = eng_upload(e, bm, &h.bm))) return rc;
    if ((rc = eng_upload(e, sbm_off, &h.sbm_off))) return rc;
    {
        std::vector<uint32_t> spd(kShortPrimeWords * lq.size(), 0);
        std::vector<double> spe(lq.size());
        for (size_t p = 0; p < lq.size(); p++) {
            const uint32_t q = lq[p];
            const double c[3] = {(double)q, 1.0 / (double)q, (double)ldinv[p]};
            uint32_t* k = &spd[kShortPrimeWords * p];
            memcpy(k, c, sizeof c);
            k[6] = q;
            k[7] = sbm_off[p];
            spe[p] = (double)(((1ull << 32) % q) * ldinv[p] % q);
        }
        if ((rc = eng_upload(e, spd, &h.spd))) return rc;
        if ((rc = eng_upload(e, spe, &h.spe))) return rc;
    }
    if ((rc = eng_upload(e, sbm, &h.sbm))) return rc;
    if ((rc = eng_upload(e, tr_end, &h.tr_end))) return rc;
    if (e->parent) {
        // a prefix wheel tests the same round primes with its parent's table
        const DevEngine& ph = e->parent->host;
        if (ph.round_first != h.round_first || ph.round_end != h.round_end ||
            ph.nround != h.nround)
            return fail(FK_E_ARG, "prefix wheel's rounds differ from its engine's");
        h.kt = ph.kt;
        h.kidx = ph.kidx;
        h.kstride = ph.kstride;
        // the row order is per round: shared when the rounds are the same
        h.kperm = std::equal(ph.rb, ph.rb + ph.nround + 1, h.rb) ? ph.kperm : nullptr;
    } else {
        // kill sets of the round primes, built on the device
        const uint32_t np = e->round_end - e->round_first;
        const size_t nt = (size_t)np * h.kstride + kRowTail;
        uint32_t* kt = nullptr;
        HIPCHK(hipMalloc(&kt, nt * sizeof(uint32_t)));
        HIPCHK(hipMemset(kt, 0xff, nt * sizeof(uint32_t)));   // kKillPad
        e->allocs.push_back(kt);
        e->device_bytes += nt * sizeof(uint32_t);
        HIPCHK(launch_kill_table(ctx->d_primes, e->round_first, np, d.family,
                                 d.n, h.d, h.kidx != 0, h.kstride, kt, nullptr));
        HIPCHK(hipDeviceSynchronize());
        h.kt = kt;
        h.kperm = nullptr;
        if (!h.kidx && np > 0) {
            // each round's rows by groups of four residues (1, 2, 3, more),
            // stable: a wave of the lane = prime kernel then mostly holds
            // rows of one group count and compares no more groups than its
            // rows need
            std::vector<uint32_t> cnt(np), perm(np);
            HIPCHK(hipMemcpy2D(cnt.data(), sizeof(uint32_t),
                               kt + offsetof(RoundRow, cnt) / sizeof(uint32_t),
                               h.kstride * sizeof(uint32_t), sizeof(uint32_t), np,
                               hipMemcpyDeviceToHost));
            for (uint32_t r = 0; r < h.nround; r++) {
                const uint32_t lo = h.rb[r] - e->round_first;
                const uint32_t hi = h.rb[r + 1] - e->round_first;
                uint32_t at = lo;
                for (uint32_t g = 1; g <= 4; g++)
                    for (uint32_t i = lo; i < hi; i++) {
                        const uint32_t gi = std::min(4u, std::max(1u, (cnt[i] + 3u) / 4u));
                        if (gi == g) perm[at++] = i;
                    }
            }
            if ((rc = eng_upload(e, perm, &h.kperm))) return rc;
        }
    }
    std::vector<DevEngine> one(1, h);
    const DevEngine* de = nullptr;
    if ((rc = eng_upload(e, one, &de))) return rc;
    e->d_engine = const_cast<DevEngine*>(de);
    return FK_OK;
}

// Prefix wheels: by default up to kAltWheels, of at most kAltLines
// residues each (their host tables are built at engine creation).  2^25
// admits the Euler 31# wheel (29,937,600 residues at n = 17, 240 MB) under
// the 37# one: a 5e14 window (the reference's SCORE shapes) holds 67
// positions on each 37# line and 2,493 on each 31# line, and sweeps in 13
// against 27 ms on it.
static constexpr int kAltWheels = 4;
static constexpr uint64_t kAltLines = 1ull << 25;

// The engine's prefix wheels (fk_engine_desc.alt_wheels): the wheel without
// its last 1, 2, ... primes, each an engine of its own sieve-side tables
// over the parent's round table.  Only when every wheel prime is below
// q_med, so that the primes a prefix drops fall to its LDS stage and the
// rounds start where the parent's do.  A prefix that cannot be built (its
// LDS tables do not fit) is skipped: the full wheel always works.
static int build_alts(fk_engine* e)
{
    const int want = e->desc.alt_wheels == 0 ? kAltWheels : e->desc.alt_wheels;
    if (want < 0 || e->wheel.size() < 2 || e->wheel.back() >= e->q_med ||
        e->short_words == 0)
        return FK_OK;
    for (size_t len = e->wheel.size() - 1; len >= 1 && (int)e->alts.size() < want;
         len--) {
        std::vector<uint32_t> prefix(e->wheel.begin(), e->wheel.begin() + len);
        bool ok = false;
        const uint64_t count = wheel_count(e->family, e->n, prefix, &ok);
        if (!ok || count == 0 || count > kAltLines) continue;
        fk_engine* a = new fk_engine();
        a->desc = e->desc;
        std::memset(a->desc.wheel_primes, 0, sizeof a->desc.wheel_primes);
        for (size_t i = 0; i < len; i++) a->desc.wheel_primes[i] = prefix[i];
        a->desc.nwheel = (int32_t)len;
        a->desc.alt_wheels = -1;
        a->parent = e;
        const std::string keep = g_err;
        const int rc = build_engine(a);
        if (rc || a->short_words == 0) {
            g_err = keep;
            engine_free(a);
            delete a;
            continue;
        }
        e->device_bytes += a->device_bytes;
        e->alts.push_back(a);
    }
    return FK_OK;
}

// ---------------------------------------------------------------- batch

// One step of a batch's launch sequence.  batch_enqueue records the whole
// sequence first (every allocation and occupancy query done), then either
// issues it on the batch's stream, with stage events, or captures it once
// into a HIP graph and replays the graph while the recorded sequence stays
// byte-identical (same grids, same argument blocks, same buffers; only the
// device-side inputs uploaded before it differ).  Steps are zero-filled
// before they are set, so that byte comparison is exact.
struct Step {
    enum Kind : int32_t { kMark, kReadback, kSieve, kRound, kCompact,
                          kGather, kClassify };
    int32_t kind;
    int32_t ev;                      // kMark: event index
    int32_t side;                    // kRound: launched on the side stream,
                                     // beside the round's other launch
    uint32_t grid, gy;
    int32_t threads, wpt;
    size_t lds;
    int32_t lane_prime;
    SieveArgs sa;
    RoundArgs ra;
    CompactArgs ca;
    GatherArgs ga;
    ClassifyArgs cla;
    ReadbackArgs rb;
};

static Step new_step(int32_t kind)
{
    Step st;
    std::memset(&st, 0, sizeof st);
    st.kind = kind;
    return st;
}

static bool same_plan(const std::vector<Step>& a, const std::vector<Step>& b)
{
    return a.size() == b.size() &&
           (a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(Step)) == 0);
}

// The launch step's kind as fk_batch_step_times reports it.
static int32_t step_kind(const Step& st)
{
    switch (st.kind) {
    case Step::kSieve: return FK_STEP_SIEVE;
    case Step::kRound:
        return st.lane_prime ? FK_STEP_ROUND_PRIME : FK_STEP_ROUND_CANDIDATE;
    case Step::kCompact: return FK_STEP_COMPACT;
    case Step::kGather: return FK_STEP_GATHER;
    case Step::kClassify: return FK_STEP_CLASSIFY;
    case Step::kReadback: return FK_STEP_READBACK;
    default: return 0;
    }
}

// The side stream of a batch whose rounds test both kinds of row: a round's
// two launches touch disjoint queue segments, so one runs on the side
// stream (forked from the batch's stream before the other is issued) and
// the next step waits for it.  In a captured graph they are two parallel
// nodes.
struct SideStream {
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};

// Issue the recorded steps on `s`; stage events only when `events`, and
// with `step_ev` (two events per launch step) a start and a stop event for
// every launch step.  The sieve, round and compaction kernels carry theirs
// into the launch, which stamps them at the kernel's own start and end; the
// other steps (several kernels or a copy) are bracketed by recorded events.
static hipError_t emit_steps(const std::vector<Step>& plan, hipEvent_t* ev,
                             bool events, hipStream_t s, const SideStream& side,
                             const std::vector<hipEvent_t>* step_ev = nullptr)
{
    hipError_t e = hipSuccess;
    size_t k = 0;
    // a parallel section: side launches of one kind (forked once, in a row
    // on the side stream), then the batch stream's launches of that kind;
    // the first step of another kind waits for the side stream
    bool joined = true;
    int32_t section = 0;
    for (const Step& st : plan) {
        const bool on_side = (st.kind == Step::kRound || st.kind == Step::kSieve) &&
                             st.side && side.stream;
        if (!joined && !(st.kind == section && (on_side || !st.side))) {
            e = hipStreamWaitEvent(s, side.join, 0);
            if (e != hipSuccess) return e;
            joined = true;
        }
        if (st.kind == Step::kMark) {
            if (events) e = hipEventRecord(ev[st.ev], s);
            if (e != hipSuccess) return e;
            continue;
        }
        LaunchEvents le;
        if (step_ev) {
            le.start = (*step_ev)[2 * k];
            le.stop = (*step_ev)[2 * k + 1];
            k++;
        }
        const LaunchEvents* lp = step_ev ? &le : nullptr;
        if (on_side) {
            if (joined) {
                e = hipEventRecord(side.fork, s);
                if (e == hipSuccess) e = hipStreamWaitEvent(side.stream, side.fork, 0);
            }
            if (e == hipSuccess)
                e = st.kind == Step::kRound
                        ? launch_round_test(st.ra, st.lane_prime != 0, st.grid,
                                            st.gy, side.stream, lp)
                        : (st.wpt == 0 ? launch_sieve_short(st.sa, st.threads, st.grid,
                                                            st.lds, side.stream, lp)
                                       : launch_sieve(st.sa, st.threads, st.wpt, st.grid,
                                                      st.lds, side.stream, lp));
            if (e == hipSuccess) e = hipEventRecord(side.join, side.stream);
            if (e != hipSuccess) return e;
            joined = false;
            section = st.kind;
            continue;
        }
        const bool bracket = step_ev && (st.kind == Step::kGather ||
                                         st.kind == Step::kClassify ||
                                         st.kind == Step::kReadback);
        if (bracket) e = hipEventRecord(le.start, s);
        if (e != hipSuccess) return e;
        switch (st.kind) {
        case Step::kSieve:
            e = st.wpt == 0 ? launch_sieve_short(st.sa, st.threads, st.grid,
                                                 st.lds, s, lp)
                            : launch_sieve(st.sa, st.threads, st.wpt, st.grid,
                                           st.lds, s, lp);
            break;
        case Step::kRound:
            e = launch_round_test(st.ra, st.lane_prime != 0, st.grid, st.gy, s,
                                  lp);
            break;
        case Step::kCompact:
            e = launch_round_compact(st.ca, st.grid, s, lp);
            break;
        case Step::kGather:
            e = launch_gather(st.ga, st.grid, s);
            break;
        case Step::kClassify:
            e = launch_classify(st.cla, st.grid, s);
            break;
        case Step::kReadback:
            e = launch_readback(st.rb, s);
            break;
        default:
            break;
        }
        if (e == hipSuccess && bracket) e = hipEventRecord(le.stop, s);
        if (e != hipSuccess) return e;
    }
    if (!joined) e = hipStreamWaitEvent(s, side.join, 0);
    return e;
}

struct fk_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    SideStream side;                  // made when a plan first needs it
    hipEvent_t ev[6] = {};
    // eager submissions: a start and a stop event for every launch step
    // (grow-only), the steps' kinds and, once collected, their durations
    std::vector<hipEvent_t> step_ev;
    std::vector<int32_t> step_kinds;
    std::vector<float> step_ms;
    bool pending = false;
    bool throughput = false;       // FK_BATCH_THROUGHPUT
    bool have_results = false;
    // device buffers (grow-only)
    // the per-submission descriptors live in one device area, uploaded in
    // one copy from a pinned staging area of the same layout
    uint8_t* d_up = nullptr;
    uint8_t* h_up = nullptr;
    size_t up_bytes = 0;
    size_t o_engines = 0, o_pieces = 0, o_win_engine = 0, o_wins = 0, o_segs = 0;
    size_t o_counts = 0;
    DevEngine* d_engines = nullptr;   size_t cap_engines = 0;
    DevPiece* d_pieces = nullptr;     size_t cap_pieces = 0;
    uint32_t* d_win_engine = nullptr; size_t cap_win = 0;
    DevWinInfo* d_wins = nullptr;
    DevRec* d_q[2] = {nullptr, nullptr}; size_t cap_q = 0;
    uint32_t* d_flags = nullptr;
    uint64_t* d_fail2 = nullptr;      // per record: base-2 failures
    uint64_t* d_failx = nullptr;      // per record: other-base failures
    uint32_t* d_counts = nullptr;     size_t cap_counts = 0;  // in d_up
    // pinned host mirrors
    uint32_t* h_counts = nullptr;
    DevRec* h_recs = nullptr;         size_t cap_hrecs = 0;
    // the submitted job, kept for results() and re-runs
    std::vector<fk_engine*> engines;
    std::vector<uint32_t> engine_of;
    std::vector<fk_window> windows;   // what this submission sweeps
    // the job behind the submissions: each window's requested end and its
    // cursor (the start of what is not yet swept; advanced when a
    // submission's results are collected), and the sweep budget
    std::vector<u128> job_hi, cursor;
    u128 budget = 0;                  // line values per submission (0: all)
    std::vector<fk_classify_spec> cls;
    uint32_t final_slot = 0;          // which counter holds the final count
    std::vector<uint32_t> round_slots;  // counter of each executed round
    std::vector<uint32_t> stage_counts; // queue after the sieve and rounds
    uint32_t final_q = 0;             // which queue holds the final records
    uint32_t queue_cap = 0;           // entries per queue (all segments)
    std::vector<uint32_t> seg_caps;   // per engine: its queue segment
    uint2* d_segs = nullptr;          size_t cap_segs = 0;
    // graph mode: the captured launch sequence and the plan it was made from
    bool use_graph = false;
    bool graph_at_once = false;       // capture a new sequence at once
    bool graph_ran = false;           // the last submission ran as a graph
    bool staged = false;              // the last submission recorded stage events
    std::vector<Step> prev_plan;      // graph mode: the last sequence run eagerly
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::vector<Step> graph_plan;
    uint64_t graph_captures = 0, graph_launches = 0;
    uint64_t reruns = 0;              // batches rerun after a queue overflow
    uint32_t copied = 0;
    uint64_t tiles = 0, candidates = 0;
    std::vector<uint64_t> win_wheel;  // per window: the wheel period sieved on
    std::vector<fk_record> out;       // sorted results
    std::vector<uint64_t> counts;
    float ms[5] = {0, 0, 0, 0, 0};
};

static int batch_free(fk_batch* b)
{
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
    (void)hipSetDevice(b->device);
    void* ds[] = {b->d_up, b->d_q[0], b->d_q[1], b->d_flags,
                  b->d_fail2, b->d_failx};
    for (void* p : ds) if (p) (void)hipFree(p);
    if (b->h_up) (void)hipHostFree(b->h_up);
    if (b->h_counts) (void)hipHostFree(b->h_counts);
    if (b->h_recs) (void)hipHostFree(b->h_recs);
    for (auto& e : b->ev) if (e) (void)hipEventDestroy(e);
    for (auto& e : b->step_ev) if (e) (void)hipEventDestroy(e);
    b->step_ev.clear();
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->side.stream) (void)hipStreamDestroy(b->side.stream);
    if (b->side.fork) (void)hipEventDestroy(b->side.fork);
    if (b->side.join) (void)hipEventDestroy(b->side.join);
    b->side = SideStream{};
    return FK_OK;
}

template <class T>
static int grow(T** p, size_t* cap, size_t need)
{
    if (need <= *cap && *p) return FK_OK;
    size_t nc = std::max<size_t>(need, *cap * 3 / 2);
    nc = std::max<size_t>(nc, 16);
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    HIPCHK(hipMalloc(p, nc * sizeof(T)));
    *cap = nc;
    return FK_OK;
}

// Lay out the descriptor areas by grow-only capacities, so the device
// pointers (and a captured graph's kernel arguments) stay put while the
// submissions fit.
static int grow_uploads(fk_batch* b, size_t neng, size_t npieces,
                        size_t nwin, size_t nseg, size_t ncount)
{
    auto need = [](size_t cap, size_t n) {
        return n > cap ? std::max<size_t>(std::max<size_t>(n, cap * 3 / 2), 16)
                       : cap;
    };
    const size_t ce = need(b->cap_engines, neng), cp = need(b->cap_pieces, npieces),
                 cw = need(b->cap_win, nwin), cs = need(b->cap_segs, nseg),
                 cc = need(b->cap_counts, ncount);
    if (b->d_up && ce == b->cap_engines && cp == b->cap_pieces &&
        cw == b->cap_win && cs == b->cap_segs && cc == b->cap_counts)
        return FK_OK;
    auto align = [](size_t x) { return (x + 255) & ~(size_t)255; };
    size_t off = 0;
    b->o_engines = off;    off = align(off + ce * sizeof(DevEngine));
    b->o_pieces = off;     off = align(off + cp * sizeof(DevPiece));
    b->o_win_engine = off; off = align(off + cw * sizeof(uint32_t));
    b->o_wins = off;       off = align(off + cw * sizeof(DevWinInfo));
    b->o_segs = off;       off = align(off + cs * sizeof(uint2));
    b->o_counts = off;     off = align(off + cc * sizeof(uint32_t));
    if (b->d_up) (void)hipFree(b->d_up);
    if (b->h_up) (void)hipHostFree(b->h_up);
    if (b->h_counts) (void)hipHostFree(b->h_counts);
    b->h_counts = nullptr;
    b->d_up = nullptr;
    b->h_up = nullptr;
    HIPCHK(hipMalloc(&b->d_up, off));
    HIPCHK(hipHostMalloc(&b->h_up, off));
    // coherent: the readback kernel's stores are seen by the host once the
    // stream is synchronized, without passing through the GPU's caches
    HIPCHK(hipHostMalloc(&b->h_counts, cc * sizeof(uint32_t),
                         hipHostMallocCoherent));
    b->up_bytes = off;
    b->d_engines = reinterpret_cast<DevEngine*>(b->d_up + b->o_engines);
    b->d_pieces = reinterpret_cast<DevPiece*>(b->d_up + b->o_pieces);
    b->d_win_engine = reinterpret_cast<uint32_t*>(b->d_up + b->o_win_engine);
    b->d_wins = reinterpret_cast<DevWinInfo*>(b->d_up + b->o_wins);
    b->d_segs = reinterpret_cast<uint2*>(b->d_up + b->o_segs);
    b->d_counts = reinterpret_cast<uint32_t*>(b->d_up + b->o_counts);
    b->cap_engines = ce;
    b->cap_pieces = cp;
    b->cap_win = cw;
    b->cap_segs = cs;
    b->cap_counts = cc;
    return FK_OK;
}

// A batch's stream.  HIP maps streams onto a few hardware queues (four by
// default) and two streams on one queue run one after the other; a
// stream created with a CU mask gets a hardware queue of its own, so the
// slots of a pipeline (dedicated) are made that way, with every CU in the
// mask.  Such a stream synchronizes with the null stream, which the
// engine uses only when it grows a buffer.
static hipError_t create_stream(int device, bool dedicated, hipStream_t* s)
{
    if (!dedicated) return hipStreamCreateWithFlags(s, hipStreamNonBlocking);
    int ncu = 0;
    hipError_t e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                                         device);
    if (e != hipSuccess) return e;
    if (ncu <= 0) return hipErrorInvalidValue;
    std::vector<uint32_t> mask((size_t)(ncu + 31) / 32, 0u);
    for (int c = 0; c < ncu; c++) mask[(size_t)c / 32] |= 1u << (c % 32);
    return hipExtStreamCreateWithCUMask(s, (uint32_t)mask.size(), mask.data());
}

static int create_batch(int device, uint32_t flags, fk_batch** out)
{
    const bool dedicated = (flags & FK_BATCH_DEDICATED_QUEUE) != 0;
    fk_batch* b = new fk_batch();
    b->device = device;
    b->throughput = (flags & FK_BATCH_THROUGHPUT) != 0;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = create_stream(device, dedicated, &b->stream);
    for (int i = 0; i < 6 && e == hipSuccess; i++)
        e = hipEventCreate(&b->ev[i]);
    if (e != hipSuccess) {
        batch_free(b);
        delete b;
        return fail(FK_E_DEVICE, "batch setup failed: %s",
                    hipGetErrorString(e));
    }
    *out = b;
    return FK_OK;
}

// Expected survivors of the LDS stage for a window: wheel density times
// the product of (1 - |K_q|/q) over the LDS primes.
static double expected_lds(const fk_engine* e, double span)
{
    return span * e->line_density * e->surv_lds;
}

// Round boundaries by prime value: q_med * growth^r.
static double round_survival(const fk_engine* e, double lo, double hi)
{
    if (hi <= lo) return 1.0;
    // sum over primes in [lo, hi) of w/q ~ w ln(ln hi / ln lo), with about
    // n killing residues per prime in every family
    const double w = e->n;
    return std::pow(std::log(lo) / std::log(hi), w);
}

// records x chain cap at or below which classification runs every base in
// one pass
static constexpr uint32_t kSingleLimit = 8192;

// short-line sieve: chunk sizes up to kShortChunkMax words per lane (3, 6
// and 12 divide lines of 24 words, which a 2^32-j Dickson window has),
// chosen per window by instructions per lane: about kShortStartCost to form
// a chunk's start in a prime's coordinates (one fma and a double-precision
// residue), kShortWordCost per word.  Measured on the 256-window mixed
// batch: start costs 8 and 12 give the same time, 16 is 4% slower and the
// 28 of the 32-bit Barrett start 10% slower.
static constexpr uint32_t kShortChunks[] = {1, 2, 3, 4, 6, 8, 12, 16};
static constexpr uint64_t kShortStartCost = 12, kShortWordCost = 5;

// A window on wheel t with short lines: its base, the chunk with the least
// work per line, its (line, chunk) items and the whole window's work (lane
// instructions: per LDS prime a start per chunk and a bitmap window per
// word, padded words included).  False when t's lines are too long for the
// short-line sieve.
struct ShortPlan {
    u128 B;
    uint64_t H;
    uint32_t chunk;
    uint64_t items;
    double work;
};

static bool short_plan(const fk_engine* t, u128 lo, u128 hi, ShortPlan* sp)
{
    if (!t->short_words) return false;
    const u128 B = lo - lo % t->D;
    const uint64_t H = hi > B ? (uint64_t)(hi - B) : 0;
    const uint64_t npw = (H / t->D + 1 + 63) / 64;
    if (npw > t->short_words) return false;
    uint64_t best = ~0ull, nch_best = 0;
    uint32_t chunk = 0;
    for (uint32_t c : kShortChunks) {
        const uint64_t nch = (npw + c - 1) / c;
        const uint64_t cost = nch * (kShortStartCost + c * kShortWordCost);
        if (cost < best && t->nlines * nch < (1ull << 32)) {
            best = cost;
            chunk = c;
            nch_best = nch;
        }
    }
    if (!chunk) return false;
    sp->B = B;
    sp->H = H;
    sp->chunk = chunk;
    sp->items = t->nlines * nch_best;
    sp->work = (double)best * (double)t->nlines *
               (double)std::max<uint32_t>(t->host.nlds, 1u);
    return true;
}

// A prefix wheel takes a window only if it leaves the window at least this
// many (line, chunk) items (or as many as the full wheel): fewer lanes
// than two tiles would run the sieve as a few long chains.
static constexpr uint64_t kAltMinItems = 512;

// lane-per-candidate rounds: workgroups to aim for (candidate blocks times
// prime blocks).  A workgroup's set-up (records, flags, the rows' touch,
// the final ballot) costs about 100 VALU per wave, so fewer, longer prime
// blocks spend more of the VALU on tests, while more, shorter ones finish
// a lone window sooner.  A batch that runs beside others
// (FK_BATCH_THROUGHPUT, the slots of a pipeline) takes 1024: the Dickson
// sweep went from 0.152-0.164 to 0.139-0.152 ms per window with four in
// flight (512: 0.16, too few workgroups to fill the chip).  Alone, 4096:
// 1024 cost the Euler SCORE_WIDE range 7% and the one-submission mixed
// batch 20%.
static constexpr uint32_t kRoundWorkgroups = 4096;
static constexpr uint32_t kRoundWorkgroupsThroughput = 1024;
// ... and only for rounds that start with at least this many expected
// candidates: a round of few is latency-bound either way (this kept the
// Dickson sweep's rate and took the plain engine on the mixed windows of
// 2^20..2^30 values from 1.1-1.3e12 to 1.5-1.7e12 candidates/s)
static constexpr double kThroughputAlive = 2048.0;

// Late rounds of few candidates are launch-bound: a round and its
// compaction cost two launches (about 8 us alone) whatever they test.
// (Submissions of one engine only; see nused in batch_enqueue.)
// group_rounds joins a round to the one before it while the candidates the
// earlier rounds of the group kill, tested again against the joined
// round's primes, stay within kGroupExtraTests tests (about a microsecond
// of the chip's VALU).  Config 2's AP windows (candidates entering the
// rounds 372k, 65k, 13k, 3k, 795, 227, 62, 20, 9, 3, 1, 0, ...) keep 6 of
// their 15 rounds; the Dickson sweep's rounds kill more per round (every
// join would add 6-19M tests) and keep all 12.
static constexpr double kGroupExtraTests = 4.0e6;

// Lane = prime for AP index rows in rounds of few candidates and many
// primes (a joined tail: 3 candidates against a million primes leaves
// lane = candidate 3 lanes of 64); residue rows stay with lane = candidate,
// which tests their count-nested residues for fewer VALU.
static constexpr double kSparseAlive = 320.0;
static constexpr uint32_t kSparsePrimes = 8192;

// h's rounds grouped for a submission that expects `alive` LDS-stage
// survivors of engine e (see kGroupExtraTests).
static void group_rounds(const fk_engine* e, const DeviceCtx* ctx, double alive,
                         DevEngine& h)
{
    auto prime_at = [&](uint32_t i) {
        return i < ctx->primes.size() ? (double)ctx->primes[i] : (double)e->q2;
    };
    uint32_t rb[33];
    rb[0] = h.rb[0];
    uint32_t nr = 0;
    double a = alive, a0 = alive;          // entering round r, the group
    for (uint32_t r = 0; r < h.nround; r++) {
        const double np = (double)(h.rb[r + 1] - h.rb[r]);
        if (nr > 0 && (a0 - a) * np <= kGroupExtraTests) {
            rb[nr] = h.rb[r + 1];           // joins the group
        } else {
            rb[++nr] = h.rb[r + 1];         // starts one
            a0 = a;
        }
        a *= round_survival(e, prime_at(h.rb[r]), prime_at(h.rb[r + 1]));
    }
    for (uint32_t i = 0; i <= nr; i++) h.rb[i] = rb[i];
    for (uint32_t i = nr + 1; i < 33; i++) h.rb[i] = rb[nr];
    h.nround = nr;
}

// A submission whose windows expect fewer LDS-stage survivors than
// kMergeAlive in all sweeps with every other round boundary of each
// engine's table (round growth squared): its rounds are launch- and
// latency-bound, and half as many cost less than the tests they add.  (A
// submission of many candidates keeps every boundary: its launches serve
// all its engines at once.)  The records do not depend on the boundaries.
static constexpr double kMergeAlive = 4096.0;

static int batch_enqueue(fk_batch* b, const std::vector<uint32_t>* caps_in)
{
    const size_t nwin = b->windows.size();
    // the engine with the most candidates sets the tile threads, the
    // longest tile and the round-kernel crossover of the whole batch
    fk_engine* e0 = b->engines[0];
    {
        std::vector<double> work(b->engines.size(), 0.0);
        for (size_t w = 0; w < nwin; w++) {
            const fk_engine* e = b->engines[b->engine_of[w]];
            const u128 lo = mk128(b->windows[w].lo), hi = mk128(b->windows[w].hi);
            if (hi > lo) work[b->engine_of[w]] += (double)(hi - lo) * e->line_density;
        }
        size_t best = 0;
        for (size_t k = 1; k < work.size(); k++)
            if (work[k] > work[best]) best = k;
        e0 = b->engines[best];
    }
    DeviceCtx* ctx = nullptr;
    uint32_t need_limit = 0;
    for (fk_engine* e : b->engines) need_limit = std::max(need_limit, e->q2);
    int rc = ctx_for(b->device, need_limit, &ctx);
    if (rc) return rc;
    HIPCHK(hipSetDevice(b->device));

    // ---- pieces, grouped by tile shape: each window's tile is no longer
    // than its lines need (words_per_thread halves while half a tile still
    // holds a line), and each group of windows with one shape is one sieve
    // launch over a contiguous range of tiles
    const uint32_t T = e0->threads;
    // tab: the engine whose wheel and LDS tables sieve the window (the
    // window's own, or one of its prefix wheels)
    struct WinPlan { u128 B; uint64_t H; uint32_t wpt, chunk; const fk_engine* tab; };
    std::vector<WinPlan> wp(nwin);
    for (size_t w = 0; w < nwin; w++) {
        const fk_engine* e = b->engines[b->engine_of[w]];
        const u128 lo = mk128(b->windows[w].lo), hi = mk128(b->windows[w].hi);
        const u128 B = lo - lo % e->D;
        wp[w].B = B;
        wp[w].H = hi > B ? (uint64_t)(hi - B) : 0;
        wp[w].tab = e;
        const uint64_t npos = wp[w].H / e->D + 1;
        // lines of a few words: one lane per chunk of a line (k_sieve_short),
        // the chunk with the least work per line: a start per prime and
        // chunk, a bitmap window per prime and word (padded words included).
        // Lines of a word or less on a wide wheel are mostly padding (a
        // 2^20-j Dickson window has 0.4 positions on each of the 66,759
        // lines of its 2,800,733-wide wheel): a prefix wheel, fewer and
        // longer lines, may sieve the window with a small part of the work.
        ShortPlan sp;
        wp[w].chunk = 0;
        if (short_plan(e, lo, hi, &sp)) {
            const uint64_t min_items = std::min(kAltMinItems, sp.items);
            for (const fk_engine* a : e->alts) {
                ShortPlan ap;
                if (short_plan(a, lo, hi, &ap) && ap.items >= min_items &&
                    ap.work < sp.work) {
                    sp = ap;
                    wp[w].tab = a;
                }
            }
            wp[w].B = sp.B;
            wp[w].H = sp.H;
            wp[w].chunk = sp.chunk;
            wp[w].wpt = 0;
            continue;
        }
        // the tile length (words per thread up to the engine's) with the
        // least work per line: padded positions plus a per-tile charge for
        // the tile's fixed stages (2048 positions' worth)
        uint32_t wpt = 1;
        uint64_t best = ~0ull;
        for (uint32_t k = 1; k <= e0->wpt; k++) {
            if (!sieve_shape_ok((int)T, (int)k)) continue;
            const uint64_t S = (uint64_t)64 * T * k;
            const uint64_t nt = (npos + S - 1) / S;
            const uint64_t cost = nt * S + nt * 2048;
            if (cost < best || (cost == best && k > wpt)) {
                best = cost;
                wpt = k;
            }
        }
        wp[w].wpt = wpt;
    }
    std::vector<uint32_t> order(nwin);
    for (size_t w = 0; w < nwin; w++) order[w] = (uint32_t)w;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
        return wp[x].wpt != wp[y].wpt ? wp[x].wpt > wp[y].wpt
                                      : wp[x].chunk > wp[y].chunk;
    });
    struct Group { uint32_t wpt, chunk; uint64_t tile_begin, tile_end; };
    std::vector<Group> groups;

    // the engine array the kernels index: the submission's engines, then
    // the prefix wheels its windows use (sieve-side tables only; the
    // queues, counters and rounds are the submission's engines')
    std::vector<const fk_engine*> tabs(b->engines.begin(), b->engines.end());
    auto tab_index = [&](const fk_engine* t, uint32_t own) {
        if (t == b->engines[own]) return own;
        for (size_t i = b->engines.size(); i < tabs.size(); i++)
            if (tabs[i] == t) return (uint32_t)i;
        tabs.push_back(t);
        return (uint32_t)(tabs.size() - 1);
    };
    b->win_wheel.assign(nwin, 0);

    std::vector<DevPiece> pieces;
    std::vector<DevWinInfo> wins(nwin);
    std::vector<uint32_t> win_engine(nwin);
    uint64_t tiles = 0;
    uint32_t nlds_max = 0, pat_max = 0, bm_max = 0;
    uint32_t s_nlds_max = 0, s_bm_max = 0;      // over short windows
    double expect = 0.0, cand = 0.0;
    std::vector<double> e_alive(b->engines.size(), 0.0);  // expected per engine
    for (fk_engine* e : b->engines) {
        nlds_max = std::max(nlds_max, e->host.nlds);
        pat_max = std::max(pat_max, e->pat_words);
        bm_max = std::max(bm_max, e->host.bm_words);
    }
    for (size_t w = 0; w < nwin; w++) {
        win_engine[w] = b->engine_of[w];
        wins[w].engine = b->engine_of[w];
        wins[w].mode = b->cls[w].mode;
        wins[w].cap = b->cls[w].cap;
        wins[w].exact_from = b->cls[w].exact_from - b->cls[w].first;
        wins[w].first = b->cls[w].first;
    }
    for (uint32_t w : order) {
        const fk_engine* e = b->engines[b->engine_of[w]];
        const u128 lo = mk128(b->windows[w].lo), hi = mk128(b->windows[w].hi);
        if (hi <= lo) continue;
        b->win_wheel[w] = wp[w].tab->D;
        if (wp[w].wpt == 0) {
            // short lines: one piece, tiles of 256 (line, chunk) items, on
            // the window's wheel (its engine's or a prefix wheel)
            const fk_engine* t = wp[w].tab;
            if (groups.empty() || groups.back().wpt != 0 ||
                groups.back().chunk != wp[w].chunk)
                groups.push_back({0u, wp[w].chunk, tiles, tiles});
            const uint64_t npos_all = wp[w].H / t->D + 1;
            DevPiece P{};
            for (int i = 0; i < 4; i++) P.base[i] = (uint32_t)(wp[w].B >> (32 * i));
            P.lo_rel = lo > wp[w].B ? (uint64_t)(lo - wp[w].B) : 0;
            P.hi_rel = (uint64_t)(hi - wp[w].B);
            P.npos = (uint32_t)npos_all;
            P.npw = (uint32_t)((npos_all + 63) / 64);
            const uint64_t nch = (P.npw + wp[w].chunk - 1) / wp[w].chunk;
            P.nseg = (uint32_t)((t->nlines * nch + 255) / 256);
            P.tile0 = (uint32_t)tiles;
            P.engine = tab_index(t, b->engine_of[w]);
            P.queue = b->engine_of[w];
            P.window = (uint32_t)w;
            if (tiles + P.nseg >= (1ull << 32))
                return fail(FK_E_ARG, "batch too large: more than 2^32 tiles");
            tiles += P.nseg;
            pieces.push_back(P);
            groups.back().tile_end = tiles;
            s_nlds_max = std::max(s_nlds_max, t->host.nlds);
            s_bm_max = std::max(s_bm_max, t->host.sbm_words);
            const double span = (double)(hi - lo);
            expect += expected_lds(e, span);
            e_alive[b->engine_of[w]] += expected_lds(e, span);
            cand += span * e->line_density;
            continue;
        }
        const uint64_t S = (uint64_t)64 * T * wp[w].wpt;
        const uint64_t PIECE = std::max<uint64_t>(S, ((uint64_t)1 << 31) / S * S);
        if (groups.empty() || groups.back().wpt != wp[w].wpt)
            groups.push_back({wp[w].wpt, 0u, tiles, tiles});
        const uint64_t npos_all = wp[w].H / e->D + 1;
        for (uint64_t p0 = 0; p0 < npos_all; p0 += PIECE) {
            DevPiece P{};
            const u128 base = wp[w].B + (u128)e->D * p0;
            for (int i = 0; i < 4; i++) P.base[i] = (uint32_t)(base >> (32 * i));
            P.lo_rel = lo > base ? (uint64_t)(lo - base) : 0;
            P.hi_rel = (uint64_t)(hi - base);
            P.npos = (uint32_t)std::min<uint64_t>(PIECE, npos_all - p0);
            P.nseg = (uint32_t)((P.npos + S - 1) / S);
            P.tile0 = (uint32_t)tiles;
            P.engine = b->engine_of[w];
            P.queue = b->engine_of[w];
            P.window = (uint32_t)w;
            const uint64_t t = (uint64_t)P.nseg * e->nlines;
            if (tiles + t >= (1ull << 32))
                return fail(FK_E_ARG, "batch too large: more than 2^32 tiles");
            tiles += t;
            pieces.push_back(P);
        }
        groups.back().tile_end = tiles;
        const double span = (double)(hi - lo);
        expect += expected_lds(e, span);
        e_alive[b->engine_of[w]] += expected_lds(e, span);
        cand += span * e->line_density;
    }
    b->tiles = tiles;
    b->candidates = (uint64_t)cand;
    std::vector<char> used(b->engines.size(), 0);
    for (size_t w = 0; w < nwin; w++) used[b->engine_of[w]] = 1;

    // ---- queue segments: each engine's from its expected LDS-stage
    // survivors (or from the counts of a run that overflowed)
    const size_t ne = b->engines.size();
    std::vector<uint32_t> caps(ne, 0);
    std::vector<uint2> segs(ne);
    uint64_t qtot = 0;
    for (size_t k = 0; k < ne; k++) {
        if (used[k]) {
            double c = 1.5 * e_alive[k] + 8.0 * std::sqrt(e_alive[k]) + 1024.0;
            if (e_alive[k] > 0.9 * kSegmentMax)
                return fail(FK_E_RANGE, "engine %zu's windows in this batch "
                            "are expected to leave %.3g sieve survivors, "
                            "more than one queue segment holds (%u): split "
                            "them", k, e_alive[k], kSegmentMax);
            if (b->engines[k]->segment_cap) c = b->engines[k]->segment_cap;
            if (caps_in && k < caps_in->size() && (*caps_in)[k])
                c = (*caps_in)[k];
            caps[k] = (uint32_t)std::min(c, (double)kSegmentMax);
            caps[k] = std::max(256u, (caps[k] + 255u) & ~255u);
        }
        segs[k] = make_uint2((uint32_t)qtot, caps[k]);
        qtot += caps[k];
    }
    if (qtot >= 0xf0000000ull)
        return fail(FK_E_OVERFLOW, "survivor queues need %llu entries",
                    (unsigned long long)qtot);
    const uint32_t queue_cap = (uint32_t)std::max<uint64_t>(qtot, 256);
    b->queue_cap = queue_cap;
    b->seg_caps = caps;

    // ---- rounds: round r of every engine together (each engine keeps its
    // own q_med * growth^r boundaries, DevEngine::rb; an engine of few
    // expected candidates in this submission takes every other boundary)
    std::vector<DevEngine> eff(b->engines.size());
    uint32_t nround_max = 0;
    // the late-round grouping and the sparse lane = prime rounds serve
    // submissions of one engine (plain sweeps, one engine's requests):
    // with several, a round launch serves them all and its grid and kernel
    // follow the sum of their expected candidates
    const size_t nused = (size_t)std::count(used.begin(), used.end(), 1);
    double alive_all = 0.0;
    for (size_t k = 0; k < b->engines.size(); k++)
        if (used[k]) alive_all += e_alive[k];
    for (size_t k = 0; k < b->engines.size(); k++) {
        DevEngine& h = eff[k];
        h = b->engines[k]->host;
        if (used[k] && alive_all < kMergeAlive && h.nround > 1) {
            uint32_t nr = 0;
            for (uint32_t r = 0; r < h.nround; r += 2)
                h.rb[++nr] = h.rb[std::min(r + 2, h.nround)];
            for (uint32_t i = nr + 1; i < 33; i++) h.rb[i] = h.rb[nr];
            h.nround = nr;
        }
        if (used[k] && nused == 1 && h.nround > 1)
            group_rounds(b->engines[k], ctx, e_alive[k], h);
        if (used[k]) nround_max = std::max(nround_max, h.nround);
    }
    // counters: per stage (the sieve, then each round) one per engine,
    // then the packed total
    const size_t nstage = 1 + nround_max;
    const size_t ncounts = nstage * ne + 1;

    // ---- buffers
    if ((rc = grow_uploads(b, tabs.size(), std::max<size_t>(pieces.size(), 1),
                           nwin, ne, ncounts)))
        return rc;
    if (queue_cap > b->cap_q || !b->d_q[0]) {
        size_t c = 0;
        for (int i = 0; i < 2; i++) {
            c = 0;
            if (b->d_q[i]) (void)hipFree(b->d_q[i]);
            b->d_q[i] = nullptr;
            if ((rc = grow(&b->d_q[i], &c, queue_cap))) return rc;
        }
        size_t cf = 0;
        if (b->d_flags) (void)hipFree(b->d_flags);
        b->d_flags = nullptr;
        if ((rc = grow(&b->d_flags, &cf, c))) return rc;
        for (uint64_t** m : {&b->d_fail2, &b->d_failx}) {
            size_t cm = 0;
            if (*m) (void)hipFree(*m);
            *m = nullptr;
            if ((rc = grow(m, &cm, c))) return rc;
            // the classification passes leave these zero after every use
            HIPCHK(hipMemset(*m, 0, cm * sizeof(uint64_t)));
        }
        b->cap_q = c;
    }

    // ---- uploads (small): staged in pinned memory, one asynchronous copy.
    // The staging area is rewritten only by the next submission, which
    // starts after this one's work (and so this copy) has been waited for.
    hipStream_t s = b->stream;
    {
        uint8_t* h = b->h_up;
        DevEngine* he = reinterpret_cast<DevEngine*>(h + b->o_engines);
        for (size_t k = 0; k < b->engines.size(); k++) he[k] = eff[k];
        for (size_t k = b->engines.size(); k < tabs.size(); k++) he[k] = tabs[k]->host;
        if (!pieces.empty())
            std::memcpy(h + b->o_pieces, pieces.data(), pieces.size() * sizeof(DevPiece));
        std::memcpy(h + b->o_win_engine, win_engine.data(), nwin * sizeof(uint32_t));
        std::memcpy(h + b->o_wins, wins.data(), nwin * sizeof(DevWinInfo));
        std::memcpy(h + b->o_segs, segs.data(), ne * sizeof(uint2));
        // the stage counters start at zero: the copy clears them
        std::memset(h + b->o_counts, 0, ncounts * sizeof(uint32_t));
        const size_t used = b->o_counts + ncounts * sizeof(uint32_t);
        HIPCHK(hipMemcpyAsync(b->d_up, h, used, hipMemcpyHostToDevice, s));
    }

    std::vector<Step> plan;
    auto mark = [&](int i) {
        Step st = new_step(Step::kMark);
        st.ev = i;
        plan.push_back(st);
    };
    mark(0);
    // the upload zeroed the counters; the sieve clears each emitted
    // record's flag (the compactions leave the flags they read cleared)
    mark(1);

    // ---- sieve: one launch per tile shape
    for (const Group& g : groups) {
        if (g.tile_end == g.tile_begin) continue;
        const uint64_t S = (uint64_t)64 * T * g.wpt;
        Step st = new_step(Step::kSieve);
        SieveArgs& a = st.sa;
        if (g.wpt == 0) {
            // short lines: two per-prime arrays and the kill bitmaps
            a.engines = b->d_engines;
            a.pieces = b->d_pieces;
            a.npieces = (uint32_t)pieces.size();
            a.tile_begin = (uint32_t)g.tile_begin;
            a.ntiles = (uint32_t)g.tile_end;
            a.nlds_max = s_nlds_max;
            a.pat_max = 0;
            a.bm_max = s_bm_max;
            a.segs = b->d_segs;
            a.out = b->d_q[0];
            a.counts = b->d_counts;
            a.flags = b->d_flags;
            const size_t lds = short_lds(s_nlds_max, s_bm_max);
            if (lds > 160 * 1024)
                return fail(FK_E_ARG, "LDS need %zu bytes > 160 KiB (lower q_med"
                            " or q_pat)", lds);
            int occ = sieve_short_occupancy((int)g.chunk, lds);
            if (occ < 1) occ = 1;
            uint64_t grid = (uint64_t)occ * ctx->cus;
            grid = std::min<uint64_t>(grid, g.tile_end - g.tile_begin);
            st.threads = (int)g.chunk;      // short lines: the chunk
            st.wpt = 0;
            st.grid = (uint32_t)grid;
            st.lds = lds;
            plan.push_back(st);
            continue;
        }
        a.engines = b->d_engines;
        a.pieces = b->d_pieces;
        a.npieces = (uint32_t)pieces.size();
        a.tile_begin = (uint32_t)g.tile_begin;
        a.ntiles = (uint32_t)g.tile_end;
        a.nlds_max = nlds_max;
        a.pat_max = pat_max;
        a.bm_max = bm_max;
        a.segs = b->d_segs;
        a.out = b->d_q[0];
        a.counts = b->d_counts;               // stage 0
        a.flags = b->d_flags;
        // layout of k_sieve's dynamic LDS: bits, s0, per-prime info,
        // pattern words, kill bitmaps
        const size_t lds = S / 8 + 16 * (size_t)((nlds_max + 3) / 4) +
                           16 * (size_t)nlds_max + 8 * (size_t)pat_max +
                           4 * (size_t)((bm_max + 3) & ~3u) +
                           2 * std::min<size_t>(S / 16, 4096) *
                               (S <= 65536 ? 2 : 4);
        if (lds > 160 * 1024)
            return fail(FK_E_ARG, "LDS need %zu bytes > 160 KiB (lower q_med"
                        " or q_pat)", lds);
        int occ = sieve_occupancy((int)T, (int)g.wpt, lds);
        if (occ < 1) occ = 1;
        uint64_t grid = (uint64_t)occ * ctx->cus;
        grid = std::min<uint64_t>(grid, g.tile_end - g.tile_begin);
        st.threads = (int)T;
        st.wpt = (int)g.wpt;
        st.grid = (uint32_t)grid;
        st.lds = lds;
        plan.push_back(st);
    }
    // the sieve launches write disjoint records through atomic slots, so
    // they need no order: the one of most tiles runs on the batch's stream,
    // the others (issued first) beside it on the side stream
    {
        size_t first = plan.size();
        while (first > 0 && plan[first - 1].kind == Step::kSieve) first--;
        if (plan.size() - first >= 2) {
            size_t big = first;
            auto tiles = [](const Step& st) { return st.sa.ntiles - st.sa.tile_begin; };
            for (size_t i = first; i < plan.size(); i++)
                if (tiles(plan[i]) > tiles(plan[big])) big = i;
            std::rotate(plan.begin() + big, plan.begin() + big + 1, plan.end());
            for (size_t i = first; i + 1 < plan.size(); i++) plan[i].side = 1;
        }
    }
    mark(2);

    // ---- rounds
    uint32_t cur = 0, cstage = 0;
    b->round_slots.clear();
    double alive = expect;
    for (uint32_t r = 0; r < nround_max; r++) {
        // one test launch per row kind (residue rows, AP index rows) among
        // the engines with a round r: each kernel then compiles only its
        // own path, at its own register count
        bool any_round = false;
        double alive_r = 0.0;
        for (uint32_t kind : {kRowsResidue, kRowsIndex}) {
            // the primes of round r of every engine of this kind
            uint32_t P0 = 0xffffffffu, P1 = 0;
            uint32_t kg = 1;
            alive = 0.0;
            for (size_t k = 0; k < b->engines.size(); k++) {
                const DevEngine& h = eff[k];
                if (!used[k] || r >= h.nround || h.rb[r + 1] <= h.rb[r]) continue;
                if ((h.kidx ? kRowsIndex : kRowsResidue) != kind) continue;
                P0 = std::min(P0, h.rb[r]);
                P1 = std::max(P1, h.rb[r + 1]);
                alive += e_alive[k];
                // residue groups per row held in registers: three when
                // every engine's kill sets fit twelve residues (both rows
                // of a pass then load together), else five
                if (!h.kidx) kg = std::max(kg, (kill_max(h.family, h.n) + 3u) / 4u);
            }
            if (P0 >= P1) continue;
            any_round = true;
            alive_r += alive;
            Step rst = new_step(Step::kRound);
            RoundArgs& ra = rst.ra;
            ra.engines = b->d_engines;
            ra.nengines = (uint32_t)ne;
            ra.segs = b->d_segs;
            ra.primed = ctx->d_primed;
            ra.recs = b->d_q[cur];
            ra.counts = b->d_counts + cstage * ne;
            ra.flags = b->d_flags;
            ra.round = r;
            ra.p_begin = P0;
            ra.p_end = P1;
            ra.q_min = ctx->primes[P0];
            ra.l22 = ctx->primes[P1 - 1] < (1u << 24) ? 1u : 0u;
            ra.kgroups = kg <= 3 ? 3u : 5u;
            ra.rows = kind;
            const uint32_t np = P1 - P0;
            // Many primes and enough candidates: lane = prime (256 primes
            // per workgroup column, the queue cut into slices of about 256
            // candidates, so that no wave walks a long chain of
            // candidates).  Otherwise lane = candidate (candidate blocks
            // times prime blocks of at least 16, about kRoundWorkgroups
            // workgroups: a workgroup pulls its block's rows toward the
            // caches once, then its waves walk them with a cache latency
            // per pair).  Both read the true count on the device; grids
            // sized from an estimate that is off still cover every
            // candidate.
            const bool lane_prime =
                (nused == 1 && kind == kRowsIndex && np >= kSparsePrimes &&
                 alive <= kSparseAlive) ||
                (np >= e0->round_lane_prime && alive >= e0->round_lane_prime / 8.0);
            uint32_t gx, gy;
            if (lane_prime) {
                gx = (np + 255) / 256;
                gy = (uint32_t)std::min(std::ceil(alive / kRoundSlice),
                                        std::max(1.0, std::floor((1u << 22) / (double)gx)));
                gy = std::max(1u, std::min(gy, 65535u));
            } else {
                gx = (uint32_t)std::min(1024.0, std::max(1.0, std::ceil(alive / (256.0 * kRoundKC))));
                const uint32_t target =
                    b->throughput && alive >= kThroughputAlive
                        ? kRoundWorkgroupsThroughput : kRoundWorkgroups;
                gy = std::min((np + 15) / 16, (target + gx - 1) / gx);
                gy = std::max(1u, std::min(gy, 65535u));
                ra.pblock = (np + gy - 1) / gy;
                gy = (np + ra.pblock - 1) / ra.pblock;
            }
            rst.lane_prime = lane_prime ? 1 : 0;
            rst.grid = gx;
            rst.gy = gy;
            plan.push_back(rst);
        }
        if (!any_round) continue;
        // a round of both row kinds is two launches on disjoint segments:
        // the first goes to the side stream, beside the second
        if (plan.size() >= 2 && plan[plan.size() - 2].kind == Step::kRound &&
            plan[plan.size() - 2].ra.round == r &&
            plan.back().kind == Step::kRound && plan.back().ra.round == r)
            plan[plan.size() - 2].side = 1;
        alive = alive_r;
        Step cst = new_step(Step::kCompact);
        CompactArgs& ca = cst.ca;
        ca.nengines = (uint32_t)ne;
        ca.segs = b->d_segs;
        ca.in = b->d_q[cur];
        ca.in_counts = b->d_counts + cstage * ne;
        ca.flags = b->d_flags;
        ca.out = b->d_q[cur ^ 1];
        ca.out_counts = b->d_counts + (size_t)(r + 1) * ne;
        uint32_t cg = (uint32_t)std::min(1024.0, std::ceil(alive / 1024.0) + ne);
        cst.grid = cg;
        plan.push_back(cst);
        cur ^= 1;
        cstage = r + 1;
        b->round_slots.push_back(cstage);
        for (size_t k = 0; k < b->engines.size(); k++) {
            const fk_engine* e = b->engines[k];
            const DevEngine& h = eff[k];
            if (r >= h.nround) continue;
            const double lo = ctx->primes[h.rb[r]];
            const double hi = h.rb[r + 1] < ctx->primes.size()
                                  ? (double)ctx->primes[h.rb[r + 1]] : (double)e->q2;
            e_alive[k] *= round_survival(e, lo, hi);
        }
    }
    alive = 0.0;
    for (double x : e_alive) alive += x;
    // ---- pack the segments for the classification and the copy back.  A
    // single engine's segment already starts the queue and is dense after
    // the last compaction: its last counter is the total, no copy needed.
    uint32_t total_slot = (uint32_t)(nstage * ne);
    if (ne == 1) {
        total_slot = (uint32_t)(cstage * ne);
    } else {
        Step gst = new_step(Step::kGather);
        GatherArgs& ga = gst.ga;
        ga.nengines = (uint32_t)ne;
        ga.segs = b->d_segs;
        ga.in = b->d_q[cur];
        ga.counts = b->d_counts + cstage * ne;
        ga.out = b->d_q[cur ^ 1];
        ga.total = b->d_counts + nstage * ne;
        uint32_t gg = (uint32_t)std::min(1024.0, std::ceil(alive / 256.0) + ne);
        gst.grid = gg;
        plan.push_back(gst);
        cur ^= 1;
    }
    mark(3);

    // ---- classification
    bool any_cls = false;
    int cap_max = 1;
    for (size_t w = 0; w < nwin; w++) {
        any_cls = any_cls || b->cls[w].mode != 0;
        if (b->cls[w].mode) cap_max = std::max(cap_max, b->cls[w].cap);
    }
    if (any_cls) {
        Step kst = new_step(Step::kClassify);
        ClassifyArgs& ca = kst.cla;
        ca.engines = b->d_engines;
        ca.wins = b->d_wins;
        ca.recs = b->d_q[cur];
        ca.count = b->d_counts + total_slot;
        ca.cap = queue_cap;
        ca.cap_max = (uint32_t)cap_max;
        ca.fail2 = b->d_fail2;
        ca.failx = b->d_failx;
        // few candidates: every base at once, one latency instead of two
        // (the kernels decide from the true count)
        ca.single_limit = kSingleLimit;
        const double work = alive * cap_max * 7.0;
        uint32_t g = (uint32_t)std::min(16384.0,
                                        std::max(1024.0, std::ceil(work / 128.0) + 4));
        kst.grid = g;
        plan.push_back(kst);
    }
    mark(4);

    // ---- readback: counters and the first `copy` final records
    uint32_t copy = (uint32_t)std::min<double>(queue_cap,
                                               4.0 * alive + 4096.0);
    if (copy > b->cap_hrecs || !b->h_recs) {
        if (b->h_recs) (void)hipHostFree(b->h_recs);
        b->h_recs = nullptr;
        size_t c = std::max<size_t>(copy, b->cap_hrecs * 3 / 2);
        HIPCHK(hipHostMalloc(&b->h_recs, c * sizeof(DevRec),
                             hipHostMallocCoherent));
        b->cap_hrecs = c;
    }
    {
        // one kernel writes both into the pinned buffers: the counters and
        // as many of the `copy` leading records as the queue holds
        Step st = new_step(Step::kReadback);
        ReadbackArgs& r = st.rb;
        r.counts = b->d_counts;
        r.ncounts = (uint32_t)ncounts;
        r.counts_host = b->h_counts;
        r.recs = b->d_q[cur];
        r.count = b->d_counts + total_slot;
        r.cap = copy;
        r.recs_host = b->h_recs;
        plan.push_back(st);
    }
    mark(5);

    // ---- run: eagerly with stage events; in graph mode, as a graph once a
    // launch sequence repeats (captured on its second submission in a row,
    // replayed while it stays the same), and eagerly without stage events
    // while it has not repeated, so one-off sequences pay no capture
    bool need_side = false;
    for (const Step& st : plan) need_side = need_side || st.side;
    if (need_side && !b->side.stream) {
        HIPCHK(hipStreamCreateWithFlags(&b->side.stream, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&b->side.fork, hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&b->side.join, hipEventDisableTiming));
    }
    b->graph_ran = false;
    b->staged = false;
    if (b->use_graph) {
        const bool replay = b->graph_exec && same_plan(plan, b->graph_plan);
        const bool repeat = !replay && (b->graph_at_once ||
                                        same_plan(plan, b->prev_plan));
        if (repeat) {
            if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
            if (b->graph) (void)hipGraphDestroy(b->graph);
            b->graph_exec = nullptr;
            b->graph = nullptr;
            b->graph_plan.clear();
            HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            const hipError_t ce = emit_steps(plan, b->ev, false, s, b->side);
            hipGraph_t g = nullptr;
            const hipError_t ee = hipStreamEndCapture(s, &g);
            if (ce != hipSuccess || ee != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return fail(FK_E_DEVICE, "graph capture failed: %s",
                            hipGetErrorString(ce != hipSuccess ? ce : ee));
            }
            b->graph = g;
            HIPCHK(hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0));
            b->graph_plan = plan;
            b->graph_captures++;
        }
        HIPCHK(hipEventRecord(b->ev[0], s));
        if (replay || repeat) {
            HIPCHK(hipGraphLaunch(b->graph_exec, s));
            b->graph_launches++;
            b->graph_ran = true;
        } else {
            HIPCHK(emit_steps(plan, b->ev, false, s, b->side));
        }
        HIPCHK(hipEventRecord(b->ev[5], s));
        b->prev_plan = plan;
    } else {
        b->step_kinds.clear();
        for (const Step& st : plan)
            if (st.kind != Step::kMark) b->step_kinds.push_back(step_kind(st));
        while (b->step_ev.size() < 2 * b->step_kinds.size()) {
            hipEvent_t e = nullptr;
            HIPCHK(hipEventCreate(&e));
            b->step_ev.push_back(e);
        }
        HIPCHK(emit_steps(plan, b->ev, true, s, b->side, &b->step_ev));
        b->staged = true;
    }
    b->final_slot = total_slot;
    b->final_q = cur;
    b->copied = copy;
    b->pending = true;
    b->have_results = false;
    return FK_OK;
}

static int batch_collect(fk_batch* b)
{
    if (!b->pending) {
        if (b->have_results) return FK_OK;
        return fail(FK_E_ARG, "nothing submitted");
    }
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipEventSynchronize(b->ev[5]));
    b->pending = false;
    const size_t ne = b->seg_caps.size();
    // A segment the sieve overfilled (its counter passed its size; nothing
    // was written past it) is resized to the true count and the batch run
    // again.  The counts are exact, so one rerun suffices; a second
    // overflow means the counts cannot be trusted and is an error.
    for (int attempt = 0;; attempt++) {
        std::vector<uint32_t> caps = b->seg_caps;
        bool over = false;
        for (size_t k = 0; k < ne; k++) {
            const uint32_t c = b->h_counts[k];      // stage 0: the sieve
            if (c <= caps[k]) continue;
            if (c >= kSegmentMax)
                return fail(FK_E_RANGE, "engine %zu's windows left %u sieve "
                            "survivors, more than one queue segment holds: "
                            "split them", k, c);
            caps[k] = (uint32_t)std::min<double>(1.25 * c + 65536.0,
                                                 kSegmentMax);
            over = true;
        }
        if (!over) break;
        if (attempt >= 1)
            return fail(FK_E_OVERFLOW, "survivor queue overflowed again "
                        "after resizing to the measured counts");
        b->reruns++;
        int rc = batch_enqueue(b, &caps);
        if (rc) return rc;
        HIPCHK(hipEventSynchronize(b->ev[5]));
        b->pending = false;
    }
    const uint32_t n = b->h_counts[b->final_slot];
    auto stage_total = [&](uint32_t st) {
        uint64_t t = 0;
        for (size_t k = 0; k < ne; k++)
            t += std::min(b->h_counts[st * ne + k], b->seg_caps[k]);
        return (uint32_t)t;
    };
    b->stage_counts.assign(1, stage_total(0));
    for (uint32_t st : b->round_slots) b->stage_counts.push_back(stage_total(st));
    b->step_ms.clear();
    if (!b->staged) {
        // graph mode records no stage events: only the whole is timed
        for (int i = 0; i < 4; i++) b->ms[i] = 0.0f;
    } else {
        for (size_t k = 0; k < b->step_kinds.size(); k++) {
            float t = 0.0f;
            HIPCHK(hipEventElapsedTime(&t, b->step_ev[2 * k],
                                       b->step_ev[2 * k + 1]));
            b->step_ms.push_back(t);
        }
        HIPCHK(hipEventElapsedTime(&b->ms[0], b->ev[0], b->ev[1]));
        HIPCHK(hipEventElapsedTime(&b->ms[1], b->ev[1], b->ev[2]));
        HIPCHK(hipEventElapsedTime(&b->ms[2], b->ev[2], b->ev[3]));
        HIPCHK(hipEventElapsedTime(&b->ms[3], b->ev[3], b->ev[4]));
    }
    HIPCHK(hipEventElapsedTime(&b->ms[4], b->ev[0], b->ev[5]));
    std::vector<DevRec> extra;
    const DevRec* recs = b->h_recs;
    if (n > b->copied) {
        extra.resize(n);
        HIPCHK(hipMemcpy(extra.data(), b->d_q[b->final_q], (size_t)n * sizeof(DevRec),
                         hipMemcpyDeviceToHost));
        recs = extra.data();
    }
    const size_t nwin = b->windows.size();
    b->out.resize(n);
    b->counts.assign(nwin, 0);
    for (uint32_t i = 0; i < n; i++) {
        const DevRec& r = recs[i];
        u128 t = ((u128)r.t[3] << 96) | ((u128)r.t[2] << 64) |
                 ((u128)r.t[1] << 32) | r.t[0];
        const fk_engine* e = b->engines[b->engine_of[r.window]];
        if (e->family == FK_FAMILY_DICKSON) t /= e->W;
        fk_record& o = b->out[i];
        o.value[0] = (uint64_t)t;
        o.value[1] = (uint64_t)(t >> 64);
        o.run = r.run;
        o.window = r.window;
        b->counts[r.window]++;
    }
    std::sort(b->out.begin(), b->out.end(),
              [](const fk_record& x, const fk_record& y) {
                  if (x.window != y.window) return x.window < y.window;
                  if (x.value[1] != y.value[1]) return x.value[1] < y.value[1];
                  return x.value[0] < y.value[0];
              });
    for (size_t w = 0; w < nwin && w < b->cursor.size(); w++)
        b->cursor[w] = mk128(b->windows[w].hi);
    b->have_results = true;
    return FK_OK;
}

// The next submission of the job: every window from its cursor to its end,
// cut so that the submission sweeps at most `budget` line values, taken
// window by window in order.
static int batch_start(fk_batch* b)
{
    const size_t nwin = b->job_hi.size();
    u128 left = b->budget ? b->budget : ~(u128)0;
    u128 total = 0;
    for (size_t w = 0; w < nwin; w++) {
        const u128 lo = b->cursor[w], hi = b->job_hi[w];
        const u128 take = hi > lo ? std::min(hi - lo, left) : 0;
        left -= take;
        total += take;
        fk_window& win = b->windows[w];
        win.lo[0] = (uint64_t)lo;
        win.lo[1] = (uint64_t)(lo >> 64);
        win.hi[0] = (uint64_t)(lo + take);
        win.hi[1] = (uint64_t)((lo + take) >> 64);
    }
    if (total == 0) {
        // nothing left to sweep: an empty result, nothing on the device
        b->out.clear();
        b->win_wheel.assign(nwin, 0);
        b->counts.assign(nwin, 0);
        b->stage_counts.clear();
        b->have_results = true;
        b->pending = false;
        return FK_OK;
    }
    return batch_enqueue(b, nullptr);
}

static int batch_submit(fk_batch* b, fk_engine* const* engines,
                        size_t nengines, const uint32_t* engine_of,
                        const fk_window* windows, const fk_classify_spec* cls,
                        size_t nwin)
{
    if (b->pending) return fail(FK_E_BUSY, "batch has work in flight");
    if (nengines == 0 || nwin == 0)
        return fail(FK_E_ARG, "empty batch");
    for (size_t i = 0; i < nengines; i++) {
        if (!engines[i]) return fail(FK_E_ARG, "null engine %zu", i);
        if (engines[i]->device != b->device)
            return fail(FK_E_ARG, "engine %zu lives on another device", i);
    }
    b->engines.assign(engines, engines + nengines);
    b->engine_of.assign(engine_of, engine_of + nwin);
    b->windows.assign(windows, windows + nwin);
    b->cls.resize(nwin);
    for (size_t w = 0; w < nwin; w++) {
        if (b->engine_of[w] >= nengines)
            return fail(FK_E_ARG, "window %zu names engine %u", w,
                        b->engine_of[w]);
        const fk_engine* e = engines[b->engine_of[w]];
        const u128 lo = mk128(windows[w].lo), hi = mk128(windows[w].hi);
        if (hi < lo) return fail(FK_E_RANGE, "window %zu has hi < lo", w);
        if ((hi >> 126) != 0)
            return fail(FK_E_RANGE, "window %zu reaches past 2^126", w);
        const u128 B = lo - lo % e->D;
        if (((hi - B) >> 63) != 0)
            return fail(FK_E_RANGE, "window %zu is wider than 2^63", w);
        u128 top = hi;
        if (e->family == FK_FAMILY_DICKSON) top = hi * e->W;
        if ((top >> 127) != 0)
            return fail(FK_E_RANGE, "window %zu values exceed 127 bits", w);
        fk_classify_spec c{0, 0, 0, 0};
        if (cls) c = cls[w];
        if (c.mode && (c.first < 0 || c.first > (1 << 20)))
            return fail(FK_E_ARG, "chain index %d outside [0, 2^20]", c.first);
        if (c.mode < 0 || c.mode > 2)
            return fail(FK_E_ARG, "classification mode %d", c.mode);
        if (c.mode && (c.cap < 0 || c.cap > 64))
            return fail(FK_E_ARG, "classification cap %d outside [0, 64]",
                        c.cap);
        b->cls[w] = c;
    }
    b->job_hi.resize(nwin);
    b->cursor.resize(nwin);
    for (size_t w = 0; w < nwin; w++) {
        b->cursor[w] = mk128(windows[w].lo);
        b->job_hi[w] = std::max(mk128(windows[w].hi), b->cursor[w]);
    }
    return batch_start(b);
}

// ------------------------------------------------------------------ ABI

extern "C" {

int fk_abi_version(void) { return FK_ABI_VERSION; }

int fk_last_error(char* buf, size_t len)
{
    if (!buf || len == 0) return FK_E_ARG;
    snprintf(buf, len, "%s", g_err.c_str());
    return FK_OK;
}

int fk_device_count(int* count)
{
    if (!count) return fail(FK_E_ARG, "null count");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return FK_OK;
}

int fk_engine_create(const fk_engine_desc* desc, fk_engine** out)
{
    if (!desc || !out) return fail(FK_E_ARG, "null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(FK_E_NODEV, "no HIP device visible");
    if (desc->device < 0 || desc->device >= ndev)
        return fail(FK_E_NODEV, "device %d not present (%d visible)",
                    desc->device, ndev);
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, desc->device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(FK_E_NODEV, "device %d is %s, this library is built for "
                    "gfx950", desc->device, prop.gcnArchName);
    fk_engine* e = new fk_engine();
    e->desc = *desc;
    int rc = build_engine(e);
    if (rc == FK_OK) rc = build_alts(e);
    if (rc == FK_OK) rc = create_batch(e->device, 0u, &e->own);
    // sweeps of one window shape repeat one launch sequence: the engine's
    // own batch replays it as a graph (fk_engine_batch + fk_batch_set_graph
    // turn this off, for stage timings)
    if (rc == FK_OK) e->own->use_graph = true;
    if (rc) {
        std::string keep = g_err;
        engine_free(e);
        delete e;
        g_err = keep;
        return rc;
    }
    *out = e;
    return FK_OK;
}

int fk_engine_destroy(fk_engine* eng)
{
    if (!eng) return FK_OK;
    if (eng->own) {
        batch_free(eng->own);
        delete eng->own;
    }
    engine_free(eng);
    delete eng;
    return FK_OK;
}

int fk_engine_get_info(const fk_engine* e, fk_engine_info* info)
{
    if (!e || !info) return fail(FK_E_ARG, "null argument");
    std::memset(info, 0, sizeof *info);
    info->family = e->family;
    info->n = e->n;
    info->q2 = e->q2;
    info->nwheel = (uint32_t)e->wheel.size();
    for (size_t i = 0; i < e->wheel.size() && i < 16; i++)
        info->wheel_primes[i] = e->wheel[i];
    info->line_step = e->D;
    info->nlines = e->nlines;
    info->W = e->W;
    info->d[0] = (uint64_t)e->dval;
    info->d[1] = (uint64_t)(e->dval >> 64);
    info->q_pat = e->q_pat;
    info->q_mark = e->q_mark;
    info->q_med = e->q_med;
    info->threads = e->threads;
    info->words_per_thread = e->wpt;
    info->n_lds_primes = e->host.nlds;
    info->n_pattern_primes = e->npat;
    info->n_round_primes = e->round_end - e->round_first;
    info->n_rounds = e->nrounds;
    info->short_words = e->short_words;
    info->device_bytes = e->device_bytes;
    info->survival_lds = e->surv_lds;
    info->mark_density = e->mark_density;
    return FK_OK;
}

int fk_engine_sweep(fk_engine* eng, const fk_window* win,
                    const fk_classify_spec* cls, fk_record* out,
                    size_t out_cap, size_t* n_out)
{
    if (!eng || !win || !n_out) return fail(FK_E_ARG, "null argument");
    fk_batch* b = eng->own;
    uint32_t zero = 0;
    fk_engine* es[1] = {eng};
    int rc = batch_submit(b, es, 1, &zero, win, cls, 1);
    if (rc) return rc;
    rc = batch_collect(b);
    if (rc) return rc;
    *n_out = b->out.size();
    if (b->out.size() > out_cap)
        return fail(FK_E_OVERFLOW, "%zu survivors, room for %zu",
                    b->out.size(), out_cap);
    if (!b->out.empty())
        std::memcpy(out, b->out.data(), b->out.size() * sizeof(fk_record));
    return FK_OK;
}

}  // extern "C"

// Device buffers of one classification pass: the records, the window
// record, the count and the two failure masks per record.
struct ClsBuffers {
    DevRec* recs;
    DevWinInfo* win;
    uint32_t* cnt;
    uint64_t* fail;              // 2 n words
};

// One classification pass on the device in the given buffers: chain values
// win.first .. win.first + win.cap - 1 of each value, the run within the
// pass into runs.  Allocates no device memory.
static int classify_in(fk_engine* e, const std::vector<u128>& vals,
                       const DevWinInfo& win, std::vector<int32_t>& runs,
                       const ClsBuffers& B)
{
    const size_t n = vals.size();
    std::vector<DevRec> recs(n);
    for (size_t i = 0; i < n; i++) {
        for (int k = 0; k < 4; k++) recs[i].t[k] = (uint32_t)(vals[i] >> (32 * k));
        recs[i].window = 0;
        recs[i].run = -1;
    }
    const uint32_t cnt = (uint32_t)n;
    HIPCHK(hipSetDevice(e->device));
    hipStream_t s = e->own->stream;
    hipError_t err = launch_zero_u64(B.fail, 2 * n, s);
    if (err == hipSuccess)
        err = hipMemcpyAsync(B.recs, recs.data(), n * sizeof(DevRec),
                             hipMemcpyHostToDevice, s);
    if (err == hipSuccess)
        err = hipMemcpyAsync(B.win, &win, sizeof win, hipMemcpyHostToDevice, s);
    if (err == hipSuccess)
        err = hipMemcpyAsync(B.cnt, &cnt, sizeof cnt, hipMemcpyHostToDevice, s);
    if (err == hipSuccess) {
        ClassifyArgs ca{};
        ca.engines = e->d_engine;
        ca.wins = B.win;
        ca.recs = B.recs;
        ca.count = B.cnt;
        ca.cap = cnt;
        ca.cap_max = (uint32_t)std::max(1, win.cap);
        ca.fail2 = B.fail;
        ca.failx = B.fail + n;
        ca.single_limit = kSingleLimit;
        const uint64_t work = n * (uint64_t)ca.cap_max *
                              (n * (uint64_t)ca.cap_max <= kSingleLimit ? 7 : 1);
        uint32_t g = (uint32_t)std::min<uint64_t>(16384, (work + 127) / 128 + 4);
        err = launch_classify(ca, g, s);
    }
    if (err == hipSuccess)
        err = hipMemcpyAsync(recs.data(), B.recs, n * sizeof(DevRec),
                             hipMemcpyDeviceToHost, s);
    if (err == hipSuccess) err = hipStreamSynchronize(s);
    if (err != hipSuccess)
        return fail(FK_E_DEVICE, "classification failed: %s",
                    hipGetErrorString(err));
    runs.resize(n);
    for (size_t i = 0; i < n; i++) runs[i] = recs[i].run;
    return FK_OK;
}

// The same pass in device buffers allocated for it.
static int classify_pass(fk_engine* e, const std::vector<u128>& vals,
                         const DevWinInfo& win, std::vector<int32_t>& runs)
{
    const size_t n = vals.size();
    HIPCHK(hipSetDevice(e->device));
    ClsBuffers B{nullptr, nullptr, nullptr, nullptr};
    hipError_t err = hipMalloc(&B.recs, n * sizeof(DevRec));
    if (err == hipSuccess) err = hipMalloc(&B.win, sizeof(DevWinInfo));
    if (err == hipSuccess) err = hipMalloc(&B.cnt, sizeof(uint32_t));
    if (err == hipSuccess) err = hipMalloc(&B.fail, 2 * n * sizeof(uint64_t));
    int rc = err == hipSuccess
                 ? classify_in(e, vals, win, runs, B)
                 : fail(err == hipErrorOutOfMemory ? FK_E_OOM : FK_E_DEVICE,
                        "classification buffers: %s", hipGetErrorString(err));
    if (B.recs) (void)hipFree(B.recs);
    if (B.win) (void)hipFree(B.win);
    if (B.cnt) (void)hipFree(B.cnt);
    if (B.fail) (void)hipFree(B.fail);
    return rc;
}

// Chain values one pass covers (the per-record failure masks are 64 bits).
static constexpr int kPassValues = 64;

// The chain classification behind fk_engine_classify and
// fk_engine_classify_static: pass k looks at chain values first + 64k ..
// first + 64k + 63, for the values whose chain ran through the whole of
// pass k - 1.  `pass` runs one pass on the device.
template <class Pass>
static int classify_chains(const uint64_t* values, size_t n,
                           const fk_classify_spec* cls, int32_t* runs,
                           Pass pass)
{
    if ((!values && n) || !cls || (!runs && n))
        return fail(FK_E_ARG, "null argument");
    if (n == 0) return FK_OK;
    if (cls->mode < 1 || cls->mode > 2 || cls->cap < 0 ||
        cls->cap > FK_CLASSIFY_CAP_MAX)
        return fail(FK_E_ARG, "bad classification spec (mode %d, cap %d; "
                    "cap at most %d)", cls->mode, cls->cap,
                    FK_CLASSIFY_CAP_MAX);
    if (cls->first < 0 || cls->first > (1 << 20))
        return fail(FK_E_ARG, "chain index %d outside [0, 2^20]", cls->first);
    std::vector<size_t> live(n);
    for (size_t i = 0; i < n; i++) {
        live[i] = i;
        runs[i] = 0;
    }
    for (int done = 0; done < cls->cap && !live.empty(); done += kPassValues) {
        const int pcap = std::min(kPassValues, cls->cap - done);
        const int first = cls->first + done;
        DevWinInfo win{0, cls->mode, pcap, cls->exact_from - first, first};
        std::vector<u128> vals(live.size());
        for (size_t j = 0; j < live.size(); j++)
            vals[j] = ((u128)values[2 * live[j] + 1] << 64) | values[2 * live[j]];
        std::vector<int32_t> got;
        int rc = pass(vals, win, got);
        if (rc) return rc;
        std::vector<size_t> on;
        for (size_t j = 0; j < live.size(); j++) {
            runs[live[j]] = done + got[j];
            if (got[j] == pcap) on.push_back(live[j]);
        }
        live.swap(on);
    }
    return FK_OK;
}

extern "C" {

int fk_engine_classify(fk_engine* e, const uint64_t* values, size_t n,
                       const fk_classify_spec* cls, int32_t* runs)
{
    if (!e) return fail(FK_E_ARG, "null argument");
    if (n > (1u << 26)) return fail(FK_E_ARG, "at most 2^26 values per call");
    return classify_chains(values, n, cls, runs,
                           [&](const std::vector<u128>& v, const DevWinInfo& w,
                               std::vector<int32_t>& r) {
                               return classify_pass(e, v, w, r);
                           });
}

}  // extern "C"

// ------------------------------------------- classification, no allocation
//
// fk_engine_classify_static runs the same passes in a fixed device array:
// per value a record and two failure masks, plus the window record and the
// count.  Each device holds its own copy of the array for the life of the
// process; host code reaches it only through the address looked up on that
// device (the array's name is not a device address on the host).

static constexpr size_t kStaticValues = FK_CLASSIFY_STATIC_MAX;
static constexpr size_t kStaticHead = 256;       // DevWinInfo, then the count
static constexpr size_t kStaticRecs = kStaticValues * sizeof(DevRec);
static constexpr size_t kStaticBytes =
    kStaticHead + kStaticRecs + 2 * kStaticValues * sizeof(uint64_t);
static_assert(kStaticRecs % 256 == 0, "the masks start 256-byte aligned");

static __device__ __align__(256) uint8_t scratch[kStaticBytes];

static std::mutex g_scratch_mu;
static std::vector<std::pair<int, uint8_t*>> g_scratch;   // device, address

// The scratch array's address on `device` (which must be current): looked
// up on the first call per device, kept after.
static int scratch_for(int device, uint8_t** out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    for (const auto& d : g_scratch)
        if (d.first == device) {
            *out = d.second;
            return FK_OK;
        }
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(FK_E_DEVICE, "scratch array address on device %d: %s",
                    device, hipGetErrorString(e));
    g_scratch.emplace_back(device, static_cast<uint8_t*>(ptr));
    *out = static_cast<uint8_t*>(ptr);
    return FK_OK;
}

extern "C" {

int fk_engine_classify_static(fk_engine* e, const uint64_t* values, size_t n,
                              const fk_classify_spec* cls, int32_t* runs)
{
    if (!e) return fail(FK_E_ARG, "null argument");
    if (n > kStaticValues)
        return fail(FK_E_ARG, "%zu values: the fixed classification array "
                    "holds %zu per call", n, kStaticValues);
    if (n == 0)          // argument checks only: no pass runs
        return classify_chains(values, n, cls, runs,
                               [](const std::vector<u128>&, const DevWinInfo&,
                                  std::vector<int32_t>&) { return FK_OK; });
    HIPCHK(hipSetDevice(e->device));
    uint8_t* base = nullptr;
    int rc = scratch_for(e->device, &base);
    if (rc) return rc;
    ClsBuffers B;
    B.win = reinterpret_cast<DevWinInfo*>(base);
    B.cnt = reinterpret_cast<uint32_t*>(base + 64);
    B.recs = reinterpret_cast<DevRec*>(base + kStaticHead);
    B.fail = reinterpret_cast<uint64_t*>(base + kStaticHead + kStaticRecs);
    return classify_chains(values, n, cls, runs,
                           [&](const std::vector<u128>& v, const DevWinInfo& w,
                               std::vector<int32_t>& r) {
                               return classify_in(e, v, w, r, B);
                           });
}

int fk_batch_create(int device, fk_batch** out)
{
    return fk_batch_create_ex(device, 0u, out);
}

int fk_batch_create_ex(int device, uint32_t flags, fk_batch** out)
{
    if (!out) return fail(FK_E_ARG, "null argument");
    if (flags & ~(uint32_t)(FK_BATCH_DEDICATED_QUEUE | FK_BATCH_THROUGHPUT))
        return fail(FK_E_ARG, "unknown batch flags 0x%x", flags);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return fail(FK_E_NODEV, "device %d not present", device);
    return create_batch(device, flags, out);
}

int fk_batch_destroy(fk_batch* b)
{
    if (!b) return FK_OK;
    if (b->pending) (void)hipEventSynchronize(b->ev[5]);
    batch_free(b);
    delete b;
    return FK_OK;
}

int fk_batch_submit(fk_batch* b, fk_engine* const* engines, size_t nengines,
                    const uint32_t* engine_of, const fk_window* windows,
                    const fk_classify_spec* cls, size_t nwin)
{
    if (!b || !engines || !engine_of || !windows)
        return fail(FK_E_ARG, "null argument");
    return batch_submit(b, engines, nengines, engine_of, windows, cls, nwin);
}

int fk_batch_query(fk_batch* b, int* done)
{
    if (!b || !done) return fail(FK_E_ARG, "null argument");
    if (!b->pending) { *done = 1; return FK_OK; }
    hipError_t e = hipEventQuery(b->ev[5]);
    if (e == hipSuccess) *done = 1;
    else if (e == hipErrorNotReady) *done = 0;
    else return fail(FK_E_DEVICE, "event query: %s", hipGetErrorString(e));
    return FK_OK;
}

int fk_batch_results(fk_batch* b, fk_record* out, size_t out_cap,
                     size_t* n_out, uint64_t* counts)
{
    if (!b || !n_out) return fail(FK_E_ARG, "null argument");
    int rc = batch_collect(b);
    if (rc) return rc;
    *n_out = b->out.size();
    if (counts)
        std::memcpy(counts, b->counts.data(), b->counts.size() * sizeof(uint64_t));
    if (b->out.size() > out_cap)
        return fail(FK_E_OVERFLOW, "%zu records, room for %zu", b->out.size(),
                    out_cap);
    if (!b->out.empty())
        std::memcpy(out, b->out.data(), b->out.size() * sizeof(fk_record));
    return FK_OK;
}

int fk_batch_timing(fk_batch* b, float* ms5, uint64_t* tiles,
                    uint64_t* candidates)
{
    if (!b) return fail(FK_E_ARG, "null argument");
    if (ms5) std::memcpy(ms5, b->ms, sizeof b->ms);
    if (tiles) *tiles = b->tiles;
    if (candidates) *candidates = b->candidates;
    return FK_OK;
}

int fk_batch_set_graph(fk_batch* b, int enable)
{
    if (!b) return fail(FK_E_ARG, "null argument");
    if (b->pending) return fail(FK_E_BUSY, "batch has work in flight");
    b->use_graph = enable != 0;
    b->graph_at_once = enable == 2;
    return FK_OK;
}

int fk_batch_graph_stats(fk_batch* b, uint64_t* captures, uint64_t* launches)
{
    if (!b) return fail(FK_E_ARG, "null argument");
    if (captures) *captures = b->graph_captures;
    if (launches) *launches = b->graph_launches;
    return FK_OK;
}

int fk_batch_counts(fk_batch* b, uint32_t* counts, size_t cap, size_t* n)
{
    if (!b || !n) return fail(FK_E_ARG, "null argument");
    *n = b->stage_counts.size();
    if (counts)
        std::memcpy(counts, b->stage_counts.data(),
                    std::min(cap, b->stage_counts.size()) * sizeof(uint32_t));
    return cap < b->stage_counts.size() && counts
               ? fail(FK_E_OVERFLOW, "%zu stage counts, room for %zu",
                      b->stage_counts.size(), cap)
               : FK_OK;
}

int fk_batch_step_times(fk_batch* b, int32_t* kinds, float* ms, size_t cap,
                        size_t* n)
{
    if (!b || !n) return fail(FK_E_ARG, "null argument");
    if (b->pending) return fail(FK_E_BUSY, "batch has work in flight");
    *n = b->step_ms.size();
    const size_t m = std::min(cap, b->step_ms.size());
    for (size_t k = 0; k < m; k++) {
        if (kinds) kinds[k] = b->step_kinds[k];
        if (ms) ms[k] = b->step_ms[k];
    }
    return cap < b->step_ms.size() && (kinds || ms)
               ? fail(FK_E_OVERFLOW, "%zu launches, room for %zu",
                      b->step_ms.size(), cap)
               : FK_OK;
}

int fk_probe_peaks(int device, double* valu_f64_per_s, double* valu_int_per_s,
                   double* lds_bytes_per_s)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(FK_E_NODEV, "no HIP device visible");
    if (device < 0 || device >= ndev)
        return fail(FK_E_NODEV, "device %d not present (%d visible)", device,
                    ndev);
    HIPCHK(hipSetDevice(device));
    int ncu = 0;
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount,
                                 device));
    hipStream_t s = nullptr;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    double f = 0.0, i = 0.0, l = 0.0;
    const hipError_t err = probe_peaks(ncu, s, &f, &i, &l);
    (void)hipStreamDestroy(s);
    if (err != hipSuccess)
        return fail(FK_E_DEVICE, "peak probes: %s", hipGetErrorString(err));
    if (valu_f64_per_s) *valu_f64_per_s = f;
    if (valu_int_per_s) *valu_int_per_s = i;
    if (lds_bytes_per_s) *lds_bytes_per_s = l;
    return FK_OK;
}

int fk_batch_window_wheel(fk_batch* b, size_t window, uint64_t* line_step)
{
    if (!b || !line_step) return fail(FK_E_ARG, "null argument");
    if (b->pending) return fail(FK_E_BUSY, "batch has work in flight");
    if (window >= b->windows.size())
        return fail(FK_E_ARG, "window %zu of %zu", window, b->windows.size());
    *line_step = window < b->win_wheel.size() ? b->win_wheel[window] : 0;
    return FK_OK;
}

int fk_batch_reruns(fk_batch* b, uint64_t* reruns)
{
    if (!b || !reruns) return fail(FK_E_ARG, "null argument");
    *reruns = b->reruns;
    return FK_OK;
}

int fk_batch_set_budget(fk_batch* b, const uint64_t values[2])
{
    if (!b) return fail(FK_E_ARG, "null argument");
    if (b->pending) return fail(FK_E_BUSY, "batch has work in flight");
    b->budget = values ? mk128(values) : 0;
    return FK_OK;
}

int fk_batch_cursor_get(fk_batch* b, size_t window, uint64_t next[2])
{
    if (!b || !next) return fail(FK_E_ARG, "null argument");
    if (b->pending) return fail(FK_E_BUSY, "batch has work in flight");
    if (window >= b->cursor.size())
        return fail(FK_E_ARG, "window %zu of a job of %zu", window,
                    b->cursor.size());
    next[0] = (uint64_t)b->cursor[window];
    next[1] = (uint64_t)(b->cursor[window] >> 64);
    return FK_OK;
}

int fk_batch_cursor_set(fk_batch* b, size_t window, const uint64_t next[2])
{
    if (!b || !next) return fail(FK_E_ARG, "null argument");
    if (b->pending) return fail(FK_E_BUSY, "batch has work in flight");
    if (window >= b->cursor.size())
        return fail(FK_E_ARG, "window %zu of a job of %zu", window,
                    b->cursor.size());
    const u128 v = mk128(next);
    if (v > b->job_hi[window])
        return fail(FK_E_RANGE, "cursor past the end of window %zu", window);
    b->cursor[window] = v;
    return FK_OK;
}

int fk_batch_resume(fk_batch* b)
{
    if (!b) return fail(FK_E_ARG, "null argument");
    if (b->pending) return fail(FK_E_BUSY, "batch has work in flight");
    if (b->job_hi.empty()) return fail(FK_E_ARG, "nothing submitted");
    return batch_start(b);
}

int fk_engine_batch(fk_engine* eng, fk_batch** out)
{
    if (!eng || !out) return fail(FK_E_ARG, "null argument");
    *out = eng->own;
    return FK_OK;
}

}  // extern "C"
