// fk_kernels.hip -- the device side of the sieve-and-test engine (gfx950).
//
// Pipeline for a batch of window pieces (see fk_device.h for the line
// model):
//   k_prep          B mod q for every (piece, LDS prime)
//   k_sieve         one LDS tile (S line positions of one wheel line) per
//                   workgroup iteration: pattern words for the smallest
//                   primes, ds_or stride marking for the medium primes,
//                   then popcount/prefix compaction of the surviving bits
//                   into the candidate queue
//   k_round_test    per-candidate kill tests for one range of large primes
//                   (lane = prime, candidates staged through LDS), killed
//                   candidates flagged
//   k_round_compact wave-ballot compaction of the unflagged candidates
//   k_cls_*         strong-probable-prime chains (Montgomery, 2..8 limbs),
//                   parallel over (candidate, value, base)
//
// Reference counterparts: euler-prime-runs/euler_gpu.py:438-575 and :108-273
// (bit-sieve, rounds, cold stage), dickson-ladders/ladder_gpu.py:254-407
// (sieve1a, sieve1b, sieve1b_warp), primorial-ap/ap_gpu.py:155-392
// (build_tab, mark_shared, mark, compact); the strong test restates
// huntlib/primes.py:23-80.  None of them is translated: the tile layout,
// the shared line model and the on-device classification are this build's.
#include <hip/hip_runtime.h>

#include "fk_device.h"
#include "fk_launch.h"

namespace fk {

// ------------------------------------------------------------------ prep
__global__ void k_prep(const DevEngine* __restrict__ engines,
                       const DevPiece* __restrict__ pieces, uint32_t npieces,
                       uint32_t* __restrict__ bq)
{
    for (uint32_t pi = blockIdx.y; pi < npieces; pi += gridDim.y) {
        const DevPiece P = pieces[pi];
        const DevEngine& E = engines[P.engine];
        for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < E.nlds;
             p += gridDim.x * blockDim.x) {
            const uint32_t q = E.lds_q[p];
            const uint64_t m64 = E.lds_m64[p];
            uint32_t r = red64(P.base[3], q, m64);
            r = red64(((uint64_t)r << 32) | P.base[2], q, m64);
            r = red64(((uint64_t)r << 32) | P.base[1], q, m64);
            r = red64(((uint64_t)r << 32) | P.base[0], q, m64);
            bq[P.bq_off + p] = r;
        }
    }
}

// ----------------------------------------------------------------- sieve

__device__ __forceinline__ uint32_t find_piece(const DevPiece* pieces,
                                               uint32_t npieces,
                                               uint32_t tile)
{
    uint32_t lo = 0, hi = npieces;          // last piece with tile0 <= tile
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (pieces[mid].tile0 <= tile) lo = mid; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ uint64_t ceil_div_u64(uint64_t a, uint64_t b)
{
    return a / b + (a % b != 0);
}

// Dead-bit mask for the 64 positions [pos0, pos0 + 64) of a line whose
// valid positions are [vlo, vend).
__device__ __forceinline__ uint64_t edge_mask(uint64_t pos0, uint64_t vlo,
                                              uint64_t vend)
{
    uint64_t m = 0;
    if (vlo > pos0) {
        uint64_t c = vlo - pos0;
        m |= (c >= 64) ? ~0ull : ((1ull << c) - 1);
    }
    uint64_t hc = (vend > pos0) ? (vend - pos0) : 0;
    if (hc < 64) m |= ~0ull << hc;
    return m;
}

// BYTES = 0: medium primes set bits with ds_or (atomic read-modify-write).
// BYTES = 1: medium primes store a 1 into a byte per position (plain
//            ds_write_b8: idempotent, no read, two VALU per mark); the
//            pattern stage keeps the bit array and compaction consults the
//            bytes only for the few positions the patterns left alive.
template <int T, int WPT, int BYTES>
__global__ __launch_bounds__(T) void k_sieve(SieveArgs a)
{
    extern __shared__ uint64_t lds64[];
    constexpr uint32_t S = 64u * T * WPT;          // line positions per tile
    constexpr uint32_t MARK64 = BYTES ? S / 8 : 0;  // u64 words of marks
    uint32_t* bits = reinterpret_cast<uint32_t*>(lds64);     // S/32 words
    uint8_t* marks = reinterpret_cast<uint8_t*>(lds64 + S / 64);  // [S]
    uint32_t* s0 = reinterpret_cast<uint32_t*>(lds64 + S / 64 + MARK64);
    uint64_t* pat = lds64 + (S / 64) + MARK64 + ((a.nlds_max + 1) / 2);
    __shared__ uint32_t sh_count, sh_base;

    const uint32_t tid = threadIdx.x;
    uint32_t cur_engine = 0xffffffffu;

    for (uint32_t tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
        const uint32_t pi = find_piece(a.pieces, a.npieces, tile);
        const DevPiece& P = a.pieces[pi];
        const DevEngine& E = a.engines[P.engine];
        const uint32_t local = tile - P.tile0;
        const uint32_t line = local / P.nseg;
        const uint32_t seg = local - line * P.nseg;
        const uint64_t v0 = (uint64_t)seg * S;
        const uint64_t len = (P.npos - v0 < S) ? (P.npos - v0) : S;
        const uint64_t off = E.offs[line];
        const uint64_t D = E.D;
        // valid positions of this line: lo_rel <= off + D v < hi_rel
        uint64_t vlo = (off >= P.lo_rel) ? 0 : ceil_div_u64(P.lo_rel - off, D);
        uint64_t vhi = (P.hi_rel > off) ? ceil_div_u64(P.hi_rel - off, D) : 0;
        uint64_t vend = (vhi < v0 + len) ? vhi : v0 + len;
        if (vlo >= vend || vhi <= v0 || vlo >= v0 + len) continue;

        if (P.engine != cur_engine) {
            __syncthreads();
            for (uint32_t w = tid; w < E.pat_words; w += T) pat[w] = E.pat[w];
            cur_engine = P.engine;
        }
        // s0[p] = (v0 + t_p) mod q with t_p = (B + off) D^-1 mod q
        for (uint32_t p = tid; p < E.nlds; p += T) {
            const uint32_t q = E.lds_q[p];
            const uint64_t m64 = E.lds_m64[p];
            uint32_t t = a.bq[P.bq_off + p] + red64(off, q, m64);
            if (t >= q) t -= q;
            t = mulmod32(t, E.lds_dinv[p], q, m64);
            s0[p] = red64(v0 + t, q, m64);
        }
        __syncthreads();

        // ---- pattern words: OR of every small prime's precomputed word
        uint64_t acc[WPT];
#pragma unroll
        for (int k = 0; k < WPT; k++)
            acc[k] = edge_mask(v0 + 64ull * (tid + (uint32_t)k * T), vlo, vend);
        for (uint32_t p = 0; p < E.npat; p++) {
            const uint32_t q = E.lds_q[p];
            const uint32_t m32 = E.lds_m32[p];
            const uint32_t step = red32(64u * T, q, m32);
            const uint64_t* tab = pat + E.pat_off[p];
            uint32_t s = red32(s0[p] + 64u * tid, q, m32);
#pragma unroll
            for (int k = 0; k < WPT; k++) {
                acc[k] |= tab[s];
                s += step;
                if (s >= q) s -= q;
            }
        }
#pragma unroll
        for (int k = 0; k < WPT; k++) {
            const uint32_t w = tid + (uint32_t)k * T;
            bits[2 * w] = (uint32_t)acc[k];
            bits[2 * w + 1] = (uint32_t)(acc[k] >> 32);
        }
        if (BYTES) {
            uint4* m4 = reinterpret_cast<uint4*>(marks);
            const uint32_t nz = ((uint32_t)len + 15) / 16;
            for (uint32_t i = tid; i < nz; i += T) m4[i] = make_uint4(0, 0, 0, 0);
        }
        if (tid == 0) sh_count = 0;
        __syncthreads();

        // ---- medium primes: stride marking with LDS atomics
        const uint32_t ulen = (uint32_t)len;
        for (uint32_t i = tid; i < E.npairs; i += T) {
            const uint32_t pr = E.pairs[i];
            const uint32_t pidx = pr >> 16;
            const uint32_t res = pr & 0xffffu;
            const uint32_t q = E.lds_q[pidx];
            const uint32_t s = s0[pidx];
            uint32_t pos = (res >= s) ? res - s : res + q - s;
            if (BYTES) {
                for (; pos < ulen; pos += q) marks[pos] = 1;
            } else {
                for (; pos < ulen; pos += q)
                    atomicOr(&bits[pos >> 5], 1u << (pos & 31));
            }
        }
        __syncthreads();

        // ---- compaction: surviving bits -> candidate records
        uint64_t live[WPT];
        uint32_t cnt = 0;
#pragma unroll
        for (int k = 0; k < WPT; k++) {
            const uint32_t w = tid + (uint32_t)k * T;
            live[k] = ~(((uint64_t)bits[2 * w + 1] << 32) | bits[2 * w]);
            if (BYTES) {
                uint64_t m = live[k];
                while (m) {
                    const int b = __builtin_ctzll(m);
                    m &= m - 1;
                    if (marks[64 * w + b]) live[k] &= ~(1ull << b);
                }
            }
            cnt += __popcll(live[k]);
        }
        uint32_t mine = 0;
        if (cnt) mine = atomicAdd(&sh_count, cnt);
        __syncthreads();
        if (tid == 0 && sh_count) sh_base = atomicAdd(a.out_count, sh_count);
        __syncthreads();
        if (cnt) {
            uint32_t slot = sh_base + mine;
            const unsigned __int128 B =
                ((unsigned __int128)(((uint64_t)P.base[3] << 32) | P.base[2]) << 64)
                | (((uint64_t)P.base[1] << 32) | P.base[0]);
#pragma unroll
            for (int k = 0; k < WPT; k++) {
                uint64_t m = live[k];
                while (m) {
                    const int b = __builtin_ctzll(m);
                    m &= m - 1;
                    const uint64_t v = v0 + 64ull * (tid + (uint32_t)k * T) + b;
                    unsigned __int128 x = B + off + (unsigned __int128)D * v;
                    if (E.family == kDickson) x *= E.W;
                    if (slot < a.out_cap) {
                        DevRec r;
                        r.t[0] = (uint32_t)x;
                        r.t[1] = (uint32_t)(x >> 32);
                        r.t[2] = (uint32_t)(x >> 64);
                        r.t[3] = (uint32_t)(x >> 96);
                        r.window = P.window;
                        r.run = -1;
                        a.out[slot] = r;
                    }
                    slot++;
                }
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- rounds

// A candidate staged for the kill tests, with the few engine constants the
// tests read, so the per-prime loop touches LDS only.
struct Staged {
    uint32_t t[4];
    int32_t family, n;
    uint32_t round_first, round_end, xxmax;
    uint32_t d[4];
    const uint32_t* xxbits;
};

// Does prime q divide one of the family's values for this candidate?
__device__ __forceinline__ bool kill_test(const Staged& c, const PrimeK& pk)
{
    const uint32_t q = pk.q;
    const uint32_t r = red128(c.t, pk);
    if (c.family == kEuler) {
        // q | x^2+x+p  <=>  (q - p mod q) mod q == x^2+x (mod q)
        if (q > c.xxmax) {
            const uint32_t dd = r ? q - r : 0;
            return dd <= c.xxmax && ((c.xxbits[dd >> 5] >> (dd & 31)) & 1);
        }
        for (int x = 0; x < c.n; x++) {
            const uint64_t v = (uint64_t)x * x + x + r;
            if (v % q == 0) return true;
        }
        return false;
    } else if (c.family == kDickson) {
        // r = k mod q; q | m k^2 + 1 for some m <= n
        const uint32_t kk = mulmod32(r, r, q, pk.m64);
        if (kk == 0) return false;
        uint32_t v = kk + 1;
        if (v >= q) v -= q;
        for (int m = 1; m <= c.n; m++) {
            if (v == 0) return true;
            v += kk;
            if (v >= q) v -= q;
        }
        return false;
    } else {
        // q | p + j d for some j < n
        const uint32_t dq = red128(c.d, pk);
        if (dq == 0) return r == 0;
        uint32_t f = r;
        for (int j = 0; j < c.n; j++) {
            if (f == 0) return true;
            f += dq;
            if (f >= q) f -= q;
        }
        return false;
    }
}

// A workgroup owns 256 consecutive primes of the round (one per lane, its
// constants held in registers) and walks the candidate queue RT_RECS
// records at a time, staged in LDS; every lane tests its prime against each
// staged record, so no lane runs a serial chain of dependent loads.
constexpr int RT_RECS = 16;

__global__ __launch_bounds__(256) void k_round_test(RoundArgs a)
{
    __shared__ Staged st[RT_RECS];
    __shared__ uint32_t slive[RT_RECS];
    const uint32_t count = min(*a.count, a.cap);
    const uint32_t tid = threadIdx.x;
    for (uint32_t pb = blockIdx.x; pb < a.nchunks; pb += gridDim.x) {
        const uint32_t p = a.p_begin + pb * 256 + tid;
        const bool have = p < a.p_end;
        const PrimeK pk = have ? a.primes[p] : a.primes[a.p_begin];
        for (uint32_t r0 = blockIdx.y * RT_RECS; r0 < count;
             r0 += gridDim.y * RT_RECS) {
            __syncthreads();
            if (tid < RT_RECS) {
                const uint32_t i = r0 + tid;
                uint32_t live = 0;
                if (i < count && a.flags[i] == 0) {
                    const DevRec rec = a.recs[i];
                    const DevEngine& E = a.engines[a.win_engine[rec.window]];
                    Staged c;
                    for (int k = 0; k < 4; k++) {
                        c.t[k] = rec.t[k];
                        c.d[k] = E.d[k];
                    }
                    c.family = E.family;
                    c.n = E.n;
                    c.round_first = E.round_first;
                    c.round_end = E.round_end;
                    c.xxmax = E.xxmax;
                    c.xxbits = E.xxbits;
                    st[tid] = c;
                    live = 1;
                }
                slive[tid] = live;
            }
            __syncthreads();
            if (!have) continue;
            for (int k = 0; k < RT_RECS; k++) {
                if (!slive[k]) continue;
                const Staged& c = st[k];
                if (p >= c.round_first && p < c.round_end && kill_test(c, pk))
                    a.flags[r0 + k] = 1;
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_round_compact(CompactArgs a)
{
    const uint32_t count = min(*a.in_count, a.cap);
    const uint32_t lane = __lane_id();
    for (uint32_t base = blockIdx.x * blockDim.x; base < count;
         base += gridDim.x * blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        bool alive = false;
        if (i < count) {
            alive = a.flags[i] == 0;
            a.flags[i] = 0;
        }
        const uint64_t mask = __ballot(alive);
        if (mask == 0) continue;
        const int leader = __ffsll((unsigned long long)mask) - 1;
        uint32_t wbase = 0;
        if (lane == (uint32_t)leader)
            wbase = atomicAdd(a.out_count, (uint32_t)__popcll(mask));
        wbase = __shfl(wbase, leader);
        if (alive) {
            const uint32_t below = __builtin_amdgcn_mbcnt_hi(
                (uint32_t)(mask >> 32),
                __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
            a.out[wbase + below] = a.in[i];
        }
    }
}

// -------------------------------------------------------------- classify

// The x-th / m-th / j-th value of the family's chain for one candidate.
__device__ __forceinline__ void chain_value(const DevEngine& E,
                                            const uint32_t t[4],
                                            const uint32_t kk[8], int i,
                                            uint32_t v[8])
{
    if (E.family == kEuler) {                 // p + x^2 + x
        const uint64_t xx = (uint64_t)i * i + i;
        uint32_t add[8] = {(uint32_t)xx, (uint32_t)(xx >> 32), 0, 0,
                           0, 0, 0, 0};
        uint32_t p8[8] = {t[0], t[1], t[2], t[3], 0, 0, 0, 0};
        mw_add<8>(v, p8, add);
    } else if (E.family == kDickson) {        // (i + 1) k^2 + 1
        mw_mul_small(v, kk, 8, (uint32_t)(i + 1));
        uint32_t one[8] = {1, 0, 0, 0, 0, 0, 0, 0};
        mw_add<8>(v, v, one);
    } else {                                  // p + i d
        uint32_t jd[8];
        mw_mul_small(jd, E.d, 8, (uint32_t)i);
        uint32_t p8[8] = {t[0], t[1], t[2], t[3], 0, 0, 0, 0};
        mw_add<8>(v, p8, jd);
    }
}

// Classification in three passes, parallel over (record, value, base):
//   k_cls_base2   every value i < cap of every record, base 2 alone;
//                 failures set bit i of fail2[record]
//   k_cls_full    bases 2..7 (indices 1..6) for the values before the first
//                 base-2 failure; failures set bit i of failx[record]
//   k_cls_finish  run = first failing value (FULL), or the base-2 bound
//                 when it is below exact_from (TWO_PASS); clears the masks
// A sequential chain stops at its first composite; here the later values'
// base-2 tests run too, in parallel, and are ignored.  The verdicts are the
// sequential chain's exactly, because the seven-base test is the AND of the
// single-base tests (huntlib/primes.py:31-47).

__device__ __forceinline__ void rec_chain_value(const DevEngine& E,
                                                const DevRec& rec, int i,
                                                uint32_t v[8])
{
    uint32_t kk[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (E.family == kDickson) mw_mul44(kk, rec.t, rec.t);
    chain_value(E, rec.t, kk, i, v);
}

__device__ __forceinline__ int first_set(uint64_t m, int cap)
{
    return m ? min(cap, __builtin_ctzll(m)) : cap;
}

__global__ __launch_bounds__(128) void k_cls_base2(ClassifyArgs a)
{
    const uint32_t count = min(*a.count, a.cap);
    const uint64_t total = (uint64_t)count * a.cap_max;
    for (uint64_t idx = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
         idx < total; idx += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t r = (uint32_t)(idx / a.cap_max);
        const int i = (int)(idx - (uint64_t)r * a.cap_max);
        const DevRec rec = a.recs[r];
        const DevWinInfo W = a.wins[rec.window];
        if (W.mode == 0 || i >= W.cap) continue;
        const DevEngine& E = a.engines[W.engine];
        uint32_t v[8];
        rec_chain_value(E, rec, i, v);
        if (!is_sprp(v, 0, 1))
            atomicOr((unsigned long long*)&a.fail2[r], 1ull << i);
    }
}

__global__ __launch_bounds__(128) void k_cls_full(ClassifyArgs a)
{
    const uint32_t count = min(*a.count, a.cap);
    // single pass: all seven bases (base 2 into fail2) with no dependence
    // on a base-2 pass; otherwise bases 3..7 after it
    const uint32_t nb = a.single ? 7 : 6;
    const uint64_t per = (uint64_t)a.cap_max * nb;
    const uint64_t total = (uint64_t)count * per;
    for (uint64_t idx = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
         idx < total; idx += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t r = (uint32_t)(idx / per);
        const uint32_t rem = (uint32_t)(idx - (uint64_t)r * per);
        const int i = (int)(rem / nb);
        const int b = (int)(rem - (uint32_t)i * nb) + (a.single ? 0 : 1);
        if (b == 0) {
            const DevRec rec = a.recs[r];
            const DevWinInfo W = a.wins[rec.window];
            if (W.mode == 0 || i >= W.cap) continue;
            const DevEngine& E = a.engines[W.engine];
            uint32_t v[8];
            rec_chain_value(E, rec, i, v);
            if (!is_sprp(v, 0, 1))
                atomicOr((unsigned long long*)&a.fail2[r], 1ull << i);
            continue;
        }
        const DevRec rec = a.recs[r];
        const DevWinInfo W = a.wins[rec.window];
        if (W.mode == 0 || i >= W.cap) continue;
        if (!a.single) {
            const int u = first_set(a.fail2[r], W.cap);
            if (i >= u) continue;
            if (W.mode == 2 && u < W.exact_from) continue;
        }
        const DevEngine& E = a.engines[W.engine];
        uint32_t v[8];
        rec_chain_value(E, rec, i, v);
        if (!is_sprp(v, b, 1))
            atomicOr((unsigned long long*)&a.failx[r], 1ull << i);
    }
}

__global__ __launch_bounds__(256) void k_cls_finish(ClassifyArgs a)
{
    const uint32_t count = min(*a.count, a.cap);
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < count;
         r += gridDim.x * blockDim.x) {
        const DevWinInfo W = a.wins[a.recs[r].window];
        const uint64_t f2 = a.fail2[r], fx = a.failx[r];
        a.fail2[r] = 0;
        a.failx[r] = 0;
        if (W.mode == 0) continue;
        const int u = first_set(f2, W.cap);
        int run;
        if (W.mode == 2 && u < W.exact_from) run = u;
        else run = min(u, first_set(fx, W.cap));
        a.recs[r].run = run;
    }
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += gridDim.x * blockDim.x)
        p[i] = 0;
}

// --------------------------------------------------------------- launchers

#define FK_SIEVE_SHAPES(X)                                                 \
    X(256, 1) X(256, 2) X(256, 4) X(256, 8) X(256, 16)                       \
    X(512, 1) X(512, 2) X(512, 4) X(512, 8)                                  \
    X(1024, 1) X(1024, 2) X(1024, 4)

hipError_t launch_sieve(const SieveArgs& a, int threads, int wpt, int bytes,
                        uint32_t grid, size_t lds_bytes, hipStream_t stream)
{
#define FK_SIEVE_CASE(TT, WW)                                              \
    if (threads == TT && wpt == WW) {                                      \
        auto fn = bytes ? k_sieve<TT, WW, 1> : k_sieve<TT, WW, 0>;         \
        if (lds_bytes > 65536)                                             \
            (void)hipFuncSetAttribute((const void*)fn,                     \
                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes); \
        hipLaunchKernelGGL(fn, dim3(grid), dim3(TT), lds_bytes, stream, a); \
        return hipGetLastError();                                          \
    }
    FK_SIEVE_SHAPES(FK_SIEVE_CASE)
#undef FK_SIEVE_CASE
    return hipErrorInvalidValue;
}

int sieve_occupancy(int threads, int wpt, int bytes, size_t lds_bytes)
{
    int blocks = 0;
#define FK_OCC_CASE(TT, WW)                                                 \
    if (threads == TT && wpt == WW) {                                       \
        auto fn = bytes ? k_sieve<TT, WW, 1> : k_sieve<TT, WW, 0>;          \
        if (lds_bytes > 65536)                                              \
            (void)hipFuncSetAttribute((const void*)fn,                      \
                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);  \
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(                   \
                &blocks, fn, TT, lds_bytes) != hipSuccess) blocks = 0;      \
        return blocks;                                                      \
    }
    FK_SIEVE_SHAPES(FK_OCC_CASE)
#undef FK_OCC_CASE
    return 0;
}

bool sieve_shape_ok(int threads, int wpt)
{
#define FK_OK_CASE(TT, WW) if (threads == TT && wpt == WW) return true;
    FK_SIEVE_SHAPES(FK_OK_CASE)
#undef FK_OK_CASE
    return false;
}

hipError_t launch_prep(const DevEngine* engines, const DevPiece* pieces,
                       uint32_t npieces, uint32_t max_nlds, uint32_t* bq,
                       hipStream_t stream)
{
    if (npieces == 0 || max_nlds == 0) return hipSuccess;
    uint32_t gx = (max_nlds + 255) / 256;
    if (gx > 64) gx = 64;
    const uint32_t gy = npieces < 32768 ? npieces : 32768;
    hipLaunchKernelGGL(k_prep, dim3(gx, gy), dim3(256), 0, stream,
                       engines, pieces, npieces, bq);
    return hipGetLastError();
}

hipError_t launch_round_test(const RoundArgs& a, uint32_t gy,
                             hipStream_t stream)
{
    uint32_t gx = a.nchunks ? a.nchunks : 1;
    if (gx > 1u << 20) gx = 1u << 20;
    if (gy < 1) gy = 1;
    if (gy > 65535) gy = 65535;
    hipLaunchKernelGGL(k_round_test, dim3(gx, gy), dim3(256), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_round_compact(const CompactArgs& a, uint32_t grid,
                                hipStream_t stream)
{
    hipLaunchKernelGGL(k_round_compact, dim3(grid ? grid : 1), dim3(256), 0,
                       stream, a);
    return hipGetLastError();
}

hipError_t launch_classify(const ClassifyArgs& a, uint32_t grid,
                           hipStream_t stream)
{
    if (grid < 1) grid = 1;
    hipError_t e;
    if (!a.single) {
        hipLaunchKernelGGL(k_cls_base2, dim3(grid), dim3(128), 0, stream, a);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_cls_full, dim3(grid), dim3(128), 0, stream, a);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    uint32_t g2 = grid / 4 + 1;
    hipLaunchKernelGGL(k_cls_finish, dim3(g2), dim3(256), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_zero(uint32_t* p, uint32_t n, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    uint32_t grid = (n + 255) / 256;
    if (grid > 1024) grid = 1024;
    hipLaunchKernelGGL(k_zero_u32, dim3(grid), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

}  // namespace fk
