// fk_kernels.hip -- the device side of the sieve-and-test engine (gfx950).
//
// Pipeline for a batch of window pieces (see fk_device.h for the line
// model):
//   k_prep          B mod q for every (piece, LDS prime)
//   k_sieve         one LDS tile (S line positions of one wheel line) per
//                   workgroup iteration: pattern words for the smallest
//                   primes, optional ds_or stride marking, per-survivor
//                   tests against LDS kill bitmaps for the medium primes,
//                   then compaction of the surviving bits into the queue
//   k_round         per-candidate kill tests for one range of large primes:
//                   lane = candidate, the prime uniform across the wave (its
//                   constants and kill set in scalar registers), the residue
//                   formed exactly in double precision; killed candidates
//                   flagged
//   k_kill_table    (engine creation) the value-domain kill set of every
//                   round prime
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

#include <type_traits>

#include "fk_device.h"
#include "fk_kill.h"
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

// One tile = S consecutive positions of one wheel line, in LDS as bits
// (1 = dead).  Per tile:
//   1. s0[p] = line position of the tile start, shifted into prime p's
//      kill coordinates, for every LDS-stage prime
//   2. pattern stage: the first npat primes OR in precomputed 64-bit words
//   3. marking stage: primes [npat, nmark) stride-marked with ds_or
//   4. test stage: the positions still alive are listed in LDS and run
//      through in-tile rounds over the primes [nmark, nlds); in a round
//      every lane tests the same prime (broadcast LDS reads) against its
//      own candidate -- one Barrett step and one bit of the prime's
//      line-domain kill bitmap -- and the survivors are compacted with a
//      wave ballot into the other list
//   5. the survivors become candidate records
// Per-engine LDS (loaded when a workgroup meets a new engine): pattern
// words, {q, floor(2^32/q), bitmap offset} per prime, kill bitmaps.
template <uint32_t S>
struct TileList {
    typedef typename std::conditional<(S <= 65536), uint16_t, uint32_t>::type T;
};

__device__ __forceinline__ void emit_record(const SieveArgs& a,
                                            const DevPiece& P,
                                            const DevEngine& E,
                                            unsigned __int128 B, uint64_t off,
                                            uint64_t v, uint32_t slot)
{
    unsigned __int128 x = B + off + (unsigned __int128)E.D * v;
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
}

template <int T, int WPT>
__global__ __launch_bounds__(T) void k_sieve(SieveArgs a)
{
    extern __shared__ uint64_t lds64[];
    constexpr uint32_t S = 64u * T * WPT;          // line positions per tile
    // listed survivors per tile: past this the tile tests word by word
    constexpr uint32_t LCAP = (S / 16 < 4096) ? S / 16 : 4096;
    typedef typename TileList<S>::T PosT;
    uint32_t* bits = reinterpret_cast<uint32_t*>(lds64);          // S/32
    uint32_t* s0 = reinterpret_cast<uint32_t*>(lds64 + S / 64);   // nlds_max
    uint4* pinfo = reinterpret_cast<uint4*>(lds64 + S / 64 +
                                            ((a.nlds_max + 3) / 4) * 2);
    uint64_t* pat = reinterpret_cast<uint64_t*>(pinfo + a.nlds_max);
    uint32_t* bm = reinterpret_cast<uint32_t*>(pat + a.pat_max);
    PosT* listA = reinterpret_cast<PosT*>(bm + ((a.bm_max + 3) & ~3u));
    PosT* listB = listA + LCAP;
    __shared__ uint32_t sh_count, sh_base, sh_next, sh_n, sh_m;

    const uint32_t tid = threadIdx.x;
    const uint32_t lane = __lane_id();
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
            for (uint32_t w = tid; w < E.bm_words; w += T) bm[w] = E.bm[w];
            for (uint32_t p = tid; p < E.nlds; p += T)
                pinfo[p] = make_uint4(E.lds_q[p], E.lds_m32[p],
                                      p >= E.nmark ? E.bm_off[p - E.nmark] : 0,
                                      0);
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
        if (tid == 0) { sh_count = 0; sh_next = 0; sh_n = 0; }
        __syncthreads();

        // ---- 2. pattern words: OR of every small prime's precomputed word
        uint64_t acc[WPT];
#pragma unroll
        for (int k = 0; k < WPT; k++)
            acc[k] = edge_mask(v0 + 64ull * (tid + (uint32_t)k * T), vlo, vend);
        for (uint32_t p = 0; p < E.npat; p++) {
            const uint4 pq = pinfo[p];
            const uint32_t q = pq.x, m32 = pq.y;
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
        __syncthreads();

        // ---- 3. marking stage: stride marking with LDS atomics
        const uint32_t ulen = (uint32_t)len;
        if (E.npairs) {
            for (uint32_t i = tid; i < E.npairs; i += T) {
                const uint32_t pr = E.pairs[i];
                const uint32_t pidx = pr >> 16;
                const uint32_t res = pr & 0xffffu;
                const uint32_t q = pinfo[pidx].x;
                const uint32_t s = s0[pidx];
                uint32_t pos = (res >= s) ? res - s : res + q - s;
                for (; pos < ulen; pos += q)
                    atomicOr(&bits[pos >> 5], 1u << (pos & 31));
            }
            __syncthreads();
        }

        uint64_t live[WPT];
        uint32_t cnt = 0;
#pragma unroll
        for (int k = 0; k < WPT; k++) {
            const uint32_t w = tid + (uint32_t)k * T;
            live[k] = ~(((uint64_t)bits[2 * w + 1] << 32) | bits[2 * w]);
            cnt += __popcll(live[k]);
        }

        // ---- 4. test stage
        bool listed = false;
        PosT* cur = listA;
        uint32_t n = 0;
        if (E.nmark < E.nlds) {
            // 4a. list the live positions
            const uint32_t at = cnt ? atomicAdd(&sh_n, cnt) : 0;
            if (cnt && at + cnt <= LCAP) {
                uint32_t k2 = at;
#pragma unroll
                for (int k = 0; k < WPT; k++) {
                    uint64_t m = live[k];
                    const uint32_t w = tid + (uint32_t)k * T;
                    while (m) {
                        const int b = __builtin_ctzll(m);
                        m &= m - 1;
                        cur[k2++] = (PosT)(64u * w + b);
                    }
                }
            }
            __syncthreads();
            n = sh_n;
            if (n <= LCAP) {
                listed = true;
                // 4b. in-tile rounds over the listed candidates
                PosT* nxt = listB;
                uint32_t p = E.nmark;
                for (uint32_t r = 0; r < E.ntr && n; r++) {
                    const uint32_t pend = E.tr_end[r];
                    if (tid == 0) sh_m = 0;
                    __syncthreads();
                    for (uint32_t b0 = 0; b0 < n; b0 += T) {
                        const uint32_t i = b0 + tid;
                        bool alive = false;
                        uint32_t pos = 0;
                        if (i < n) {
                            pos = cur[i];
                            alive = true;
                            for (uint32_t pp = p; pp < pend; pp++) {
                                const uint4 pq = pinfo[pp];
                                const uint32_t x = red32(s0[pp] + pos, pq.x, pq.y);
                                if ((bm[pq.z + (x >> 5)] >> (x & 31)) & 1) {
                                    alive = false;
                                    break;
                                }
                            }
                        }
                        const uint64_t mask = __ballot(alive);
                        if (mask) {
                            const int leader = __ffsll((unsigned long long)mask) - 1;
                            uint32_t wb = 0;
                            if (lane == (uint32_t)leader)
                                wb = atomicAdd(&sh_m, (uint32_t)__popcll(mask));
                            wb = __shfl(wb, leader);
                            if (alive) {
                                const uint32_t below = __builtin_amdgcn_mbcnt_hi(
                                    (uint32_t)(mask >> 32),
                                    __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
                                nxt[wb + below] = (PosT)pos;
                            }
                        }
                    }
                    __syncthreads();
                    n = sh_m;
                    PosT* t = cur; cur = nxt; nxt = t;
                    p = pend;
                    __syncthreads();        // every lane has read sh_m
                }
            } else {
                // 4c. too many to list: test word by word (divergent but
                // correct), then fall through to the bit compaction
                const uint32_t nwords = (ulen + 63) / 64;
                const uint32_t p0 = E.nmark, p1 = E.nlds;
                for (;;) {
                    const uint32_t w = atomicAdd(&sh_next, 1u);
                    if (w >= nwords) break;
                    uint64_t lv = ~(((uint64_t)bits[2 * w + 1] << 32) | bits[2 * w]);
                    uint64_t dead = 0, m = lv;
                    while (m) {
                        const int b = __builtin_ctzll(m);
                        m &= m - 1;
                        const uint32_t pos = 64u * w + b;
                        for (uint32_t pp = p0; pp < p1; pp++) {
                            const uint4 pq = pinfo[pp];
                            const uint32_t x = red32(s0[pp] + pos, pq.x, pq.y);
                            if ((bm[pq.z + (x >> 5)] >> (x & 31)) & 1) {
                                dead |= 1ull << b;
                                break;
                            }
                        }
                    }
                    if (dead) {
                        const uint64_t nw = ~(lv & ~dead);
                        bits[2 * w] = (uint32_t)nw;
                        bits[2 * w + 1] = (uint32_t)(nw >> 32);
                    }
                }
                __syncthreads();
                cnt = 0;
#pragma unroll
                for (int k = 0; k < WPT; k++) {
                    const uint32_t w = tid + (uint32_t)k * T;
                    live[k] = ~(((uint64_t)bits[2 * w + 1] << 32) | bits[2 * w]);
                    cnt += __popcll(live[k]);
                }
            }
        }

        // ---- 5. survivors -> candidate records
        const unsigned __int128 B =
            ((unsigned __int128)(((uint64_t)P.base[3] << 32) | P.base[2]) << 64)
            | (((uint64_t)P.base[1] << 32) | P.base[0]);
        if (listed) {
            if (tid == 0 && n) sh_base = atomicAdd(a.out_count, n);
            __syncthreads();
            for (uint32_t i = tid; i < n; i += T)
                emit_record(a, P, E, B, off, v0 + cur[i], sh_base + i);
        } else {
            uint32_t mine = 0;
            if (cnt) mine = atomicAdd(&sh_count, cnt);
            __syncthreads();
            if (tid == 0 && sh_count) sh_base = atomicAdd(a.out_count, sh_count);
            __syncthreads();
            if (cnt) {
                uint32_t slot = sh_base + mine;
#pragma unroll
                for (int k = 0; k < WPT; k++) {
                    uint64_t m = live[k];
                    while (m) {
                        const int b = __builtin_ctzll(m);
                        m &= m - 1;
                        emit_record(a, P, E, B, off,
                                    v0 + 64ull * (tid + (uint32_t)k * T) + b,
                                    slot++);
                    }
                }
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- rounds

// Kill-set table of each engine's round primes, built once at engine
// creation: one thread per prime.  Row layout: [count, residues...,
// kKillPad...], stride words per row.
__global__ void k_kill_table(const PrimeK* __restrict__ primes, uint32_t p0,
                             uint32_t np, int family, int n, uint4 d,
                             uint32_t stride, uint32_t* __restrict__ kt)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const uint32_t dl[4] = {d.x, d.y, d.z, d.w};
    uint32_t* row = kt + (size_t)i * stride;
    row[0] = kill_values(family, n, dl, primes[p0 + i], row + 1, stride - 1);
}

// Read-only tables the kernels index with wave-uniform indices, viewed in
// the constant address space: loads through it become scalar loads into
// SGPRs (the compiler cannot otherwise prove that the kernel's own stores
// leave them alone).
#define FK_CONST __attribute__((address_space(4)))
template <class T>
__device__ __forceinline__ const FK_CONST T* uniform_view(const T* p)
{
    return (const FK_CONST T*)p;
}

// The candidate's residue mod the prime from its NL 16-bit limbs.
template <int NL>
__device__ __forceinline__ uint32_t limb_residue(const double (&a)[8],
                                                 const FK_CONST PrimeD* P)
{
    double s = a[0];
#pragma unroll
    for (int k = 1; k < NL; k++) s = fma(a[k], P->c[k - 1], s);
    return resid_d(s, P->qd, P->qi, P->q);
}

__device__ __forceinline__ uint4 load4(const FK_CONST uint32_t* p)
{
    return make_uint4(p[0], p[1], p[2], p[3]);
}

// Lanes whose residue r matches one of the four words.
__device__ __forceinline__ uint64_t match4(uint32_t r, uint4 v)
{
    return __ballot(r == v.x) | __ballot(r == v.y) | __ballot(r == v.z) |
           __ballot(r == v.w);
}

// Primes [lo, hi) of one engine against this lane's candidate; returns the
// mask of lanes killed.  The prime index is uniform across the wave, so the
// prime's constants and kill-set row are scalar loads; the lane does the
// reduction and the compares (one v_cmp into a lane mask per residue).  The
// first 16 words of a row (count and up to 15 residues, padded) are read
// unconditionally; longer rows continue in a loop.
template <int NL>
__device__ __forceinline__ uint64_t round_scan(const double (&a)[8],
                                               const FK_CONST PrimeD* pd,
                                               const FK_CONST uint32_t* kt,
                                               uint32_t stride, uint32_t rf,
                                               uint32_t lo, uint32_t hi)
{
    uint64_t hit = 0;
    for (uint32_t p = lo; p < hi; p++) {
        const FK_CONST uint32_t* row = kt + (size_t)(p - rf) * stride;
        const uint4 v0 = load4(row), v1 = load4(row + 4), v2 = load4(row + 8),
                    v3 = load4(row + 12);
        const uint32_t r = limb_residue<NL>(a, pd + p);
        hit |= __ballot(r == v0.y) | __ballot(r == v0.z) |
               __ballot(r == v0.w) | match4(r, v1) | match4(r, v2) |
               match4(r, v3);
        for (uint32_t c = 16; c <= v0.x; c += 4) hit |= match4(r, load4(row + c));
    }
    return hit;
}

// One round: every queued candidate against the primes [p_begin, p_end),
// lane = candidate, blockIdx.y = a block of pblock primes.  A wave whose
// lanes come from several engines takes them one engine at a time (the
// engine index is made uniform with readlane), so the prime-side data
// stays uniform.  A killed candidate's flag is set; the compaction that
// follows drops it.
__global__ __launch_bounds__(256) void k_round(RoundArgs a)
{
    const uint32_t count = min(*a.count, a.cap);
    const uint32_t pb0 = a.p_begin + blockIdx.y * a.pblock;
    const uint32_t pb1 = min(pb0 + a.pblock, a.p_end);
    if (pb0 >= pb1) return;
    for (uint32_t base = blockIdx.x * blockDim.x; base < count;
         base += gridDim.x * blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        const bool live = i < count && a.flags[i] == 0;
        uint32_t t[4] = {0u, 0u, 0u, 0u};
        uint32_t eidx = 0;
        if (live) {
            const DevRec rec = a.recs[i];
            for (int k = 0; k < 4; k++) t[k] = rec.t[k];
            eidx = a.win_engine[rec.window];
        }
        double l[8];
#pragma unroll
        for (int k = 0; k < 8; k++)
            l[k] = (double)((t[k >> 1] >> ((k & 1) * 16)) & 0xffffu);
        const bool wide = __ballot(live && (t[2] | t[3]) != 0) != 0;
        bool killed = false;
        uint64_t todo = __ballot(live);
        while (todo) {
            const int lead = __ffsll((unsigned long long)todo) - 1;
            const uint32_t e = __builtin_amdgcn_readlane(eidx, lead);
            const bool mine = live && eidx == e;
            todo &= ~(uint64_t)__ballot(mine);
            const FK_CONST DevEngine* E = uniform_view(a.engines) + e;
            const uint32_t rf = E->round_first;
            const uint32_t lo = max(pb0, rf);
            const uint32_t hi = min(pb1, E->round_end);
            const FK_CONST PrimeD* pd = uniform_view(a.primed);
            const FK_CONST uint32_t* kt = uniform_view(E->kt);
            if (mine && lo < hi) {
                const uint64_t m =
                    wide ? round_scan<8>(l, pd, kt, E->kstride, rf, lo, hi)
                         : round_scan<4>(l, pd, kt, E->kstride, rf, lo, hi);
                killed = killed || ((m >> __lane_id()) & 1);
            }
        }
        if (killed) a.flags[i] = 1;
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

hipError_t launch_sieve(const SieveArgs& a, int threads, int wpt,
                        uint32_t grid, size_t lds_bytes, hipStream_t stream)
{
#define FK_SIEVE_CASE(TT, WW)                                              \
    if (threads == TT && wpt == WW) {                                      \
        auto fn = k_sieve<TT, WW>;                                         \
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

int sieve_occupancy(int threads, int wpt, size_t lds_bytes)
{
    int blocks = 0;
#define FK_OCC_CASE(TT, WW)                                                 \
    if (threads == TT && wpt == WW) {                                       \
        auto fn = k_sieve<TT, WW>;                                          \
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

hipError_t launch_round_test(const RoundArgs& a, uint32_t gx, uint32_t gy,
                             hipStream_t stream)
{
    // every prime block must be launched: the host sizes pblock for gy
    if (gx < 1 || gy < 1 || gy > 65535 || a.pblock == 0 ||
        (uint64_t)gy * a.pblock < (uint64_t)(a.p_end - a.p_begin))
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_round, dim3(gx, gy), dim3(256), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_kill_table(const PrimeK* primes, uint32_t p0, uint32_t np,
                             int family, int n, const uint32_t d[4],
                             uint32_t stride, uint32_t* kt, hipStream_t stream)
{
    if (np == 0) return hipSuccess;
    if (stride < 16 || (stride & 3) || stride < 1 + kill_max(family, n))
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_kill_table, dim3((np + 255) / 256), dim3(256), 0,
                       stream, primes, p0, np, family, n,
                       make_uint4(d[0], d[1], d[2], d[3]), stride, kt);
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
