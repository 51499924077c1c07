// fk_kernels.hip -- the device side of the sieve-and-test engine (gfx950).
//
// Pipeline for a batch of window pieces (see fk_device.h for the line
// model):
//   k_sieve         one LDS tile (S line positions of one wheel line) per
//                   workgroup iteration: pattern words for the smallest
//                   primes, optional ds_or stride marking, per-survivor
//                   tests against LDS kill bitmaps for the medium primes,
//                   then compaction of the surviving bits into the queue
//   k_sieve_short   windows whose lines hold a few words: one lane per
//                   chunk of 1..16 words of a line, every LDS prime a
//                   64-bit window of its kill bitmap
//   k_round_c/_p    per-candidate kill tests for one range of large primes:
//                   lane = candidate (the prime uniform across the wave, its
//                   constants and kill set in scalar registers) or lane =
//                   prime, the residue formed exactly in double precision;
//                   killed candidates flagged
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
// the shared line model and the on-device classification have no
// counterpart there.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <type_traits>
#include <mutex>
#include <vector>

#include "fk_device.h"
#include "fk_kill.h"
#include "fk_launch.h"

namespace fk {

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

// Four consecutive words through a uniform view (one scalar x4 load).
__device__ __forceinline__ uint4 load4(const FK_CONST uint32_t* p)
{
    return make_uint4(p[0], p[1], p[2], p[3]);
}

// ------------------------------------------------------------------ prep
// ----------------------------------------------------------------- sieve

__device__ __forceinline__ uint32_t find_piece(const FK_CONST DevPiece* pieces,
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

// B mod q for a piece's base B (four 32-bit limbs, most significant first)
__device__ __forceinline__ uint32_t base_mod(const FK_CONST DevPiece& P,
                                             uint32_t q, uint64_t m64)
{
    uint32_t r = red64(P.base[3], q, m64);
    r = red64(((uint64_t)r << 32) | P.base[2], q, m64);
    r = red64(((uint64_t)r << 32) | P.base[1], q, m64);
    return red64(((uint64_t)r << 32) | P.base[0], q, m64);
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
//   2b. bitmap stage: primes [nmark, nbits) OR in 64-bit windows of their
//      periodic kill bitmaps
//   3. marking stage: primes [npat, nmark) stride-marked with ds_or
//   4. test stage: the positions still alive are listed in LDS and run
//      through in-tile rounds over the primes [nbits, nlds); in a round
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
                                            const FK_CONST DevPiece& P,
                                            const FK_CONST DevEngine& E,
                                            unsigned __int128 B, uint64_t off,
                                            uint64_t v, uint32_t slot)
{
    unsigned __int128 x = B + off + (unsigned __int128)E.D * v;
    if (E.family == kDickson) x *= E.W;
    const uint2 seg = a.segs[P.queue];
    if (slot < seg.y) {
        a.flags[seg.x + slot] = 0;        // live for the first round
        DevRec r;
        r.t[0] = (uint32_t)x;
        r.t[1] = (uint32_t)(x >> 32);
        r.t[2] = (uint32_t)(x >> 64);
        r.t[3] = (uint32_t)(x >> 96);
        r.window = P.window;
        r.run = -1;
        a.out[seg.x + slot] = r;
    }
}

// Does one of the LDS test primes [p, pend) kill tile position pos?  Each
// prime's kill bitmap is read at (s0 + pos) mod q, in order, stopping at
// the first kill.
__device__ __forceinline__ bool tile_kills(const uint4* pinfo,
                                           const uint32_t* s0,
                                           const uint32_t* bm, uint32_t p,
                                           uint32_t pend, uint32_t pos)
{
    for (uint32_t pp = p; pp < pend; pp++) {
        const uint4 pq = pinfo[pp];
        const uint32_t x = red32(s0[pp] + pos, pq.x, pq.y);
        if ((bm[pq.z + (x >> 5)] >> (x & 31)) & 1) return true;
    }
    return false;
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

    // a contiguous run of tiles per workgroup: consecutive tiles are
    // consecutive lines of one window, so the engine tables loaded into LDS
    // change only at window boundaries
    const uint64_t tiles_per =
        ((uint64_t)(a.ntiles - a.tile_begin) + gridDim.x - 1) / gridDim.x;
    const uint64_t tile_lo = a.tile_begin + (uint64_t)blockIdx.x * tiles_per;
    const uint64_t tile_hi = min((uint64_t)a.ntiles, tile_lo + tiles_per);
    for (uint32_t tile = (uint32_t)tile_lo; tile < tile_hi; tile++) {
        // the tile's piece and engine are uniform: scalar loads, cached
        const uint32_t pi = find_piece(uniform_view(a.pieces), a.npieces, tile);
        const FK_CONST DevPiece& P = uniform_view(a.pieces)[pi];
        const FK_CONST DevEngine& E = uniform_view(a.engines)[P.engine];
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
            // per LDS prime: q, 2^32/q, its kill bitmap's first word (test
            // primes) and its pattern table's first word (pattern primes)
            for (uint32_t p = tid; p < E.nlds; p += T)
                pinfo[p] = make_uint4(E.lds_q[p], E.lds_m32[p],
                                      p >= E.nmark ? E.bm_off[p - E.nmark] : 0,
                                      p < E.npat ? E.pat_off[p] : 0);
            cur_engine = P.engine;
        }
        // s0[p] = (v0 + t_p) mod q with t_p = (B + off) D^-1 mod q
        for (uint32_t p = tid; p < E.nlds; p += T) {
            const uint32_t q = E.lds_q[p];
            const uint64_t m64 = E.lds_m64[p];
            uint32_t t = base_mod(P, q, m64) + red64(off, q, m64);
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
        // a thread whose first word is past the line's end has only dead
        // words (edge_mask); short lines leave whole waves idle here
        const bool words_here = v0 + 64ull * tid < vend;
        const uint32_t npat_here = words_here ? E.npat : 0u;
        for (uint32_t p = 0; p < npat_here; p++) {
            const uint4 pq = pinfo[p];
            const uint32_t q = pq.x, m32 = pq.y;
            const uint32_t step = red32(64u * T, q, m32);
            const uint64_t* tab = pat + pq.w;
            uint32_t s = red32(s0[p] + 64u * tid, q, m32);
#pragma unroll
            for (int k = 0; k < WPT; k++) {
                acc[k] |= tab[s];
                s += step;
                if (s >= q) s -= q;
            }
        }
        // ---- 2b. bitmap windows: the 64 kill bits of word w are bits
        // [x, x + 64) of the prime's periodic bitmap, x = (s0 + 64 w) mod q,
        // funnel-shifted out of three consecutive words
        const uint32_t nbits_here = words_here ? E.nbits : E.nmark;
        for (uint32_t p = E.nmark; p < nbits_here; p++) {
            const uint4 pq = pinfo[p];
            const uint32_t q = pq.x, m32 = pq.y;
            const uint32_t step = red32(64u * T, q, m32);
            const uint32_t* b = bm + pq.z;
            uint32_t x = red32(s0[p] + 64u * tid, q, m32);
#pragma unroll
            for (int k = 0; k < WPT; k++) {
                const uint32_t i = x >> 5, o = x & 31u;
                const uint32_t w0 = b[i], w1 = b[i + 1], w2 = b[i + 2];
                const uint32_t lo = __builtin_amdgcn_alignbit(w1, w0, o);
                const uint32_t hi = __builtin_amdgcn_alignbit(w2, w1, o);
                acc[k] |= ((uint64_t)hi << 32) | lo;
                x += step;
                if (x >= q) x -= q;
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
        if (E.nbits < E.nlds) {
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
                uint32_t p = E.nbits;
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
                            alive = !tile_kills(pinfo, s0, bm, p, pend, pos);
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
                const uint32_t p0 = E.nbits, p1 = E.nlds;
                for (;;) {
                    const uint32_t w = atomicAdd(&sh_next, 1u);
                    if (w >= nwords) break;
                    uint64_t lv = ~(((uint64_t)bits[2 * w + 1] << 32) | bits[2 * w]);
                    uint64_t dead = 0, m = lv;
                    while (m) {
                        const int b = __builtin_ctzll(m);
                        m &= m - 1;
                        const uint32_t pos = 64u * w + b;
                        if (tile_kills(pinfo, s0, bm, p0, p1, pos))
                            dead |= 1ull << b;
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
            if (tid == 0 && n) sh_base = atomicAdd(a.counts + P.queue, n);
            __syncthreads();
            for (uint32_t i = tid; i < n; i += T)
                emit_record(a, P, E, B, off, v0 + cur[i], sh_base + i);
        } else {
            uint32_t mine = 0;
            if (cnt) mine = atomicAdd(&sh_count, cnt);
            __syncthreads();
            if (tid == 0 && sh_count)
                sh_base = atomicAdd(a.counts + P.queue, sh_count);
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

// ---------------------------------------------------- sieve, short lines

// Windows whose lines hold only a few 64-position words (a short window
// over a large wheel): one lane per chunk of C consecutive words of one
// line instead of one workgroup per line.  A tile here is 256 consecutive
// (line, chunk) items of one piece, lines major.  For every LDS prime the
// lane forms its first word's position x in the prime's kill coordinates
// (the work the line kernel does once per tile); its C words are then the
// 64-bit windows at x, x + 64, ... of the prime's periodic kill bitmap,
// which repeats its period through kShortTailBits bits past q: 2C + 1
// consecutive LDS words and two funnel shifts per word, with no reduction
// mod q.  Bitmaps for every prime (no 64-bit pattern tables, which take 64
// times the LDS) keep the workgroup's LDS to a few tens of KB, so that
// many workgroups share a CU.  Surviving bits become records with one
// atomic per wave.  No survivor lists; barriers only where the tile's
// piece changes.

// The chunk's C words ORed with every LDS prime's bitmap windows.  x is
// (v0 + (B + off) D^-1) mod q, formed in double precision: with
// bqd = (B D^-1) mod q from LDS (one broadcast word per prime),
// s = off D^-1 + bqd + v0 is an integer congruent to x, and x = s mod q by
// resid_d1.  For D <= 2^32, off D^-1 < 2^32 2^16 (LDS primes are below
// 2^16), so s < 2^49 and one fma forms it exactly.  WIDE periods split off
// into 32-bit halves: s = off_hi e32 + off_lo D^-1 + bqd + v0 with
// e32 = (2^32 D^-1) mod q, below 2^47 + 2^48 + 2^33 < 2^49, two exact fmas.
// Either way s < 2^49 meets resid_d1's condition for every q >= 2: five or
// six double operations and a conversion per prime, where 32-bit Barrett
// reductions took three reductions and a multiply.  The prime's constants
// come from the engine's table with scalar loads (the index is
// wave-uniform): LDS bandwidth goes to the bitmap windows.
template <int C, bool WIDE>
__device__ __forceinline__ void short_words(const FK_CONST uint32_t* sp,
                                            const FK_CONST double* se,
                                            const double* bqd,
                                            const uint32_t* bm, uint32_t nlds,
                                            double offlo, double offhi,
                                            double v0d, uint64_t (&acc)[C])
{
    static_assert(C <= (int)kShortChunkMax, "chunk past the bitmap tail");
    // A wave walks the primes as one chain of dependent latencies (scalar
    // load, residue, LDS reads); unrolled, several primes' chains are in
    // flight at once.  The workgroups' LDS already caps a SIMD at a few
    // waves, so the registers this takes are free (the wide form, for
    // wheels past 2^32, is not unrolled, and sixteen-word chunks unroll by
    // three: more would take the kernel past 128 registers, three waves
    // per SIMD instead of four).
    constexpr int kUnroll = WIDE ? 1 : (C >= 16 ? 3 : 4);
#pragma unroll kUnroll
    for (uint32_t p = 0; p < nlds; p++) {
        // {q, 1/q, D^-1} as doubles, q, bitmap word
        const FK_CONST uint32_t* k = sp + kShortPrimeWords * p;
        const uint4 c0 = load4(k);
        const uint4 c1 = load4(k + 4);
        const double qd = __hiloint2double((int)c0.y, (int)c0.x);
        const double qi = __hiloint2double((int)c0.w, (int)c0.z);
        const double di = __hiloint2double((int)c1.y, (int)c1.x);
        double t = bqd[p] + v0d;
        if (WIDE) t = fma(offhi, se[p], t);
        const uint32_t x = resid_d1(fma(offlo, di, t), qd, qi, c1.z);
        const uint32_t* b = bm + c1.w + (x >> 5);
        const uint32_t o = x & 31u;
        uint32_t w[2 * C + 1];
#pragma unroll
        for (int j = 0; j < 2 * C + 1; j++) w[j] = b[j];
#pragma unroll
        for (int k = 0; k < C; k++)
            acc[k] |= ((uint64_t)__builtin_amdgcn_alignbit(w[2 * k + 2], w[2 * k + 1], o)
                       << 32) |
                      __builtin_amdgcn_alignbit(w[2 * k + 1], w[2 * k], o);
    }
}

// A workgroup gathers its tiles' survivors, as (line, position) pairs, in
// LDS and reserves their queue slots with one atomic per flush of the
// buffer (once it is half full, when the piece changes, and at the end),
// instead of one atomic per wave per tile: on a wide wheel with short lines
// (the Euler SCORE range: 2.3M tiles, nine million atomics on one counter)
// those serialised on the counter's L2 line and took most of the launch.
// A tile of more survivors than the buffer has room for reserves its own
// slots, one atomic per wave.
constexpr uint32_t kShortBuf = 512;

template <int C>
__global__ __launch_bounds__(256) void k_sieve_short(SieveArgs a)
{
    extern __shared__ uint64_t lds64[];
    __shared__ uint2 sbuf[kShortBuf];
    __shared__ uint32_t sh_wtot[2][4];      // per wave, tiles alternating
    __shared__ uint32_t sh_base;
    // [nlds_max] (B D^-1) mod q per LDS prime as a double, then the bitmaps
    double* bqd = reinterpret_cast<double*>(lds64);
    uint32_t* bm = reinterpret_cast<uint32_t*>(lds64 + ((a.nlds_max + 1) & ~1u));
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = __lane_id();
    const uint32_t wave = tid >> 6;
    uint32_t cur_engine = 0xffffffffu, cur_piece = 0xffffffffu;
    // entries in the buffer: every thread keeps the same count
    uint32_t nbuf = 0;
    // the buffered survivors of piece pb out to the queue (every thread of
    // the workgroup calls this at the same point)
    auto flush = [&](uint32_t pb) {
        const uint32_t n = nbuf;
        if (n == 0) return;
        nbuf = 0;
        __syncthreads();                            // entries complete
        const FK_CONST DevPiece& Q = uniform_view(a.pieces)[pb];
        const FK_CONST DevEngine& F = uniform_view(a.engines)[Q.engine];
        if (tid == 0) sh_base = atomicAdd(a.counts + Q.queue, n);
        __syncthreads();
        const uint32_t base = sh_base;
        const unsigned __int128 B =
            ((unsigned __int128)(((uint64_t)Q.base[3] << 32) | Q.base[2]) << 64)
            | (((uint64_t)Q.base[1] << 32) | Q.base[0]);
        for (uint32_t i = tid; i < n; i += 256) {
            const uint2 e = sbuf[i];
            emit_record(a, Q, F, B, F.offs[e.x], e.y, base + i);
        }
        __syncthreads();                            // entries consumed
    };
    const uint64_t tiles_per =
        ((uint64_t)(a.ntiles - a.tile_begin) + gridDim.x - 1) / gridDim.x;
    const uint64_t tile_lo = a.tile_begin + (uint64_t)blockIdx.x * tiles_per;
    const uint64_t tile_hi = min((uint64_t)a.ntiles, tile_lo + tiles_per);
    for (uint32_t tile = (uint32_t)tile_lo; tile < tile_hi; tile++) {
        const uint32_t pi = find_piece(uniform_view(a.pieces), a.npieces, tile);
        const FK_CONST DevPiece& P = uniform_view(a.pieces)[pi];
        const FK_CONST DevEngine& E = uniform_view(a.engines)[P.engine];
        // the buffer is emptied between tiles, while no words are live (a
        // flush beside them took the sixteen-word kernel past 128
        // registers): when half full, and when the piece changes
        if (pi == cur_piece && nbuf >= kShortBuf / 2) flush(pi);
        if (pi != cur_piece) {
            // the last piece's survivors out; the engine's tables when the
            // engine changes, the piece's B mod q always
            if (cur_piece != 0xffffffffu) flush(cur_piece);
            __syncthreads();
            if (P.engine != cur_engine) {
                // the bitmaps in 16-byte pieces, eight loads in flight per
                // thread (one workgroup usually sieves one tile: this copy
                // is its prologue)
                const uint32_t n4 = E.sbm_words / 4;
                const uint4* src = reinterpret_cast<const uint4*>(E.sbm);
                uint4* dst = reinterpret_cast<uint4*>(bm);
#pragma unroll 8
                for (uint32_t w = tid; w < n4; w += 256) dst[w] = src[w];
                for (uint32_t w = 4 * n4 + tid; w < E.sbm_words; w += 256)
                    bm[w] = E.sbm[w];
            }
            for (uint32_t p = tid; p < E.nlds; p += 256) {
                const uint32_t q = E.lds_q[p];
                const uint32_t bq = base_mod(P, q, E.lds_m64[p]);
                bqd[p] = (double)((bq * E.lds_dinv[p]) % q);   // < 2^32
            }
            cur_engine = P.engine;
            cur_piece = pi;
            __syncthreads();
        }
        const uint32_t nch = (P.npw + C - 1) / C;
        const uint32_t item = (tile - P.tile0) * 256u + tid;
        const uint32_t line = item / nch;
        const uint32_t v0 = 64u * C * (item - line * nch);
        // the item's valid positions [vlo, vend); none (every word dead)
        // for an item past its piece or outside the window.  The words
        // start clear and take their edge masks after the sieve: an
        // all-dead start for the items with no positions, merged with the
        // edge masks before it, kept a second copy of the C words in
        // registers.
        uint64_t off = 0, vlo = 0, vend = 0;
        if (line < E.nlines && v0 < P.npos) {
            off = E.offs[line];
            const uint64_t D = E.D;
            const uint64_t lo =
                (off >= P.lo_rel) ? 0 : ceil_div_u64(P.lo_rel - off, D);
            const uint64_t vhi =
                (P.hi_rel > off) ? ceil_div_u64(P.hi_rel - off, D) : 0;
            const uint64_t end = (vhi < (uint64_t)P.npos) ? vhi : (uint64_t)P.npos;
            if (lo < end && v0 < end && v0 + 64 * C > lo) {
                vlo = lo;
                vend = end;
            }
        }
        const bool here = vlo < vend;
        uint64_t acc[C];
#pragma unroll
        for (int k = 0; k < C; k++) acc[k] = 0;
        if (here) {
            const FK_CONST uint32_t* sp = uniform_view(E.spd);
            const FK_CONST double* se = uniform_view(E.spe);
            if ((E.D >> 32) == 0)
                short_words<C, false>(sp, se, bqd, bm, E.nlds, (double)off,
                                      0.0, (double)v0, acc);
            else
                short_words<C, true>(sp, se, bqd, bm, E.nlds,
                                     (double)(uint32_t)off,
                                     (double)(off >> 32), (double)v0, acc);
        }
        // records: a wave-wide prefix of the per-lane counts, then the
        // workgroup's: into the LDS buffer, or straight out for a tile that
        // overflows it
        uint32_t c = 0;
#pragma unroll
        for (int k = 0; k < C; k++) {
            acc[k] = ~(acc[k] | edge_mask(v0 + 64u * k, vlo, vend));  // live
            c += __popcll(acc[k]);
        }
        uint32_t incl = c;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t o = __shfl_up(incl, d);
            if (lane >= (uint32_t)d) incl += o;
        }
        const uint32_t total = __shfl(incl, 63);
        // (a wave writes the other tile parity's slots next, after this
        // tile's barrier, which every wave passes only once it has read
        // these)
        const uint32_t par = tile & 1u;
        if (lane == 63) sh_wtot[par][wave] = total;
        __syncthreads();
        uint32_t woff = 0, ttot = 0;
#pragma unroll
        for (uint32_t w = 0; w < 4; w++) {
            const uint32_t t = sh_wtot[par][w];
            woff += w < wave ? t : 0u;
            ttot += t;
        }
        if (nbuf + ttot <= kShortBuf) {
            uint32_t at = nbuf + woff + incl - c;
#pragma unroll
            for (int k = 0; k < C; k++) {
                uint64_t m = acc[k];
                while (m) {
                    const int bit = __builtin_ctzll(m);
                    m &= m - 1;
                    sbuf[at++] = make_uint2(line, v0 + 64u * k + (uint32_t)bit);
                }
            }
            nbuf += ttot;
        } else {
            uint32_t base = 0;
            if (total) {
                if (lane == 63) base = atomicAdd(a.counts + P.queue, total);
                base = __shfl(base, 63);
            }
            if (c) {
                const unsigned __int128 B =
                    ((unsigned __int128)(((uint64_t)P.base[3] << 32) | P.base[2]) << 64)
                    | (((uint64_t)P.base[1] << 32) | P.base[0]);
                uint32_t slot = base + incl - c;
#pragma unroll
                for (int k = 0; k < C; k++) {
                    uint64_t m = acc[k];
                    while (m) {
                        const int bit = __builtin_ctzll(m);
                        m &= m - 1;
                        emit_record(a, P, E, B, off, (uint64_t)v0 + 64u * k + bit,
                                    slot++);
                    }
                }
            }
        }
    }
    if (cur_piece != 0xffffffffu) flush(cur_piece);
}

// ---------------------------------------------------------------- rounds

// Round table of each engine, built once at engine creation: one thread
// per round prime writes its RoundRow and kill set, or its IndexRow.
__global__ void k_kill_table(const PrimeK* __restrict__ primes, uint32_t p0,
                             uint32_t np, int family, int n, uint4 d,
                             uint32_t index, uint32_t stride,
                             uint32_t* __restrict__ kt)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const uint32_t dl[4] = {d.x, d.y, d.z, d.w};
    if (index)
        index_row(n, dl, primes[p0 + i], kt + (size_t)i * stride);
    else
        round_row(family, n, dl, primes[p0 + i], kt + (size_t)i * stride, stride);
}


// Global block b of a stage, in blocks of `per` entries of each engine's
// segment, engine by engine: the engine, the block within it, the engine's
// entry count and the entries of the engines before it.  False past the
// last block.  The loop is over the batch's engines (a few dozen), with
// wave-uniform scalar loads.
struct SegBlock {
    uint32_t e, lb, n, before;
};

__device__ __forceinline__ bool seg_block(const FK_CONST uint2* segs,
                                          const FK_CONST uint32_t* counts,
                                          uint32_t ne, uint32_t per,
                                          uint32_t b, SegBlock& out)
{
    uint32_t acc = 0, before = 0;
    for (uint32_t e = 0; e < ne; e++) {
        const uint32_t n = min(counts[e], segs[e].y);
        const uint32_t nb = (n + per - 1) / per;
        if (b < acc + nb) {
            out.e = e;
            out.lb = b - acc;
            out.n = n;
            out.before = before;
            return true;
        }
        acc += nb;
        before += n;
    }
    return false;
}

__device__ __forceinline__ double load_f64(const FK_CONST uint32_t* p)
{
    return __hiloint2double((int)p[1], (int)p[0]);
}

// ---- lane = candidate (many candidates, few primes): the prime is
// wave-uniform, its row is read with scalar loads into SGPRs.

// One round prime's row as scalar values: the header and the first 4 KG
// killing residues (12 + 4 KG words; a table carries kRowTail words past its
// last row, so reading a fixed size never leaves it).
template <int KG>
struct RowView {
    double qd, qi, c16, c32, c48;
    uint32_t q, cnt;
    uint4 k[KG];
};

template <int KG>
__device__ __forceinline__ RowView<KG> load_row(const FK_CONST uint32_t* row)
{
    RowView<KG> v;
    v.qd = load_f64(row + 0);
    v.qi = load_f64(row + 2);
    v.c16 = load_f64(row + 4);
    v.c32 = load_f64(row + 6);
    v.c48 = load_f64(row + 8);
    v.q = row[10];
    v.cnt = row[11];
#pragma unroll
    for (int i = 0; i < KG; i++) v.k[i] = load4(row + kRowHead + 4 * i);
    return v;
}

// The lane's candidate residue mod the row's prime from its 16-bit limbs;
// limbs past the fourth take 2^64..2^112 mod q from the prime table.  (This
// kernel waits on its scalar row loads, not on VALU: feeding a value below
// 2^53 to the residue directly, as the lane = prime kernel does, measured
// no faster here.)
// NL == 22: a[0..3] are 22-bit limbs (values below 2^88, primes below
// 2^24), against 2^22, 2^44 and 2^66 mod q from the prime table: three
// FMAs where 16-bit limbs past 2^64 take seven.
template <int NL, bool FOLD, int KG>
__device__ __forceinline__ uint32_t row_residue(const double (&a)[8],
                                                const RowView<KG>& v,
                                                const FK_CONST PrimeD* P)
{
    double s;
    if (NL == 22)
        s = fma(a[3], P->c22[2], fma(a[2], P->c22[1], fma(a[1], P->c22[0], a[0])));
    else
        s = fma(a[3], v.c48, fma(a[2], v.c32, fma(a[1], v.c16, a[0])));
    if (NL > 4 && NL != 22) {
#pragma unroll
        for (int k = 4; k < NL; k++) s = fma(a[k], P->c[k - 1], s);
    }
    // folded rows (Dickson) compare min(r, q - r), formed directly
    return FOLD ? resid_fold_d(s, v.qd, v.qi, v.q) : resid_d1(s, v.qd, v.qi, v.q);
}

// The lane's residue against the row's kill set, kept in a per-lane
// minimum of r ^ k that falls to zero on a match (see narrow_test): only
// the four-residue groups the row's count reaches (nested uniform
// branches), the first 4 KG residues from registers, longer rows in a
// loop.  (Folded rows get a folded residue from row_residue.)
__device__ __forceinline__ uint32_t min4x(uint32_t acc, uint32_t r, uint4 k)
{
    acc = min(min(acc, r ^ k.x), r ^ k.y);         // v_min3_u32
    return min(min(acc, r ^ k.z), r ^ k.w);
}

template <int KG>
__device__ __forceinline__ uint32_t row_min(uint32_t acc, uint32_t r,
                                            const RowView<KG>& v,
                                            const FK_CONST uint32_t* row)
{
    acc = min4x(acc, r, v.k[0]);
    if (v.cnt > 4u) {
        acc = min4x(acc, r, v.k[1]);
        if (KG > 2 && v.cnt > 8u) {
            acc = min4x(acc, r, v.k[2]);
#pragma unroll
            for (int g = 3; g < KG; g++)
                if (v.cnt > 4u * g) acc = min4x(acc, r, v.k[g]);
        }
    }
    if (v.cnt > 4u * KG) {
        for (uint32_t c = 4u * KG; c < v.cnt; c += 4)
            acc = min4x(acc, r, load4(row + kRowHead + c));
    }
    return acc;
}

// KC candidates per lane against the primes [lo, hi), two primes per pass;
// hit[c] collects the lanes whose c-th candidate some prime kills.
template <int NL, int KG, bool FOLD, int KC>
__device__ __forceinline__ void round_scan(const double (&a)[KC][8],
                                           const FK_CONST PrimeD* pd,
                                           const FK_CONST uint32_t* kt,
                                           uint32_t stride, uint32_t rf,
                                           uint32_t lo, uint32_t hi,
                                           uint64_t (&hit)[KC])
{
    uint32_t acc[KC];
#pragma unroll
    for (int c = 0; c < KC; c++) acc[c] = 0xffffffffu;
    uint32_t p = lo;
    for (; p + 1 < hi; p += 2) {
        const FK_CONST uint32_t* r0 = kt + (size_t)(p - rf) * stride;
        const FK_CONST uint32_t* r1 = r0 + stride;
        const RowView<KG> v0 = load_row<KG>(r0);
        const RowView<KG> v1 = load_row<KG>(r1);
        if (KG == 3) {
            // both rows in scalar registers before any use: one wait per
            // pass instead of one per residue group (the compiler would
            // otherwise sink each group's load into its count branch)
            asm volatile("" ::"s"(v0.k[1].x), "s"(v0.k[1].y), "s"(v0.k[1].z),
                         "s"(v0.k[1].w), "s"(v0.k[2].x), "s"(v0.k[2].y),
                         "s"(v0.k[2].z), "s"(v0.k[2].w), "s"(v1.q), "s"(v1.cnt),
                         "s"(v1.k[1].x), "s"(v1.k[1].y), "s"(v1.k[1].z),
                         "s"(v1.k[1].w), "s"(v1.k[2].x), "s"(v1.k[2].y),
                         "s"(v1.k[2].z), "s"(v1.k[2].w));
        }
        uint32_t x0[KC], x1[KC];
#pragma unroll
        for (int c = 0; c < KC; c++) {
            x0[c] = row_residue<NL, FOLD>(a[c], v0, pd + p);
            x1[c] = row_residue<NL, FOLD>(a[c], v1, pd + p + 1);
        }
#pragma unroll
        for (int c = 0; c < KC; c++)
            acc[c] = row_min<KG>(row_min<KG>(acc[c], x0[c], v0, r0), x1[c], v1, r1);
    }
    if (p < hi) {
        const FK_CONST uint32_t* r0 = kt + (size_t)(p - rf) * stride;
        const RowView<KG> v0 = load_row<KG>(r0);
#pragma unroll
        for (int c = 0; c < KC; c++)
            acc[c] = row_min<KG>(acc[c], row_residue<NL, FOLD>(a[c], v0, pd + p), v0, r0);
    }
#pragma unroll
    for (int c = 0; c < KC; c++) hit[c] |= __ballot(acc[c] == 0u);
}

// ---- Euler rows of primes above (n - 1) n.  q kills p when p + x(x + 1)
// is 0 mod q for some x < n, i.e. when u = (-p) mod q equals some
// x(x + 1) <= (n - 1) n < q (a residue below q equals such a product only
// as an integer).  The products x(x + 1) are exactly the u with 4u + 1 an
// odd square, so these rows are tested without their residue lists: u is
// clamped to (n - 1) n + 1 (odd, never such a product) and 4u + 1 < 4 n^2
// is tested for a square with the hardware square root (within an ulp:
// rounded to the nearest integer it is the exact root of a square).  About
// ten VALU per test where the row's n residues take 1.5 per residue.  The
// result is zero when q kills.
__device__ __forceinline__ uint32_t pronic_d(uint32_t r, uint32_t q, uint32_t pmax)
{
    const uint32_t u = min(q - r, pmax + 1u);      // r = 0: u = q, clamped
    const uint32_t w = 4u * u + 1u;
    const uint32_t t = (uint32_t)(__builtin_amdgcn_sqrtf((float)w) + 0.5f);
    const uint32_t d = (t * t) ^ w;
    return r == 0u ? 0u : d;                       // x = 0: q divides p
}

// KC limb sets per lane against the rows [lo, hi) by pronic_d.
template <int NL, int KC>
__device__ __forceinline__ void pronic_scan(const double (&a)[KC][8],
                                            const FK_CONST PrimeD* pd,
                                            const FK_CONST uint32_t* kt,
                                            uint32_t stride, uint32_t rf,
                                            uint32_t lo, uint32_t hi,
                                            uint32_t pmax, uint64_t (&hit)[KC])
{
    uint32_t acc[KC];
#pragma unroll
    for (int c = 0; c < KC; c++) acc[c] = 0xffffffffu;
#pragma unroll 2
    for (uint32_t p = lo; p < hi; p++) {
        const RowView<1> v = load_row<1>(kt + (size_t)(p - rf) * stride);
#pragma unroll
        for (int c = 0; c < KC; c++)
            acc[c] = min(acc[c], pronic_d(row_residue<NL, false>(a[c], v, pd + p),
                                          v.q, pmax));
    }
#pragma unroll
    for (int c = 0; c < KC; c++) hit[c] |= __ballot(acc[c] == 0u);
}

// ---- values below 2^53 (every live lane of the wave): the residue comes
// from the value itself, one exact double, with no limb sum (resid_d1 and
// resid_fold_d hold for s / q < 2^48: values below 2^53, primes of at least
// 32), and a row needs only q, 1/q, the count and the residues.
//
// The loop is bound by instruction issue (VALU and SALU in about equal
// number per test) more than by the row loads, which hit the scalar cache
// after the block's touch; what the loads cost is the wait each one adds
// when the compiler sinks it into the count branch that uses it.  So the
// first 24 words of both rows of a pass (header and twelve residues) are
// loaded by one asm block that issues the four loads back to back and
// waits for them there (no load is left in flight past the block, and the
// outputs never share registers with the addresses).  A table carries
// kRowTail (32) words past its last row, so 24 words from any row start
// stay inside it; words past a row's own stride are never compared, since
// only residues below the row's count in groups of four are.
typedef uint32_t u32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x8 __attribute__((ext_vector_type(8)));

struct NarrowRow {
    double qd, qi;
    uint32_t q, cnt;
    uint32_t k[12];
};

__device__ __forceinline__ NarrowRow narrow_row(const u32x16& a, const u32x8& b)
{
    NarrowRow v;
    v.qd = __hiloint2double((int)a[1], (int)a[0]);
    v.qi = __hiloint2double((int)a[3], (int)a[2]);
    v.q = a[10];
    v.cnt = a[11];
#pragma unroll
    for (int i = 0; i < 4; i++) v.k[i] = a[12 + i];
#pragma unroll
    for (int i = 0; i < 8; i++) v.k[4 + i] = b[i];
    return v;
}

// Both rows of a pass, loaded together (and one row, for an odd count).
__device__ __forceinline__ void load_narrow_pair(const FK_CONST uint32_t* r0,
                                                 const FK_CONST uint32_t* r1,
                                                 NarrowRow& v0, NarrowRow& v1)
{
    static_assert(kRowTail >= 24, "a pass reads 24 words from a row start");
    u32x16 a0, a1;
    u32x8 b0, b1;
    asm volatile("s_load_dwordx16 %0, %4, 0x0\n\t"
                 "s_load_dwordx8 %1, %4, 0x40\n\t"
                 "s_load_dwordx16 %2, %5, 0x0\n\t"
                 "s_load_dwordx8 %3, %5, 0x40\n\t"
                 "s_waitcnt lgkmcnt(0)"
                 : "=&s"(a0), "=&s"(b0), "=&s"(a1), "=&s"(b1)
                 : "s"(r0), "s"(r1));
    v0 = narrow_row(a0, b0);
    v1 = narrow_row(a1, b1);
}

__device__ __forceinline__ NarrowRow load_narrow_one(const FK_CONST uint32_t* r0)
{
    u32x16 a0;
    u32x8 b0;
    asm volatile("s_load_dwordx16 %0, %2, 0x0\n\t"
                 "s_load_dwordx8 %1, %2, 0x40\n\t"
                 "s_waitcnt lgkmcnt(0)"
                 : "=&s"(a0), "=&s"(b0)
                 : "s"(r0));
    return narrow_row(a0, b0);
}

// The lane's residue (folded for folded rows) against the row's residues,
// kept in a per-lane minimum of r ^ k that falls to zero when r is one of
// them (padding kKillPad never matches a residue below 2^30): exclusive
// ors and three-way minima on the vector unit, where ballots would cost a
// scalar OR per compare and the scalar unit was the busier of the two.
// Groups of four below the count, nested, so a row of at most four
// residues costs one scalar compare and branch.
__device__ __forceinline__ uint32_t min4x(uint32_t acc, uint32_t r, uint32_t a,
                                         uint32_t b, uint32_t c, uint32_t d)
{
    return min4x(acc, r, make_uint4(a, b, c, d));
}

template <bool FOLD>
__device__ __forceinline__ uint32_t narrow_test(double s, const NarrowRow& v,
                                                const FK_CONST uint32_t* row,
                                                uint32_t acc)
{
    const uint32_t r = FOLD ? resid_fold_d(s, v.qd, v.qi, v.q)
                            : resid_d1(s, v.qd, v.qi, v.q);
    acc = min4x(acc, r, v.k[0], v.k[1], v.k[2], v.k[3]);
    if (v.cnt > 4u) {
        acc = min4x(acc, r, v.k[4], v.k[5], v.k[6], v.k[7]);
        if (v.cnt > 8u) {
            acc = min4x(acc, r, v.k[8], v.k[9], v.k[10], v.k[11]);
            for (uint32_t c = 12; c < v.cnt; c++) acc = min(acc, r ^ row[kRowHead + c]);
        }
    }
    return acc;
}

// KC values per lane against the rows [lo, hi), two per pass; hit[c]
// gets the lanes whose c-th value some row kills.
template <bool FOLD, int KC>
__device__ __forceinline__ void narrow_scan(const double (&s)[KC],
                                            const FK_CONST uint32_t* kt,
                                            uint32_t stride, uint32_t rf,
                                            uint32_t lo, uint32_t hi,
                                            uint64_t (&hit)[KC])
{
    uint32_t acc[KC];
#pragma unroll
    for (int c = 0; c < KC; c++) acc[c] = 0xffffffffu;
    // running row pointers: one 64-bit add per pass, not a product
    const FK_CONST uint32_t* r0 = kt + (size_t)(lo - rf) * stride;
    const size_t step = 2 * (size_t)stride;
    uint32_t p = lo;
    for (; p + 2 <= hi; p += 2, r0 += step) {
        const FK_CONST uint32_t* r1 = r0 + stride;
        NarrowRow v0, v1;
        load_narrow_pair(r0, r1, v0, v1);
#pragma unroll
        for (int c = 0; c < KC; c++)
            acc[c] = narrow_test<FOLD>(s[c], v1, r1, narrow_test<FOLD>(s[c], v0, r0, acc[c]));
    }
    if (p < hi) {
        const NarrowRow v0 = load_narrow_one(r0);
#pragma unroll
        for (int c = 0; c < KC; c++) acc[c] = narrow_test<FOLD>(s[c], v0, r0, acc[c]);
    }
#pragma unroll
    for (int c = 0; c < KC; c++) hit[c] |= __ballot(acc[c] == 0u);
}

// The same values against Euler rows of primes above (n - 1) n (pronic_d).
template <int KC>
__device__ __forceinline__ void narrow_pronic(const double (&s)[KC],
                                              const FK_CONST uint32_t* kt,
                                              uint32_t stride, uint32_t rf,
                                              uint32_t lo, uint32_t hi,
                                              uint32_t pmax, uint64_t (&hit)[KC])
{
    uint32_t acc[KC];
#pragma unroll
    for (int c = 0; c < KC; c++) acc[c] = 0xffffffffu;
    const FK_CONST uint32_t* r0 = kt + (size_t)(lo - rf) * stride;
#pragma unroll 2
    for (uint32_t p = lo; p < hi; p++, r0 += stride) {
        const double qd = load_f64(r0), qi = load_f64(r0 + 2);
        const uint32_t q = r0[10];
#pragma unroll
        for (int c = 0; c < KC; c++)
            acc[c] = min(acc[c], pronic_d(resid_d1(s[c], qd, qi, q), q, pmax));
    }
#pragma unroll
    for (int c = 0; c < KC; c++) hit[c] |= __ballot(acc[c] == 0u);
}

// ---- index rows (AP): the lane's t = -p / d mod q from the row's linear
// form in p's 16-bit limbs (fk_kill.h IndexRow); q kills the lanes with
// t < n.  One compare per (candidate, prime) instead of one per residue.
struct IdxView {
    double qd, qi, u0, u1, u2, u3;
    uint32_t q, n;
};

__device__ __forceinline__ IdxView load_idx(const FK_CONST uint32_t* row)
{
    IdxView v;
    v.qd = load_f64(row + 0);
    v.qi = load_f64(row + 2);
    v.u0 = load_f64(row + kIndexU);
    v.u1 = load_f64(row + kIndexU + 2);
    v.u2 = load_f64(row + kIndexU + 4);
    v.u3 = load_f64(row + kIndexU + 6);
    v.q = row[12];
    v.n = row[13];
    return v;
}

// The lane's u for the row: s = sum a_k u_k (each product below 2^46 for
// q < 2^30, the sum of eight below 2^49, all exact in a double), then u =
// s - floor(s/q - 1/8) q, exact and in [0, 2q) as in resid_d1 (s / q <
// 2^19).  t = s mod q is min(u, u - q) in unsigned arithmetic.
template <int NL>
__device__ __forceinline__ uint32_t idx_u(const double (&a)[8], const IdxView& v,
                                          const FK_CONST uint32_t* row)
{
    double s = fma(a[3], v.u3, fma(a[2], v.u2, fma(a[1], v.u1, a[0] * v.u0)));
    if (NL > 4) {
#pragma unroll
        for (int k = 4; k < NL; k++)
            s = fma(a[k], load_f64(row + kIndexUW + 2 * (k - 4)), s);
    }
    const double qt = floor(fma(s, v.qi, -0.125));
    return (uint32_t)fma(-qt, v.qd, s);
}

// KC candidates per lane against the index rows [lo, hi): R rows per pass
// (their scalar loads in flight together), four for values below 2^64.
// A row kills the lane when t < n (n, the engine's depth, is every row's
// bound): the lane keeps min over the rows of u and u - q (one three-way
// minimum per row, the minimum of the t's), compared with n once at the
// end and one ballot.
template <int NL, int KC>
__device__ __forceinline__ void index_scan(const double (&a)[KC][8],
                                           const FK_CONST uint32_t* kt,
                                           uint32_t rf, uint32_t lo, uint32_t hi,
                                           uint32_t n, uint64_t (&hit)[KC])
{
    constexpr uint32_t R = NL > 4 ? 2 : 4;
    uint32_t acc[KC];
#pragma unroll
    for (int c = 0; c < KC; c++) acc[c] = 0xffffffffu;
    uint32_t p = lo;
    for (; p + R <= hi; p += R) {
        const FK_CONST uint32_t* r0 = kt + (size_t)(p - rf) * kIndexStride;
        IdxView v[R];
#pragma unroll
        for (uint32_t i = 0; i < R; i++) v[i] = load_idx(r0 + i * kIndexStride);
#pragma unroll
        for (uint32_t i = 0; i < R; i++)
#pragma unroll
            for (int c = 0; c < KC; c++) {
                const uint32_t u = idx_u<NL>(a[c], v[i], r0 + i * kIndexStride);
                acc[c] = min(min(acc[c], u), u - v[i].q);      // v_min3_u32
            }
    }
    for (; p < hi; p++) {
        const FK_CONST uint32_t* r0 = kt + (size_t)(p - rf) * kIndexStride;
        const IdxView v = load_idx(r0);
#pragma unroll
        for (int c = 0; c < KC; c++) {
            const uint32_t u = idx_u<NL>(a[c], v, r0);
            acc[c] = min(min(acc[c], u), u - v.q);
        }
    }
#pragma unroll
    for (int c = 0; c < KC; c++) hit[c] |= __ballot(acc[c] < n);
}

// The same test in 22-bit limbs (launches whose primes are all below 2^24;
// values below 2^88): NL limbs a_k < 2^22 against the row's u 2^(22k) mod q,
// three for values below 2^66, four below 2^88.
struct IdxView22 {
    double qd, qi, v0, v1, v2, v3;
    uint32_t q;
};

__device__ __forceinline__ IdxView22 load_idx22(const FK_CONST uint32_t* row)
{
    IdxView22 v;
    v.qd = load_f64(row + 0);
    v.qi = load_f64(row + 2);
    v.v0 = load_f64(row + kIndexV);
    v.v1 = load_f64(row + kIndexV + 2);
    v.v2 = load_f64(row + kIndexV + 4);
    v.v3 = load_f64(row + kIndexV + 6);
    v.q = row[12];
    return v;
}

template <int NL>
__device__ __forceinline__ uint32_t idx_u22(const double (&a)[4], const IdxView22& v)
{
    double s = fma(a[2], v.v2, fma(a[1], v.v1, a[0] * v.v0));
    if (NL > 3) s = fma(a[3], v.v3, s);
    const double qt = floor(fma(s, v.qi, -0.125));
    return (uint32_t)fma(-qt, v.qd, s);
}

template <int NL, int KC>
__device__ __forceinline__ void index_scan22(const double (&a)[KC][4],
                                             const FK_CONST uint32_t* kt,
                                             uint32_t rf, uint32_t lo, uint32_t hi,
                                             uint32_t n, uint64_t (&hit)[KC])
{
    constexpr uint32_t R = 4;
    uint32_t acc[KC];
#pragma unroll
    for (int c = 0; c < KC; c++) acc[c] = 0xffffffffu;
    uint32_t p = lo;
    for (; p + R <= hi; p += R) {
        const FK_CONST uint32_t* r0 = kt + (size_t)(p - rf) * kIndexStride;
        IdxView22 v[R];
#pragma unroll
        for (uint32_t i = 0; i < R; i++) v[i] = load_idx22(r0 + i * kIndexStride);
#pragma unroll
        for (uint32_t i = 0; i < R; i++)
#pragma unroll
            for (int c = 0; c < KC; c++) {
                const uint32_t u = idx_u22<NL>(a[c], v[i]);
                acc[c] = min(min(acc[c], u), u - v[i].q);
            }
    }
    for (; p < hi; p++) {
        const IdxView22 v = load_idx22(kt + (size_t)(p - rf) * kIndexStride);
#pragma unroll
        for (int c = 0; c < KC; c++) {
            const uint32_t u = idx_u22<NL>(a[c], v);
            acc[c] = min(min(acc[c], u), u - v.q);
        }
    }
#pragma unroll
    for (int c = 0; c < KC; c++) hit[c] |= __ballot(acc[c] < n);
}

// A value's 22-bit limbs (below 2^88: t[3] == 0, t[2] < 2^24).
__device__ __forceinline__ void limbs22(const uint32_t (&t)[4], double (&a)[4])
{
    constexpr uint32_t M = 0x3fffffu;
    a[0] = (double)(t[0] & M);
    a[1] = (double)(((t[0] >> 22) | (t[1] << 10)) & M);
    a[2] = (double)(((t[1] >> 12) | (t[2] << 20)) & M);
    a[3] = (double)((t[2] >> 2) & M);
}

// One round, lane = candidate: blockIdx.x walks the engines' segments in
// blocks of 256 KC candidates (grid-stride; lane l takes entries l,
// l + 256, ...), blockIdx.y takes a block of pblock primes.  The engine is
// uniform per block, so the prime-side data is too, and each prime row read
// serves KC candidates per lane.  A killed candidate's flag is set.

template <int KG, uint32_t ROWS>
__global__ __launch_bounds__(256) void k_round_c(RoundArgs a)
{
    constexpr int KC = kRoundKC;
    const uint32_t pb0 = a.p_begin + blockIdx.y * a.pblock;
    const uint32_t pb1 = min(pb0 + a.pblock, a.p_end);
    if (pb0 >= pb1) return;
    const FK_CONST uint2* segs = uniform_view(a.segs);
    const FK_CONST uint32_t* counts = uniform_view(a.counts);
    uint32_t touch = 0;
    SegBlock sb;
    for (uint32_t blk = blockIdx.x;
         seg_block(segs, counts, a.nengines, 256 * KC, blk, sb); blk += gridDim.x) {
        const FK_CONST DevEngine* E = uniform_view(a.engines) + sb.e;
        if (a.round >= E->nround) continue;
        // engines with the other kind of row have a launch of their own
        if ((E->kidx ? kRowsIndex : kRowsResidue) != ROWS) continue;
        const uint32_t lo = max(pb0, E->rb[a.round]);
        const uint32_t hi = min(pb1, E->rb[a.round + 1]);
        if (lo >= hi) continue;
        // touch this block's rows (one vector load per 64-byte line) so that
        // the scalar row loads below find them in the caches: vector loads
        // complete in order, so the records' wait below also covers these,
        // one memory latency for the whole block instead of one per pair
        // (the first 256 lines of the block: a hint, not a requirement)
        uint32_t tv = 0;
        {
            const uint32_t* rows = E->kt + (size_t)(lo - E->round_first) * E->kstride;
            const uint32_t nlines = ((hi - lo) * E->kstride + 15u) / 16u;
            if (threadIdx.x < nlines) tv = rows[16u * threadIdx.x];
        }
        // the flags and the records are loaded together (one latency)
        uint32_t t[KC][4], idx[KC];
        bool live[KC];
        bool any = false, wide = false, big = false;
#pragma unroll
        for (int c = 0; c < KC; c++) {
            const uint32_t li = sb.lb * (256u * KC) + 256u * c + threadIdx.x;
            idx[c] = segs[sb.e].x + li;
            uint32_t flag = 1;
            for (int k = 0; k < 4; k++) t[c][k] = 0u;
            if (li < sb.n) {
                flag = a.flags[idx[c]];
                const DevRec rec = a.recs[idx[c]];
                for (int k = 0; k < 4; k++) t[c][k] = rec.t[k];
            }
            live[c] = flag == 0;
            any = any || live[c];
            wide = wide || (live[c] && (t[c][2] | t[c][3]) != 0);
            big = big || (live[c] && ((t[c][2] | t[c][3]) != 0 ||
                                      t[c][1] >= (1u << 21)));
        }
        if (!__ballot(any)) {
            touch ^= tv;
            continue;
        }
        if (ROWS == kRowsResidue && !__ballot(big) && a.q_min >= 32u) {
            // every live value below 2^53: the value as one double
            double s[KC];
#pragma unroll
            for (int c = 0; c < KC; c++)
                s[c] = (double)(((uint64_t)t[c][1] << 32) | t[c][0]);
            const FK_CONST uint32_t* kt = uniform_view(E->kt);
            uint64_t m[KC];
#pragma unroll
            for (int c = 0; c < KC; c++) m[c] = 0;
            // Euler rows past (n - 1) n without their residues
            const uint32_t pf = min(max(lo, E->pronic_first), hi);
            if (kill_folded(E->family))
                narrow_scan<true, KC>(s, kt, E->kstride, E->round_first, lo, hi, m);
            else if (pf < hi) {
                if (lo < pf)
                    narrow_scan<false, KC>(s, kt, E->kstride, E->round_first, lo, pf, m);
                narrow_pronic<KC>(s, kt, E->kstride, E->round_first, pf, hi,
                                  (uint32_t)(E->n - 1) * (uint32_t)E->n, m);
            } else
                narrow_scan<false, KC>(s, kt, E->kstride, E->round_first, lo, hi, m);
#pragma unroll
            for (int c = 0; c < KC; c++)
                if (live[c] && ((m[c] >> __lane_id()) & 1)) a.flags[idx[c]] = 1;
            touch ^= tv;
            continue;
        }
        if (ROWS == kRowsIndex && a.l22) {
            // every prime of the launch below 2^24: 22-bit limbs, three or
            // four, unless some live value reaches 2^88
            bool huge = false, w66 = false;
#pragma unroll
            for (int c = 0; c < KC; c++) {
                huge = huge || (live[c] && (t[c][3] != 0 || t[c][2] >= (1u << 24)));
                w66 = w66 || (live[c] && t[c][2] >= 4u);
            }
            if (!__ballot(huge)) {
                double l22[KC][4];
#pragma unroll
                for (int c = 0; c < KC; c++) limbs22(t[c], l22[c]);
                const FK_CONST uint32_t* kt = uniform_view(E->kt);
                const uint32_t n = (uint32_t)E->n;
                uint64_t m[KC];
#pragma unroll
                for (int c = 0; c < KC; c++) m[c] = 0;
                if (__ballot(w66)) index_scan22<4, KC>(l22, kt, E->round_first, lo, hi, n, m);
                else index_scan22<3, KC>(l22, kt, E->round_first, lo, hi, n, m);
#pragma unroll
                for (int c = 0; c < KC; c++)
                    if (live[c] && ((m[c] >> __lane_id()) & 1)) a.flags[idx[c]] = 1;
                touch ^= tv;
                continue;
            }
        }
        double l[KC][8];
#pragma unroll
        for (int c = 0; c < KC; c++)
#pragma unroll
            for (int k = 0; k < 8; k++)
                l[c][k] = (double)((t[c][k >> 1] >> ((k & 1) * 16)) & 0xffffu);
        const bool wwide = __ballot(wide) != 0;
        // values past 2^64 in 22-bit limbs where every prime of the launch
        // is below 2^24 and every live value below 2^88
        bool huge88 = false;
#pragma unroll
        for (int c = 0; c < KC; c++)
            huge88 = huge88 || (live[c] && (t[c][3] != 0 || t[c][2] >= (1u << 24)));
        const bool w22 = ROWS == kRowsResidue && a.l22 && wwide && !__ballot(huge88);
        if (w22) {
#pragma unroll
            for (int c = 0; c < KC; c++) {
                double l22[4];
                limbs22(t[c], l22);
#pragma unroll
                for (int k = 0; k < 4; k++) l[c][k] = l22[k];
            }
        }
        const FK_CONST PrimeD* pd = uniform_view(a.primed);
        const FK_CONST uint32_t* kt = uniform_view(E->kt);
        const uint32_t rf = E->round_first;
        const uint32_t st = E->kstride;
        uint64_t m[KC];
#pragma unroll
        for (int c = 0; c < KC; c++) m[c] = 0;
        if (ROWS == kRowsIndex) {
            const uint32_t n = (uint32_t)E->n;
            if (wwide) index_scan<8, KC>(l, kt, rf, lo, hi, n, m);
            else index_scan<4, KC>(l, kt, rf, lo, hi, n, m);
        } else if (kill_folded(E->family)) {
            if (w22) round_scan<22, KG, true, KC>(l, pd, kt, st, rf, lo, hi, m);
            else if (wwide) round_scan<8, KG, true, KC>(l, pd, kt, st, rf, lo, hi, m);
            else round_scan<4, KG, true, KC>(l, pd, kt, st, rf, lo, hi, m);
        } else {
            // Euler rows past (n - 1) n without their residues
            const uint32_t pf = min(max(lo, E->pronic_first), hi);
            if (lo < pf) {
                if (w22) round_scan<22, KG, false, KC>(l, pd, kt, st, rf, lo, pf, m);
                else if (wwide) round_scan<8, KG, false, KC>(l, pd, kt, st, rf, lo, pf, m);
                else round_scan<4, KG, false, KC>(l, pd, kt, st, rf, lo, pf, m);
            }
            if (pf < hi) {
                const uint32_t pmax = (uint32_t)(E->n - 1) * (uint32_t)E->n;
                if (w22) pronic_scan<22, KC>(l, pd, kt, st, rf, pf, hi, pmax, m);
                else if (wwide) pronic_scan<8, KC>(l, pd, kt, st, rf, pf, hi, pmax, m);
                else pronic_scan<4, KC>(l, pd, kt, st, rf, pf, hi, pmax, m);
            }
        }
#pragma unroll
        for (int c = 0; c < KC; c++)
            if (live[c] && ((m[c] >> __lane_id()) & 1)) a.flags[idx[c]] = 1;
        touch ^= tv;
    }
    // keeps the touch loads; round == ~0 never comes from the host
    if (a.round == 0xffffffffu && touch == 1u) a.flags[0] = touch;
}

// ---- lane = prime (many primes, few candidates): each lane holds one
// prime's row in VGPRs and the wave walks the candidates, staged in LDS.

// One round prime's row in a lane's registers: the residue constants and
// up to kRegResidues killing residues (kKillPad past the count).
constexpr int kRegResidues = 32;

struct LaneRow {
    double qd, qi, c16, c32, c48;
    double cw[4];                 // 2^64 .. 2^112 mod q (values past 2^64)
    double c22[3];                // 2^22, 2^44, 2^66 mod q (22-bit limbs)
    uint32_t q, cnt;
    uint32_t k[kRegResidues];
    const uint32_t* row;          // the row in memory (residues past 32)
};

// The row's header and the residue groups a wave compares from registers:
// three groups (twelve residues) when the family's kill sets fit them
// (ncmp <= 12), five otherwise; the constants for values past 2^53 only
// when the slice holds such values (wide: 16-bit limbs past 2^64, w22:
// 22-bit limbs), each a load of its own per lane.
__device__ __forceinline__ void load_lane_row(LaneRow& L, const uint32_t* row,
                                              const PrimeD* pd, uint32_t stride,
                                              uint32_t ncmp, bool wide, bool w22)
{
    RoundRow h;
    const uint4* r4 = reinterpret_cast<const uint4*>(row);
    uint4 w0 = r4[0], w1 = r4[1], w2 = r4[2];
    const uint32_t hw[kRowHead] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y,
                                   w1.z, w1.w, w2.x, w2.y, w2.z, w2.w};
    __builtin_memcpy(&h, hw, sizeof h);
    L.qd = h.qd;
    L.qi = h.qi;
    L.c16 = h.c16;
    L.c32 = h.c32;
    L.c48 = h.c48;
    L.q = h.q;
    L.cnt = h.cnt;
    L.row = row;
    const int ng = ncmp <= 12u ? 3 : 5;
#pragma unroll
    for (int i = 0; i < kRegResidues / 4; i++) {
        // residue groups inside the row only (rows are 16..32 words here)
        const uint4 v = (i < ng && 4 * i + 4 <= 20 && kRowHead + 4u * i + 4u <= stride)
                            ? r4[3 + i]
                            : make_uint4(kKillPad, kKillPad, kKillPad, kKillPad);
        L.k[4 * i] = v.x;
        L.k[4 * i + 1] = v.y;
        L.k[4 * i + 2] = v.z;
        L.k[4 * i + 3] = v.w;
    }
    if (wide && !w22) {
#pragma unroll
        for (int i = 0; i < 4; i++) L.cw[i] = pd->c[3 + i];
    }
    if (w22) {
#pragma unroll
        for (int i = 0; i < 3; i++) L.c22[i] = pd->c22[i];
    }
}

// Round prime rows hold at least 20 residues (32 words); rows of engines
// with larger kill sets (Dickson n > 10, the others n > 20) carry more.
__device__ __forceinline__ void load_lane_row_wide(LaneRow& L, uint32_t kmax)
{
    const uint32_t* res = L.row + kRowHead;
#pragma unroll
    for (int i = 20; i < kRegResidues; i++)
        L.k[i] = ((uint32_t)i < kmax) ? res[i] : kKillPad;
}

// The lane's prime against a candidate of residue-limbs a[] (wave
// uniform, read from LDS): zero when the prime kills it.  NG > 0: the
// wave's rows hold at most 4 NG residues, NG groups of four compared from
// registers, unrolled.  NG == 0: ncmp (wave uniform, a multiple of 4)
// residues from registers, residues past 32 (kill sets larger than 32)
// from memory.
template <int NL, bool FOLD, int NG>
__device__ __forceinline__ uint32_t lane_kills(const LaneRow& L, const double* a,
                                               uint32_t ncmp)
{
    // NL == 1: a[0] is the value itself (below 2^53, primes at least 32);
    // NL == 22: a[0..3] are 22-bit limbs (below 2^88, primes below 2^24)
    double s = a[0];
    if (NL == 22)
        s = fma(a[3], L.c22[2], fma(a[2], L.c22[1], fma(a[1], L.c22[0], a[0])));
    else if (NL > 1)
        s = fma(a[3], L.c48, fma(a[2], L.c32, fma(a[1], L.c16, a[0])));
    if (NL > 4 && NL != 22) {
#pragma unroll
        for (int k = 0; k < 4; k++) s = fma(a[4 + k], L.cw[k], s);
    }
    const uint32_t r = FOLD ? resid_fold_d(s, L.qd, L.qi, L.q)
                            : resid_d1(s, L.qd, L.qi, L.q);
    // zero iff r is one of the residues: a minimum of r ^ k in vector
    // registers (padding 0xffffffff never matches a residue below 2^30)
    uint32_t d = 0xffffffffu;
    if (NG > 0) {
#pragma unroll
        for (int i = 0; i < 4 * NG; i += 4)
            d = min(d, min(min(r ^ L.k[i], r ^ L.k[i + 1]),
                           min(r ^ L.k[i + 2], r ^ L.k[i + 3])));
        return d;
    }
#pragma unroll
    for (int i = 0; i < kRegResidues; i += 4) {
        if ((uint32_t)i < ncmp)
            d = min(d, min(min(r ^ L.k[i], r ^ L.k[i + 1]),
                           min(r ^ L.k[i + 2], r ^ L.k[i + 3])));
    }
    if (ncmp > (uint32_t)kRegResidues) {
        for (uint32_t c = kRegResidues; c < L.cnt; c++)
            d = min(d, r ^ L.row[kRowHead + c]);
    }
    return d;
}

// One round, lane = prime: blockIdx.x picks 256 primes (one per lane),
// blockIdx.y walks slices of kRoundSlice candidates of the engines'
// segments (grid-stride).  Per slice the workgroup stages every live
// candidate's 16-bit limbs (as doubles) in LDS once; each wave then walks
// them with broadcast LDS reads, every lane testing its own prime: the
// exact double-precision residue (resid_d1) and one compare per killing
// residue.  A candidate that some lane's prime kills is flagged by every
// such lane (each storing 1).  There is no global memory access per
// (candidate, prime).

// One slice's candidates against the lane's prime; candidate c's input is
// a + step * c, and kills(x) says whether the lane's prime kills the
// candidate of input x.  Every lane computes (lanes past the round's
// primes hold a row that kills nothing), with no per-lane branch around
// the test, and the flags are written only when one of two candidates in
// a row is killed, which is rare: one ballot per two candidates.
template <typename Kills>
__device__ __forceinline__ void lane_prime_walk(Kills kills, const double* a,
                                                uint32_t step,
                                                const uint32_t* live_idx,
                                                uint32_t nlive, uint32_t* flags)
{
    // two candidates per step: two independent residue chains in flight
    // per lane (four hold 120 VGPRs in the index form, half the waves)
    constexpr uint32_t U = 2;
    uint32_t c = 0;
    for (; c + U <= nlive; c += U) {
        bool kill[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) kill[u] = kills(a + step * (c + u));
        if (__ballot(kill[0] | kill[1])) {
#pragma unroll
            for (uint32_t u = 0; u < U; u++)
                if (kill[u]) flags[live_idx[c + u]] = 1;
        }
    }
    for (; c < nlive; c++)
        if (kills(a + step * c)) flags[live_idx[c]] = 1;
}

// Residue rows: 16-bit limbs (8 past 2^64, else 4), or with every value
// below 2^53 the value itself (NL == 1).
// miss: 0 for a lane whose prime is in the round, 1 (the test never
// reaches zero) for the others.  NG as in lane_kills.
template <bool FOLD, int NG>
__device__ __forceinline__ void lane_prime_slice(const LaneRow& L, uint32_t miss,
                                                 uint32_t ncmp, bool wide,
                                                 bool narrow, bool w22,
                                                 const double* limbs,
                                                 const double* vals,
                                                 const double* l22,
                                                 const uint32_t* live_idx,
                                                 uint32_t nlive, uint32_t* flags)
{
    if (w22)
        lane_prime_walk([&](const double* x) {
                            return (lane_kills<22, FOLD, NG>(L, x, ncmp) | miss) == 0u; },
                        l22, 4, live_idx, nlive, flags);
    else if (wide)
        lane_prime_walk([&](const double* x) {
                            return (lane_kills<8, FOLD, NG>(L, x, ncmp) | miss) == 0u; },
                        limbs, 8, live_idx, nlive, flags);
    else if (narrow)
        lane_prime_walk([&](const double* x) {
                            return (lane_kills<1, FOLD, NG>(L, x, ncmp) | miss) == 0u; },
                        vals, 1, live_idx, nlive, flags);
    else
        lane_prime_walk([&](const double* x) {
                            return (lane_kills<4, FOLD, NG>(L, x, ncmp) | miss) == 0u; },
                        limbs, 8, live_idx, nlive, flags);
}

// lane_prime_slice with the wave's group count ng as a template argument.
template <bool FOLD>
__device__ __forceinline__ void lane_prime_groups(uint32_t ng, const LaneRow& L,
                                                  uint32_t miss, uint32_t ncmp,
                                                  bool wide, bool narrow, bool w22,
                                                  const double* limbs,
                                                  const double* vals,
                                                  const double* l22,
                                                  const uint32_t* live_idx,
                                                  uint32_t nlive, uint32_t* flags)
{
    if (ng == 1)
        lane_prime_slice<FOLD, 1>(L, miss, ncmp, wide, narrow, w22, limbs, vals, l22,
                                  live_idx, nlive, flags);
    else if (ng == 2)
        lane_prime_slice<FOLD, 2>(L, miss, ncmp, wide, narrow, w22, limbs, vals, l22,
                                  live_idx, nlive, flags);
    else if (ng == 3)
        lane_prime_slice<FOLD, 3>(L, miss, ncmp, wide, narrow, w22, limbs, vals, l22,
                                  live_idx, nlive, flags);
    else
        lane_prime_slice<FOLD, 0>(L, miss, ncmp, wide, narrow, w22, limbs, vals, l22,
                                  live_idx, nlive, flags);
}

// An index row (AP) in a lane's registers.
struct LaneIdx {
    double qd, qi;
    double u[8];                  // u 2^(16k) mod q
    double v[4];                  // u 2^(22k) mod q
    uint32_t q, n;
};

// half: the slice's values take the 22-bit form, which reads only the
// row's first 64 bytes (the 16-bit constants are left zero)
__device__ __forceinline__ void load_lane_idx(LaneIdx& I, const uint32_t* row,
                                              bool half)
{
    IndexRow h;
    const uint4* r4 = reinterpret_cast<const uint4*>(row);
    uint32_t hw[kIndexStride];
#pragma unroll
    for (int i = 0; i < (int)kIndexStride / 4; i++) {
        const uint4 v = (!half || i < (int)kIndexHalf / 4) ? r4[i]
                                                           : make_uint4(0, 0, 0, 0);
        hw[4 * i] = v.x;
        hw[4 * i + 1] = v.y;
        hw[4 * i + 2] = v.z;
        hw[4 * i + 3] = v.w;
    }
    __builtin_memcpy(&h, hw, sizeof h);
    I.qd = h.qd;
    I.qi = h.qi;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        I.u[k] = h.u[k];
        I.u[4 + k] = h.uw[k];
        I.v[k] = h.v[k];
    }
    I.q = h.q;
    I.n = h.n;
}

// t < nlim: nlim is the row's n for a lane whose prime is in the round, 0
// (nothing killed) for the others.
template <int NL>
__device__ __forceinline__ bool lane_idx_kills(const LaneIdx& I, const double* a,
                                               uint32_t nlim)
{
    double s = a[0] * I.u[0];
#pragma unroll
    for (int k = 1; k < NL; k++) s = fma(a[k], I.u[k], s);
    return resid_d1(s, I.qd, I.qi, I.q) < nlim;
}

// The same in 22-bit limbs (every prime of the launch below 2^24).
template <int NL>
__device__ __forceinline__ bool lane_idx_kills22(const LaneIdx& I, const double* a,
                                                 uint32_t nlim)
{
    double s = a[0] * I.v[0];
#pragma unroll
    for (int k = 1; k < NL; k++) s = fma(a[k], I.v[k], s);
    return resid_d1(s, I.qd, I.qi, I.q) < nlim;
}

template <uint32_t ROWS>
__global__ __launch_bounds__(256) void k_round_p(RoundArgs a)
{
    __shared__ double sh_limbs[kRoundSlice * 8];
    __shared__ double sh_val[kRoundSlice];      // the value, when below 2^53
    __shared__ double sh_l22[kRoundSlice * 4];
    __shared__ uint32_t sh_idx[kRoundSlice];
    __shared__ uint32_t sh_n, sh_wide, sh_wide53, sh_w66, sh_huge;
    const uint32_t p = a.p_begin + blockIdx.x * 256 + threadIdx.x;
    const uint32_t blk_p0 = a.p_begin + blockIdx.x * 256;
    if (blk_p0 >= a.p_end) return;                // whole workgroup past it
    const FK_CONST uint2* segs = uniform_view(a.segs);
    const FK_CONST uint32_t* counts = uniform_view(a.counts);
    const FK_CONST DevEngine* engs = uniform_view(a.engines);
    LaneRow L;
    SegBlock sb;
    for (uint32_t blk = blockIdx.y;
         seg_block(segs, counts, a.nengines, kRoundSlice, blk, sb);
         blk += gridDim.y) {
        const uint32_t e = sb.e;
        if (a.round >= engs[e].nround) continue;
        // engines with the other kind of row have a launch of their own
        if ((engs[e].kidx ? kRowsIndex : kRowsResidue) != ROWS) continue;
        const uint32_t rlo = engs[e].rb[a.round], rhi = engs[e].rb[a.round + 1];
        // workgroup-uniform: does any of its 256 primes belong to e's round?
        if (blk_p0 + 256 <= rlo || blk_p0 >= rhi) continue;
        const bool mine = p >= rlo && p < rhi && p < a.p_end;
        // stage the slice's live candidates: limbs and queue index
        if (threadIdx.x == 0) {
            sh_n = 0;
            sh_wide = 0;
            sh_wide53 = 0;
            sh_w66 = 0;
            sh_huge = 0;
        }
        __syncthreads();
        {
            const uint32_t li = sb.lb * kRoundSlice + threadIdx.x;
            if (li < sb.n) {
                const uint32_t i = segs[e].x + li;
                if (a.flags[i] == 0) {
                    const DevRec rec = a.recs[i];
                    const uint32_t slot = atomicAdd(&sh_n, 1u);
#pragma unroll
                    for (int k = 0; k < 8; k++)
                        sh_limbs[8 * slot + k] =
                            (double)((rec.t[k >> 1] >> ((k & 1) * 16)) & 0xffffu);
                    sh_val[slot] =
                        (double)(((uint64_t)rec.t[1] << 32) | rec.t[0]);
                    sh_idx[slot] = i;
                    if ((rec.t[2] | rec.t[3]) != 0) sh_wide = 1;
                    if ((rec.t[2] | rec.t[3]) != 0 || rec.t[1] >= (1u << 21))
                        sh_wide53 = 1;
                    if (a.l22) {
                        const uint32_t tt[4] = {rec.t[0], rec.t[1], rec.t[2], rec.t[3]};
                        double l22[4];
                        limbs22(tt, l22);
#pragma unroll
                        for (int k = 0; k < 4; k++) sh_l22[4 * slot + k] = l22[k];
                        if (rec.t[2] >= 4u || rec.t[3] != 0) sh_w66 = 1;
                        if (rec.t[3] != 0 || rec.t[2] >= (1u << 24)) sh_huge = 1;
                    }
                }
            }
        }
        __syncthreads();
        const uint32_t nlive = sh_n;
        const bool wide = sh_wide != 0;
        const bool below53 = sh_wide53 == 0;
        if (nlive != 0) {
            // compares per test: the family's largest kill set, in fours
            const uint32_t ncmp = (kill_max(engs[e].family, engs[e].n) + 3u) & ~3u;
            const uint32_t stride = engs[e].kstride;
            // lanes past the round's primes hold the row of its last prime
            // (in the table) and are masked in the test: no lane branches
            // around it (the workgroup holds some of the round's primes)
            uint32_t pc = min(max(p, rlo), rhi - 1u);
            // residue rows in the round's count order (a permutation of
            // the round): the waves' rows then mostly need as many groups
            // of four compares as each other
            if (ROWS == kRowsResidue && engs[e].kperm)
                pc = engs[e].round_first + engs[e].kperm[pc - engs[e].round_first];
            // the one-value form needs every value below 2^53 and every
            // prime of the round at least 32
            const bool narrow = below53 && a.q_min >= 32u;
            // values past 2^64 in 22-bit limbs (primes below 2^24, values
            // below 2^88)
            const bool w22 = a.l22 && wide && sh_huge == 0;
            if (ROWS == kRowsResidue && __ballot(mine)) {
                load_lane_row(L, engs[e].kt + (size_t)(pc - engs[e].round_first) * stride,
                              a.primed + pc, stride, ncmp, wide, w22);
                if (ncmp > 20) load_lane_row_wide(L, stride - kRowHead);
            }
            if (ROWS == kRowsIndex) {
                if (__ballot(mine)) {
                    LaneIdx I;
                    load_lane_idx(I, engs[e].kt + (size_t)(pc - engs[e].round_first) *
                                                      kIndexStride,
                                  a.l22 && sh_huge == 0);
                    const uint32_t nlim = mine ? I.n : 0u;
                    if (a.l22 && sh_huge == 0) {
                        if (sh_w66)
                            lane_prime_walk([&](const double* x) {
                                                return lane_idx_kills22<4>(I, x, nlim); },
                                            sh_l22, 4, sh_idx, nlive, a.flags);
                        else
                            lane_prime_walk([&](const double* x) {
                                                return lane_idx_kills22<3>(I, x, nlim); },
                                            sh_l22, 4, sh_idx, nlive, a.flags);
                    } else if (wide)
                        lane_prime_walk([&](const double* x) {
                                            return lane_idx_kills<8>(I, x, nlim); },
                                        sh_limbs, 8, sh_idx, nlive, a.flags);
                    else
                        lane_prime_walk([&](const double* x) {
                                            return lane_idx_kills<4>(I, x, nlim); },
                                        sh_limbs, 8, sh_idx, nlive, a.flags);
                }
            } else if (__ballot(mine)) {
                const uint32_t miss = mine ? 0u : 1u;
                // groups of four the wave's largest row needs (every lane
                // compares them all: the residues past a row's count are
                // padding); 0: more than twelve, the general loop
                const uint32_t ng = __ballot(L.cnt > 12u) ? 0u
                                    : __ballot(L.cnt > 8u) ? 3u
                                    : __ballot(L.cnt > 4u) ? 2u : 1u;
                if (kill_folded(engs[e].family))
                    lane_prime_groups<true>(ng, L, miss, ncmp, wide, narrow, w22,
                                            sh_limbs, sh_val, sh_l22, sh_idx, nlive,
                                            a.flags);
                else
                    lane_prime_groups<false>(ng, L, miss, ncmp, wide, narrow, w22,
                                             sh_limbs, sh_val, sh_l22, sh_idx, nlive,
                                             a.flags);
            }
        }
        __syncthreads();                           // before the next staging
    }
}

// Compaction of the unflagged candidates, within each engine's segment:
// each workgroup takes 1024 entries of one segment (four per thread,
// coalesced), counts its survivors with ballots, and reserves room for all
// of them in the segment with one atomic.  One atomic per wave serialises
// on the counter when the queue holds many records.
constexpr uint32_t CP_PER = 4;

__global__ __launch_bounds__(256) void k_round_compact(CompactArgs a)
{
    __shared__ uint32_t wsum[4];
    __shared__ uint32_t wbase;
    const FK_CONST uint2* segs = uniform_view(a.segs);
    const FK_CONST uint32_t* counts = uniform_view(a.in_counts);
    const uint32_t lane = __lane_id();
    const uint32_t wave = threadIdx.x >> 6;
    SegBlock sb;
    for (uint32_t blk = blockIdx.x;
         seg_block(segs, counts, a.nengines, 256 * CP_PER, blk, sb);
         blk += gridDim.x) {
        const uint32_t base = segs[sb.e].x;
        const uint32_t l0 = sb.lb * (256 * CP_PER);
        // the records (24 B: three 8-byte words) are loaded with their
        // flags, one memory latency; the stores below then wait only for the
        // atomic
        uint32_t flag[CP_PER];
        uint2 rec[CP_PER][3];
#pragma unroll
        for (uint32_t k = 0; k < CP_PER; k++) {
            const uint32_t li = l0 + k * 256 + threadIdx.x;
            flag[k] = 1;
            if (li < sb.n) {
                flag[k] = a.flags[base + li];
                const uint2* src = reinterpret_cast<const uint2*>(a.in + base + li);
                rec[k][0] = src[0];
                rec[k][1] = src[1];
                rec[k][2] = src[2];
            }
        }
        uint64_t m[CP_PER];
        uint32_t tot = 0;
#pragma unroll
        for (uint32_t k = 0; k < CP_PER; k++) {
            const uint32_t li = l0 + k * 256 + threadIdx.x;
            if (li < sb.n) a.flags[base + li] = 0;
            m[k] = __ballot(flag[k] == 0);
            tot += (uint32_t)__popcll((unsigned long long)m[k]);
        }
        if (lane == 0) wsum[wave] = tot;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t run = 0;
            for (int w = 0; w < 4; w++) {
                const uint32_t t = wsum[w];
                wsum[w] = run;
                run += t;
            }
            wbase = run ? atomicAdd(a.out_counts + sb.e, run) : 0u;
        }
        __syncthreads();
        uint32_t off = base + wbase + wsum[wave];
#pragma unroll
        for (uint32_t k = 0; k < CP_PER; k++) {
            if ((m[k] >> lane) & 1) {
                const uint32_t below = __builtin_amdgcn_mbcnt_hi(
                    (uint32_t)(m[k] >> 32),
                    __builtin_amdgcn_mbcnt_lo((uint32_t)m[k], 0u));
                uint2* dst = reinterpret_cast<uint2*>(a.out + off + below);
                dst[0] = rec[k][0];
                dst[1] = rec[k][1];
                dst[2] = rec[k][2];
            }
            off += (uint32_t)__popcll((unsigned long long)m[k]);
        }
        __syncthreads();           // wsum and wbase are reused
    }
}

// Pack the engines' segments into one array for the classification and
// the copy back.
__global__ __launch_bounds__(256) void k_gather(GatherArgs a)
{
    const FK_CONST uint2* segs = uniform_view(a.segs);
    const FK_CONST uint32_t* counts = uniform_view(a.counts);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        uint32_t tot = 0;
        for (uint32_t e = 0; e < a.nengines; e++) tot += min(counts[e], segs[e].y);
        *a.total = tot;
    }
    SegBlock sb;
    for (uint32_t blk = blockIdx.x;
         seg_block(segs, counts, a.nengines, 256, blk, sb); blk += gridDim.x) {
        const uint32_t li = sb.lb * 256 + threadIdx.x;
        if (li < sb.n) a.out[sb.before + li] = a.in[segs[sb.e].x + li];
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

// Few candidates: every base in one pass (one latency instead of two),
// decided from the true count.
__device__ __forceinline__ bool cls_single(const ClassifyArgs& a,
                                           uint32_t count)
{
    return (uint64_t)count * a.cap_max <= a.single_limit;
}

__global__ __launch_bounds__(128) void k_cls_base2(ClassifyArgs a)
{
    const uint32_t count = min(*a.count, a.cap);
    if (cls_single(a, count)) return;
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
        rec_chain_value(E, rec, W.first + i, v);
        if (!is_sprp(v, 0, 1))
            atomicOr((unsigned long long*)&a.fail2[r], 1ull << i);
    }
}

__global__ __launch_bounds__(128) void k_cls_full(ClassifyArgs a)
{
    const uint32_t count = min(*a.count, a.cap);
    // single pass: all seven bases (base 2 into fail2) with no dependence
    // on a base-2 pass; otherwise bases 3..7 after it.  One call site of
    // the strong test, so a wave whose lanes hold different bases runs one
    // instruction stream.
    const bool single = cls_single(a, count);
    const uint32_t nb = single ? 7 : 6;
    const uint64_t per = (uint64_t)a.cap_max * nb;
    const uint64_t total = (uint64_t)count * per;
    for (uint64_t idx = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
         idx < total; idx += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t r = (uint32_t)(idx / per);
        const uint32_t rem = (uint32_t)(idx - (uint64_t)r * per);
        const int i = (int)(rem / nb);
        const int b = (int)(rem - (uint32_t)i * nb) + (single ? 0 : 1);
        const DevRec rec = a.recs[r];
        const DevWinInfo W = a.wins[rec.window];
        if (W.mode == 0 || i >= W.cap) continue;
        if (!single) {
            const int u = first_set(a.fail2[r], W.cap);
            if (i >= u) continue;
            if (W.mode == 2 && u < W.exact_from) continue;
        }
        const DevEngine& E = a.engines[W.engine];
        uint32_t v[8];
        rec_chain_value(E, rec, W.first + i, v);
        if (!is_sprp(v, b, 1))
            atomicOr((unsigned long long*)(b == 0 ? &a.fail2[r] : &a.failx[r]),
                     1ull << i);
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


// --------------------------------------------------------------- launchers

// A launch, stamped with the caller's start/stop events when it gives them
// (hipExtLaunchKernelGGL records them at the kernel's own start and end, so
// no separate event packet sits between two launches).
#define FK_LAUNCH(EV, FN, GRID, BLOCK, LDS, STREAM, ...)                     \
    do {                                                                     \
        if ((EV) && (EV)->start)                                             \
            hipExtLaunchKernelGGL(FN, GRID, BLOCK, (uint32_t)(LDS), STREAM,  \
                                  (EV)->start, (EV)->stop, 0, __VA_ARGS__);  \
        else                                                                 \
            hipLaunchKernelGGL(FN, GRID, BLOCK, LDS, STREAM, __VA_ARGS__);   \
    } while (0)

// Tile shapes (threads, words per thread): a tile is 64 * T * WPT line
// positions.  The odd word counts let a window's tiles fit its lines more
// tightly (the planner picks the shape with the fewest padded positions).
#define FK_SIEVE_SHAPES(X)                                                 \
    X(256, 1) X(256, 2) X(256, 3) X(256, 4) X(256, 6) X(256, 8) X(256, 16)   \
    X(512, 1) X(512, 2) X(512, 3) X(512, 4) X(512, 8)                        \
    X(1024, 1) X(1024, 2) X(1024, 4)

// Dynamic LDS above 64 KiB must be allowed per kernel (and device) before
// a launch.  The call costs tens of microseconds of host time, so it is
// made once per kernel, device and larger size, not per launch.
static void allow_lds(const void* fn, size_t lds_bytes)
{
    if (lds_bytes <= 65536) return;
    struct Allowed { const void* fn; int device; size_t bytes; };
    static std::mutex mu;
    static std::vector<Allowed> done;
    int dev = 0;
    (void)hipGetDevice(&dev);
    std::lock_guard<std::mutex> lock(mu);
    for (Allowed& a : done) {
        if (a.fn == fn && a.device == dev) {
            if (a.bytes >= lds_bytes) return;
            if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds_bytes) == hipSuccess)
                a.bytes = lds_bytes;
            return;
        }
    }
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds_bytes) == hipSuccess)
        done.push_back({fn, dev, lds_bytes});
}

static const void* sieve_short_fn(int chunk)
{
    switch (chunk) {
    case 1: return (const void*)k_sieve_short<1>;
    case 2: return (const void*)k_sieve_short<2>;
    case 3: return (const void*)k_sieve_short<3>;
    case 4: return (const void*)k_sieve_short<4>;
    case 6: return (const void*)k_sieve_short<6>;
    case 8: return (const void*)k_sieve_short<8>;
    case 12: return (const void*)k_sieve_short<12>;
    case 16: return (const void*)k_sieve_short<16>;
    default: return nullptr;
    }
}

bool sieve_short_chunk_ok(int chunk) { return sieve_short_fn(chunk) != nullptr; }

hipError_t launch_sieve_short(const SieveArgs& a, int chunk, uint32_t grid,
                              size_t lds_bytes, hipStream_t stream,
                              const LaunchEvents* ev)
{
    const void* fn = sieve_short_fn(chunk);
    if (!fn) return hipErrorInvalidValue;
    allow_lds(fn, lds_bytes);
#define FK_SHORT_CASE(CC)                                                  \
    case CC:                                                               \
        FK_LAUNCH(ev, k_sieve_short<CC>, dim3(grid), dim3(256), lds_bytes,  \
                  stream, a);                                              \
        break;
    switch (chunk) {
        FK_SHORT_CASE(1) FK_SHORT_CASE(2) FK_SHORT_CASE(3) FK_SHORT_CASE(4)
        FK_SHORT_CASE(6) FK_SHORT_CASE(8) FK_SHORT_CASE(12) FK_SHORT_CASE(16)
    }
#undef FK_SHORT_CASE
    return hipGetLastError();
}

int sieve_short_occupancy(int chunk, size_t lds_bytes)
{
    const void* fn = sieve_short_fn(chunk);
    if (!fn) return 0;
    int blocks = 0;
    allow_lds(fn, lds_bytes);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, fn, 256,
                                                     lds_bytes) != hipSuccess)
        blocks = 0;
    return blocks;
}

hipError_t launch_sieve(const SieveArgs& a, int threads, int wpt,
                        uint32_t grid, size_t lds_bytes, hipStream_t stream,
                        const LaunchEvents* ev)
{
#define FK_SIEVE_CASE(TT, WW)                                              \
    if (threads == TT && wpt == WW) {                                      \
        auto fn = k_sieve<TT, WW>;                                         \
        allow_lds((const void*)fn, lds_bytes);                             \
        FK_LAUNCH(ev, fn, dim3(grid), dim3(TT), lds_bytes, stream, a);     \
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
        allow_lds((const void*)fn, lds_bytes);                              \
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

hipError_t launch_round_test(const RoundArgs& a, bool lane_prime, uint32_t gx,
                             uint32_t gy, hipStream_t stream,
                             const LaunchEvents* ev)
{
    const uint64_t np = a.p_end - a.p_begin;
    if (gx < 1 || gy < 1 || gy > 65535) return hipErrorInvalidValue;
    if (lane_prime) {
        // every prime must have its lane: 256 primes per blockIdx.x
        if ((uint64_t)gx * 256 < np) return hipErrorInvalidValue;
        if (a.rows == kRowsResidue)
            FK_LAUNCH(ev, k_round_p<kRowsResidue>, dim3(gx, gy), dim3(256), 0,
                      stream, a);
        else if (a.rows == kRowsIndex)
            FK_LAUNCH(ev, k_round_p<kRowsIndex>, dim3(gx, gy), dim3(256), 0,
                      stream, a);
        else
            return hipErrorInvalidValue;
    } else {
        // every prime block must be launched
        if (a.pblock == 0 || (uint64_t)gy * a.pblock < np)
            return hipErrorInvalidValue;
        if (a.rows == kRowsIndex)
            FK_LAUNCH(ev, (k_round_c<3, kRowsIndex>), dim3(gx, gy), dim3(256),
                      0, stream, a);
        else if (a.rows == kRowsResidue && a.kgroups == 3)
            FK_LAUNCH(ev, (k_round_c<3, kRowsResidue>), dim3(gx, gy), dim3(256),
                      0, stream, a);
        else if (a.rows == kRowsResidue && a.kgroups == 5)
            FK_LAUNCH(ev, (k_round_c<5, kRowsResidue>), dim3(gx, gy), dim3(256),
                      0, stream, a);
        else
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_kill_table(const PrimeK* primes, uint32_t p0, uint32_t np,
                             int family, int n, const uint32_t d[4], bool index,
                             uint32_t stride, uint32_t* kt, hipStream_t stream)
{
    if (np == 0) return hipSuccess;
    if (index ? (!kill_indexable(family) || stride != kIndexStride)
              : stride != round_row_stride(family, n))
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_kill_table, dim3((np + 255) / 256), dim3(256), 0,
                       stream, primes, p0, np, family, n,
                       make_uint4(d[0], d[1], d[2], d[3]), index ? 1u : 0u,
                       stride, kt);
    return hipGetLastError();
}

hipError_t launch_round_compact(const CompactArgs& a, uint32_t grid,
                                hipStream_t stream, const LaunchEvents* ev)
{
    FK_LAUNCH(ev, k_round_compact, dim3(grid ? grid : 1), dim3(256), 0,
              stream, a);
    return hipGetLastError();
}

hipError_t launch_classify(const ClassifyArgs& a, uint32_t grid,
                           hipStream_t stream)
{
    if (grid < 1) grid = 1;
    hipError_t e;
    // exits at once when the count makes it a single pass
    hipLaunchKernelGGL(k_cls_base2, dim3(grid), dim3(128), 0, stream, a);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_cls_full, dim3(grid), dim3(128), 0, stream, a);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    uint32_t g2 = grid / 4 + 1;
    hipLaunchKernelGGL(k_cls_finish, dim3(g2), dim3(256), 0, stream, a);
    return hipGetLastError();
}

// The readback kernel: one launch in place of two copies; only the records
// the queue holds cross the bus (8-byte words, grid-stride).
// A factored wheel expanded into its residues: for each residue b of the
// wheel without its last prime L (period Mb), the nj residues b + j Mb with
// j from b's class (b mod L) of the lift table, written at b's index times
// nj.  One thread per base residue, grid-stride.
__global__ __launch_bounds__(256) void k_wheel_expand(
    const uint64_t* __restrict__ base, uint64_t nbase,
    const uint32_t* __restrict__ lifts, uint32_t nj, uint32_t L, uint64_t Mb,
    uint64_t* __restrict__ out)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
         i < nbase; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t b = base[i];
        const uint32_t* row = lifts + (size_t)(b % L) * nj;
        uint64_t* dst = out + i * nj;
        for (uint32_t t = 0; t < nj; t++) dst[t] = b + (uint64_t)row[t] * Mb;
    }
}

hipError_t launch_wheel_expand(const uint64_t* base, uint64_t nbase,
                               const uint32_t* lifts, uint32_t nj, uint32_t L,
                               uint64_t Mb, uint64_t* out, hipStream_t stream)
{
    const uint32_t grid = (uint32_t)std::min<uint64_t>((nbase + 255) / 256, 65536);
    hipLaunchKernelGGL(k_wheel_expand, dim3(std::max(grid, 1u)), dim3(256), 0,
                       stream, base, nbase, lifts, nj, L, Mb, out);
    return hipGetLastError();
}

// Zero n words (a fill kernel rather than a memset, so that a caller may
// capture it in a graph).
__global__ __launch_bounds__(256) void k_zero_u64(uint64_t* p, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

hipError_t launch_zero_u64(uint64_t* p, size_t n, hipStream_t stream)
{
    if (n == 0) return hipSuccess;
    const uint32_t grid = (uint32_t)std::min<size_t>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(k_zero_u64, dim3(grid), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_readback(ReadbackArgs a)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t step = gridDim.x * blockDim.x;
    for (uint32_t i = t; i < a.ncounts; i += step) a.counts_host[i] = a.counts[i];
    const uint32_t n = min(*a.count, a.cap);
    const uint2* src = reinterpret_cast<const uint2*>(a.recs);
    uint2* dst = reinterpret_cast<uint2*>(a.recs_host);
    for (uint32_t i = t; i < 3u * n; i += step) dst[i] = src[i];
}

hipError_t launch_readback(const ReadbackArgs& a, hipStream_t stream)
{
    uint32_t grid = (3u * a.cap + 255) / 256;
    grid = grid < 1 ? 1 : (grid > 256 ? 256 : grid);
    hipLaunchKernelGGL(k_readback, dim3(grid), dim3(256), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_gather(const GatherArgs& a, uint32_t grid,
                         hipStream_t stream)
{
    hipLaunchKernelGGL(k_gather, dim3(grid ? grid : 1), dim3(256), 0, stream, a);
    return hipGetLastError();
}

// ------------------------------------------------------------ peak probes
// The rooflines' denominators, measured: the vector ALU's issue rate on
// two dependent-chain loops (eight independent chains per lane, one VALU
// instruction per step: a double FMA, as the round kernels' residues, and
// a 32-bit funnel shift, as the sieve's words) and the LDS read bandwidth
// (16-byte reads, consecutive lanes on consecutive addresses, no bank
// conflicts).  The results stay in registers; a store no input can reach
// keeps the loops.

__global__ __launch_bounds__(256) void k_probe_valu_f64(uint32_t iters, double seed,
                                                        double* out)
{
    double x[8];
    const double m = 1.0 - 0x1p-40;
#pragma unroll
    for (int i = 0; i < 8; i++) x[i] = (double)(threadIdx.x + i) + seed;
    for (uint32_t it = 0; it < iters; it++) {
#pragma unroll
        for (int i = 0; i < 8; i++) x[i] = fma(x[i], m, seed);
    }
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 8; i++) s += x[i];
    if (s == -1.0) out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_probe_valu_int(uint32_t iters, uint32_t seed,
                                                        uint32_t* out)
{
    uint32_t x[8];
#pragma unroll
    for (int i = 0; i < 8; i++) x[i] = threadIdx.x * 2654435761u + i + seed;
    for (uint32_t it = 0; it < iters; it++) {
#pragma unroll
        for (int i = 0; i < 8; i++) x[i] = __builtin_amdgcn_alignbit(x[i], seed, 7);
    }
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) s ^= x[i];
    if (s == 0x9e3779b9u && seed == 0u) out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

constexpr uint32_t kProbeLdsEntries = 2048;      // 16-byte entries: 32 KB

__global__ __launch_bounds__(256) void k_probe_lds(uint32_t iters, uint32_t* out)
{
    __shared__ uint4 buf[kProbeLdsEntries];
    for (uint32_t i = threadIdx.x; i < kProbeLdsEntries; i += 256)
        buf[i] = make_uint4(i, 3 * i, 5 * i, 7 * i);
    __syncthreads();
    uint4 acc = make_uint4(0, 0, 0, 0);
    for (uint32_t it = 0; it < iters; it++) {
        // eight reads at 2 KB strides from a base that moves each pass
        const uint4* p = buf + ((threadIdx.x + it * 64u) & 1023u);
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const uint4 v = p[k * 128];
            acc.x ^= v.x;
            acc.y ^= v.y;
            acc.z ^= v.z;
            acc.w ^= v.w;
        }
    }
    if ((acc.x ^ acc.y ^ acc.z ^ acc.w) == 0x9e3779b9u && iters == 0u)
        out[blockIdx.x * blockDim.x + threadIdx.x] = acc.x;
}

hipError_t probe_peaks(int ncu, hipStream_t s, double* valu_f64, double* valu_int,
                       double* lds)
{
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t err = hipEventCreate(&e0);
    if (err == hipSuccess) err = hipEventCreate(&e1);
    const uint32_t grid = (uint32_t)ncu * 8u;
    const uint32_t it_valu = 8192, it_lds = 4096;
    auto timed = [&](auto launch, double work, double* rate) {
        for (int rep = 0; rep < 2 && err == hipSuccess; rep++) {
            err = hipEventRecord(e0, s);
            if (err == hipSuccess) {
                launch();
                err = hipGetLastError();
            }
            if (err == hipSuccess) err = hipEventRecord(e1, s);
            if (err == hipSuccess) err = hipEventSynchronize(e1);
            float ms = 0.0f;
            if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
            if (err == hipSuccess && ms > 0.0f) *rate = work / (ms * 1e-3);
        }
    };
    const double lanes = (double)grid * 256.0;
    timed([&] { hipLaunchKernelGGL(k_probe_valu_f64, dim3(grid), dim3(256), 0, s,
                                   it_valu, 0.5, (double*)nullptr); },
          lanes * 8.0 * it_valu, valu_f64);
    timed([&] { hipLaunchKernelGGL(k_probe_valu_int, dim3(grid), dim3(256), 0, s,
                                   it_valu, 12345u, (uint32_t*)nullptr); },
          lanes * 8.0 * it_valu, valu_int);
    timed([&] { hipLaunchKernelGGL(k_probe_lds, dim3(grid), dim3(256), 0, s,
                                   it_lds, (uint32_t*)nullptr); },
          lanes * 8.0 * 16.0 * it_lds, lds);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return err;
}

}  // namespace fk
