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
//                   (workgroup = prime chunk, lanes = candidates), killed
//                   candidates flagged
//   k_round_compact wave-ballot compaction of the unflagged candidates
//   k_classify      strong-probable-prime chains (Montgomery, 2..8 limbs)
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

template <int T, int WPT>
__global__ __launch_bounds__(T) void k_sieve(SieveArgs a)
{
    extern __shared__ uint64_t lds64[];
    constexpr uint32_t S = 64u * T * WPT;          // line positions per tile
    uint32_t* bits = reinterpret_cast<uint32_t*>(lds64);     // S/32 words
    uint32_t* s0 = bits + S / 32;                             // [nlds_max]
    uint64_t* pat = lds64 + (S / 64) + ((a.nlds_max + 1) / 2);  // patterns
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
            for (; pos < ulen; pos += q)
                atomicOr(&bits[pos >> 5], 1u << (pos & 31));
        }
        __syncthreads();

        // ---- compaction: surviving bits -> candidate records
        uint64_t live[WPT];
        uint32_t cnt = 0;
#pragma unroll
        for (int k = 0; k < WPT; k++) {
            const uint32_t w = tid + (uint32_t)k * T;
            live[k] = ~(((uint64_t)bits[2 * w + 1] << 32) | bits[2 * w]);
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

// Does prime p divide one of the family's values for this candidate?
__device__ __forceinline__ bool kill_test(const DevEngine& E,
                                          const uint32_t t[4],
                                          const PrimeK& pk)
{
    const uint32_t q = pk.q;
    const uint32_t r = red128(t, pk);
    if (E.family == kEuler) {
        // q | x^2+x+p  <=>  (q - p mod q) mod q == x^2+x (mod q)
        if (q > E.xxmax) {
            const uint32_t dd = r ? q - r : 0;
            return dd <= E.xxmax && ((E.xxbits[dd >> 5] >> (dd & 31)) & 1);
        }
        for (int x = 0; x < E.n; x++) {
            const uint64_t v = (uint64_t)x * x + x + r;
            if (v % q == 0) return true;
        }
        return false;
    } else if (E.family == kDickson) {
        // t = k mod q; q | m k^2 + 1 for some m <= n
        const uint32_t kk = mulmod32(r, r, q, pk.m64);
        if (kk == 0) return false;
        uint32_t v = kk + 1;
        if (v >= q) v -= q;
        for (int m = 1; m <= E.n; m++) {
            if (v == 0) return true;
            v += kk;
            if (v >= q) v -= q;
        }
        return false;
    } else {
        // q | p + j d for some j < n
        const uint32_t dq = red128(E.d, pk);
        if (dq == 0) return r == 0;
        uint32_t f = r;
        for (int j = 0; j < E.n; j++) {
            if (f == 0) return true;
            f += dq;
            if (f >= q) f -= q;
        }
        return false;
    }
}

__global__ __launch_bounds__(256) void k_round_test(RoundArgs a)
{
    const uint32_t count = min(*a.count, a.cap);
    for (uint32_t chunk = blockIdx.x; chunk < a.nchunks;
         chunk += gridDim.x) {
        const uint32_t c0 = a.p_begin + chunk * a.chunk;
        const uint32_t c1 = min(c0 + a.chunk, a.p_end);
        for (uint32_t i = blockIdx.y * blockDim.x + threadIdx.x; i < count;
             i += gridDim.y * blockDim.x) {
            if (a.flags[i]) continue;
            const DevRec rec = a.recs[i];
            const DevEngine& E = a.engines[a.win_engine[rec.window]];
            const uint32_t lo = max(c0, E.round_first);
            const uint32_t hi = min(c1, E.round_end);
            for (uint32_t p = lo; p < hi; p++) {
                if (kill_test(E, rec.t, a.primes[p])) {
                    a.flags[i] = 1;
                    break;
                }
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

__device__ int chain_length(const DevEngine& E, const uint32_t t[4],
                            const uint32_t kk[8], int cap, int nbases)
{
    int i = 0;
    while (i < cap) {
        uint32_t v[8];
        chain_value(E, t, kk, i, v);
        if (!is_sprp(v, nbases)) break;
        i++;
    }
    return i;
}

__global__ __launch_bounds__(128) void k_classify(ClassifyArgs a)
{
    const uint32_t count = min(*a.count, a.cap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += gridDim.x * blockDim.x) {
        DevRec rec = a.recs[i];
        const DevWinInfo W = a.wins[rec.window];
        if (W.mode == 0) continue;
        const DevEngine& E = a.engines[W.engine];
        uint32_t kk[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (E.family == kDickson) mw_mul44(kk, rec.t, rec.t);
        int run;
        if (W.mode == 2) {
            run = chain_length(E, rec.t, kk, W.cap, 1);
            if (run >= W.exact_from) run = chain_length(E, rec.t, kk, W.cap, 7);
        } else {
            run = chain_length(E, rec.t, kk, W.cap, 7);
        }
        a.recs[i].run = run;
    }
}

// Standalone classification of caller-supplied values (the wrapper's
// run_length / chain_depth entry points).
__global__ __launch_bounds__(128) void k_classify_values(
    const DevEngine* __restrict__ E, const uint32_t* __restrict__ vals,
    uint32_t n, int mode, int cap, int exact_from, int32_t* __restrict__ out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += gridDim.x * blockDim.x) {
        uint32_t t[4] = {vals[4 * i], vals[4 * i + 1], vals[4 * i + 2],
                         vals[4 * i + 3]};
        uint32_t kk[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (E->family == kDickson) mw_mul44(kk, t, t);
        int run;
        if (mode == 2) {
            run = chain_length(*E, t, kk, cap, 1);
            if (run >= exact_from) run = chain_length(*E, t, kk, cap, 7);
        } else {
            run = chain_length(*E, t, kk, cap, 7);
        }
        out[i] = run;
    }
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += gridDim.x * blockDim.x)
        p[i] = 0;
}

// --------------------------------------------------------------- launchers

#define FK_SIEVE_CASE(TT, WW)                                              \
    if (threads == TT && wpt == WW) {                                      \
        auto fn = k_sieve<TT, WW>;                                         \
        if (lds_bytes > 65536)                                             \
            (void)hipFuncSetAttribute((const void*)fn,                     \
                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes); \
        hipLaunchKernelGGL(fn, dim3(grid), dim3(TT), lds_bytes, stream, a); \
        return hipGetLastError();                                          \
    }

hipError_t launch_sieve(const SieveArgs& a, int threads, int wpt,
                        uint32_t grid, size_t lds_bytes, hipStream_t stream)
{
    FK_SIEVE_CASE(256, 1) FK_SIEVE_CASE(256, 2) FK_SIEVE_CASE(256, 4)
    FK_SIEVE_CASE(256, 8) FK_SIEVE_CASE(256, 16)
    FK_SIEVE_CASE(512, 1) FK_SIEVE_CASE(512, 2) FK_SIEVE_CASE(512, 4)
    FK_SIEVE_CASE(512, 8)
    FK_SIEVE_CASE(1024, 1) FK_SIEVE_CASE(1024, 2) FK_SIEVE_CASE(1024, 4)
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
    FK_OCC_CASE(256, 1) FK_OCC_CASE(256, 2) FK_OCC_CASE(256, 4)
    FK_OCC_CASE(256, 8) FK_OCC_CASE(256, 16)
    FK_OCC_CASE(512, 1) FK_OCC_CASE(512, 2) FK_OCC_CASE(512, 4)
    FK_OCC_CASE(512, 8)
    FK_OCC_CASE(1024, 1) FK_OCC_CASE(1024, 2) FK_OCC_CASE(1024, 4)
#undef FK_OCC_CASE
    return 0;
}

bool sieve_shape_ok(int threads, int wpt)
{
    return sieve_occupancy(threads, wpt, 1024) >= 0 &&
           ((threads == 256 && (wpt == 1 || wpt == 2 || wpt == 4 ||
                                wpt == 8 || wpt == 16)) ||
            (threads == 512 && (wpt == 1 || wpt == 2 || wpt == 4 ||
                                wpt == 8)) ||
            (threads == 1024 && (wpt == 1 || wpt == 2 || wpt == 4)));
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
    if (gx > 65535) gx = 65535;
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
    hipLaunchKernelGGL(k_classify, dim3(grid ? grid : 1), dim3(128), 0,
                       stream, a);
    return hipGetLastError();
}

hipError_t launch_classify_values(const DevEngine* E, const uint32_t* vals,
                                  uint32_t n, int mode, int cap,
                                  int exact_from, int32_t* out,
                                  hipStream_t stream)
{
    uint32_t grid = (n + 127) / 128;
    if (grid < 1) grid = 1;
    if (grid > 4096) grid = 4096;
    hipLaunchKernelGGL(k_classify_values, dim3(grid), dim3(128), 0, stream,
                       E, vals, n, mode, cap, exact_from, out);
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
