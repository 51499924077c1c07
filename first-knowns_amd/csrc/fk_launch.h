// fk_launch.h -- kernel argument blocks and the host-side launchers.
#pragma once

#include <hip/hip_runtime.h>

#include "fk_device.h"

namespace fk {

// Classification request of one window (mode 0: none, 1: all seven bases,
// 2: base-2 chain first, redone with all bases from exact_from up).
struct DevWinInfo {
    uint32_t engine;
    int32_t mode;
    int32_t cap;              // chain values looked at (at most 64 per pass)
    int32_t exact_from;       // relative to `first`
    int32_t first;            // chain index of the pass's first value
};

// Queue segments: engine e owns entries [segs[e].x, segs[e].x + segs[e].y)
// of both queues (and of the flag array) for a whole batch, so that every
// wave of the round kernels sees one engine.  A stage's counters hold one
// count per engine (capped at the segment size where read).
struct SieveArgs {
    const DevEngine* engines;
    const DevPiece* pieces;
    uint32_t npieces;
    uint32_t tile_begin;      // this launch's tiles: [tile_begin, ntiles)
    uint32_t ntiles;
    uint32_t nlds_max;        // LDS room for per-prime arrays (max over engines)
    uint32_t pat_max;         // LDS room for pattern words (max over engines)
    uint32_t bm_max;          // LDS room for kill-bitmap words (max over engines)
    const uint2* segs;        // per engine: queue segment (base, size)
    DevRec* out;
    uint32_t* counts;         // per engine: records emitted
    uint32_t* flags;          // per queue entry: cleared for every record
                              // emitted (the rounds' kill flags)
};

// lane = candidate rounds: candidates per lane (a workgroup takes 256 of
// them per lane slot).  Two per lane halve the scalar row loads per test but
// spill scalar registers and halve the waves; on the Dickson n = 10 sweep
// the rounds took 0.35 ms per window against 0.28 ms with one.
constexpr int kRoundKC = 1;

struct RoundArgs {
    const DevEngine* engines;
    uint32_t nengines;
    const uint2* segs;
    const PrimeD* primed;     // the device-wide prime table (residue form)
    const DevRec* recs;
    const uint32_t* counts;   // per engine: candidates in the queue
    uint32_t* flags;
    uint32_t round;           // round index (each engine's rb[round..])
    uint32_t p_begin, p_end;  // global prime indices: the union of the
                              // engines' round-`round` primes
    uint32_t pblock;          // lane = candidate: primes per blockIdx.y
    uint32_t kgroups;         // lane = candidate: groups of four killing
                              // residues held in registers per row (3 or 5;
                              // rows with more are finished in a loop)
    uint32_t q_min;           // the round's smallest prime (the one-value
                              // residue form needs q >= 32)
    uint32_t rows;            // the launch tests the engines whose round
                              // rows are of this kind (kRowsResidue or
                              // kRowsIndex) and skips the others
    uint32_t l22;             // every prime of the launch below 2^24: the
                              // 22-bit limb forms apply (values below 2^88)
};
constexpr uint32_t kRowsResidue = 1, kRowsIndex = 2;
// lane = prime rounds: the most candidates a workgroup stages per slice
constexpr uint32_t kRoundSlice = 256;

struct CompactArgs {
    uint32_t nengines;
    const uint2* segs;
    const DevRec* in;
    const uint32_t* in_counts;
    uint32_t* flags;
    DevRec* out;
    uint32_t* out_counts;
};

// The final queue packed for classification and the copy back: every
// engine's segment, in engine order, from out[0]; *total = the record count.
struct GatherArgs {
    uint32_t nengines;
    const uint2* segs;
    const DevRec* in;
    const uint32_t* counts;
    DevRec* out;
    uint32_t* total;
};

// The readback: the stage counters, then the first min(*count, cap)
// records of the final queue, written straight into pinned host memory.
struct ReadbackArgs {
    const uint32_t* counts;
    uint32_t ncounts;
    uint32_t* counts_host;
    const DevRec* recs;
    const uint32_t* count;    // the final record count
    uint32_t cap;             // records the host buffer holds
    DevRec* recs_host;
};

struct ClassifyArgs {
    const DevEngine* engines;
    const DevWinInfo* wins;
    DevRec* recs;
    const uint32_t* count;
    uint32_t cap;
    uint32_t cap_max;         // largest chain cap of any window
    uint64_t* fail2;          // per record: values failing base 2
    uint64_t* failx;          // per record: values failing bases 3..7
    uint32_t single_limit;    // records x cap_max at or below which all
                              // bases run in one pass
};

// Events the launch itself stamps at its kernel's start and end (timed
// eager submissions; null otherwise, and never under graph capture).
struct LaunchEvents {
    hipEvent_t start = nullptr;
    hipEvent_t stop = nullptr;
};

hipError_t launch_sieve(const SieveArgs& a, int threads, int wpt,
                        uint32_t grid, size_t lds_bytes, hipStream_t stream,
                        const LaunchEvents* ev = nullptr);
int sieve_occupancy(int threads, int wpt, size_t lds_bytes);
// windows with short lines: one lane per chunk of `chunk` words of a line
// (1, 2, 4, 8 or 16), 256-item tiles
// LDS bytes of k_sieve_short: one 8-byte word per LDS prime for the
// piece's base term (whole 16-byte groups), then the kill bitmaps.
inline size_t short_lds(size_t nlds, size_t bm_words)
{
    return 8 * ((nlds + 1) & ~(size_t)1) + 4 * bm_words;
}
hipError_t launch_sieve_short(const SieveArgs& a, int chunk, uint32_t grid,
                              size_t lds_bytes, hipStream_t stream,
                              const LaunchEvents* ev = nullptr);
int sieve_short_occupancy(int chunk, size_t lds_bytes);
bool sieve_short_chunk_ok(int chunk);
bool sieve_shape_ok(int threads, int wpt);
hipError_t launch_round_test(const RoundArgs& a, bool lane_prime, uint32_t gx,
                             uint32_t gy, hipStream_t stream,
                             const LaunchEvents* ev = nullptr);
hipError_t launch_kill_table(const PrimeK* primes, uint32_t p0, uint32_t np,
                             int family, int n, const uint32_t d[4], bool index,
                             uint32_t stride, uint32_t* kt, hipStream_t stream);
hipError_t launch_round_compact(const CompactArgs& a, uint32_t grid,
                                hipStream_t stream,
                                const LaunchEvents* ev = nullptr);
hipError_t launch_readback(const ReadbackArgs& a, hipStream_t stream);
hipError_t launch_gather(const GatherArgs& a, uint32_t grid,
                         hipStream_t stream);
hipError_t launch_classify(const ClassifyArgs& a, uint32_t grid,
                           hipStream_t stream);
hipError_t launch_zero_u64(uint64_t* p, size_t n, hipStream_t stream);
// Measured peaks (one probe launch each, timed twice with HIP events on
// `s`): VALU lane-operations per second for double FMAs and 32-bit funnel
// shifts, LDS bytes read per second.
hipError_t probe_peaks(int ncu, hipStream_t s, double* valu_f64, double* valu_int,
                       double* lds);
hipError_t launch_wheel_expand(const uint64_t* base, uint64_t nbase,
                               const uint32_t* lifts, uint32_t nj, uint32_t L,
                               uint64_t Mb, uint64_t* out, hipStream_t stream);

}  // namespace fk
