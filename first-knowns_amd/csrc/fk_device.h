// fk_device.h -- the structures the host planner hands to the kernels.
//
// Vocabulary.  Every family is a sieve over one integer line (Euler: p,
// Dickson: j with k = W j, AP: p).  A WHEEL of the first few kill primes of
// the family is enumerated rather than sieved: its admissible residues
// ("lines") off_i < D, with D the wheel period, leave candidates
// value = base + off_i + D * v, and line position v is what a tile's bit
// stands for.  For a sieve prime q the killed values F_q (family-specific)
// become the killed line positions K_q = F_q * D^-1 (mod q), shifted per
// line and window by t = (base + off_i) * D^-1 (mod q).  Small primes are
// OR-ed in as precomputed 64-bit pattern words, medium primes are either
// stride-marked or tested per surviving position against a kill bitmap,
// both in LDS, and primes above the LDS cutoff are tested per surviving
// candidate in compaction rounds.
#pragma once

#include <stdint.h>

#include "fk_arith.h"

namespace fk {

enum Family : int32_t { kEuler = 1, kDickson = 2, kAP = 3 };

// Per-engine constants and table pointers (device memory), read by every
// kernel through one pointer per window.
struct DevEngine {
    int32_t family;
    int32_t n;
    uint32_t nlines;          // wheel residues
    uint32_t nlds;            // LDS-stage primes (pattern, marking, test)
    uint32_t npat;            // [0, npat): pattern words
    uint32_t nmark;           // [npat, nmark): stride marking
    uint32_t nbits;           // [nmark, nbits): 64-bit windows of periodic
                              // kill bitmaps ORed into the tile's words;
                              // [nbits, nlds): per-survivor bitmap tests
    uint32_t npairs;          // (prime, residue) marking pairs
    uint32_t pat_words;       // total pattern words (sum of q)
    uint32_t bm_words;        // total kill-bitmap words of the test primes
    uint32_t round_first;     // global prime-table index of the first round prime
    uint32_t round_end;       // one past the last prime below q2
    uint32_t kstride;         // round-table row stride in words
    uint32_t kidx;            // 1: the round rows are AP index rows (no
                              // round prime divides d), 0: residue rows
    uint32_t nround;          // compaction rounds of this engine
    uint32_t rb[33];          // [nround + 1] prime-table index where each
                              // round starts (rb[0] = round_first, the last
                              // = round_end): q_med * growth^r
    uint64_t D;               // wheel period of the line
    uint32_t W;               // Dickson: k = W j (product of primes <= n+1)
    uint32_t d[8];            // AP: P(n), the common difference, in limbs
    const uint64_t* offs;     // [nlines] line residues off_i < D
    const uint32_t* lds_q;    // [nlds] primes
    const uint32_t* lds_m32;  // [nlds] floor(2^32 / q)
    const uint64_t* lds_m64;  // [nlds] floor(2^64 / q)
    const uint32_t* lds_dinv; // [nlds] D^-1 mod q
    const uint32_t* pat_off;  // [npat] first word of each prime's table
    const uint64_t* pat;      // [pat_words] pattern words
    const uint32_t* pairs;    // [npairs] (prime index << 16) | residue
    const uint32_t* bm_off;   // [nlds - nmark] first bitmap word per prime
    const uint32_t* bm;       // [bm_words] line-domain kill bitmaps, each
                              // q bits then its first 64 bits again
    uint32_t sbm_words;       // kill bitmaps of every LDS prime with a
    const uint32_t* sbm_off;  // tail of kShortTailBits: [nlds] first words,
    const uint32_t* sbm;      // [sbm_words] bitmaps (short-line sieve)
    const uint32_t* spd;      // [kShortPrimeWords nlds] the short-line
                              // sieve's per-prime constants, read with
                              // scalar loads
    const double* spe;        // [nlds] (2^32 D^-1) mod q, the same sieve's
                              // constant for wheel periods past 2^32
    uint32_t ntr;             // in-tile rounds over [nmark, nlds)
    const uint32_t* tr_end;   // [ntr] LDS-prime index where round r ends
    const uint32_t* kt;       // [round_end - round_first][kstride] round
                              // table: per round prime a RoundRow header
                              // (fk_kill.h), then the killing residues of
                              // the value, padded; or an IndexRow
    uint32_t pronic_first;    // Euler: prime-table index of the first round
                              // prime above (n - 1) n, whose rows the lane =
                              // candidate kernel tests without their
                              // residues; round_end for the other families
    const uint32_t* kperm;    // residue rows: [round_end - round_first] the
                              // rows of each round ordered by their groups
                              // of four residues (a permutation within
                              // every round), for the lane = prime kernel;
                              // null for index rows
};

// One piece of one window: a contiguous range of line positions of every
// line, small enough that positions fit 32 bits.
struct DevPiece {
    uint32_t base[4];         // value of (line 0 residue 0, position 0): B
    uint64_t lo_rel;          // window start minus B (values below are out)
    uint64_t hi_rel;          // window end minus B (values at/above are out)
    uint32_t npos;            // positions per line in this piece
    uint32_t nseg;            // tiles per line
    uint32_t tile0;           // first global tile index
    uint32_t engine;          // index into the engine array: the wheel and
                              // LDS tables the piece is sieved with
    uint32_t window;          // caller's window index
    uint32_t npw;             // short-line pieces: 64-bit words per line,
                              // tiles of 256 (line, chunk) items; 0: one
                              // tile per S positions of a line
    uint32_t queue;           // the engine whose queue segment and counter
                              // take the survivors: `engine` itself, or the
                              // engine whose prefix wheel `engine` is
};

// Short-line sieve: at most this many words per lane, and the bits its
// bitmaps repeat past q, so that a lane's words are the windows at x,
// x + 64, ... with no reduction mod q.
constexpr uint32_t kShortChunkMax = 16;
constexpr uint32_t kShortTailBits = 64 * kShortChunkMax;
// Words per LDS prime in the short-line sieve's constant table: q, 1/q and
// D^-1 mod q as doubles, q and the prime's first bitmap word (two 16-byte
// scalar loads).
constexpr uint32_t kShortPrimeWords = 8;

// A candidate in flight: the value the kill tests and the strong test need
// (Euler p, Dickson k = W j, AP p) and its window.
struct DevRec {
    uint32_t t[4];
    uint32_t window;
    int32_t run;
};

}  // namespace fk
