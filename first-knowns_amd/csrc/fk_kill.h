// fk_kill.h -- value-domain kill sets of the round primes.
//
// A round prime q kills a candidate when q divides one of the family's n
// values built from it; that depends only on the candidate's value mod q:
//   Euler   (value p)      q | x^2 + x + p   for some 0 <= x < n
//   Dickson (value k=W j)  q | m k^2 + 1     for some 1 <= m <= n
//   AP      (value p)      q | p + j d       for some 0 <= j < n, d = P(n)
// (euler-prime-runs/euler_search.py:85-87, dickson-ladders/
// ladder_gpu.py:597-614, primorial-ap/ap_gpu.py:88-97 state the same
// conditions).  The engine stores, per round prime, the killing residues
// without repeats, padded to a fixed stride with kKillPad (never a residue),
// and the kernels compare a candidate's residue against them.
//
// Dickson kill sets are closed under negation (the roots of m k^2 = -1 come
// in pairs r, q - r, and 0 is never one), so they are stored folded: one
// value min(r, q - r) per pair, and a residue x is tested as min(x, q - x).
#pragma once

#include <stddef.h>

#include "fk_arith.h"

namespace fk {

constexpr uint32_t kKillPad = 0xffffffffu;

// Largest stored kill set: one residue per value (Dickson: one folded root
// pair per m).
FK_HD uint32_t kill_max(int family, int n)
{
    (void)family;
    return (uint32_t)n;
}

// Dickson rows hold folded residues.
FK_HD bool kill_folded(int family) { return family == 2; }

FK_HD uint32_t pow_mod(uint32_t a, uint32_t e, const PrimeK& p)
{
    uint32_t r = 1u % p.q, b = a % p.q;
    while (e) {
        if (e & 1) r = mulmod32(r, b, p.q, p.m64);
        b = mulmod32(b, b, p.q, p.m64);
        e >>= 1;
    }
    return r;
}

// A square root of the quadratic residue a (0 < a < q) modulo an odd prime
// q, by Tonelli-Shanks.  Every loop is bounded: the exponent of 2 in q - 1
// is below 32, so is M, and M falls on each pass.
FK_HD uint32_t sqrt_mod(uint32_t a, const PrimeK& p)
{
    const uint32_t q = p.q;
    if ((q & 3u) == 3u) return pow_mod(a, (q + 1) / 4, p);
    uint32_t Q = q - 1, S = 0;
    while (!(Q & 1u)) {
        Q >>= 1;
        S++;
    }
    uint32_t z = 2;
    while (z < q && pow_mod(z, (q - 1) / 2, p) != q - 1) z++;
    uint32_t M = S;
    uint32_t c = pow_mod(z, Q, p);
    uint32_t t = pow_mod(a, Q, p);
    uint32_t R = pow_mod(a, (Q + 1) / 2, p);
    for (int pass = 0; pass < 32 && t != 1; pass++) {
        uint32_t i = 0, tt = t;
        while (tt != 1 && i < M) {
            tt = mulmod32(tt, tt, q, p.m64);
            i++;
        }
        if (i >= M) break;                 // a was not a residue
        uint32_t b = c;
        for (uint32_t k = 0; k + i + 1 < M; k++) b = mulmod32(b, b, q, p.m64);
        M = i;
        c = mulmod32(b, b, q, p.m64);
        t = mulmod32(t, c, q, p.m64);
        R = mulmod32(R, b, q, p.m64);
    }
    return R;
}

FK_HD void kill_add(uint32_t* out, uint32_t& cnt, uint32_t v)
{
    for (uint32_t i = 0; i < cnt; i++)
        if (out[i] == v) return;
    out[cnt++] = v;
}

// The kill set of the odd prime q (> n) for family/n, d = P(n) in four
// limbs (AP only), folded for Dickson.  Writes out[0, count), pads
// out[count, stride) and returns count.
FK_HD uint32_t kill_values(int family, int n, const uint32_t d[4],
                           const PrimeK& p, uint32_t* out, uint32_t stride)
{
    const uint32_t q = p.q;
    uint32_t cnt = 0;
    if (family == 1) {
        for (int x = 0; x < n; x++) {
            const uint32_t v = red64((uint64_t)x * x + (uint64_t)x, q, p.m64);
            kill_add(out, cnt, v ? q - v : 0u);
        }
    } else if (family == 2) {
        for (int m = 1; m <= n; m++) {
            const uint32_t mq = (uint32_t)m % q;
            if (mq == 0) continue;             // m k^2 + 1 == 1 (mod q)
            const uint32_t a = q - pow_mod(mq, q - 2, p);    // -1/m
            if (pow_mod(a, (q - 1) / 2, p) != 1) continue;
            const uint32_t r = sqrt_mod(a, p);
            kill_add(out, cnt, r < q - r ? r : q - r);
        }
    } else {
        const uint32_t dq = red128(d, p);
        for (int j = 0; j < n; j++) {
            const uint32_t v = mulmod32((uint32_t)j % q, dq, q, p.m64);
            kill_add(out, cnt, v ? q - v : 0u);
        }
    }
    for (uint32_t i = cnt; i < stride; i++) out[i] = kKillPad;
    return cnt;
}

// One round prime's row in an engine's round table: everything a wave
// needs to test its candidates against the prime, contiguous, so that two
// 64-byte scalar loads fetch it.  Words [0, 12) are this header, words
// [12, stride) the killing residues (kKillPad past `cnt`).
struct RoundRow {
    double qd;        // q
    double qi;        // nearest double to 1/q
    double c16;       // 2^16 mod q
    double c32;       // 2^32 mod q
    double c48;       // 2^48 mod q
    uint32_t q;
    uint32_t cnt;     // killing residues
};
constexpr uint32_t kRowHead = 12;           // words of RoundRow
static_assert(sizeof(RoundRow) == 4 * kRowHead, "RoundRow layout");

// Row stride in words: the header plus the largest kill set in whole groups
// of four residues (the round kernels compare four at a time), and whole
// 64-byte lines past 32 words.  Dickson n <= 12 rows are 24 words, so the
// two rows of a lane = candidate pass fit the scalar registers together.
FK_HD uint32_t round_row_stride(int family, int n)
{
    const uint32_t w = kRowHead + ((kill_max(family, n) + 3u) & ~3u);
    return w <= 32 ? w : (w + 15u) & ~15u;
}

// Words a round table carries past its last row: the kernels read rows in
// a fixed size (up to 32 words) and ignore words past the row's count.
constexpr uint32_t kRowTail = 32;

FK_HD void round_row(int family, int n, const uint32_t d[4], const PrimeK& p,
                     uint32_t* row, uint32_t stride)
{
    const PrimeD pd = make_primed(p.q);
    RoundRow h;
    h.qd = pd.qd;
    h.qi = pd.qi;
    h.c16 = pd.c[0];
    h.c32 = pd.c[1];
    h.c48 = pd.c[2];
    h.q = p.q;
    h.cnt = kill_values(family, n, d, p, row + kRowHead, stride - kRowHead);
    uint32_t hw[kRowHead];
    __builtin_memcpy(hw, &h, sizeof h);          // no type punning
    for (uint32_t i = 0; i < kRowHead; i++) row[i] = hw[i];
}

// ---- index rows (AP): the kill set of a prime q that does not divide d
// is {-j d mod q : 0 <= j < n}, n distinct residues, so q kills p exactly
// when t = -p / d mod q is below n.  t is a linear form in p's 16-bit
// limbs: t = sum_k a_k (u 2^(16k) mod q) mod q with u = -1/d mod q, the
// same fused-multiply-add sum and exact reduction as a residue, so one
// compare replaces n.  (A prime dividing d kills p iff q | p; an engine
// with such a round prime keeps the residue rows.)  For q < 2^24 the same
// form in 22-bit limbs is exact too (each product below 2^46, four of them
// below 2^48) and takes p to 2^88 in four limbs (three below 2^66), where
// 16-bit limbs take eight past 2^64.
// The first 64 bytes hold all that the 22-bit form reads, so a round of
// 22-bit tests fetches half of each row (a million-prime round of a few
// candidates is bound by those fetches).
struct IndexRow {
    double qd;        // q
    double qi;        // nearest double to 1/q
    double v[4];      // u 2^(22k) mod q, k = 0..3 (22-bit limbs, q < 2^24)
    uint32_t q;
    uint32_t n;       // t < n kills
    uint32_t pad[2];
    double u[4];      // u 2^(16k) mod q, k = 0..3
    double uw[4];     // u 2^(16k) mod q, k = 4..7 (values past 2^64)
};
constexpr uint32_t kIndexStride = 32;       // words of IndexRow
constexpr uint32_t kIndexV = 4;             // first word of v
constexpr uint32_t kIndexU = 16;            // first word of u
constexpr uint32_t kIndexUW = 24;           // first word of uw
constexpr uint32_t kIndexHalf = 16;         // words the 22-bit form reads
static_assert(sizeof(IndexRow) == 4 * kIndexStride, "IndexRow layout");
static_assert(offsetof(IndexRow, v) == 4 * kIndexV, "IndexRow layout");
static_assert(offsetof(IndexRow, q) == 48 && offsetof(IndexRow, n) == 52,
              "IndexRow layout");
static_assert(offsetof(IndexRow, u) == 4 * kIndexU, "IndexRow layout");
static_assert(offsetof(IndexRow, uw) == 4 * kIndexUW, "IndexRow layout");

// Families whose rows may be index rows.
FK_HD bool kill_indexable(int family) { return family == 3; }

// The index row of q (q must not divide d).
FK_HD void index_row(int n, const uint32_t d[4], const PrimeK& p, uint32_t* row)
{
    const PrimeD pd = make_primed(p.q);
    const uint32_t dq = red128(d, p);                   // d mod q
    const uint32_t di = pow_mod(dq, p.q - 2, p);          // 1/d, q prime
    uint32_t c = di ? p.q - di : 0u;                      // u = -1/d
    const uint32_t two16 = (uint32_t)((1ull << 16) % p.q);
    const uint32_t two22 = (uint32_t)((1ull << 22) % p.q);
    IndexRow h;
    h.qd = pd.qd;
    h.qi = pd.qi;
    uint32_t c22 = c;
    for (int k = 0; k < 4; k++) {
        h.v[k] = (double)c22;
        c22 = mulmod32(c22, two22, p.q, p.m64);
    }
    for (int k = 0; k < 8; k++) {
        if (k < 4) h.u[k] = (double)c;
        else h.uw[k - 4] = (double)c;
        c = mulmod32(c, two16, p.q, p.m64);
    }
    h.q = p.q;
    h.n = (uint32_t)n;
    h.pad[0] = h.pad[1] = 0;
    uint32_t hw[kIndexStride];
    __builtin_memcpy(hw, &h, sizeof h);          // no type punning
    for (uint32_t i = 0; i < kIndexStride; i++) row[i] = hw[i];
}

}  // namespace fk
