// fk_arith.h -- exact integer arithmetic shared by the device kernels and
// the host planner: Barrett reduction by 32-bit primes, multiword values in
// 32-bit limbs, Montgomery multiplication and the strong-probable-prime
// (Miller-Rabin) test with the reference's seven bases.
//
// Everything here is `__host__ __device__` so the host planner and the
// host-side self test use the very code the kernels run.  Limbs are
// little-endian uint32_t; a multiword value of L limbs lives in a plain
// array.  The 32x32->64 multiply-add pattern `(uint64_t)a * b + c + d` is
// what gfx950 executes as v_mad_u64_u32.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define FK_HD __host__ __device__ __forceinline__
#else
#define FK_HD inline
#endif

namespace fk {

// huntlib/primes.py:17-20
constexpr int kNumBases = 7;
constexpr int kNumSmall = 12;

FK_HD uint32_t mr_base(int i)
{
    // 2, 325, 9375, 28178, 450775, 9780504, 1795265022
    return i == 0 ? 2u : i == 1 ? 325u : i == 2 ? 9375u : i == 3 ? 28178u
         : i == 4 ? 450775u : i == 5 ? 9780504u : 1795265022u;
}

FK_HD uint32_t small_prime(int i)
{
    // 2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37
    return i < 6 ? (i < 3 ? (i == 0 ? 2u : i == 1 ? 3u : 5u)
                          : (i == 3 ? 7u : i == 4 ? 11u : 13u))
                 : (i < 9 ? (i == 6 ? 17u : i == 7 ? 19u : 23u)
                          : (i == 9 ? 29u : i == 10 ? 31u : 37u));
}

FK_HD uint64_t mulhi64(uint64_t a, uint64_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// ------------------------------------------------------------ Barrett
// For a prime q < 2^30, m64 = floor(2^64 / q).  qhat = mulhi(x, m64) is at
// most two below floor(x / q) (huntlib/gpu.py:8-28 states the same bound),
// so two conditional subtracts make the remainder exact.
struct PrimeK {
    uint32_t q;
    uint32_t c1, c2, c3;   // 2^32, 2^64, 2^96 mod q
    uint64_t m64;          // floor(2^64 / q)
};

FK_HD uint32_t red64(uint64_t x, uint32_t q, uint64_t m64)
{
    uint64_t qh = mulhi64(x, m64);
    uint64_t r = x - qh * q;
    if (r >= q) r -= q;
    if (r >= q) r -= q;
    return (uint32_t)r;
}

FK_HD uint32_t mulmod32(uint32_t a, uint32_t b, uint32_t q, uint64_t m64)
{
    return red64((uint64_t)a * b, q, m64);
}

// A 128-bit value in four limbs, reduced mod q < 2^30: each limb times
// 2^(32i) mod q stays below 2^62, so the sum of four fits in 64 bits.
FK_HD uint32_t red128(const uint32_t t[4], const PrimeK& p)
{
    uint64_t acc = (uint64_t)t[0] + (uint64_t)t[1] * p.c1
                 + (uint64_t)t[2] * p.c2 + (uint64_t)t[3] * p.c3;
    return red64(acc, p.q, p.m64);
}

// 32-bit Barrett for x < 2^32: m32 = floor(2^32 / q), one correction.
FK_HD uint32_t red32(uint32_t x, uint32_t q, uint32_t m32)
{
    uint32_t qh = (uint32_t)(((uint64_t)x * m32) >> 32);
    uint32_t r = x - qh * q;
    if (r >= q) r -= q;
    if (r >= q) r -= q;
    return r;
}

FK_HD PrimeK make_primek(uint32_t q)
{
    PrimeK p;
    p.q = q;
    uint64_t t = ((uint64_t)1 << 32) % q;
    p.c1 = (uint32_t)t;
    p.c2 = (uint32_t)((t * t) % q);
    p.c3 = (uint32_t)(((uint64_t)p.c2 * t) % q);
    p.m64 = (q == 1) ? 0 : (uint64_t)(~(uint64_t)0) / q;
    // floor((2^64 - 1) / q) == floor(2^64 / q) unless q | 2^64, i.e. never
    // for an odd prime; q = 2 gives 2^63 either way.
    if (q == 2) p.m64 = (uint64_t)1 << 63;
    return p;
}

// ------------------------------------------- floating-point residues
// A value split into 16-bit limbs a_i is congruent mod q to
// s = sum a_i * c_i with c_i = 2^(16 i) mod q.  For q < 2^30 and at most
// eight limbs every product is below 2^46 and s below 2^49, so fused
// multiply-adds in double precision form s exactly.  With qi the double
// nearest 1/q, floor(s * qi) is within one of floor(s / q) (the relative
// error of s * qi is below 2^-52 and s / q below 2^49), so s - floor(s*qi) q,
// formed exactly by one more fma, lies in [-q, 2q) and two unsigned min
// steps bring it into [0, q).  The double-precision pipe runs these at full
// rate, where 32-bit integer multiplies are quarter rate.
struct PrimeD {
    uint32_t q;
    uint32_t pad;
    double qd;             // q
    double qi;             // nearest double to 1/q
    double c[7];           // 2^16, 2^32, ..., 2^112 mod q
};

FK_HD PrimeD make_primed(uint32_t q)
{
    PrimeD p;
    p.qd = (double)q;
    p.qi = 1.0 / (double)q;
    uint64_t c = 1;
    for (int i = 0; i < 7; i++) {
        c = (c << 16) % q;
        p.c[i] = (double)c;
    }
    p.q = q;
    p.pad = 0;
    return p;
}

// s (an integer below 2^49 held exactly) mod q, for q < 2^30.
FK_HD uint32_t resid_d(double s, double qd, double qi, uint32_t q)
{
    const double qt = floor(s * qi);
    const double r = fma(-qt, qd, s);            // exact, in [-q, 2q)
    uint32_t u = (uint32_t)(int32_t)r;
    const uint32_t up = u + q;                   // fixes [-q, 0)
    u = u < up ? u : up;
    const uint32_t dn = u - q;                   // fixes [q, 2q)
    return u < dn ? u : dn;
}

// ----------------------------------------------------- multiword values
// value = sum limb[i] * 2^(32 i); L is the limb count.

template <int L>
FK_HD int mw_cmp(const uint32_t* a, const uint32_t* b)
{
    for (int i = L - 1; i >= 0; i--) {
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    }
    return 0;
}

template <int L>
FK_HD uint32_t mw_sub(uint32_t* r, const uint32_t* a, const uint32_t* b)
{
    uint64_t borrow = 0;
    for (int i = 0; i < L; i++) {
        uint64_t d = (uint64_t)a[i] - b[i] - borrow;
        r[i] = (uint32_t)d;
        borrow = (d >> 32) & 1;
    }
    return (uint32_t)borrow;
}

template <int L>
FK_HD uint32_t mw_add(uint32_t* r, const uint32_t* a, const uint32_t* b)
{
    uint64_t c = 0;
    for (int i = 0; i < L; i++) {
        uint64_t s = (uint64_t)a[i] + b[i] + c;
        r[i] = (uint32_t)s;
        c = s >> 32;
    }
    return (uint32_t)c;
}

template <int L>
FK_HD bool mw_is_zero(const uint32_t* a)
{
    uint32_t o = 0;
    for (int i = 0; i < L; i++) o |= a[i];
    return o == 0;
}

// value mod a small constant divisor (Horner over the limbs)
template <int L>
FK_HD uint32_t mw_mod_small(const uint32_t* a, uint32_t d)
{
    uint64_t r = 0;
#pragma unroll
    for (int i = L - 1; i >= 0; i--) r = ((r << 32) | a[i]) % d;
    return (uint32_t)r;
}

// -inverse of an odd a modulo 2^32 (Newton: each step doubles the bits)
FK_HD uint32_t neg_inv32(uint32_t a)
{
    uint32_t x = a;                 // correct to 3 bits for odd a
    for (int i = 0; i < 5; i++) x *= 2u - a * x;
    return 0u - x;
}

// Montgomery product r = a * b * 2^(-32 L) mod m, for odd m < 2^(32 L)
// (coarsely integrated operand scanning).  a, b < m; r < m.
template <int L>
FK_HD void mont_mul(uint32_t* r, const uint32_t* a, const uint32_t* b,
                    const uint32_t* m, uint32_t minv)
{
    uint32_t t[L + 2];
    for (int i = 0; i < L + 2; i++) t[i] = 0;
    for (int i = 0; i < L; i++) {
        uint64_t c = 0;
        const uint32_t bi = b[i];
        for (int j = 0; j < L; j++) {
            uint64_t s = (uint64_t)a[j] * bi + t[j] + c;
            t[j] = (uint32_t)s;
            c = s >> 32;
        }
        uint64_t s = (uint64_t)t[L] + c;
        t[L] = (uint32_t)s;
        t[L + 1] = (uint32_t)(s >> 32);
        const uint32_t u = t[0] * minv;
        c = ((uint64_t)u * m[0] + t[0]) >> 32;
        for (int j = 1; j < L; j++) {
            uint64_t s2 = (uint64_t)u * m[j] + t[j] + c;
            t[j - 1] = (uint32_t)s2;
            c = s2 >> 32;
        }
        s = (uint64_t)t[L] + c;
        t[L - 1] = (uint32_t)s;
        t[L] = t[L + 1] + (uint32_t)(s >> 32);
    }
    // t < 2m: one conditional subtract
    uint32_t d[L];
    uint32_t borrow = mw_sub<L>(d, t, m);
    const bool use_t = (t[L] == 0) && borrow;
    for (int i = 0; i < L; i++) r[i] = use_t ? t[i] : d[i];
}

// Montgomery square r = a^2 * 2^(-32 L) mod m: the cross products are
// formed once and doubled, then the 2L-limb square is reduced limb by limb
// (separated operand scanning).  a < m; r < m.
template <int L>
FK_HD void mont_sqr(uint32_t* r, const uint32_t* a, const uint32_t* m,
                    uint32_t minv)
{
    uint32_t t[2 * L + 1];
#pragma unroll
    for (int i = 0; i < 2 * L + 1; i++) t[i] = 0;
    // cross products a_i a_j, i < j
#pragma unroll
    for (int i = 0; i < L; i++) {
        uint64_t c = 0;
#pragma unroll
        for (int j = i + 1; j < L; j++) {
            uint64_t s = (uint64_t)a[i] * a[j] + t[i + j] + c;
            t[i + j] = (uint32_t)s;
            c = s >> 32;
        }
        t[i + L] = (uint32_t)c;
    }
    // double them
    uint32_t top = 0;
#pragma unroll
    for (int i = 1; i < 2 * L; i++) {
        uint32_t nt = t[i] >> 31;
        t[i] = (t[i] << 1) | top;
        top = nt;
    }
    // add the squares a_i^2
    uint64_t c = 0;
#pragma unroll
    for (int i = 0; i < L; i++) {
        uint64_t sq = (uint64_t)a[i] * a[i];
        uint64_t s0 = (uint64_t)t[2 * i] + (uint32_t)sq + c;
        t[2 * i] = (uint32_t)s0;
        uint64_t s1 = (uint64_t)t[2 * i + 1] + (sq >> 32) + (s0 >> 32);
        t[2 * i + 1] = (uint32_t)s1;
        c = s1 >> 32;
    }
    // reduce: add u_i m 2^(32 i) so the low L limbs vanish
#pragma unroll
    for (int i = 0; i < L; i++) {
        const uint32_t u = t[i] * minv;
        uint64_t cc = 0;
#pragma unroll
        for (int j = 0; j < L; j++) {
            uint64_t s = (uint64_t)u * m[j] + t[i + j] + cc;
            t[i + j] = (uint32_t)s;
            cc = s >> 32;
        }
#pragma unroll
        for (int k = i + L; k < 2 * L + 1; k++) {
            uint64_t s = (uint64_t)t[k] + cc;
            t[k] = (uint32_t)s;
            cc = s >> 32;
        }
    }
    // result t[L..2L] < 2m: one conditional subtract
    uint32_t d[L];
    uint32_t borrow = mw_sub<L>(d, t + L, m);
    const bool use_t = (t[2 * L] == 0) && borrow;
#pragma unroll
    for (int i = 0; i < L; i++) r[i] = use_t ? t[L + i] : d[i];
}

// r = 2 r mod m for r < m (m < 2^(32 L))
template <int L>
FK_HD void mw_dbl_mod(uint32_t* r, const uint32_t* m)
{
    const uint32_t top = r[L - 1] >> 31;
#pragma unroll
    for (int i = L - 1; i > 0; i--) r[i] = (r[i] << 1) | (r[i - 1] >> 31);
    r[0] <<= 1;
    uint32_t d[L];
    const uint32_t borrow = mw_sub<L>(d, r, m);
    const bool sub = top || !borrow;
#pragma unroll
    for (int i = 0; i < L; i++) r[i] = sub ? d[i] : r[i];
}

// r = (r + b) mod m for r, b < m
template <int L>
FK_HD void mw_add_mod(uint32_t* r, const uint32_t* b, const uint32_t* m)
{
    const uint32_t carry = mw_add<L>(r, r, b);
    uint32_t d[L];
    const uint32_t borrow = mw_sub<L>(d, r, m);
    const bool sub = carry || !borrow;
#pragma unroll
    for (int i = 0; i < L; i++) r[i] = sub ? d[i] : r[i];
}

// R mod m with R = 2^(32 L): start from the top bit of m (2^(B-1) < m for
// odd m > 1 of bit length B) and double the remaining 32 L - B + 1 times.
template <int L>
FK_HD void mont_one(uint32_t* r, const uint32_t* m)
{
    int B = 0;
#pragma unroll
    for (int i = 0; i < L; i++)
        if (m[i]) B = 32 * i + 32 - __builtin_clz(m[i]);
    const int tb = B - 1;
#pragma unroll
    for (int i = 0; i < L; i++) r[i] = (i == (tb >> 5)) ? (1u << (tb & 31)) : 0u;
    for (int k = 0; k < 32 * L - tb; k++) mw_dbl_mod<L>(r, m);
}

// Shift right by `bits` (< 32 L), with only compile-time limb indexes so
// the value stays in registers on the device.
template <int L>
FK_HD void mw_shr(uint32_t* a, int bits)
{
#pragma unroll
    for (int k = 0; k < L - 1; k++) {
        if (bits >= 32) {
#pragma unroll
            for (int i = 0; i < L - 1; i++) a[i] = a[i + 1];
            a[L - 1] = 0;
            bits -= 32;
        }
    }
    if (bits) {
#pragma unroll
        for (int i = 0; i < L - 1; i++)
            a[i] = (a[i] >> bits) | (a[i + 1] << (32 - bits));
        a[L - 1] >>= bits;
    }
}

// Strong probable-prime test of odd m > 37 (no factor <= 37) to the bases
// first..first+nbases-1 of the seven.  Returns true when m passes every
// base asked for.  huntlib/primes.py:31-47 (the loop over bases) and
// :61-80 (base 2 alone) are the two uses.  In the Montgomery domain
// (R = 2^(32 L)): one = R mod m, a R mod m by double-and-add of one, and
// for base 2 the "multiply by the base" step of the ladder is a doubling.
template <int L>
FK_HD bool sprp_bases(const uint32_t* m, int first, int nbases)
{
    const uint32_t minv = neg_inv32(m[0]);
    uint32_t one[L], mone[L], d[L];
    mont_one<L>(one, m);
    mw_sub<L>(mone, m, one);            // (m - 1) R mod m == m - (R mod m)
    // m - 1 = d * 2^s
#pragma unroll
    for (int i = 0; i < L; i++) d[i] = m[i];
    d[0] -= 1;                          // m odd: no borrow
    int s = 0;
    bool found = false;
#pragma unroll
    for (int i = 0; i < L; i++) {
        if (!found) {
            if (d[i] != 0) { s += __builtin_ctz(d[i]); found = true; }
            else s += 32;
        }
    }
    mw_shr<L>(d, s);
    bool one_limb = true;
#pragma unroll
    for (int i = 1; i < L; i++) one_limb = one_limb && (m[i] == 0);

    for (int b = first; b < first + nbases; b++) {
        // a %= m; a == 0 -> skip (a < 2^31, so only a one-limb m reduces it)
        uint32_t a = mr_base(b);
        if (one_limb) a = a % m[0];
        if (a == 0) continue;
        uint32_t am[L], x[L];
#pragma unroll
        for (int i = 0; i < L; i++) { am[i] = 0; x[i] = 0; }
        // a R mod m by double-and-add over the bits of a
        for (int k = 31 - __builtin_clz(a); k >= 0; k--) {
            mw_dbl_mod<L>(am, m);
            if ((a >> k) & 1) mw_add_mod<L>(am, one, m);
        }
        const bool two = (a == 2);
        bool started = false;
#pragma unroll
        for (int w = L - 1; w >= 0; w--) {
            const uint32_t word = d[w];
            for (int bit = 31; bit >= 0; bit--) {
                const bool set = (word >> bit) & 1;
                if (started) {
                    mont_sqr<L>(x, x, m, minv);
                    if (set) {
                        if (two) mw_dbl_mod<L>(x, m);
                        else mont_mul<L>(x, x, am, m, minv);
                    }
                } else if (set) {
                    started = true;
#pragma unroll
                    for (int i = 0; i < L; i++) x[i] = am[i];
                }
            }
        }
        if (mw_cmp<L>(x, one) == 0 || mw_cmp<L>(x, mone) == 0) continue;
        bool ok = false;
        for (int r = 1; r < s; r++) {
            mont_sqr<L>(x, x, m, minv);
            if (mw_cmp<L>(x, mone) == 0) { ok = true; break; }
        }
        if (!ok) return false;
    }
    return true;
}

// huntlib/primes.py:23-47 (bases 0..6) and :50-80 (base 0 alone) for a
// value of up to 8 limbs: the small-prime pre-check answers m == sp, then
// the strong test to bases first..first+nbases-1 runs at the narrowest
// limb count that holds m.  The seven-base verdict is the AND of the
// single-base verdicts, which is what lets the device test bases in
// parallel.
FK_HD bool is_sprp(const uint32_t* v8, int first, int nbases)
{
    int top = 0;
#pragma unroll
    for (int i = 1; i < 8; i++) if (v8[i] != 0) top = i;
    if (top == 0 && v8[0] < 2) return false;
#pragma unroll
    for (int i = 0; i < kNumSmall; i++) {
        const uint32_t sp = small_prime(i);
        if (mw_mod_small<8>(v8, sp) == 0) return top == 0 && v8[0] == sp;
    }
    if (top < 2) return sprp_bases<2>(v8, first, nbases);
    if (top < 4) return sprp_bases<4>(v8, first, nbases);
    if (top < 6) return sprp_bases<6>(v8, first, nbases);
    return sprp_bases<8>(v8, first, nbases);
}

// --------------------------------------------- small multiword helpers
FK_HD void mw8_zero(uint32_t* v) { for (int i = 0; i < 8; i++) v[i] = 0; }

// v8 = a4 * b (b < 2^32), widened to 8 limbs
FK_HD void mw_mul_small(uint32_t* r8, const uint32_t* a, int la, uint32_t b)
{
    uint64_t c = 0;
    for (int i = 0; i < 8; i++) {
        uint64_t s = (i < la ? (uint64_t)a[i] * b : 0) + c;
        r8[i] = (uint32_t)s;
        c = s >> 32;
    }
}

// r8 = a * b for 4-limb operands (full 256-bit product)
FK_HD void mw_mul44(uint32_t* r8, const uint32_t* a, const uint32_t* b)
{
    for (int i = 0; i < 8; i++) r8[i] = 0;
    for (int i = 0; i < 4; i++) {
        uint64_t c = 0;
        for (int j = 0; j < 4; j++) {
            uint64_t s = (uint64_t)a[j] * b[i] + r8[i + j] + c;
            r8[i + j] = (uint32_t)s;
            c = s >> 32;
        }
        r8[i + 4] = (uint32_t)c;
    }
}

// r8 += a8 (mod 2^256); returns carry
FK_HD uint32_t mw_add8(uint32_t* r8, const uint32_t* a8)
{
    return mw_add<8>(r8, r8, a8);
}

}  // namespace fk
