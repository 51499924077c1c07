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
    double c22[3];         // 2^22, 2^44, 2^66 mod q (22-bit limbs, q < 2^24)
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
    c = 1;
    for (int i = 0; i < 3; i++) {
        c = (c << 22) % q;
        p.c22[i] = (double)c;
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

// The same residue with one fix instead of two: the quotient estimate is
// taken 1/8 low.  For an integer s held exactly (s < 2^53) with
// s / q < 2^48 -- s < 2^49 and q >= 2, or s < 2^53 and q >= 32 --
// fma(s, qi, -1/8) differs from s / q - 1/8 by at most 2^-5 (qi's relative
// error 2^-53 times s / q) plus 2^-5 (its own rounding at magnitude below
// 2^48), so its floor is floor(s / q) or one less, and s - floor * q, below
// 2q and formed by one fma, is exact and lies in [0, 2q).
FK_HD uint32_t resid_d1(double s, double qd, double qi, uint32_t q)
{
    const double qt = floor(fma(s, qi, -0.125));
    const uint32_t u = (uint32_t)fma(-qt, qd, s);    // exact, in [0, 2q)
    const uint32_t dn = u - q;
    return u < dn ? u : dn;
}

// The folded residue min(s mod q, q - s mod q) directly, for the rows that
// hold folded kill sets.  The quotient qt is the nearest integer to s * qi;
// under resid_d1's condition (s exact, s / q < 2^48) s * qi is within 2^-4
// of s / q, so |s / q - qt| <= 1/2 + 2^-4 and t = s - qt * q (exact, one
// fma) has |t| < q.  t is then congruent to u = |t| or to -u, and both
// fold to min(u, q - u).
FK_HD uint32_t resid_fold_d(double s, double qd, double qi, uint32_t q)
{
    const double qt = rint(s * qi);
    const uint32_t u = (uint32_t)fabs(fma(-qt, qd, s));   // exact, < q
    const uint32_t d = q - u;
    return u < d ? u : d;
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

// ------------------------------------------ radix-2^29 Montgomery residues
// A residue modulo an odd m of bit length k is held in N = ceil((k + 2) / 29)
// limbs of 29 bits (uint32_t each, little-endian), so R = 2^(29 N) > 4 m;
// N <= 9 covers 256-bit values.  Every limb product is below 2^59 (2^58
// for two limbs, 2^59 against a doubled limb in the square), and a column
// holds at most N products and N - 1 reduction terms: below 2^62.2 with the
// carry in.  So a column is summed in 64-bit accumulators with one
// v_mad_u64_u32 per term and no carry handling, and the carry out is one
// shift.  Residues are kept lazily, below 2 m rather than m, inside the
// exponentiation; reduce_w makes them exact where they are compared.

constexpr int kLimbBits = 29;
constexpr uint32_t kLimbMask = (1u << kLimbBits) - 1;
constexpr int kMaxLimbs = 9;                 // 29 * 9 >= 256 + 2

template <int N>
struct ModW {
    uint32_t m[N];
    uint32_t minv;       // -m^-1 mod 2^29
    double inv_top;      // 1 / top_w<N>(m)
};

// The top min(N, 4) limbs of x as a double (the same scale for every x of
// N limbs).  m's top four limbs hold at least 85 of its bits, so ratios of
// these approximate ratios of the full values to about 2^-50.
template <int N>
FK_HD double top_w(const uint32_t* x)
{
    constexpr int T = N < 4 ? N : 4;
    double r = 0.0;
#pragma unroll
    for (int i = N - 1; i >= N - T; i--)
        r = r * (double)(1u << kLimbBits) + (double)x[i];
    return r;
}

// 29-bit limbs of the value in eight 32-bit words (bits past word 7 are 0)
template <int N>
FK_HD void to_w(uint32_t* r, const uint32_t* v8)
{
#pragma unroll
    for (int i = 0; i < N; i++) {
        const int bit = kLimbBits * i, w = bit >> 5, sh = bit & 31;
        const uint64_t lo = w < 8 ? v8[w] : 0u;
        const uint64_t hi = w + 1 < 8 ? v8[w + 1] : 0u;
        r[i] = (uint32_t)((lo | (hi << 32)) >> sh) & kLimbMask;
    }
}

// x = x - m when x >= m (x < 2m on entry, exact residue on exit)
template <int N>
FK_HD void reduce_w(uint32_t* x, const uint32_t* m)
{
    uint32_t d[N];
    int32_t borrow = 0;
#pragma unroll
    for (int i = 0; i < N; i++) {
        const int32_t t = (int32_t)x[i] - (int32_t)m[i] + borrow;
        d[i] = (uint32_t)t & kLimbMask;
        borrow = t >> kLimbBits;             // 0 or -1
    }
    const bool ge = borrow == 0;
#pragma unroll
    for (int i = 0; i < N; i++) x[i] = ge ? d[i] : x[i];
}

template <int N>
FK_HD bool eq_w(const uint32_t* a, const uint32_t* b)
{
    uint32_t o = 0;
#pragma unroll
    for (int i = 0; i < N; i++) o |= a[i] ^ b[i];
    return o == 0;
}

// r = x y R^-1 mod m, up to one m (r < 2m), for x, y < 2m (limbs < 2^29).
// Product scanning with the reduction interleaved column by column: u_i
// clears column i, and the top N columns are the result (R > 4m bounds it
// by (4 m^2 + R m) / R < 2 m).  Two accumulators per column, the products
// and the reduction terms, so the two multiply-add chains run side by side.
template <int N>
FK_HD void mont_mul_w(uint32_t* r, const uint32_t* x, const uint32_t* y,
                      const ModW<N>& M)
{
    uint32_t u[N], o[N];
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < N; i++) {
        uint64_t red = 0;
#pragma unroll
        for (int j = 0; j <= i; j++) acc += (uint64_t)x[j] * y[i - j];
#pragma unroll
        for (int j = 0; j < i; j++) red += (uint64_t)u[j] * M.m[i - j];
        acc += red;
        u[i] = ((uint32_t)acc * M.minv) & kLimbMask;
        acc += (uint64_t)u[i] * M.m[0];
        acc >>= kLimbBits;
    }
#pragma unroll
    for (int i = N; i < 2 * N - 1; i++) {
        uint64_t red = 0;
#pragma unroll
        for (int j = i - N + 1; j < N; j++) {
            acc += (uint64_t)x[j] * y[i - j];
            red += (uint64_t)u[j] * M.m[i - j];
        }
        acc += red;
        o[i - N] = (uint32_t)acc & kLimbMask;
        acc >>= kLimbBits;
    }
    o[N - 1] = (uint32_t)acc;
#pragma unroll
    for (int i = 0; i < N; i++) r[i] = o[i];
}

// r = x^2 R^-1 mod m, up to one m, for x < 2m: mont_mul_w with each cross
// product x_j x_k (j < k) formed once against 2 x_j.
template <int N>
FK_HD void mont_sqr_w(uint32_t* r, const uint32_t* x, const ModW<N>& M)
{
    uint32_t u[N], o[N], x2[N];
#pragma unroll
    for (int j = 0; j < N; j++) x2[j] = x[j] << 1;
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < 2 * N - 1; i++) {
        uint64_t red = 0;
#pragma unroll
        for (int j = (i < N ? 0 : i - N + 1); 2 * j < i; j++)
            acc += (uint64_t)x2[j] * x[i - j];
        if ((i & 1) == 0) acc += (uint64_t)x[i >> 1] * x[i >> 1];
        if (i < N) {
#pragma unroll
            for (int j = 0; j < i; j++) red += (uint64_t)u[j] * M.m[i - j];
            acc += red;
            u[i] = ((uint32_t)acc * M.minv) & kLimbMask;
            acc += (uint64_t)u[i] * M.m[0];
        } else {
#pragma unroll
            for (int j = i - N + 1; j < N; j++)
                red += (uint64_t)u[j] * M.m[i - j];
            acc += red;
            o[i - N] = (uint32_t)acc & kLimbMask;
        }
        acc >>= kLimbBits;
    }
    o[N - 1] = (uint32_t)acc;
#pragma unroll
    for (int i = 0; i < N; i++) r[i] = o[i];
}

// r = a x mod m, up to one m (r < 2m), for x < 2m and a < 2^31.  The
// quotient q of a x by m is below 2^32; the double estimate, scaled down by
// 1 - 2^-40, is within 2^-8 + 2^-17 below a x / m, so its floor is q or
// q - 1 and a x - floor * m lies in [0, 2m).  One signed pass per limb
// (every term below 2^61 in magnitude).
template <int N>
FK_HD void mul_small_w(uint32_t* r, const uint32_t* x, uint32_t a,
                       const ModW<N>& M)
{
    const double qf = (double)a * top_w<N>(x) * M.inv_top * (1.0 - 0x1p-40);
    const uint32_t q = (uint32_t)qf;
    int64_t c = 0;
#pragma unroll
    for (int j = 0; j < N; j++) {
        const int64_t v = (int64_t)((uint64_t)a * x[j])
                        - (int64_t)((uint64_t)q * M.m[j]) + c;
        r[j] = (uint32_t)v & kLimbMask;
        c = v >> kLimbBits;
    }
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

// Strong probable-prime test of odd m > 37 with no factor <= 37, of bit
// length k <= 29 N - 2, to the bases first..first+nbases-1 of the seven.
// huntlib/primes.py:31-47 (the loop over bases) and :61-80 (base 2 alone)
// are the two uses.  In the Montgomery domain (R = 2^(29 N)) the ladder
// starts from one = R mod m and, per bit of d from the top, squares and
// then multiplies by the bit's factor (the base or 1) as a plain scalar:
// one instruction stream for every base, with no table and no branch.
template <int N>
FK_HD bool sprp_w(const uint32_t* v8, int first, int nbases)
{
    ModW<N> M;
    to_w<N>(M.m, v8);
    M.minv = neg_inv32(v8[0]) & kLimbMask;
    M.inv_top = 1.0 / top_w<N>(M.m);
    // one = 2^(29 N) mod m: 2^(k-1) < m, times 2^(29 N - k + 1) in factors
    // of at most 2^28 (one factor unless N was raised to its minimum of 2)
    int k = 0;
#pragma unroll
    for (int i = 0; i < 8; i++)
        if (v8[i]) k = 32 * i + 32 - __builtin_clz(v8[i]);
    const int tb = k - 1;
    uint32_t one[N], mone[N];
#pragma unroll
    for (int i = 0; i < N; i++)
        one[i] = (i == tb / kLimbBits) ? 1u << (tb % kLimbBits) : 0u;
    for (int sh = kLimbBits * N - tb; sh > 0; sh -= 28)
        mul_small_w<N>(one, one, 1u << (sh < 28 ? sh : 28), M);
    reduce_w<N>(one, M.m);
    int32_t borrow = 0;                   // mone = m - one = (m - 1) R mod m
#pragma unroll
    for (int i = 0; i < N; i++) {
        const int32_t t = (int32_t)M.m[i] - (int32_t)one[i] + borrow;
        mone[i] = (uint32_t)t & kLimbMask;
        borrow = t >> kLimbBits;
    }
    // m - 1 = d 2^s
    constexpr int W = (kLimbBits * N + 31) / 32 < 8 ? (kLimbBits * N + 31) / 32 : 8;
    uint32_t d[W];
#pragma unroll
    for (int i = 0; i < W; i++) d[i] = v8[i];
    d[0] -= 1;                            // m odd: no borrow
    int s = 0;
    bool found = false;
#pragma unroll
    for (int i = 0; i < W; i++) {
        if (!found) {
            if (d[i] != 0) { s += __builtin_ctz(d[i]); found = true; }
            else s += 32;
        }
    }
    mw_shr<W>(d, s);
    int dtop = 0;
#pragma unroll
    for (int i = 0; i < W; i++) if (d[i]) dtop = i;

    const bool one_word = k <= 32;
    for (int b = first; b < first + nbases; b++) {
        // a mod m == 0 skips the base (a < 2^31, so only m < 2^32 can
        // divide it); otherwise the unreduced a gives the same powers
        const uint32_t a = mr_base(b);
        if (one_word && a % v8[0] == 0) continue;
        uint32_t x[N];
#pragma unroll
        for (int i = 0; i < N; i++) x[i] = one[i];
#pragma unroll
        for (int w = W - 1; w >= 0; w--) {
            if (w > dtop) continue;
            const uint32_t word = d[w];
            const int nb = w == dtop ? 32 - __builtin_clz(word) : 32;
            for (int bit = nb - 1; bit >= 0; bit--) {
                mont_sqr_w<N>(x, x, M);
                mul_small_w<N>(x, x, ((word >> bit) & 1u) ? a : 1u, M);
            }
        }
        reduce_w<N>(x, M.m);
        if (eq_w<N>(x, one) || eq_w<N>(x, mone)) continue;
        bool ok = false;
        for (int r = 1; r < s; r++) {
            mont_sqr_w<N>(x, x, M);
            uint32_t e[N];
#pragma unroll
            for (int i = 0; i < N; i++) e[i] = x[i];
            reduce_w<N>(e, M.m);
            if (eq_w<N>(e, mone)) { ok = true; break; }
        }
        if (!ok) return false;
    }
    return true;
}

// huntlib/primes.py:23-47 (bases 0..6) and :50-80 (base 0 alone) for a
// value of up to 8 limbs: the small-prime pre-check answers m == sp, then
// the strong test to bases first..first+nbases-1 runs at the narrowest
// radix-2^29 limb count that leaves R > 4 m.  The seven-base verdict is the
// AND of the single-base verdicts, which is what lets the device test bases
// in parallel.
FK_HD bool is_sprp(const uint32_t* v8, int first, int nbases)
{
    int top = 0;
#pragma unroll
    for (int i = 1; i < 8; i++) if (v8[i] != 0) top = i;
    if (top == 0 && v8[0] < 2) return false;
    // a factor among the twelve small primes: then m is prime only if it is
    // that prime.  The residues come from two 8-limb Horner passes, modulo
    // 3*5*...*29 (below 2^32) and 31*37, and 2 from the low bit.
    const uint32_t r1 = mw_mod_small<8>(v8, 3234846615u);
    const uint32_t r2 = mw_mod_small<8>(v8, 1147u);
    bool factor = (v8[0] & 1u) == 0;
#pragma unroll
    for (int i = 1; i < 10; i++) factor = factor || r1 % small_prime(i) == 0;
    factor = factor || r2 % 31u == 0 || r2 % 37u == 0;
    if (factor) {
        bool is_small = false;
#pragma unroll
        for (int i = 0; i < kNumSmall; i++)
            is_small = is_small || (top == 0 && v8[0] == small_prime(i));
        return is_small;
    }
    int k = 0;                            // bit length, no dynamic index
#pragma unroll
    for (int i = 0; i < 8; i++)
        if (v8[i]) k = 32 * i + 32 - __builtin_clz(v8[i]);
    const int n = (k + 2 + kLimbBits - 1) / kLimbBits;
    switch (n) {
    case 0: case 1: case 2: return sprp_w<2>(v8, first, nbases);
    case 3: return sprp_w<3>(v8, first, nbases);
    case 4: return sprp_w<4>(v8, first, nbases);
    case 5: return sprp_w<5>(v8, first, nbases);
    case 6: return sprp_w<6>(v8, first, nbases);
    case 7: return sprp_w<7>(v8, first, nbases);
    case 8: return sprp_w<8>(v8, first, nbases);
    default: return sprp_w<9>(v8, first, nbases);
    }
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
