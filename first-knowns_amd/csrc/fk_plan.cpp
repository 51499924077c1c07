// fk_plan.cpp -- the families' mathematics, computed on the host once per
// engine.  Written from the definitions (cited per function); shares no
// code with the CPU oracle under oracle/.
#include "fk_plan.h"

#include <algorithm>

namespace fk {

void append_kill_bitmap(const std::vector<uint32_t>& kills, uint32_t q,
                        uint32_t tail, std::vector<uint32_t>& offs,
                        std::vector<uint32_t>& words)
{
    offs.push_back((uint32_t)words.size());
    const size_t base = words.size();
    words.resize(base + ((uint64_t)q - 1 + tail) / 32 + 1, 0u);
    for (uint32_t r : kills)
        for (uint64_t b = r; b < (uint64_t)q + tail; b += q)
            words[base + (b >> 5)] |= 1u << (b & 31);
}

std::vector<uint32_t> primes_below(uint32_t limit)
{
    std::vector<uint32_t> out;
    if (limit < 3) return out;
    // odd-only sieve of Eratosthenes
    const uint64_t half = limit / 2;
    std::vector<uint8_t> comp(half + 1, 0);
    out.push_back(2);
    for (uint64_t i = 1; i < half; i++) {
        if (comp[i]) continue;
        const uint64_t p = 2 * i + 1;
        if (p >= limit) break;
        out.push_back((uint32_t)p);
        for (uint64_t j = (p * p) / 2; j < half; j += p) comp[j] = 1;
    }
    return out;
}

uint64_t powmod64(uint64_t a, uint64_t e, uint64_t m)
{
    u128 r = 1 % m, b = a % m;
    while (e) {
        if (e & 1) r = r * b % m;
        b = b * b % m;
        e >>= 1;
    }
    return (uint64_t)r;
}

uint32_t inv_mod(uint32_t a, uint32_t q)
{
    return (uint32_t)powmod64(a % q, q - 2, q);
}

uint64_t dickson_wheel(int n)
{
    uint64_t w = 1;
    for (uint32_t q : primes_below((uint32_t)n + 2)) w *= q;
    return w;
}

u128 ap_difference(int n)
{
    u128 d = 1;
    int c = 0;
    for (uint32_t q : primes_below(200)) {
        if (c == n) break;
        d *= q;
        c++;
    }
    return d;
}

uint32_t first_kill_prime(int family, int n)
{
    // Dickson: primes <= n+1 divide W and never divide m k^2 + 1 for k on
    // the wheel (ladder_reference.py WHEEL argument); the others start at 2.
    return family == 2 ? (uint32_t)n + 2 : 2;
}

// Square root of a quadratic residue a mod an odd prime q (Tonelli-Shanks).
static uint64_t sqrt_mod_prime(uint64_t a, uint64_t q)
{
    if (a == 0) return 0;
    uint64_t Q = q - 1, S = 0;
    while ((Q & 1) == 0) { Q >>= 1; S++; }
    if (S == 1) return powmod64(a, (q + 1) / 4, q);
    uint64_t z = 2;
    while (powmod64(z, (q - 1) / 2, q) != q - 1) z++;
    uint64_t M = S, c = powmod64(z, Q, q), t = powmod64(a, Q, q);
    uint64_t R = powmod64(a, (Q + 1) / 2, q);
    while (t != 1) {
        uint64_t i = 0, tt = t;
        while (tt != 1) { tt = (uint64_t)((u128)tt * tt % q); i++; }
        uint64_t b = c;
        for (uint64_t j = 0; j + i + 1 < M; j++) b = (uint64_t)((u128)b * b % q);
        M = i;
        c = (uint64_t)((u128)b * b % q);
        t = (uint64_t)((u128)t * c % q);
        R = (uint64_t)((u128)R * b % q);
    }
    return R;
}

std::vector<uint32_t> kill_set(int family, int n, uint32_t q)
{
    std::vector<uint32_t> out;
    if (family == 1) {
        for (int x = 0; x < n; x++) {
            const uint64_t v = ((uint64_t)x * x + x) % q;
            out.push_back((uint32_t)((q - v) % q));
        }
    } else if (family == 2) {
        // j is killed when (W j)^2 == -1/m (mod q): two roots per m whose
        // -1/m is a square, none otherwise.
        const uint64_t W = dickson_wheel(n) % q;
        const uint32_t winv = inv_mod((uint32_t)W, q);
        for (int m = 1; m <= n; m++) {
            const uint64_t rhs = (q - inv_mod((uint32_t)(m % q), q)) % q;
            if (q == 2) {
                out.push_back((uint32_t)((u128)rhs * winv % q));
                continue;
            }
            if (powmod64(rhs, (q - 1) / 2, q) != 1) continue;
            const uint64_t r = sqrt_mod_prime(rhs, q);
            out.push_back((uint32_t)((u128)r * winv % q));
            out.push_back((uint32_t)((u128)((q - r) % q) * winv % q));
        }
    } else {
        const u128 d = ap_difference(n);
        const uint64_t dq = (uint64_t)(d % q);
        for (int i = 0; i < n; i++)
            out.push_back((uint32_t)((q - (u128)i * dq % q) % q));
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    return out;
}

std::vector<uint64_t> wheel_offsets(int family, int n,
                                    const std::vector<uint32_t>& wheel,
                                    uint64_t* modulus, bool* ok)
{
    std::vector<uint64_t> offs(1, 0);
    u128 m = 1;
    *ok = true;
    for (uint32_t q : wheel) {
        std::vector<uint32_t> ks = kill_set(family, n, q);
        std::vector<uint8_t> dead(q, 0);
        for (uint32_t r : ks) dead[r] = 1;
        std::vector<uint64_t> nx;
        nx.reserve(offs.size() * (q - ks.size()));
        const uint64_t mq = (uint64_t)(m % q);
        for (uint32_t j = 0; j < q; j++) {
            const uint64_t shift = (uint64_t)j * (uint64_t)m;
            const uint64_t sq = (uint64_t)(((u128)j * mq) % q);
            for (uint64_t o : offs) {
                uint64_t r = o % q + sq;
                if (r >= q) r -= q;
                if (!dead[r]) nx.push_back(o + shift);
            }
        }
        offs.swap(nx);
        m *= q;
        if (m >> 63) { *ok = false; return {}; }
    }
    std::sort(offs.begin(), offs.end());
    *modulus = (uint64_t)m;
    return offs;
}

uint64_t wheel_count(int family, int n, const std::vector<uint32_t>& wheel,
                     bool* ok)
{
    u128 m = 1, c = 1;
    *ok = true;
    for (uint32_t q : wheel) {
        c *= q - kill_set(family, n, q).size();
        m *= q;
        if (m >> 63) {
            *ok = false;
            return 0;
        }
    }
    return (uint64_t)c;
}

std::vector<uint32_t> wheel_lifts(int family, int n, uint32_t L, uint64_t Mb,
                                  uint32_t* nj)
{
    std::vector<uint32_t> ks = kill_set(family, n, L);
    std::vector<uint8_t> dead(L, 0);
    for (uint32_t r : ks) dead[r] = 1;
    *nj = L - (uint32_t)ks.size();
    const uint64_t mbl = Mb % L;
    std::vector<uint32_t> lifts;
    lifts.reserve((size_t)L * *nj);
    for (uint32_t c = 0; c < L; c++)
        for (uint32_t j = 0; j < L; j++)
            if (!dead[(c + (uint64_t)j * mbl) % L]) lifts.push_back(j);
    return lifts;
}

}  // namespace fk
