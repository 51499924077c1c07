// fk_plan.h -- host-side mathematics of the three families: prime lists,
// killed residue sets, the wheel and the per-prime LDS tables.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

namespace fk {

typedef unsigned __int128 u128;

std::vector<uint32_t> primes_below(uint32_t limit);

uint64_t powmod64(uint64_t a, uint64_t e, uint64_t m);
uint32_t inv_mod(uint32_t a, uint32_t q);          // a^-1 mod prime q

// W(n) = product of primes <= n + 1 (dickson-ladders/ladder_reference.py:60-66)
uint64_t dickson_wheel(int n);

// P(n) = product of the first n primes (primorial-ap/ap_reference.py:99-110)
u128 ap_difference(int n);

// The first kill prime of a family: primes below it never kill.
uint32_t first_kill_prime(int family, int n);

// Killed residues of the family's line value mod prime q, sorted:
//   Euler   {-(x^2 + x) mod q : x < n}           (euler_search.py:85-87)
//   Dickson {j : q | m (W j)^2 + 1, some m <= n}  (ladder_search.py:71-89)
//   AP      {-i P(n) mod q : i < n}               (ap_search.py:102-104)
std::vector<uint32_t> kill_set(int family, int n, uint32_t q);

// Appends prime q's periodic kill bitmap to `words` and its first word's
// index to `offs`: bit b is set iff b mod q is in `kills`, for
// b < q + tail (tail a multiple of 64), so that the 64-bit windows at
// x, x + 64, ... below q + tail - 64 lie inside it, for any x < q, in
// (q - 1 + tail) / 32 + 1 words (one more for a window's third word).
void append_kill_bitmap(const std::vector<uint32_t>& kills, uint32_t q,
                        uint32_t tail, std::vector<uint32_t>& offs,
                        std::vector<uint32_t>& words);

// Admissible residues mod prod(wheel) (CRT lifting, sorted) and that
// product.  Empty result with *ok = false if the product overflows.
std::vector<uint64_t> wheel_offsets(int family, int n,
                                    const std::vector<uint32_t>& wheel,
                                    uint64_t* modulus, bool* ok);

// Number of admissible residues mod prod(wheel), from the kill-set sizes
// (the residues themselves are not formed); *ok = false if the product of
// the wheel primes passes 2^63.
uint64_t wheel_count(int family, int n, const std::vector<uint32_t>& wheel,
                     bool* ok);

// The wheel factored at its last prime L: a residue base + j Mb (base an
// admissible residue of the wheel without L, whose period is Mb; j < L) is
// admissible iff (base + j Mb) mod L avoids L's kill set.  For each class
// c = base mod L the admissible j, ascending, are lifts[c * nj + t],
// t < nj, and nj = L - |kill set| is the same for every class.
std::vector<uint32_t> wheel_lifts(int family, int n, uint32_t L, uint64_t Mb,
                                  uint32_t* nj);

}  // namespace fk
