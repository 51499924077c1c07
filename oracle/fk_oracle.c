/*
 * fk_oracle.c -- TEST INFRASTRUCTURE, not product code.
 *
 * A plain-C restatement of the reference's CPU sieve engines, used only as
 * the checker: by tests/, by __graft_entry__.smoke() and by bench.py's CPU
 * baseline leg.  Nothing in the shipped library links it, calls it or falls
 * back to it.  It is deliberately written the way the reference CPU engines
 * are written (plain `%`, byte arrays, strided kills, per-candidate LUT
 * tests), and shares no code and no construction with the HIP engine:
 *
 *   Euler    euler-prime-runs/euler_search.py:85-87   forbidden()
 *            euler-prime-runs/euler_search.py:215-249 build_wheel() (CRT lift)
 *            euler-prime-runs/euler_search.py:252-255 stage_primes()
 *            euler-prime-runs/euler_search.py:306-348 survivors_pre_mr()
 *   Dickson  dickson-ladders/ladder_search.py:71-89   _forbidden_j()
 *            dickson-ladders/ladder_search.py:92-112  survivors_j()
 *   AP       primorial-ap/ap_search.py:95-106          __init__ (forb tables)
 *            primorial-ap/ap_search.py:109-141         survivors()
 *
 * Values that can pass 2^64 (Euler p to 1e24, AP p to 1e26, Dickson j to the
 * folded reach) are carried as unsigned __int128; every function returns the
 * total survivor count (which may exceed `cap`; only `cap` are written) or a
 * negative error code.  Output is ascending.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

typedef unsigned __int128 u128;

#define ORC_EARG   (-1)
#define ORC_ENOMEM (-2)

static u128 mk(uint64_t lo, uint64_t hi) { return ((u128)hi << 64) | lo; }

/* ------------------------------------------------------------------ primes */

static uint32_t *primes_below(uint32_t limit, int64_t *count)
{
    uint8_t *comp = calloc(limit + 1, 1);
    int64_t c = 0;
    uint32_t *out;
    if (!comp) return NULL;
    for (uint64_t i = 2; i < limit; i++) {
        if (comp[i]) continue;
        c++;
        for (uint64_t j = i * i; j < limit; j += i) comp[j] = 1;
    }
    out = malloc((c ? c : 1) * sizeof(uint32_t));
    if (!out) { free(comp); return NULL; }
    c = 0;
    for (uint64_t i = 2; i < limit; i++)
        if (!comp[i]) out[c++] = (uint32_t)i;
    free(comp);
    *count = c;
    return out;
}

int64_t fko_primes(uint32_t limit, uint32_t *out, int64_t cap)
{
    int64_t c = 0;
    uint32_t *p = primes_below(limit, &c);
    if (!p) return ORC_ENOMEM;
    for (int64_t i = 0; i < c && i < cap; i++) out[i] = p[i];
    free(p);
    return c;
}

static uint64_t powmod64(uint64_t a, uint64_t e, uint64_t m)
{
    u128 r = 1, b = a % m;
    while (e) {
        if (e & 1) r = r * b % m;
        b = b * b % m;
        e >>= 1;
    }
    return (uint64_t)(r % m);
}

static int cmp_u128(const void *a, const void *b)
{
    u128 x = *(const u128 *)a, y = *(const u128 *)b;
    return x < y ? -1 : x > y;
}

/* Per-thread growable survivor lists, merged and sorted at the end. */
typedef struct { u128 *v; int64_t n, cap; } vec128;

static int vpush(vec128 *s, u128 x)
{
    if (s->n == s->cap) {
        int64_t nc = s->cap ? 2 * s->cap : 1024;
        u128 *nv = realloc(s->v, nc * sizeof(u128));
        if (!nv) return -1;
        s->v = nv;
        s->cap = nc;
    }
    s->v[s->n++] = x;
    return 0;
}

static int64_t emit_sorted(vec128 *parts, int nparts, uint64_t *out,
                           int64_t cap, int pairs)
{
    int64_t total = 0, k = 0;
    u128 *all;
    for (int i = 0; i < nparts; i++) total += parts[i].n;
    all = malloc((total ? total : 1) * sizeof(u128));
    if (!all) return ORC_ENOMEM;
    for (int i = 0; i < nparts; i++) {
        memcpy(all + k, parts[i].v, parts[i].n * sizeof(u128));
        k += parts[i].n;
        free(parts[i].v);
    }
    qsort(all, total, sizeof(u128), cmp_u128);
    for (int64_t i = 0; i < total && i < cap; i++) {
        if (pairs) {
            out[2 * i] = (uint64_t)all[i];
            out[2 * i + 1] = (uint64_t)(all[i] >> 64);
        } else {
            out[i] = (uint64_t)all[i];
        }
    }
    free(all);
    return total;
}

/* ------------------------------------------------------------------- Euler */

/* euler_search.py:85-87: residues r with q | x^2+x+r for some x < n. */
static void euler_forbidden(uint32_t q, int n, uint8_t *lut)
{
    memset(lut, 0, q);
    for (int x = 0; x < n; x++) {
        uint64_t v = ((uint64_t)x * x + x) % q;
        lut[(q - v) % q] = 1;
    }
}

/* euler_search.py:215-249: admissible offsets mod prod(wheel), CRT lifting. */
static uint64_t *euler_wheel(int n, const uint32_t *wheel, int nwheel,
                             int64_t *count, uint64_t *modulus)
{
    uint64_t *offs = malloc(sizeof(uint64_t)), m = 1;
    int64_t c = 1;
    if (!offs) return NULL;
    offs[0] = 0;
    for (int w = 0; w < nwheel; w++) {
        uint32_t q = wheel[w];
        uint8_t *lut = malloc(q);
        uint64_t *nx = malloc((size_t)c * q * sizeof(uint64_t));
        int64_t k = 0;
        if (!lut || !nx) { free(lut); free(nx); free(offs); return NULL; }
        euler_forbidden(q, n, lut);
        for (uint32_t j = 0; j < q; j++)
            for (int64_t i = 0; i < c; i++) {
                uint64_t lift = offs[i] + (uint64_t)j * m;
                if (!lut[lift % q]) nx[k++] = lift;
            }
        free(lut);
        free(offs);
        offs = nx;
        c = k;
        m *= q;
    }
    /* sorted, as build_wheel returns it */
    {
        u128 *tmp = malloc((c ? c : 1) * sizeof(u128));
        if (!tmp) { free(offs); return NULL; }
        for (int64_t i = 0; i < c; i++) tmp[i] = offs[i];
        qsort(tmp, c, sizeof(u128), cmp_u128);
        for (int64_t i = 0; i < c; i++) offs[i] = (uint64_t)tmp[i];
        free(tmp);
    }
    *count = c;
    *modulus = m;
    return offs;
}

int64_t fko_euler_wheel(int n, const uint32_t *wheel, int nwheel,
                        uint64_t *out, int64_t cap, uint64_t *modulus)
{
    int64_t c = 0;
    uint64_t *offs = euler_wheel(n, wheel, nwheel, &c, modulus);
    if (!offs) return ORC_ENOMEM;
    for (int64_t i = 0; i < c && i < cap; i++) out[i] = offs[i];
    free(offs);
    return c;
}

/*
 * euler_search.py:306-348.  Candidates (k, off) with p = k*M + off for the
 * wheel's admissible offsets; stage 1 (primes in (wheel max, q1)) asks the
 * forbidden-residue LUT, stage 2 (primes in [q1, q2)) asks exact
 * divisibility of the n values.  Survivors are the p in [lo, hi).
 */
int64_t fko_euler_survivors(int n, const uint32_t *wheel, int nwheel,
                            uint32_t q1, uint32_t q2,
                            uint64_t lo_lo, uint64_t lo_hi,
                            uint64_t hi_lo, uint64_t hi_hi,
                            uint64_t *out, int64_t cap, int nthreads)
{
    u128 lo = mk(lo_lo, lo_hi), hi = mk(hi_lo, hi_hi);
    int64_t noffs = 0, nall = 0, ns1 = 0, ns2 = 0;
    uint64_t M = 0, *offs;
    uint32_t *all, *s1, *s2, after;
    uint8_t **lut;
    uint32_t *xx;
    vec128 *parts;
    u128 k0, k1;
    int fail = 0;
    if (n < 1 || nwheel < 1 || hi < lo || q1 > q2) return ORC_EARG;
    offs = euler_wheel(n, wheel, nwheel, &noffs, &M);
    if (!offs) return ORC_ENOMEM;
    all = primes_below(q2, &nall);
    if (!all) { free(offs); return ORC_ENOMEM; }
    after = wheel[nwheel - 1];
    s1 = malloc((nall + 1) * sizeof(uint32_t));
    s2 = malloc((nall + 1) * sizeof(uint32_t));
    lut = calloc(nall + 1, sizeof(uint8_t *));
    xx = malloc(n * sizeof(uint32_t));
    for (int64_t i = 0; i < nall; i++) {
        if (all[i] <= after) continue;
        if (all[i] < q1) s1[ns1++] = all[i];
        else s2[ns2++] = all[i];
    }
    for (int64_t i = 0; i < ns1; i++) {
        lut[i] = malloc(s1[i]);
        euler_forbidden(s1[i], n, lut[i]);
    }
    for (int x = 0; x < n; x++) xx[x] = (uint32_t)((uint64_t)x * x + x);

    k0 = lo / M;
    k1 = hi / M + 1;
    if (nthreads < 1) nthreads = omp_get_max_threads();
    parts = calloc(nthreads, sizeof(vec128));
#pragma omp parallel num_threads(nthreads)
    {
        int tid = omp_get_thread_num();
        uint32_t *kq1 = malloc((ns1 + 1) * sizeof(uint32_t));
        uint32_t *kq2 = malloc((ns2 + 1) * sizeof(uint32_t));
        int64_t nk = (int64_t)(k1 - k0);
#pragma omp for schedule(dynamic, 1)
        for (int64_t dk = 0; dk < nk; dk++) {
            u128 k = k0 + (u128)dk;
            /* (k mod q) * (M mod q) mod q, once per period */
            for (int64_t i = 0; i < ns1; i++)
                kq1[i] = (uint32_t)((uint64_t)(k % s1[i]) * (M % s1[i]) % s1[i]);
            for (int64_t i = 0; i < ns2; i++)
                kq2[i] = (uint32_t)((uint64_t)(k % s2[i]) * (M % s2[i]) % s2[i]);
            for (int64_t o = 0; o < noffs; o++) {
                u128 p = k * M + offs[o];
                int dead = 0;
                if (p < lo || p >= hi) continue;
                for (int64_t i = 0; i < ns1 && !dead; i++) {
                    uint32_t q = s1[i];
                    uint32_t r = (uint32_t)((kq1[i] + offs[o] % q) % q);
                    dead = lut[i][r];
                }
                for (int64_t i = 0; i < ns2 && !dead; i++) {
                    uint32_t q = s2[i];
                    uint64_t r = (kq2[i] + offs[o] % q) % q;
                    for (int x = 0; x < n; x++)
                        if ((r + xx[x]) % q == 0) { dead = 1; break; }
                }
                if (!dead && vpush(&parts[tid], p)) fail = 1;
            }
        }
        free(kq1);
        free(kq2);
    }
    for (int64_t i = 0; i < ns1; i++) free(lut[i]);
    free(lut); free(xx); free(s1); free(s2); free(all); free(offs);
    if (fail) {
        for (int i = 0; i < nthreads; i++) free(parts[i].v);
        free(parts);
        return ORC_ENOMEM;
    }
    {
        int64_t r = emit_sorted(parts, nthreads, out, cap, 1);
        free(parts);
        return r;
    }
}

/* ----------------------------------------------------------------- Dickson */

static int legendre(uint64_t a, uint64_t q)
{
    uint64_t r = powmod64(a, (q - 1) / 2, q);
    return r == 1 ? 1 : (r == 0 ? 0 : -1);
}

/* Tonelli-Shanks square root of a quadratic residue a mod odd prime q. */
static uint64_t sqrt_mod(uint64_t a, uint64_t q)
{
    uint64_t Q = q - 1, S = 0, z = 2, M, c, t, R;
    if (a == 0) return 0;
    while ((Q & 1) == 0) { Q >>= 1; S++; }
    if (S == 1) return powmod64(a, (q + 1) / 4, q);
    while (legendre(z, q) != -1) z++;
    M = S;
    c = powmod64(z, Q, q);
    t = powmod64(a, Q, q);
    R = powmod64(a, (Q + 1) / 2, q);
    while (t != 1) {
        uint64_t i = 0, tt = t, b;
        while (tt != 1) { tt = (uint64_t)((u128)tt * tt % q); i++; }
        b = c;
        for (uint64_t j = 0; j + i + 1 < M; j++) b = (uint64_t)((u128)b * b % q);
        M = i;
        c = (uint64_t)((u128)b * b % q);
        t = (uint64_t)((u128)t * c % q);
        R = (uint64_t)((u128)R * b % q);
    }
    return R;
}

static int cmp_u32(const void *a, const void *b)
{
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : x > y;
}

/*
 * ladder_search.py:71-89: killed j residues mod q = roots of
 * k^2 == -1/m (mod q) over m = 1..n, times W^-1.  Writes a sorted,
 * de-duplicated list into out (room for 2n) and returns its length.
 */
static int dickson_forbidden(uint32_t q, int n, uint64_t W, uint32_t *out)
{
    uint64_t winv = powmod64(W % q, q - 2, q);
    int c = 0;
    for (int m = 1; m <= n; m++) {
        uint64_t rhs = (q - powmod64((uint64_t)m, q - 2, q)) % q;
        uint64_t r;
        if (legendre(rhs, q) != 1) continue;
        r = sqrt_mod(rhs, q);
        out[c++] = (uint32_t)((u128)r * winv % q);
        out[c++] = (uint32_t)((u128)((q - r) % q) * winv % q);
    }
    qsort(out, c, sizeof(uint32_t), cmp_u32);
    {
        int k = 0;
        for (int i = 0; i < c; i++)
            if (k == 0 || out[k - 1] != out[i]) out[k++] = out[i];
        c = k;
    }
    return c;
}

int fko_dickson_forbidden(uint32_t q, int n, uint64_t W, uint32_t *out)
{
    if (q < 3 || (W % q) == 0) return ORC_EARG;
    return dickson_forbidden(q, n, W, out);
}

static uint64_t wheel_modulus(int n)
{
    uint64_t w = 1;
    for (int q = 2; q <= n + 1; q++) {
        int pr = 1;
        for (int d = 2; d * d <= q; d++) if (q % d == 0) { pr = 0; break; }
        if (pr) w *= (uint64_t)q;
    }
    return w;
}

/*
 * ladder_search.py:56-69 (__init__: primes and the root table) as a reusable
 * context, so a timed sweep excludes table construction exactly as the
 * reference's CpuEngine builds its tables once in __init__.
 */
typedef struct {
    int n;
    int64_t np;
    uint32_t *pr, *roots, *rcnt;
} dickson_ctx;

void *fko_dickson_ctx_create(int n, uint32_t q2)
{
    uint64_t W = wheel_modulus(n);
    int64_t nall = 0;
    uint32_t *all;
    dickson_ctx *c;
    if (n < 1) return NULL;
    all = primes_below(q2, &nall);
    if (!all) return NULL;
    c = calloc(1, sizeof *c);
    c->n = n;
    c->pr = malloc((nall + 1) * sizeof(uint32_t));
    for (int64_t i = 0; i < nall; i++)
        if (all[i] >= (uint32_t)(n + 2)) c->pr[c->np++] = all[i];
    free(all);
    c->roots = malloc((c->np + 1) * 2 * n * sizeof(uint32_t));
    c->rcnt = malloc((c->np + 1) * sizeof(uint32_t));
    for (int64_t i = 0; i < c->np; i++)
        c->rcnt[i] = dickson_forbidden(c->pr[i], n, W, c->roots + i * 2 * n);
    return c;
}

void fko_dickson_ctx_free(void *ctx)
{
    dickson_ctx *c = ctx;
    if (!c) return;
    free(c->pr); free(c->roots); free(c->rcnt); free(c);
}

/*
 * ladder_search.py:92-112: segmented byte sieve over j, one block at a time,
 * alive[first::q] = False for every killed residue of every prime in
 * [n+2, q2).  Blocks run in parallel; output is the ascending j.
 */
int64_t fko_dickson_ctx_survivors(void *ctx, uint64_t jlo_lo, uint64_t jlo_hi,
                                  uint64_t jhi_lo, uint64_t jhi_hi,
                                  uint64_t *out, int64_t cap, int nthreads,
                                  uint32_t block)
{
    dickson_ctx *c = ctx;
    u128 jlo = mk(jlo_lo, jlo_hi), jhi = mk(jhi_lo, jhi_hi);
    int64_t nblocks;
    vec128 *parts;
    int fail = 0, n;
    if (!c || jhi < jlo || block == 0) return ORC_EARG;
    n = c->n;
    nblocks = (int64_t)((jhi - jlo + block - 1) / block);
    parts = calloc(nblocks ? nblocks : 1, sizeof(vec128));
    if (nthreads < 1) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
    {
        uint8_t *alive = malloc(block);
#pragma omp for schedule(dynamic, 1)
        for (int64_t b = 0; b < nblocks; b++) {
            u128 j0 = jlo + (u128)b * block;
            uint64_t size = (uint64_t)((jhi - j0) < block ? (jhi - j0) : block);
            memset(alive, 1, size);
            for (int64_t i = 0; i < c->np; i++) {
                uint32_t q = c->pr[i];
                uint32_t start = (uint32_t)(j0 % q);
                const uint32_t *t = c->roots + i * 2 * n;
                for (uint32_t r = 0; r < c->rcnt[i]; r++) {
                    uint64_t first = (t[r] + q - start) % q;
                    for (uint64_t x = first; x < size; x += q) alive[x] = 0;
                }
            }
            for (uint64_t x = 0; x < size; x++)
                if (alive[x] && vpush(&parts[b], j0 + x)) fail = 1;
        }
        free(alive);
    }
    if (fail) {
        for (int64_t i = 0; i < nblocks; i++) free(parts[i].v);
        free(parts);
        return ORC_ENOMEM;
    }
    {
        int64_t r = emit_sorted(parts, (int)nblocks, out, cap, 1);
        free(parts);
        return r;
    }
}

int64_t fko_dickson_survivors(int n, uint32_t q2,
                              uint64_t jlo_lo, uint64_t jlo_hi,
                              uint64_t jhi_lo, uint64_t jhi_hi,
                              uint64_t *out, int64_t cap, int nthreads,
                              uint32_t block)
{
    void *c = fko_dickson_ctx_create(n, q2);
    int64_t r;
    if (!c) return ORC_EARG;
    r = fko_dickson_ctx_survivors(c, jlo_lo, jlo_hi, jhi_lo, jhi_hi, out,
                                  cap, nthreads, block);
    fko_dickson_ctx_free(c);
    return r;
}

/* ---------------------------------------------------------------------- AP */

/*
 * ap_search.py:95-106 and :109-141.  Primes 7..q2 with killed residues
 * {(-j*d) mod q : j < n} (which is {0} when q | d), a flat byte array per
 * block with the mod-30 wheel tiled in from the block's phase, and strided
 * kills.  base is a 128-bit value, offsets are relative to it.
 */
int64_t fko_ap_survivors(int n, uint32_t q2, uint64_t base_lo,
                         uint64_t base_hi, uint64_t span, uint64_t *out,
                         int64_t cap, int nthreads, uint32_t block)
{
    static const uint32_t first_primes[26] = {
        2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61,
        67, 71, 73, 79, 83, 89, 97, 101};
    static const uint8_t wheel30[30] = {
        0,1,0,0,0,0,0,1,0,0,0,1,0,1,0,0,0,1,0,1,0,0,0,1,0,0,0,0,0,1};
    u128 base = mk(base_lo, base_hi);
    int64_t nall = 0, np = 0, nblocks;
    uint32_t *all, *pr, *forb, *fcnt;
    vec128 *parts;
    int fail = 0;
    if (n < 1 || n > 26 || block == 0) return ORC_EARG;
    all = primes_below(q2, &nall);
    if (!all) return ORC_ENOMEM;
    pr = malloc((nall + 1) * sizeof(uint32_t));
    for (int64_t i = 0; i < nall; i++)
        if (all[i] >= 7) pr[np++] = all[i];
    forb = malloc((np + 1) * n * sizeof(uint32_t));
    fcnt = malloc((np + 1) * sizeof(uint32_t));
    for (int64_t i = 0; i < np; i++) {
        uint32_t q = pr[i];
        uint64_t dq = 1;
        uint32_t *f = forb + i * n;
        int c = 0;
        /* d = P(n) = product of the first n primes, reduced mod q */
        for (int k = 0; k < n; k++) dq = dq * first_primes[k] % q;
        for (int j = 0; j < n; j++) {
            uint32_t v = (uint32_t)((q - (uint64_t)j * dq % q) % q);
            int seen = 0;
            for (int t = 0; t < c; t++) if (f[t] == v) { seen = 1; break; }
            if (!seen) f[c++] = v;
        }
        qsort(f, c, sizeof(uint32_t), cmp_u32);
        fcnt[i] = c;
    }
    free(all);
    nblocks = (int64_t)((span + block - 1) / block);
    parts = calloc(nblocks ? nblocks : 1, sizeof(vec128));
    if (nthreads < 1) nthreads = omp_get_max_threads();
#pragma omp parallel num_threads(nthreads)
    {
        uint8_t *alive = malloc(block);
#pragma omp for schedule(dynamic, 1)
        for (int64_t b = 0; b < nblocks; b++) {
            uint64_t off = (uint64_t)b * block;
            uint64_t size = (span - off) < block ? (span - off) : block;
            u128 bb = base + off;
            uint32_t phase = (uint32_t)(bb % 30);
            for (uint64_t x = 0; x < size; x++)
                alive[x] = wheel30[(phase + x) % 30];
            for (int64_t i = 0; i < np; i++) {
                uint32_t q = pr[i];
                uint32_t bm = (uint32_t)(bb % q);
                const uint32_t *f = forb + i * n;
                for (uint32_t r = 0; r < fcnt[i]; r++) {
                    uint64_t start = (f[r] + q - bm) % q;
                    for (uint64_t x = start; x < size; x += q) alive[x] = 0;
                }
            }
            for (uint64_t x = 0; x < size; x++)
                if (alive[x] && vpush(&parts[b], (u128)(off + x))) fail = 1;
        }
        free(alive);
    }
    free(pr); free(forb); free(fcnt);
    if (fail) {
        for (int64_t i = 0; i < nblocks; i++) free(parts[i].v);
        free(parts);
        return ORC_ENOMEM;
    }
    {
        int64_t r = emit_sorted(parts, (int)nblocks, out, cap, 0);
        free(parts);
        return r;
    }
}
