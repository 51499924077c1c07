// Host build of the round-prime arithmetic the kernels use
// (first-knowns_amd/csrc/fk_kill.h, fk_arith.h).  Reads lines
//   K family n q d0 d1 d2 d3      -> count and the kill set (value domain)
//   W family n q d0 d1 d2 d3      -> the round-table row of q
//   R q nl t0 t1 t2 t3            -> the residue of the value t (four hex
//                                    32-bit limbs) by the 16-bit-limb
//                                    double-precision path with nl limbs,
//                                    two-fix and one-fix forms, and folded
//   N q vlo vhi                   -> one-fix and folded residues of the value
//                                    (vhi:vlo < 2^53) held as one double
//   I n q nl d0 d1 d2 d3 t0 t1 t2 t3 -> the AP index -t/d mod q of the value t
//                                    from q's index row (nl limbs), and the
//                                    row's words
#include "../first-knowns_amd/csrc/fk_kill.h"
#include <cstdio>
#include <cstring>
#include <vector>
int main()
{
    char line[512];
    while (fgets(line, sizeof line, stdin)) {
        if (line[0] == 'K') {
            int fam, n;
            unsigned q;
            uint32_t d[4];
            if (sscanf(line + 1, "%d %d %u %x %x %x %x", &fam, &n, &q, &d[0],
                       &d[1], &d[2], &d[3]) != 7) continue;
            const fk::PrimeK pk = fk::make_primek(q);
            const uint32_t stride = fk::kill_max(fam, n) + 3;
            std::vector<uint32_t> out(stride);
            const uint32_t c = fk::kill_values(fam, n, d, pk, out.data(), stride);
            printf("%u", c);
            for (uint32_t i = 0; i < stride; i++) printf(" %u", out[i]);
            printf("\n");
        } else if (line[0] == 'W') {
            int fam, n;
            unsigned q;
            uint32_t d[4];
            if (sscanf(line + 1, "%d %d %u %x %x %x %x", &fam, &n, &q, &d[0],
                       &d[1], &d[2], &d[3]) != 7) continue;
            const fk::PrimeK pk = fk::make_primek(q);
            const uint32_t stride = fk::round_row_stride(fam, n);
            std::vector<uint32_t> row(stride);
            fk::round_row(fam, n, d, pk, row.data(), stride);
            fk::RoundRow h;
            memcpy(&h, row.data(), sizeof h);
            printf("%u %.17g %.17g %.17g %.17g %.17g %u %u", stride, h.qd, h.qi,
                   h.c16, h.c32, h.c48, h.q, h.cnt);
            for (uint32_t i = fk::kRowHead; i < stride; i++) printf(" %u", row[i]);
            printf("\n");
        } else if (line[0] == 'N') {
            unsigned q;
            uint32_t lo, hi;
            if (sscanf(line + 1, "%u %x %x", &q, &lo, &hi) != 3) continue;
            const fk::PrimeD P = fk::make_primed(q);
            const double s = (double)(((uint64_t)hi << 32) | lo);
            printf("%u %u\n", fk::resid_d1(s, P.qd, P.qi, P.q),
                   fk::resid_fold_d(s, P.qd, P.qi, P.q));
        } else if (line[0] == 'I') {
            int n, nl;
            unsigned q;
            uint32_t d[4], t[4];
            if (sscanf(line + 1, "%d %u %d %x %x %x %x %x %x %x %x", &n, &q, &nl,
                       &d[0], &d[1], &d[2], &d[3], &t[0], &t[1], &t[2],
                       &t[3]) != 11) continue;
            const fk::PrimeK pk = fk::make_primek(q);
            uint32_t row[fk::kIndexStride];
            fk::index_row(n, d, pk, row);
            fk::IndexRow h;
            memcpy(&h, row, sizeof h);
            double s = 0;
            for (int k = 0; k < nl; k++) {
                const double a = (double)((t[k >> 1] >> ((k & 1) * 16)) & 0xffffu);
                s = fma(a, k < 4 ? h.u[k] : h.uw[k - 4], s);
            }
            // the 22-bit limb form (q < 2^24, values below 2^88)
            uint32_t t22 = 0xffffffffu;
            if (q < (1u << 24) && t[3] == 0 && t[2] < (1u << 24)) {
                const uint32_t M = 0x3fffffu;
                const double a22[4] = {(double)(t[0] & M),
                                       (double)(((t[0] >> 22) | (t[1] << 10)) & M),
                                       (double)(((t[1] >> 12) | (t[2] << 20)) & M),
                                       (double)((t[2] >> 2) & M)};
                double s22 = a22[0] * h.v[0];
                for (int k = 1; k < 4; k++) s22 = fma(a22[k], h.v[k], s22);
                t22 = fk::resid_d1(s22, h.qd, h.qi, h.q);
            }
            printf("%u %u %u %u", fk::resid_d1(s, h.qd, h.qi, h.q), h.q, h.n, t22);
            for (uint32_t i = 0; i < fk::kIndexStride; i++) printf(" %u", row[i]);
            printf("\n");
        } else if (line[0] == 'R') {
            unsigned q;
            int nl;
            uint32_t t[4];
            if (sscanf(line + 1, "%u %d %x %x %x %x", &q, &nl, &t[0], &t[1],
                       &t[2], &t[3]) != 6) continue;
            const fk::PrimeD P = fk::make_primed(q);
            double a[8];
            for (int k = 0; k < 8; k++)
                a[k] = (double)((t[k >> 1] >> ((k & 1) * 16)) & 0xffffu);
            double s = a[0];
            for (int k = 1; k < nl; k++) s = fma(a[k], P.c[k - 1], s);
            // the 22-bit limb form (q < 2^24, values below 2^88)
            uint32_t t22 = 0xffffffffu, f22 = 0xffffffffu;
            if (q < (1u << 24) && t[3] == 0 && t[2] < (1u << 24)) {
                const uint32_t M = 0x3fffffu;
                const double a22[4] = {(double)(t[0] & M),
                                       (double)(((t[0] >> 22) | (t[1] << 10)) & M),
                                       (double)(((t[1] >> 12) | (t[2] << 20)) & M),
                                       (double)((t[2] >> 2) & M)};
                const double s22 = fma(a22[3], P.c22[2],
                                       fma(a22[2], P.c22[1], fma(a22[1], P.c22[0], a22[0])));
                t22 = fk::resid_d1(s22, P.qd, P.qi, P.q);
                f22 = fk::resid_fold_d(s22, P.qd, P.qi, P.q);
            }
            printf("%u %u %u %u %u\n", fk::resid_d(s, P.qd, P.qi, P.q),
                   fk::resid_d1(s, P.qd, P.qi, P.q),
                   fk::resid_fold_d(s, P.qd, P.qi, P.q), t22, f22);
        }
    }
}
