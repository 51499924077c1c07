// Host build of the engine's table builders (first-knowns_amd/csrc/
// fk_plan.cpp).  Reads lines
//   B tail q r1 r2 ... ; q2 s1 s2 ...
//        -> kill bitmaps with that tail appended to one table:
//           "offsets o1 o2 ..." then the words
//   W family n q1 q2 ... qk
//        -> "count c" (wheel_count), the wheel's residues (wheel_offsets),
//           then the residues of the factored form: the wheel without qk
//           lifted over qk (wheel_lifts), in the device kernel's order
#include "../first-knowns_amd/csrc/fk_plan.h"
#include <cstdio>
#include <sstream>
#include <string>
#include <vector>
int main()
{
    std::string line;
    char buf[1 << 16];
    while (fgets(buf, sizeof buf, stdin)) {
        line = buf;
        if (!line.empty() && line[0] == 'W') {
            std::stringstream ss(line.substr(1));
            int family, n;
            ss >> family >> n;
            std::vector<uint32_t> wheel;
            uint32_t q;
            while (ss >> q) wheel.push_back(q);
            bool ok = false;
            printf("count %llu\n",
                   (unsigned long long)fk::wheel_count(family, n, wheel, &ok));
            uint64_t M = 0, Mb = 0;
            for (uint64_t o : fk::wheel_offsets(family, n, wheel, &M, &ok))
                printf("%llu ", (unsigned long long)o);
            printf("\n");
            std::vector<uint32_t> head(wheel.begin(), wheel.end() - 1);
            std::vector<uint64_t> base = fk::wheel_offsets(family, n, head, &Mb, &ok);
            uint32_t nj = 0;
            std::vector<uint32_t> lifts = fk::wheel_lifts(family, n, wheel.back(), Mb, &nj);
            for (uint64_t b : base)
                for (uint32_t t = 0; t < nj; t++)
                    printf("%llu ", (unsigned long long)(b + (uint64_t)lifts[(b % wheel.back()) * nj + t] * Mb));
            printf("\n");
            fflush(stdout);
            continue;
        }
        if (line.empty() || line[0] != 'B') continue;
        std::vector<uint32_t> offs, words;
        std::stringstream all(line.substr(1));
        uint32_t tail = 0;
        all >> tail;
        std::string part;
        while (std::getline(all, part, ';')) {
            std::stringstream ss(part);
            uint32_t q, r;
            if (!(ss >> q)) continue;
            std::vector<uint32_t> kills;
            while (ss >> r) kills.push_back(r);
            fk::append_kill_bitmap(kills, q, tail, offs, words);
        }
        printf("offsets");
        for (uint32_t o : offs) printf(" %u", o);
        printf("\n");
        for (uint32_t w : words) printf("%u ", w);
        printf("\n");
        fflush(stdout);
    }
    return 0;
}
