// Host build of the engine's table builders (first-knowns_amd/csrc/
// fk_plan.cpp).  Reads lines
//   B tail q r1 r2 ... ; q2 s1 s2 ...
//        -> kill bitmaps with that tail appended to one table:
//           "offsets o1 o2 ..." then the words
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
