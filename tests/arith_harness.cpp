// Reads "first nbases limb0..limb7" lines (hex limbs, little-endian) and
// prints 1/0 for the product strong test fk::is_sprp(v, first, nbases).
#include "../first-knowns_amd/csrc/fk_arith.h"
#include <cstdio>
#include <cstdlib>
#include <cstring>
int main(){
  char line[512];
  while (fgets(line, sizeof line, stdin)) {
    uint32_t v[8]; int first, nb;
    if (sscanf(line, "%d %d %x %x %x %x %x %x %x %x", &first, &nb, &v[0],&v[1],&v[2],&v[3],&v[4],&v[5],&v[6],&v[7]) != 10) continue;
    printf("%d\n", fk::is_sprp(v, first, nb) ? 1 : 0);
  }
}
