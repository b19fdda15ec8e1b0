/* Writes the oracle's closed-form jump table in the reference's text format
 * (test infrastructure: the reference build reads ./jump_matrix.txt). */
#include <stdio.h>
#include <stdlib.h>

#include "mrg8_oracle.h"

int main(int argc, char** argv) {
  if (argc != 2) {
    fprintf(stderr, "usage: %s OUT\n", argv[0]);
    return 2;
  }
  uint32_t* tab = (uint32_t*)malloc(sizeof(uint32_t) * 64 * MRG8O_NJUMP);
  if (!tab) return 1;
  mrg8o_jump_table(tab);
  int rc = mrg8o_write_jump_table(tab, argv[1]);
  free(tab);
  return rc ? 1 : 0;
}
