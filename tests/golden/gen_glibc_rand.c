/* Writes known outputs of glibc rand() after srand(123456789), the seed of
 * the reference's initial conditions (Solver/src/solver.c:1121):
 * the first 2048 values, then every 50000th value up to index 1000000.
 * Build and run:  gcc -O2 -o /tmp/gen_glibc_rand tests/golden/gen_glibc_rand.c
 *                 /tmp/gen_glibc_rand > tests/golden/glibc_rand_123456789.txt */
#include <stdio.h>
#include <stdlib.h>

int main(void) {
  srand(123456789u);
  printf("# index value  (glibc rand() after srand(123456789); RAND_MAX = %d)\n", RAND_MAX);
  for (long i = 0; i <= 1000000; ++i) {
    const int v = rand();
    if (i < 2048 || i % 50000 == 0) printf("%ld %d\n", i, v);
  }
  return 0;
}
