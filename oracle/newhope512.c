/* NewHope-512 (TEST INFRASTRUCTURE ONLY): newhope.c with N = 512, functions
 * named oracle_newhope512_*. */
#define NH_N 512
#include "newhope.c"
