/*
 * orc_mt19937.c -- MT19937 restatement of the oracle (test infrastructure, see oracle.h).
 */
#include "orc_internal.h"

#include <stdlib.h>

void* orc_mt19937_create(void) { return calloc(1, 16); }
void  orc_mt19937_destroy(void* m) { free(m); }
int   orc_mt19937_generate(void* m, uint64_t seed, int* initialized, int kind, void* out, size_t n, uint64_t addr,
                           double a, double b)
{
    (void)m; (void)seed; (void)initialized; (void)kind; (void)out; (void)n; (void)addr; (void)a; (void)b;
    return ST_TYPE_ERROR;
}
