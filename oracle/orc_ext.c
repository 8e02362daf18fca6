/* orc_ext.c -- MTGP32, MT19937 and Sobol restatements of the oracle (test infrastructure). */
#include "orc_internal.h"

#include <stdlib.h>

struct orc_ext
{
    int unused;
};

int orc_ext_create(orc_gen* g)
{
    (void)g;
    return -1;
}
void orc_ext_destroy(orc_gen* g) { (void)g; }
int  orc_ext_generate(orc_gen* g, int kind, void* out, size_t n, uint64_t addr, double a, double b)
{
    (void)g; (void)kind; (void)out; (void)n; (void)addr; (void)a; (void)b;
    return ST_TYPE_ERROR;
}
int orc_sobol_load(const char* p) { (void)p; return -1; }
const unsigned int*       orc_sobol32_vectors(void) { return NULL; }
const unsigned long long* orc_sobol64_vectors(void) { return NULL; }
const unsigned int*       orc_scramble32_constants(void) { return NULL; }
const unsigned long long* orc_scramble64_constants(void) { return NULL; }
