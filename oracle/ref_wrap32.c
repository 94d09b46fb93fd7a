/*
 * TEST INFRASTRUCTURE ONLY.  Compiles the reference's own CatmullClark.h
 * (found through -I pointing at the reference tree) into oracle/_ref/ so the
 * tests can pin the oracle against the real implementation.  Nothing of the
 * reference is copied here; this file only instantiates the header.
 */
#include <omp.h>
#define CC_IMPLEMENTATION
#include "CatmullClark.h"

/* thread count of the reference's OpenMP loops (1 = deterministic scatter) */
void ref_set_threads(int n) { omp_set_num_threads(n); }
int ref_real_size(void) { return (int)sizeof(cc_VertexPoint) / 3; }
