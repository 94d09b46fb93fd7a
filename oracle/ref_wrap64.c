/*
 * TEST INFRASTRUCTURE ONLY.  The reference's CatmullClark.h instantiated with
 * every float replaced by double: the FP64 run of the reference's own CPU
 * code that single-precision results are measured against.  All system
 * headers the reference needs are included before the substitution.
 */
#include <assert.h>
#include <stdbool.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#define float double
#define CC_IMPLEMENTATION
#include "CatmullClark.h"
#undef float

void ref_set_threads(int n) { omp_set_num_threads(n); }
int ref_real_size(void) { return (int)sizeof(cc_VertexPoint) / 3; }
