/*
 * TEST INFRASTRUCTURE ONLY.  Compiles the reference's example application
 * examples/subd_cpu.c (found through -I pointing at the reference tree) into
 * oracle/_ref/ so the tests can compare the library's OBJ export with the
 * reference's own ExportToObj on the same refined arrays.  Nothing of the
 * reference is copied here; this file only instantiates the example with its
 * main renamed and exposes its static exporter.
 */
#define main ref_subd_cpu_main
#include "subd_cpu.c"
#undef main

void ref_export_obj(const cc_Subd *subd, int32_t depth, const char *filename)
{
    ExportToObj(subd, depth, filename);
}
