/*
 * Harness around the reference preprocessor, compiled from the reference's
 * own preprocess.c where it lies (the recipe in oracle/Makefile puts that
 * directory on the include path; nothing is copied).  The reference main
 * (preprocess.c:195-204) is hard-coded to a 400^3 "flag.bin"; this harness
 * renames it and calls the reference's build_V/save_V with the dimensions
 * and file names given on the command line.
 *
 * TEST INFRASTRUCTURE ONLY: used to produce golden fixtures and to pin the
 * oracle's build_V restatement.
 *
 * usage: ref_preprocess.x np nx ny nz flagfile outfile
 */
#define main plb_reference_preprocess_main
#include "preprocess.c"
#undef main

int main(int argc, char *argv[])
{
    if (argc != 7) {
        fprintf(stderr, "usage: %s np nx ny nz flagfile outfile\n", argv[0]);
        return 10;
    }
    int num_p = atoi(argv[1]);
    V_DESC v = build_V((unsigned short)num_p, atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), argv[5]);
    save_V(&v, argv[6]);
    free_v_desc(&v);
    return 0;
}
