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
 * Prints "PLB_REF_PREPROCESS nodes= build_s= save_s=" (wall time of the
 * reference's build_V, which includes reading flagfile, and of save_V).
 */
#define _POSIX_C_SOURCE 199309L
#define main plb_reference_preprocess_main
#include "preprocess.c"
#undef main

#include <time.h>

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int main(int argc, char *argv[])
{
    if (argc != 7) {
        fprintf(stderr, "usage: %s np nx ny nz flagfile outfile\n", argv[0]);
        return 10;
    }
    int num_p = atoi(argv[1]);
    const double t0 = now_s();
    V_DESC v = build_V((unsigned short)num_p, atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), argv[5]);
    const double t1 = now_s();
    save_V(&v, argv[6]);
    const double t2 = now_s();
    printf("PLB_REF_PREPROCESS nodes=%d build_s=%.6f save_s=%.6f\n", v.num_node, t1 - t0, t2 - t1);
    free_v_desc(&v);
    return 0;
}
