/*
 * preprocess.x -- drop-in for the reference's preprocessor
 * (preprocess.c:195-204): `./preprocess.x np` reads a 400^3 flag.bin and
 * writes V.bin.  The sparse index is built on the GPU (plb_preprocess).
 * Optional trailing arguments give the dimensions and file names the
 * reference hard-codes:  ./preprocess.x np [nx ny nz [flagfile [vfile]]]
 */
#include <stdio.h>
#include <stdlib.h>

#include "plb.h"
#include "plb_files.h"

int main(int argc, char *argv[])
{
    if (argc != 2 && argc != 5 && argc != 6 && argc != 7) {
        fprintf(stderr, "Error, usage: ./preprocess.x np [nx ny nz [flagfile [vfile]]]\n");
        exit(-1);
    }
    int num_p = atoi(argv[1]);
    int nx = 400, ny = 400, nz = 400; /* preprocess.c:200 */
    const char *flagfile = "flag.bin", *vfile = "V.bin";
    if (argc >= 5) {
        nx = atoi(argv[2]);
        ny = atoi(argv[3]);
        nz = atoi(argv[4]);
    }
    if (argc >= 6) flagfile = argv[5];
    if (argc >= 7) vfile = argv[6];
    const char *dev_env = getenv("PLB_DEVICE");
    int device = dev_env ? atoi(dev_env) : 0;
    if (nx <= 0 || ny <= 0 || nz <= 0) {
        fprintf(stderr, "Error, bad dimensions\n");
        exit(-1);
    }
    int64_t nvox = (int64_t)nx * ny * nz;
    uint8_t *flag = (uint8_t *)malloc((size_t)nvox);
    if (!flag) {
        fprintf(stderr, "Memory allocating failed!\n");
        exit(-3);
    }
    int rc = plb_read_flag(flagfile, nvox, flag);
    if (rc == -1) {
        fprintf(stderr, "flag data file opening error!\n");
        exit(-1);
    }
    if (rc == -2) {
        fprintf(stderr, "flag data file reading error!\n");
        exit(-2);
    }
    int64_t n = 0;
    plb_vfile v;
    v.num_p = num_p;
    v.nx = nx;
    v.ny = ny;
    v.nz = nz;
    rc = plb_preprocess(flag, nx, ny, nz, num_p, device, NULL, NULL, &n);
    if (rc == 0 && n >= INT32_MAX) rc = PLB_ERR_RANGE;
    if (rc) {
        fprintf(stderr, "preprocess failed: %s\n", plb_strerror(rc));
        free(flag);
        exit(rc);
    }
    printf("num_nodes = %d\n", (int)n);
    v.num_node = (int32_t)n;
    v.V_sub_start = (int32_t *)calloc((size_t)num_p, sizeof(int32_t));
    v.V = (plb_node_info *)malloc(sizeof(plb_node_info) * (size_t)(n ? n : 1));
    if (!v.V_sub_start || !v.V) {
        fprintf(stderr, "Memory allocating failed!\n");
        exit(-3);
    }
    rc = plb_preprocess(flag, nx, ny, nz, num_p, device, v.V, v.V_sub_start, &n);
    free(flag);
    if (rc) {
        fprintf(stderr, "preprocess failed: %s\n", plb_strerror(rc));
        exit(rc);
    }
    if (plb_write_vfile(vfile, &v)) {
        fprintf(stderr, "File writing failed\n");
        exit(-1);
    }
    printf("V.bin written %ld Bytes\n", (long)(v.num_node * (long)sizeof(plb_node_info)));
    printf("NODE_INFO_size: %ld Bytes\n", (long)sizeof(plb_node_info));
    plb_free_vfile(&v);
    return 0;
}
