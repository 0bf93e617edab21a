/*
 * mkflag.x -- writes a synthetic flag.bin (u8, x fastest, 1 = pore).
 *   ./mkflag.x noise   nx ny nz sigma porosity seed [file]
 *   ./mkflag.x spheres nx ny nz rmin rmax porosity seed [file]
 *   ./mkflag.x bcc     n chi [file]
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "plb.h"
#include "plb_files.h"

static void usage(void)
{
    fprintf(stderr,
            "usage: ./mkflag.x noise nx ny nz sigma porosity seed [file]\n"
            "       ./mkflag.x spheres nx ny nz rmin rmax porosity seed [file]\n"
            "       ./mkflag.x bcc n chi [file]\n");
    exit(-1);
}

int main(int argc, char *argv[])
{
    if (argc < 2) usage();
    const char *out = "flag.bin";
    int nx, ny, nz, rc;
    uint8_t *flag;
    if (!strcmp(argv[1], "noise")) {
        if (argc != 8 && argc != 9) usage();
        nx = atoi(argv[2]), ny = atoi(argv[3]), nz = atoi(argv[4]);
        if (argc == 9) out = argv[8];
        flag = (uint8_t *)malloc((size_t)nx * ny * nz);
        rc = plb_mask_noise(flag, nx, ny, nz, atof(argv[5]), atof(argv[6]),
                            strtoull(argv[7], NULL, 10));
    } else if (!strcmp(argv[1], "spheres")) {
        if (argc != 9 && argc != 10) usage();
        nx = atoi(argv[2]), ny = atoi(argv[3]), nz = atoi(argv[4]);
        if (argc == 10) out = argv[9];
        flag = (uint8_t *)malloc((size_t)nx * ny * nz);
        rc = plb_mask_spheres(flag, nx, ny, nz, atof(argv[5]), atof(argv[6]), atof(argv[7]),
                              strtoull(argv[8], NULL, 10));
    } else if (!strcmp(argv[1], "bcc")) {
        if (argc != 4 && argc != 5) usage();
        nx = ny = nz = atoi(argv[2]);
        if (argc == 5) out = argv[4];
        flag = (uint8_t *)malloc((size_t)nx * ny * nz);
        rc = plb_mask_bcc(flag, nx, atof(argv[3]));
    } else {
        usage();
        return -1;
    }
    if (rc) {
        fprintf(stderr, "mkflag: %s\n", plb_strerror(rc));
        return rc;
    }
    int64_t nvox = (int64_t)nx * ny * nz, pores = 0;
    for (int64_t k = 0; k < nvox; k++) pores += flag[k];
    if (plb_write_flag(out, nvox, flag)) {
        fprintf(stderr, "mkflag: cannot write %s\n", out);
        return -1;
    }
    printf("%s: %d x %d x %d, %lld pores, porosity %.6f\n", out, nx, ny, nz, (long long)pores,
           (double)pores / (double)nvox);
    free(flag);
    return 0;
}
