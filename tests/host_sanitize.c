/*
 * Host-side code under AddressSanitizer / UndefinedBehaviorSanitizer: the
 * seeded mask generators and the flag / V.bin / part-file I/O of the command
 * lines, including short-file error paths.  Built with gcc by
 * tests/test_sanitizers.py (no GPU code is involved).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "plb.h"
#include "plb_files.h"

static int fails = 0;
#define CHECK(c)                                                   \
    do {                                                           \
        if (!(c)) {                                                \
            fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            fails++;                                               \
        }                                                          \
    } while (0)

static double porosity(const uint8_t *f, long n)
{
    long p = 0;
    for (long i = 0; i < n; i++) p += f[i] == 1;
    return (double)p / (double)n;
}

int main(void)
{
    const int nx = 24, ny = 20, nz = 16;
    const long n = (long)nx * ny * nz;
    uint8_t *a = malloc(n), *b = malloc(n);
    CHECK(plb_mask_noise(a, nx, ny, nz, 2.0, 0.3, 7) == 0);
    CHECK(fabs(porosity(a, n) - 0.3) < 1.0 / n + 1e-12);
    CHECK(plb_mask_spheres(b, nx, ny, nz, 2.0, 4.0, 0.4, 9) == 0);
    CHECK(porosity(b, n) <= 0.4 && porosity(b, n) > 0.2);
    uint8_t *c = malloc(16 * 16 * 16);
    CHECK(plb_mask_bcc(c, 16, 0.6) == 0);

    /* flag file round trip and a short read */
    CHECK(plb_write_flag("flag.bin", n, a) == 0);
    memset(b, 0, n);
    CHECK(plb_read_flag("flag.bin", n, b) == 0);
    CHECK(memcmp(a, b, n) == 0);
    CHECK(plb_read_flag("flag.bin", n + 1, b) != 0);
    CHECK(plb_read_flag("missing.bin", n, b) == -1);

    /* V.bin round trip, then truncated */
    plb_vfile v = {0}, w = {0};
    v.num_node = 5;
    v.num_p = 2;
    v.nx = nx;
    v.ny = ny;
    v.nz = nz;
    int32_t start[2] = {0, 3};
    v.V_sub_start = start;
    v.V = calloc(5, sizeof(plb_node_info));
    for (int i = 0; i < 5; i++) {
        v.V[i].x = i;
        v.V[i].y = 2 * i;
        v.V[i].z = 3 * i;
    }
    CHECK(plb_write_vfile("V.bin", &v) == 0);
    CHECK(plb_read_vfile("V.bin", &w) == 0);
    CHECK(w.num_node == 5 && w.num_p == 2 && w.V_sub_start[1] == 3 && w.V[4].z == 12);
    plb_free_vfile(&w);
    FILE *fp = fopen("V.bin", "rb"), *out = fopen("Vshort.bin", "wb");
    char buf[64];
    size_t got = fread(buf, 1, sizeof buf, fp);
    fwrite(buf, 1, got, out);
    fclose(fp);
    fclose(out);
    CHECK(plb_read_vfile("Vshort.bin", &w) == -5);
    CHECK(plb_read_vfile("missing.bin", &w) == -1);
    free(v.V);

    /* part file round trip */
    double r[4][6];
    for (int k = 0; k < 4; k++)
        for (int i = 0; i < 6; i++) r[k][i] = k + 0.25 * i;
    CHECK(plb_write_part(1, 700, 6, r[0], r[1], r[2], r[3]) == 0);
    double s[4][6];
    CHECK(plb_read_part(1, 700, 6, s[0], s[1], s[2], s[3]) == 0);
    CHECK(memcmp(r, s, sizeof r) == 0);
    CHECK(plb_read_part(1, 700, 7, s[0], s[1], s[2], s[3]) != 0);

    free(a);
    free(b);
    free(c);
    printf("host checks: %d failed\n", fails);
    return fails != 0;
}
