/* plb_files.c -- readers and writers of flag.bin, V.bin and part files. */
#include "plb_files.h"

#include <stdio.h>
#include <stdlib.h>

int plb_read_flag(const char *path, int64_t nvox, uint8_t *out)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) return -1;
    size_t got = fread(out, 1, (size_t)nvox, fp);
    fclose(fp);
    return got == (size_t)nvox ? 0 : -2;
}

int plb_write_flag(const char *path, int64_t nvox, const uint8_t *flag)
{
    FILE *fp = fopen(path, "wb");
    if (!fp) return -1;
    size_t put = fwrite(flag, 1, (size_t)nvox, fp);
    fclose(fp);
    return put == (size_t)nvox ? 0 : -1;
}

int plb_read_vfile_header(const char *path, plb_vfile *v)
{
    v->V_sub_start = NULL;
    v->V = NULL;
    FILE *fp = fopen(path, "rb");
    if (!fp) return -1;
    int32_t hdr[5];
    if (fread(hdr, sizeof(int32_t), 5, fp) != 5) {
        fclose(fp);
        return -5;
    }
    v->num_node = hdr[0];
    v->num_p = hdr[1];
    v->nx = hdr[2];
    v->ny = hdr[3];
    v->nz = hdr[4];
    if (v->num_node < 0 || v->num_p < 1) {
        fclose(fp);
        return -5;
    }
    v->V_sub_start = (int32_t *)calloc((size_t)v->num_p, sizeof(int32_t));
    if (!v->V_sub_start) {
        fclose(fp);
        return -4;
    }
    if (fread(v->V_sub_start, sizeof(int32_t), (size_t)v->num_p, fp) != (size_t)v->num_p) {
        fclose(fp);
        plb_free_vfile(v);
        return -5;
    }
    fclose(fp);
    return 0;
}

int plb_read_vfile_records(const char *path, plb_vfile *v)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) return -1;
    v->V = (plb_node_info *)malloc(sizeof(plb_node_info) * (size_t)(v->num_node ? v->num_node : 1));
    if (!v->V) {
        fclose(fp);
        return -4;
    }
    /* the records follow the 5-word header and V_sub_start */
    if (fseek(fp, (long)(sizeof(int32_t) * (5 + (size_t)v->num_p)), SEEK_SET) != 0 ||
        fread(v->V, sizeof(plb_node_info), (size_t)v->num_node, fp) != (size_t)v->num_node) {
        fclose(fp);
        free(v->V);
        v->V = NULL;
        return -5;
    }
    fclose(fp);
    return 0;
}

int plb_read_vfile(const char *path, plb_vfile *v)
{
    int rc = plb_read_vfile_header(path, v);
    if (rc) return rc;
    rc = plb_read_vfile_records(path, v);
    if (rc) plb_free_vfile(v);
    return rc;
}

int plb_write_vfile(const char *path, const plb_vfile *v)
{
    FILE *fp = fopen(path, "wb");
    if (!fp) return -1;
    int32_t hdr[5] = {v->num_node, v->num_p, v->nx, v->ny, v->nz};
    int ok = fwrite(hdr, sizeof(int32_t), 5, fp) == 5 &&
             fwrite(v->V_sub_start, sizeof(int32_t), (size_t)v->num_p, fp) == (size_t)v->num_p &&
             fwrite(v->V, sizeof(plb_node_info), (size_t)v->num_node, fp) == (size_t)v->num_node;
    ok &= fclose(fp) == 0;
    return ok ? 0 : -1;
}

void plb_free_vfile(plb_vfile *v)
{
    free(v->V_sub_start);
    free(v->V);
    v->V_sub_start = NULL;
    v->V = NULL;
}

int plb_write_part(int pid, int it, int64_t n, const double *rho, const double *ux,
                   const double *uy, const double *uz)
{
    char name[100];
    snprintf(name, sizeof name, "part_%03d_%07d.bin", pid, it);
    FILE *fp = fopen(name, "wb");
    if (!fp) return -3;
    /* the reference leaves short writes unchecked; a full disk is an error here */
    int ok = fwrite(rho, sizeof(double), (size_t)n, fp) == (size_t)n &&
             fwrite(ux, sizeof(double), (size_t)n, fp) == (size_t)n &&
             fwrite(uy, sizeof(double), (size_t)n, fp) == (size_t)n &&
             fwrite(uz, sizeof(double), (size_t)n, fp) == (size_t)n;
    ok &= fclose(fp) == 0;
    return ok ? 0 : -3;
}

int plb_read_part(int pid, int it, int64_t n, double *rho, double *ux, double *uy, double *uz)
{
    char name[100];
    snprintf(name, sizeof name, "part_%03d_%07d.bin", pid, it);
    FILE *fp = fopen(name, "rb");
    if (!fp) return -1;
    int ok = fread(rho, sizeof(double), (size_t)n, fp) == (size_t)n &&
             fread(ux, sizeof(double), (size_t)n, fp) == (size_t)n &&
             fread(uy, sizeof(double), (size_t)n, fp) == (size_t)n &&
             fread(uz, sizeof(double), (size_t)n, fp) == (size_t)n;
    fclose(fp);
    return ok ? 0 : -2;
}
