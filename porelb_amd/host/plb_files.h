/* plb_files.h -- the reference's file formats, for the command lines.
 *   flag.bin : u8 [nz][ny][nx], x fastest, 1 = pore (preprocess.c:61-69)
 *   V.bin    : i32 num_node, num_p, nx, ny, nz; i32 V_sub_start[num_p];
 *              NODE_INFO[num_node] (preprocess.c:166-193)
 *   part_PPP_TTTTTTT.bin : f64 Rho[N_p], Ux[N_p], Uy[N_p], Uz[N_p] of rank P
 *              (porelb.c:375-391)
 */
#ifndef PLB_FILES_H
#define PLB_FILES_H
#include <stdint.h>

#include "plb.h"

typedef struct {
    int32_t num_node, num_p, nx, ny, nz;
    int32_t *V_sub_start;
    plb_node_info *V;
} plb_vfile;

/* 0, -1 (open), -2 (short read), -3 (memory) */
int plb_read_flag(const char *path, int64_t nvox, uint8_t *out);
int plb_write_flag(const char *path, int64_t nvox, const uint8_t *flag);
/* reads every rank's records; 0, -1 (open), -5 (short read), -4 (memory) */
int plb_read_vfile(const char *path, plb_vfile *v);
/* the same in two parts: the header and V_sub_start only (v->V = NULL), then
 * the records of a vfile whose header is already read */
int plb_read_vfile_header(const char *path, plb_vfile *v);
int plb_read_vfile_records(const char *path, plb_vfile *v);
int plb_write_vfile(const char *path, const plb_vfile *v);
void plb_free_vfile(plb_vfile *v);
/* part_%03d_%07d.bin for rank pid; 0 or -3 (open) */
int plb_write_part(int pid, int it, int64_t n, const double *rho, const double *ux,
                   const double *uy, const double *uz);
int plb_read_part(int pid, int it, int64_t n, double *rho, double *ux, double *uy, double *uz);
#endif
