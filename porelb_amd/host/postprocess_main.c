/*
 * postprocess.x -- the reference's post-processor (postprocess.c:87-319)
 * without the proprietary TECIO library:
 *
 *     ./postprocess.x A t1 t2 ...        full field  -> full_TTTTTTT.dat
 *     ./postprocess.x X|Y|Z n t1 t2 ...  one slice   -> Slice_[XYZ]_TTTTTTT.dat
 *     ./postprocess.x K t1 t2 ...        permeability report on stdout
 *
 * Reads V.bin, flag.bin and every rank's part_PPP_TTTTTTT.bin and scatters
 * the macros to the voxel grid by V's coordinates (postprocess.c:160-214).
 * Slices are the reference's ASCII format (postprocess.c:19-85; the flag
 * column is printed as a number, where the reference passes an int to %e).
 * `A` writes the same variables "X Y Z rho vx vy vz flag" as an ASCII
 * Tecplot BLOCK zone instead of TECIO's binary file.
 *
 * `K` is the permeability report the reference leaves to getK.c's analytic
 * estimate: from the part files alone it prints, per time, the FP64 sums
 * over all pore nodes in V.bin order, <u_x> = sum u_x / N, the superficial
 * velocity q = sum u_x / (nx ny nz) and K = nu q / g_x with
 * nu = (tau - 1/2)/3 (porelb.c:89).  tau and g_x come from PLB_TAU and
 * PLB_GX, with porelb.x's defaults (the reference's porelb.c:83-94).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "plb.h"
#include "plb_files.h"

static int nx, ny, nz;
static double *rho, *u, *v_, *w;
static unsigned char *flag;

static void print_help(void)
{
    printf("Usage: ./postprocess.x  [A|X|Y|Z] [n] time_list\n");
    printf("Output 3D full data     : ./postprocess.x A time_list\n");
    printf("Output slice at  X = 20 : ./postprocess.x X 20  time_list\n");
    printf("Permeability report     : ./postprocess.x K time_list\n");
}

static double env_double(const char *name, double dflt)
{
    const char *s = getenv(name);
    return (s && *s) ? atof(s) : dflt;
}

/* Sums over every rank's part file for time t, in V.bin record order. */
static void report_K(const plb_vfile *V, int t, double tau, double gx)
{
    double s_rho = 0.0, s_ux = 0.0;
    for (int p = 0; p < V->num_p; p++) {
        const long b = V->V_sub_start[p];
        const long N = (p + 1 < V->num_p ? V->V_sub_start[p + 1] : V->num_node) - b;
        double *buf = malloc(sizeof(double) * 4 * (N ? N : 1));
        if (!buf || plb_read_part(p, t, N, buf, buf + N, buf + 2 * N, buf + 3 * N)) {
            fprintf(stderr, "Part file openning error!\n");
            exit(-1);
        }
        for (long k = 0; k < N; k++) {
            s_rho += buf[k];
            s_ux += buf[N + k];
        }
        free(buf);
    }
    const double nvox = (double)V->nx * V->ny * V->nz;
    const double mean = V->num_node ? s_ux / V->num_node : 0.0;
    const double q = s_ux / nvox;
    const double nu = 1.0 / 3.0 * (tau - 0.5); /* porelb.c:89 */
    printf("t= %d N= %d sum_rho= %.17e sum_ux= %.17e mean_ux= %.17e q= %.17e K= %.17e\n", t,
           V->num_node, s_rho, s_ux, mean, q, nu * q / gx);
}

static void slice(char axis, int n, int t)
{
    char name[100];
    snprintf(name, sizeof name, "Slice_%c_%07d.dat", axis, t);
    FILE *fp = fopen(name, "w");
    if (!fp) {
        fprintf(stderr, "Slice file open error!\n");
        exit(-3);
    }
    fprintf(fp, "TITLE = \" %c slice at %04d\"\n", axis == 'Z' ? 'Y' : axis, n);
    fprintf(fp, "Variables = \"X\", \"Y\", \"Rho\", \"Ux\", \"Uy\", \"Uz\", \"flag\"\n");
    if (axis == 'X')
        fprintf(fp, "ZONE I = %4d, J = %4d, F=POINT\n", ny, nz);
    else
        fprintf(fp, "ZONE I = %4d, J = %4d, F=POINT\n", nz, nx);
    const long plane = (long)nx * ny;
    if (axis == 'X') {
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++) {
                long loc = k * plane + (long)j * nx + n;
                fprintf(fp, "%f %f % 21.14e  % 21.14E  % 21.14E  % 21.14E  % 8.1e\n", (double)k,
                        (double)j, rho[loc], u[loc], v_[loc], w[loc], (double)flag[loc]);
            }
    } else if (axis == 'Y') {
        for (int k = 0; k < nz; k++)
            for (int i = 0; i < nx; i++) {
                long loc = k * plane + (long)n * nx + i;
                fprintf(fp, "%f %f % 21.14e  % 21.14E  % 21.14E  % 21.14E  % 8.1e\n", (double)k,
                        (double)i, rho[loc], u[loc], v_[loc], w[loc], (double)flag[loc]);
            }
    } else {
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                long loc = n * plane + (long)j * nx + i;
                fprintf(fp, "%f %f % 21.14e  % 21.14E  % 21.14E  % 21.14E  % 8.1e\n", (double)j,
                        (double)i, rho[loc], u[loc], v_[loc], w[loc], (double)flag[loc]);
            }
    }
    if (fclose(fp) != 0) {
        fprintf(stderr, "Slice file write error!\n");
        exit(-3);
    }
}

static void full(int t)
{
    char name[100];
    snprintf(name, sizeof name, "full_%07d.dat", t);
    FILE *fp = fopen(name, "w");
    if (!fp) {
        fprintf(stderr, "Full file open error!\n");
        exit(-3);
    }
    const long size = (long)nx * ny * nz;
    fprintf(fp, "TITLE = \"SIMPLE DATASET\"\n");
    fprintf(fp, "VARIABLES = \"X\" \"Y\" \"Z\" \"rho\" \"vx\" \"vy\" \"vz\" \"flag\"\n");
    fprintf(fp, "ZONE T=\"Simple Zone\", I=%d, J=%d, K=%d, DATAPACKING=BLOCK, SOLUTIONTIME=%d\n",
            nx, ny, nz, t);
    for (int var = 0; var < 8; var++)
        for (long loc = 0; loc < size; loc++) {
            const long x = loc % nx, y = (loc / nx) % ny, z = loc / ((long)nx * ny);
            double val = var == 0 ? x : var == 1 ? y : var == 2 ? z : var == 3 ? rho[loc]
                       : var == 4 ? u[loc] : var == 5 ? v_[loc] : var == 6 ? w[loc] : flag[loc];
            fprintf(fp, "%.14e\n", val);
        }
    if (fclose(fp) != 0) {
        fprintf(stderr, "Full file write error!\n");
        exit(-3);
    }
}

int main(int argc, char *argv[])
{
    if (argc < 2) {
        print_help();
        exit(-2);
    }
    const char mode = argv[1][0];
    if (strlen(argv[1]) != 1 ||
        (mode != 'A' && mode != 'X' && mode != 'Y' && mode != 'Z' && mode != 'K')) {
        fprintf(stderr, "Not implemented yet!\n");
        exit(-3);
    }
    int first = (mode == 'A' || mode == 'K') ? 2 : 3, n_slice = 0;
    if (first == 3) {
        printf("argc = %d\n", argc); /* postprocess.c:107 */
        if (argc < 3) {
            print_help();
            exit(-2);
        }
        n_slice = atoi(argv[2]);
    }
    plb_vfile V;
    if (plb_read_vfile("V.bin", &V)) {
        fprintf(stderr, "Node map file openning error!\n");
        exit(-1);
    }
    printf(">>> Total num of fluid  nodes = %d\n", V.num_node);
    printf(">>> num_p = %d\n", V.num_p);
    nx = V.nx;
    ny = V.ny;
    nz = V.nz;
    if (mode == 'K') {
        const double tau = env_double("PLB_TAU", 0.8), gx = env_double("PLB_GX", 1e-3);
        printf("Total %d time seriese data\n", argc - first);
        for (int a = first; a < argc; a++) report_K(&V, atoi(argv[a]), tau, gx);
        plb_free_vfile(&V);
        return 0;
    }
    const long size = (long)nx * ny * nz;
    if ((mode == 'X' && (n_slice < 0 || n_slice >= nx)) || (mode == 'Y' && (n_slice < 0 || n_slice >= ny)) ||
        (mode == 'Z' && (n_slice < 0 || n_slice >= nz))) {
        fprintf(stderr, "Slice index out of range!\n");
        exit(-3);
    }
    rho = calloc(size, sizeof(double));
    u = calloc(size, sizeof(double));
    v_ = calloc(size, sizeof(double));
    w = calloc(size, sizeof(double));
    flag = calloc(size, 1);
    if (plb_read_flag("flag.bin", size, flag) == -1) {
        fprintf(stderr, "Flag file openning error!\n");
        exit(-1);
    }
    printf("Total %d time seriese data\n", argc - first);
    for (int a = first; a < argc; a++) {
        const int t = atoi(argv[a]);
        for (int p = 0; p < V.num_p; p++) {
            const long b = V.V_sub_start[p];
            const long N = (p + 1 < V.num_p ? V.V_sub_start[p + 1] : V.num_node) - b;
            double *buf = malloc(sizeof(double) * 4 * (N ? N : 1));
            if (plb_read_part(p, t, N, buf, buf + N, buf + 2 * N, buf + 3 * N)) {
                fprintf(stderr, "Part file openning error!\n");
                exit(-1);
            }
            for (long k = 0; k < N; k++) {
                const plb_node_info *r = &V.V[b + k];
                const long loc = ((long)r->z * ny + r->y) * nx + r->x;
                rho[loc] = buf[k];
                u[loc] = buf[N + k];
                v_[loc] = buf[2 * N + k];
                w[loc] = buf[3 * N + k];
            }
            free(buf);
        }
        if (mode == 'A')
            full(t);
        else
            slice(mode, n_slice, t);
    }
    plb_free_vfile(&V);
    free(rho);
    free(u);
    free(v_);
    free(w);
    free(flag);
    return 0;
}
