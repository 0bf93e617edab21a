/*
 * porelb.x -- drop-in for the reference solver (porelb.c:402-443):
 *
 *     [mpirun -np P] ./porelb.x it_max
 *
 * reads V.bin (every rank's records), runs it_max D3Q19 steps on the GPU
 * through libplb, and writes part_PPP_TTTTTTT.bin for every rank P of the
 * V.bin at it % save == 0 and after the last step, like save_data
 * (porelb.c:375-391, 436, 440).  Messages and exit codes follow the
 * reference (-10 argc, -1 V.bin, -2 rank count, -3/-4 check_V, -6 density).
 * Under an MPI launcher (PMI_* or OMPI_COMM_WORLD_* in the environment) the
 * launcher's size must equal V.bin's num_p; rank 0 drives the GPU for all
 * ranks and the other ranks exit at once.
 *
 * Parameters the reference hard-codes (porelb.c:83-94) come from the
 * environment; defaults are the reference's values:
 *   PLB_TAU (0.8)  PLB_GX (1e-3)  PLB_RHO0 (1)  PLB_SAVE (1000)
 *   PLB_PRECISION fp64|fp32|fp16 (fp64)  PLB_BB refcompat|halfway (refcompat)
 *   PLB_FP16_ROUND nearest|stochastic (stochastic): rounding of FP16 stores
 *   PLB_RESERVE_GIB (0): device memory held below the lattice (placement_reserve_gib)
 *   PLB_LAYOUT soa|aosoa (aosoa)  PLB_ORDER reference|brick|pencil (pencil)
 *   PLB_BRICK bx,by,bz (pencil: nx,64,1)  PLB_INDEX i32|tile8|bitmap (tile8)
 *   PLB_BLOCK (128 below 4 M pore nodes, else 256)  PLB_GRAPH (0: steps per graph replay)
 *   PLB_WAVES (step-kernel waves per SIMD, 0 = no cap; 5 for FP64/FP32 from
 *   4 M pore nodes, else 0)
 *   PLB_DEVICE (0)  PLB_PRINT 1|0 (1: the per-step line of porelb.c:435)
 *   PLB_AUTOTUNE cachefile: time index encoding x layout x node order, then
 *   block size x nodes per thread x waves per SIMD, on this V.bin once; reuse
 *   the result from cachefile after
 *   PLB_TOL tol, PLB_CHECK m (off; 1000): stop before it_max once <u_x> changes by
 *   at most tol * max(|<u_x>|, 1e-3 |gx|) between checks m steps apart (the
 *   rule of plb_run_to_convergence); the final part files carry the step reached
 *   PLB_ASYNC_SAVE 1|0 (1): part files are written by a worker thread from one
 *   of two host buffers while the GPU keeps stepping (same files)
 *   PLB_CHECKPOINT file: the stored population state written after the last
 *   step, bit for bit in every precision (header: step, node count,
 *   precision, FP16 scale, rho0)
 *   PLB_RESTART file: start from such a checkpoint; it must come from the
 *   same V.bin, precision and rho0 (else exit -5).  Together they fill the
 *   empty save_state/recovery_state stubs of porelb.c:367-373
 */
#define _POSIX_C_SOURCE 199309L
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "plb.h"
#include "plb_files.h"
#include "plb_tune.h"

static int env_int(const char *k, int def)
{
    const char *v = getenv(k);
    return v && *v ? atoi(v) : def;
}
static double env_dbl(const char *k, double def)
{
    const char *v = getenv(k);
    return v && *v ? atof(v) : def;
}
static int env_choice(const char *k, const char *a, const char *b, const char *c, int def)
{
    const char *v = getenv(k);
    if (!v || !*v) return def;
    if (a && !strcmp(v, a)) return 0;
    if (b && !strcmp(v, b)) return 1;
    if (c && !strcmp(v, c)) return 2;
    fprintf(stderr, ">>> bad value %s=%s\n", k, v);
    exit(-10);
}
static double wtime(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* Checkpoint file: this header, then the stored elements (plb_read_state),
 * AoS [N][19] in V order, elem_bytes each. */
#define CKPT_MAGIC 0x32424c50u /* "PLB2" */
typedef struct {
    uint32_t magic;
    int32_t precision;       /* PLB_FP64 | PLB_FP32 | PLB_FP16 */
    int32_t elem_bytes;      /* 8, 4 or 2 */
    int32_t fp16_scale_log2; /* FP16 storage scale 2^k, else 0 */
    int64_t step;            /* the next step to run */
    int64_t n_nodes;
    double rho0;             /* shift of FP32/FP16 storage */
} ckpt_header;

/* save_data (porelb.c:375-391) for every rank: the macros of step `it` go
 * to one host buffer, and its part files are written either at once or by a
 * worker thread while the GPU runs the next steps.  Two buffers alternate; a
 * buffer is reused only after its writer has finished. */
typedef struct {
    const plb_vfile *v;
    double *buf;
    int it;
    int failed;
    int running;
    pthread_t th;
} plb_saver;

static void *save_worker(void *arg)
{
    plb_saver *s = (plb_saver *)arg;
    const plb_vfile *v = s->v;
    const int64_t N = v->num_node;
    const double *buf = s->buf;
    for (int p = 0; p < v->num_p && !s->failed; p++) {
        int64_t b = v->V_sub_start[p];
        int64_t e = p + 1 < v->num_p ? v->V_sub_start[p + 1] : N;
        if (plb_write_part(p, s->it, e - b, buf + b, buf + N + b, buf + 2 * N + b, buf + 3 * N + b))
            s->failed = 1;
    }
    return NULL;
}

static void saver_wait(plb_saver *s)
{
    if (s->running) {
        pthread_join(s->th, NULL);
        s->running = 0;
    }
    if (s->failed) {
        fprintf(stderr, "Macro binary file open error!\n");
        exit(-3);
    }
}

static void save_all(plb_ctx *c, plb_saver sv[2], int *k, int async, int it)
{
    plb_saver *s = &sv[*k];
    saver_wait(s);
    const int64_t N = s->v->num_node;
    if (plb_read_macros(c, s->buf, s->buf + N, s->buf + 2 * N, s->buf + 3 * N)) {
        fprintf(stderr, "Macro read error!\n");
        exit(-3);
    }
    s->it = it;
    s->failed = 0;
    if (async && pthread_create(&s->th, NULL, save_worker, s) == 0) {
        s->running = 1;
    } else {
        save_worker(s);
        saver_wait(s);
    }
    if (async) *k ^= 1;
}

int main(int argc, char *argv[])
{
    int rank = env_int("PMI_RANK", env_int("OMPI_COMM_WORLD_RANK", 0));
    int size = env_int("PMI_SIZE", env_int("OMPI_COMM_WORLD_SIZE", -1));
    if (argc != 2) {
        fprintf(stderr, ">>> Error argc, usage ./porelb.x it_max\n");
        exit(-10);
    }
    int it_max = atoi(argv[1]);

    /* The header first (porelb.c:24-46): the rank-count check and the exit of
     * ranks other than 0 come before the records are read. */
    plb_vfile v;
    int rc = plb_read_vfile_header("V.bin", &v);
    if (rc == -1) {
        fprintf(stderr, "Node map file openning error!\n");
        exit(-1);
    }
    if (rc) {
        fprintf(stderr, "%d >>> Error read V \n", rank);
        exit(rc);
    }
    if (rank == 0) printf("[0]>>> Total num of fluid  nodes = %d\n", v.num_node);
    if (size > 0 && size != v.num_p) {
        if (rank == 0)
            fprintf(stderr, "[0]>>> MPI process num %d differs from num_p %d in V\n", size, v.num_p);
        exit(-2);
    }
    if (rank != 0) return 0;
    rc = plb_read_vfile_records("V.bin", &v);
    if (rc) {
        fprintf(stderr, "%d >>> Error read V \n", rank);
        exit(rc);
    }
    for (int p = 0; p < v.num_p; p++) {
        int64_t e = p + 1 < v.num_p ? v.V_sub_start[p + 1] : v.num_node;
        fprintf(stderr, "%d >>> My num of fluid nodes: %d \n", p, (int)(e - v.V_sub_start[p]));
    }

    plb_config cfg;
    plb_config_default(&cfg);
    cfg.tau = env_dbl("PLB_TAU", cfg.tau);
    cfg.gx = env_dbl("PLB_GX", cfg.gx);
    cfg.rho0 = env_dbl("PLB_RHO0", cfg.rho0);
    cfg.precision = env_choice("PLB_PRECISION", "fp64", "fp32", "fp16", PLB_FP64);
    cfg.bb_mode = env_choice("PLB_BB", "refcompat", "halfway", NULL, PLB_BB_REFCOMPAT);
    /* Default shape: the one measured fastest across the configurations
     * (AoSoA, pencils of 64 full rows, TILE8; 128 threads on small lattices).
     * Every shape gives bit-identical results. */
    cfg.layout = env_choice("PLB_LAYOUT", "soa", "aosoa", NULL, PLB_LAYOUT_AOSOA);
    cfg.ordering = env_choice("PLB_ORDER", "reference", "brick", "pencil", PLB_ORDER_PENCIL);
    if (cfg.ordering == PLB_ORDER_PENCIL) {
        cfg.brick[0] = v.nx;
        cfg.brick[1] = 64;
        cfg.brick[2] = 1;
    }
    {
        const char *b = getenv("PLB_BRICK"); /* "bx,by,bz" */
        if (b && *b && sscanf(b, "%d,%d,%d", &cfg.brick[0], &cfg.brick[1], &cfg.brick[2]) != 3) {
            fprintf(stderr, ">>> bad value PLB_BRICK=%s\n", b);
            exit(-10);
        }
    }
    cfg.index_kind = env_choice("PLB_INDEX", "i32", "tile8", "bitmap", PLB_INDEX_TILE8);
    cfg.block_size = env_int("PLB_BLOCK", v.num_node < (1 << 22) ? 128 : 256);
    cfg.graph_steps = env_int("PLB_GRAPH", cfg.graph_steps);
    /* At most 5 step-kernel waves per SIMD on large FP64/FP32 lattices (measured
     * faster there; FP16 runs best uncapped). */
    cfg.waves_per_simd =
        env_int("PLB_WAVES", cfg.precision != PLB_FP16 && v.num_node >= (1 << 22) ? 5 : 0);
    cfg.fp16_rounding = env_choice("PLB_FP16_ROUND", "nearest", "stochastic", NULL, PLB_ROUND_STOCHASTIC);
    cfg.placement_reserve_gib = env_int("PLB_RESERVE_GIB", 0);
    /* one population buffer (AA pattern): needs PLB_BB=halfway */
    cfg.stream_mode = env_choice("PLB_STREAM", "pull", "aa", NULL, PLB_STREAM_PULL);
    cfg.device = env_int("PLB_DEVICE", 0);
    const int save_every = env_int("PLB_SAVE", 1000); /* porelb.c:93 */
    const int print = env_int("PLB_PRINT", 1);
    if (save_every < 1) {
        fprintf(stderr, ">>> PLB_SAVE must be >= 1\n");
        exit(-10);
    }
    const double tol = env_dbl("PLB_TOL", 0.0);
    const int check_every = env_int("PLB_CHECK", 1000);
    if (!(tol >= 0.0) || check_every < 1) {
        fprintf(stderr, ">>> PLB_TOL must be >= 0 and PLB_CHECK >= 1\n");
        exit(-10);
    }

    const char *tune = getenv("PLB_AUTOTUNE");
    if (tune && *tune) {
        /* The autotuner works on the voxel mask; V.bin's coordinates give it. */
        const int64_t nvox = (int64_t)v.nx * v.ny * v.nz;
        uint8_t *mask = (uint8_t *)calloc((size_t)nvox, 1);
        if (!mask) {
            fprintf(stderr, ">>> out of memory for the autotuner mask\n");
            exit(-4);
        }
        for (int64_t k = 0; k < v.num_node; k++) {
            const plb_node_info *r = &v.V[k];
            if (r->x < v.nx && r->y < v.ny && r->z < v.nz)
                mask[((int64_t)r->z * v.ny + r->y) * v.nx + r->x] = 1;
        }
        plb_config best;
        rc = plb_autotune(&cfg, mask, v.nx, v.ny, v.nz, v.num_p, 20, tune, &best, NULL, 0, NULL);
        free(mask);
        if (rc) {
            fprintf(stderr, ">>> autotune: %s\n", plb_strerror(rc));
            exit(rc);
        }
        cfg = best;
        printf(">>> autotune: index %d layout %d order %d brick %d,%d,%d block %d nodes/thread %d "
               "waves/SIMD %d\n",
               cfg.index_kind, cfg.layout, cfg.ordering, cfg.brick[0], cfg.brick[1], cfg.brick[2],
               cfg.block_size, cfg.nodes_per_thread, cfg.waves_per_simd);
    }

    plb_ctx *c = NULL;
    rc = plb_create(&c, &cfg, v.V, v.num_node, v.V_sub_start, v.num_p, v.nx, v.ny, v.nz);
    if (rc == PLB_ERR_BOUNDS) {
        fprintf(stderr, "%d >>> Check V, bound error!\n", 0);
        exit(-3);
    }
    if (rc == PLB_ERR_PID) {
        fprintf(stderr, "%d >>> Check V, pe error!\n", 0);
        exit(-4);
    }
    if (rc) {
        fprintf(stderr, ">>> libplb: %s\n", plb_strerror(rc));
        exit(rc);
    }
    for (int p = 0; p < v.num_p; p++) printf("%d >>> Check V past\n", p);

    const int64_t N = v.num_node;
    int it0 = 0;
    const char *restart = getenv("PLB_RESTART");
    if (restart && *restart) {
        plb_state_info si;
        plb_get_state_info(c, &si);
        ckpt_header h;
        void *f = plb_host_alloc((size_t)si.elem_bytes * 19 * (size_t)(N ? N : 1));
        FILE *fp = fopen(restart, "rb");
        if (!f || !fp || fread(&h, sizeof h, 1, fp) != 1 || h.magic != CKPT_MAGIC || h.n_nodes != N) {
            fprintf(stderr, ">>> cannot read checkpoint %s\n", restart);
            exit(-5);
        }
        /* The stored elements are only meaningful in the storage they were
         * written from: same precision, FP16 scale and shift. */
        if (h.precision != si.precision || h.elem_bytes != si.elem_bytes ||
            h.fp16_scale_log2 != si.fp16_scale_log2 || h.rho0 != si.rho0) {
            fprintf(stderr,
                    ">>> checkpoint %s holds precision %d (scale 2^%d, rho0 %g), this run %d (2^%d, %g)\n",
                    restart, h.precision, h.fp16_scale_log2, h.rho0, si.precision, si.fp16_scale_log2,
                    si.rho0);
            exit(-5);
        }
        if (fread(f, (size_t)si.elem_bytes * 19, (size_t)N, fp) != (size_t)N) {
            fprintf(stderr, ">>> cannot read checkpoint %s\n", restart);
            exit(-5);
        }
        fclose(fp);
        it0 = (int)h.step;
        if (plb_write_state(c, f)) exit(-5);
        plb_host_free(f);
    }

    const int async = env_int("PLB_ASYNC_SAVE", 1);
    plb_saver sv[2];
    memset(sv, 0, sizeof sv);
    for (int b = 0; b < (async ? 2 : 1); b++) {
        sv[b].v = &v;
        /* page-locked, so the read-back runs at the PCIe rate */
        sv[b].buf = (double *)plb_host_alloc(sizeof(double) * 4 * (size_t)(N ? N : 1));
        if (!sv[b].buf) {
            fprintf(stderr, "Macro memory allocation error!\n");
            exit(-4);
        }
    }
    int kbuf = 0;
    double t1 = wtime();
    int it = it0;
    /* convergence checks (PLB_TOL > 0) every check_every steps from it0 */
    long long next_check = tol > 0.0 ? (long long)it0 + check_every : -1;
    double prev_ux = NAN;
    int converged = 0;
    while (it < it_max && !converged) {
        /* run up to and including the next save step (or to the end) */
        int next_save = ((it + save_every - 1) / save_every) * save_every;
        int last = next_save < it_max ? next_save : it_max - 1;
        if (next_check > 0 && next_check - 1 < last) last = (int)(next_check - 1);
        rc = plb_step(c, it, (int64_t)last - it + 1);
        if (rc == PLB_ERR_DENSITY) {
            plb_fault ft;
            plb_last_fault(c, &ft);
            if (print)
                for (int s = it; s < ft.step; s++) printf(">>> Step : %20d \n", s);
            fprintf(stderr, "%d >>> at %d (%d, %d, %d),  Macro denstity  = %g , < 0.5 !\n", ft.pid,
                    (int)ft.local_id, ft.x, ft.y, ft.z, ft.rho);
            saver_wait(&sv[0]);
            saver_wait(&sv[1]);
            exit(-6);
        }
        if (rc) {
            fprintf(stderr, ">>> libplb: %s\n", plb_strerror(rc));
            saver_wait(&sv[0]);
            saver_wait(&sv[1]);
            exit(rc);
        }
        if (print)
            for (int s = it; s <= last; s++) printf(">>> Step : %20d \n", s);
        if (last % save_every == 0) save_all(c, sv, &kbuf, async, last);
        it = last + 1;
        if (next_check > 0 && it == next_check) {
            plb_reduction r;
            rc = plb_reduce(c, &r);
            if (rc) {
                fprintf(stderr, ">>> libplb: %s\n", plb_strerror(rc));
                saver_wait(&sv[0]);
                saver_wait(&sv[1]);
                exit(rc);
            }
            const double scale = fabs(r.mean_ux) > 1e-3 * fabs(cfg.gx) ? fabs(r.mean_ux) : 1e-3 * fabs(cfg.gx);
            if (!isnan(prev_ux) && fabs(r.mean_ux - prev_ux) <= tol * scale) {
                converged = 1;
                printf(">>> Converged at step %d: <u_x> = %.10e, K = %.10e\n", it, r.mean_ux, r.K);
            }
            prev_ux = r.mean_ux;
            next_check += check_every;
        }
    }
    double t2 = wtime();
    printf("Elapsed time is %f\n", t2 - t1);
    if (t2 > t1 && it > it0)
        printf("MLUPS %.3f (%lld pore nodes x %d steps)\n",
               (double)N * (it - it0) / (t2 - t1) / 1e6, (long long)N, it - it0);
    save_all(c, sv, &kbuf, async, it);
    saver_wait(&sv[0]);
    saver_wait(&sv[1]);

    const char *ckpt = getenv("PLB_CHECKPOINT");
    if (ckpt && *ckpt) {
        plb_state_info si;
        plb_get_state_info(c, &si);
        ckpt_header h;
        memset(&h, 0, sizeof h);
        h.magic = CKPT_MAGIC;
        h.precision = si.precision;
        h.elem_bytes = si.elem_bytes;
        h.fp16_scale_log2 = si.fp16_scale_log2;
        h.step = it; /* the step reached: it_max, or earlier on convergence */
        h.n_nodes = N;
        h.rho0 = si.rho0;
        void *f = plb_host_alloc((size_t)si.elem_bytes * 19 * (size_t)(N ? N : 1));
        FILE *fp = fopen(ckpt, "wb");
        if (!f || !fp || plb_read_state(c, f)) {
            fprintf(stderr, ">>> cannot write checkpoint %s\n", ckpt);
            exit(-3);
        }
        int bad = fwrite(&h, sizeof h, 1, fp) != 1 ||
                  fwrite(f, (size_t)si.elem_bytes * 19, (size_t)N, fp) != (size_t)N;
        bad |= fclose(fp) != 0;
        plb_host_free(f);
        if (bad) {
            fprintf(stderr, ">>> cannot write checkpoint %s\n", ckpt);
            exit(-3);
        }
    }
    plb_host_free(sv[0].buf);
    plb_host_free(sv[1].buf);
    plb_destroy(c);
    plb_free_vfile(&v);
    return 0;
}
