/*
 * plb.h -- C ABI of libplb, the MI355X (gfx950) implementation of porelb's
 * sparse-lattice D3Q19 time step.
 *
 * The reference (zhulianhua/porelb) has no library API: its solver is a set
 * of `void f(SIMULATION*)` functions called from porelb.c main
 * (porelb.h:127-165, porelb.c:402-443), its sparse index build is
 * preprocess.c build_V/save_V (preprocess.c:29-193), and its halo exchange is
 * MPI point-to-point (porelb.c:143-312).  Each entry point below names the
 * reference code it replaces.  Plain C types only: pointers, sizes, codes.
 *
 * Ownership: the caller owns every host buffer; the library owns all device
 * memory of a context.  Threading: one host thread per context; each context
 * runs on its own HIP stream.  Errors: negative return codes (plb_strerror);
 * the library never exits the process.
 */
#ifndef PLB_H
#define PLB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PLB_Q 19
#define PLB_ABI_VERSION 1

/* ---- error codes (negative).  Where the reference exits with a code, the
 * same value is used (porelb.c:18-22, 29-34, 120-137, 338-342, 406-410). */
#define PLB_OK 0
#define PLB_ERR_FILE (-1)      /* input file missing (porelb.c:18-22)          */
#define PLB_ERR_NUM_P (-2)     /* rank-count / z-slab mismatch (porelb.c:29-34) */
#define PLB_ERR_BOUNDS (-3)    /* node coordinate out of range (porelb.c:120-126) */
#define PLB_ERR_PID (-4)       /* EXTERN neighbour pid invalid (porelb.c:128-137) */
#define PLB_ERR_READ (-5)      /* short read (porelb.c:65-69)                  */
#define PLB_ERR_DENSITY (-6)   /* density guard rho>1.2|rho<0.5 (porelb.c:338-342) */
#define PLB_ERR_ARGS (-10)     /* bad command line (porelb.c:406-410)          */
#define PLB_ERR_INVALID (-20)  /* invalid argument or configuration            */
#define PLB_ERR_HIP (-21)      /* HIP runtime error                            */
#define PLB_ERR_NOMEM (-22)    /* device or host allocation failed             */
#define PLB_ERR_RANGE (-23)    /* node count beyond int32 node ids             */
#define PLB_ERR_NODEVICE (-24) /* no usable GPU                                */
#define PLB_ERR_FLAG (-25)     /* voxel flag outside {0,1}                     */

/* ---- V.bin record (porelb.h:55-76; preprocess.h:10-22).  8-byte neighbour
 * entries, 152-byte node records, little-endian; nb[j-1] describes the
 * periodic neighbour at x+e_j. */
enum { PLB_TYPE_FLUID = 0, PLB_TYPE_SOLID = 1, PLB_TYPE_EXTERN = 2 };
typedef struct {
    uint8_t node_type;
    uint16_t node_pid;
    uint32_t node_id;
} plb_neighbor_info;
typedef struct {
    uint16_t x, y, z;
    plb_neighbor_info nb_info[PLB_Q - 1];
} plb_node_info;

/* ---- configuration */
enum { PLB_FP64 = 0, PLB_FP32 = 1, PLB_FP16 = 2 };
/* Bounce-back: REFCOMPAT reproduces the reference's stream/bounce/receive
 * overwrite order (porelb.c:285-312) exactly; HALFWAY is textbook half-way
 * bounce-back (mass conserving). */
enum { PLB_BB_REFCOMPAT = 0, PLB_BB_HALFWAY = 1 };
/* Device layout of populations and pull table: SOA = [dir][node];
 * AOSOA = tiles of 64 nodes, [tile][dir][64]. */
enum { PLB_LAYOUT_SOA = 0, PLB_LAYOUT_AOSOA = 1 };
/* Device node order: REFERENCE = V order (z,y,x lexicographic per rank);
 * BRICK = bricks of brick[0] x brick[1] x brick[2] voxels in lexicographic
 * brick order (x fastest), z,y,x inside a brick; PENCIL = the same bricks
 * visited z fastest, then y, then x, so with bricks of (nx, Py, 1) the nodes
 * run through pencils of Py rows plane by plane and a node's z-neighbours are
 * Py rows (not a whole plane) away.  Outputs are always in V order. */
enum { PLB_ORDER_REFERENCE = 0, PLB_ORDER_BRICK = 1, PLB_ORDER_PENCIL = 2 };
/* Pull-table encoding on the device: I32 = one int32 per (node, direction);
 * TILE8 = per 64-node tile and direction one int32 base plus a u8 offset per
 * node (254 = constant, 255 = bounce), with int32 entries for the rare
 * (tile, direction) pairs whose sources span more than 253 ids;
 * BITMAP = no table: one pore bit per voxel in node order with the pore count
 * before every 32-voxel word, plus each node's voxel position, from which the
 * step kernel derives every source (half-way bounce-back only; node order
 * REFERENCE, or PENCIL with brick[0] >= nx and brick[2] == 1). */
enum { PLB_INDEX_I32 = 0, PLB_INDEX_TILE8 = 1, PLB_INDEX_BITMAP = 2 };

typedef struct {
    double tau;          /* relaxation time (reference: 0.8, porelb.c:87)         */
    double gx;           /* body force along x (reference: 1e-3, porelb.c:91)     */
    double rho0;         /* initial density (reference: 1, porelb.c:228)          */
    int precision;       /* PLB_FP64 | PLB_FP32 | PLB_FP16                        */
    int bb_mode;         /* PLB_BB_REFCOMPAT | PLB_BB_HALFWAY                     */
    int layout;          /* PLB_LAYOUT_SOA | PLB_LAYOUT_AOSOA                     */
    int ordering;        /* PLB_ORDER_REFERENCE | PLB_ORDER_BRICK | PLB_ORDER_PENCIL */
    int brick[3];        /* brick edge lengths x,y,z for PLB_ORDER_BRICK / PENCIL */
    int block_size;      /* threads per block of the step kernel (64..1024)       */
    int device;          /* HIP device ordinal                                    */
    int fp16_scale_log2; /* FP16: stored = (f - w rho0) * 2^k; 0 = derive from gx */
    int index_kind;      /* PLB_INDEX_I32 | PLB_INDEX_TILE8 | PLB_INDEX_BITMAP    */
    int nodes_per_thread; /* 1, or 2: a wave advances two 64-node tiles at once
                             (block_size <= 512; single lattices only)          */
    int graph_steps;     /* 0: one kernel launch per step; G > 0 (even, <= 4096):
                            runs of G steps replay one captured HIP graph
                            (single lattices only; bit-identical results)      */
    int waves_per_simd;  /* 0 (or 8): as many resident waves as registers
                            allow; 1..7: about this many step-kernel waves per
                            SIMD.  Each block reserves share = LDS per CU /
                            max(1, 4*w*64/block_size) bytes of LDS, rounded
                            down to 512 and at most 64 KB, so a CU holds
                            floor(LDS per CU / share) blocks (fewer concurrent
                            streams per CU; results unchanged)                 */
    int fp16_rounding;   /* FP16 storage: PLB_ROUND_NEAREST (round to nearest
                            even) or PLB_ROUND_STOCHASTIC (round up with
                            probability equal to the fraction of an FP16 ulp,
                            from a hash of node id and step: unbiased, so
                            increments below half an ulp are not lost)        */
    int placement_reserve_gib; /* 0..160: GiB of device memory the context allocates
                            before any of its own buffers and holds until
                            plb_destroy, so its buffers are placed above it.
                            On a freshly booted MI355X the first ~40 GiB the
                            first process allocates stream about 6 % slower
                            for as long as they live; 40 places a lattice
                            clear of them.  0: no reserve                    */
    int stream_mode;     /* PLB_STREAM_PULL: two population buffers, every
                            step pulls through the table.  PLB_STREAM_AA: one
                            buffer, AA pattern -- even steps update each node
                            in place (no table read), odd steps gather from
                            and scatter to the neighbours; the same results,
                            half the population memory.  AA needs half-way
                            bounce-back, nodes_per_thread 1, graph_steps 0
                            and a single lattice (not a batch)               */
} plb_config;
enum { PLB_ROUND_NEAREST = 0, PLB_ROUND_STOCHASTIC = 1 };
enum { PLB_STREAM_PULL = 0, PLB_STREAM_AA = 1 };

/* Fills the reference's parameters (porelb.c:83-94) and FP64/REFCOMPAT/SOA/
 * REFERENCE order/I32 index, 256 threads, one node per thread, device 0. */
void plb_config_default(plb_config *cfg);

typedef struct plb_ctx plb_ctx;

/* ---- context creation.  Replaces alloc_memory + check_V +
 * build_send_recv_info (porelb.c:15-81, 115-222): every rank's records of a
 * V.bin are taken at once; links to other ranks (EXTERN) become global node
 * ids in the pull table, so no halo pack/unpack exists on one GPU.  The
 * checks of check_V are applied to all records (-3 / -4). */
int plb_create(plb_ctx **out, const plb_config *cfg, const plb_node_info *V, int64_t n_nodes,
               const int32_t *V_sub_start, int num_p, int nx, int ny, int nz);

/* Same, straight from a voxel mask (u8 [nz][ny][nx], 1 = pore, 0 = solid,
 * x fastest): the sparse index is built on the device (preprocess.c:29-164
 * semantics with num_p z-slabs) without V records. */
int plb_create_from_flag(plb_ctx **out, const plb_config *cfg, const uint8_t *flag, int nx,
                         int ny, int nz, int num_p);

void plb_destroy(plb_ctx *c);

/* init_dist_macro (porelb.c:224-238): rho = rho0, u = 0, f = feq. */
int plb_init_equilibrium(plb_ctx *c);

/* nsteps time steps it0 .. it0+nsteps-1 of porelb.c main (porelb.c:426-437):
 * stream + bounce-back + receive + collision fused in one kernel per step.
 * `it` enters only through the reference's half-force shift for it > 0
 * (porelb.c:347-348).  Macros are stored on the last step.  Blocks until the
 * steps finish; returns PLB_ERR_DENSITY if the density guard tripped (the
 * state is then frozen at the end of the faulting step). */
int plb_step(plb_ctx *c, int64_t it0, int64_t nsteps);

/* As plb_step, but enqueues on the context stream and returns at once;
 * plb_sync waits and reports the density guard. */
int plb_step_async(plb_ctx *c, int64_t it0, int64_t nsteps);
int plb_sync(plb_ctx *c);

/* Same as plb_step with HIP events recorded on the context stream around the
 * launches; *ms = elapsed device time of the nsteps steps. */
int plb_step_timed(plb_ctx *c, int64_t it0, int64_t nsteps, double *ms);

/* Macros of the last stored step, V order (porelb.c:350-356; save_data
 * writes them, porelb.c:375-391). Any pointer may be NULL. */
int plb_read_macros(plb_ctx *c, double *rho, double *ux, double *uy, double *uz);

/* Populations f0 after the last step, AoS [n][19] in V order, as FP64
 * (shifted storage is un-shifted). */
int plb_read_populations(plb_ctx *c, double *f_aos);

/* Overwrites the populations (AoS [n][19], V order) and clears the density
 * guard; used by tests and to start from a given FP64 state.  With FP32/FP16
 * storage the values are shifted and rounded to the stored type, so an
 * FP64 round trip through plb_read_populations is exact for FP64 storage
 * only: checkpoints use plb_read_state / plb_write_state. */
int plb_write_populations(plb_ctx *c, const double *f_aos);

/* Stored population state, bit for bit: checkpoint/resume (new; fills the
 * empty save_state/recovery_state stubs of porelb.c:367-373).  The elements
 * the step kernel keeps, AoS [n][19] in V order: FP64 populations (8 B),
 * FP32 deviations f - w_j rho0 (4 B), or the FP16 bits of
 * (f - w_j rho0) 2^k (2 B).  Writing back what plb_read_state gave, into a
 * context of the same mask and precision, continues the run bit for bit.
 * plb_write_state also clears the density guard. */
typedef struct {
    int precision;       /* PLB_FP64 | PLB_FP32 | PLB_FP16 */
    int elem_bytes;      /* 8, 4 or 2 bytes per stored population */
    int fp16_scale_log2; /* k of the FP16 storage scale 2^k (0 otherwise) */
    double rho0;         /* the shift w_j rho0 of FP32/FP16 storage */
    int64_t n_nodes;     /* the state holds n_nodes * 19 elements */
} plb_state_info;
int plb_get_state_info(const plb_ctx *c, plb_state_info *out);
int plb_read_state(plb_ctx *c, void *buf);
int plb_write_state(plb_ctx *c, const void *buf);

/* The pull table [18][n] in V order with V-order node ids; entry codes:
 * >= 0 source node (pull direction j from it), -1 bounce (own node,
 * direction re[j]), -2 constant tp[j]*rho0. */
int plb_read_table(plb_ctx *c, int32_t *tab);

int64_t plb_num_nodes(const plb_ctx *c);

/* FP64 reductions (new: the reference has none; getK.c is an analytic
 * calculator).  mass = sum of all populations; sum_ux = sum of the Ux macro;
 * mean_ux = sum_ux / n; q = sum_ux / (nx ny nz); K = nu q / gx with
 * nu = (tau - 1/2)/3 (porelb.c:89). */
typedef struct {
    double mass;
    double sum_ux;
    double mean_ux;
    double q;
    double K;
    int64_t n_nodes;
} plb_reduction;
int plb_reduce(plb_ctx *c, plb_reduction *out);
/* Steady-state run (new; the reference always runs it_max steps): steps from
 * it0 in runs of check_every, and after each run compares <u_x> with the one
 * check_every steps before; stops when |d<u_x>| <= tol * max(|<u_x>|,
 * 1e-3 |gx|) (the rule of plb_batch_run) or after max_steps steps.
 * *steps_done = steps taken, *out = the reductions of the last step (whose
 * macros are stored).  tol = 0 runs all max_steps.  A density-guard fault
 * returns PLB_ERR_DENSITY as plb_step does. */
int plb_run_to_convergence(plb_ctx *c, int64_t it0, int64_t max_steps, int64_t check_every, double tol,
                           int64_t *steps_done, plb_reduction *out);

/* Density-guard record: first step with a violation, and at that step the
 * lowest V-order node, its rank pid and local id, coordinates and rho
 * (the reference's message, porelb.c:340). */
typedef struct {
    int tripped;
    int64_t step;
    int64_t node;
    int pid;
    int64_t local_id;
    int x, y, z;
    double rho;
} plb_fault;
int plb_last_fault(plb_ctx *c, plb_fault *out);

/* Bytes per node update the step kernel moves as algorithmic traffic:
 * populations read + written (2*19*element size) plus the pull-table bytes
 * of the context's encoding per node (and, for FP16 storage with stochastic
 * rounding in a node order other than V order, the 4-byte rounding key). */
double plb_bytes_per_update(const plb_ctx *c);

const char *plb_strerror(int code);

/* Page-locked host memory (new; the reference writes from malloc'd arrays,
 * porelb.c:375-391).  plb_read_macros / plb_read_populations copy into such a
 * buffer by DMA at the full PCIe rate instead of through a staging copy.
 * Returns NULL on failure; release with plb_host_free. */
void *plb_host_alloc(size_t bytes);
void plb_host_free(void *p);

/* plb_read_macros for the V-order nodes [first, first + n) without any
 * device allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  The gather to V order is staged through
 * a fixed device array of PLB_NOALLOC_MAX_NODES x 4 doubles (32 MiB) that each
 * GPU the library runs on holds until the process ends; n beyond that
 * returns PLB_ERR_RANGE.  Calls that overlap on one GPU race on that array. */
#define PLB_NOALLOC_MAX_NODES (1 << 20)
int plb_read_macros_noalloc(plb_ctx *c, int64_t first, int64_t n, double *rho, double *ux, double *uy,
                            double *uz);

/* ---- preprocess on the device (preprocess.c:29-164 build_V): flag ->
 * V records (V order) and V_sub_start.  Call with V_out == NULL to get the
 * pore count only.  Fields preprocess.c leaves uninitialised (SOLID pid/id,
 * padding) are zero. */
int plb_preprocess(const uint8_t *flag, int nx, int ny, int nz, int num_p, int device,
                   plb_node_info *V_out, int32_t *V_sub_start_out, int64_t *n_nodes_out);

/* ---- batched samples: many independent porous samples of uneven pore
 * count advanced by one launch per step over all active samples, with
 * per-sample convergence stopping.  cfg supplies precision, bounce-back,
 * layout, index encoding (I32 or TILE8), node order (PENCIL uses each sample's own x extent
 * as the pencil width) and block size; tau and gx come per sample.  With FP16 storage and fp16_scale_log2 == 0 the scale is
 * derived from the largest |gx| of the samples. */
typedef struct {
    const uint8_t *flag; /* u8 [nz][ny][nx] */
    int nx, ny, nz;
    double tau, gx;
} plb_sample;
typedef struct plb_batch plb_batch;
int plb_batch_create(plb_batch **out, const plb_config *cfg, const plb_sample *samples,
                     int n_samples);
/* Runs until every sample has converged (|d<u_x>| <= tol * max(|<u_x>|,
 * 1e-3 |g_x|) between checks check_every steps apart; the floor stops samples
 * without a percolating path) or max_steps.  steps_done[s] = steps the
 * sample ran; updates[0] = total node updates performed; ms = device time. */
int plb_batch_run(plb_batch *b, int64_t max_steps, int64_t check_every, double tol,
                  int64_t *steps_done, int64_t *updates, double *ms);
/* The same run as one persistent launch: a resident grid (cooperative
 * launch) advances every sample through the same steps, checks and
 * convergence rule as plb_batch_run, with a grid barrier between steps and
 * the per-sample u_x sums and stop decisions made on the device (no host
 * round trip per check).  The per-sample sums use the same fixed order in
 * both drivers, so both stop every sample at the same step and leave the
 * same states.  The blocks of each XCD take the chunks of their share of
 * the batch one at a time from a device counter.  A grid that turns out not
 * to be all resident leaves before any work and is launched again with one
 * block fewer per CU; a barrier wait longer than 2 s ends the launch and
 * returns PLB_ERR_HIP.  ms covers the last launch.  At most
 * PLB_PERSIST_MAX_SAMPLES samples (PLB_ERR_INVALID beyond). */
#define PLB_PERSIST_MAX_SAMPLES 2048
int plb_batch_run_persistent(plb_batch *b, int64_t max_steps, int64_t check_every, double tol,
                             int64_t *steps_done, int64_t *updates, double *ms);
/* Per-sample reductions after a run (n_samples entries). */
int plb_batch_reduce(plb_batch *b, plb_reduction *out);
/* rho, u_x, u_y, u_z of one sample in its V order (z,y,x lexicographic, one
 * rank), plb_batch_num_nodes(b, sample) values each: the sample's last step
 * of the last run (rho0 and 0 before any run) -- the arrays save_data writes
 * (porelb.c:375-391), per sample.  Any pointer may be NULL. */
int plb_batch_read_macros(plb_batch *b, int sample, double *rho, double *ux, double *uy, double *uz);
int64_t plb_batch_num_nodes(const plb_batch *b, int sample);
/* After a run returned PLB_ERR_DENSITY: the sample whose node tripped the
 * density guard first (*sample) and the fault as plb_last_fault gives it,
 * with node and local_id the V-order id within that sample and pid 0.
 * *sample = -1 and out->tripped = 0 if no guard tripped. */
int plb_batch_last_fault(plb_batch *b, int *sample, plb_fault *out);
void plb_batch_destroy(plb_batch *b);

/* ---- synthetic voxel masks (inputs for tests and benchmarks; the reference
 * reads a given flag.bin, preprocess.c:61-69).  Deterministic from seed.
 * noise: i.i.d. N(0,1), periodic Gaussian blur sigma, the lowest
 * round(porosity * V) values are pore (exact porosity).
 * spheres: overlapping solid spheres, radius U[rmin,rmax], centres U[0,n)^3,
 * periodic, added until porosity <= target.
 * bcc: two solid spheres of radius sqrt(3) (n-1) chi / 4 per periodic cell of
 * n^3 voxels at the corner and the centre (the getK.c geometry). */
int plb_mask_noise(uint8_t *flag, int nx, int ny, int nz, double sigma, double porosity,
                   uint64_t seed);
int plb_mask_spheres(uint8_t *flag, int nx, int ny, int nz, double rmin, double rmax,
                     double porosity, uint64_t seed);
int plb_mask_bcc(uint8_t *flag, int n, double chi);

#ifdef __cplusplus
}
#endif
#endif
