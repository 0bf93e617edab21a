/*
 * plb_oracle.h -- CPU oracle for porelb's sparse D3Q19 time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * solver, used as the checker by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py.  Nothing in the product (libplb, the CLIs)
 * links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  How each one is
 * pinned (reference binary built from its own sources, golden fixtures,
 * closed forms, brute force) is stated beside it.
 */
#ifndef PLB_ORACLE_H
#define PLB_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_Q 19

/* NEIGHBOR_TYPE / NEIGHBOR_INFO / NODE_INFO (porelb.h:55-76).  The layout is
 * the V.bin record layout: 8-byte neighbour entries, 152-byte node records. */
enum { ORC_FLUID = 0, ORC_SOLID = 1, ORC_EXTERN = 2 };
typedef struct {
    uint8_t node_type;
    uint16_t node_pid;
    uint32_t node_id;
} orc_nb_t;
typedef struct {
    uint16_t x, y, z;
    orc_nb_t nb[ORC_Q - 1];
} orc_node_t;

/* Pull-table entry codes (a source node id is >= 0). */
#define ORC_TAB_BOUNCE (-1)
#define ORC_TAB_CONST (-2)
enum { ORC_BB_REFCOMPAT = 0, ORC_BB_HALFWAY = 1 };

/* Lattice tables in the reference order (porelb.h:10-50). */
extern const int orc_e[ORC_Q][3];
extern const double orc_tp[ORC_Q];
extern const int orc_re[ORC_Q];

/* Number of voxels with flag==1 (preprocess.c:74-82). */
int64_t orc_count_pores(const uint8_t *flag, int nx, int ny, int nz);

/* build_V (preprocess.c:29-164).  V must hold orc_count_pores() records,
 * V_sub_start num_p entries.  Fields the reference leaves uninitialised
 * (SOLID pid/id, padding) are written as zero.  Returns 0, or -2 when the
 * z-slab rule pid = z/(nz/num_p) would leave [0,num_p) (preprocess.c:107),
 * -3 for a flag value other than 0/1 (preprocess.c:78 and :143 disagree on
 * those). */
int orc_build_V(const uint8_t *flag, int nx, int ny, int nz, int num_p,
                orc_node_t *V, int32_t *V_sub_start);

/* Pull table [18][N] (entry for direction j at tab[(j-1)*N + g]) derived from
 * V by the rule that reproduces stream_inner_nodes + update_recv_dist
 * (porelb.c:285-312) in ref-compat mode, or textbook half-way bounce-back. */
int orc_pull_table(const orc_node_t *V, int64_t N, const int32_t *V_sub_start,
                   int num_p, int mode, int32_t *tab);

/* init_dist_macro (porelb.c:224-238): f0 = f1 = feq(rho0, 0). AoS [N][19]. */
void orc_init_equilibrium(int64_t N, double rho0, double *f0, double *f1);

/* Sequential push restatement of one or more time steps of porelb.c main
 * (porelb.c:426-437) with num_p emulated ranks: stream_inner_nodes per rank,
 * then every rank's update_recv_dist, then collision.  f0/f1 are AoS [N][19]
 * in V order and carry state between calls.  macros (may be NULL) receives
 * Rho,Ux,Uy,Uz [4][N] of the last step.  mode ORC_BB_HALFWAY swaps in a
 * textbook push-form half-way bounce-back.  Returns 0, or -6 for the density
 * guard (porelb.c:338-342) with *fault_node = first offending node. */
int orc_push_run(const orc_node_t *V, int64_t N, const int32_t *V_sub_start,
                 int num_p, int mode, double tau, double gx, int64_t it0,
                 int64_t nsteps, double *f0, double *f1, double *macros,
                 int64_t *fault_node, int64_t *fault_step);

/* The same steps through a pull table: f_in(i,j) from tab, then collision.
 * f0 AoS [N][19] in/out; scratch AoS [N][19]. */
int orc_pull_run(const int32_t *tab, int64_t N, double tau, double gx,
                 double rho0, int64_t it0, int64_t nsteps, double *f0,
                 double *scratch, double *macros, int64_t *fault_node,
                 int64_t *fault_step);

/* FP64 reductions over a population state and the Ux macro.
 * out[0] = mass = sum_i sum_j f0[i][j]
 * out[1] = sum_i Ux[i]
 * out[2] = <u_x>_pore = out[1]/N
 * out[3] = superficial q = out[1]/(nx ny nz)
 * out[4] = K = nu q / gx, nu = (tau - 1/2)/3 (porelb.c:89) */
void orc_reduce(const double *f0, const double *ux, int64_t N, int nx, int ny,
                int nz, double tau, double gx, double *out);

/* getK.c:39-72: analytic permeability of a BCC sphere array. out[0]=K1
 * (Sangani-Acrivos drag series), out[1]=K2 (Kozeny-Carman, C=5),
 * out[2]=porosity epsilon, out[3]=sphere radius a. */
void orc_getK_bcc(double chi, int nx, double *out);

#ifdef __cplusplus
}
#endif
#endif
