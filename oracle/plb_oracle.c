/*
 * plb_oracle.c -- CPU oracle for porelb's sparse D3Q19 time step.
 *
 * TEST INFRASTRUCTURE ONLY (see plb_oracle.h).  Compiled with
 * -ffp-contract=off so that every FP64 expression rounds exactly as written,
 * in the same order as the reference's C source.
 *
 * Pinning:
 *  - orc_build_V      : V.bin files written by the reference preprocess.c
 *                       (built from its own sources into oracle/_ref) and the
 *                       committed fixtures in tests/golden; brute-force
 *                       invariants on tiny masks.
 *  - orc_push_run     : part files and population dumps written by the
 *                       reference porelb.c (oracle/_ref, fixtures), bitwise;
 *                       all-fluid closed form for the momentum.
 *  - orc_pull_table /
 *    orc_pull_run     : bitwise equality with orc_push_run for 1..4 emulated
 *                       ranks (ref-compat) and with the push-form half-way
 *                       bounce-back (halfway); mass conservation in halfway.
 *  - orc_reduce       : definition; mass closed form; K of a BCC array
 *                       against orc_getK_bcc within a few percent.
 *  - orc_getK_bcc     : hand-evaluated series values at chi = 0 and the
 *                       dilute limit K1 -> 2a^2/(9(1-eps)).
 */
#include "plb_oracle.h"

#include <stdlib.h>
#include <string.h>

#define Q ORC_Q

/* porelb.h:10-34 -- velocity set in the reference's order. */
const int orc_e[Q][3] = {
    {0, 0, 0},
    {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1},
    {0, 1, 1}, {0, -1, -1}, {0, -1, 1}, {0, 1, -1},
    {-1, 0, -1}, {1, 0, 1}, {-1, 0, 1}, {1, 0, -1},
    {-1, 1, 0}, {1, -1, 0}, {-1, -1, 0}, {1, 1, 0}};
/* porelb.h:36-42 */
const double orc_tp[Q] = {1.0 / 3,
                          1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18,
                          1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36,
                          1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36,
                          1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};
/* porelb.h:44-50 */
const int orc_re[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};

/* feq (porelb.c:6-13), same evaluation order. */
static double orc_feq(int k, double rho, double u, double v, double w)
{
    double uv, eu, x;
    eu = orc_e[k][0] * u + orc_e[k][1] * v + orc_e[k][2] * w;
    uv = u * u + v * v + w * w;
    x = orc_tp[k] * rho * (1.0 + 3 * eu + 4.5 * eu * eu - 1.5 * uv);
    return x;
}

int64_t orc_count_pores(const uint8_t *flag, int nx, int ny, int nz)
{
    int64_t n = 0, size = (int64_t)nx * ny * nz;
    for (int64_t k = 0; k < size; k++)
        n += (flag[k] == 1);
    return n;
}

/* build_V (preprocess.c:29-164).  Pass 1 numbers the pores in z,y,x order
 * (:74-82); pass 2 assigns the z-slab rank and the position within the rank
 * (:97-117); pass 3 classifies the 18 periodic neighbours at x+e_j
 * (:124-157). */
int orc_build_V(const uint8_t *flag, int nx, int ny, int nz, int num_p,
                orc_node_t *V, int32_t *V_sub_start)
{
    if (num_p < 1 || nz / num_p == 0) return -2;
    int64_t size = (int64_t)nx * ny * nz;
    int slab = nz / num_p;
    if ((nz - 1) / slab >= num_p) return -2; /* preprocess.c:107 overflows */

    for (int64_t k = 0; k < size; k++)
        if (flag[k] > 1) return -3; /* preprocess.c handles 0/1 only (:78, :143) */
    int64_t *x_to_n = (int64_t *)malloc(sizeof(int64_t) * (size_t)size);
    int64_t *count = (int64_t *)calloc((size_t)num_p, sizeof(int64_t));
    int64_t n = 0;
    for (int64_t k = 0; k < size; k++) {
        x_to_n[k] = -1;
        if (flag[k] == 1) x_to_n[k] = n++;
    }
    int32_t *node_pid = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
    int64_t *node_pos = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n ? n : 1));
    n = 0;
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                int64_t k = ((int64_t)z * ny + y) * nx + x;
                if (flag[k] != 1) continue;
                int p = z / slab;
                node_pid[n] = p;
                node_pos[n] = count[p]++;
                n++;
            }
    V_sub_start[0] = 0;
    for (int p = 1; p < num_p; p++) V_sub_start[p] = (int32_t)(V_sub_start[p - 1] + count[p - 1]);

    n = 0;
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                int64_t k = ((int64_t)z * ny + y) * nx + x;
                if (flag[k] != 1) continue;
                orc_node_t rec;
                memset(&rec, 0, sizeof rec);
                rec.x = (uint16_t)x;
                rec.y = (uint16_t)y;
                rec.z = (uint16_t)z;
                int p = node_pid[n];
                for (int i = 0; i < Q - 1; i++) {
                    int xp = (x + orc_e[i + 1][0] + nx) % nx;
                    int yp = (y + orc_e[i + 1][1] + ny) % ny;
                    int zp = (z + orc_e[i + 1][2] + nz) % nz;
                    int64_t kp = ((int64_t)zp * ny + yp) * nx + xp;
                    /* preprocess.c:143 tests flag==0 for SOLID; every other
                     * value is a pore for the neighbour classification. */
                    if (flag[kp] == 0) {
                        rec.nb[i].node_type = ORC_SOLID;
                    } else {
                        int64_t np_ = x_to_n[kp];
                        int pp = node_pid[np_];
                        rec.nb[i].node_id = (uint32_t)node_pos[np_];
                        rec.nb[i].node_pid = (uint16_t)pp;
                        rec.nb[i].node_type = (pp == p) ? ORC_FLUID : ORC_EXTERN;
                    }
                }
                V[V_sub_start[p] + node_pos[n]] = rec;
                n++;
            }
    free(x_to_n);
    free(count);
    free(node_pid);
    free(node_pos);
    return 0;
}

/* Rank that owns global record g: the last p with V_sub_start[p] <= g. */
static int owner_rank(const int32_t *V_sub_start, int num_p, int64_t g)
{
    int p = 0;
    for (int q = 0; q < num_p; q++)
        if (V_sub_start[q] <= g) p = q;
    return p;
}

static int64_t gid_of(const orc_nb_t *nb, int own_p, const int32_t *V_sub_start)
{
    int p = (nb->node_type == ORC_FLUID) ? own_p : nb->node_pid;
    return (int64_t)V_sub_start[p] + nb->node_id;
}

/* The pull form of stream_inner_nodes + update_recv_dist (porelb.c:285-312).
 * For slot (i,j), j>=1: dn = nb at x+e_j, up = nb at x-e_j = x+e_re[j].
 * up pushes f0[up][j] into (i,j) -- during the stream loop when up is FLUID
 * (local order decides against the bounce at i), after the loop when up is
 * EXTERN (the receive overrides).  i bounces f0[i][re[j]] into (i,j) when dn
 * is SOLID.  With no writer the slot keeps its initial value tp[j]*rho0. */
int orc_pull_table(const orc_node_t *V, int64_t N, const int32_t *V_sub_start,
                   int num_p, int mode, int32_t *tab)
{
    for (int64_t g = 0; g < N; g++) {
        int p = owner_rank(V_sub_start, num_p, g);
        int64_t k = g - V_sub_start[p];
        for (int j = 1; j < Q; j++) {
            const orc_nb_t *dn = &V[g].nb[j - 1];
            const orc_nb_t *up = &V[g].nb[orc_re[j] - 1];
            int32_t v;
            if (mode == ORC_BB_HALFWAY) {
                v = (up->node_type != ORC_SOLID) ? (int32_t)gid_of(up, p, V_sub_start)
                                                 : ORC_TAB_BOUNCE;
            } else {
                if (up->node_type != ORC_SOLID &&
                    (dn->node_type != ORC_SOLID || up->node_type == ORC_EXTERN ||
                     (int64_t)up->node_id > k))
                    v = (int32_t)gid_of(up, p, V_sub_start);
                else if (dn->node_type == ORC_SOLID)
                    v = ORC_TAB_BOUNCE;
                else
                    v = ORC_TAB_CONST;
            }
            tab[(int64_t)(j - 1) * N + g] = v;
        }
    }
    return 0;
}

void orc_init_equilibrium(int64_t N, double rho0, double *f0, double *f1)
{
    for (int64_t i = 0; i < N; i++)
        for (int j = 0; j < Q; j++) {
            double v = orc_feq(j, rho0, 0.0, 0.0, 0.0);
            f0[i * Q + j] = v;
            if (f1) f1[i * Q + j] = v;
        }
}

/* collision (porelb.c:318-365) for one node: f1 row in, f0 row out. */
static int collide_node(const double *fin, double *fout, double omega, double Xforce,
                        int64_t it, double *rho_o, double *u_o, double *v_o, double *w_o)
{
    double rho = 0.0, rhou = 0.0, rhov = 0.0, rhow = 0.0;
    for (int j = 0; j < Q; j++) {
        rho += fin[j];
        rhou += orc_e[j][0] * fin[j];
        rhov += orc_e[j][1] * fin[j];
        rhow += orc_e[j][2] * fin[j];
    }
    int bad = (rho > 1.2) | (rho < 0.5);
    if (it > 0) rhou += 0.5 * Xforce;
    double u = rhou / rho, v = rhov / rho, w = rhow / rho;
    *rho_o = rho;
    *u_o = u;
    *v_o = v;
    *w_o = w;
    for (int j = 0; j < Q; j++)
        fout[j] = (1.0 - omega) * fin[j] +
                  (omega + (1.0 - 0.5 * omega) * 3.0 * (orc_e[j][0] - u) * Xforce) *
                      orc_feq(j, rho, u, v, w);
    return bad;
}

/* Collision over all nodes; records the first density-guard violation of the
 * step.  The reference exits at that node (porelb.c:338-342); the oracle
 * finishes the step and reports it. */
static int collide_all(const double *f1, double *f0, int64_t N, double omega, double gx,
                       int64_t it, double *macros, int64_t *fault_node)
{
    int fault = 0;
    for (int64_t i = 0; i < N; i++) {
        double r, u, v, w;
        int bad = collide_node(f1 + i * Q, f0 + i * Q, omega, gx, it, &r, &u, &v, &w);
        if (macros) {
            macros[i] = r;
            macros[N + i] = u;
            macros[2 * N + i] = v;
            macros[3 * N + i] = w;
        }
        if (bad && !fault) {
            fault = 1;
            if (fault_node) *fault_node = i;
        }
    }
    return fault;
}

int orc_push_run(const orc_node_t *V, int64_t N, const int32_t *V_sub_start,
                 int num_p, int mode, double tau, double gx, int64_t it0,
                 int64_t nsteps, double *f0, double *f1, double *macros,
                 int64_t *fault_node, int64_t *fault_step)
{
    double omega = 1.0 / tau; /* porelb.c:88 */
    for (int64_t s = 0; s < nsteps; s++) {
        int64_t it = it0 + s;
        if (mode == ORC_BB_HALFWAY) {
            /* Textbook push form: the post-collision population leaving x
             * along e_j lands at x+e_j, or returns to x reversed when x+e_j
             * is solid. */
            for (int64_t g = 0; g < N; g++) {
                int p = owner_rank(V_sub_start, num_p, g);
                f1[g * Q] = f0[g * Q];
                for (int j = 1; j < Q; j++) {
                    const orc_nb_t *dn = &V[g].nb[j - 1];
                    if (dn->node_type == ORC_SOLID)
                        f1[g * Q + orc_re[j]] = f0[g * Q + j];
                    else
                        f1[gid_of(dn, p, V_sub_start) * Q + j] = f0[g * Q + j];
                }
            }
        } else {
            /* stream_inner_nodes (porelb.c:285-300), rank by rank in local
             * node order ... */
            for (int p = 0; p < num_p; p++) {
                int64_t b = V_sub_start[p], e_ = (p + 1 < num_p) ? V_sub_start[p + 1] : N;
                for (int64_t g = b; g < e_; g++) {
                    f1[g * Q] = f0[g * Q];
                    for (int j = 1; j < Q; j++) {
                        const orc_nb_t *dn = &V[g].nb[j - 1];
                        if (dn->node_type == ORC_FLUID)
                            f1[((int64_t)V_sub_start[p] + dn->node_id) * Q + j] = f0[g * Q + j];
                        else if (dn->node_type == ORC_SOLID)
                            f1[g * Q + j] = f0[g * Q + orc_re[j]];
                    }
                }
            }
            /* ... then update_send_dist / MPI / update_recv_dist
             * (porelb.c:240-312): every EXTERN link delivers f0[own][j] to
             * f1[nid][j] on the peer after its stream loop. */
            for (int64_t g = 0; g < N; g++)
                for (int j = 1; j < Q; j++) {
                    const orc_nb_t *dn = &V[g].nb[j - 1];
                    if (dn->node_type == ORC_EXTERN)
                        f1[((int64_t)V_sub_start[dn->node_pid] + dn->node_id) * Q + j] =
                            f0[g * Q + j];
                }
        }
        int64_t fn = -1;
        int bad = collide_all(f1, f0, N, omega, gx, it, macros, &fn);
        if (bad) {
            if (fault_node) *fault_node = fn;
            if (fault_step) *fault_step = it;
            return -6;
        }
    }
    return 0;
}

int orc_pull_run(const int32_t *tab, int64_t N, double tau, double gx, double rho0,
                 int64_t it0, int64_t nsteps, double *f0, double *scratch,
                 double *macros, int64_t *fault_node, int64_t *fault_step)
{
    double omega = 1.0 / tau;
    for (int64_t s = 0; s < nsteps; s++) {
        int64_t it = it0 + s;
        for (int64_t i = 0; i < N; i++) {
            scratch[i * Q] = f0[i * Q];
            for (int j = 1; j < Q; j++) {
                int32_t t = tab[(int64_t)(j - 1) * N + i];
                double v;
                if (t >= 0)
                    v = f0[(int64_t)t * Q + j];
                else if (t == ORC_TAB_BOUNCE)
                    v = f0[i * Q + orc_re[j]];
                else
                    v = orc_feq(j, rho0, 0.0, 0.0, 0.0);
                scratch[i * Q + j] = v;
            }
        }
        int64_t fn = -1;
        int bad = collide_all(scratch, f0, N, omega, gx, it, macros, &fn);
        if (bad) {
            if (fault_node) *fault_node = fn;
            if (fault_step) *fault_step = it;
            return -6;
        }
    }
    return 0;
}

void orc_reduce(const double *f0, const double *ux, int64_t N, int nx, int ny, int nz,
                double tau, double gx, double *out)
{
    double mass = 0.0, sux = 0.0;
    for (int64_t i = 0; i < N; i++) {
        for (int j = 0; j < Q; j++) mass += f0[i * Q + j];
        if (ux) sux += ux[i];
    }
    double nu = 1.0 / 3.0 * (tau - 0.5);
    double q = sux / ((double)nx * ny * nz);
    out[0] = mass;
    out[1] = sux;
    out[2] = N ? sux / (double)N : 0.0;
    out[3] = q;
    out[4] = nu * q / gx;
}

/* Sangani & Acrivos (1982) BCC drag-series coefficients, as tabulated in
 * getK.c:5-37 (published data). */
static const double orc_qs_bcc[31] = {
    1.0,       1.575834, 2.483254, 3.233022, 4.022864, 4.650320, 5.281412, 5.826374,
    6.258376,  6.544504, 6.878396, 7.190839, 7.268068, 7.304025, 7.301217, 7.236410,
    7.298014,  7.369849, 7.109497, 6.228418, 5.235796, 4.476874, 3.541982, 2.939353,
    3.935484,  5.179097, 3.959872, 2.227627, 3.393390, 4.491369, 2.200686};

/* getK.c:39-72. L = nx-1, a = sqrt(3) L chi / 4, eps = 1 - (8/3) pi a^3/L^3,
 * S = 8 pi a^2 / L^3, d* = sum qs_i chi^i, K1 = 2 a^2 / (9 d* (1-eps)),
 * K2 = eps^3 / (5 S^2). */
void orc_getK_bcc(double chi, int nx, double *out)
{
    const double pi = 3.14159265358979323846;
    double L = (double)(nx - 1);
    double a = 1.7320508075688772 * L * chi / 4.0;
    double eps = 1.0 - 8.0 / 3 * pi * a * a * a / L / L / L;
    double S = 2 * 4 * pi * a * a / L / L / L;
    double dstar = 0.0, p = 1.0;
    for (int i = 0; i < 31; i++) {
        dstar += orc_qs_bcc[i] * p;
        p *= chi;
    }
    out[0] = 2.0 / 9 * a * a / dstar / (1 - eps);
    out[1] = eps * eps * eps / 5.0 / S / S;
    out[2] = eps;
    out[3] = a;
}
