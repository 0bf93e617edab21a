// plb_api.hip -- the extern "C" boundary of libplb (declared in include/plb.h).
//
// One context = one pore-node lattice resident on one GPU, one HIP stream.
// The reference's SIMULATION lifecycle (porelb.c:402-443) maps as:
//   alloc_memory/check_V/build_send_recv_info -> plb_create / plb_create_from_flag
//   init_dist_macro                            -> plb_init_equilibrium
//   loop of send/stream/recv/collision         -> plb_step
//   save_data                                  -> plb_read_macros (+ host file write)
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>

#include "plb_internal.hpp"

using namespace plb;

void plb_set_hip_error(hipError_t e, const char* what, const char* file, int line)
{
    fprintf(stderr, "libplb: HIP error %d (%s) in %s at %s:%d\n", (int)e, hipGetErrorString(e),
            what, file, line);
}

extern "C" void plb_config_default(plb_config* cfg)
{
    memset(cfg, 0, sizeof *cfg);
    cfg->tau = 0.8;   // porelb.c:87
    cfg->gx = 1.0e-3; // porelb.c:91
    cfg->rho0 = 1.0;  // porelb.c:228
    cfg->precision = PLB_FP64;
    cfg->bb_mode = PLB_BB_REFCOMPAT;
    cfg->layout = PLB_LAYOUT_SOA;
    cfg->ordering = PLB_ORDER_REFERENCE;
    cfg->brick[0] = 16;
    cfg->brick[1] = 4;
    cfg->brick[2] = 4;
    cfg->block_size = 256;
    cfg->device = 0;
    cfg->fp16_scale_log2 = 0;
    cfg->index_kind = PLB_INDEX_I32;
    cfg->nodes_per_thread = 1;
    cfg->graph_steps = 0;
    cfg->waves_per_simd = 0;
    cfg->fp16_rounding = PLB_ROUND_NEAREST;
    cfg->stream_mode = PLB_STREAM_PULL;
}

extern "C" const char* plb_strerror(int code)
{
    switch (code) {
    case PLB_OK: return "ok";
    case PLB_ERR_FILE: return "input file could not be opened";
    case PLB_ERR_NUM_P: return "rank count does not fit the V.bin / z-slab rule";
    case PLB_ERR_BOUNDS: return "node coordinate or neighbour id out of range";
    case PLB_ERR_PID: return "EXTERN neighbour with invalid rank id";
    case PLB_ERR_READ: return "short read";
    case PLB_ERR_DENSITY: return "density guard: rho > 1.2 or rho < 0.5";
    case PLB_ERR_ARGS: return "bad command line";
    case PLB_ERR_INVALID: return "invalid argument";
    case PLB_ERR_HIP: return "HIP runtime error";
    case PLB_ERR_NOMEM: return "out of memory";
    case PLB_ERR_RANGE: return "too many pore nodes for int32 node ids";
    case PLB_ERR_NODEVICE: return "no usable GPU";
    case PLB_ERR_FLAG: return "voxel flag value other than 0 or 1";
    default: return "unknown error";
    }
}

int plb::check_cfg(const plb_config* cfg)
{
    if (!cfg) return PLB_ERR_INVALID;
    if (!(cfg->tau > 0.0) || !isfinite(cfg->gx) || !(cfg->rho0 > 0.0)) return PLB_ERR_INVALID;
    if (cfg->precision < PLB_FP64 || cfg->precision > PLB_FP16) return PLB_ERR_INVALID;
    if (cfg->bb_mode != PLB_BB_REFCOMPAT && cfg->bb_mode != PLB_BB_HALFWAY) return PLB_ERR_INVALID;
    if (cfg->layout != PLB_LAYOUT_SOA && cfg->layout != PLB_LAYOUT_AOSOA) return PLB_ERR_INVALID;
    if (cfg->ordering != PLB_ORDER_REFERENCE && cfg->ordering != PLB_ORDER_BRICK &&
        cfg->ordering != PLB_ORDER_PENCIL)
        return PLB_ERR_INVALID;
    if (cfg->ordering != PLB_ORDER_REFERENCE &&
        (cfg->brick[0] < 1 || cfg->brick[1] < 1 || cfg->brick[2] < 1 ||
         cfg->brick[0] > 65535 || cfg->brick[1] > 65535 || cfg->brick[2] > 65535))
        return PLB_ERR_INVALID;
    if (cfg->block_size < 64 || cfg->block_size > 1024 || cfg->block_size % 64) return PLB_ERR_INVALID;
    if (cfg->fp16_scale_log2 < 0 || cfg->fp16_scale_log2 > 100) return PLB_ERR_INVALID;
    if (cfg->index_kind != PLB_INDEX_I32 && cfg->index_kind != PLB_INDEX_TILE8 &&
        cfg->index_kind != PLB_INDEX_BITMAP)
        return PLB_ERR_INVALID;
    // BITMAP derives sources from voxel positions: half-way bounce-back, and
    // node orders that are whole x-rows (checked against nx at build time).
    if (cfg->index_kind == PLB_INDEX_BITMAP &&
        (cfg->bb_mode != PLB_BB_HALFWAY ||
         (cfg->ordering != PLB_ORDER_REFERENCE && cfg->ordering != PLB_ORDER_PENCIL) ||
         (cfg->ordering == PLB_ORDER_PENCIL && cfg->brick[2] != 1)))
        return PLB_ERR_INVALID;
    if (cfg->nodes_per_thread != 1 && cfg->nodes_per_thread != 2) return PLB_ERR_INVALID;
    if (cfg->nodes_per_thread == 2 && cfg->block_size > 512) return PLB_ERR_INVALID;
    if (cfg->graph_steps < 0 || cfg->graph_steps > 4096 || cfg->graph_steps % 2) return PLB_ERR_INVALID;
    if (cfg->waves_per_simd < 0 || cfg->waves_per_simd > 8) return PLB_ERR_INVALID;
    if (cfg->fp16_rounding != PLB_ROUND_NEAREST && cfg->fp16_rounding != PLB_ROUND_STOCHASTIC)
        return PLB_ERR_INVALID;
    if (cfg->placement_reserve_gib < 0 || cfg->placement_reserve_gib > 160) return PLB_ERR_INVALID;
    if (cfg->stream_mode != PLB_STREAM_PULL && cfg->stream_mode != PLB_STREAM_AA) return PLB_ERR_INVALID;
    if (cfg->stream_mode == PLB_STREAM_AA &&
        (cfg->bb_mode != PLB_BB_HALFWAY || cfg->nodes_per_thread != 1 || cfg->graph_steps != 0))
        return PLB_ERR_INVALID;
    return PLB_OK;
}

// FP16 storage scale 2^k: deviations of order 1000*|gx| map to ~2048, far
// from both the FP16 overflow (65504) and its subnormal range (< 6.1e-5).
int plb::fp16_scale_log2_for(double gx)
{
    const double g = fabs(gx) > 0 ? fabs(gx) : 1e-9;
    const int k = (int)floor(log2(2048.0 / (1000.0 * g)));
    return std::max(1, std::min(k, 60));
}

int plb::ctx_begin(plb_ctx* c, const plb_config* cfg, int nx, int ny, int nz, int num_p)
{
    c->cfg = *cfg;
    c->nx = nx;
    c->ny = ny;
    c->nz = nz;
    c->num_p = num_p;
    c->elem = cfg->precision == PLB_FP64 ? 8 : (cfg->precision == PLB_FP32 ? 4 : 2);
    c->omega = 1.0 / cfg->tau;  // porelb.c:88
    const int k = cfg->fp16_scale_log2 ? cfg->fp16_scale_log2 : fp16_scale_log2_for(cfg->gx);
    c->scale_log2 = cfg->precision == PLB_FP16 ? k : 0;
    c->scale = cfg->precision == PLB_FP16 ? ldexpf(1.f, k) : 1.f;
    c->inv_scale = cfg->precision == PLB_FP16 ? ldexpf(1.f, -k) : 1.f;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return PLB_ERR_NODEVICE;
    if (cfg->device < 0 || cfg->device >= ndev) return PLB_ERR_INVALID;
    c->device = cfg->device;
    PLB_HIP(hipSetDevice(c->device));
    PLB_HIP(hipDeviceGetAttribute(&c->lds_per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                                  c->device));
    const int rc = noalloc_setup(c->device);  // the no-alloc read-back's staging array
    if (rc) return rc;
    PLB_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    if (cfg->placement_reserve_gib > 0)  // before every buffer of the context
        PLB_HIP(hipMalloc(&c->reserve, (size_t)cfg->placement_reserve_gib << 30));
    PLB_HIP(hipEventCreate(&c->ev0));
    PLB_HIP(hipEventCreate(&c->ev1));
    return PLB_OK;
}

int plb::ctx_alloc_state(plb_ctx* c)
{
    const size_t nf = (size_t)c->Npad * Q * c->elem;
    PLB_HIP(hipMalloc(&c->f[0], nf));
    if (c->cfg.stream_mode == PLB_STREAM_AA)
        c->f[1] = c->f[0];  // one buffer; c->cur is the AA phase
    else
        PLB_HIP(hipMalloc(&c->f[1], nf));
    PLB_HIP(hipMalloc(&c->macros, sizeof(double) * 4 * c->Npad));
    PLB_HIP(hipMemsetAsync(c->macros, 0, sizeof(double) * 4 * c->Npad, c->stream));
    PLB_HIP(hipMalloc(&c->fault, sizeof(unsigned long long)));
    PLB_HIP(hipMemsetAsync(c->fault, 0xff, sizeof(unsigned long long), c->stream));
    int dev_cus = 256;
    (void)hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, c->device);
    c->red_blocks = std::max(1, std::min<int>(dev_cus * 4, (int)((c->N + 255) / 256)));
    PLB_HIP(hipMalloc(&c->red, sizeof(double) * (2 * c->red_blocks + 2)));
    PLB_HIP(hipMalloc(&c->it_dev, sizeof(int64_t)));
    PLB_HIP(hipStreamSynchronize(c->stream));
    if (c->cfg.precision == PLB_FP16 && c->cfg.fp16_rounding == PLB_ROUND_STOCHASTIC) {
        c->key_bytes = 4.0;
        const int rc = build_key_runs(c);
        if (rc) return rc;
    }
    return plb_init_equilibrium(c);
}

extern "C" void plb_destroy(plb_ctx* c)
{
    if (!c) return;
    if (c->stream) {
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
    }
    void* ptrs[] = {c->f[0], c->f[1] != c->f[0] ? c->f[1] : nullptr, c->tab, c->tdesc, c->toff, c->toff16, c->twide, c->bm, c->vox, c->perm,
                    c->inv_perm, c->key_perm, c->kseg, c->node_xyz, c->macros, c->fault, c->red, c->reserve};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (hipGraphExec_t& g : c->gexec)
        if (g) (void)hipGraphExecDestroy(g);
    if (c->it_dev) (void)hipFree(c->it_dev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" int plb_create(plb_ctx** out, const plb_config* cfg, const plb_node_info* V,
                          int64_t n_nodes, const int32_t* V_sub_start, int num_p, int nx, int ny,
                          int nz)
{
    if (!out) return PLB_ERR_INVALID;
    *out = nullptr;
    Range range("plb_create");
    int rc = check_cfg(cfg);
    if (rc) return rc;
    if (n_nodes < 0 || (n_nodes > 0 && !V) || !V_sub_start) return PLB_ERR_INVALID;
    if (nx <= 0 || ny <= 0 || nz <= 0 || nx > 65535 || ny > 65535 || nz > 65535)
        return PLB_ERR_INVALID;
    if (num_p < 1 || num_p > 65535) return PLB_ERR_NUM_P;
    if (n_nodes >= INT32_MAX) return PLB_ERR_RANGE;
    // V_sub_start must partition [0, n_nodes) into consecutive rank slices.
    if (V_sub_start[0] != 0) return PLB_ERR_NUM_P;
    for (int p = 1; p < num_p; p++)
        if (V_sub_start[p] < V_sub_start[p - 1]) return PLB_ERR_NUM_P;
    if (V_sub_start[num_p - 1] > n_nodes) return PLB_ERR_NUM_P;

    plb_ctx* c = new plb_ctx();
    rc = ctx_begin(c, cfg, nx, ny, nz, num_p);
    if (!rc) {
        c->V_sub_start.assign(V_sub_start, V_sub_start + num_p);
        c->N = n_nodes;
        c->Npad = std::max<int64_t>(64, (n_nodes + 63) / 64 * 64);
        rc = build_table_from_V(c, V, n_nodes);
    }
    if (!rc) rc = ctx_alloc_state(c);
    if (rc) {
        plb_destroy(c);
        return rc;
    }
    *out = c;
    return PLB_OK;
}

extern "C" int plb_create_from_flag(plb_ctx** out, const plb_config* cfg, const uint8_t* flag,
                                    int nx, int ny, int nz, int num_p)
{
    if (!out) return PLB_ERR_INVALID;
    *out = nullptr;
    Range range("plb_create_from_flag");
    int rc = check_cfg(cfg);
    if (rc) return rc;
    if (!flag || nx <= 0 || ny <= 0 || nz <= 0 || nx > 65535 || ny > 65535 || nz > 65535)
        return PLB_ERR_INVALID;
    // z-slab rule pid = z/(nz/num_p) must stay inside [0, num_p) (preprocess.c:107).
    if (num_p < 1 || num_p > 65535 || nz / num_p == 0) return PLB_ERR_NUM_P;
    if ((nz - 1) / (nz / num_p) >= num_p) return PLB_ERR_NUM_P;
    plb_ctx* c = new plb_ctx();
    rc = ctx_begin(c, cfg, nx, ny, nz, num_p);
    if (!rc) rc = build_table_from_flag(c, flag);
    if (!rc) rc = ctx_alloc_state(c);
    if (rc) {
        plb_destroy(c);
        return rc;
    }
    *out = c;
    return PLB_OK;
}

__global__ void k_fill_macros(double* m, int64_t Npad, double rho0)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Npad) return;
    m[i] = rho0;
    m[Npad + i] = 0.0;
    m[2 * Npad + i] = 0.0;
    m[3 * Npad + i] = 0.0;
}

extern "C" int plb_init_equilibrium(plb_ctx* c)
{
    if (!c) return PLB_ERR_INVALID;
    PLB_HIP(hipSetDevice(c->device));
    // AA: the rest state (every slot w_j rho0) is the local-step form, so the
    // first step is a neighbour step, like a pull step
    c->cur = c->cfg.stream_mode == PLB_STREAM_AA ? 1 : 0;
    int rc = launch_init_eq(c);
    if (rc) return rc;
    hipLaunchKernelGGL(k_fill_macros, dim3((unsigned)((c->Npad + 255) / 256)), dim3(256), 0,
                       c->stream, c->macros, c->Npad, c->cfg.rho0);
    PLB_HIP(hipGetLastError());
    PLB_HIP(hipMemsetAsync(c->fault, 0xff, sizeof(unsigned long long), c->stream));
    PLB_HIP(hipStreamSynchronize(c->stream));
    c->pending_n = 0;
    return PLB_OK;
}

static unsigned long long read_fault_key(plb_ctx* c, int* rc)
{
    unsigned long long key = NO_FAULT;
    hipError_t e = hipMemcpyAsync(&key, c->fault, sizeof key, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        plb_set_hip_error(e, "read fault", __FILE__, __LINE__);
        *rc = PLB_ERR_HIP;
    }
    return key;
}

// A state written in from outside starts with a clear density guard: the
// fault word is reset and nothing is pending.
static int clear_fault(plb_ctx* c)
{
    PLB_HIP(hipMemsetAsync(c->fault, 0xff, sizeof(unsigned long long), c->stream));
    PLB_HIP(hipStreamSynchronize(c->stream));
    c->pending_n = 0;
    return PLB_OK;
}

extern "C" int plb_step_async(plb_ctx* c, int64_t it0, int64_t nsteps)
{
    if (!c || nsteps < 0 || it0 < 0 || it0 + nsteps >= ((int64_t)1 << 31)) return PLB_ERR_INVALID;
    Range range("plb_step");
    PLB_HIP(hipSetDevice(c->device));
    if (c->pending_n == 0) {
        c->pending_it0 = it0;
        c->cur_before = c->cur;
    }
    int64_t s = 0;
    // Runs of graph_steps steps replay a captured graph; the last step of the
    // call is always a plain launch (it writes the macros).
    for (const int64_t G = c->N > 0 ? c->cfg.graph_steps : 0; G > 0 && nsteps - s > G; s += G) {
        int rc = launch_graph_steps(c, it0 + s, G);
        if (rc) return rc;
    }
    for (; s < nsteps; s++) {
        int rc = launch_step(c, it0 + s, s == nsteps - 1);
        if (rc) return rc;
    }
    c->pending_n += nsteps;
    return PLB_OK;
}

extern "C" int plb_sync(plb_ctx* c)
{
    if (!c) return PLB_ERR_INVALID;
    PLB_HIP(hipSetDevice(c->device));
    int rc = PLB_OK;
    const unsigned long long key = read_fault_key(c, &rc);
    if (rc) return rc;
    const int64_t it0 = c->pending_it0, n = c->pending_n;
    c->pending_n = 0;
    if (key == NO_FAULT) return PLB_OK;
    const int64_t fs = (int64_t)(key >> 32);
    if (n > 0) {
        // Steps after the faulting one did not run: the faulting step's
        // output buffer holds the current state.  A fault from an earlier
        // call made every launch of this one a no-op.
        if (fs >= it0 && fs < it0 + n)
            c->cur = c->cur_before ^ (int)((fs - it0 + 1) & 1);
        else if (fs < it0)
            c->cur = c->cur_before;
    }
    return PLB_ERR_DENSITY;
}

extern "C" int plb_step(plb_ctx* c, int64_t it0, int64_t nsteps)
{
    int rc = plb_step_async(c, it0, nsteps);
    if (rc) return rc;
    return plb_sync(c);
}

extern "C" int plb_step_timed(plb_ctx* c, int64_t it0, int64_t nsteps, double* ms)
{
    if (!c || !ms) return PLB_ERR_INVALID;
    PLB_HIP(hipSetDevice(c->device));
    PLB_HIP(hipEventRecord(c->ev0, c->stream));
    int rc = plb_step_async(c, it0, nsteps);
    if (rc) return rc;
    PLB_HIP(hipEventRecord(c->ev1, c->stream));
    PLB_HIP(hipEventSynchronize(c->ev1));
    float t = 0.f;
    PLB_HIP(hipEventElapsedTime(&t, c->ev0, c->ev1));
    *ms = t;
    return plb_sync(c);
}

extern "C" int64_t plb_num_nodes(const plb_ctx* c) { return c ? c->N : -1; }

extern "C" void* plb_host_alloc(size_t bytes)
{
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) return nullptr;
    return p;
}

extern "C" void plb_host_free(void* p)
{
    if (p) (void)hipHostFree(p);
}

extern "C" double plb_bytes_per_update(const plb_ctx* c)
{
    if (!c) return 0.0;
    // FP16 storage with stochastic rounding also reads each node's V-order key:
    // 16 bytes per tile of runs, plus 4 per node in tiles that load their keys
    const bool key = c->cfg.precision == PLB_FP16 && c->cfg.fp16_rounding == PLB_ROUND_STOCHASTIC &&
                     (c->perm || c->key_perm);
    // AA streaming reads the table on odd steps only
    const double idx = c->cfg.stream_mode == PLB_STREAM_AA ? 0.5 * c->index_bytes : c->index_bytes;
    return 2.0 * Q * (double)c->elem + idx + (key ? c->key_bytes : 0.0);
}

static const int64_t CHUNK_NODES = 1 << 20;


extern "C" int plb_read_macros(plb_ctx* c, double* rho, double* ux, double* uy, double* uz)
{
    if (!c) return PLB_ERR_INVALID;
    Range range("plb_read_macros");
    PLB_HIP(hipSetDevice(c->device));
    double* outs[4] = {rho, ux, uy, uz};
    const int64_t chunk = std::min<int64_t>(std::max<int64_t>(c->N, 1), 8 * CHUNK_NODES);
    double* d = nullptr;
    PLB_HIP(hipMalloc(&d, sizeof(double) * chunk));
    int rc = PLB_OK;
    for (int comp = 0; comp < 4 && !rc; comp++) {
        if (!outs[comp]) continue;
        for (int64_t b = 0; b < c->N && !rc; b += chunk) {
            const int64_t n = std::min(chunk, c->N - b);
            rc = launch_read_macros(c, comp, b, n, d);
            if (!rc && hipMemcpyAsync(outs[comp] + b, d, sizeof(double) * n, hipMemcpyDeviceToHost,
                                      c->stream) != hipSuccess)
                rc = PLB_ERR_HIP;
            if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = PLB_ERR_HIP;
        }
    }
    (void)hipFree(d);
    return rc;
}

extern "C" int plb_read_populations(plb_ctx* c, double* f_aos)
{
    if (!c || (!f_aos && c->N)) return PLB_ERR_INVALID;
    PLB_HIP(hipSetDevice(c->device));
    const int64_t chunk = std::min<int64_t>(std::max<int64_t>(c->N, 1), CHUNK_NODES);
    double* d = nullptr;
    PLB_HIP(hipMalloc(&d, sizeof(double) * Q * chunk));
    int rc = PLB_OK;
    for (int64_t b = 0; b < c->N && !rc; b += chunk) {
        const int64_t n = std::min(chunk, c->N - b);
        rc = launch_read_pop(c, b, n, d);
        if (!rc && hipMemcpyAsync(f_aos + b * Q, d, sizeof(double) * Q * n, hipMemcpyDeviceToHost,
                                  c->stream) != hipSuccess)
            rc = PLB_ERR_HIP;
        if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = PLB_ERR_HIP;
    }
    (void)hipFree(d);
    return rc;
}

extern "C" int plb_write_populations(plb_ctx* c, const double* f_aos)
{
    if (!c || (!f_aos && c->N)) return PLB_ERR_INVALID;
    PLB_HIP(hipSetDevice(c->device));
    const int64_t chunk = std::min<int64_t>(std::max<int64_t>(c->N, 1), CHUNK_NODES);
    double* d = nullptr;
    PLB_HIP(hipMalloc(&d, sizeof(double) * Q * chunk));
    int rc = PLB_OK;
    for (int64_t b = 0; b < c->N && !rc; b += chunk) {
        const int64_t n = std::min(chunk, c->N - b);
        if (hipMemcpyAsync(d, f_aos + b * Q, sizeof(double) * Q * n, hipMemcpyHostToDevice,
                           c->stream) != hipSuccess)
            rc = PLB_ERR_HIP;
        if (!rc) rc = launch_write_pop(c, b, n, d);
        if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = PLB_ERR_HIP;
    }
    (void)hipFree(d);
    if (!rc) rc = clear_fault(c);
    return rc;
}

extern "C" int plb_get_state_info(const plb_ctx* c, plb_state_info* out)
{
    if (!c || !out) return PLB_ERR_INVALID;
    memset(out, 0, sizeof *out);
    out->precision = c->cfg.precision;
    out->elem_bytes = (int)c->elem;
    out->fp16_scale_log2 = c->scale_log2;
    out->rho0 = c->cfg.rho0;
    out->n_nodes = c->N;
    return PLB_OK;
}

// Stored elements between host AoS [n][19] and the device layout, through a
// device staging chunk.
static int copy_state(plb_ctx* c, void* host, bool to_device)
{
    if (!c || (!host && c->N)) return PLB_ERR_INVALID;
    PLB_HIP(hipSetDevice(c->device));
    const int64_t chunk = std::min<int64_t>(std::max<int64_t>(c->N, 1), CHUNK_NODES);
    const size_t rec = Q * c->elem;
    void* d = nullptr;
    PLB_HIP(hipMalloc(&d, rec * chunk));
    int rc = PLB_OK;
    char* h = (char*)host;
    for (int64_t b = 0; b < c->N && !rc; b += chunk) {
        const int64_t n = std::min(chunk, c->N - b);
        if (to_device) {
            if (hipMemcpyAsync(d, h + b * rec, rec * n, hipMemcpyHostToDevice, c->stream) != hipSuccess)
                rc = PLB_ERR_HIP;
            if (!rc) rc = launch_copy_state(c, b, n, d, true);
        } else {
            rc = launch_copy_state(c, b, n, d, false);
            if (!rc && hipMemcpyAsync(h + b * rec, d, rec * n, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
                rc = PLB_ERR_HIP;
        }
        if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = PLB_ERR_HIP;
    }
    (void)hipFree(d);
    return rc;
}

extern "C" int plb_read_state(plb_ctx* c, void* buf) { return copy_state(c, buf, false); }

extern "C" int plb_write_state(plb_ctx* c, const void* buf)
{
    int rc = copy_state(c, const_cast<void*>(buf), true);
    if (!rc) rc = clear_fault(c);
    return rc;
}

extern "C" int plb_read_table(plb_ctx* c, int32_t* tab)
{
    if (!c || (!tab && c->N)) return PLB_ERR_INVALID;
    PLB_HIP(hipSetDevice(c->device));
    const int64_t chunk = std::min<int64_t>(std::max<int64_t>(c->N, 1), CHUNK_NODES);
    int32_t* d = nullptr;
    PLB_HIP(hipMalloc(&d, sizeof(int32_t) * (Q - 1) * chunk));
    int rc = PLB_OK;
    for (int64_t b = 0; b < c->N && !rc; b += chunk) {
        const int64_t n = std::min(chunk, c->N - b);
        rc = launch_read_table(c, b, n, d);
        for (int j = 0; j < Q - 1 && !rc; j++)
            if (hipMemcpyAsync(tab + (int64_t)j * c->N + b, d + (int64_t)j * n,
                               sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
                rc = PLB_ERR_HIP;
        if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = PLB_ERR_HIP;
    }
    (void)hipFree(d);
    return rc;
}

extern "C" int plb_reduce(plb_ctx* c, plb_reduction* out)
{
    if (!c || !out) return PLB_ERR_INVALID;
    Range range("plb_reduce");
    PLB_HIP(hipSetDevice(c->device));
    double r[2] = {0.0, 0.0};
    if (c->N > 0) {
        int rc = launch_reduce(c, r);
        if (rc) return rc;
    }
    double mass = r[0];
    if (c->cfg.precision != PLB_FP64) {
        double w = 0.0;
        for (int j = 0; j < Q; j++) w += TP[j] * c->cfg.rho0;
        mass += w * (double)c->N;
    }
    out->mass = mass;
    out->sum_ux = r[1];
    out->mean_ux = c->N ? r[1] / (double)c->N : 0.0;
    out->q = r[1] / ((double)c->nx * c->ny * c->nz);
    const double nu = 1.0 / 3.0 * (c->cfg.tau - 0.5);  // porelb.c:89
    out->K = c->cfg.gx != 0.0 ? nu * out->q / c->cfg.gx : 0.0;
    out->n_nodes = c->N;
    return PLB_OK;
}

extern "C" int plb_run_to_convergence(plb_ctx* c, int64_t it0, int64_t max_steps, int64_t check_every,
                                      double tol, int64_t* steps_done, plb_reduction* out)
{
    if (!c || !steps_done || !out || it0 < 0 || max_steps < 0 || check_every < 1 || !(tol >= 0.0))
        return PLB_ERR_INVALID;
    Range range("plb_run_to_convergence");
    *steps_done = 0;
    int rc = plb_reduce(c, out);  // the state before the first run (macros as stored)
    if (rc) return rc;
    double prev = NAN;
    const double umin = 1e-3 * fabs(c->cfg.gx);
    for (int64_t s = 0; s < max_steps;) {
        const int64_t m = std::min(check_every, max_steps - s);
        rc = plb_step(c, it0 + s, m);
        if (rc) return rc;
        s += m;
        *steps_done = s;
        rc = plb_reduce(c, out);
        if (rc) return rc;
        const double mean = out->mean_ux;
        if (tol > 0.0 && !isnan(prev) && fabs(mean - prev) <= tol * std::max(fabs(mean), umin)) break;
        prev = mean;
    }
    return PLB_OK;
}

extern "C" int plb_last_fault(plb_ctx* c, plb_fault* out)
{
    if (!c || !out) return PLB_ERR_INVALID;
    PLB_HIP(hipSetDevice(c->device));
    memset(out, 0, sizeof *out);
    int rc = PLB_OK;
    const unsigned long long key = read_fault_key(c, &rc);
    if (rc) return rc;
    if (key == NO_FAULT) return PLB_OK;
    out->tripped = 1;
    out->step = (int64_t)(key >> 32);
    out->node = (int64_t)(key & 0xffffffffull);
    if (out->node >= c->N) return PLB_ERR_BOUNDS;
    int p = 0;
    for (int q = 0; q < c->num_p; q++)
        if (c->V_sub_start[q] <= out->node) p = q;
    out->pid = p;
    out->local_id = out->node - c->V_sub_start[p];
    int64_t k = out->node;
    if (c->inv_perm) {
        uint32_t kk = 0;
        PLB_HIP(hipMemcpy(&kk, c->inv_perm + out->node, sizeof kk, hipMemcpyDeviceToHost));
        if ((int64_t)kk >= c->N) return PLB_ERR_BOUNDS;
        k = kk;
    }
    uint64_t xyz = 0;
    PLB_HIP(hipMemcpy(&xyz, c->node_xyz + k, sizeof xyz, hipMemcpyDeviceToHost));
    PLB_HIP(hipMemcpy(&out->rho, c->macros + k, sizeof(double), hipMemcpyDeviceToHost));
    out->x = (int)(xyz & 0xffff);
    out->y = (int)((xyz >> 16) & 0xffff);
    out->z = (int)((xyz >> 32) & 0xffff);
    return PLB_OK;
}
