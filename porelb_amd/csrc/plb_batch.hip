// plb_batch.hip -- many independent porous samples in one lattice.
//
// The samples' pore nodes are concatenated, each sample padded to a multiple
// of the block size, so every block of the step kernel lies in one sample and
// takes that sample's omega and body force.  One launch per time step covers
// all active samples; the grid enumerates the blocks of active samples only
// (a device list rebuilt when a sample stops), so a converged sample costs
// nothing.  Every `check_every` steps the ux macros are stored and summed per
// sample in FP64; a sample stops when |d<u_x>| / |<u_x>| < tol between two
// checks.  A stopped sample's state stays in the ping-pong buffer that held
// it at its last step (recorded per sample).  The macros are stored on every
// check step, so after a run they hold each sample's last step; the batch
// context's inv_perm maps each sample's V order into the concatenation.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "plb_internal.hpp"
#include "plb_tune.h"

using namespace plb;

struct plb_batch {
    plb_ctx* c = nullptr;
    int S = 0;
    std::vector<int64_t> off, n;
    std::vector<int> nx, ny, nz;
    std::vector<double> tau, gx;
    std::vector<int32_t> block_sample;
    std::vector<int> sample_cur;
    std::vector<char> active;
    int32_t* d_block_sample = nullptr;
    int32_t* d_active = nullptr;
    int64_t* d_end = nullptr;
    int64_t* d_off = nullptr;
    int64_t* d_n = nullptr;
    double* d_omega = nullptr;
    double* d_gx = nullptr;
    double* d_sums = nullptr;
    double* d_partial = nullptr;  // per-chunk u_x partial sums
};

extern "C" void plb_batch_destroy(plb_batch* b)
{
    if (!b) return;
    if (b->c) {
        (void)hipSetDevice(b->c->device);
        (void)hipStreamSynchronize(b->c->stream);
    }
    void* ptrs[] = {b->d_block_sample, b->d_active, b->d_end, b->d_off, b->d_n,
                    b->d_omega, b->d_gx, b->d_sums, b->d_partial};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    plb_destroy(b->c);
    delete b;
}

// inv[off + g] = off + (sample order of V-order node g), identity without a
// sample order.
__global__ void k_batch_inv(uint32_t* inv, const uint32_t* sample_inv, int64_t n, int64_t off)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    inv[off + g] = (uint32_t)(off + (sample_inv ? (int64_t)sample_inv[g] : g));
}

// For the node at internal index k = off + (sample order of V-order node g):
// key[k] = g, its V-order id within its sample (keys the FP16
// stochastic-rounding stream), and perm[k] = off + g, its V-order id in the
// concatenation (the density guard reports the lowest of these, as the
// reference reports the first node of its V-order loop).
__global__ void k_batch_key(uint32_t* key, uint32_t* perm, const uint32_t* sample_inv, int64_t n, int64_t off)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int64_t k = off + (sample_inv ? (int64_t)sample_inv[g] : g);
    key[k] = (uint32_t)g;
    perm[k] = (uint32_t)(off + g);
}

static int batch_build(plb_batch* b, const plb_config* cfg, const plb_sample* smp)
{
    plb_ctx* c = b->c;
    const int bs = cfg->block_size;
    std::vector<int32_t*> tabs(b->S, nullptr);
    std::vector<uint64_t*> xyzs(b->S, nullptr);
    std::vector<uint32_t*> invs(b->S, nullptr);
    int rc = PLB_OK;
    int64_t total = 0;
    for (int s = 0; s < b->S && !rc; s++) {
        int64_t ns = 0;
        std::vector<int32_t> start;
        rc = build_ref_from_flag(smp[s].flag, smp[s].nx, smp[s].ny, smp[s].nz, 1, cfg->bb_mode,
                                 c->stream, &tabs[s], &xyzs[s], &ns, &start);
        if (!rc && cfg->ordering != PLB_ORDER_REFERENCE && ns > 0) {
            // The sample's node order (bricks or pencils with the sample's own
            // x extent for pencils); its V-order map is kept for read-backs.
            plb_config scfg = *cfg;
            if (scfg.ordering == PLB_ORDER_PENCIL) scfg.brick[0] = smp[s].nx;
            int32_t* t2 = nullptr;
            rc = compute_order(scfg, smp[s].nx, smp[s].ny, smp[s].nz, ns, xyzs[s], c->stream, &invs[s],
                               nullptr, nullptr);
            if (!rc && hipMalloc(&t2, sizeof(int32_t) * (Q - 1) * ns) != hipSuccess) rc = PLB_ERR_NOMEM;
            if (!rc) rc = permute_ref_table(tabs[s], ns, invs[s], t2, c->stream);
            if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = PLB_ERR_HIP;
            if (!rc) {
                (void)hipFree(tabs[s]);
                tabs[s] = t2;
            } else if (t2) {
                (void)hipFree(t2);
            }
        }
        b->off[s] = total;
        b->n[s] = ns;
        total += std::max<int64_t>(bs, (ns + bs - 1) / bs * bs);
    }
    if (!rc && total >= INT32_MAX) rc = PLB_ERR_RANGE;
    if (!rc) {
        c->N = total;
        c->Npad = total;
        hipError_t e = hipMalloc(&c->tab, sizeof(int32_t) * (Q - 1) * total);
        if (e == hipSuccess) e = hipMemsetAsync(c->tab, 0, sizeof(int32_t) * (Q - 1) * total, c->stream);
        if (e == hipSuccess) e = hipMalloc(&c->node_xyz, sizeof(uint64_t) * total);
        if (e == hipSuccess) e = hipMalloc(&c->inv_perm, sizeof(uint32_t) * total);
        if (e == hipSuccess) e = hipMalloc(&c->key_perm, sizeof(uint32_t) * total);
        if (e == hipSuccess) e = hipMemsetAsync(c->key_perm, 0, sizeof(uint32_t) * total, c->stream);
        if (e == hipSuccess) e = hipMalloc(&c->perm, sizeof(uint32_t) * total);
        if (e == hipSuccess) e = hipMemsetAsync(c->perm, 0, sizeof(uint32_t) * total, c->stream);
        if (e != hipSuccess) rc = e == hipErrorOutOfMemory ? PLB_ERR_NOMEM : PLB_ERR_HIP;
    }
    for (int s = 0; s < b->S && !rc; s++) {
        rc = copy_table_offset(cfg->layout, tabs[s], b->n[s], b->off[s], c->tab, c->Npad, c->stream);
        if (!rc && b->n[s] &&
            hipMemcpyAsync(c->node_xyz + b->off[s], xyzs[s], sizeof(uint64_t) * b->n[s],
                           hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
            rc = PLB_ERR_HIP;
        if (!rc && b->n[s] > 0) {
            hipLaunchKernelGGL(k_batch_inv, dim3((unsigned)((b->n[s] + 255) / 256)), dim3(256), 0, c->stream,
                               c->inv_perm, invs[s], b->n[s], b->off[s]);
            hipLaunchKernelGGL(k_batch_key, dim3((unsigned)((b->n[s] + 255) / 256)), dim3(256), 0, c->stream,
                               c->key_perm, c->perm, invs[s], b->n[s], b->off[s]);
            if (hipGetLastError() != hipSuccess) rc = PLB_ERR_HIP;
        }
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess && !rc) rc = PLB_ERR_HIP;
    for (int s = 0; s < b->S; s++) {
        if (tabs[s]) (void)hipFree(tabs[s]);
        if (xyzs[s]) (void)hipFree(xyzs[s]);
        if (invs[s]) (void)hipFree(invs[s]);
    }
    c->index_bytes = (Q - 1) * 4.0;
    if (!rc && cfg->index_kind == PLB_INDEX_TILE8) rc = compress_tile8(c);
    return rc;
}

extern "C" int plb_batch_create(plb_batch** out, const plb_config* cfg, const plb_sample* samples,
                                int n_samples)
{
    if (!out) return PLB_ERR_INVALID;
    *out = nullptr;
    int rc = check_cfg(cfg);
    if (rc) return rc;
    if (n_samples < 1 || !samples) return PLB_ERR_INVALID;
    if (cfg->nodes_per_thread != 1) return PLB_ERR_INVALID;
    if (cfg->index_kind == PLB_INDEX_BITMAP) return PLB_ERR_INVALID;  // samples have their own dims
    if (cfg->stream_mode != PLB_STREAM_PULL) return PLB_ERR_INVALID;  // AA: single lattices only
    for (int s = 0; s < n_samples; s++) {
        const plb_sample& p = samples[s];
        if (!p.flag || p.nx <= 0 || p.ny <= 0 || p.nz <= 0 || p.nx > 65535 || p.ny > 65535 ||
            p.nz > 65535 || !(p.tau > 0.0) || !isfinite(p.gx))
            return PLB_ERR_INVALID;
    }
    plb_batch* b = new plb_batch();
    b->S = n_samples;
    b->off.assign(n_samples, 0);
    b->n.assign(n_samples, 0);
    b->c = new plb_ctx();
    plb_config bcfg = *cfg;
    if (bcfg.precision == PLB_FP16 && bcfg.fp16_scale_log2 == 0) {
        // one storage scale for all samples: the largest body force decides
        double gmax = 0.0;
        for (int s = 0; s < n_samples; s++) gmax = std::max(gmax, fabs(samples[s].gx));
        bcfg.fp16_scale_log2 = fp16_scale_log2_for(gmax);
    }
    rc = ctx_begin(b->c, &bcfg, samples[0].nx, samples[0].ny, samples[0].nz, 1);
    if (!rc) rc = batch_build(b, &bcfg, samples);
    if (!rc) rc = ctx_alloc_state(b->c);
    if (!rc) {
        const int bs = cfg->block_size;
        const int64_t nblocks = b->c->Npad / bs;
        b->block_sample.assign(nblocks, 0);
        std::vector<int64_t> end(n_samples);
        std::vector<double> om(n_samples), gx(n_samples);
        for (int s = 0; s < n_samples; s++) {
            const int64_t b0 = b->off[s] / bs;
            const int64_t b1 = (s + 1 < n_samples ? b->off[s + 1] : b->c->Npad) / bs;
            for (int64_t k = b0; k < b1; k++) b->block_sample[k] = s;
            end[s] = b->off[s] + b->n[s];
            om[s] = 1.0 / samples[s].tau;
            gx[s] = samples[s].gx;
            b->nx.push_back(samples[s].nx);
            b->ny.push_back(samples[s].ny);
            b->nz.push_back(samples[s].nz);
            b->tau.push_back(samples[s].tau);
            b->gx.push_back(samples[s].gx);
        }
        hipError_t e = hipMalloc(&b->d_block_sample, sizeof(int32_t) * nblocks);
        if (e == hipSuccess) e = hipMalloc(&b->d_active, sizeof(int32_t) * nblocks);
        if (e == hipSuccess) e = hipMalloc(&b->d_end, sizeof(int64_t) * n_samples);
        if (e == hipSuccess) e = hipMalloc(&b->d_off, sizeof(int64_t) * n_samples);
        if (e == hipSuccess) e = hipMalloc(&b->d_n, sizeof(int64_t) * n_samples);
        if (e == hipSuccess) e = hipMalloc(&b->d_omega, sizeof(double) * n_samples);
        if (e == hipSuccess) e = hipMalloc(&b->d_gx, sizeof(double) * n_samples);
        if (e == hipSuccess) e = hipMalloc(&b->d_sums, sizeof(double) * n_samples);
        if (e == hipSuccess) e = hipMalloc(&b->d_partial, sizeof(double) * nblocks);
        if (e == hipSuccess) e = hipMemset(b->d_partial, 0, sizeof(double) * nblocks);
        if (e == hipSuccess)
            e = hipMemcpy(b->d_block_sample, b->block_sample.data(), sizeof(int32_t) * nblocks,
                          hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(b->d_end, end.data(), sizeof(int64_t) * n_samples, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(b->d_off, b->off.data(), sizeof(int64_t) * n_samples,
                          hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(b->d_n, b->n.data(), sizeof(int64_t) * n_samples, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(b->d_omega, om.data(), sizeof(double) * n_samples, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(b->d_gx, gx.data(), sizeof(double) * n_samples, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            plb_set_hip_error(e, "plb_batch_create", __FILE__, __LINE__);
            rc = e == hipErrorOutOfMemory ? PLB_ERR_NOMEM : PLB_ERR_HIP;
        }
        b->sample_cur.assign(n_samples, 0);
        b->active.assign(n_samples, 1);
    }
    if (rc) {
        plb_batch_destroy(b);
        return rc;
    }
    *out = b;
    return PLB_OK;
}

extern "C" int64_t plb_batch_num_nodes(const plb_batch* b, int s)
{
    if (!b || s < 0 || s >= b->S) return -1;
    return b->n[s];
}

static int upload_active(plb_batch* b, int* n_active)
{
    std::vector<int32_t> list;
    const int bs = b->c->cfg.block_size;
    for (int s = 0; s < b->S; s++) {
        if (!b->active[s]) continue;
        const int64_t b0 = b->off[s] / bs;
        const int64_t b1 = (b->off[s] + b->n[s] + bs - 1) / bs;
        for (int64_t k = b0; k < b1; k++) list.push_back((int32_t)k);
    }
    *n_active = (int)list.size();
    if (!list.empty())
        PLB_HIP(hipMemcpyAsync(b->d_active, list.data(), sizeof(int32_t) * list.size(),
                               hipMemcpyHostToDevice, b->c->stream));
    PLB_HIP(hipStreamSynchronize(b->c->stream));
    return PLB_OK;
}

extern "C" int plb_batch_run(plb_batch* b, int64_t max_steps, int64_t check_every, double tol,
                             int64_t* steps_done, int64_t* updates, double* ms)
{
    if (!b || max_steps < 0 || check_every < 1 || !(tol >= 0.0)) return PLB_ERR_INVALID;
    Range range("plb_batch_run");
    plb_ctx* c = b->c;
    PLB_HIP(hipSetDevice(c->device));
    int rc = plb_init_equilibrium(c);
    if (rc) return rc;
    b->active.assign(b->S, 1);
    for (int s = 0; s < b->S; s++)
        if (b->n[s] == 0) b->active[s] = 0;
    std::vector<int64_t> done(b->S, 0);
    std::vector<double> prev(b->S, NAN), sums(b->S, 0.0);
    int n_active = 0;
    rc = upload_active(b, &n_active);
    if (rc) return rc;
    int64_t upd = 0;
    int64_t it = 0;
    PLB_HIP(hipEventRecord(c->ev0, c->stream));
    while (n_active > 0 && it < max_steps) {
        const int64_t m = std::min(check_every, max_steps - it);
        for (int64_t k = 0; k < m; k++) {
            rc = launch_step_batch(c, it + k, k == m - 1, b->d_active, n_active, b->d_block_sample,
                                   b->d_end, b->d_omega, b->d_gx);
            if (rc) return rc;
        }
        it += m;
        for (int s = 0; s < b->S; s++)
            if (b->active[s]) {
                done[s] += m;
                upd += m * b->n[s];
            }
        rc = launch_sample_sums(c, b->d_active, n_active, b->d_block_sample, b->d_end, b->d_off, b->d_n, b->S,
                                b->d_partial, b->d_sums);
        if (rc) return rc;
        unsigned long long key = NO_FAULT;
        PLB_HIP(hipMemcpyAsync(sums.data(), b->d_sums, sizeof(double) * b->S, hipMemcpyDeviceToHost,
                               c->stream));
        PLB_HIP(hipMemcpyAsync(&key, c->fault, sizeof key, hipMemcpyDeviceToHost, c->stream));
        PLB_HIP(hipStreamSynchronize(c->stream));
        if (key != NO_FAULT) return PLB_ERR_DENSITY;
        bool changed = false;
        for (int s = 0; s < b->S; s++) {
            if (!b->active[s]) continue;
            const double mean = sums[s] / (double)b->n[s];
            // relative change, with a floor of 1e-3 |g_x| on the scale so a
            // sample without a percolating path (mean u_x -> 0) also stops
            const double scale = fmax(fabs(mean), 1e-3 * fabs(b->gx[s]));
            if (!isnan(prev[s]) && fabs(mean - prev[s]) <= tol * scale) {
                b->active[s] = 0;
                b->sample_cur[s] = c->cur;
                changed = true;
            }
            prev[s] = mean;
        }
        if (changed) {
            rc = upload_active(b, &n_active);
            if (rc) return rc;
        }
    }
    for (int s = 0; s < b->S; s++)
        if (b->active[s]) b->sample_cur[s] = c->cur;
    PLB_HIP(hipEventRecord(c->ev1, c->stream));
    PLB_HIP(hipEventSynchronize(c->ev1));
    float t = 0.f;
    PLB_HIP(hipEventElapsedTime(&t, c->ev0, c->ev1));
    if (ms) *ms = t;
    if (steps_done)
        for (int s = 0; s < b->S; s++) steps_done[s] = done[s];
    if (updates) *updates = upd;
    return PLB_OK;
}

// The whole run as one persistent launch (k_batch_persistent): the same
// steps, check schedule, convergence rule and per-sample sums as
// plb_batch_run, with the checks made on the device.
extern "C" int plb_batch_run_persistent(plb_batch* b, int64_t max_steps, int64_t check_every, double tol,
                                        int64_t* steps_done, int64_t* updates, double* ms)
{
    if (!b || max_steps < 0 || check_every < 1 || !(tol >= 0.0)) return PLB_ERR_INVALID;
    if (b->S > PLB_PERSIST_MAX_SAMPLES) return PLB_ERR_INVALID;  // running flags live in LDS
    Range range("plb_batch_run_persistent");
    plb_ctx* c = b->c;
    PLB_HIP(hipSetDevice(c->device));
    int rc = plb_init_equilibrium(c);
    if (rc) return rc;
    const int S = b->S;
    const int bs = c->cfg.block_size;
    std::vector<int> act(S), cur(S, c->cur);
    std::vector<int64_t> steps(S, 0);
    std::vector<double> prev(S, NAN);
    int n_active = 0;
    for (int s = 0; s < S; s++) {
        act[s] = b->n[s] > 0 ? 1 : 0;
        n_active += act[s];
    }
    // per-run device state: active, cur, steps, prev, n_active, abort, barrier and claim counters
    DevBuf<int> d_act, d_cur, d_flags;
    DevBuf<int64_t> d_steps;
    DevBuf<double> d_prev;
    DevBuf<unsigned long long> d_bar;
    PLB_HIP(d_act.alloc(S));
    PLB_HIP(d_cur.alloc(S));
    PLB_HIP(d_steps.alloc(S));
    PLB_HIP(d_prev.alloc(S));
    PLB_HIP(d_flags.alloc(2));
    PLB_HIP(d_bar.alloc(34 * 16));
    const int flags[2] = {n_active, 0};
    PLB_HIP(hipMemcpyAsync(d_act, act.data(), sizeof(int) * S, hipMemcpyHostToDevice, c->stream));
    PLB_HIP(hipMemcpyAsync(d_cur, cur.data(), sizeof(int) * S, hipMemcpyHostToDevice, c->stream));
    PLB_HIP(hipMemcpyAsync(d_steps, steps.data(), sizeof(int64_t) * S, hipMemcpyHostToDevice, c->stream));
    PLB_HIP(hipMemcpyAsync(d_prev, prev.data(), sizeof(double) * S, hipMemcpyHostToDevice, c->stream));
    PLB_HIP(hipMemcpyAsync(d_flags, flags, sizeof flags, hipMemcpyHostToDevice, c->stream));
    PLB_HIP(hipMemsetAsync(d_bar, 0, sizeof(unsigned long long) * 34 * 16, c->stream));
    PersistArgs p;
    p.f[0] = c->f[0];
    p.f[1] = c->f[1];
    p.cur0 = c->cur;
    p.nchunks = c->Npad / bs;
    p.max_steps = max_steps;
    p.check_every = check_every;
    p.tol = tol;
    p.nsamples = S;
    p.chunk_sample = b->d_block_sample;
    p.sample_end = b->d_end;
    p.sample_off = b->d_off;
    p.sample_n = b->d_n;
    p.omega = b->d_omega;
    p.gx = b->d_gx;
    p.active = d_act;
    p.cur = d_cur;
    p.steps = d_steps;
    p.prev = d_prev;
    p.partial = b->d_partial;
    p.bar = d_bar;
    p.n_active = d_flags.p;
    p.abort = d_flags.p + 1;
    int clock_khz = 0;
    PLB_HIP(hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, c->device));
    if (clock_khz < 1000) clock_khz = 100000;  // no usable rate reported: the 100 MHz of gfx9
    p.wait_ticks = (uint64_t)clock_khz * 2000ull;  // 2 s of wall_clock64
    p.start_ticks = (uint64_t)clock_khz * 50ull;   // 50 ms
    int cus = 0;
    PLB_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    int blocks = 0, cap = 0;
    int hflags[2] = {0, 0};
    float t = 0.f;
    // A grid that is not all resident leaves before any work with abort = 2
    // (the occupancy figure is an estimate); it is launched again with one
    // block fewer per CU, with the barrier state cleared.
    for (;;) {
        PLB_HIP(hipEventRecord(c->ev0, c->stream));
        if (n_active > 0 && max_steps > 0) {
            rc = launch_batch_persistent(c, p, cap, &blocks);
            if (rc) return rc;
        }
        PLB_HIP(hipEventRecord(c->ev1, c->stream));
        PLB_HIP(hipEventSynchronize(c->ev1));
        PLB_HIP(hipEventElapsedTime(&t, c->ev0, c->ev1));
        PLB_HIP(hipMemcpy(hflags, d_flags, sizeof hflags, hipMemcpyDeviceToHost));
        if (hflags[1] != 2 || blocks <= cus) break;
        cap = blocks - cus;
        PLB_HIP(hipMemset(d_bar, 0, sizeof(unsigned long long) * 34 * 16));
        PLB_HIP(hipMemcpy(d_flags, flags, sizeof flags, hipMemcpyHostToDevice));
    }
    unsigned long long key = NO_FAULT;
    PLB_HIP(hipMemcpy(&key, c->fault, sizeof key, hipMemcpyDeviceToHost));
    PLB_HIP(hipMemcpy(cur.data(), d_cur, sizeof(int) * S, hipMemcpyDeviceToHost));
    PLB_HIP(hipMemcpy(steps.data(), d_steps, sizeof(int64_t) * S, hipMemcpyDeviceToHost));
    if (hflags[1]) {
        fprintf(stderr,
                hflags[1] == 2 ? "libplb: persistent batch run: the grid did not start whole (%d blocks, wall clock %d kHz)\n"
                               : "libplb: persistent batch run: a grid barrier wait exceeded 2 s (%d blocks, wall clock %d kHz)\n",
                blocks, clock_khz);
        return PLB_ERR_HIP;
    }
    if (key != NO_FAULT) return PLB_ERR_DENSITY;
    int64_t upd = 0;
    for (int s = 0; s < S; s++) {
        b->sample_cur[s] = cur[s];
        upd += steps[s] * b->n[s];
    }
    // the context's buffer parity after the last step any sample ran
    int64_t smax = 0;
    for (int s = 0; s < S; s++) smax = std::max(smax, steps[s]);
    c->cur = (int)((c->cur + smax) & 1);
    if (ms) *ms = t;
    if (steps_done)
        for (int s = 0; s < S; s++) steps_done[s] = steps[s];
    if (updates) *updates = upd;
    return PLB_OK;
}

extern "C" int plb_batch_last_fault(plb_batch* b, int* sample, plb_fault* out)
{
    if (!b || !sample || !out) return PLB_ERR_INVALID;
    plb_ctx* c = b->c;
    PLB_HIP(hipSetDevice(c->device));
    memset(out, 0, sizeof *out);
    *sample = -1;
    unsigned long long key = NO_FAULT;
    PLB_HIP(hipMemcpy(&key, c->fault, sizeof key, hipMemcpyDeviceToHost));
    if (key == NO_FAULT) return PLB_OK;
    // the key holds the node's V-order index in the concatenation
    const int64_t gc = (int64_t)(key & 0xffffffffull);
    int s = -1;
    for (int q = 0; q < b->S; q++)
        if (gc >= b->off[q] && gc < b->off[q] + b->n[q]) s = q;
    if (s < 0) return PLB_ERR_BOUNDS;
    const int64_t g = gc - b->off[s];  // V-order id within the sample
    uint32_t i = 0;                    // its internal index (the macros' position)
    PLB_HIP(hipMemcpy(&i, c->inv_perm + gc, sizeof i, hipMemcpyDeviceToHost));
    if ((int64_t)i >= c->Npad) return PLB_ERR_BOUNDS;
    uint64_t xyz = 0;  // node_xyz holds each sample's coordinates in its V order
    PLB_HIP(hipMemcpy(&xyz, c->node_xyz + b->off[s] + g, sizeof xyz, hipMemcpyDeviceToHost));
    PLB_HIP(hipMemcpy(&out->rho, c->macros + i, sizeof(double), hipMemcpyDeviceToHost));
    *sample = s;
    out->tripped = 1;
    out->step = (int64_t)(key >> 32);
    out->node = g;
    out->pid = 0;
    out->local_id = g;
    out->x = (int)(xyz & 0xffff);
    out->y = (int)((xyz >> 16) & 0xffff);
    out->z = (int)((xyz >> 32) & 0xffff);
    return PLB_OK;
}

extern "C" int plb_batch_read_macros(plb_batch* b, int sample, double* rho, double* ux, double* uy,
                                     double* uz)
{
    if (!b || sample < 0 || sample >= b->S) return PLB_ERR_INVALID;
    plb_ctx* c = b->c;
    PLB_HIP(hipSetDevice(c->device));
    const int64_t n = b->n[sample], off = b->off[sample];
    if (n == 0) return PLB_OK;
    double* outs[4] = {rho, ux, uy, uz};
    const int64_t chunk = std::min<int64_t>(n, (int64_t)8 << 20);
    double* d = nullptr;
    PLB_HIP(hipMalloc(&d, sizeof(double) * chunk));
    int rc = PLB_OK;
    for (int comp = 0; comp < 4 && !rc; comp++) {
        if (!outs[comp]) continue;
        for (int64_t k = 0; k < n && !rc; k += chunk) {
            const int64_t m = std::min(chunk, n - k);
            rc = launch_read_macros(c, comp, off + k, m, d);
            if (!rc && hipMemcpyAsync(outs[comp] + k, d, sizeof(double) * m, hipMemcpyDeviceToHost,
                                      c->stream) != hipSuccess)
                rc = PLB_ERR_HIP;
            if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = PLB_ERR_HIP;
        }
    }
    (void)hipFree(d);
    return rc;
}

extern "C" int plb_batch_reduce(plb_batch* b, plb_reduction* out)
{
    if (!b || !out) return PLB_ERR_INVALID;
    plb_ctx* c = b->c;
    PLB_HIP(hipSetDevice(c->device));
    double wsum = 0.0;
    for (int j = 0; j < Q; j++) wsum += TP[j] * c->cfg.rho0;
    for (int s = 0; s < b->S; s++) {
        double r[2] = {0.0, 0.0};
        if (b->n[s] > 0) {
            int rc = launch_reduce_range(c, b->sample_cur[s], b->off[s], b->n[s], true, r);
            if (rc) return rc;
        }
        plb_reduction& o = out[s];
        o.mass = r[0] + (c->cfg.precision != PLB_FP64 ? wsum * (double)b->n[s] : 0.0);
        o.sum_ux = r[1];
        o.n_nodes = b->n[s];
        o.mean_ux = b->n[s] ? r[1] / (double)b->n[s] : 0.0;
        o.q = r[1] / ((double)b->nx[s] * b->ny[s] * b->nz[s]);
        const double nu = 1.0 / 3.0 * (b->tau[s] - 0.5);
        o.K = b->gx[s] != 0.0 ? nu * o.q / b->gx[s] : 0.0;
    }
    return PLB_OK;
}

// Launch-shape search for a batch: every candidate gives the same per-sample
// results (the shapes change the node order inside each sample, the index
// encoding and the launch, never a source or a rounding), so only the time of
// `steps` fixed steps (tol 0: no sample stops) is compared.  Stage 1 builds
// one batch per index encoding x layout x node order at the base block size;
// stage 2 one batch per block size for the winner; stage 3 times the wave
// caps on the stage-2 winner's batch, since they change only the launch.
extern "C" int plb_batch_autotune(const plb_config* base, const plb_sample* samples, int n_samples, int steps,
                                  int persistent, plb_config* best, plb_tune_result* results, int max_results,
                                  int* n_results)
{
    if (!base || !samples || n_samples < 1 || !best || steps < 1 || max_results < 0) return PLB_ERR_INVALID;
    if (persistent && n_samples > PLB_PERSIST_MAX_SAMPLES) return PLB_ERR_INVALID;
    auto run = persistent ? plb_batch_run_persistent : plb_batch_run;
    Range range("plb_batch_autotune");
    int rc = check_cfg(base);
    if (rc) return rc;
    if (n_results) *n_results = 0;
    *best = *base;
    best->nodes_per_thread = 1;
    best->graph_steps = 0;
    double best_mlups = -1.0;
    int nres = 0;
    auto record = [&](const plb_config& cfg, int64_t upd, double ms) {
        const double m = ms > 0.0 ? (double)upd / (ms * 1e-3) / 1e6 : 0.0;
        if (results && nres < max_results) results[nres] = {cfg, m, ms};
        nres++;
        if (m > best_mlups) {
            best_mlups = m;
            *best = cfg;
        }
    };
    // One warm-up run, then the timed run; returns the open batch in *keep
    // when asked (stage 3 reuses it).
    auto time_batch = [&](const plb_config& cfg, plb_batch** keep) -> int {
        plb_batch* b = nullptr;
        int r = plb_batch_create(&b, &cfg, samples, n_samples);
        if (r) return r;
        int64_t upd = 0;
        double ms = 0.0;
        r = run(b, std::max(2, steps / 4), steps, 0.0, nullptr, &upd, &ms);
        if (!r) r = run(b, steps, steps, 0.0, nullptr, &upd, &ms);
        if (!r) record(cfg, upd, ms);
        if (r || !keep)
            plb_batch_destroy(b);
        else
            *keep = b;
        return r;
    };
    int max_nx = 1;
    for (int s = 0; s < n_samples; s++) max_nx = std::max(max_nx, samples[s].nx);
    // node orders: V order; pencils of 16, 32 and 64 full rows (a pencil
    // takes each sample's own x extent)
    const int rows[] = {0, 16, 32, 64};
    for (int idx : {PLB_INDEX_I32, PLB_INDEX_TILE8})
        for (int layout = 0; layout < 2; layout++)
            for (int py : rows) {
                plb_config cfg = *best;
                cfg.index_kind = idx;
                cfg.layout = layout;
                cfg.block_size = base->block_size;
                cfg.waves_per_simd = 0;
                cfg.ordering = py ? PLB_ORDER_PENCIL : PLB_ORDER_REFERENCE;
                if (py) {
                    cfg.brick[0] = max_nx;
                    cfg.brick[1] = py;
                    cfg.brick[2] = 1;
                }
                rc = time_batch(cfg, nullptr);
                if (rc) return rc;
            }
    const plb_config stage1 = *best;
    plb_batch* keep = nullptr;
    for (int bs : {128, 256, 512}) {
        if (bs == stage1.block_size) continue;
        plb_config cfg = stage1;
        cfg.block_size = bs;
        rc = time_batch(cfg, nullptr);
        if (rc) return rc;
    }
    const plb_config stage2 = *best;
    rc = time_batch(stage2, &keep);  // the winner's batch, timed again beside its wave caps
    if (rc) return rc;
    for (int waves : {6, 5, 4}) {
        plb_config cfg = stage2;
        cfg.waves_per_simd = waves;
        keep->c->cfg.waves_per_simd = waves;
        int64_t upd = 0;
        double ms = 0.0;
        rc = run(keep, steps, steps, 0.0, nullptr, &upd, &ms);
        if (rc) break;
        record(cfg, upd, ms);
    }
    plb_batch_destroy(keep);
    if (rc) return rc;
    if (n_results) *n_results = std::min(nres, max_results);
    return PLB_OK;
}
