// plb_tune.hip -- HBM copy ceiling and the launch-shape autotuner.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "plb_internal.hpp"
#include "plb_tune.h"

using namespace plb;

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// One 16-byte element per thread, non-temporal both ways: the fastest plain
// copy form measured on MI355X (grid-stride and unrolled forms are slower).
__global__ void __launch_bounds__(256) k_copy16(const u32x4* __restrict__ in, u32x4* __restrict__ out,
                                                int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(in + i), out + i);
}

extern "C" int plb_copy_bandwidth(int device, int64_t bytes, int iters, double* gbs)
{
    if (!gbs || bytes < 16 * 1024 || bytes >= ((int64_t)1 << 39) || iters < 1) return PLB_ERR_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return PLB_ERR_NODEVICE;
    if (device < 0 || device >= ndev) return PLB_ERR_INVALID;
    PLB_HIP(hipSetDevice(device));
    const int64_t n = bytes / 16;
    u32x4 *a = nullptr, *b = nullptr;
    PLB_HIP(hipMalloc(&a, n * 16));
    PLB_HIP(hipMalloc(&b, n * 16));
    PLB_HIP(hipMemset(a, 1, n * 16));
    hipEvent_t e0, e1;
    PLB_HIP(hipEventCreate(&e0));
    PLB_HIP(hipEventCreate(&e1));
    const unsigned grid = (unsigned)((n + 255) / 256);
    float best = 1e30f;
    for (int k = 0; k <= iters; k++) {
        PLB_HIP(hipEventRecord(e0, 0));
        hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, 0, a, b, n);
        PLB_HIP(hipEventRecord(e1, 0));
        PLB_HIP(hipEventSynchronize(e1));
        float t = 0.f;
        PLB_HIP(hipEventElapsedTime(&t, e0, e1));
        if (k > 0) best = std::min(best, t);
    }
    *gbs = 2.0 * (double)(n * 16) / (best * 1e-3) / 1e9;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return PLB_OK;
}

static uint64_t mask_hash(const uint8_t* flag, int64_t n, int nx, int ny, int nz)
{
    // FNV-1a over the voxels and the dimensions.
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](uint64_t v) {
        h ^= v;
        h *= 1099511628211ull;
    };
    mix((uint64_t)nx);
    mix((uint64_t)ny);
    mix((uint64_t)nz);
    const uint64_t* w = (const uint64_t*)flag;
    const int64_t nw = n / 8;
    for (int64_t i = 0; i < nw; i++) {
        uint64_t v;
        memcpy(&v, w + i, 8);
        mix(v);
    }
    for (int64_t i = nw * 8; i < n; i++) mix(flag[i]);
    return h;
}

static bool cache_lookup(const char* path, uint64_t key, int prec, int bb, plb_config* best)
{
    FILE* fp = fopen(path, "r");
    if (!fp) return false;
    char line[512];
    bool found = false;
    while (fgets(line, sizeof line, fp)) {
        unsigned long long k;
        int p, b, layout, order, bx, by, bz, bs, npt, idx, waves;
        if (sscanf(line, "%llx %d %d %d %d %d %d %d %d %d %d %d", &k, &p, &b, &idx, &layout, &order, &bx,
                   &by, &bz, &bs, &npt, &waves) == 12 &&
            k == key && p == prec && b == bb) {
            best->index_kind = idx;
            best->layout = layout;
            best->ordering = order;
            best->brick[0] = bx;
            best->brick[1] = by;
            best->brick[2] = bz;
            best->block_size = bs;
            best->nodes_per_thread = npt;
            best->waves_per_simd = waves;
            found = true;  // the last matching line wins
        }
    }
    fclose(fp);
    return found;
}

extern "C" int plb_autotune(const plb_config* base, const uint8_t* flag, int nx, int ny, int nz,
                            int num_p, int steps, const char* cache_path, plb_config* best,
                            plb_tune_result* results, int max_results, int* n_results)
{
    if (!base || !flag || !best || steps < 1) return PLB_ERR_INVALID;
    Range range("plb_autotune");
    int rc = check_cfg(base);
    if (rc) return rc;
    if (n_results) *n_results = 0;
    *best = *base;
    const int64_t nvox = (int64_t)nx * ny * nz;
    // AA streaming is tuned and cached apart from the pull mode
    const uint64_t key = mask_hash(flag, nvox, nx, ny, nz) ^
                         (base->stream_mode == PLB_STREAM_AA ? 0x9e3779b97f4a7c15ull : 0ull);
    if (cache_path && cache_lookup(cache_path, key, base->precision, base->bb_mode, best))
        return PLB_OK;

    // Stage 1: index encoding x layout x node order at 256 threads, one node
    // per thread, no occupancy cap.  Stage 2: block size x nodes per thread x
    // waves per SIMD for the winner.
    // Every index encoding gives the same sources, so the choice changes speed
    // only, never results.
    double best_mlups = -1.0;
    int nres = 0;
    auto time_cfg = [&](const plb_config& cfg, double* mlups) -> int {
        plb_ctx* c = nullptr;
        int r = plb_create_from_flag(&c, &cfg, flag, nx, ny, nz, num_p);
        if (r) return r;
        double ms = 0.0;
        r = plb_step(c, 0, std::max(2, steps / 4));
        if (!r) r = plb_step_timed(c, 0, steps, &ms);
        const double n = (double)plb_num_nodes(c);
        plb_destroy(c);
        if (r) return r;
        *mlups = n * steps / (ms * 1e-3) / 1e6;
        if (results && nres < max_results) results[nres] = {cfg, *mlups, ms};
        nres++;
        if (*mlups > best_mlups) {
            best_mlups = *mlups;
            *best = cfg;
        }
        return PLB_OK;
    };
    // node orders: reference; 8^3 bricks; pencils of 64 and 128 full rows
    const int orders[][4] = {{PLB_ORDER_REFERENCE, 0, 0, 0},
                             {PLB_ORDER_BRICK, 8, 8, 8},
                             {PLB_ORDER_PENCIL, nx, 64, 1},
                             {PLB_ORDER_PENCIL, nx, 128, 1}};
    std::vector<int> kinds = {PLB_INDEX_I32, PLB_INDEX_TILE8};
    if (base->bb_mode == PLB_BB_HALFWAY) kinds.push_back(PLB_INDEX_BITMAP);
    for (int idx : kinds)
        for (int layout = 0; layout < 2; layout++)
            for (const auto& o : orders) {
                if (idx == PLB_INDEX_BITMAP && o[0] == PLB_ORDER_BRICK) continue;
                plb_config cfg = *base;
                cfg.index_kind = idx;
                cfg.layout = layout;
                cfg.ordering = o[0];
                if (o[0] != PLB_ORDER_REFERENCE) {
                    cfg.brick[0] = o[1];
                    cfg.brick[1] = o[2];
                    cfg.brick[2] = o[3];
                }
                cfg.block_size = 256;
                cfg.nodes_per_thread = 1;
                cfg.graph_steps = 0;
                cfg.waves_per_simd = 0;
                double m = 0.0;
                rc = time_cfg(cfg, &m);
                if (rc) return rc;
            }
    // Stage 2 changes only the launch (block size, nodes per thread, waves
    // per SIMD), so every candidate runs on one context: the comparison is
    // free of the allocation-placement spread between contexts.
    const plb_config stage1 = *best;
    plb_ctx* c = nullptr;
    rc = plb_create_from_flag(&c, &stage1, flag, nx, ny, nz, num_p);
    if (rc) return rc;
    const double n_nodes = (double)plb_num_nodes(c);
    int64_t it = 0;
    for (int bs : {128, 256, 512})
        for (int npt : {1, 2})
            for (int waves : {0, 6, 5, 4}) {
                if (npt == 2 && stage1.stream_mode == PLB_STREAM_AA) continue;  // AA: one node per thread
                plb_config cfg = stage1;
                cfg.block_size = bs;
                cfg.nodes_per_thread = npt;
                cfg.waves_per_simd = waves;
                c->cfg = cfg;
                double ms = 0.0;
                const int warm = std::max(2, steps / 4);
                rc = plb_step(c, it, warm);
                if (!rc) rc = plb_step_timed(c, it + warm, steps, &ms);
                it += warm + steps;
                if (rc) {
                    plb_destroy(c);
                    return rc;
                }
                const double m = n_nodes * steps / (ms * 1e-3) / 1e6;
                if (results && nres < max_results) results[nres] = {cfg, m, ms};
                nres++;
                if (m > best_mlups) {
                    best_mlups = m;
                    *best = cfg;
                }
            }
    plb_destroy(c);
    if (n_results) *n_results = std::min(nres, max_results);
    if (cache_path) {
        FILE* fp = fopen(cache_path, "a");
        if (fp) {
            fprintf(fp, "%016llx %d %d %d %d %d %d %d %d %d %d %d %.1f\n", (unsigned long long)key,
                    base->precision, base->bb_mode, best->index_kind, best->layout, best->ordering,
                    best->brick[0], best->brick[1], best->brick[2], best->block_size,
                    best->nodes_per_thread, best->waves_per_simd, best_mlups);
            fclose(fp);
        }
    }
    return PLB_OK;
}
