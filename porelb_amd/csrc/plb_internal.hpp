// plb_internal.hpp -- shared definitions of libplb (gfx950).
//
// Device data layout (HBM):
//   populations  f[2]   two ping-pong buffers of Npad*19 storage elements
//                       (double, or float/half holding f - w_j*rho0)
//   pull table   tab    Npad*18 int32: >=0 source node, -1 bounce, -2 const
//                or TILE8: tdesc [T][18] int32 (base >= 0, or -(wide slot+1)),
//                toff [T][4][64] u32 + toff16 [T][64] u16, u8 offsets
//                (0..253 offset, 254 const, 255 bounce),
//                twide [W][64] int32 entries of the wide (tile, direction) pairs
//                or BITMAP: bm [words] {pore bits, pores before} per 32 voxels
//                of the node order, vox [Npad] u32 voxel position per node
//   macros              [4][Npad] double (rho, ux, uy, uz), written on
//                       request (last step of a call, convergence checks)
//   perm/inv_perm       internal id <-> V-order id, only for BRICK order
//   node_xyz            packed coordinates per internal id (x|y<<16|z<<32)
// Layouts: SOA = [dir][node], AOSOA = [node/64][dir][64].
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <rocprofiler-sdk-roctx/roctx.h>
#include <stdint.h>

#include <vector>

#include "plb.h"

namespace plb {

constexpr int Q = 19;

// D3Q19 in the reference's direction order (porelb.h:10-50).
constexpr int EX[Q] = {0, 1, -1, 0, 0, 0, 0, 0, 0, 0, 0, -1, 1, -1, 1, -1, 1, -1, 1};
constexpr int EY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, -1, 1};
constexpr int EZ[Q] = {0, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1, 0, 0, 0, 0};
constexpr int RE[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
constexpr double TP[Q] = {1.0 / 3,  1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18,
                          1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36,
                          1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};

constexpr int32_t TAB_BOUNCE = -1;
constexpr int32_t TAB_CONST = -2;
constexpr unsigned long long NO_FAULT = ~0ull;
constexpr int TILE8_WORDS = 5;  // words of packed u8 offsets per node: 4 u32 + one u16
constexpr int TILE8_FULL = 4;   // the u32 words; directions 17 and 18 live in the u16

// Position of population (dir j, node n).
template <int LAYOUT>
__host__ __device__ __forceinline__ int64_t fpos(int j, int64_t n, int64_t npad)
{
    if (LAYOUT == PLB_LAYOUT_SOA) return (int64_t)j * npad + n;
    return (n >> 6) * (Q * 64) + (int64_t)j * 64 + (n & 63);
}
// Position of pull-table entry (dir j = 1..18 stored at slot j-1, node n).
template <int LAYOUT>
__host__ __device__ __forceinline__ int64_t tpos(int jm1, int64_t n, int64_t npad)
{
    if (LAYOUT == PLB_LAYOUT_SOA) return (int64_t)jm1 * npad + n;
    return (n >> 6) * ((Q - 1) * 64) + (int64_t)jm1 * 64 + (n & 63);
}

// Work-groups are dispatched round-robin over the 8 XCDs of an MI355X, each
// with its own L2.  Remap so that every XCD walks one contiguous stretch of
// the node range: neighbouring tiles (which share cache lines of the gathered
// populations) then meet in the same L2.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t nb)
{
    const int64_t per = nb >> 3, rem = nb & 7;
    const int64_t x = b & 7, slot = b >> 3;
    return x * per + (x < rem ? x : rem) + slot;
}

// Division of a u32 by a run-time constant d (1 <= d < 2^31) with one
// high multiply: l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1,
// n / d = (t + ((n - t) >> min(l,1))) >> max(l-1,0) with t = mulhi(m, n)
// (Granlund & Montgomery, PLDI 1994, section 4).
struct FastDiv {
    uint32_t m = 1, s1 = 0, s2 = 0;
};
inline FastDiv make_fastdiv(uint32_t d)
{
    FastDiv f;
    uint32_t l = 0;
    while ((1ull << l) < d) l++;
    f.m = (uint32_t)(((1ull << 32) * ((1ull << l) - d)) / d + 1);
    f.s1 = l > 0 ? 1 : 0;
    f.s2 = l > 0 ? l - 1 : 0;
    return f;
}
__device__ __forceinline__ uint32_t fastdiv(uint32_t n, const FastDiv& f)
{
    const uint32_t t = __umulhi(f.m, n);
    return (t + ((n - t) >> f.s1)) >> f.s2;
}

__host__ __device__ __forceinline__ uint64_t pack_xyz(uint64_t x, uint64_t y, uint64_t z)
{
    return x | (y << 16) | (z << 32);
}

}  // namespace plb

struct plb_ctx {
    plb_config cfg;
    int device = 0;
    int lds_per_cu = 0;         // bytes of LDS per CU (occupancy cap of the step kernel)
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int64_t N = 0, Npad = 0;
    int nx = 0, ny = 0, nz = 0, num_p = 1;
    std::vector<int32_t> V_sub_start;
    size_t elem = 8;            // bytes per stored population
    void* f[2] = {nullptr, nullptr};
    int cur = 0;                // f[cur] holds the current post-collision state
    int32_t* tab = nullptr;     // internal layout (I32 index)
    int32_t* tdesc = nullptr;   // TILE8 index
    uint32_t* toff = nullptr;
    uint16_t* toff16 = nullptr;
    int32_t* twide = nullptr;
    int64_t n_wide = 0;
    uint2* bm = nullptr;        // BITMAP index: {pore bits, pores before} per 32 voxels
    uint32_t* vox = nullptr;    // BITMAP index: node -> voxel position in node order
    int py = 0;                 // BITMAP index: rows per pencil (ny for the V order)
    int64_t n_bm_words = 0;
    int64_t* it_dev = nullptr;  // first step number of a graph replay
    hipGraphExec_t gexec[2] = {nullptr, nullptr};  // G-step graphs by starting buffer
    double index_bytes = 72.0;  // pull-table bytes per node
    uint32_t* perm = nullptr;   // internal -> V order (null: identity)
    uint32_t* key_perm = nullptr;  // batches: internal -> V order within the node's sample
    // FP16 stochastic rounding: per 64-node tile, the rounding keys as up to
    // three runs of consecutive keys {base0, base1, base2, start1 | start2 << 8},
    // key = base_s + lane in run s; bit 31 of the last word: load key_perm/perm.
    uint4* kseg = nullptr;       // null: keys loaded per node
    double key_bytes = 0.0;     // rounding-key bytes per node update (FP16 stochastic)
    uint32_t* inv_perm = nullptr;
    uint64_t* node_xyz = nullptr;  // internal order
    double* macros = nullptr;   // [4][Npad]
    void* reserve = nullptr;    // placement reserve (cfg.placement_reserve_gib)
    unsigned long long* fault = nullptr;
    double* red = nullptr;      // reduction partials
    int red_blocks = 0;
    double omega = 0.0;
    float scale = 1.f, inv_scale = 1.f;  // FP16 storage scale 2^scale_log2
    int scale_log2 = 0;
    int64_t last_it = -1;       // last step issued
    int64_t pending_it0 = 0, pending_n = 0;
    int cur_before = 0;
};

// Error helpers (host).
#define PLB_HIP(expr)                                                   \
    do {                                                                \
        hipError_t e_ = (expr);                                         \
        if (e_ != hipSuccess) {                                         \
            plb_set_hip_error(e_, #expr, __FILE__, __LINE__);           \
            return e_ == hipErrorOutOfMemory ? PLB_ERR_NOMEM : PLB_ERR_HIP; \
        }                                                               \
    } while (0)

void plb_set_hip_error(hipError_t e, const char* what, const char* file, int line);

namespace plb {
// roctx range around a library call (shows up under rocprofv3 --marker-trace;
// a no-op without a profiler).
struct Range {
    explicit Range(const char* name) { roctxRangePushA(name); }
    ~Range() { roctxRangePop(); }
    Range(const Range&) = delete;
    Range& operator=(const Range&) = delete;
};

// Owner of a temporary device array: freed on every return path.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(int64_t n) { return hipMalloc((void**)&p, sizeof(T) * (size_t)(n > 0 ? n : 1)); }
    T* release()
    {
        T* q = p;
        p = nullptr;
        return q;
    }
    operator T*() const { return p; }
};
}  // namespace plb


// Internal entry points shared between translation units.
namespace plb {
// Arguments of the persistent batch kernel (plb_step.hip, k_batch_persistent).
struct PersistArgs {
    void* f[2];
    int cur0;                     // buffer holding every sample's state at step 0
    int64_t nchunks;
    int64_t max_steps, check_every;
    double tol;
    int nsamples;
    const int32_t* chunk_sample;  // chunk -> sample
    const int64_t* sample_end;    // first node after each sample
    const int64_t* sample_off;
    const int64_t* sample_n;
    const double* omega;
    const double* gx;
    int* active;                  // per sample: 1 while running
    int* cur;                     // per sample: buffer of its state after its last step
    int64_t* steps;               // per sample: steps run
    double* prev;                 // per sample: <u_x> at the previous check (NaN before)
    double* partial;              // per chunk: u_x sum on check steps
    unsigned long long* bar;      // grid-barrier, chunk-claim and start counters: 34 lines of 16 words, zeroed
    int* abort;                   // 1: a barrier wait timed out; 2: the grid was not all resident
    int* n_active;                // samples still running
    uint64_t wait_ticks;          // wall_clock64 ticks a barrier wait may take
    uint64_t start_ticks;         // ticks the first block waits for the whole grid to start
};
// max_blocks > 0 caps the grid below the computed residency.
int launch_batch_persistent(plb_ctx* c, const PersistArgs& p, int max_blocks, int* blocks);
int build_table_from_V(plb_ctx* c, const plb_node_info* V, int64_t n);
int build_table_from_flag(plb_ctx* c, const uint8_t* flag);
int launch_step(plb_ctx* c, int64_t it, bool write_macros, const int64_t* it_base = nullptr);
int launch_graph_steps(plb_ctx* c, int64_t it0, int64_t G);
int launch_reduce(plb_ctx* c, double* host_out2);
int exclusive_scan_u64(const uint64_t* in, uint64_t* out, int64_t n, hipStream_t s);
int finalize_table(plb_ctx* c, int32_t* tabref, const uint64_t* xyz_ref);
int build_ref_from_flag(const uint8_t* flag, int nx, int ny, int nz, int num_p, int mode,
                        hipStream_t s, int32_t** tabref_out, uint64_t** xyz_out, int64_t* n_out,
                        std::vector<int32_t>* start_out);
int copy_table_offset(int layout, const int32_t* tabref, int64_t n, int64_t off, int32_t* out,
                      int64_t npad, hipStream_t s);
int launch_step_batch(plb_ctx* c, int64_t it, bool write_macros, const int32_t* active,
                      int n_active, const int32_t* block_sample, const int64_t* sample_end,
                      const double* omega, const double* gx);
int launch_init_eq(plb_ctx* c);
int launch_read_pop(plb_ctx* c, int64_t first, int64_t n, double* dout);
int launch_write_pop(plb_ctx* c, int64_t first, int64_t n, const double* din);
int launch_read_macros(plb_ctx* c, int comp, int64_t first, int64_t n, double* dout);
int launch_copy_state(plb_ctx* c, int64_t first, int64_t n, void* aos, bool to_device);
int launch_read_table(plb_ctx* c, int64_t first, int64_t n, int32_t* dout);
int launch_reduce_range(plb_ctx* c, int buf, int64_t first, int64_t n, bool with_mass,
                        double* out2_host);
int launch_sample_sums(plb_ctx* c, const int32_t* d_active, int n_active, const int32_t* chunk_sample,
                       const int64_t* sample_end, const int64_t* d_off, const int64_t* d_n, int nsamples,
                       double* d_partial, double* d_out);
int check_cfg(const plb_config* cfg);
int fp16_scale_log2_for(double gx);
int compress_tile8(plb_ctx* c);
int build_key_runs(plb_ctx* c);
int build_bitmap(plb_ctx* c);
int compute_order(const plb_config& cfg, int nx, int ny, int nz, int64_t N, const uint64_t* xyz_ref,
                  hipStream_t s, uint32_t** inv_perm, uint32_t** perm, uint64_t* xyz_int);
int permute_ref_table(const int32_t* tabref, int64_t n, const uint32_t* inv_perm, int32_t* out,
                      hipStream_t s);
int ctx_begin(plb_ctx* c, const plb_config* cfg, int nx, int ny, int nz, int num_p);
int noalloc_setup(int device);
int ctx_alloc_state(plb_ctx* c);
}  // namespace plb
