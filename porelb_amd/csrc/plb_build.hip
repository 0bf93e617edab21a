// plb_build.hip -- sparse index build on the device.
//
// Replaces, for the one-GPU case, three pieces of the reference:
//  * preprocess.c build_V (preprocess.c:29-164): pore compaction in z,y,x
//    order, z-slab rank = z/(nz/num_p), periodic neighbour classification;
//  * the per-step decoding of 152-byte NODE_INFO records in
//    stream_inner_nodes (porelb.c:285-300);
//  * the halo bookkeeping of build_send_recv_info / update_recv_dist
//    (porelb.c:143-222, 302-312): a link to another rank becomes a global
//    node id in the pull table.
//
// Pull-table rule for slot (i, j), j >= 1, with dn = neighbour at x+e_j and
// up = neighbour at x-e_j (V record entries nb[j-1] and nb[re[j]-1]):
//   REFCOMPAT: up pore and (dn pore, or up on another rank, or up after i in
//              its rank's local order)       -> stream from up
//              else dn solid                 -> bounce (i, re[j])
//              else                          -> constant tp[j]*rho0
//   HALFWAY:   up pore -> stream from up, else bounce (i, re[j])
// REFCOMPAT reproduces the last-writer-wins order of the reference's push
// stream, its bounce rule f1[i][j] = f0[i][re[j]] (porelb.c:296-297), and
// the receive that runs after the stream loop (porelb.c:428-433).
//
// Voxel masks are held as one bit per voxel (wave-wide __ballot over 64
// voxels) plus an exclusive prefix of per-word pore counts, so the pore id
// of any voxel is prefix[w] + popcount(bits[w] & below) -- 1.5 bits per
// voxel instead of the 4-byte X_to_n array of preprocess.c:50.
#include <stdio.h>

#include <algorithm>

#include "plb_internal.hpp"

namespace plb {

// ---------------------------------------------------------------- scan
constexpr int SCAN_T = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int64_t SCAN_PER_BLOCK = SCAN_T * SCAN_ITEMS;

__global__ void __launch_bounds__(SCAN_T)
k_scan_block(const uint64_t* __restrict__ in, uint64_t* __restrict__ out, uint64_t* bsum,
             int64_t n)
{
    __shared__ uint64_t s[SCAN_T];
    const int tid = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * SCAN_PER_BLOCK + (int64_t)tid * SCAN_ITEMS;
    uint64_t v[SCAN_ITEMS];
    uint64_t sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        sum += v[k];
    }
    s[tid] = sum;
    __syncthreads();
    for (int off = 1; off < SCAN_T; off <<= 1) {
        uint64_t t = tid >= off ? s[tid - off] : 0;
        __syncthreads();
        s[tid] += t;
        __syncthreads();
    }
    uint64_t excl = s[tid] - sum;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++) {
        if (base + k < n) out[base + k] = excl;
        excl += v[k];
    }
    if (bsum && tid == SCAN_T - 1) bsum[blockIdx.x] = s[tid];
}

__global__ void k_scan_add(uint64_t* out, const uint64_t* boff, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += boff[i / SCAN_PER_BLOCK];
}

int exclusive_scan_u64(const uint64_t* in, uint64_t* out, int64_t n, hipStream_t s)
{
    if (n <= 0) return PLB_OK;
    int64_t nb = (n + SCAN_PER_BLOCK - 1) / SCAN_PER_BLOCK;
    if (nb == 1) {
        hipLaunchKernelGGL(k_scan_block, dim3(1), dim3(SCAN_T), 0, s, in, out, nullptr, n);
        PLB_HIP(hipGetLastError());
        return PLB_OK;
    }
    DevBuf<uint64_t> bsum, boff;
    PLB_HIP(bsum.alloc(nb));
    PLB_HIP(boff.alloc(nb));
    hipLaunchKernelGGL(k_scan_block, dim3((unsigned)nb), dim3(SCAN_T), 0, s, in, out, bsum.p, n);
    PLB_HIP(hipGetLastError());
    int rc = exclusive_scan_u64(bsum, boff, nb, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, out, boff.p,
                       n);
    PLB_HIP(hipGetLastError());
    PLB_HIP(hipStreamSynchronize(s));
    return PLB_OK;
}

// ---------------------------------------------------------------- bitmask helpers
__device__ __forceinline__ bool is_pore(const uint64_t* bits, int64_t v)
{
    return (bits[v >> 6] >> (v & 63)) & 1ull;
}
__device__ __forceinline__ int64_t pore_id(const uint64_t* bits, const uint64_t* wpre, int64_t v)
{
    const int64_t w = v >> 6;
    const uint64_t below = (1ull << (v & 63)) - 1ull;
    return (int64_t)wpre[w] + __popcll(bits[w] & below);
}

// flag (u8, x fastest) -> bits + per-word counts.  blockDim multiple of 64.
__global__ void k_flag_bits(const uint8_t* __restrict__ flag, int64_t nvox, int64_t nwords,
                            uint64_t* __restrict__ bits, uint64_t* __restrict__ wcount, int* err)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint8_t f = v < nvox ? flag[v] : 0;
    if (f > 1) atomicCAS(err, 0, PLB_ERR_FLAG);
    const uint64_t m = __ballot(f == 1);
    if ((threadIdx.x & 63) == 0) {
        const int64_t w = v >> 6;
        if (w < nwords) {
            bits[w] = m;
            wcount[w] = __popcll(m);
        }
    }
}

__global__ void k_nodes_from_bits(const uint64_t* __restrict__ bits,
                                  const uint64_t* __restrict__ wpre, int64_t nvox, int nx, int ny,
                                  uint64_t* __restrict__ xyz)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox || !is_pore(bits, v)) return;
    const int64_t id = pore_id(bits, wpre, v);
    const int64_t plane = (int64_t)nx * ny;
    const int64_t z = v / plane, r = v - z * plane;
    const int64_t y = r / nx, x = r - y * nx;
    xyz[id] = pack_xyz(x, y, z);
}

struct GridArgs {
    const uint64_t* bits;
    const uint64_t* wpre;
    const uint64_t* xyz;
    int64_t N;
    int nx, ny, nz, slab;
};

__device__ __forceinline__ void unpack(uint64_t p, int& x, int& y, int& z)
{
    x = (int)(p & 0xffff);
    y = (int)((p >> 16) & 0xffff);
    z = (int)((p >> 32) & 0xffff);
}
__device__ __forceinline__ int64_t lin(const GridArgs& a, int x, int y, int z)
{
    return ((int64_t)z * a.ny + y) * a.nx + x;
}
__device__ __forceinline__ int wrap(int v, int n) { return (v + n) % n; }

// Pull table straight from the voxel bitmask (V order, [18][N]).
__global__ void k_table_from_bits(GridArgs a, int mode, int32_t* __restrict__ tab)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.N) return;
    int x, y, z;
    unpack(a.xyz[g], x, y, z);
    const int p = z / a.slab;
#pragma unroll
    for (int j = 1; j < Q; j++) {
        const int64_t vd = lin(a, wrap(x + EX[j], a.nx), wrap(y + EY[j], a.ny), wrap(z + EZ[j], a.nz));
        const int zu = wrap(z - EZ[j], a.nz);
        const int64_t vu = lin(a, wrap(x - EX[j], a.nx), wrap(y - EY[j], a.ny), zu);
        const bool up = is_pore(a.bits, vu);
        int32_t t;
        if (mode == PLB_BB_HALFWAY) {
            t = up ? (int32_t)pore_id(a.bits, a.wpre, vu) : TAB_BOUNCE;
        } else {
            const bool dn = is_pore(a.bits, vd);
            const int64_t uid = up ? pore_id(a.bits, a.wpre, vu) : -1;
            if (up && (dn || (zu / a.slab) != p || uid > g))
                t = (int32_t)uid;
            else if (!dn)
                t = TAB_BOUNCE;
            else
                t = TAB_CONST;
        }
        tab[(int64_t)(j - 1) * a.N + g] = t;
    }
}

// V records from the voxel bitmask (device preprocess; preprocess.c:124-157).
__global__ void k_V_from_bits(GridArgs a, const int32_t* __restrict__ start,
                              uint64_t* __restrict__ Vw)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.N) return;
    int x, y, z;
    unpack(a.xyz[g], x, y, z);
    const int p = z / a.slab;
    uint64_t* rec = Vw + g * 19;
    rec[0] = pack_xyz(x, y, z);
#pragma unroll
    for (int j = 1; j < Q; j++) {
        const int zd = wrap(z + EZ[j], a.nz);
        const int64_t vd = lin(a, wrap(x + EX[j], a.nx), wrap(y + EY[j], a.ny), zd);
        uint64_t w;
        if (!is_pore(a.bits, vd)) {
            w = (uint64_t)PLB_TYPE_SOLID;
        } else {
            const int pp = zd / a.slab;
            const int64_t local = pore_id(a.bits, a.wpre, vd) - start[pp];
            const uint64_t type = (pp == p) ? PLB_TYPE_FLUID : PLB_TYPE_EXTERN;
            w = type | ((uint64_t)pp << 16) | ((uint64_t)local << 32);
        }
        rec[j] = w;
    }
}

__global__ void k_slab_starts(GridArgs a, int num_p, int32_t* start)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= num_p) return;
    const int64_t v = (int64_t)p * a.slab * a.nx * a.ny;
    start[p] = (int32_t)pore_id(a.bits, a.wpre, v);
}

// ---------------------------------------------------------------- from V records
struct VArgs {
    const uint64_t* V;  // records as 19 words
    int64_t N;
    const int32_t* start;
    int num_p;
    int nx, ny, nz;
    int mode;
    int32_t* tab;
    uint64_t* xyz;
    int* err;
};

__device__ __forceinline__ int owner_rank(const int32_t* start, int num_p, int64_t g)
{
    int lo = 0, hi = num_p - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((int64_t)start[mid] <= g)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__global__ void k_table_from_V(VArgs a)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.N) return;
    const uint64_t* rec = a.V + g * 19;
    const uint64_t w0 = rec[0];
    const int x = (int)(w0 & 0xffff), y = (int)((w0 >> 16) & 0xffff), z = (int)((w0 >> 32) & 0xffff);
    // check_V (porelb.c:115-141), applied to every rank's records.
    if (x >= a.nx || y >= a.ny || z >= a.nz) atomicCAS(a.err, 0, PLB_ERR_BOUNDS);
    a.xyz[g] = pack_xyz(x, y, z);
    const int p = owner_rank(a.start, a.num_p, g);
    const int64_t k = g - a.start[p];
    auto count = [&](int q) -> int64_t {
        return (q + 1 < a.num_p ? (int64_t)a.start[q + 1] : a.N) - a.start[q];
    };
    uint64_t nb[Q];
#pragma unroll
    for (int j = 1; j < Q; j++) nb[j] = rec[j];
    // Validate every entry once.
#pragma unroll
    for (int j = 1; j < Q; j++) {
        const int type = (int)(nb[j] & 0xff);
        const int pid = (int)((nb[j] >> 16) & 0xffff);
        const int64_t id = (int64_t)(nb[j] >> 32);
        if (type == PLB_TYPE_EXTERN) {
            if (pid > a.num_p - 1 || pid == p) atomicCAS(a.err, 0, PLB_ERR_PID);
            else if (id >= count(pid)) atomicCAS(a.err, 0, PLB_ERR_BOUNDS);
        } else if (type == PLB_TYPE_FLUID) {
            if (id >= count(p)) atomicCAS(a.err, 0, PLB_ERR_BOUNDS);
        } else if (type != PLB_TYPE_SOLID) {
            atomicCAS(a.err, 0, PLB_ERR_INVALID);
        }
    }
#pragma unroll
    for (int j = 1; j < Q; j++) {
        const uint64_t dn = nb[j], up = nb[RE[j]];
        const int dtype = (int)(dn & 0xff), utype = (int)(up & 0xff);
        const int upid = (int)((up >> 16) & 0xffff);
        const int64_t uid = (int64_t)(up >> 32);
        int64_t ug = -1;
        if (utype == PLB_TYPE_FLUID) ug = a.start[p] + uid;
        if (utype == PLB_TYPE_EXTERN && upid < a.num_p) ug = a.start[upid] + uid;
        if (ug >= a.N) ug = -1;  // invalid entries were flagged above
        int32_t t;
        if (a.mode == PLB_BB_HALFWAY) {
            t = (utype != PLB_TYPE_SOLID && ug >= 0) ? (int32_t)ug : TAB_BOUNCE;
        } else {
            if (utype != PLB_TYPE_SOLID && ug >= 0 &&
                (dtype != PLB_TYPE_SOLID || utype == PLB_TYPE_EXTERN || uid > k))
                t = (int32_t)ug;
            else if (dtype == PLB_TYPE_SOLID)
                t = TAB_BOUNCE;
            else
                t = TAB_CONST;
        }
        a.tab[(int64_t)(j - 1) * a.N + g] = t;
    }
}

// ---------------------------------------------------------------- ordering + layout
struct BrickArgs {
    int nx, ny, nz, bx, by, bz, nbx, nby, nbz;
    int zfast;  // PENCIL: bricks visited z fastest, then y, then x
};
__device__ __forceinline__ int64_t brick_lin(const BrickArgs& b, int x, int y, int z)
{
    const int64_t brick = b.zfast ? ((int64_t)(x / b.bx) * b.nby + (y / b.by)) * b.nbz + (z / b.bz)
                                  : ((int64_t)(z / b.bz) * b.nby + (y / b.by)) * b.nbx + (x / b.bx);
    const int64_t inner = ((int64_t)(z % b.bz) * b.by + (y % b.by)) * b.bx + (x % b.bx);
    return brick * ((int64_t)b.bx * b.by * b.bz) + inner;
}

__global__ void k_brick_bits(const uint64_t* __restrict__ xyz, int64_t N, BrickArgs b,
                             unsigned long long* bits)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= N) return;
    int x, y, z;
    unpack(xyz[g], x, y, z);
    const int64_t v = brick_lin(b, x, y, z);
    atomicOr(bits + (v >> 6), 1ull << (v & 63));
}

__global__ void k_popc_words(const uint64_t* bits, uint64_t* cnt, int64_t nw)
{
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w < nw) cnt[w] = __popcll(bits[w]);
}

__global__ void k_brick_ids(const uint64_t* __restrict__ xyz_ref, int64_t N, BrickArgs b,
                            const uint64_t* bits, const uint64_t* wpre, uint32_t* inv_perm,
                            uint32_t* perm, uint64_t* xyz_int)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= N) return;
    int x, y, z;
    const uint64_t p = xyz_ref[g];
    unpack(p, x, y, z);
    const int64_t k = pore_id(bits, wpre, brick_lin(b, x, y, z));
    inv_perm[g] = (uint32_t)k;
    perm[k] = (uint32_t)g;
    xyz_int[k] = p;
}

template <int LAYOUT>
__global__ void k_write_table(const int32_t* __restrict__ tabref, int64_t N,
                              const uint32_t* __restrict__ inv_perm, int32_t* __restrict__ out,
                              int64_t npad)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= N) return;
    const int64_t k = inv_perm ? (int64_t)inv_perm[g] : g;
#pragma unroll
    for (int j = 0; j < Q - 1; j++) {
        int32_t t = tabref[(int64_t)j * N + g];
        if (t >= 0 && inv_perm) t = (int32_t)inv_perm[t];
        out[tpos<LAYOUT>(j, k, npad)] = t;
    }
}

static unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Node order of a lattice: inv_perm[g] = internal id of V-order node g,
// perm = its inverse, xyz_int = coordinates in internal order.  For the
// reference order nothing is allocated (identity).
int compute_order(const plb_config& cfg, int nx, int ny, int nz, int64_t N, const uint64_t* xyz_ref,
                  hipStream_t s, uint32_t** inv_perm, uint32_t** perm, uint64_t* xyz_int)
{
    *inv_perm = nullptr;
    if (perm) *perm = nullptr;
    if (cfg.ordering == PLB_ORDER_REFERENCE || N == 0) {
        if (xyz_int && N > 0)
            PLB_HIP(hipMemcpyAsync(xyz_int, xyz_ref, sizeof(uint64_t) * N, hipMemcpyDeviceToDevice, s));
        return PLB_OK;
    }
    BrickArgs b;
    b.zfast = cfg.ordering == PLB_ORDER_PENCIL ? 1 : 0;
    b.nx = nx;
    b.ny = ny;
    b.nz = nz;
    // a brick edge longer than the domain only pads the bitmask
    b.bx = std::min(cfg.brick[0], nx);
    b.by = std::min(cfg.brick[1], ny);
    b.bz = std::min(cfg.brick[2], nz);
    b.nbx = (nx + b.bx - 1) / b.bx;
    b.nby = (ny + b.by - 1) / b.by;
    b.nbz = (nz + b.bz - 1) / b.bz;
    const int64_t nvb = (int64_t)b.nbx * b.nby * b.nbz * b.bx * b.by * b.bz;
    const int64_t nw = (nvb + 63) / 64;
    DevBuf<uint64_t> bits, cnt, wpre, xyz_tmp;
    DevBuf<uint32_t> inv, p;
    PLB_HIP(bits.alloc(nw));
    PLB_HIP(cnt.alloc(nw));
    PLB_HIP(wpre.alloc(nw));
    PLB_HIP(hipMemsetAsync(bits, 0, sizeof(uint64_t) * nw, s));
    hipLaunchKernelGGL(k_brick_bits, dim3(nblk(N, 256)), dim3(256), 0, s, xyz_ref, N, b,
                       (unsigned long long*)bits.p);
    hipLaunchKernelGGL(k_popc_words, dim3(nblk(nw, 256)), dim3(256), 0, s, bits.p, cnt.p, nw);
    PLB_HIP(hipGetLastError());
    int rc = exclusive_scan_u64(cnt, wpre, nw, s);
    if (rc) return rc;
    // Every node must own a voxel of its own: two records with the same
    // coordinates would get the same internal id, race on its table row and
    // leave another id unassigned.  The set bits count the distinct voxels.
    uint64_t last[2] = {0, 0};
    PLB_HIP(hipMemcpyAsync(&last[0], wpre.p + nw - 1, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    PLB_HIP(hipMemcpyAsync(&last[1], cnt.p + nw - 1, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    PLB_HIP(hipStreamSynchronize(s));
    if ((int64_t)(last[0] + last[1]) != N) return PLB_ERR_BOUNDS;
    PLB_HIP(inv.alloc(N));
    PLB_HIP(p.alloc(N));
    if (!xyz_int) PLB_HIP(xyz_tmp.alloc(N));
    hipLaunchKernelGGL(k_brick_ids, dim3(nblk(N, 256)), dim3(256), 0, s, xyz_ref, N, b, bits.p, wpre.p,
                       inv.p, p.p, xyz_int ? xyz_int : xyz_tmp.p);
    PLB_HIP(hipGetLastError());
    PLB_HIP(hipStreamSynchronize(s));
    *inv_perm = inv.release();
    if (perm) *perm = p.release();
    return PLB_OK;
}

// V-order table [18][n] -> the same table in internal order with internal
// ids, still [18][n] SoA (used for batch samples).
int permute_ref_table(const int32_t* tabref, int64_t n, const uint32_t* inv_perm, int32_t* out,
                      hipStream_t s)
{
    if (n == 0) return PLB_OK;
    hipLaunchKernelGGL(k_write_table<PLB_LAYOUT_SOA>, dim3(nblk(n, 256)), dim3(256), 0, s, tabref, n,
                       inv_perm, out, n);
    PLB_HIP(hipGetLastError());
    return PLB_OK;
}

int finalize_table(plb_ctx* c, int32_t* tabref, const uint64_t* xyz_ref)
{
    const int64_t N = c->N;
    hipStream_t s = c->stream;
    const int64_t tab_elems = (Q - 1) * c->Npad;
    PLB_HIP(hipMalloc(&c->tab, sizeof(int32_t) * tab_elems));
    PLB_HIP(hipMemsetAsync(c->tab, 0, sizeof(int32_t) * tab_elems, s));
    PLB_HIP(hipMalloc(&c->node_xyz, sizeof(uint64_t) * (N ? N : 1)));
    int rc = compute_order(c->cfg, c->nx, c->ny, c->nz, N, xyz_ref, s, &c->inv_perm, &c->perm,
                           c->node_xyz);
    if (rc) return rc;
    if (N > 0) {
        if (c->cfg.layout == PLB_LAYOUT_SOA)
            hipLaunchKernelGGL(k_write_table<PLB_LAYOUT_SOA>, dim3(nblk(N, 256)), dim3(256), 0, s,
                               tabref, N, c->inv_perm, c->tab, c->Npad);
        else
            hipLaunchKernelGGL(k_write_table<PLB_LAYOUT_AOSOA>, dim3(nblk(N, 256)), dim3(256), 0,
                               s, tabref, N, c->inv_perm, c->tab, c->Npad);
        PLB_HIP(hipGetLastError());
    }
    PLB_HIP(hipStreamSynchronize(s));
    c->index_bytes = (Q - 1) * 4.0;
    if (c->cfg.index_kind == PLB_INDEX_TILE8) return compress_tile8(c);
    if (c->cfg.index_kind == PLB_INDEX_BITMAP) return build_bitmap(c);
    return PLB_OK;
}

// ---------------------------------------------------------------- TILE8 index
// One 64-thread block per tile: wave-wide min/max of the stream sources of
// each direction decide between a u8 offset from a base and a wide (int32)
// slot.
constexpr int TILE8_SPAN = 253;

__device__ __forceinline__ int32_t wave_min(int32_t v)
{
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v = min(v, __shfl_xor(v, m, 64));
    return v;
}
__device__ __forceinline__ int32_t wave_max(int32_t v)
{
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v = max(v, __shfl_xor(v, m, 64));
    return v;
}

template <int LAYOUT>
__device__ __forceinline__ void tile_span(const int32_t* tab, int64_t Npad, int64_t N, int64_t n,
                                          int jm1, int32_t& t, int32_t& mn, int32_t& mx)
{
    t = n < N ? tab[tpos<LAYOUT>(jm1, n, Npad)] : TAB_CONST;
    mn = wave_min(t >= 0 ? t : INT32_MAX);
    mx = wave_max(t >= 0 ? t : -1);
}

template <int LAYOUT>
__global__ void __launch_bounds__(64)
k_tile8_flags(const int32_t* __restrict__ tab, int64_t Npad, int64_t N, uint64_t* __restrict__ wide)
{
    const int64_t tile = blockIdx.x;
    const int64_t n = tile * 64 + threadIdx.x;
    for (int jm1 = 0; jm1 < Q - 1; jm1++) {
        int32_t t, mn, mx;
        tile_span<LAYOUT>(tab, Npad, N, n, jm1, t, mn, mx);
        if (threadIdx.x == 0) wide[tile * (Q - 1) + jm1] = (mx >= 0 && mx - mn > TILE8_SPAN) ? 1 : 0;
    }
}

// Offsets are packed four directions per 32-bit word: word k of a lane holds
// directions 4k+1..4k+4 in bytes 0..3 ([tile][4][64] words), and directions
// 17 and 18 sit in a u16 per lane ([tile][64]): 18 bytes per node, read with
// four dword loads and one short load.
template <int LAYOUT>
__global__ void __launch_bounds__(64)
k_tile8_write(const int32_t* __restrict__ tab, int64_t Npad, int64_t N,
              const uint64_t* __restrict__ wide, const uint64_t* __restrict__ slot,
              int32_t* __restrict__ desc, uint32_t* __restrict__ off, uint16_t* __restrict__ off16,
              int32_t* __restrict__ wtab)
{
    const int64_t tile = blockIdx.x;
    const int lane = threadIdx.x;
    const int64_t n = tile * 64 + lane;
    uint32_t words[TILE8_WORDS];
#pragma unroll
    for (int k = 0; k < TILE8_WORDS; k++) words[k] = 0xfefefefeu;
    for (int jm1 = 0; jm1 < Q - 1; jm1++) {
        int32_t t, mn, mx;
        tile_span<LAYOUT>(tab, Npad, N, n, jm1, t, mn, mx);
        const int64_t k = tile * (Q - 1) + jm1;
        uint32_t o;
        if (wide[k]) {
            wtab[(int64_t)slot[k] * 64 + lane] = t;
            if (lane == 0) desc[k] = -(int32_t)slot[k] - 1;
            o = 254;
        } else {
            const int32_t base = mx >= 0 ? mn : 0;
            if (lane == 0) desc[k] = base;
            o = t >= 0 ? (uint32_t)(t - base) : (t == TAB_BOUNCE ? 255u : 254u);
        }
        const int sh = 8 * (jm1 & 3);
        words[jm1 >> 2] = (words[jm1 >> 2] & ~(0xffu << sh)) | (o << sh);
    }
#pragma unroll
    for (int k = 0; k < TILE8_FULL; k++) off[(tile * TILE8_FULL + k) * 64 + lane] = words[k];
    off16[tile * 64 + lane] = (uint16_t)(words[TILE8_FULL] & 0xffffu);
}

// Replaces the int32 table of a context by the TILE8 encoding.
int compress_tile8(plb_ctx* c)
{
    hipStream_t s = c->stream;
    const int64_t T = c->Npad / 64;
    const int64_t nk = T * (Q - 1);
    DevBuf<uint64_t> wide, slot;
    PLB_HIP(wide.alloc(nk));
    PLB_HIP(slot.alloc(nk));
    if (c->cfg.layout == PLB_LAYOUT_SOA)
        hipLaunchKernelGGL(k_tile8_flags<PLB_LAYOUT_SOA>, dim3((unsigned)T), dim3(64), 0, s, c->tab,
                           c->Npad, c->N, wide.p);
    else
        hipLaunchKernelGGL(k_tile8_flags<PLB_LAYOUT_AOSOA>, dim3((unsigned)T), dim3(64), 0, s, c->tab,
                           c->Npad, c->N, wide.p);
    PLB_HIP(hipGetLastError());
    int rc = exclusive_scan_u64(wide, slot, nk, s);
    if (rc) return rc;
    uint64_t last_slot = 0, last_wide = 0;
    PLB_HIP(hipMemcpyAsync(&last_slot, slot.p + nk - 1, 8, hipMemcpyDeviceToHost, s));
    PLB_HIP(hipMemcpyAsync(&last_wide, wide.p + nk - 1, 8, hipMemcpyDeviceToHost, s));
    PLB_HIP(hipStreamSynchronize(s));
    c->n_wide = (int64_t)(last_slot + last_wide);
    PLB_HIP(hipMalloc(&c->tdesc, sizeof(int32_t) * nk));
    PLB_HIP(hipMalloc(&c->toff, sizeof(uint32_t) * T * TILE8_FULL * 64));
    PLB_HIP(hipMalloc(&c->toff16, sizeof(uint16_t) * T * 64));
    PLB_HIP(hipMalloc(&c->twide, sizeof(int32_t) * 64 * (c->n_wide ? c->n_wide : 1)));
    if (c->cfg.layout == PLB_LAYOUT_SOA)
        hipLaunchKernelGGL(k_tile8_write<PLB_LAYOUT_SOA>, dim3((unsigned)T), dim3(64), 0, s, c->tab,
                           c->Npad, c->N, wide.p, slot.p, c->tdesc, c->toff, c->toff16, c->twide);
    else
        hipLaunchKernelGGL(k_tile8_write<PLB_LAYOUT_AOSOA>, dim3((unsigned)T), dim3(64), 0, s, c->tab,
                           c->Npad, c->N, wide.p, slot.p, c->tdesc, c->toff, c->toff16, c->twide);
    PLB_HIP(hipGetLastError());
    PLB_HIP(hipStreamSynchronize(s));
    (void)hipFree(c->tab);
    c->tab = nullptr;
    c->index_bytes = c->N ? ((double)T * (TILE8_FULL * 256 + 128) + (double)nk * 4 + (double)c->n_wide * 256) /
                                (double)c->N
                          : 0.0;
    return PLB_OK;
}

// ---------------------------------------------------------------- FP16 rounding keys
// The stochastic-rounding key of a node is its V-order id.  In pencil order
// the rows of one pencil in one z-plane follow each other as in the V order,
// so a 64-node tile's keys are one or two runs of consecutive values (a run
// ends where the pencil moves on to the next plane).  A tile with at most
// three runs (the builder checks every lane) is stored as the runs' bases
// (key - lane at the run's first lane) and start lanes, 16 bytes that the
// step reads with one scalar load instead of 4 bytes per node; other tiles
// are marked (bit 31 of the last word) and keep loading their keys.  Lanes at
// or past N are not stepped and are skipped.
__global__ void k_key_runs(const uint32_t* key, int64_t N, int64_t T, uint4* out, unsigned long long* n_marked)
{
    const int64_t tile = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tile >= T) return;
    const int64_t i0 = tile * 64;
    uint32_t base[3] = {0u, 0u, 0u}, start[3] = {0u, 64u, 64u};
    int runs = 0;
    bool fits = true;
    uint32_t prev = 0u;
    for (int l = 0; l < 64 && i0 + l < N; l++) {
        const uint32_t k = key[i0 + l];
        if (l == 0 || k != prev + 1u) {
            if (runs == 3) {
                fits = false;
                break;
            }
            start[runs] = (uint32_t)l;
            base[runs] = k - (uint32_t)l;
            runs++;
        }
        prev = k;
    }
    if (fits) {
        out[tile] = make_uint4(base[0], base[1], base[2], start[1] | (start[2] << 8));
    } else {
        out[tile] = make_uint4(0u, 0u, 0u, 0x80000000u);
        atomicAdd(n_marked, 1ull);
    }
}

int build_key_runs(plb_ctx* c)
{
    const uint32_t* key = c->key_perm ? c->key_perm : c->perm;
    if (!key || c->N == 0) return PLB_OK;
    const int64_t T = (c->Npad + 63) / 64;
    DevBuf<unsigned long long> marked;
    PLB_HIP(marked.alloc(1));
    PLB_HIP(hipMemsetAsync(marked.p, 0, sizeof(unsigned long long), c->stream));
    PLB_HIP(hipMalloc(&c->kseg, sizeof(uint4) * T));
    hipLaunchKernelGGL(k_key_runs, dim3((unsigned)((T + 255) / 256)), dim3(256), 0, c->stream, key, c->N, T,
                       c->kseg, marked.p);
    PLB_HIP(hipGetLastError());
    unsigned long long nm = 0;
    PLB_HIP(hipMemcpyAsync(&nm, marked.p, sizeof nm, hipMemcpyDeviceToHost, c->stream));
    PLB_HIP(hipStreamSynchronize(c->stream));
    c->key_bytes = (16.0 * (double)T + 256.0 * (double)nm) / (double)c->N;
    return PLB_OK;
}

// ---------------------------------------------------------------- BITMAP index
// Node order as a voxel position: pencils of py rows, z fastest across
// pencils, V = x + nx*(ly + py*(z + nz*yb)) with y = yb*py + ly.  With
// py = ny this is the V-file order z,y,x.  Node ids are the ranks of the
// pores in this order, so id(v) = pores before v's 32-voxel word + pores
// below v inside the word.
struct PencilMap {
    uint32_t nx, nz, py;
};

__global__ void k_bm_bits(const uint64_t* __restrict__ xyz, int64_t N, PencilMap m, uint32_t* bits,
                          uint32_t* vox)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= N) return;
    int x, y, z;
    unpack(xyz[g], x, y, z);
    const uint32_t yb = (uint32_t)y / m.py, ly = (uint32_t)y - yb * m.py;
    const uint32_t v = ((yb * m.nz + (uint32_t)z) * m.py + ly) * m.nx + (uint32_t)x;
    vox[g] = v;
    atomicOr(bits + (v >> 5), 1u << (v & 31u));
}

__global__ void k_bm_count(const uint32_t* bits, int64_t nw, uint64_t* cnt)
{
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w < nw) cnt[w] = (uint64_t)__popc(bits[w]);
}

__global__ void k_bm_pack(const uint32_t* bits, const uint64_t* pre, int64_t nw, uint2* bm)
{
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (w < nw) bm[w] = make_uint2(bits[w], (uint32_t)pre[w]);
}

// Every node must find itself: its voxel is a pore whose rank is its id.
__global__ void k_bm_check(const uint2* bm, const uint32_t* vox, int64_t N, int* err)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= N) return;
    const uint32_t v = vox[g];
    const uint2 wd = bm[v >> 5];
    const uint32_t b = v & 31u;
    const uint32_t id = wd.y + (uint32_t)__popc(wd.x & ((1u << b) - 1u));
    if (!((wd.x >> b) & 1u) || (int64_t)id != g) *err = 1;
}

// Replaces the int32 table of a context by the BITMAP index.
int build_bitmap(plb_ctx* c)
{
    hipStream_t s = c->stream;
    const plb_config& cfg = c->cfg;
    uint32_t py = (uint32_t)c->ny;
    if (cfg.ordering == PLB_ORDER_PENCIL) {
        if (cfg.brick[0] < c->nx || cfg.brick[2] != 1) return PLB_ERR_INVALID;
        py = (uint32_t)std::min(cfg.brick[1], c->ny);
    } else if (cfg.ordering != PLB_ORDER_REFERENCE) {
        return PLB_ERR_INVALID;
    }
    const uint64_t nyb = ((uint64_t)c->ny + py - 1) / py;
    const uint64_t nvox = (uint64_t)c->nx * nyb * py * (uint64_t)c->nz;
    if (nvox > 0xffffffffull) return PLB_ERR_RANGE;
    const int64_t nw = (int64_t)((nvox + 31) / 32);
    DevBuf<uint32_t> bits;
    DevBuf<uint64_t> cnt, pre;
    DevBuf<int> derr;
    PLB_HIP(bits.alloc(nw));
    PLB_HIP(cnt.alloc(nw));
    PLB_HIP(pre.alloc(nw));
    PLB_HIP(derr.alloc(1));
    PLB_HIP(hipMemsetAsync(bits, 0, sizeof(uint32_t) * nw, s));
    PLB_HIP(hipMemsetAsync(derr, 0, sizeof(int), s));
    PLB_HIP(hipMalloc(&c->vox, sizeof(uint32_t) * c->Npad));
    PLB_HIP(hipMemsetAsync(c->vox, 0, sizeof(uint32_t) * c->Npad, s));
    PLB_HIP(hipMalloc(&c->bm, sizeof(uint2) * nw));
    PencilMap m;
    m.nx = (uint32_t)c->nx;
    m.nz = (uint32_t)c->nz;
    m.py = py;
    if (c->N > 0)
        hipLaunchKernelGGL(k_bm_bits, dim3(nblk(c->N, 256)), dim3(256), 0, s, c->node_xyz, c->N, m,
                           bits.p, c->vox);
    hipLaunchKernelGGL(k_bm_count, dim3(nblk(nw, 256)), dim3(256), 0, s, bits.p, nw, cnt.p);
    PLB_HIP(hipGetLastError());
    int rc = exclusive_scan_u64(cnt, pre, nw, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_bm_pack, dim3(nblk(nw, 256)), dim3(256), 0, s, bits.p, pre.p, nw, c->bm);
    if (c->N > 0)
        hipLaunchKernelGGL(k_bm_check, dim3(nblk(c->N, 256)), dim3(256), 0, s, c->bm, c->vox, c->N,
                           derr.p);
    PLB_HIP(hipGetLastError());
    int herr = 0;
    PLB_HIP(hipMemcpyAsync(&herr, derr, sizeof(int), hipMemcpyDeviceToHost, s));
    PLB_HIP(hipStreamSynchronize(s));
    if (herr) return PLB_ERR_INVALID;  // node order and voxel order disagree
    (void)hipFree(c->tab);
    c->tab = nullptr;
    c->py = (int)py;
    c->n_bm_words = nw;
    c->index_bytes = c->N ? (4.0 * (double)c->N + 8.0 * (double)nw) / (double)c->N : 0.0;
    return PLB_OK;
}

int build_table_from_V(plb_ctx* c, const plb_node_info* V, int64_t n)
{
    hipStream_t s = c->stream;
    DevBuf<uint64_t> xyz;
    DevBuf<int32_t> dstart, tabref;
    DevBuf<int> derr;
    int herr = 0;
    {
        DevBuf<uint64_t> dV;  // 152-byte records as 19 words; freed before the final table exists
        PLB_HIP(dV.alloc(n * 19));
        PLB_HIP(dstart.alloc(c->num_p));
        PLB_HIP(tabref.alloc((Q - 1) * n));
        PLB_HIP(xyz.alloc(n));
        PLB_HIP(derr.alloc(1));
        PLB_HIP(hipMemsetAsync(derr, 0, sizeof(int), s));
        if (n) PLB_HIP(hipMemcpyAsync(dV, V, sizeof(plb_node_info) * n, hipMemcpyHostToDevice, s));
        PLB_HIP(hipMemcpyAsync(dstart, c->V_sub_start.data(), sizeof(int32_t) * c->num_p,
                               hipMemcpyHostToDevice, s));
        if (n) {
            VArgs a;
            a.V = dV;
            a.N = n;
            a.start = dstart;
            a.num_p = c->num_p;
            a.nx = c->nx;
            a.ny = c->ny;
            a.nz = c->nz;
            a.mode = c->cfg.bb_mode;
            a.tab = tabref;
            a.xyz = xyz;
            a.err = derr;
            hipLaunchKernelGGL(k_table_from_V, dim3(nblk(n, 256)), dim3(256), 0, s, a);
            PLB_HIP(hipGetLastError());
        }
        PLB_HIP(hipMemcpyAsync(&herr, derr, sizeof(int), hipMemcpyDeviceToHost, s));
        PLB_HIP(hipStreamSynchronize(s));
    }
    return herr ? herr : finalize_table(c, tabref, xyz);
}

// Device-side mask analysis shared by plb_create_from_flag and plb_preprocess.
struct FlagIndex {
    DevBuf<uint64_t> bits, wpre, xyz;
    int64_t N = 0, nvox = 0, nwords = 0;
};

static int index_flag(const uint8_t* flag, int nx, int ny, int nz, hipStream_t s, FlagIndex& fi)
{
    fi.nvox = (int64_t)nx * ny * nz;
    fi.nwords = (fi.nvox + 63) / 64;
    uint64_t last_pre = 0, last_cnt = 0;
    int herr = 0;
    {
        DevBuf<uint8_t> dflag;
        DevBuf<uint64_t> cnt;
        DevBuf<int> derr;
        PLB_HIP(dflag.alloc(fi.nvox));
        PLB_HIP(fi.bits.alloc(fi.nwords));
        PLB_HIP(cnt.alloc(fi.nwords));
        PLB_HIP(fi.wpre.alloc(fi.nwords));
        PLB_HIP(derr.alloc(1));
        PLB_HIP(hipMemsetAsync(derr, 0, sizeof(int), s));
        PLB_HIP(hipMemcpyAsync(dflag, flag, fi.nvox, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_flag_bits, dim3(nblk(fi.nvox, 256)), dim3(256), 0, s, dflag.p, fi.nvox,
                           fi.nwords, fi.bits.p, cnt.p, derr.p);
        PLB_HIP(hipGetLastError());
        int rc = exclusive_scan_u64(cnt, fi.wpre, fi.nwords, s);
        if (rc) return rc;
        PLB_HIP(hipMemcpyAsync(&last_pre, fi.wpre.p + fi.nwords - 1, 8, hipMemcpyDeviceToHost, s));
        PLB_HIP(hipMemcpyAsync(&last_cnt, cnt.p + fi.nwords - 1, 8, hipMemcpyDeviceToHost, s));
        PLB_HIP(hipMemcpyAsync(&herr, derr, sizeof(int), hipMemcpyDeviceToHost, s));
        PLB_HIP(hipStreamSynchronize(s));
    }
    if (herr) return herr;
    fi.N = (int64_t)(last_pre + last_cnt);
    if (fi.N >= INT32_MAX) return PLB_ERR_RANGE;
    PLB_HIP(fi.xyz.alloc(fi.N));
    hipLaunchKernelGGL(k_nodes_from_bits, dim3(nblk(fi.nvox, 256)), dim3(256), 0, s, fi.bits.p,
                       fi.wpre.p, fi.nvox, nx, ny, fi.xyz.p);
    PLB_HIP(hipGetLastError());
    return PLB_OK;
}

// V-order pull table [18][N] and node coordinates straight from a voxel mask,
// with num_p z-slabs (preprocess.c:97-117 rank rule).  Caller frees.
int build_ref_from_flag(const uint8_t* flag, int nx, int ny, int nz, int num_p, int mode,
                        hipStream_t s, int32_t** tabref_out, uint64_t** xyz_out, int64_t* n_out,
                        std::vector<int32_t>* start_out)
{
    *tabref_out = nullptr;
    *xyz_out = nullptr;
    FlagIndex fi;
    int rc = index_flag(flag, nx, ny, nz, s, fi);
    if (rc) return rc;
    GridArgs a;
    a.bits = fi.bits;
    a.wpre = fi.wpre;
    a.xyz = fi.xyz;
    a.N = fi.N;
    a.nx = nx;
    a.ny = ny;
    a.nz = nz;
    a.slab = nz / num_p;
    DevBuf<int32_t> dstart, tabref;
    PLB_HIP(dstart.alloc(num_p));
    hipLaunchKernelGGL(k_slab_starts, dim3(nblk(num_p, 64)), dim3(64), 0, s, a, num_p, dstart.p);
    start_out->assign(num_p, 0);
    PLB_HIP(hipMemcpyAsync(start_out->data(), dstart, sizeof(int32_t) * num_p, hipMemcpyDeviceToHost, s));
    PLB_HIP(tabref.alloc((Q - 1) * fi.N));
    if (fi.N) {
        hipLaunchKernelGGL(k_table_from_bits, dim3(nblk(fi.N, 256)), dim3(256), 0, s, a, mode, tabref.p);
        PLB_HIP(hipGetLastError());
    }
    PLB_HIP(hipStreamSynchronize(s));
    *tabref_out = tabref.release();
    *xyz_out = fi.xyz.release();
    *n_out = fi.N;
    return PLB_OK;
}

int build_table_from_flag(plb_ctx* c, const uint8_t* flag)
{
    DevBuf<int32_t> tabref;
    DevBuf<uint64_t> xyz;
    int64_t n = 0;
    int rc = build_ref_from_flag(flag, c->nx, c->ny, c->nz, c->num_p, c->cfg.bb_mode, c->stream,
                                 &tabref.p, &xyz.p, &n, &c->V_sub_start);
    if (rc) return rc;
    c->N = n;
    c->Npad = std::max<int64_t>(64, (n + 63) / 64 * 64);
    return finalize_table(c, tabref, xyz);
}

template <int LAYOUT>
__global__ void k_copy_table_offset(const int32_t* __restrict__ tabref, int64_t n, int64_t off,
                                    int32_t* __restrict__ out, int64_t npad)
{
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
#pragma unroll
    for (int j = 0; j < Q - 1; j++) {
        int32_t t = tabref[(int64_t)j * n + g];
        if (t >= 0) t += (int32_t)off;
        out[tpos<LAYOUT>(j, off + g, npad)] = t;
    }
}

// Places one sample's V-order table at node offset `off` of a batch table.
int copy_table_offset(int layout, const int32_t* tabref, int64_t n, int64_t off, int32_t* out,
                      int64_t npad, hipStream_t s)
{
    if (n == 0) return PLB_OK;
    if (layout == PLB_LAYOUT_SOA)
        hipLaunchKernelGGL(k_copy_table_offset<PLB_LAYOUT_SOA>, dim3(nblk(n, 256)), dim3(256), 0, s,
                           tabref, n, off, out, npad);
    else
        hipLaunchKernelGGL(k_copy_table_offset<PLB_LAYOUT_AOSOA>, dim3(nblk(n, 256)), dim3(256), 0,
                           s, tabref, n, off, out, npad);
    PLB_HIP(hipGetLastError());
    return PLB_OK;
}

}  // namespace plb

using namespace plb;

// Device preprocess on its own stream; the stream is destroyed on every path.
static int preprocess_on(hipStream_t s, const uint8_t* flag, int nx, int ny, int nz, int num_p,
                         plb_node_info* V_out, int32_t* V_sub_start_out, int64_t* n_nodes_out)
{
    FlagIndex fi;
    int rc = index_flag(flag, nx, ny, nz, s, fi);
    if (rc) return rc;
    if (n_nodes_out) *n_nodes_out = fi.N;
    GridArgs a;
    a.bits = fi.bits;
    a.wpre = fi.wpre;
    a.xyz = fi.xyz;
    a.N = fi.N;
    a.nx = nx;
    a.ny = ny;
    a.nz = nz;
    a.slab = nz / num_p;
    DevBuf<int32_t> dstart;
    PLB_HIP(dstart.alloc(num_p));
    hipLaunchKernelGGL(k_slab_starts, dim3(nblk(num_p, 64)), dim3(64), 0, s, a, num_p, dstart.p);
    if (V_sub_start_out)
        PLB_HIP(hipMemcpyAsync(V_sub_start_out, dstart, sizeof(int32_t) * num_p, hipMemcpyDeviceToHost, s));
    if (V_out && fi.N) {
        // Records are produced in chunks so the device copy stays bounded.
        const int64_t chunk = 16 * 1024 * 1024;
        DevBuf<plb_node_info> dV;
        PLB_HIP(dV.alloc(fi.N < chunk ? fi.N : chunk));
        for (int64_t b0 = 0; b0 < fi.N; b0 += chunk) {
            const int64_t m = (fi.N - b0 < chunk) ? fi.N - b0 : chunk;
            GridArgs ac = a;
            ac.xyz = fi.xyz.p + b0;
            ac.N = m;
            hipLaunchKernelGGL(k_V_from_bits, dim3(nblk(m, 256)), dim3(256), 0, s, ac, dstart.p,
                               (uint64_t*)dV.p);
            PLB_HIP(hipGetLastError());
            PLB_HIP(hipMemcpyAsync(V_out + b0, dV, sizeof(plb_node_info) * m, hipMemcpyDeviceToHost, s));
            PLB_HIP(hipStreamSynchronize(s));
        }
    }
    PLB_HIP(hipStreamSynchronize(s));
    return PLB_OK;
}

extern "C" int plb_preprocess(const uint8_t* flag, int nx, int ny, int nz, int num_p, int device,
                              plb_node_info* V_out, int32_t* V_sub_start_out,
                              int64_t* n_nodes_out)
{
    if (!flag || nx <= 0 || ny <= 0 || nz <= 0 || nx > 65535 || ny > 65535 || nz > 65535)
        return PLB_ERR_INVALID;
    if (num_p < 1 || num_p > 65535 || nz / num_p == 0) return PLB_ERR_NUM_P;
    if ((nz - 1) / (nz / num_p) >= num_p) return PLB_ERR_NUM_P;
    Range range("plb_preprocess");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return PLB_ERR_NODEVICE;
    if (device < 0 || device >= ndev) return PLB_ERR_INVALID;
    PLB_HIP(hipSetDevice(device));
    hipStream_t s;
    PLB_HIP(hipStreamCreate(&s));
    const int rc = preprocess_on(s, flag, nx, ny, nz, num_p, V_out, V_sub_start_out, n_nodes_out);
    (void)hipStreamDestroy(s);
    return rc;
}
