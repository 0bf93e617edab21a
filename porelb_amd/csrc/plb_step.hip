// plb_step.hip -- the fused D3Q19 time step on gfx950, plus the kernels that
// move populations and macros in and out of the device layouts and the FP64
// reductions.
//
// One kernel per time step, one thread per pore node:
//   1. 18 pull-table entries (int32, TILE8 bases + packed u8 offsets, or
//      derived from the BITMAP index);
//   2. 19 population gathers issued back to back (stream: f[j][src];
//      bounce and const: f[re[j]][self], const replaced after the load);
//   3. collision with body force (porelb.c:318-365);
//   4. 19 coalesced non-temporal stores into the other ping-pong buffer.
// This fuses stream_inner_nodes, update_recv_dist and collision
// (porelb.c:285-365); the reference's f1 array never exists.
//
// FP64 (PLB_FP64): unshifted populations, arithmetic in exactly the
// reference's evaluation order (porelb.c:6-13, 326-362); the library is
// compiled with -ffp-contract=off so results match the reference bit for bit.
// FP32/FP16 storage: populations are stored as deviations d_j = f_j - w_j*rho0
// (FP16 scaled by 2^k), and the collision runs in FP32 on the deviations:
//   d'_j = (1-omega) d_j + omega*dfeq_j + S_j*feq_j,
//   dfeq_j = w_j (drho + rho*P_j), P_j = 3 eu + 4.5 eu^2 - 1.5 u^2,
//   S_j = (1 - omega/2)*3*(e_jx - u)*F,
// which is the reference update minus w_j*rho0 on both sides.
#include <algorithm>
#include <type_traits>

#include "plb_internal.hpp"

namespace plb {

struct StepArgs {
    const void* fin;
    void* fout;
    const int32_t* tab;
    const int32_t* tdesc;
    const uint32_t* toff;
    const uint16_t* toff16;
    const int32_t* twide;
    int64_t N, Npad;
    double omega, gx, rho0;
    int64_t it;
    int write_macros;
    double* macros;  // [4][Npad]
    unsigned long long* fault;
    const uint32_t* perm;
    float scale, inv_scale;
    // BITMAP index: voxel position V = x + nx*(ly + py*(z + nz*yb)), y = yb*py + ly
    const uint2* bm;
    const uint32_t* vox;
    uint32_t nx, ny, nz, py;
    FastDiv div_nx, div_py, div_nz;
    // Graph replays: the step number is it + *it_base (set before each replay).
    const int64_t* it_base;
    int stochastic;  // FP16 stores: stochastic rounding
    int has_const;   // the table may hold constant entries (reference bounce order)
    // Node id that keys the rounding stream: the V-order id of the node
    // (within its sample for batches), so every node order and a batch give
    // the same results as a single lattice in the V order.  Null: identity.
    const uint32_t* key_perm;
    // The same keys as runs per 64-node tile (null: load key_perm per node).
    const uint4* kseg;
};

__device__ __forceinline__ StepArgs with_step(const StepArgs& a)
{
    StepArgs b = a;
    if (a.it_base) b.it += *a.it_base;
    return b;
}

// ---------------------------------------------------------------- FP64, reference order
__device__ __forceinline__ double feq_ref(int k, double rho, double u, double v, double w)
{
    // porelb.c:8-12, same operations in the same order.
    double uv, eu, x;
    eu = EX[k] * u + EY[k] * v + EZ[k] * w;
    uv = u * u + v * v + w * w;
    x = TP[k] * rho * (1.0 + 3 * eu + 4.5 * eu * eu - 1.5 * uv);
    return x;
}

__device__ __forceinline__ void record_fault(const StepArgs& a, int64_t i, double rho)
{
    const uint64_t g = a.perm ? (uint64_t)a.perm[i] : (uint64_t)i;
    const uint64_t key = ((uint64_t)a.it << 32) | g;
    atomicMin(a.fault, (unsigned long long)key);
    a.macros[i] = rho;  // rho of the offending node, read back by plb_last_fault
}

// A step after the first faulting step stores nothing, so the state stays
// that of the faulting step (the reference exits there).  The fault word is
// loaded at kernel entry but only consulted before the stores, so it adds no
// latency in front of the table and population loads.
__device__ __forceinline__ bool frozen(unsigned long long fk, int64_t it)
{
    return fk != NO_FAULT && (int64_t)(fk >> 32) < it;
}

// ---------------------------------------------------------------- phases of one node
// A node update runs in four phases so a thread can interleave two nodes
// (issue both nodes' index loads, then both nodes' gathers, before waiting):
//   issue_entries  -> index loads (I32: 18 words; TILE8: 4 words + 1 short + 18 scalar
//                     descriptors of the wave's tile)
//   decode_entries -> t[1..18]: >= 0 source node, -1 bounce, -2 constant
//   gather         -> the 19 stored populations
//   collide_*      -> collision and the 19 stores

template <int IDX>
struct Entries {
    int32_t t[Q];                // I32: raw entries; after decode: decoded
    uint32_t w[TILE8_WORDS];     // TILE8: packed offsets of this lane
    int32_t desc[Q];             // TILE8: the tile's descriptors (uniform)
    uint32_t v;                  // BITMAP: the node's voxel position
};

// ---------------------------------------------------------------- BITMAP sources
// Periodic neighbour coordinate c - e (e in {-1, 0, 1}) on [0, n).
__device__ __forceinline__ uint32_t wrap_sub(uint32_t c, int e, uint32_t n)
{
    if (e > 0) return c == 0 ? n - 1 : c - 1;
    if (e < 0) return c + 1 == n ? 0 : c + 1;
    return c;
}

// Voxel position of the pull source x - e_j, for any node.
__device__ __forceinline__ uint32_t bitmap_src_general(const StepArgs& a, int j, uint32_t x, uint32_t y,
                                                       uint32_t z)
{
    const uint32_t xs = wrap_sub(x, EX[j], a.nx);
    const uint32_t ys = wrap_sub(y, EY[j], a.ny);
    const uint32_t zs = wrap_sub(z, EZ[j], a.nz);
    const uint32_t yb = fastdiv(ys, a.div_py);
    const uint32_t ly = ys - yb * a.py;
    return ((yb * a.nz + zs) * a.py + ly) * a.nx + xs;
}

__device__ __forceinline__ void bitmap_coords(const StepArgs& a, uint32_t v, uint32_t& x, uint32_t& y,
                                              uint32_t& z, uint32_t& ly)
{
    const uint32_t r1 = fastdiv(v, a.div_nx);
    x = v - r1 * a.nx;
    const uint32_t r2 = fastdiv(r1, a.div_py);
    ly = r1 - r2 * a.py;
    const uint32_t yb = fastdiv(r2, a.div_nz);
    z = r2 - yb * a.nz;
    y = yb * a.py + ly;
}

// Node id of the pore at voxel position v, or TAB_BOUNCE for a solid voxel
// (half-way bounce-back): the word's pore count plus the pores below v.
__device__ __forceinline__ int32_t bitmap_lookup(uint2 wd, uint32_t v)
{
    const uint32_t b = v & 31u;
    const uint32_t below = __builtin_amdgcn_ubfe(wd.x, 0, b);
    const int32_t id = (int32_t)(wd.y + __builtin_popcount(below));
    return ((wd.x >> b) & 1u) ? id : TAB_BOUNCE;
}

// A read through the constant address space, for data no kernel of the
// launch writes (sample tables, pull-table descriptors): it stays a scalar
// load where the address is uniform, also in kernels that store elsewhere.
template <typename T>
__device__ __forceinline__ T ld_const(const T* q)
{
    return *(const __attribute__((address_space(4))) T*)q;
}

// CONST_DESC: read the TILE8 descriptors through the constant address space
// (for loops that also store; one-node kernels get scalar loads anyway).
template <int LAYOUT, int IDX, bool CONST_DESC = false>
__device__ __forceinline__ void issue_entries(const StepArgs& a, int64_t i, Entries<IDX>& e)
{
    // Index words are read once per step: non-temporal loads keep them from
    // displacing population lines that neighbouring tiles read again.
    if constexpr (IDX == PLB_INDEX_I32) {
#pragma unroll
        for (int j = 1; j < Q; j++) e.t[j] = __builtin_nontemporal_load(a.tab + tpos<LAYOUT>(j - 1, i, a.Npad));
    } else if constexpr (IDX == PLB_INDEX_BITMAP) {
        e.v = __builtin_nontemporal_load(a.vox + i);
    } else {
        // TILE8: the wave is one 64-node tile, so its 18 descriptors are
        // uniform (scalar loads); each lane reads 4 words and a short of packed
        // offsets.
        const int64_t tile = __builtin_amdgcn_readfirstlane((int)(i >> 6));
        const uint32_t* o = a.toff + tile * TILE8_FULL * 64 + (i & 63);
#pragma unroll
        for (int k = 0; k < TILE8_FULL; k++) e.w[k] = __builtin_nontemporal_load(o + k * 64);
        e.w[TILE8_FULL] = __builtin_nontemporal_load(a.toff16 + tile * 64 + (i & 63));
        const int32_t* d = a.tdesc + tile * (Q - 1);
        if constexpr (CONST_DESC) {
            // never written while steps run: scalar loads even in a loop
            // that stores, where the compiler could not prove it otherwise
#pragma unroll
            for (int j = 1; j < Q; j++) e.desc[j] = ld_const(d + j - 1);
        } else {
#pragma unroll
            for (int j = 1; j < Q; j++) e.desc[j] = d[j - 1];
        }
    }
}

template <int IDX>
__device__ __forceinline__ void decode_entries(const StepArgs& a, int64_t i, Entries<IDX>& e)
{
    if constexpr (IDX == PLB_INDEX_BITMAP) {
        const uint32_t v = e.v;
        uint32_t x, y, z, ly;
        bitmap_coords(a, v, x, y, z, ly);
        // Away from pencil edges and the periodic y/z boundary every source is
        // a fixed offset in node order; x wraps are a per-lane correction.
        const bool inner = (int)ly >= 1 && (int)ly <= (int)a.py - 2 && (int)y >= 1 &&
                           (int)y <= (int)a.ny - 2 && (int)z >= 1 && (int)z <= (int)a.nz - 2;
        uint32_t vs[Q];
        if (__builtin_amdgcn_ballot_w64(!inner) == 0) {
            const uint32_t xm = x == 0 ? a.nx : 0u;            // source at x-1 wraps to nx-1
            const uint32_t xp = x + 1 == a.nx ? a.nx : 0u;     // source at x+1 wraps to 0
            const uint32_t row = a.nx, plane = a.nx * a.py;
#pragma unroll
            for (int j = 1; j < Q; j++) {
                uint32_t d = v - (uint32_t)EX[j] - (uint32_t)EY[j] * row - (uint32_t)EZ[j] * plane;
                if (EX[j] > 0) d += xm;
                if (EX[j] < 0) d -= xp;
                vs[j] = d;
            }
        } else {
#pragma unroll
            for (int j = 1; j < Q; j++) vs[j] = bitmap_src_general(a, j, x, y, z);
        }
        uint2 wd[Q];
#pragma unroll
        for (int j = 1; j < Q; j++) wd[j] = a.bm[vs[j] >> 5];
#pragma unroll
        for (int j = 1; j < Q; j++) e.t[j] = bitmap_lookup(wd[j], vs[j]);
    } else if constexpr (IDX == PLB_INDEX_TILE8) {
        int32_t any = 0;
#pragma unroll
        for (int j = 1; j < Q; j++) {
            // one select: 0..253 -> base + v, 254 -> -2 (const), 255 -> -1 (bounce)
            const int32_t v = (int32_t)((e.w[(j - 1) >> 2] >> (8 * ((j - 1) & 3))) & 0xffu);
            e.t[j] = v < 254 ? e.desc[j] + v : v - 256;
            any |= e.desc[j];
        }
        if (any < 0) {  // a (tile, direction) pair of this tile is wide
#pragma unroll
            for (int j = 1; j < Q; j++)
                if (e.desc[j] < 0)
                    e.t[j] = __builtin_nontemporal_load(a.twide + (int64_t)(-e.desc[j] - 1) * 64 + (i & 63));
        }
    }
}

// 19 gathers: rest population of the node, then per direction the source
// node's slot j (stream) or the node's own slot re[j] (bounce; constant
// entries read that slot too and are replaced afterwards).
template <typename S, int LAYOUT>
__device__ __forceinline__ void gather(const StepArgs& a, int64_t i, const int32_t t[Q], S raw[Q])
{
    const S* __restrict__ fin = (const S*)a.fin;
    raw[0] = fin[fpos<LAYOUT>(0, i, a.Npad)];
#pragma unroll
    for (int j = 1; j < Q; j++) {
        const int64_t src = t[j] >= 0 ? fpos<LAYOUT>(j, t[j], a.Npad) : fpos<LAYOUT>(RE[j], i, a.Npad);
        raw[j] = fin[src];
    }
}

// Where the post-collision population j of node i is stored.  MODE 0: the
// other buffer, slot j (pull, two buffers).  One buffer with the AA pattern:
// MODE 1 (local steps) the node's own slot re[j]; MODE 2 (neighbour steps)
// slot j of the node downstream, t[re[j]] (x + e_j), or the node's own slot
// re[j] when that neighbour is solid (half-way bounce-back).  The rest
// population stays in slot 0.
template <int LAYOUT, int MODE>
__device__ __forceinline__ int64_t out_pos(const StepArgs& a, int64_t i, const int32_t t[Q], int j)
{
    if constexpr (MODE == 0) return fpos<LAYOUT>(j, i, a.Npad);
    else if constexpr (MODE == 1) return fpos<LAYOUT>(RE[j], i, a.Npad);
    else {
        if (j == 0) return fpos<LAYOUT>(0, i, a.Npad);
        return t[RE[j]] >= 0 ? fpos<LAYOUT>(j, t[RE[j]], a.Npad) : fpos<LAYOUT>(RE[j], i, a.Npad);
    }
}

// FP64 collision in the reference's evaluation order (porelb.c:326-362).
template <int LAYOUT, int MODE = 0>
__device__ __forceinline__ void collide_f64(const StepArgs& a, int64_t i, const int32_t t[Q],
                                            double f[Q], unsigned long long fk)
{
    double* __restrict__ fout = (double*)a.fout;
    if constexpr (MODE == 0) {  // constant entries: reference bounce order, two buffers only
#pragma unroll
        for (int j = 1; j < Q; j++)
            if (t[j] == TAB_CONST) f[j] = TP[j] * a.rho0 * 1.0;  // feq(rho0, u=0) = initial f1
    }

    double rho = 0.0, rhou = 0.0, rhov = 0.0, rhow = 0.0;
#pragma unroll
    for (int j = 0; j < Q; j++) {
        rho += f[j];
        rhou += EX[j] * f[j];
        rhov += EY[j] * f[j];
        rhow += EZ[j] * f[j];
    }
    if (frozen(fk, a.it)) return;
    if ((rho > 1.2) | (rho < 0.5)) record_fault(a, i, rho);
    if (a.it > 0) rhou += 0.5 * a.gx;
    const double u = rhou / rho, v = rhov / rho, w = rhow / rho;
    if (a.write_macros) {
        a.macros[i] = rho;
        a.macros[a.Npad + i] = u;
        a.macros[2 * a.Npad + i] = v;
        a.macros[3 * a.Npad + i] = w;
    }
    const double omega = a.omega, Xforce = a.gx;
#pragma unroll
    for (int j = 0; j < Q; j++) {
        const double o = (1.0 - omega) * f[j] +
                         (omega + (1.0 - 0.5 * omega) * 3.0 * (EX[j] - u) * Xforce) *
                             feq_ref(j, rho, u, v, w);
        __builtin_nontemporal_store(o, fout + out_pos<LAYOUT, MODE>(a, i, t, j));
    }
}

// ---------------------------------------------------------------- shifted FP32 / FP16 storage
// The collision works in the stored scale: ds_j = d_j * sc (sc = 2^k for
// FP16, 1 for FP32).  Scaling by a power of two commutes with every FP32
// rounding, so the moments unscaled once and the stores with the scale
// folded into the collision constants are bit-identical to unscaling each
// population and scaling each store.
__device__ __forceinline__ float to_float(float v) { return v; }
__device__ __forceinline__ float to_float(__half v) { return __half2float(v); }
__device__ __forceinline__ void st_dev(float* p, int64_t k, float v)
{
    __builtin_nontemporal_store(v, p + k);
}
__device__ __forceinline__ void st_dev(__half* p, int64_t k, float v)
{
    const unsigned short b = __half_as_ushort(__float2half_rn(v));
    __builtin_nontemporal_store(b, (unsigned short*)p + k);
}

// Stochastic rounding of FP16 stores.  v_cvt_sr_f16_f32 rounds up with
// probability equal to the dropped fraction of an FP16 ulp, taken from the
// random word r; the expected stored value equals the FP32 value, so the
// sub-ulp increments of a flow near steady state accumulate instead of
// being rounded away.  r for direction j of a node: a 32-bit hash of the
// node id and the step (lowbias32, C. Wellons, "Prospecting for hash
// functions", 2018) plus j times the golden-ratio constant, which keeps every
// direction's word uniformly distributed.
typedef _Float16 plb_half2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t mix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
__device__ __forceinline__ void st_dev_sr(float* p, int64_t k, float v, uint32_t) { st_dev(p, k, v); }
__device__ __forceinline__ void st_dev_sr(__half* p, int64_t k, float v, uint32_t r)
{
    const plb_half2 h = __builtin_amdgcn_cvt_sr_f16_f32(plb_half2{0, 0}, v, r, false);
    __builtin_nontemporal_store(__builtin_bit_cast(unsigned short, h[0]), (unsigned short*)p + k);
}

template <typename S, int LAYOUT, int MODE = 0>
__device__ __forceinline__ void collide_shift(const StepArgs& a, int64_t i, const int32_t t[Q],
                                              const S raw[Q], unsigned long long fk, uint32_t key)
{
    S* __restrict__ fout = (S*)a.fout;
    constexpr bool scaled = sizeof(S) == 2;
    const float inv = a.inv_scale, sc = a.scale;
    float ds[Q];
#pragma unroll
    for (int j = 0; j < Q; j++) ds[j] = to_float(raw[j]);
    if constexpr (MODE == 0) {
        if (__builtin_amdgcn_readfirstlane(a.has_const)) {  // constant entries: reference bounce order only
#pragma unroll
            for (int j = 1; j < Q; j++)
                if (t[j] == TAB_CONST) ds[j] = 0.f;
        }
    }

    float drho = 0.f, jx = 0.f, jy = 0.f, jz = 0.f;
#pragma unroll
    for (int j = 0; j < Q; j++) {
        drho += ds[j];
        if (EX[j]) jx += EX[j] > 0 ? ds[j] : -ds[j];
        if (EY[j]) jy += EY[j] > 0 ? ds[j] : -ds[j];
        if (EZ[j]) jz += EZ[j] > 0 ? ds[j] : -ds[j];
    }
    if constexpr (scaled) {
        drho *= inv;
        jx *= inv;
        jy *= inv;
        jz *= inv;
    }
    const float rho0 = (float)a.rho0;
    const float rho = rho0 + drho;
    if (frozen(fk, a.it)) return;
    if ((rho > 1.2f) | (rho < 0.5f)) record_fault(a, i, (double)rho);
    const float F = (float)a.gx;
    if (a.it > 0) jx += 0.5f * F;
    const float rinv = 1.0f / rho;
    const float u = jx * rinv, v = jy * rinv, w = jz * rinv;
    if (a.write_macros) {
        a.macros[i] = (double)rho;
        a.macros[a.Npad + i] = (double)u;
        a.macros[2 * a.Npad + i] = (double)v;
        a.macros[3 * a.Npad + i] = (double)w;
    }
    const float omega = (float)a.omega;
    const float c1 = 1.0f - omega;
    const float c2 = (1.0f - 0.5f * omega) * 3.0f * F;
    // Fused multiply-adds are written out (the library is built without
    // contraction for the FP64 path's sake); this path is held to a
    // tolerance against FP64, not to bitwise equality.
    const float usq = 1.5f * fmaf(u, u, fmaf(v, v, w * w));
    const float fu = -c2 * u;  // force term c2 (e_jx - u) for e_jx = 0, then +-c2
    // the store scale folded into the constants of the dfeq and force terms
    const float omega_s = scaled ? omega * sc : omega;
    const float f0s = scaled ? fu * sc : fu, fps = scaled ? (fu + c2) * sc : fu + c2,
                fms = scaled ? (fu - c2) * sc : fu - c2;
    S* __restrict__ onode = fout + fpos<LAYOUT>(0, i, a.Npad);
    // The rounding mode is uniform: one loop per mode under a scalar branch.
    auto relax_and_store = [&](auto stochastic) {
        constexpr bool SR = decltype(stochastic)::value;
        uint32_t rnd = 0;
        if constexpr (SR) rnd = mix32(key ^ mix32((uint32_t)a.it * 0x9e3779b9u + 0x632be5abu));
#pragma unroll
        for (int j = 0; j < Q; j++) {
            float eu = 0.f;
            if (EX[j]) eu = EX[j] > 0 ? u : -u;
            if (EY[j]) eu += EY[j] > 0 ? v : -v;
            if (EZ[j]) eu += EZ[j] > 0 ? w : -w;
            const float P = fmaf(eu, fmaf(4.5f, eu, 3.0f), -usq);
            const float wj = (float)TP[j];
            const float dfeq = wj * fmaf(rho, P, drho);
            const float feq = fmaf(wj, rho0, dfeq);
            const float frc = EX[j] > 0 ? fps : (EX[j] < 0 ? fms : f0s);
            const float o = fmaf(c1, ds[j], fmaf(omega_s, dfeq, frc * feq));
            if constexpr (MODE == 0) {
                if constexpr (SR)
                    st_dev_sr(onode, fpos<LAYOUT>(j, 0, a.Npad), o, rnd + (uint32_t)j * 0x9e3779b9u);
                else
                    st_dev(onode, fpos<LAYOUT>(j, 0, a.Npad), o);
            } else {
                const int64_t k = out_pos<LAYOUT, MODE>(a, i, t, j);
                if constexpr (SR)
                    st_dev_sr(fout, k, o, rnd + (uint32_t)j * 0x9e3779b9u);
                else
                    st_dev(fout, k, o);
            }
        }
    };
    if constexpr (scaled) {
        if (__builtin_amdgcn_readfirstlane(a.stochastic))
            relax_and_store(std::true_type{});
        else
            relax_and_store(std::false_type{});
    } else {
        relax_and_store(std::false_type{});
    }
}

// Key of node i's stochastic-rounding stream (FP16 storage only): from the
// tile's runs of consecutive keys (one scalar load per wave), or per node in
// the tiles marked as not fitting three runs.  Called after the gathers are
// issued, so the scalar load's latency hides behind theirs.
template <typename S>
__device__ __forceinline__ uint32_t sr_key(const StepArgs& a, int64_t i)
{
    if constexpr (sizeof(S) == 2)
        if (a.stochastic && a.key_perm) {
            if (a.kseg) {
                const int64_t tile = __builtin_amdgcn_readfirstlane((int)(i >> 6));
                const uint32_t* d = (const uint32_t*)(a.kseg + tile);
                const uint32_t b0 = ld_const(d), b1 = ld_const(d + 1), b2 = ld_const(d + 2), st = ld_const(d + 3);
                const uint32_t lane = (uint32_t)i & 63u;
                const uint32_t base = lane >= ((st >> 8) & 0xffu) ? b2 : (lane >= (st & 0xffu) ? b1 : b0);
                if (!(st >> 31)) return base + lane;
            }
            return a.key_perm[i];
        }
    return (uint32_t)i;
}

template <typename S, int LAYOUT, int MODE = 0>
__device__ __forceinline__ void collide_any(const StepArgs& a, int64_t i, const int32_t t[Q], S raw[Q],
                                            unsigned long long fk, uint32_t key)
{
    if constexpr (sizeof(S) == 8)
        collide_f64<LAYOUT, MODE>(a, i, t, (double*)raw, fk);
    else
        collide_shift<S, LAYOUT, MODE>(a, i, t, raw, fk, key);
}

template <typename S, int LAYOUT, int IDX, bool CONST_DESC = false>
__device__ __forceinline__ void node_any(const StepArgs& a, int64_t i, unsigned long long fk)
{
    Entries<IDX> e;
    issue_entries<LAYOUT, IDX, CONST_DESC>(a, i, e);
    decode_entries<IDX>(a, i, e);
    S raw[Q];
    gather<S, LAYOUT>(a, i, e.t, raw);
    const uint32_t key = sr_key<S>(a, i);
    collide_any<S, LAYOUT>(a, i, e.t, raw, fk, key);
}

// Single lattice, one node per thread, XCD-contiguous block order.
template <typename S, int LAYOUT, int IDX>
__global__ void __launch_bounds__(1024) k_step(const StepArgs a)
{
    const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
    const int64_t i = blk * blockDim.x + threadIdx.x;
    if (i >= a.N) return;
    node_any<S, LAYOUT, IDX>(with_step(a), i, *a.fault);
}

// Single lattice, one buffer, AA pattern (half-way bounce-back).  The
// buffer alternates between two forms of the post-collision populations
// f*_j(y) of the last step: after a neighbour step (and at rest), f*_j(y)
// lies at out_pos<2>(y, j), which is the pre-collision population of the
// node downstream -- so the next step reads each node's own 19 slots (MODE
// 1, no pull table) and leaves f*_j(y) in the node's own slot re[j]
// (out_pos<1>).  The step after that (MODE 2) reads population j of node x
// from slot re[j] of the upstream node t[j] (x - e_j; its own slot j at a
// wall) and stores by out_pos<2>.  A node of a MODE 2 step reads and writes
// the same 19 locations, so nodes never race; every step stores the values
// the pull step computes.  c->cur names the form (1: out_pos<1>, 0:
// out_pos<2>); read-backs map through it.
template <typename S, int LAYOUT, int IDX, int MODE>
__device__ __forceinline__ void step_aa(const StepArgs& a0)
{
    const StepArgs a = with_step(a0);
    const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
    const int64_t i = blk * blockDim.x + threadIdx.x;
    if (i >= a.N) return;
    const S* __restrict__ f = (const S*)a.fin;
    S raw[Q];
    if constexpr (MODE == 1) {
#pragma unroll
        for (int j = 0; j < Q; j++) raw[j] = f[fpos<LAYOUT>(j, i, a.Npad)];
        const uint32_t key = sr_key<S>(a, i);
        const int32_t none[Q] = {};
        collide_any<S, LAYOUT, 1>(a, i, none, raw, *a.fault, key);
    } else {
        Entries<IDX> e;
        issue_entries<LAYOUT, IDX>(a, i, e);
        decode_entries<IDX>(a, i, e);
        raw[0] = f[fpos<LAYOUT>(0, i, a.Npad)];
#pragma unroll
        for (int j = 1; j < Q; j++) {
            const int64_t src = e.t[j] >= 0 ? fpos<LAYOUT>(RE[j], e.t[j], a.Npad) : fpos<LAYOUT>(j, i, a.Npad);
            raw[j] = f[src];
        }
        const uint32_t key = sr_key<S>(a, i);
        collide_any<S, LAYOUT, 2>(a, i, e.t, raw, *a.fault, key);
    }
}
template <typename S, int LAYOUT, int IDX>
__global__ void __launch_bounds__(1024) k_step_aa_even(const StepArgs a)
{
    step_aa<S, LAYOUT, IDX, 1>(a);
}
template <typename S, int LAYOUT, int IDX>
__global__ void __launch_bounds__(1024) k_step_aa_odd(const StepArgs a)
{
    step_aa<S, LAYOUT, IDX, 2>(a);
}

// Single lattice, two nodes per thread: each wave covers two consecutive
// 64-node tiles and keeps both tiles' index loads and gathers in flight
// together (more bytes in flight per wave at the cost of registers).
template <typename S, int LAYOUT, int IDX>
__global__ void __launch_bounds__(512) k_step2(const StepArgs a0)
{
    const StepArgs a = with_step(a0);
    const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
    const int64_t t = blk * blockDim.x + threadIdx.x;
    const int64_t i0 = (t >> 6) * 128 + (t & 63);
    const int64_t i1 = i0 + 64;
    if (i0 >= a.N) return;
    const bool two = i1 < a.N;
    const unsigned long long fk = *a.fault;
    Entries<IDX> e0, e1;
    issue_entries<LAYOUT, IDX>(a, i0, e0);
    if (two) issue_entries<LAYOUT, IDX>(a, i1, e1);
    decode_entries<IDX>(a, i0, e0);
    if (two) decode_entries<IDX>(a, i1, e1);
    S r0[Q], r1[Q];
    gather<S, LAYOUT>(a, i0, e0.t, r0);
    if (two) gather<S, LAYOUT>(a, i1, e1.t, r1);
    const uint32_t k0 = sr_key<S>(a, i0), k1 = two ? sr_key<S>(a, i1) : 0u;
    collide_any<S, LAYOUT>(a, i0, e0.t, r0, fk, k0);
    if (two) collide_any<S, LAYOUT>(a, i1, e1.t, r1, fk, k1);
}

// Batch of samples: the grid enumerates the blocks of active samples
// (active[]), each block lies inside one sample (samples are padded to a
// multiple of the block size), and the sample supplies omega and gx.
struct BatchArgs {
    const int32_t* active;       // global block ids of active samples
    const int32_t* block_sample; // global block -> sample
    const int64_t* sample_end;   // first node after each sample's nodes
    const double* omega;         // per sample
    const double* gx;            // per sample
};

template <typename S, int LAYOUT, int IDX>
__global__ void __launch_bounds__(1024) k_step_batch(StepArgs a, const BatchArgs b)
{
    const int64_t slot = xcd_block(blockIdx.x, gridDim.x);
    const int64_t gb = b.active[slot];
    const int s = b.block_sample[gb];
    const int64_t i = gb * blockDim.x + threadIdx.x;
    if (i >= b.sample_end[s]) return;
    a.omega = b.omega[s];
    a.gx = b.gx[s];
    node_any<S, LAYOUT, IDX>(a, i, *a.fault);
}

// ---------------------------------------------------------------- fixed-order sums
// Sum of one double per thread of the block, in a fixed order; the result
// is valid in thread 0.  lds holds blockDim.x / 64 doubles.
__device__ __forceinline__ double block_sum_fixed(double v, double* lds)
{
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += lds[w];
    __syncthreads();
    return t;
}

// Sum of a sample's chunk partials by one wave (lane 0 holds the result).
__device__ __forceinline__ double wave_sum_chunks(const double* partial, int64_t c0, int64_t nc)
{
    const int lane = threadIdx.x & 63;
    double acc = 0.0;
    for (int64_t k = lane; k < nc; k += 64) acc += partial[c0 + k];
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) acc += __shfl_xor(acc, m, 64);
    return acc;
}

// ---------------------------------------------------------------- persistent batch
// One launch runs every sample of a batch from its first step to its own
// convergence stop.  The grid is resident (cooperative launch); each
// work-group walks the chunks (blocks of the batch layout) of the active
// samples, with each XCD taking contiguous runs of chunks as in xcd_block.
// A grid barrier separates time steps.  On check steps every chunk stores
// its macros and the fixed-order sum of its u_x; after a barrier, one wave
// per sample forms the sample's sum (the same order as k_sample_sums),
// applies the convergence rule and retires the sample; a second barrier
// publishes the decisions.
//
// Barriers follow the cooperative grid-sync pattern: a block barrier, then
// one thread per block releases the block's stores at agent scope and
// arrives; after the whole grid has arrived it acquires, and a second block
// barrier lets the block go on.  Arrivals and waits are spread over eight
// groups of blocks (block id mod 8, the XCD a block is dispatched to), each
// on its own 128-byte line: the last arrival of a group carries the group
// to a grid-wide counter, and the last group releases every group's flag.
// All counters are monotonic (barrier `gen` completes at gen x members), so
// nothing is reset between barriers.  A wait is bounded by wall time
// (wall_clock64 ticks): a block that waits longer sets `abort`, every block
// leaves the loop at its next barrier, and the host reports the failure, so
// the kernel always ends.  The samples' running flags are copied to LDS
// after every barrier that can change them, so the chunk loop reads them
// without a memory round trip.
constexpr int PERSIST_MAX_SAMPLES = PLB_PERSIST_MAX_SAMPLES;
constexpr int BAR_STRIDE = 16;  // u64 words between counters: one 128-byte line each

__device__ __forceinline__ bool grid_barrier(const PersistArgs& p, unsigned long long& gen)
{
    __shared__ int ok;
    __syncthreads();
    gen++;
    if (threadIdx.x == 0) {
        const unsigned G = gridDim.x, g = blockIdx.x & 7u;
        const unsigned long long members = (G - 1 - g) / 8 + 1, groups = G < 8 ? G : 8;
        unsigned long long* cnt = p.bar + g * BAR_STRIDE;
        unsigned long long* all = p.bar + 8 * BAR_STRIDE;
        unsigned long long* rel = p.bar + 9 * BAR_STRIDE;
        int good = 1;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        if (__hip_atomic_fetch_add(cnt, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1 == gen * members) {
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
            if (__hip_atomic_fetch_add(all, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1 == gen * groups) {
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
                for (unsigned k = 0; k < groups; k++)
                    __hip_atomic_store(rel + k * BAR_STRIDE, gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        const uint64_t t0 = wall_clock64();
        while (__hip_atomic_load(rel + g * BAR_STRIDE, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < gen) {
            if (__hip_atomic_load(p.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                good = 0;
                break;
            }
            if (wall_clock64() - t0 > p.wait_ticks) {
                __hip_atomic_store(p.abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                good = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(8);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        ok = good;
    }
    __syncthreads();
    return ok;
}

template <typename S, int LAYOUT, int IDX>
__global__ void __launch_bounds__(1024) k_batch_persistent(StepArgs a, const PersistArgs p)
{
    __shared__ double lds[16];
    // aprefix[s] = active chunks of the samples before s (so sample s runs
    // iff aprefix[s + 1] > aprefix[s]); rebuilt after every decision barrier.
    __shared__ int aprefix[PERSIST_MAX_SAMPLES + 1];
    __shared__ int stop;
    __shared__ int64_t claim[2];   // the chunks this block claimed, in its stretch (alternating slots)
    const int bs = blockDim.x;
    // The active chunks, in sample order, are split into eight contiguous
    // stretches, one per XCD (block id mod 8).  The blocks of a stretch take
    // its chunks one at a time from the stretch's counter for this step
    // (line 17 + 2 * group + step parity of p.bar), so the XCD sweeps its
    // stretch in order and a block that finishes early takes more, as the
    // hardware dispatcher does in a launch per step.  Each block claims its
    // next chunk before it processes the current one, which hides the
    // atomic's round trip.  (Claims of 2, 4 or 8 chunks, spread so that the
    // chunks in flight stay consecutive, were slower.)
    const unsigned G = gridDim.x, grp = blockIdx.x & 7u, slot = blockIdx.x >> 3;
    const int64_t ngrp = G < 8 ? G : 8;   // groups that have blocks
    unsigned long long gen = 0;
    // Residency check before any work: every block counts itself in and
    // waits, briefly, for the whole grid.  One decision word settles the
    // outcome for all blocks: the first block to see the full count sets it
    // to "go", the first to wait past start_ticks sets it to "leave" (a grid
    // larger than what the CUs really hold never completes the count), and
    // every block then follows it.  On "leave" the blocks exit before any
    // work with abort = 2, and the host launches a smaller grid.
    if (threadIdx.x == 0) {
        unsigned long long* started = p.bar + 33 * BAR_STRIDE;
        unsigned long long* decision = started + 1;   // 0 open, 1 go, 2 leave
        __hip_atomic_fetch_add(started, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t t0 = wall_clock64();
        unsigned long long d;
        while ((d = __hip_atomic_load(decision, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == 0) {
            unsigned long long open = 0;
            if (__hip_atomic_load(started, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= G)
                __hip_atomic_compare_exchange_strong(decision, &open, 1ull, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
            else if (wall_clock64() - t0 > p.start_ticks)
                __hip_atomic_compare_exchange_strong(decision, &open, 2ull, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
            else
                __builtin_amdgcn_s_sleep(8);
        }
        if (d == 2) __hip_atomic_store(p.abort, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        stop = d != 1;
    }
    __syncthreads();
    if (stop) return;
    auto rebuild = [&]() {
        if (threadIdx.x == 0) {
            int acc = 0;
            for (int k = 0; k < p.nsamples; k++) {
                aprefix[k] = acc;
                if (p.active[k]) acc += (int)((p.sample_n[k] + bs - 1) / bs);
            }
            aprefix[p.nsamples] = acc;
        }
        __syncthreads();
    };
    rebuild();
    for (int64_t t = 0; t < p.max_steps; t++) {
        const bool check = (t + 1) % p.check_every == 0 || t + 1 == p.max_steps;
        const int par = (p.cur0 + (int)(t & 1)) & 1;
        a.fin = p.f[par];
        a.fout = p.f[par ^ 1];
        a.it = t;
        a.write_macros = check ? 1 : 0;
        const int64_t total = aprefix[p.nsamples];
        const int64_t j0 = grp * total / ngrp, j1 = (grp + 1) * total / ngrp;
        unsigned long long* claim_cnt = p.bar + (17 + 2 * grp + (t & 1)) * BAR_STRIDE;
        // This step's counter was zeroed in the step before (or by the host);
        // the other parity's is zeroed now for the next step, by one block of
        // the group: the barrier at the end of this step orders that store
        // before any claim of the next step.
        if (slot == 0 && threadIdx.x == 0)
            __hip_atomic_store(p.bar + (17 + 2 * grp + ((t + 1) & 1)) * BAR_STRIDE, 0ull, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        const int64_t len = j1 - j0;
        if (threadIdx.x == 0)
            claim[0] = (int64_t)__hip_atomic_fetch_add(claim_cnt, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        int cs = 0;   // claim slot being processed
        int64_t mine = claim[0];
        int s = 0;
        while (mine < len) {
            // The other slot was last read before the previous barrier.
            if (threadIdx.x == 0)
                claim[cs ^ 1] = (int64_t)__hip_atomic_fetch_add(claim_cnt, 1ull, __ATOMIC_RELAXED,
                                                                __HIP_MEMORY_SCOPE_AGENT);
            const int64_t j = j0 + mine;
            while (aprefix[s + 1] <= j) s++;   // uniform per block; a block's claims only grow
            const int64_t chunk = ld_const(p.sample_off + s) / bs + (j - aprefix[s]);
            const int64_t i = chunk * bs + threadIdx.x;
            const bool live = i < ld_const(p.sample_end + s);
            if (live) {
                a.omega = ld_const(p.omega + s);
                a.gx = ld_const(p.gx + s);
                node_any<S, LAYOUT, IDX, true>(a, i, NO_FAULT);
            }
            if (check) {
                const double t_sum = block_sum_fixed(live ? a.macros[a.Npad + i] : 0.0, lds);
                if (threadIdx.x == 0) p.partial[chunk] = t_sum;
            }
            __syncthreads();   // the next claim has landed in LDS
            cs ^= 1;
            mine = claim[cs];
        }
        if (!grid_barrier(p, gen)) return;
        // Density guard: stop after the faulting step.  Only a fault of this
        // step or earlier counts: a faster block may already be in the next
        // step, and every block must leave at the same barrier.
        if (threadIdx.x == 0) {
            const unsigned long long fk = __hip_atomic_load(a.fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            stop = fk != NO_FAULT && (int64_t)(fk >> 32) <= t;
        }
        __syncthreads();
        if (stop) return;
        if (!check) continue;
        if (threadIdx.x < 64) {
            for (int64_t q = xcd_block(blockIdx.x, G); q < p.nsamples; q += G) {
                if (aprefix[q + 1] == aprefix[q]) continue;   // not running
                const double sum = wave_sum_chunks(p.partial, p.sample_off[q] / bs, (p.sample_n[q] + bs - 1) / bs);
                if (threadIdx.x == 0) {
                    const double mean = sum / (double)p.sample_n[q];
                    const double scale = fmax(fabs(mean), 1e-3 * fabs(p.gx[q]));
                    const double prev = p.prev[q];
                    p.steps[q] = t + 1;
                    p.cur[q] = par ^ 1;
                    if (!isnan(prev) && fabs(mean - prev) <= p.tol * scale) {
                        p.active[q] = 0;
                        __hip_atomic_fetch_add(p.n_active, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                    p.prev[q] = mean;
                }
            }
        }
        if (!grid_barrier(p, gen)) return;
        if (threadIdx.x == 0) stop = __hip_atomic_load(p.n_active, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
        __syncthreads();
        if (stop) return;
        rebuild();
    }
}

// ---------------------------------------------------------------- init / IO kernels
template <typename S, int LAYOUT>
__global__ void k_init_eq(void* f, int64_t N, int64_t Npad, double rho0)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Npad) return;
#pragma unroll
    for (int j = 0; j < Q; j++) {
        const int64_t k = fpos<LAYOUT>(j, i, Npad);
        if constexpr (sizeof(S) == 8)
            ((double*)f)[k] = feq_ref(j, rho0, 0.0, 0.0, 0.0);  // init_dist_macro
        else if constexpr (sizeof(S) == 4)
            ((float*)f)[k] = 0.f;
        else
            ((__half*)f)[k] = __float2half(0.f);
    }
}

// AoS [n][19] FP64 in V order  <->  device layout.  `first` = first V-order
// node of the chunk, `n` nodes.
template <typename S, int LAYOUT>
__global__ void k_read_pop(const void* f, int64_t Npad, const uint32_t* inv_perm, int64_t first,
                           int64_t n, double rho0, float inv, double* out)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t g = first + r;
    const int64_t k = inv_perm ? (int64_t)inv_perm[g] : g;
#pragma unroll
    for (int j = 0; j < Q; j++) {
        const int64_t p = fpos<LAYOUT>(j, k, Npad);
        double v;
        if constexpr (sizeof(S) == 8)
            v = ((const double*)f)[p];
        else if constexpr (sizeof(S) == 4)
            v = TP[j] * rho0 + (double)((const float*)f)[p];
        else
            v = TP[j] * rho0 + (double)(__half2float(((const __half*)f)[p]) * inv);
        out[r * Q + j] = v;
    }
}

template <typename S, int LAYOUT>
__global__ void k_write_pop(void* f, int64_t Npad, const uint32_t* inv_perm, int64_t first,
                            int64_t n, double rho0, float sc, const double* in)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t g = first + r;
    const int64_t k = inv_perm ? (int64_t)inv_perm[g] : g;
#pragma unroll
    for (int j = 0; j < Q; j++) {
        const int64_t p = fpos<LAYOUT>(j, k, Npad);
        const double v = in[r * Q + j];
        if constexpr (sizeof(S) == 8)
            ((double*)f)[p] = v;
        else if constexpr (sizeof(S) == 4)
            ((float*)f)[p] = (float)(v - TP[j] * rho0);
        else
            ((__half*)f)[p] = __float2half_rn((float)(v - TP[j] * rho0) * sc);
    }
}

// Stored elements as they are, AoS [n][19] in V order <-> device layout
// (checkpoints: no conversion, so a written-back state is bit-identical).
template <typename S, int LAYOUT>
__global__ void k_copy_state(void* f, int64_t Npad, const uint32_t* inv_perm, int64_t first, int64_t n,
                             void* aos_bytes, int to_device)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t g = first + r;
    const int64_t k = inv_perm ? (int64_t)inv_perm[g] : g;
    S* dev = (S*)f;
    S* aos = (S*)aos_bytes;
#pragma unroll
    for (int j = 0; j < Q; j++) {
        const int64_t p = fpos<LAYOUT>(j, k, Npad);
        if (to_device)
            dev[p] = aos[r * Q + j];
        else
            aos[r * Q + j] = dev[p];
    }
}

// The same three copies for a one-buffer AA lattice: population j of node k
// sits at out_pos<form>(k, j), form 1 or 2 (the step kernels' store map).
template <int LAYOUT, int IDX>
__device__ __forceinline__ void aa_positions(const StepArgs& a, int64_t k, int form, int64_t pos[Q]);

template <typename S, int LAYOUT, int IDX>
__global__ void k_read_pop_aa(const StepArgs a, int form, const uint32_t* inv_perm, int64_t first, int64_t n,
                              double rho0, float inv, double* out)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t g = first + r;
    const int64_t k = inv_perm ? (int64_t)inv_perm[g] : g;
    int64_t pos[Q];
    aa_positions<LAYOUT, IDX>(a, k, form, pos);
#pragma unroll
    for (int j = 0; j < Q; j++) {
        double v;
        if constexpr (sizeof(S) == 8)
            v = ((const double*)a.fin)[pos[j]];
        else if constexpr (sizeof(S) == 4)
            v = TP[j] * rho0 + (double)((const float*)a.fin)[pos[j]];
        else
            v = TP[j] * rho0 + (double)(__half2float(((const __half*)a.fin)[pos[j]]) * inv);
        out[r * Q + j] = v;
    }
}

template <typename S, int LAYOUT, int IDX>
__global__ void k_write_pop_aa(const StepArgs a, int form, const uint32_t* inv_perm, int64_t first, int64_t n,
                               double rho0, float sc, const double* in)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t g = first + r;
    const int64_t k = inv_perm ? (int64_t)inv_perm[g] : g;
    int64_t pos[Q];
    aa_positions<LAYOUT, IDX>(a, k, form, pos);
#pragma unroll
    for (int j = 0; j < Q; j++) {
        const double v = in[r * Q + j];
        if constexpr (sizeof(S) == 8)
            ((double*)a.fout)[pos[j]] = v;
        else if constexpr (sizeof(S) == 4)
            ((float*)a.fout)[pos[j]] = (float)(v - TP[j] * rho0);
        else
            ((__half*)a.fout)[pos[j]] = __float2half_rn((float)(v - TP[j] * rho0) * sc);
    }
}

template <typename S, int LAYOUT, int IDX>
__global__ void k_copy_state_aa(const StepArgs a, int form, const uint32_t* inv_perm, int64_t first, int64_t n,
                                void* aos_bytes, int to_device)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t g = first + r;
    const int64_t k = inv_perm ? (int64_t)inv_perm[g] : g;
    int64_t pos[Q];
    aa_positions<LAYOUT, IDX>(a, k, form, pos);
    S* dev = (S*)a.fout;
    S* aos = (S*)aos_bytes;
#pragma unroll
    for (int j = 0; j < Q; j++) {
        if (to_device)
            dev[pos[j]] = aos[r * Q + j];
        else
            aos[r * Q + j] = dev[pos[j]];
    }
}

__global__ void k_read_macros(const double* macros, int64_t Npad, const uint32_t* inv_perm,
                              int64_t first, int64_t n, int comp, double* out)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t g = first + r;
    const int64_t k = inv_perm ? (int64_t)inv_perm[g] : g;
    out[r] = macros[comp * Npad + k];
}

// Pull sources t[1..18] of internal node k, decoded by one lane on its own
// (the lanes of a wave may hold unrelated nodes).
template <int LAYOUT, int IDX>
__device__ __forceinline__ void lane_sources(const StepArgs& a, int64_t k, int32_t t[Q])
{
    if constexpr (IDX == PLB_INDEX_I32) {
        Entries<IDX> e;
        issue_entries<LAYOUT, IDX>(a, k, e);
        for (int j = 1; j < Q; j++) t[j] = e.t[j];
    } else if constexpr (IDX == PLB_INDEX_BITMAP) {
        uint32_t x, y, z, ly;
        bitmap_coords(a, a.vox[k], x, y, z, ly);
        for (int j = 1; j < Q; j++) {
            const uint32_t vs = bitmap_src_general(a, j, x, y, z);
            t[j] = bitmap_lookup(a.bm[vs >> 5], vs);
        }
    } else {
        // Lanes of a wave hold unrelated nodes here: decode without the
        // per-wave descriptor shortcut.
        const int64_t tile = k >> 6;
        const int lane = (int)(k & 63);
        for (int j = 1; j < Q; j++) {
            const int32_t desc = a.tdesc[tile * (Q - 1) + j - 1];
            if (desc >= 0) {
                const int wi = (j - 1) >> 2;  // packed word holding direction j
                const uint32_t word = wi < TILE8_FULL ? a.toff[(tile * TILE8_FULL + wi) * 64 + lane]
                                                      : (uint32_t)a.toff16[tile * 64 + lane];
                const int32_t v = (int32_t)((word >> (8 * ((j - 1) & 3))) & 0xffu);
                t[j] = v == 255 ? TAB_BOUNCE : (v == 254 ? TAB_CONST : desc + v);
            } else {
                t[j] = a.twide[(int64_t)(-desc - 1) * 64 + lane];
            }
        }
    }
}

template <int LAYOUT, int IDX>
__device__ __forceinline__ void aa_positions(const StepArgs& a, int64_t k, int form, int64_t pos[Q])
{
    if (form == 1) {
#pragma unroll
        for (int j = 0; j < Q; j++) pos[j] = out_pos<LAYOUT, 1>(a, k, nullptr, j);
    } else {
        int32_t t[Q];
        t[0] = 0;
        lane_sources<LAYOUT, IDX>(a, k, t);
#pragma unroll
        for (int j = 0; j < Q; j++) pos[j] = out_pos<LAYOUT, 2>(a, k, t, j);
    }
}

template <int LAYOUT, int IDX>
__global__ void k_read_table(const StepArgs a, const uint32_t* inv_perm, const uint32_t* perm,
                             int64_t first, int64_t n, int32_t* out)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t g = first + r;
    const int64_t k = inv_perm ? (int64_t)inv_perm[g] : g;
    int32_t t[Q];
    lane_sources<LAYOUT, IDX>(a, k, t);
    for (int j = 1; j < Q; j++) {
        int32_t v = t[j];
        if (v >= 0 && perm) v = (int32_t)perm[v];
        out[(int64_t)(j - 1) * n + r] = v;
    }
}

// ---------------------------------------------------------------- FP64 reductions
constexpr int RED_T = 256;

// Block partial sums over nodes [first, first+n): mass (sum of the stored
// populations; skipped when with_mass == 0) and ux.
template <typename S, int LAYOUT>
__global__ void __launch_bounds__(RED_T)
k_reduce_partial(const void* f, const double* ux, int64_t first, int64_t n, int64_t Npad,
                 float inv, int with_mass, double* partial)
{
    __shared__ double sm[2][RED_T];
    double m = 0.0, s = 0.0;
    for (int64_t r = (int64_t)blockIdx.x * RED_T + threadIdx.x; r < n;
         r += (int64_t)gridDim.x * RED_T) {
        const int64_t i = first + r;
        if (with_mass) {
            double mi = 0.0;
#pragma unroll
            for (int j = 0; j < Q; j++) {
                const int64_t p = fpos<LAYOUT>(j, i, Npad);
                if constexpr (sizeof(S) == 8)
                    mi += ((const double*)f)[p];
                else if constexpr (sizeof(S) == 4)
                    mi += (double)((const float*)f)[p];
                else
                    mi += (double)(__half2float(((const __half*)f)[p]) * inv);
            }
            m += mi;
        }
        s += ux[i];
    }
    sm[0][threadIdx.x] = m;
    sm[1][threadIdx.x] = s;
    __syncthreads();
    for (int off = RED_T / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            sm[0][threadIdx.x] += sm[0][threadIdx.x + off];
            sm[1][threadIdx.x] += sm[1][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = sm[0][0];
        partial[2 * blockIdx.x + 1] = sm[1][0];
    }
}

// ---------------------------------------------------------------- per-sample u_x sums (batches)
// A sample's sum of u_x is formed in one fixed order by both batch drivers
// (one launch per step, and the persistent launch): per chunk (one block of
// the batch layout) a fixed tree over its nodes, then per sample the chunk
// partials, lane l of one wave taking chunks l, l+64, ... in turn, and a
// fixed butterfly over the 64 lanes.  Both drivers therefore see the same
// sums bit for bit and stop every sample at the same step.

// partial[chunk] = sum of ux over the chunk's nodes, for the listed chunks.
__global__ void __launch_bounds__(1024)
k_chunk_partials(const double* ux, const int32_t* chunks, const int32_t* chunk_sample,
                 const int64_t* sample_end, double* partial)
{
    __shared__ double lds[16];
    const int64_t chunk = chunks[blockIdx.x];
    const int64_t i = chunk * blockDim.x + threadIdx.x;
    const double v = i < sample_end[chunk_sample[chunk]] ? ux[i] : 0.0;
    const double t = block_sum_fixed(v, lds);
    if (threadIdx.x == 0) partial[chunk] = t;
}

// out[s] = sum of sample s's chunk partials (one wave per sample).
__global__ void __launch_bounds__(64)
k_sample_sums(const double* partial, const int64_t* off, const int64_t* n, int bs, double* out)
{
    const int s = blockIdx.x;
    const double t = wave_sum_chunks(partial, off[s] / bs, (n[s] + bs - 1) / bs);
    if (threadIdx.x == 0) out[s] = t;
}

__global__ void __launch_bounds__(RED_T) k_reduce_final(double* partial, int nb)
{
    __shared__ double sm[2][RED_T];
    double m = 0.0, s = 0.0;
    for (int b = threadIdx.x; b < nb; b += RED_T) {
        m += partial[2 * b];
        s += partial[2 * b + 1];
    }
    sm[0][threadIdx.x] = m;
    sm[1][threadIdx.x] = s;
    __syncthreads();
    for (int off = RED_T / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            sm[0][threadIdx.x] += sm[0][threadIdx.x + off];
            sm[1][threadIdx.x] += sm[1][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * nb] = sm[0][0];
        partial[2 * nb + 1] = sm[1][0];
    }
}

// ---------------------------------------------------------------- launchers
static unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

#define PLB_DISPATCH(PREC, LAYOUT, KERNEL, ...)                                             \
    do {                                                                                    \
        if ((PREC) == PLB_FP64) {                                                           \
            if ((LAYOUT) == PLB_LAYOUT_SOA) hipLaunchKernelGGL((KERNEL<double, PLB_LAYOUT_SOA>), __VA_ARGS__); \
            else hipLaunchKernelGGL((KERNEL<double, PLB_LAYOUT_AOSOA>), __VA_ARGS__);        \
        } else if ((PREC) == PLB_FP32) {                                                    \
            if ((LAYOUT) == PLB_LAYOUT_SOA) hipLaunchKernelGGL((KERNEL<float, PLB_LAYOUT_SOA>), __VA_ARGS__); \
            else hipLaunchKernelGGL((KERNEL<float, PLB_LAYOUT_AOSOA>), __VA_ARGS__);         \
        } else {                                                                            \
            if ((LAYOUT) == PLB_LAYOUT_SOA) hipLaunchKernelGGL((KERNEL<__half, PLB_LAYOUT_SOA>), __VA_ARGS__); \
            else hipLaunchKernelGGL((KERNEL<__half, PLB_LAYOUT_AOSOA>), __VA_ARGS__);        \
        }                                                                                   \
    } while (0)

#define PLB_DISPATCH_ONE_IDX(PREC, LAYOUT, IDXC, KERNEL, ...)                                  \
    do {                                                                                        \
        if ((PREC) == PLB_FP64) {                                                               \
            if ((LAYOUT) == PLB_LAYOUT_SOA) hipLaunchKernelGGL((KERNEL<double, PLB_LAYOUT_SOA, IDXC>), __VA_ARGS__); \
            else hipLaunchKernelGGL((KERNEL<double, PLB_LAYOUT_AOSOA, IDXC>), __VA_ARGS__);      \
        } else if ((PREC) == PLB_FP32) {                                                        \
            if ((LAYOUT) == PLB_LAYOUT_SOA) hipLaunchKernelGGL((KERNEL<float, PLB_LAYOUT_SOA, IDXC>), __VA_ARGS__); \
            else hipLaunchKernelGGL((KERNEL<float, PLB_LAYOUT_AOSOA, IDXC>), __VA_ARGS__);       \
        } else {                                                                                \
            if ((LAYOUT) == PLB_LAYOUT_SOA) hipLaunchKernelGGL((KERNEL<__half, PLB_LAYOUT_SOA, IDXC>), __VA_ARGS__); \
            else hipLaunchKernelGGL((KERNEL<__half, PLB_LAYOUT_AOSOA, IDXC>), __VA_ARGS__);      \
        }                                                                                       \
    } while (0)

// All three index kinds (single lattice).
#define PLB_DISPATCH3(PREC, LAYOUT, IDX, KERNEL, ...)                                         \
    do {                                                                                      \
        if ((IDX) == PLB_INDEX_I32) PLB_DISPATCH_ONE_IDX(PREC, LAYOUT, PLB_INDEX_I32, KERNEL, __VA_ARGS__); \
        else if ((IDX) == PLB_INDEX_TILE8) PLB_DISPATCH_ONE_IDX(PREC, LAYOUT, PLB_INDEX_TILE8, KERNEL, __VA_ARGS__); \
        else PLB_DISPATCH_ONE_IDX(PREC, LAYOUT, PLB_INDEX_BITMAP, KERNEL, __VA_ARGS__);        \
    } while (0)

// Table index kinds only (batches of samples).
#define PLB_DISPATCH2(PREC, LAYOUT, IDX, KERNEL, ...)                                         \
    do {                                                                                      \
        if ((IDX) == PLB_INDEX_I32) PLB_DISPATCH_ONE_IDX(PREC, LAYOUT, PLB_INDEX_I32, KERNEL, __VA_ARGS__); \
        else PLB_DISPATCH_ONE_IDX(PREC, LAYOUT, PLB_INDEX_TILE8, KERNEL, __VA_ARGS__);         \
    } while (0)

static StepArgs make_args(plb_ctx* c, int64_t it, bool write_macros)
{
    StepArgs a;
    a.fin = c->f[c->cur];
    a.fout = c->f[c->cur ^ 1];
    a.tab = c->tab;
    a.tdesc = c->tdesc;
    a.toff = c->toff;
    a.toff16 = c->toff16;
    a.twide = c->twide;
    a.N = c->N;
    a.Npad = c->Npad;
    a.omega = c->omega;
    a.gx = c->cfg.gx;
    a.rho0 = c->cfg.rho0;
    a.it = it;
    a.write_macros = write_macros ? 1 : 0;
    a.macros = c->macros;
    a.fault = c->fault;
    a.perm = c->perm;
    a.scale = c->scale;
    a.inv_scale = c->inv_scale;
    a.bm = c->bm;
    a.vox = c->vox;
    a.nx = (uint32_t)c->nx;
    a.ny = (uint32_t)c->ny;
    a.nz = (uint32_t)c->nz;
    a.py = (uint32_t)(c->py > 0 ? c->py : c->ny);
    a.div_nx = make_fastdiv(a.nx);
    a.div_py = make_fastdiv(a.py);
    a.div_nz = make_fastdiv(a.nz);
    a.it_base = nullptr;
    a.stochastic = c->cfg.precision == PLB_FP16 && c->cfg.fp16_rounding == PLB_ROUND_STOCHASTIC;
    a.key_perm = c->key_perm ? c->key_perm : c->perm;
    a.kseg = c->kseg;
    a.has_const = c->cfg.bb_mode == PLB_BB_REFCOMPAT;
    return a;
}

// Dynamic LDS per block that caps the step kernel at cfg.waves_per_simd waves
// per SIMD (the kernel itself uses no LDS): with w waves on each of a CU's 4
// SIMDs the CU holds 4*w*64/bs blocks, so each block reserves that share of
// the CU's LDS, rounded down to 512 bytes (at most 64 KB).  0 = no cap.
static unsigned step_lds_bytes(const plb_ctx* c, int bs)
{
    const int w = c->cfg.waves_per_simd;
    if (w <= 0 || w >= 8 || c->lds_per_cu <= 0) return 0u;
    const int blocks = std::max(1, 4 * w * 64 / bs);
    const long share = std::min<long>(65536, (long)c->lds_per_cu / blocks);
    return (unsigned)(share / 512 * 512);
}

int launch_step(plb_ctx* c, int64_t it, bool write_macros, const int64_t* it_base)
{
    if (c->N == 0) {  // a mask without pores: nothing to update
        c->cur ^= 1;
        return PLB_OK;
    }
    StepArgs a = make_args(c, it, write_macros);
    a.it_base = it_base;
    const int bs = c->cfg.block_size;
    const unsigned lds = step_lds_bytes(c, bs);
    if (c->cfg.stream_mode == PLB_STREAM_AA) {
        // one buffer: c->cur 1 = local-step store form (a neighbour step
        // next), 0 = neighbour-step store form (a local step next)
        if (c->cur == 0)
            PLB_DISPATCH_ONE_IDX(c->cfg.precision, c->cfg.layout, PLB_INDEX_I32, k_step_aa_even,
                                 dim3(nblk(c->N, bs)), dim3(bs), lds, c->stream, a);
        else
            PLB_DISPATCH3(c->cfg.precision, c->cfg.layout, c->cfg.index_kind, k_step_aa_odd, dim3(nblk(c->N, bs)),
                          dim3(bs), lds, c->stream, a);
    } else if (c->cfg.nodes_per_thread == 2) {
        const int64_t threads = (c->N + 127) / 128 * 64;
        PLB_DISPATCH3(c->cfg.precision, c->cfg.layout, c->cfg.index_kind, k_step2,
                      dim3(nblk(threads, bs)), dim3(bs), lds, c->stream, a);
    } else {
        PLB_DISPATCH3(c->cfg.precision, c->cfg.layout, c->cfg.index_kind, k_step, dim3(nblk(c->N, bs)),
                      dim3(bs), lds, c->stream, a);
    }
    PLB_HIP(hipGetLastError());
    c->cur ^= 1;
    return PLB_OK;
}

__global__ void k_set_step(int64_t* it_dev, int64_t it)
{
    if (threadIdx.x == 0) *it_dev = it;
}

// G steps (G even, so the buffers end where they started) from a graph
// captured once per starting buffer.  The step number is not baked into the
// graph: every kernel adds *it_dev, which a plain launch sets before each
// replay.  The graph holds kernel nodes only.
int launch_graph_steps(plb_ctx* c, int64_t it0, int64_t G)
{
    hipStream_t s = c->stream;
    const int p = c->cur;
    if (!c->gexec[p]) {
        PLB_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int rc = PLB_OK;
        for (int64_t k = 0; k < G && rc == PLB_OK; k++) rc = launch_step(c, k, false, c->it_dev);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s, &g);
        c->cur = p;  // capture does not run anything
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        PLB_HIP(e);
        const hipError_t ei = hipGraphInstantiate(&c->gexec[p], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        PLB_HIP(ei);
    }
    hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, s, c->it_dev, it0);
    PLB_HIP(hipGetLastError());
    PLB_HIP(hipGraphLaunch(c->gexec[p], s));
    return PLB_OK;
}

// One step over the active blocks of a batch (the context holds the
// concatenated samples; c->omega / c->cfg.gx are unused).
int launch_step_batch(plb_ctx* c, int64_t it, bool write_macros, const int32_t* active,
                      int n_active, const int32_t* block_sample, const int64_t* sample_end,
                      const double* omega, const double* gx)
{
    if (n_active > 0) {
        const StepArgs a = make_args(c, it, write_macros);
        BatchArgs b;
        b.active = active;
        b.block_sample = block_sample;
        b.sample_end = sample_end;
        b.omega = omega;
        b.gx = gx;
        PLB_DISPATCH2(c->cfg.precision, c->cfg.layout, c->cfg.index_kind, k_step_batch,
                      dim3(n_active), dim3(c->cfg.block_size), step_lds_bytes(c, c->cfg.block_size),
                      c->stream, a, b);
        PLB_HIP(hipGetLastError());
    }
    c->cur ^= 1;
    return PLB_OK;
}


// The persistent batch run as one cooperative launch of as many blocks as can
// be resident at once (at most one per chunk).  *blocks = the grid size.
template <typename S, int LAYOUT, int IDX>
static int launch_persistent_t(plb_ctx* c, const PersistArgs& p, int max_blocks, int* blocks)
{
    auto kern = k_batch_persistent<S, LAYOUT, IDX>;
    const int bs = c->cfg.block_size;
    // The wave cap's LDS share per block, less what the kernel holds itself.
    hipFuncAttributes fa;
    PLB_HIP(hipFuncGetAttributes(&fa, (const void*)kern));
    unsigned lds = step_lds_bytes(c, bs);
    lds = lds > fa.sharedSizeBytes ? (unsigned)((lds - fa.sharedSizeBytes) / 512 * 512) : 0u;
    int per_cu = 0, cus = 0;
    PLB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, bs, lds));
    PLB_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    // One block per CU below the computed residency: a grid at exactly the
    // computed limit waited at its first barrier for blocks that never
    // started (the limit is an estimate; the bounded wait reported it).
    if (per_cu > 1) per_cu--;
    int64_t g = std::min<int64_t>((int64_t)per_cu * cus, p.nchunks);
    if (max_blocks > 0) g = std::min<int64_t>(g, max_blocks);
    if (g < 1) return PLB_ERR_INVALID;
    StepArgs a = make_args(c, 0, false);
    PersistArgs pp = p;
    void* args[] = {(void*)&a, (void*)&pp};
    PLB_HIP(hipLaunchCooperativeKernel((const void*)kern, dim3((unsigned)g), dim3(bs), args, lds, c->stream));
    *blocks = (int)g;
    return PLB_OK;
}

int launch_batch_persistent(plb_ctx* c, const PersistArgs& p, int max_blocks, int* blocks)
{
    const int pr = c->cfg.precision, lay = c->cfg.layout, idx = c->cfg.index_kind;
#define PLB_PERSIST(ST, L, I) if (pr == (sizeof(ST) == 8 ? PLB_FP64 : sizeof(ST) == 4 ? PLB_FP32 : PLB_FP16) && lay == L && idx == I) return launch_persistent_t<ST, L, I>(c, p, max_blocks, blocks)
    PLB_PERSIST(double, PLB_LAYOUT_SOA, PLB_INDEX_I32);
    PLB_PERSIST(double, PLB_LAYOUT_SOA, PLB_INDEX_TILE8);
    PLB_PERSIST(double, PLB_LAYOUT_AOSOA, PLB_INDEX_I32);
    PLB_PERSIST(double, PLB_LAYOUT_AOSOA, PLB_INDEX_TILE8);
    PLB_PERSIST(float, PLB_LAYOUT_SOA, PLB_INDEX_I32);
    PLB_PERSIST(float, PLB_LAYOUT_SOA, PLB_INDEX_TILE8);
    PLB_PERSIST(float, PLB_LAYOUT_AOSOA, PLB_INDEX_I32);
    PLB_PERSIST(float, PLB_LAYOUT_AOSOA, PLB_INDEX_TILE8);
    PLB_PERSIST(__half, PLB_LAYOUT_SOA, PLB_INDEX_I32);
    PLB_PERSIST(__half, PLB_LAYOUT_SOA, PLB_INDEX_TILE8);
    PLB_PERSIST(__half, PLB_LAYOUT_AOSOA, PLB_INDEX_I32);
    PLB_PERSIST(__half, PLB_LAYOUT_AOSOA, PLB_INDEX_TILE8);
#undef PLB_PERSIST
    return PLB_ERR_INVALID;
}

int launch_init_eq(plb_ctx* c)
{
    for (int b = 0; b < 2; b++)
        PLB_DISPATCH(c->cfg.precision, c->cfg.layout, k_init_eq, dim3(nblk(c->Npad, 256)), dim3(256),
                     0, c->stream, c->f[b], c->N, c->Npad, c->cfg.rho0);
    PLB_HIP(hipGetLastError());
    return PLB_OK;
}

int launch_read_pop(plb_ctx* c, int64_t first, int64_t n, double* dout)
{
    if (c->cfg.stream_mode == PLB_STREAM_AA) {
        PLB_DISPATCH3(c->cfg.precision, c->cfg.layout, c->cfg.index_kind, k_read_pop_aa, dim3(nblk(n, 256)),
                      dim3(256), 0, c->stream, make_args(c, 0, false), c->cur == 1 ? 1 : 2, c->inv_perm, first, n,
                      c->cfg.rho0, c->inv_scale, dout);
        PLB_HIP(hipGetLastError());
        return PLB_OK;
    }
    PLB_DISPATCH(c->cfg.precision, c->cfg.layout, k_read_pop, dim3(nblk(n, 256)), dim3(256), 0,
                 c->stream, c->f[c->cur], c->Npad, c->inv_perm, first, n, c->cfg.rho0,
                 c->inv_scale, dout);
    PLB_HIP(hipGetLastError());
    return PLB_OK;
}

int launch_write_pop(plb_ctx* c, int64_t first, int64_t n, const double* din)
{
    if (c->cfg.stream_mode == PLB_STREAM_AA) {
        PLB_DISPATCH3(c->cfg.precision, c->cfg.layout, c->cfg.index_kind, k_write_pop_aa, dim3(nblk(n, 256)),
                      dim3(256), 0, c->stream, make_args(c, 0, false), c->cur == 1 ? 1 : 2, c->inv_perm, first, n,
                      c->cfg.rho0, c->scale, din);
        PLB_HIP(hipGetLastError());
        return PLB_OK;
    }
    PLB_DISPATCH(c->cfg.precision, c->cfg.layout, k_write_pop, dim3(nblk(n, 256)), dim3(256), 0,
                 c->stream, c->f[c->cur], c->Npad, c->inv_perm, first, n, c->cfg.rho0, c->scale,
                 din);
    PLB_HIP(hipGetLastError());
    return PLB_OK;
}

int launch_copy_state(plb_ctx* c, int64_t first, int64_t n, void* aos, bool to_device)
{
    if (c->cfg.stream_mode == PLB_STREAM_AA) {
        PLB_DISPATCH3(c->cfg.precision, c->cfg.layout, c->cfg.index_kind, k_copy_state_aa, dim3(nblk(n, 256)),
                      dim3(256), 0, c->stream, make_args(c, 0, false), c->cur == 1 ? 1 : 2, c->inv_perm, first, n,
                      aos, to_device ? 1 : 0);
        PLB_HIP(hipGetLastError());
        return PLB_OK;
    }
    PLB_DISPATCH(c->cfg.precision, c->cfg.layout, k_copy_state, dim3(nblk(n, 256)), dim3(256), 0,
                 c->stream, c->f[c->cur], c->Npad, c->inv_perm, first, n, aos, to_device ? 1 : 0);
    PLB_HIP(hipGetLastError());
    return PLB_OK;
}

int launch_read_macros(plb_ctx* c, int comp, int64_t first, int64_t n, double* dout)
{
    hipLaunchKernelGGL(k_read_macros, dim3(nblk(n, 256)), dim3(256), 0, c->stream, c->macros,
                       c->Npad, c->inv_perm, first, n, comp, dout);
    PLB_HIP(hipGetLastError());
    return PLB_OK;
}

#define PLB_READ_TABLE(LAYOUT, IDXC)                                                       \
    hipLaunchKernelGGL((k_read_table<LAYOUT, IDXC>), g, b, 0, c->stream, a, c->inv_perm, c->perm, \
                       first, n, dout)

int launch_read_table(plb_ctx* c, int64_t first, int64_t n, int32_t* dout)
{
    const StepArgs a = make_args(c, 0, false);
    const dim3 g(nblk(n, 256)), b(256);
    const int idx = c->cfg.index_kind;
    if (c->cfg.layout == PLB_LAYOUT_SOA) {
        if (idx == PLB_INDEX_TILE8) PLB_READ_TABLE(PLB_LAYOUT_SOA, PLB_INDEX_TILE8);
        else if (idx == PLB_INDEX_BITMAP) PLB_READ_TABLE(PLB_LAYOUT_SOA, PLB_INDEX_BITMAP);
        else PLB_READ_TABLE(PLB_LAYOUT_SOA, PLB_INDEX_I32);
    } else {
        if (idx == PLB_INDEX_TILE8) PLB_READ_TABLE(PLB_LAYOUT_AOSOA, PLB_INDEX_TILE8);
        else if (idx == PLB_INDEX_BITMAP) PLB_READ_TABLE(PLB_LAYOUT_AOSOA, PLB_INDEX_BITMAP);
        else PLB_READ_TABLE(PLB_LAYOUT_AOSOA, PLB_INDEX_I32);
    }
    PLB_HIP(hipGetLastError());
    return PLB_OK;
}

// out2[0] = sum of stored populations (raw for FP64, deviations otherwise)
// over nodes [first, first+n) of buffer `buf`, out2[1] = sum of ux macros.
// Deterministic: fixed grid, fixed tree order.
int launch_reduce_range(plb_ctx* c, int buf, int64_t first, int64_t n, bool with_mass,
                        double* out2)
{
    const int nb = std::max(1, std::min<int>(c->red_blocks, (int)((n + RED_T - 1) / RED_T)));
    PLB_DISPATCH(c->cfg.precision, c->cfg.layout, k_reduce_partial, dim3(nb), dim3(RED_T), 0,
                 c->stream, c->f[buf], c->macros + c->Npad, first, n, c->Npad, c->inv_scale,
                 with_mass ? 1 : 0, c->red);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(RED_T), 0, c->stream, c->red, nb);
    PLB_HIP(hipGetLastError());
    PLB_HIP(hipMemcpyAsync(out2, c->red + 2 * nb, 2 * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
    PLB_HIP(hipStreamSynchronize(c->stream));
    return PLB_OK;
}

int launch_reduce(plb_ctx* c, double* out2)
{
    return launch_reduce_range(c, c->cur, 0, c->N, true, out2);
}

int launch_sample_sums(plb_ctx* c, const int32_t* d_active, int n_active, const int32_t* chunk_sample,
                       const int64_t* sample_end, const int64_t* d_off, const int64_t* d_n, int nsamples,
                       double* d_partial, double* d_out)
{
    if (n_active > 0)
        hipLaunchKernelGGL(k_chunk_partials, dim3(n_active), dim3(c->cfg.block_size), 0, c->stream,
                           c->macros + c->Npad, d_active, chunk_sample, sample_end, d_partial);
    hipLaunchKernelGGL(k_sample_sums, dim3(nsamples), dim3(64), 0, c->stream, d_partial, d_off, d_n,
                       c->cfg.block_size, d_out);
    PLB_HIP(hipGetLastError());
    return PLB_OK;
}

}  // namespace plb
