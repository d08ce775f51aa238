// twopass.hpp -- the two-pass Potts kernels (Xi step, then V step, per
// iteration) with their initialisation and labelling kernels, generic over
// where the lines of the work layout come from:
//   VolumeLines  one volume (or one n-slab of it): every line has the same
//                length and the work offset of a line is line * wstride;
//   BatchLines   many independent solves of one K side by side: per-chunk
//                and per-line tables give each line's offsets, length,
//                neighbours and solve (whose labels and lambda the V step uses).
// A chunk is 64 lanes x VPL consecutive voxels of one line.
#pragma once

#include "sct_internal.hpp"
#include "potts_math.hpp"

namespace sct {
namespace tp {

constexpr uint32_t kNone = 0xffffffffu;

// Everything a wave needs to know about its chunk (uniform over the wave).
struct Chunk {
    uint32_t n0, len;          // first n of the chunk (line coordinates), length of the line
    uint32_t w;                // work offset of (line, n0)
    uint32_t nr, nc, pr, pc;   // work offsets of the +row, +col, -row, -col neighbours at n0, or kNone
    uint32_t u;                // input offset of (line, n0)
    uint32_t r;                // result offset of (line, n0)
    uint32_t na0;              // n0 in the input's coordinates (slab start added)
    uint32_t solve;
};

// Division by a run-time constant d (1 <= d < 2^31) as a multiply-high and
// shifts (Granlund and Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994, figure 4.1), exact for every 32-bit n.
struct FastDiv {
    uint32_t d = 1, m = 1, s1 = 0, s2 = 0;
    FastDiv() = default;
    explicit FastDiv(uint32_t dv) : d(dv)
    {
        uint32_t l = 0;
        while ((uint64_t(1) << l) < dv) ++l;   // l = ceil(log2 d)
        m = uint32_t((uint64_t(1) << 32) * ((uint64_t(1) << l) - dv) / dv + 1);
        s1 = l < 1 ? l : 1;
        s2 = l > 1 ? l - 1 : 0;
    }
    __host__ __device__ __forceinline__ uint32_t div(uint32_t n) const
    {
        const uint32_t t = uint32_t((uint64_t(m) * n) >> 32);
        return (t + ((n - t) >> s1)) >> s2;
    }
};

struct VolumeLines {
    const LineNbr *nbr;
    uint32_t lines, len, vchunks, wstride, in_stride, s0;
    FastDiv vdiv;   // division by vchunks

    static VolumeLines make(const LineNbr *nbr, uint32_t lines, uint32_t len, uint32_t vchunks, uint32_t wstride,
                            uint32_t in_stride, uint32_t s0)
    {
        VolumeLines L{nbr, lines, len, vchunks, wstride, in_stride, s0, FastDiv(vchunks ? vchunks : 1)};
        return L;
    }
    __host__ __device__ __forceinline__ uint64_t count() const { return uint64_t(lines) * vchunks; }
    template <int VPL>
    __device__ __forceinline__ Chunk chunk(uint64_t gw64) const
    {
        // chunk counts stay below 2^32 (lines < 2^32 / vchunks is checked on the host)
        const uint32_t gw = uint32_t(gw64);
        const uint32_t line = vdiv.div(gw);
        const uint32_t n0 = (gw - line * vchunks) * (kWave * VPL);
        const LineNbr nb = nbr[line];
        const auto at = [&](int32_t l) { return l >= 0 ? uint32_t(l) * wstride + n0 : kNone; };
        const uint32_t u = line * in_stride + s0 + n0;
        return Chunk{n0, len, line * wstride + n0, at(nb.nrow), at(nb.ncol), at(nb.prow), at(nb.pcol), u, u, s0 + n0, 0};
    }
};

// One line of a batch: offsets of its n = 0 element and its neighbours'.
struct BatchLine {
    uint32_t w, u, r, len;
    uint32_t nr, nc, pr, pc;   // work offsets of the neighbour lines, or kNone
    uint32_t solve, pad[3];
};

struct BatchLines {
    const uint2 *chunks;       // (line, n0) of every chunk
    const BatchLine *lines;
    uint64_t n_chunks;

    __host__ __device__ __forceinline__ uint64_t count() const { return n_chunks; }
    template <int VPL>
    __device__ __forceinline__ Chunk chunk(uint64_t gw) const
    {
        const uint2 t = chunks[gw];
        const BatchLine L = lines[t.x];
        const uint32_t n0 = t.y;
        const auto at = [&](uint32_t o) { return o == kNone ? kNone : o + n0; };
        return Chunk{n0, L.len, L.w + n0, at(L.nr), at(L.nc), at(L.pr), at(L.pc), L.u + n0, L.r + n0, n0, L.solve};
    }
};

// Labels and lambda of a solve: one set for a volume, a device table for a batch.
template <int K>
struct SolveParams {
    float c[K];
    float lambda;
};

template <int K>
struct OneSolve {
    SolveParams<K> p;
    __device__ __forceinline__ SolveParams<K> get(uint32_t) const { return p; }
};

template <int K>
struct SolveTable {
    const SolveParams<K> *p;
    __device__ __forceinline__ SolveParams<K> get(uint32_t s) const { return p[s]; }
};

// Every lane owns VPL consecutive voxels so that each lane moves 8 bytes per
// array access (FP32 VPL = 2, BF16 / FP16 VPL = 4): 16-bit loads and stores of
// one element per lane run the memory system at half the rate of 32-bit ones.
// Lanes past the end of a line keep their vector inside the line's padding
// (work strides are multiples of 32 elements >= the line length rounded up to
// VPL), which no valid voxel reads.
template <typename St>
constexpr int vpl_of() { return sizeof(typename St::T) == 4 ? 2 : 4; }

// VPL consecutive stored elements, one load or store.
template <typename St, int N>
struct Packed {
    using T = typename St::T;
    typedef T V __attribute__((ext_vector_type(N)));
    static __device__ __forceinline__ void ld(const T *p, float (&o)[N])
    {
        const V v = *reinterpret_cast<const V *>(p);
#pragma unroll
        for (int j = 0; j < N; ++j) o[j] = St::ld(v[j]);
    }
    static __device__ __forceinline__ void st(T *p, const float (&x)[N])
    {
        V v;
#pragma unroll
        for (int j = 0; j < N; ++j) v[j] = St::st(x[j]);
        *reinterpret_cast<V *>(p) = v;
    }
    // Non-temporal: for state that is next read a whole sweep later.
    static __device__ __forceinline__ void st_nt(T *p, const float (&x)[N])
    {
        V v;
#pragma unroll
        for (int j = 0; j < N; ++j) v[j] = St::st(x[j]);
        __builtin_nontemporal_store(v, reinterpret_cast<V *>(p));
    }
};

// Per-component base pointers; every component is addressed as
// base[comp] + 32-bit element offset.  The Xi step reads xi and vbar and
// writes xi_out; the V step reads xi, V and u and writes V_out and vbar_out.
// The solvers iterate in place (outputs = inputs, in_place()); the adjoint's
// recompute steps from one stored state into the next.
template <int K, typename St>
struct TwoPassArgs {
    typename St::T *xi[3 * (K - 1)];
    typename St::T *vbar[K - 1];
    float *V[K - 1];
    const float *u;
    typename St::T *xi_out[3 * (K - 1)];
    typename St::T *vbar_out[K - 1];
    float *V_out[K - 1];

    void in_place()
    {
        for (int j = 0; j < 3 * (K - 1); ++j) xi_out[j] = xi[j];
        for (int i = 0; i < K - 1; ++i) {
            vbar_out[i] = vbar[i];
            V_out[i] = V[i];
        }
    }
};

// threadIdx.x; with OPAQUE, hidden from the compiler.  A loop over tiles (the
// persistent batch kernel) reads the lane through the opaque form, so the
// values a tile derives from it are computed per tile rather than hoisted out
// of the loop and held in registers across both half-steps.
template <bool OPAQUE>
__device__ __forceinline__ uint32_t thread_id()
{
    uint32_t t = threadIdx.x;
    if constexpr (OPAQUE) asm volatile("" : "+v"(t));
    return t;
}

constexpr int kXiVoxelsPerLane = 4;  // per lane and workgroup: kXiVoxelsPerLane / VPL chunks per wave
// Chunks per wave of the Xi kernel.
template <typename St>
constexpr int xi_cpw() { return kXiVoxelsPerLane / vpl_of<St>(); }
// Waves per SIMD the Xi kernel is compiled for (launch-bounds hint): with
// 16-bit storage at K <= 3 the step is bound by how many workgroups overlap
// (7 per CU instead of 6: 7.8 -> 7.3 ms on the 1426 x 1024^2 volume); FP32
// keeps the compiler's choice (6; 7 would spill and ran 11% slower).
template <int K, typename St>
constexpr int xi_waves_per_simd() { return K <= 3 && sizeof(typename St::T) == 2 ? 7 : 1; }
constexpr int kXiBlockWaves = 4;     // default workgroup of the Xi kernel (its "tile": 64 BW lanes x 4 voxels)
constexpr int kXiThreads = kXiBlockWaves * 64;
constexpr int kXiWorkWaves = 4;      // at most this many waves drain the work list
// Sweeps an item of the work list gets in the first stage; the few that need
// more (about a quarter of the queued items, up to 20 sweeps) are parked and
// finish in a second stage on the fewest waves, so the long tail of sweep
// counts no longer sets the time of every draining wave.
constexpr int kXiFirstSweeps = 2;

// Per-voxel record of the Xi step's Dykstra, written by the REC variant of
// potts_xi_kernel for the adjoint: sweeps[w] = 0 when the input was feasible
// (the first sweep changes nothing), else the sweeps performed (1..20).
struct SweepRecord {
    uint8_t *sweeps;
};

// Workgroups of the one-chunk-per-wave kernels, in XCD runs of 64 / VPL
// (about 11 lines of 1426 voxels).
template <int VPL>
__device__ __forceinline__ uint64_t chunk_of_wave()
{
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x, 64 / VPL);
    const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    return uint64_t(b) * (blockDim.x / kWave) + w;
}

// init_v_kernel, potts-min-partition.cpp:56-94.
template <int K, typename St, typename Lines, typename Params>
__global__ __launch_bounds__(256) void potts_init_kernel(Lines lines, TwoPassArgs<K, St> a, Params prm)
{
    constexpr int VPL = vpl_of<St>();
    const uint64_t gw = chunk_of_wave<VPL>();
    if (gw >= lines.count()) return;
    const Chunk ch = lines.template chunk<VPL>(gw);
    const SolveParams<K> sp = prm.get(ch.solve);
    const uint32_t l0 = (threadIdx.x % kWave) * VPL;
#pragma unroll
    for (int j = 0; j < VPL; ++j) {
        if (ch.n0 + l0 + j >= ch.len) break;
        const int cl = nearest_label<K>(a.u[ch.u + l0 + j], sp.c);
#pragma unroll
        for (int i = 0; i < K - 1; ++i) {
            const float o = (i == cl) ? 1.0f : 0.0f;
            a.V[i][ch.w + l0 + j] = o;
            a.vbar[i][ch.w + l0 + j] = St::st(o);
        }
    }
}

// Xi = 0 (the reference's zero-initialised buffer, potts-min-partition.cpp:598),
// written by a kernel so that initialisation can be captured into a graph.
template <int K, typename St, typename Lines>
__global__ __launch_bounds__(256) void potts_xi_zero_kernel(Lines lines, TwoPassArgs<K, St> a)
{
    constexpr int VPL = vpl_of<St>();
    const uint64_t gw = chunk_of_wave<VPL>();
    if (gw >= lines.count()) return;
    const Chunk ch = lines.template chunk<VPL>(gw);
    const uint32_t l0 = (threadIdx.x % kWave) * VPL;
    if (ch.n0 + l0 >= ch.len) return;
    const float z[VPL] = {};
#pragma unroll
    for (int c = 0; c < 3 * (K - 1); ++c) Packed<St, VPL>::st(a.xi[c] + ch.w + l0, z);
}

// grad_kernel, potts-min-partition.cpp:96-247: Xi <- Pi_K(Xi + tau2 grad V-bar),
// in place (each voxel reads its own Xi and its neighbours' V-bar only).
//   A  every lane forms the Dykstra input of its voxels; most are feasible and
//      Dykstra stops after one sweep, the rest join a work list in LDS;
//   B  the work list is drained one sweep per loop trip, each lane taking a new
//      item as soon as its item converges, the result going back to LDS;
//   C  every lane stores the Xi+ of its voxels, VPL at a time.
// BW: waves per workgroup (launch 64 BW threads); the workgroup covers 256 BW
// voxels and its work list holds 128 BW of them.  REC: also store the
// SweepRecord of every voxel (the solver runs without).
template <int K, typename St, typename Lines, int BW, bool REC, bool OPAQUE = false>
__device__ __forceinline__ void xi_tile(const Lines &lines, const TwoPassArgs<K, St> &a, float tau2,
                                        DykstraLimits lim, SweepRecord rec, uint32_t blk)
{
    constexpr int KM = K - 1, NP = K * (K - 1) / 2;
    constexpr int VPL = vpl_of<St>(), CPW = xi_cpw<St>();
    // Work list: a region of kWaveItems items per wave, filled by that wave
    // alone (a slot is the region's start + the wave's running count + the
    // lane's rank), so no atomics; a voxel past its region (more than 120 or
    // 128 of a wave's 64 x 4 voxels queued) is solved in place.  120 for
    // 16-bit storage at K <= 3 keeps the LDS within a seventh of the CU's
    // (xi_waves_per_simd).
    constexpr int kWaveItems = xi_waves_per_simd<K, St>() == 7 ? 120 : 128;
    constexpr int kXiItems = kWaveItems * BW;
    using P = Packed<St, VPL>;
    // Parking pays for K <= 3; for K >= 4 the corrections of the parked
    // items (NP = 6..15 pairs) cost LDS, hence workgroups per CU.
    constexpr int kFirst = K <= 3 ? kXiFirstSweeps : 20;
    constexpr int kXiLong = K <= 3 ? 32 * BW : 1;   // items parked after kFirst sweeps
    __shared__ float qb[kXiItems][K][3];   // inputs, then results, of the queued voxels
    __shared__ float rb[kXiLong][NP][3];   // Dykstra corrections of the parked items
    __shared__ int lt[kXiLong];            // parked item: index | sweeps << 16
    __shared__ uint8_t sq[REC ? kXiItems : 1];   // sweeps of the work-list items
    __shared__ int nwi[BW];                // items in each wave's region
    __shared__ int next, nl;
    const uint32_t tid = thread_id<OPAQUE>();
    const int lane = int(tid % kWave), w = __builtin_amdgcn_readfirstlane(int(tid / kWave));
    if (tid == 0) {
        next = 0;
        nl = 0;
    }
    __syncthreads();
    const uint64_t total = lines.count();
    int queued_n = 0;   // items this wave has queued (uniform)
    float res[CPW][VPL][KM][3];
    int pos[CPW][VPL];
    int sfull[CPW][VPL];   // REC: sweeps of voxels not on the work list
    uint32_t wofs[CPW];
    int nv[CPW];   // valid voxels of this lane in chunk k
#pragma unroll
    for (int k = 0; k < CPW; ++k) {
        const uint64_t gw = (uint64_t(blk) * (blockDim.x / kWave) + w) * CPW + k;
        nv[k] = 0;
        wofs[k] = 0;
#pragma unroll
        for (int j = 0; j < VPL; ++j) {
            pos[k][j] = -1;
            sfull[k][j] = 0;
        }
        if (gw >= total) continue;   // uniform over the wave
        const Chunk ch = lines.template chunk<VPL>(gw);
        const uint32_t n0 = ch.n0 + lane * VPL, l0 = lane * VPL;
        nv[k] = n0 < ch.len ? int(min(uint32_t(VPL), ch.len - n0)) : 0;
        const uint32_t w0 = ch.w + l0;
        wofs[k] = nv[k] ? w0 : 0;
        // Every lane loads, from an address inside its line: a lane past the
        // line's end reads the chunk's first vector, a missing neighbour line
        // the voxel's own; neither value is used (the lane's voxels are not
        // valid, and xi_input skips a missing neighbour).
        const uint32_t wl = nv[k] ? w0 : ch.w;
        const uint32_t wr = nv[k] && ch.nr != kNone ? ch.nr + l0 : wl;
        const uint32_t wc = nv[k] && ch.nc != kNone ? ch.nc + l0 : wl;
        float vb[KM][VPL], vr[KM][VPL], vc[KM][VPL], x[KM][3][VPL], vn[KM][VPL];
#pragma unroll
        for (int i = 0; i < KM; ++i) {
            P::ld(a.vbar[i] + wl, vb[i]);
            P::ld(a.vbar[i] + wr, vr[i]);
            P::ld(a.vbar[i] + wc, vc[i]);
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) P::ld(a.xi[3 * i + ax] + wl, x[i][ax]);
            // V-bar at n + 1: the next element of the lane, the next lane's first,
            // or (last lane) the first of the next chunk
#pragma unroll
            for (int j = 0; j + 1 < VPL; ++j) vn[i][j] = vb[i][j + 1];
            vn[i][VPL - 1] = __shfl_down(vb[i][0], 1, kWave);
            if (lane == kWave - 1 && nv[k] == VPL && n0 + VPL < ch.len) vn[i][VPL - 1] = St::ld(a.vbar[i][w0 + VPL]);
        }
#pragma unroll
        for (int j = 0; j < VPL; ++j) {
            const bool valid = j < nv[k];
            const bool nx[3] = {n0 + j + 1 < ch.len, ch.nr != kNone, ch.nc != kNone};
            float vbj[KM], vbn[3][KM], xj[KM][3];
#pragma unroll
            for (int i = 0; i < KM; ++i) {
                vbj[i] = vb[i][j];
                vbn[0][i] = vn[i][j];
                vbn[1][i] = vr[i][j];
                vbn[2][i] = vc[i][j];
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) xj[i][ax] = x[i][ax][j];
            }
            float q[K][3];
            xi_input<K>(vbj, vbn, nx, xj, tau2, q);
            const bool queued = valid && !dykstra_first_sweep_done<K>(q, lim);
            const uint64_t mask = __ballot(queued);
            const int rank = queued_n + __popcll(mask & ((1ull << lane) - 1ull));
            const int p = w * kWaveItems + rank;
            queued_n += __popcll(mask);
            if (queued && rank < kWaveItems) {
#pragma unroll
                for (int i = 0; i < K; ++i)
#pragma unroll
                    for (int ax = 0; ax < 3; ++ax) qb[p][i][ax] = q[i][ax];
                pos[k][j] = p;
            } else if (queued) {
                sfull[k][j] = dykstra<K>(q, lim);   // work list full: solve in place
            }
#pragma unroll
            for (int i = 0; i < KM; ++i)
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) res[k][j][i][ax] = valid ? q[i][ax] : 0.0f;
        }
    }
    if (lane == 0) nwi[w] = min(queued_n, kWaveItems);
    __syncthreads();   // work list complete
    // item t of the drain: the regions' items one after the other
    int first[BW + 1];
    first[0] = 0;
#pragma unroll
    for (int i = 0; i < BW; ++i) first[i + 1] = first[i] + nwi[i];
    const int nwl = first[BW];
    auto slot_of = [&](int t) {
        int r = 0;
#pragma unroll
        for (int i = 1; i < BW; ++i) r += t >= first[i];
        int s = 0;
#pragma unroll
        for (int i = 0; i < BW; ++i) s = i == r ? r * kWaveItems + t - first[i] : s;
        return s;
    };
    if (w < kXiWorkWaves && w * kWave < nwl) {
        float q[K][3], r[NP][3];
        int sw = 0, t = 0;
        bool have = false, drained = false;
        while (true) {
            const bool need = !have && !drained;
            const uint64_t mask = __ballot(need);
            if (mask) {
                const int leader = __ffsll((unsigned long long)mask) - 1;
                int base = 0;
                if (lane == leader) base = atomicAdd(&next, __popcll(mask));
                base = __shfl(base, leader, kWave);
                if (need) {
                    const int g = base + __popcll(mask & ((1ull << lane) - 1ull));
                    if (g < nwl) {
                        t = slot_of(g);
#pragma unroll
                        for (int i = 0; i < K; ++i)
#pragma unroll
                            for (int ax = 0; ax < 3; ++ax) q[i][ax] = qb[t][i][ax];
#pragma unroll
                        for (int pp = 0; pp < NP; ++pp) r[pp][0] = r[pp][1] = r[pp][2] = 0.0f;
                        sw = 0;
                        have = true;
                    } else {
                        drained = true;
                    }
                }
            }
            if (!__any(have)) break;
            if (have) {
                const float m2 = dykstra_sweep<K>(q, r, lim);
                ++sw;
                if (m2 < lim.stol || sw == 20) {
#pragma unroll
                    for (int i = 0; i < KM; ++i)
#pragma unroll
                        for (int ax = 0; ax < 3; ++ax) qb[t][i][ax] = q[i][ax];
                    if constexpr (REC) sq[t] = uint8_t(sw);
                    have = false;
                } else if (kFirst < 20 && sw == kFirst) {
                    // park the item (q, r, sweep count) for the second stage
                    // if there is room; otherwise keep sweeping it here
                    const int slot = atomicAdd(&nl, 1);
                    if (slot < kXiLong) {
#pragma unroll
                        for (int i = 0; i < K; ++i)
#pragma unroll
                            for (int ax = 0; ax < 3; ++ax) qb[t][i][ax] = q[i][ax];
#pragma unroll
                        for (int pp = 0; pp < NP; ++pp)
#pragma unroll
                            for (int ax = 0; ax < 3; ++ax) rb[slot][pp][ax] = r[pp][ax];
                        lt[slot] = t | (sw << 16);
                        have = false;
                    }
                }
            }
        }
    }
    if constexpr (kFirst < 20) {
    __syncthreads();   // first stage done: the parked items continue on the fewest waves
    const int nlong = min(nl, kXiLong);
    for (int i = w * kWave + lane; w * kWave < nlong && i < nlong; i += int(blockDim.x)) {
        const int t = lt[i] & 0xffff;
        int sw = lt[i] >> 16;
        float q[K][3], r[NP][3];
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) q[k][ax] = qb[t][k][ax];
#pragma unroll
        for (int pp = 0; pp < NP; ++pp)
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) r[pp][ax] = rb[i][pp][ax];
        while (true) {
            const float m2 = dykstra_sweep<K>(q, r, lim);
            ++sw;
            if (m2 < lim.stol || sw == 20) break;
        }
#pragma unroll
        for (int k = 0; k < KM; ++k)
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) qb[t][k][ax] = q[k][ax];
        if constexpr (REC) sq[t] = uint8_t(sw);
    }
    }
    __syncthreads();   // results of the work list in LDS
#pragma unroll
    for (int k = 0; k < CPW; ++k) {
        if (!nv[k]) continue;
#pragma unroll
        for (int i = 0; i < KM; ++i)
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                float o[VPL];
#pragma unroll
                for (int j = 0; j < VPL; ++j) o[j] = pos[k][j] >= 0 ? qb[pos[k][j]][i][ax] : res[k][j][i][ax];
                P::st(a.xi_out[3 * i + ax] + wofs[k], o);
            }
        if constexpr (REC) {
#pragma unroll
            for (int j = 0; j < VPL; ++j)
                if (j < nv[k]) rec.sweeps[wofs[k] + j] = uint8_t(pos[k][j] >= 0 ? sq[pos[k][j]] : sfull[k][j]);
        }
    }
}

template <int K, typename St, typename Lines, int BW = kXiBlockWaves, bool REC = false>
__global__ __launch_bounds__(256, (xi_waves_per_simd<K, St>())) void potts_xi_kernel(Lines lines,
                                                                                      TwoPassArgs<K, St> a,
                                                                                      float tau2, DykstraLimits lim,
                                                                                      SweepRecord rec = {})
{
    xi_tile<K, St, Lines, BW, REC>(lines, a, tau2, lim, rec, xcd_block(blockIdx.x, gridDim.x, 64 / BW));
}

// div_kernel, potts-min-partition.cpp:349-488: V <- Pi_S(V + tau1 (div Xi - G)),
// V-bar <- 2 V+ - V.  With TRACE, the block maximum of sum_k dV_k^2 goes to *dq.
// One chunk per wave.  V and V-bar are stored non-temporally: the next reads
// are a whole sweep away, and keeping them out of the L2 leaves it to the
// neighbour lines (5% faster on 1426 x 1024^2; the Xi step's stores gain
// nothing from it).
// The V step of chunk gw (one wave); returns the largest sum_k dV_k^2 of its voxels.
template <int K, typename St, typename Lines, typename Params, bool OPAQUE = false>
__device__ __forceinline__ float v_chunk(const Lines &lines, const TwoPassArgs<K, St> &a, const Params &prm,
                                         float tau1, uint64_t gw)
{
    constexpr int KM = K - 1, VPL = vpl_of<St>();
    using P = Packed<St, VPL>;
    using PF = Packed<StorageF32, VPL>;
    const int lane = int(thread_id<OPAQUE>() % kWave);
    float d = 0.0f;
    if (gw < lines.count()) {   // uniform over the wave
        const Chunk ch = lines.template chunk<VPL>(gw);
        const SolveParams<K> sp = prm.get(ch.solve);
        const uint32_t n0 = ch.n0 + lane * VPL, l0 = lane * VPL;
        const int nv = n0 < ch.len ? int(min(uint32_t(VPL), ch.len - n0)) : 0;
        const uint32_t w0 = ch.w + l0;
        float xo[KM][3][VPL], xr[KM][VPL], xc[KM][VPL], xn[KM][VPL], v[KM][VPL], u[VPL];
#pragma unroll
        for (int i = 0; i < KM; ++i) {
#pragma unroll
            for (int j = 0; j < VPL; ++j) {
                xr[i][j] = xc[i][j] = v[i][j] = 0.0f;
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) xo[i][ax][j] = 0.0f;
            }
            if (nv) {
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) P::ld(a.xi[3 * i + ax] + w0, xo[i][ax]);
                if (ch.pr != kNone) P::ld(a.xi[3 * i + 1] + ch.pr + l0, xr[i]);
                if (ch.pc != kNone) P::ld(a.xi[3 * i + 2] + ch.pc + l0, xc[i]);
                PF::ld(a.V[i] + w0, v[i]);
            }
            // Xi_n at n - 1: the previous element of the lane, the previous
            // lane's last, or (first lane) the last of the previous chunk
#pragma unroll
            for (int j = 1; j < VPL; ++j) xn[i][j] = xo[i][0][j - 1];
            xn[i][0] = __shfl_up(xo[i][0][VPL - 1], 1, kWave);
            if (lane == 0 && nv && n0 > 0) xn[i][0] = St::ld(a.xi[3 * i][w0 - 1]);
        }
#pragma unroll
        for (int j = 0; j < VPL; ++j) u[j] = j < nv ? a.u[ch.u + l0 + j] : 0.0f;
        float vb[KM][VPL];
#pragma unroll
        for (int j = 0; j < VPL; ++j) {
            const bool px[3] = {n0 + j > 0, ch.pr != kNone, ch.pc != kNone};
            const bool nx[3] = {n0 + j + 1 < ch.len, ch.nr != kNone, ch.nc != kNone};
            float xoj[KM][3], xpj[3][KM], vj[KM], vbj[KM];
#pragma unroll
            for (int i = 0; i < KM; ++i) {
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) xoj[i][ax] = xo[i][ax][j];
                xpj[0][i] = xn[i][j];
                xpj[1][i] = xr[i][j];
                xpj[2][i] = xc[i][j];
                vj[i] = v[i][j];
            }
            const float dj = v_update<K>(xoj, xpj, px, nx, vj, vbj, u[j], sp.c, sp.lambda, tau1);
            if (j < nv) d = fmaxf(d, dj);
#pragma unroll
            for (int i = 0; i < KM; ++i) {
                v[i][j] = vj[i];
                vb[i][j] = vbj[i];
            }
        }
        if (nv) {
#pragma unroll
            for (int i = 0; i < KM; ++i) {
                PF::st_nt(a.V_out[i] + w0, v[i]);
                P::st_nt(a.vbar_out[i] + w0, vb[i]);
            }
        }
    }
    return d;
}

template <int K, typename St, bool TRACE, typename Lines, typename Params>
__global__ __launch_bounds__(256) void potts_v_kernel(Lines lines, TwoPassArgs<K, St> a, Params prm, float tau1,
                                                      float *dq)
{
    float d = v_chunk<K, St, Lines, Params>(lines, a, prm, tau1, chunk_of_wave<vpl_of<St>()>());
    if constexpr (TRACE) {
        __shared__ float red[4];
        d = wave_max(d);
        if ((threadIdx.x % kWave) == 0) red[threadIdx.x / kWave] = d;
        __syncthreads();
        if (threadIdx.x == 0) {
            float m = red[0];
            for (int i = 1; i < int(blockDim.x / kWave); ++i) m = fmaxf(m, red[i]);
            atomic_max_nonneg(dq, m);
        }
    }
}

// update_kernel, potts-min-partition.cpp:491-538: res = c[argmax V] for the
// voxels whose input coordinate lies in [r0, r1).
template <int K, typename St, typename Lines, typename Params>
__global__ __launch_bounds__(256) void potts_labels_kernel(Lines lines, TwoPassArgs<K, St> a, Params prm,
                                                           float *__restrict__ res, uint32_t r0, uint32_t r1)
{
    constexpr int VPL = vpl_of<St>();
    const uint64_t gw = chunk_of_wave<VPL>();
    if (gw >= lines.count()) return;
    const Chunk ch = lines.template chunk<VPL>(gw);
    const SolveParams<K> sp = prm.get(ch.solve);
    const uint32_t l0 = (threadIdx.x % kWave) * VPL;
#pragma unroll
    for (int j = 0; j < VPL; ++j) {
        const uint32_t na = ch.na0 + l0 + j;
        if (ch.n0 + l0 + j >= ch.len) break;
        if (na < r0 || na >= r1) continue;
        float v[K - 1];
#pragma unroll
        for (int i = 0; i < K - 1; ++i) v[i] = a.V[i][ch.w + l0 + j];
        res[ch.r + l0 + j] = sp.c[argmax_label<K>(v)];
    }
}

// Workgroups (256 threads) of the Xi kernel and of the one-chunk-per-wave kernels.
template <typename St, int BW = kXiBlockWaves>
inline uint32_t xi_grid(uint64_t chunks)
{
    constexpr uint64_t cpw = xi_cpw<St>();
    return uint32_t(((chunks + cpw - 1) / cpw + BW - 1) / BW);
}
inline uint32_t wave_grid(uint64_t chunks) { return uint32_t((chunks + 3) / 4); }

}  // namespace tp
}  // namespace sct
