// mpm_kernels.h -- the explicit USL step as HIP kernels for gfx950, for both particle-state
// forms (mpm_prec.h).
//
// One step (explicit_usl.cpp:48-87) is two or three launches over resident HBM state:
//   k_grid       per touched node: body force m g, ContactHuang (contact_huang.cpp:104-197),
//                CartesianBox / SmoothBox / FrictionalBox (cartesian_*box.cpp), moveGrid (explicit_usl.cpp:233-264);
//                one wavefront per 64-node block; consumes and re-zeroes the accumulators.
//   k_particles  one workgroup per particle block (persistent over the block list):
//                G2P from an LDS copy of the block's node patch -- movePoints,
//                calculateStrainRate, updateDensity, gravity, updateStress
//                (explicit_usl.cpp:266-355, isolin.cpp:50-78) -- then, for the next step,
//                the map (default_points.cpp:196-330) and P2G (explicit_usl.cpp:112-190)
//                with the contact-normal scatter gradS*m (contact_huang.cpp:80-99) into an
//                LDS accumulator patch (FP64 LDS atomics), flushed to the grid with FP64
//                global atomics. The G2P and P2G halves run as separate passes (default)
//                or fused into one; the first step of a call runs the P2G half alone, the
//                last the G2P half alone.
//
// The nodal body force: every point carries b = g (default_driver.cpp:109-119), so
// f_i = sum_p S_ip m_p g = m_i g; k_grid adds m_i g once per node instead of scattering
// m_p g from every point.
#pragma once

#include "mpm_common.h"
#include "mpm_prec.h"

namespace mpm {

// A node block's flag holds one bit per body of its simulation that scattered into it this
// step (bit b: the sim's body b, b < 8) and, above them, the simulation's index (the same from
// every marker); the first mark appends the block to the list of parity par.
__device__ __forceinline__ void mark_block(const KArgs &A, long long gblk, unsigned bits) {
    if (atomicOr(&A.nflag[gblk], bits) == 0u) {
        const int pos = atomicAdd(&A.ncount[A.par], 1);
        A.nlist[(long long)A.par * A.nblk_total + pos] = (int)gblk;
    }
}

// mark the node blocks covering nodes [lo, hi] (inclusive) of sim `s` for its body B;
// called by one thread
template <int DIM>
__device__ void mark_range(const KArgs &A, const SimDev &s, const BodyDev &B, const int *lo, const int *hi) {
    const unsigned bits = (1u << B.local) | ((unsigned)B.sim << 8);
    constexpr int L = Cfg<DIM>::LNB;
    int b0[3] = {0, 0, 0}, b1[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        b0[d] = (lo[d] + NSHIFT) >> L;  // node storage is shifted (node_addr)
        b1[d] = (hi[d] + NSHIFT) >> L;
    }
    for (int k = b0[2]; k <= b1[2]; k++)
        for (int j = b0[1]; j <= b1[1]; j++)
            for (int i = b0[0]; i <= b1[0]; i++) {
                long long lb = i + (long long)s.NBn[0] * (j + (long long)s.NBn[1] * k);
                mark_block(A, s.nblk_off + lb, bits);
            }
}

// particle-block key -> patch origin (node index of LDS patch element 0)
template <int DIM>
__device__ __forceinline__ void patch_origin(const SimDev &s, long long lkey, int *org) {
    constexpr int PB = Cfg<DIM>::PB;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const int pb = (int)(lkey % s.PNB[d]);
        lkey /= s.PNB[d];
        org[d] = pb * PB - 2;
    }
}

template <int DIM>
__host__ __device__ __forceinline__ int patch_index(const int *nijk, const int *org) {
    constexpr int P = Cfg<DIM>::P;
    int li = nijk[0] - org[0] + P * (nijk[1] - org[1]);
    if (DIM == 3) li += Cfg<DIM>::ZS * (nijk[DIM - 1] - org[DIM - 1]);
    return li;
}

// the per-point P2G payload: m, m v and V T (explicit_usl.cpp:132-161)
template <int DIM, class R> struct P2GVals {
    R m;
    R mv[DIM];
    R VT[DIM][DIM];
};

template <int DIM, class R, class St>
__device__ __forceinline__ void p2g_payload(const St &q, P2GVals<DIM, R> &pv) {
    pv.m = q.m;
#pragma unroll
    for (int d = 0; d < DIM; d++) pv.mv[d] = q.v[d] * q.m;
    const R Tf[3][3] = {{q.T6[0], q.T6[3], q.T6[4]}, {q.T6[3], q.T6[1], q.T6[5]}, {q.T6[4], q.T6[5], q.T6[2]}};
#pragma unroll
    for (int r = 0; r < DIM; r++)
#pragma unroll
        for (int c = 0; c < DIM; c++) pv.VT[r][c] = Tf[r][c] * q.V;
}

// node values of one P2G tuple: m, p[DIM], f[DIM] (stress divergence only), n[DIM]
template <int DIM, class R>
__host__ __device__ __forceinline__ void p2g_node_values(const P2GVals<DIM, R> &q, R S, const R *G, R *out, bool b1) {
    out[0] = S * q.m;
#pragma unroll
    for (int d = 0; d < DIM; d++) out[1 + d] = S * q.mv[d];
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        R f = 0;
#pragma unroll
        for (int c = 0; c < DIM; c++) f = fma(-q.VT[d][c], G[c], f);
        out[1 + DIM + d] = f;
    }
#pragma unroll
    for (int d = 0; d < DIM; d++) out[1 + 2 * DIM + d] = b1 ? G[d] * q.m : (R)0;
}

// Visit the window's slots the reference's tuples cover, one by one with the loop kept
// rolled (weights picked by selects); used on the rare global-memory path so its loads do
// not set the kernel's register budget. fn(off, S, G) with off the slot offsets from w.base.
template <class R>
__host__ __device__ __forceinline__ R pick3(const R *a, int i) { return i == 0 ? a[0] : (i == 1 ? a[1] : a[2]); }

template <int DIM, class R, bool REACH = true, class Fn>
__host__ __device__ __forceinline__ void for_each_sep(const Win<DIM, R> &w, R ext, Fn &&fn) {
    const R gscale = (R)1 / ((R)(1 << DIM) * (R)0.5 * ext);
    const R sscale = (R)1 / (R)(1 << DIM);
    constexpr int NS = DIM == 3 ? 27 : 9;
#pragma unroll 1
    for (int k = 0; k < NS; k++) {
        const int c0 = k % 3, c1 = (k / 3) % 3, c2 = k / 9;
        const int cs[3] = {c0, c1, c2};
        if (REACH && !slot_reached<DIM, R>(w, cs)) continue;
        const R sx = pick3(w.sw[0], c0), dx = pick3(w.dw[0], c0);
        const R sy = pick3(w.sw[1], c1), dy = pick3(w.dw[1], c1);
        const R sz = DIM == 3 ? pick3(w.sw[DIM - 1], c2) : (R)1;
        const R dz = DIM == 3 ? pick3(w.dw[DIM - 1], c2) : (R)0;
        const R S = sx * (sy * sz) * sscale;
        R G[DIM];
        G[0] = dx * (sy * sz) * gscale;
        G[1] = sx * (dy * sz) * gscale;
        if (DIM == 3) G[DIM - 1] = sx * (sy * dz) * gscale;
        int off[DIM];
        off[0] = c0;
        off[1] = c1;
        if (DIM == 3) off[DIM - 1] = c2;
        fn(off, S, G);
    }
}

// 1 / (2^DIM ext / 2), the gradient scale of the separable window: correctly rounded in
// FP64; in FP32 the hardware reciprocal (1 ulp), one instruction where the rounded division
// takes ten
template <int DIM, class R>
__device__ __forceinline__ R grad_scale(R ext) {
    const R den = (R)(1 << DIM) * (R)0.5 * ext;
    if constexpr (sizeof(R) == 4) return __builtin_amdgcn_rcpf(den);
    else return (R)1 / den;
}

// Row coefficients of the separable window: along a row (fixed c1, c2) the x-slot a has
//   S = sw0[a] * Su,  G0 = dw0[a] * gx,  G1 = sw0[a] * gy,  G2 = sw0[a] * gz
template <class R> struct RowW {
    R Su, gx, gy, gz;
};

// (D: the arithmetic type of the result; the window's weights are converted to it)
template <int DIM, class R, class D = R>
__device__ __forceinline__ RowW<D> row_weights(const Win<DIM, R> &w, int c1, int c2, D sscale, D gscale) {
    const D sz = DIM == 3 ? (D)w.sw[DIM - 1][c2] : (D)1;
    const D dz = DIM == 3 ? (D)w.dw[DIM - 1][c2] : (D)0;
    const D sy = (D)w.sw[1][c1], dy = (D)w.dw[1][c1];
    const D syz = sy * sz;
    RowW<D> r;
    r.Su = syz * sscale;
    r.gx = syz * gscale;
    r.gy = (dy * sz) * gscale;
    r.gz = (sy * dz) * gscale;
    return r;
}

template <int DIM, class R>
__device__ __forceinline__ bool window_in_patch(const Win<DIM, R> &w, const int *org) {
    constexpr int P = Cfg<DIM>::P;
    bool fits = true;
#pragma unroll
    for (int d = 0; d < DIM; d++) fits = fits && (w.base[d] >= org[d]) && (w.base[d] + 3 <= org[d] + P);
    return fits;
}

template <int DIM>
__device__ __forceinline__ bool node_in_patch(const int *nijk, const int *org) {
    constexpr int P = Cfg<DIM>::P;
    bool in = true;
#pragma unroll
    for (int d = 0; d < DIM; d++) in = in && nijk[d] >= org[d] && nijk[d] < org[d] + P;
    return in;
}

// G2P accumulation of one node: u += S v dt, dv += S dv, L += v (x) G
template <int DIM, class R>
__device__ __forceinline__ void g2p_acc(const R *vn, const R *dvn, R S, const R *G, R dt, R *up, R *dvp,
                                        R (*Lp)[DIM]) {
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        up[d] = fma(S, vn[d] * dt, up[d]);
        dvp[d] = fma(S, dvn[d], dvp[d]);
    }
#pragma unroll
    for (int r = 0; r < DIM; r++)
#pragma unroll
        for (int c = 0; c < DIM; c++) Lp[r][c] = fma(vn[r], G[c], Lp[r][c]);
}

// node values (v, dv) from the global grid output
template <int DIM, class PT>
__device__ __forceinline__ void fetch_node(const KArgs &A, const SimDev &s, long long gbase, const int *nijk,
                                           typename PT::R *vn, typename PT::R *dvn) {
    using G = GF<DIM>;
    const long long gi = gbase + node_addr<DIM>(s, nijk);
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        vn[d] = PT::fetch(A, (G::OV + d) * A.nt + gi);
        dvn[d] = PT::fetch(A, (G::ODV + d) * A.nt + gi);
    }
}

// Whether a split pass leaves the points outside its LDS path (window not inside the
// block's patch, or the per-tuple form) to its side kernel (k_p2g_side / k_g2p_side). That
// keeps those paths' registers out of the pass: measured on the 16M column, P2G FP64 gets
// 4 waves per SIMD instead of 3 (1.50 -> 1.39 ms + 0.06 ms side kernel) and G2P FP32 sheds
// its spills (0.90 -> 0.83 ms + 0.02 ms). Elsewhere it only costs the side kernel: the FP32
// P2G is at 4 waves per SIMD either way, and the FP64 G2P at 2 either way.
template <bool DO_G2P, bool DO_P2G, bool FP32>
constexpr bool side_pass() {
    return DO_G2P != DO_P2G && (DO_G2P ? FP32 : !FP32);
}

// Appends slot p to a side list with one atomic per wavefront (called by the lanes that
// append, inside their branch). Each slot is appended at most once per pass, so a list of
// np entries cannot overflow.
__device__ __forceinline__ void side_push(unsigned *list, unsigned *count, long long p) {
    const unsigned long long m = __ballot(1);
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)m) - 1;
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(count, (unsigned)__popcll(m));
    base = __shfl(base, leader);
    list[base + (unsigned)__popcll(m & ((1ull << lane) - 1ull))] = (unsigned)p;
}

// End of a side kernel: the last workgroup to finish clears count k for the next pass.
// Every workgroup read the count before it signs off here, so none still needs it, and
// the clearing is a kernel store (graph replays stay correct).
__device__ __forceinline__ void side_done(unsigned *rare_n, int k) {
    __shared__ int last;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        last = atomicAdd(&rare_n[2 + k], 1u) == gridDim.x - 1 ? 1 : 0;
    }
    __syncthreads();
    if (last && threadIdx.x == 0) {
        rare_n[k] = 0;
        rare_n[2 + k] = 0;
    }
}

// Occupancy hints (minimum waves per SIMD the register allocation must allow) and
// software prefetch of the next point's state, per pass and form. Compile-time tuning
// knobs; the defaults are the measured best.
#ifndef MPM_WPE_G2P64
#define MPM_WPE_G2P64 2
#endif
#ifndef MPM_WPE_P2G64
#define MPM_WPE_P2G64 4
#endif
#ifndef MPM_WPE_G2P32
#define MPM_WPE_G2P32 4
#endif
#ifndef MPM_WPE_P2G32
#define MPM_WPE_P2G32 4
#endif
#ifndef MPM_PF_G2P
#define MPM_PF_G2P 1
#endif
#ifndef MPM_PF_P2G32
#define MPM_PF_P2G32 1
#endif
#ifndef MPM_PF_P2G64
#define MPM_PF_P2G64 0
#endif
// Lane order of the P2G pass inside a wavefront, per form: 0 slot order; 1 and 2 place the
// point of in-block dual cell (x, y, z) on lane x + 4y + 16z (1) or x + 4z + 16y (2) (2D:
// x + 8y, y + 8x): one point of each dual cell there, any further ones on the lanes left
// free, so that the lanes served together by the LDS hit distinct banks. The order costs a
// second load trip (the dual cells first, then each lane's point). Measured on the 16M column:
// FP64 P2G 1.33 ms with order 2 against 1.39 in slot order; FP32 P2G 1.26 ms with order 2
// against 1.22 in slot order and 1.20 in slot order with the next point prefetched.
#ifndef MPM_P2G_ORDER32
#define MPM_P2G_ORDER32 0
#endif
#ifndef MPM_P2G_ORDER64
#define MPM_P2G_ORDER64 2
#endif
// G2P pass: load the next particle block's node patch by LDS DMA while this block's points
// run (measured on the 16M column: G2P FP32 0.865 -> 0.838 ms, FP64 1.670 -> 1.532 ms)
#ifndef MPM_G2P_ASYNC_PATCH
#define MPM_G2P_ASYNC_PATCH 1
#endif
#ifndef MPM_DYN_BLOCKS
#define MPM_DYN_BLOCKS 1
#endif
// Which passes take their particle blocks from a counter instead of a fixed stride. Measured
// on the 16M column against the fixed stride: P2G 1.25 -> 1.19 ms (FP32) and 1.44 -> 1.40 ms
// (FP64), G2P FP64 1.63 -> 1.61 ms; but G2P FP32 0.82 -> 0.85 ms (its first prefetch of a block
// waits for the block entry) and the 2D ensemble 219 -> 188 solves/s (blocks of a few points,
// where a counter round trip per block dominates), so those keep the fixed stride.
template <int DIM, bool DO_G2P, bool DO_P2G, bool FP32>
constexpr bool dyn_blocks() {
    return DIM == 3 && DO_G2P != DO_P2G && (DO_P2G || !FP32);
}

// Called by the whole wavefront: lane L holds a point with lane-order key `key` (-1: no
// point; -2: a point without a key, e.g. one leaving the domain). Returns, for every lane,
// the lane whose point it takes (-1: none). One claimant of each key gets lane `key`; the
// points left over (a second point of one dual cell, points without a key) fill the lanes no
// key claimed, in lane order, so every point is taken exactly once. ws, wl: 64 ints each of
// wavefront-private LDS.
__device__ __forceinline__ int lane_source(int *ws, int *wl, int lane, int key) {
    ws[lane] = -1;
    __builtin_amdgcn_wave_barrier();
    if (key >= 0) ws[key] = lane;
    __builtin_amdgcn_wave_barrier();
    const bool left = key == -2 || (key >= 0 && ws[key] != lane);
    const unsigned long long lm = __ballot(left);
    const unsigned long long below = (1ull << lane) - 1ull;
    int src = ws[lane];
    if (lm != 0ull) {  // wavefront-uniform
        if (left) wl[__popcll(lm & below)] = lane;
        __builtin_amdgcn_wave_barrier();
        const unsigned long long fm = __ballot(src < 0);
        if (src < 0) {
            const int r = __popcll(fm & below);
            src = r < (int)__popcll(lm) ? wl[r] : -1;
        }
    }
    __builtin_amdgcn_wave_barrier();
    return src;
}

template <bool DO_G2P, bool DO_P2G, bool FP32> struct PassTune {
    static constexpr int waves = (DO_G2P && DO_P2G) ? 2  // the fused pass
                                 : DO_G2P ? (FP32 ? MPM_WPE_G2P32 : MPM_WPE_G2P64)
                                          : (FP32 ? MPM_WPE_P2G32 : MPM_WPE_P2G64);
    static constexpr bool prefetch = DO_G2P ? MPM_PF_G2P != 0 : (FP32 ? MPM_PF_P2G32 != 0 : MPM_PF_P2G64 != 0);
};

// Particle pass over one particle block per workgroup (persistent over the block list).
//   DO_G2P: gather the grid of the current step (LDS patch), update the point
//   DO_P2G: map the (updated) point and scatter it into the next step's accumulators
// NRM: accumulate contact normals gradS*m for bodies that are body 1 of a contact.
// PT: particle-state form (Prec64 / Prec32). LDS planes are [field][padded node index]
// (Cfg::PSTR elements per plane); the accumulator patch is FP64 in both forms, the
// gathered grid patch has the grid output's type.
template <int DIM, bool DO_G2P, bool DO_P2G, bool NRM, class PT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PassTune<DO_G2P, DO_P2G, PT::FP32>::waves)))
void k_particles(KArgs A, int want_L) {
    constexpr int ORDER = (!DO_G2P && DO_P2G) ? (PT::FP32 ? MPM_P2G_ORDER32 : MPM_P2G_ORDER64) : 0;
    constexpr bool DYN = MPM_DYN_BLOCKS != 0 && dyn_blocks<DIM, DO_G2P, DO_P2G, PT::FP32>();
    constexpr bool SIDE = side_pass<DO_G2P, DO_P2G, PT::FP32>();
    constexpr bool PREFETCH = PassTune<DO_G2P, DO_P2G, PT::FP32>::prefetch;
    using R = typename PT::R;
    using GO = typename PT::GO;
    using St = typename PT::template State<DIM>;
    constexpr int P = Cfg<DIM>::P, PSTR = Cfg<DIM>::PSTR, ZS = Cfg<DIM>::ZS;
    constexpr int NACC = NRM ? GF<DIM>::NACC : 1 + 2 * DIM;
    constexpr int NOUT = GF<DIM>::NOUT;
    constexpr int NZ = DIM == 3 ? 3 : 1;
    constexpr int PBLKS = Cfg<DIM>::PBLKS;
    // P2G accumulator patch, FP64 in both forms. (Keeping the precision form's three force
    // planes in FP32 with ds_add_f32 took the 16M-column P2G from 1.25 to 3.43 ms on gfx950,
    // with 22x the LDS issue waits per wavefront at the same LDS instruction count.)
    __shared__ double lacc[DO_P2G ? NACC * PSTR : 1];
    // (16-byte aligned: the FP32 G2P layout pass moves 16 bytes per access)
    __shared__ __attribute__((aligned(16))) GO gpat[DO_G2P ? NOUT * PSTR : 1];
    // G2P pass: the next particle block's node patch, loaded by LDS DMA while this block's
    // points run, in x-pair order [y (, z) pair][field][2 node blocks x 64 nodes]
    constexpr bool ASYNC = MPM_G2P_ASYNC_PATCH != 0 && DO_G2P && !DO_P2G;
    constexpr int NPAIR = PBLKS / 2;
    __shared__ __attribute__((aligned(16))) GO graw[ASYNC ? NPAIR * NOUT * 128 : 1];
    bool pre = false;  // graw holds this block's patch (wavefront-uniform)
    bool first_blk = true;
    __shared__ int bb[2 * DIM];
    __shared__ int wslot[ORDER ? 512 : 1];  // per wavefront: the lane holding each key; left-over points
    const int tid = threadIdx.x;
    const int nent = *A.npblk;
    const R sscale = (R)1 / (R)(1 << DIM);
    // the grid update this pass gathers wrote a non-finite value (k_grid, parity par ^ 1)
    const bool grid_nf = DO_G2P && A.nonfin[A.par ^ 1] != 0u;
    // Block order. DYN: workgroups take blocks from a counter, the next index fetched one
    // block ahead (e_sh double-buffered: a slot is rewritten only after every thread read it
    // behind the end-of-block barrier), so no workgroup idles while others have blocks left.
    // Otherwise workgroup w takes blocks w, w + grid, ..., the next entry read one block ahead.
    __shared__ int e_sh[2];
    // the counter costs two contended atomics per workgroup and one per block: taken only
    // when there are many blocks per workgroup (small simulations keep the fixed stride)
    const bool dyn = DYN && nent >= 8 * (int)gridDim.x;
    int e = blockIdx.x, epar = 0;
    if (dyn) {
        if (tid == 0) e_sh[0] = (int)atomicAdd(&A.work[0], 1u);
        __syncthreads();
        e = uniform(e_sh[0]);
    }
    PBlk ent_next = {0, 0, 0, 0};
    if (!dyn && e < nent) ent_next = A.plist[e];
#pragma unroll 1
    while (e < nent) {
        if (dyn && tid == 0) e_sh[epar ^ 1] = (int)atomicAdd(&A.work[0], 1u);
        const PBlk ent = dyn ? A.plist[e] : ent_next;
        if (!dyn && e + (int)gridDim.x < nent) ent_next = A.plist[e + gridDim.x];
        const int body = uniform(ent.body);
        const int start = uniform(ent.start), count = uniform(ent.count);
        const BodyDev &B = A.bodies[body];
        const SimDev &s = A.sims[B.sim];
        const bool b1 = NRM && B.is_b1 != 0;
        const long long gbase = B.grid_base;
        const double dt = s.dt;
        const R dtr = (R)dt;
        int org[DIM];
        patch_origin<DIM>(s, uniform(ent.key) - B.pblk_off, org);
        if (DO_P2G) {
            // lacc is all zero here: cleared before the first block, and the flush clears
            // what a block wrote (every write lies inside the flushed hull)
            if (first_blk)
                for (int i = tid; i < NACC * PSTR; i += blockDim.x) lacc[i] = 0.0;
            first_blk = false;
            if (tid < DIM) {
                bb[tid] = 0x7fffffff;
                bb[DIM + tid] = -0x7fffffff;
            }
        }
        if (ASYNC && pre) {
            // the patch arrived by DMA during the previous block: lay it out like the
            // staging below (zeros outside the domain). FP32: a thread moves four
            // x-consecutive nodes of a node block, which sit at 16-byte aligned offsets in
            // both layouts (the x run of a node block is 4 or 8 nodes, and P, ZS and PSTR are
            // multiples of 4), with one 16-byte LDS read and write per field. FP64 moves one
            // node per thread. (Measured on the 16M column: FP32 G2P 0.830 -> 0.815 ms with
            // the 4-node moves; FP64 G2P 1.533 ms node by node, 1.555 with 2-node and 1.570
            // with 4-node moves.)
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            constexpr int V = sizeof(GO) == 4 ? 4 : 1;
            typedef GO GOV __attribute__((ext_vector_type(4)));
            for (int qd = tid; qd < PBLKS * 64 / V; qd += blockDim.x) {
                int nijk[DIM];
                bool inp, in;
                (void)patch_node<DIM>(s, org, qd * V, nijk, inp, in);
                if (inp) {  // the first node lies in the patch, so do the others
                    bool inr = true;  // the row's y (and z) lie in the domain
#pragma unroll
                    for (int d = 1; d < DIM; d++) inr = inr && nijk[d] >= 0 && nijk[d] <= s.N[d];
                    bool inx[V];
#pragma unroll
                    for (int k = 0; k < V; k++) inx[k] = V == 1 ? in : inr && nijk[0] + k >= 0 && nijk[0] + k <= s.N[0];
                    const int li = patch_index<DIM>(nijk, org);
                    const int b = (qd * V) >> 6;
                    const int src = (b >> 1) * NOUT * 128 + (b & 1) * 64 + ((qd * V) & 63);
#pragma unroll
                    for (int f = 0; f < NOUT; f++) {
                        if constexpr (V == 4) {
                            GOV v = *reinterpret_cast<const GOV *>(&graw[src + f * 128]);
#pragma unroll
                            for (int k = 0; k < 4; k++) v[k] = inx[k] ? v[k] : (GO)0;
                            *reinterpret_cast<GOV *>(&gpat[f * PSTR + li]) = v;
                        } else {
                            gpat[f * PSTR + li] = inx[0] ? graw[src + f * 128] : (GO)0;
                        }
                    }
                }
            }
        } else if (DO_G2P) {
            // stage the block's node patch. The patch starts on a node-block boundary, so it
            // is read as whole 64-node blocks: one node per thread, all its fields at once.
            for (int i = tid; i < PBLKS * 64; i += blockDim.x) {
                int nijk[DIM];
                bool inp, in;
                const long long gi = gbase + patch_node<DIM>(s, org, i, nijk, inp, in);
                if (inp) {
                    const int li = patch_index<DIM>(nijk, org);
#pragma unroll
                    for (int f = 0; f < NOUT; f++) gpat[f * PSTR + li] = in ? PT::fetch(A, f * A.nt + gi) : (GO)0;
                }
            }
        }
        __syncthreads();
        // hull of the LDS-path windows, per thread (min / max node index per dimension);
        // combined over the workgroup once per particle block
        int tlo[DIM], thi[DIM];
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            tlo[d] = 0x7fffffff;
            thi[d] = -0x7fffffff;
        }
        // with PREFETCH the next point's state is loaded before this one is processed, so
        // its loads are in flight during the LDS work (different slots: no ordering hazard).
        // (Carrying the prefetch over to the next particle block was measured: no gain in
        // P2G, and more registers than it saves in latency in G2P.)
        typename PT::template Raw<DIM> qn;
        bool an = false;
        // the volume is needed by P2G and by the sand's density (G2P of a sand body)
        const bool need_vol = DO_P2G || B.has_material == 2;
        // a slot's state is read whether or not the point is active, so that the reads of
        // the activity flag and of the state are in flight together
        if (PREFETCH && tid < count) {
            an = A.act[start + tid] != 0;
            PT::template load_raw<DIM>(A, start + tid, qn, need_vol);
        }
        const int lane = tid & 63;
        if (ASYNC) {
            // the workgroup's next block: DMA its patch into graw (after the prefetch above,
            // whose wait must not include these loads). A group moves the two x-adjacent
            // node blocks (contiguous in storage) of one field, 16 bytes a lane: one group
            // per instruction in FP64, two in FP32; node blocks outside the simulation's
            // block grid are not read.
            // (the block list and descriptors are read through the constant address space:
            // scalar loads, whose wait does not take in the vector loads in flight)
            pre = false;
            const int en = dyn ? uniform(e_sh[epar ^ 1]) : e + (int)gridDim.x;
            if (en < nent) {
                using CPB = const __attribute__((address_space(4))) PBlk;
                using CBD = const __attribute__((address_space(4))) BodyDev;
                using CSD = const __attribute__((address_space(4))) SimDev;
                CPB &ne = ((CPB *)A.plist)[en];
                CBD &B2 = ((CBD *)A.bodies)[ne.body];
                CSD &s2 = ((CSD *)A.sims)[B2.sim];
                int org2[DIM];
                {  // patch_origin, on the scalar copy
                    long long lk = ne.key - B2.pblk_off;
#pragma unroll
                    for (int d = 0; d < DIM; d++) {
                        const int pb = (int)(lk % s2.PNB[d]);
                        lk /= s2.PNB[d];
                        org2[d] = pb * Cfg<DIM>::PB - 2;
                    }
                }
                int bc[DIM];
#pragma unroll
                for (int d = 0; d < DIM; d++) bc[d] = (org2[d] + NSHIFT) / Cfg<DIM>::NB;
                constexpr int PER = 16 / (int)sizeof(GO);  // values per lane
                constexpr int LPJ = 128 / PER;             // lanes per (pair, field): two node blocks
                constexpr int JPI = 64 / LPJ;              // (pair, field) groups per instruction
                const int jl = lane / LPJ, l2 = lane - jl * LPJ;
                const int xb = bc[0] + (l2 >= LPJ / 2 ? 1 : 0);
                const GO *src0 = sizeof(GO) == 8 ? reinterpret_cast<const GO *>(A.out) : reinterpret_cast<const GO *>(A.out32);
                for (int j0 = (tid >> 6) * JPI; j0 < NPAIR * NOUT; j0 += (int)(blockDim.x >> 6) * JPI) {
                    // graw[j * 128] holds group j, so the groups j0 .. j0 + JPI - 1 of one
                    // instruction land contiguously, lane after lane
                    const int j = j0 + jl;
                    const int pair = j / NOUT, f = j - pair * NOUT;
                    const int by = bc[1] + (pair & 1), bz = DIM == 3 ? bc[DIM - 1] + (pair >> 1) : 0;
                    if (j < NPAIR * NOUT && xb < s2.NBn[0] && by < s2.NBn[1] && (DIM == 2 || bz < s2.NBn[DIM - 1])) {
                        const long long blk = bc[0] + (long long)s2.NBn[0] * (by + (long long)s2.NBn[1] * bz);
                        const GO *g = src0 + (long long)f * A.nt + B2.grid_base + blk * 64 + (long long)l2 * PER;
                        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)g,
                                                         (__attribute__((address_space(3))) void *)(graw + j0 * 128),
                                                         16, 0, 0);
                    }
                }
                pre = true;
            }
        }
#pragma unroll 1
        for (int kw = ORDER ? tid - lane : tid; kw < count; kw += blockDim.x) {  // ORDER: wavefront-uniform
            int k = ORDER ? kw + lane : kw;
            St q;
            bool a;
            if (ORDER) {
                // choose the lane order from the dual cells first, then load each lane's
                // point: a second round trip, but no state held across the permutation
                // (loading in slot order and moving the state across lanes with __shfl
                // costs the FP32 pass 22 VGPRs and one occupancy step; the FP64 state would
                // cost more)
                int key = -1;
                if (k < count) {
                    const bool act = A.act[start + k] != 0;
                    int dual[DIM], oct;
                    const bool in = PT::template dual_cell<DIM>(A, s, start + k, dual, oct);
                    if (act) key = in ? lane_key<DIM, ORDER>(dual) : -2;
                }
                // every lane given a source takes an active point
                const int src = lane_source(&wslot[(tid >> 6) * 64], &wslot[256 + (tid >> 6) * 64], lane, key);
                a = src >= 0;
                k = a ? kw + src : count;
                if (a) PT::template load<DIM>(A, start + k, q, DO_P2G);
            } else if (PREFETCH) {
                a = an;
                PT::template cook<DIM>(qn, q, need_vol);
                const int kn = k + blockDim.x;
                an = false;
                if (kn < count) {
                    an = A.act[start + kn] != 0;
                    PT::template load_raw<DIM>(A, start + kn, qn, need_vol);
                }
            } else {
                a = false;
                if (k < count) {
                    a = A.act[start + k] != 0;
                    PT::template load<DIM>(A, start + k, q, need_vol);
                }
            }
            if (!a) continue;
            const long long p = start + k;
            if (DO_G2P) {
                Win<DIM, R> w;
                bool cpdi;
                const int mode = PT::template window<DIM>(s, q, w, cpdi);
                if (mode == MODE_LEFT) {  // not reachable after this step's map; kept for safety
                    A.act[p] = 0;
                    continue;
                }
                // After a grid update with a non-finite node value every point takes the slot-
                // by-slot path over the slots the reference's tuples cover: the row-factored
                // gather multiplies every slot, and 0 * NaN would reach points the reference
                // spares.
                const bool lds_path = mode == MODE_SEP && window_in_patch<DIM>(w, org) && !grid_nf;
                if (SIDE && (mode == MODE_TUP || (mode == MODE_SEP && !lds_path))) {
                    side_push(A.rare, &A.rare_n[1], p);  // k_g2p_side updates it
                    continue;
                }
                R up[DIM], dvp[DIM], Lp[DIM][DIM];
#pragma unroll
                for (int d = 0; d < DIM; d++) {
                    up[d] = dvp[d] = 0;
#pragma unroll
                    for (int c = 0; c < DIM; c++) Lp[d][c] = 0;
                }
                if (lds_path) {
                    // row-factored gather: per row, U = sum sw0 v, DV = sum sw0 dv, LX = sum dw0 v
                    const R gscale = grad_scale<DIM, R>(q.ext);
                    const int i0 = patch_index<DIM>(w.base, org);
#pragma unroll
                    for (int c2 = 0; c2 < NZ; c2++) {
#pragma unroll
                        for (int c1 = 0; c1 < 3; c1++) {
                            const RowW<R> rw = row_weights<DIM, R>(w, c1, c2, sscale, gscale);
                            R U[DIM], DV[DIM], LX[DIM];
#pragma unroll
                            for (int d = 0; d < DIM; d++) U[d] = DV[d] = LX[d] = 0;
#pragma unroll
                            for (int c0 = 0; c0 < 3; c0++) {
                                const int i = i0 + c0 + P * c1 + ZS * c2;
#pragma unroll
                                for (int d = 0; d < DIM; d++) {
                                    const R vn = gpat[d * PSTR + i];
                                    const R dvn = gpat[(DIM + d) * PSTR + i];
                                    U[d] = fma(w.sw[0][c0], vn, U[d]);
                                    DV[d] = fma(w.sw[0][c0], dvn, DV[d]);
                                    LX[d] = fma(w.dw[0][c0], vn, LX[d]);
                                }
                            }
                            const R sdt = rw.Su * dtr;
#pragma unroll
                            for (int d = 0; d < DIM; d++) {
                                up[d] = fma(sdt, U[d], up[d]);
                                dvp[d] = fma(rw.Su, DV[d], dvp[d]);
                                Lp[d][0] = fma(rw.gx, LX[d], Lp[d][0]);
                                Lp[d][1] = fma(rw.gy, U[d], Lp[d][1]);
                                if (DIM == 3) Lp[d][DIM - 1] = fma(rw.gz, U[d], Lp[d][DIM - 1]);
                            }
                            // keep each row's LDS loads from being hoisted into earlier rows
                            __builtin_amdgcn_sched_barrier(0);
                        }
                    }
                } else if (!SIDE && mode == MODE_SEP) {
                    for_each_sep<DIM, R>(w, q.ext, [&](const int *off, R S, const R *G) {
                        int nijk[DIM];
#pragma unroll
                        for (int d = 0; d < DIM; d++) nijk[d] = w.base[d] + off[d];
                        R vn[DIM], dvn[DIM];
                        fetch_node<DIM, PT>(A, s, gbase, nijk, vn, dvn);
                        g2p_acc<DIM, R>(vn, dvn, S, G, dtr, up, dvp, Lp);
                    });
                } else if (!SIDE && mode == MODE_TUP) {
                    PT::template tuples<DIM>(s, q, cpdi, [&](const int *nijk, R S, const R *G) {
                        R vn[DIM], dvn[DIM];
                        if (node_in_patch<DIM>(nijk, org)) {
                            const int i = patch_index<DIM>(nijk, org);
#pragma unroll
                            for (int d = 0; d < DIM; d++) {
                                vn[d] = gpat[d * PSTR + i];
                                dvn[d] = gpat[(DIM + d) * PSTR + i];
                            }
                        } else {
                            fetch_node<DIM, PT>(A, s, gbase, nijk, vn, dvn);
                        }
                        g2p_acc<DIM, R>(vn, dvn, S, G, dtr, up, dvp, Lp);
                    });
                }
                PT::template advance<DIM>(B, s, dt, q, up, dvp, Lp);
                PT::template store<DIM>(A, p, q);
                if (B.has_material == 2) PT::template store_plastic<DIM>(A, p, q.nup, dt);
                if (want_L && A.Lout) {
                    double L9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
                    for (int r = 0; r < DIM; r++)
#pragma unroll
                        for (int c = 0; c < DIM; c++) L9[3 * r + c] = Lp[r][c];
#pragma unroll
                    for (int kk = 0; kk < 9; kk++) A.Lout[kk * A.np + p] = L9[kk];
                }
            }
            if (DO_P2G) {
                Win<DIM, R> w;
                bool cpdi;
                const int mode = PT::template window<DIM>(s, q, w, cpdi);
                if (mode == MODE_LEFT) {  // leaves the domain: deactivated (default_points.cpp:217-219)
                    A.act[p] = 0;
                    continue;
                }
                if (mode == MODE_NONE) continue;
                if (SIDE && (mode == MODE_TUP || (mode == MODE_SEP && !window_in_patch<DIM>(w, org)))) {
                    side_push(A.rare, &A.rare_n[0], p);  // k_p2g_side scatters it
                    continue;
                }
                P2GVals<DIM, R> pv;
                p2g_payload<DIM, R>(q, pv);
                auto scatter_global = [&](const int *nijk, R S, const R *G) {
                    R vals[GF<DIM>::NACC];
                    p2g_node_values<DIM, R>(pv, S, G, vals, b1);
                    const long long gi = gbase + node_addr<DIM>(s, nijk);
#pragma unroll
                    for (int f = 0; f < NACC; f++) {
                        if (f >= 1 + 2 * DIM && !b1) continue;
                        unsafeAtomicAdd(&A.acc[f * A.nt + gi], (double)vals[f]);
                    }
                };
                if (mode == MODE_SEP && window_in_patch<DIM>(w, org)) {
#pragma unroll
                    for (int d = 0; d < DIM; d++) {
                        tlo[d] = min(tlo[d], w.base[d]);
                        thi[d] = max(thi[d], w.base[d] + 2);
                    }
                    // row-factored scatter: along a row every field is a * sw0 + b * dw0. The
                    // node values are formed in FP64 from the point's values (in the precision
                    // form too: scalar FP32 and FP64 FMAs issue at one rate, and the LDS
                    // accumulators are FP64, so this saves a conversion per value)
                    using D = double;
                    const D gscale = (D)grad_scale<DIM, R>(q.ext);
                    const D dm = pv.m;
                    D dmv[DIM], dVT[DIM][DIM], sx0[3], dx0[3];
#pragma unroll
                    for (int d = 0; d < DIM; d++) {
                        dmv[d] = pv.mv[d];
#pragma unroll
                        for (int c = 0; c < DIM; c++) dVT[d][c] = pv.VT[d][c];
                    }
#pragma unroll
                    for (int c0 = 0; c0 < 3; c0++) {
                        sx0[c0] = w.sw[0][c0];
                        dx0[c0] = w.dw[0][c0];
                    }
                    const int i0 = patch_index<DIM>(w.base, org);
#pragma unroll
                    for (int c2 = 0; c2 < NZ; c2++) {
#pragma unroll
                        for (int c1 = 0; c1 < 3; c1++) {
                            // window slots no corner reaches hold exact zeros (weight and
                            // gradient): their rows and nodes add nothing and are skipped. (A
                            // point reaching P2G has a finite position; a NaN node value in G2P
                            // makes the position NaN too, and such points leave at the map.)
                            if (!(w.sw[1][c1] != (R)0 || w.dw[1][c1] != (R)0)) continue;
                            if (DIM == 3 && !(w.sw[DIM - 1][c2] != (R)0 || w.dw[DIM - 1][c2] != (R)0)) continue;
                            const RowW<D> rw = row_weights<DIM, R, D>(w, c1, c2, (D)sscale, gscale);
                            const D am = dm * rw.Su;
                            D ap[DIM], af0[DIM], af1[DIM];
#pragma unroll
                            for (int d = 0; d < DIM; d++) {
                                ap[d] = dmv[d] * rw.Su;
                                af0[d] = -(dVT[d][0] * rw.gx);
                                D t = dVT[d][1] * rw.gy;
                                if (DIM == 3) t = fma(dVT[d][DIM - 1], rw.gz, t);
                                af1[d] = -t;
                            }
                            const D an0 = dm * rw.gx, an1 = dm * rw.gy, an2 = dm * rw.gz;
#pragma unroll
                            for (int c0 = 0; c0 < 3; c0++) {
                                const int i = i0 + c0 + P * c1 + ZS * c2;
                                const D sx = sx0[c0], dx = dx0[c0];
                                if (sx == 0.0 && dx == 0.0) continue;
                                atomicAdd(&lacc[GF<DIM>::M * PSTR + i], am * sx);
#pragma unroll
                                for (int d = 0; d < DIM; d++) {
                                    atomicAdd(&lacc[(GF<DIM>::P + d) * PSTR + i], ap[d] * sx);
                                    atomicAdd(&lacc[(GF<DIM>::F + d) * PSTR + i], fma(af0[d], dx, af1[d] * sx));
                                }
                                if (NRM && b1) {
                                    constexpr int N0 = GF<DIM>::NRM;
                                    atomicAdd(&lacc[(N0 + 0) * PSTR + i], an0 * dx);
                                    atomicAdd(&lacc[(N0 + 1) * PSTR + i], an1 * sx);
                                    if (DIM == 3) atomicAdd(&lacc[(N0 + DIM - 1) * PSTR + i], an2 * sx);
                                }
                            }
                        }
                    }
                } else if (!SIDE && mode == MODE_SEP) {
                    int lo[DIM], hi[DIM];
#pragma unroll
                    for (int d = 0; d < DIM; d++) {
                        lo[d] = w.base[d];
                        hi[d] = w.base[d] + 2;
                    }
                    mark_range<DIM>(A, s, B, lo, hi);
                    for_each_sep<DIM, R, false>(w, q.ext, [&](const int *off, R S, const R *G) {
                        int nijk[DIM];
#pragma unroll
                        for (int d = 0; d < DIM; d++) nijk[d] = w.base[d] + off[d];
                        scatter_global(nijk, S, G);
                    });
                } else if (!SIDE) {  // MODE_TUP
                    PT::template tuples<DIM>(s, q, cpdi, [&](const int *nijk, R S, const R *G) {
                        if (node_in_patch<DIM>(nijk, org)) {
#pragma unroll
                            for (int d = 0; d < DIM; d++) {
                                tlo[d] = min(tlo[d], nijk[d]);
                                thi[d] = max(thi[d], nijk[d]);
                            }
                            R vals[GF<DIM>::NACC];
                            p2g_node_values<DIM, R>(pv, S, G, vals, b1);
                            const int i = patch_index<DIM>(nijk, org);
                            for (int f = 0; f < NACC; f++) {
                                if (f >= 1 + 2 * DIM && !b1) continue;
                                atomicAdd(&lacc[f * PSTR + i], (double)vals[f]);
                            }
                        } else {
                            mark_range<DIM>(A, s, B, nijk, nijk);
                            scatter_global(nijk, S, G);
                        }
                    });
                }
            }
        }
        if (DO_P2G) {
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                if (tlo[d] <= thi[d]) {
                    atomicMin(&bb[d], tlo[d]);
                    atomicMax(&bb[DIM + d], thi[d]);
                }
            }
            __syncthreads();
            // flush the touched part of the patch and mark its node blocks
            int lo[DIM], ex[DIM];
            int total = 1;
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                lo[d] = bb[d];
                ex[d] = bb[DIM + d] >= bb[d] ? bb[DIM + d] - bb[d] + 1 : 0;  // empty: no LDS-path point
                total *= ex[d];
            }
            for (int i = tid; i < total; i += blockDim.x) {
                int r = i;
                int nijk[DIM];
#pragma unroll
                for (int d = 0; d < DIM; d++) {
                    nijk[d] = lo[d] + r % ex[d];
                    r /= ex[d];
                }
                const int li = patch_index<DIM>(nijk, org);
                const long long gi = gbase + node_addr<DIM>(s, nijk);
#pragma unroll
                for (int f = 0; f < NACC; f++) {
                    if (f >= 1 + 2 * DIM && !b1) continue;
                    const double v = lacc[f * PSTR + li];
                    if (v != 0.0) {
                        unsafeAtomicAdd(&A.acc[f * A.nt + gi], v);
                        lacc[f * PSTR + li] = 0.0;
                    }
                }
            }
            if (tid == 0 && total > 0) {
                int hi[DIM];
#pragma unroll
                for (int d = 0; d < DIM; d++) hi[d] = lo[d] + ex[d] - 1;
                mark_range<DIM>(A, s, B, lo, hi);
            }
        }
        __syncthreads();
        if (dyn) {
            e = uniform(e_sh[epar ^ 1]);
            epar ^= 1;
        } else {
            e += gridDim.x;
        }
    }
    if (ASYNC) __builtin_amdgcn_s_waitcnt(0);  // no LDS DMA in flight when the workgroup ends
    if (dyn) {
        // the last workgroup to finish clears the counter for the next pass: every workgroup
        // read its final (out of range) index before it signs off here
        __syncthreads();
        if (tid == 0) {
            __threadfence();
            if (atomicAdd(&A.work[1], 1u) == gridDim.x - 1) {
                A.work[0] = 0;
                A.work[1] = 0;
            }
        }
    }
}

// ---------------------------------------------------------------------------------------
// Side kernels of the split passes: the points the LDS path left (window not inside the
// block's patch, or the per-tuple form: CPDI off, a corner outside the domain, corners over
// more than two elements), grid-stride over the device-side count, straight to and from
// global memory. Same arithmetic as the inline paths of the fused pass.

// P2G of the listed points: global FP64 atomics, node blocks marked for the grid update.
template <int DIM, bool NRM, class PT>
__global__ __launch_bounds__(256) void k_p2g_side(KArgs A) {
    using R = typename PT::R;
    using St = typename PT::template State<DIM>;
    constexpr int NACC = NRM ? GF<DIM>::NACC : 1 + 2 * DIM;
    const unsigned n = A.rare_n[0];
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const long long p = A.rare[i];
        const BodyDev &B = A.bodies[body_of_index(A, A.pid[p])];
        const SimDev &s = A.sims[B.sim];
        const bool b1 = NRM && B.is_b1 != 0;
        const long long gbase = B.grid_base;
        St q;
        PT::template load<DIM>(A, p, q, true);
        Win<DIM, R> w;
        bool cpdi;
        const int mode = PT::template window<DIM>(s, q, w, cpdi);
        P2GVals<DIM, R> pv;
        p2g_payload<DIM, R>(q, pv);
        auto scatter = [&](const int *nijk, R S, const R *G) {
            R vals[GF<DIM>::NACC];
            p2g_node_values<DIM, R>(pv, S, G, vals, b1);
            const long long gi = gbase + node_addr<DIM>(s, nijk);
            for (int f = 0; f < NACC; f++) {
                if (f >= 1 + 2 * DIM && !b1) continue;
                unsafeAtomicAdd(&A.acc[f * A.nt + gi], (double)vals[f]);
            }
        };
        if (mode == MODE_SEP) {
            int lo[DIM], hi[DIM];
            for (int d = 0; d < DIM; d++) {
                lo[d] = w.base[d];
                hi[d] = w.base[d] + 2;
            }
            mark_range<DIM>(A, s, B, lo, hi);
            for_each_sep<DIM, R>(w, q.ext, [&](const int *off, R S, const R *G) {
                int nijk[DIM];
                for (int d = 0; d < DIM; d++) nijk[d] = w.base[d] + off[d];
                scatter(nijk, S, G);
            });
        } else if (mode == MODE_TUP) {
            PT::template tuples<DIM>(s, q, cpdi, [&](const int *nijk, R S, const R *G) {
                mark_range<DIM>(A, s, B, nijk, nijk);
                scatter(nijk, S, G);
            });
        }
    }
    side_done(A.rare_n, 0);
}

// G2P of the listed points: gather from the global grid output, then movePoints, strain
// rate, density and stress update, and the stores, as in k_particles.
template <int DIM, class PT>
__global__ __launch_bounds__(256) void k_g2p_side(KArgs A, int want_L) {
    using R = typename PT::R;
    using St = typename PT::template State<DIM>;
    const unsigned n = A.rare_n[1];
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const long long p = A.rare[i];
        const BodyDev &B = A.bodies[body_of_index(A, A.pid[p])];
        const SimDev &s = A.sims[B.sim];
        const long long gbase = B.grid_base;
        const double dt = s.dt;
        const R dtr = (R)dt;
        St q;
        PT::template load<DIM>(A, p, q, B.has_material == 2);
        Win<DIM, R> w;
        bool cpdi;
        const int mode = PT::template window<DIM>(s, q, w, cpdi);
        R up[DIM], dvp[DIM], Lp[DIM][DIM];
        for (int d = 0; d < DIM; d++) {
            up[d] = dvp[d] = 0;
            for (int c = 0; c < DIM; c++) Lp[d][c] = 0;
        }
        auto gather = [&](const int *nijk, R S, const R *G) {
            R vn[DIM], dvn[DIM];
            fetch_node<DIM, PT>(A, s, gbase, nijk, vn, dvn);
            g2p_acc<DIM, R>(vn, dvn, S, G, dtr, up, dvp, Lp);
        };
        if (mode == MODE_SEP) {
            for_each_sep<DIM, R>(w, q.ext, [&](const int *off, R S, const R *G) {
                int nijk[DIM];
                for (int d = 0; d < DIM; d++) nijk[d] = w.base[d] + off[d];
                gather(nijk, S, G);
            });
        } else if (mode == MODE_TUP) {
            PT::template tuples<DIM>(s, q, cpdi, gather);
        }
        PT::template advance<DIM>(B, s, dt, q, up, dvp, Lp);
        PT::template store<DIM>(A, p, q);
        if (B.has_material == 2) PT::template store_plastic<DIM>(A, p, q.nup, dt);
        if (want_L && A.Lout) {
            double L9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int r = 0; r < DIM; r++)
                for (int c = 0; c < DIM; c++) L9[3 * r + c] = Lp[r][c];
            for (int kk = 0; kk < 9; kk++) A.Lout[kk * A.np + p] = L9[kk];
        }
    }
    side_done(A.rare_n, 1);
}

// ---------------------------------------------------------------------------------------
// grid update: body force, contact, boundary, moveGrid. One wavefront per node block.
// NBM: bodies per simulation the kernel handles (1, 2 or MAXB), chosen by the host from the
// largest simulation in the handle: the per-body node values live in registers, and sizing
// them to 1 body keeps a one-body grid update at 8 waves per SIMD instead of 2.
template <int DIM, class GO, int NBM>
__global__ __launch_bounds__(256) void k_grid(KArgs A) {
    GO *outp = nullptr;
    if constexpr (sizeof(GO) == 8) outp = reinterpret_cast<GO *>(A.out);
    else outp = reinterpret_cast<GO *>(A.out32);
    using G = GF<DIM>;
    constexpr int NB = Cfg<DIM>::NB, L = Cfg<DIM>::LNB;
    const int par = A.par;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        A.ncount[par ^ 1] = 0;
        A.nonfin[par ^ 1] = 0u;  // the flag of the update before this one
    }
    const int cnt = A.ncount[par];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int wpb = blockDim.x >> 6;
    unsigned long long active_nodes = 0;
    for (int bi = blockIdx.x * wpb + wv; bi < cnt; bi += gridDim.x * wpb) {
        const long long gblk = A.nlist[(long long)par * A.nblk_total + bi];
        // the block's flag: the bodies that scattered into it and the simulation's index
        const unsigned flag = (NBM > 1 || A.nsim > 1) ? (unsigned)uniform((int)A.nflag[gblk]) : 1u;
        const int si = (int)(flag >> 8);
        const SimDev &s = A.sims[si];
        const long long lb = gblk - s.nblk_off;
        int ijk[3] = {0, 0, 0};
        {
            long long t = lb;
            const int bx = (int)(t % s.NBn[0]);
            t /= s.NBn[0];
            ijk[0] = (bx << L) + (lane & (NB - 1)) - NSHIFT;  // storage is shifted (node_addr)
            if (DIM == 3) {
                const int by = (int)(t % s.NBn[1]);
                const int bz = (int)(t / s.NBn[1]);
                ijk[1] = (by << L) + ((lane >> L) & (NB - 1)) - NSHIFT;
                ijk[2] = (bz << L) + (lane >> (2 * L)) - NSHIFT;
            } else {
                ijk[1] = ((int)t << L) + (lane >> L) - NSHIFT;
            }
        }
        bool valid = true;
        bool wall = false;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            valid = valid && ijk[d] >= 0 && ijk[d] <= s.N[d];
            wall = wall || ijk[d] == 0 || ijk[d] == s.N[d];
        }
        const long long addr = lb * 64 + lane;
        // bodies of the sim that scattered into this block: the others hold zero accumulators
        // here, which are not read (or re-zeroed); their outputs are still written (zeros)
        const unsigned touched = NBM == 1 ? 1u : flag & 0xffu;
        if (valid) {
            const int nb = s.nbody;
            // f0: the force before contacts act, which the frictional box takes its normal
            // impulse from (cartesian_frictional_box.cpp:102-117 recomputes it from S and gradS)
            double m[NBM], p[NBM][DIM], f[NBM][DIM], f0[NBM][DIM], nr[NBM][DIM];
#pragma unroll
            for (int b = 0; b < NBM; b++) {
                m[b] = 0.0;
#pragma unroll
                for (int d = 0; d < DIM; d++) p[b][d] = f[b][d] = f0[b][d] = nr[b][d] = 0.0;
                if (b < nb && ((touched >> b) & 1u)) {
                    const BodyDev &B = A.bodies[s.body0 + b];
                    const long long gi = B.grid_base + addr;
                    m[b] = A.acc[G::M * A.nt + gi];
                    A.acc[G::M * A.nt + gi] = 0.0;
#pragma unroll
                    for (int d = 0; d < DIM; d++) {
                        p[b][d] = A.acc[(G::P + d) * A.nt + gi];
                        // nodal body force m_i g (default_driver.cpp:109-119 via explicit_usl.cpp:149-154)
                        f[b][d] = fma(m[b], s.g[d], A.acc[(G::F + d) * A.nt + gi]);
                        f0[b][d] = f[b][d];
                        A.acc[(G::P + d) * A.nt + gi] = 0.0;
                        A.acc[(G::F + d) * A.nt + gi] = 0.0;
                    }
                    if (B.is_b1) {
#pragma unroll
                        for (int d = 0; d < DIM; d++) {
                            nr[b][d] = A.acc[(G::NRM + d) * A.nt + gi];
                            A.acc[(G::NRM + d) * A.nt + gi] = 0.0;
                        }
                    }
                    if (m[b] > 0.0) active_nodes++;
                }
            }
            const double dt = s.dt;
            // ContactHuang::applyRules, contacts in configuration order
            for (int c = 0; NBM >= 2 && c < s.ncontact; c++) {  // a contact needs two bodies
                const ContactDev C = A.contacts[s.contact0 + c];
                double m1 = 0, m2 = 0, p1[DIM] = {}, p2[DIM] = {}, f1[DIM] = {}, f2[DIM] = {}, n1[DIM] = {};
#pragma unroll
                for (int b = 0; b < NBM; b++) {
                    if (b == C.b1) {
                        m1 = m[b];
#pragma unroll
                        for (int d = 0; d < DIM; d++) { p1[d] = p[b][d]; f1[d] = f[b][d]; n1[d] = nr[b][d]; }
                    }
                    if (b == C.b2) {
                        m2 = m[b];
#pragma unroll
                        for (int d = 0; d < DIM; d++) { p2[d] = p[b][d]; f2[d] = f[b][d]; }
                    }
                }
                // record: contact_normal at every node (generateRules, contact_huang.cpp:80-99:
                // 0/0 where body 1 has no mass gradient), contact_force 0 unless set below
                double *cr = nullptr;
                const long long S = s.nblk * 64;
                if (A.crec) {
                    cr = A.crec + C.rec;
                    double nn = 0.0;
#pragma unroll
                    for (int d = 0; d < DIM; d++) nn += n1[d] * n1[d];
                    const double rn = 1.0 / sqrt(nn);
#pragma unroll
                    for (int d = 0; d < DIM; d++) {
                        cr[d * S + addr] = n1[d] * rn;
                        cr[(DIM + d) * S + addr] = 0.0;
                    }
                }
                if (!(m1 > 0.0 && m2 > 0.0)) continue;
                double nn = 0.0;
#pragma unroll
                for (int d = 0; d < DIM; d++) nn += n1[d] * n1[d];
                const double rn = 1.0 / sqrt(nn);
                double nrm[DIM], mv1[DIM], mv2[DIM], vcm[DIM];
#pragma unroll
                for (int d = 0; d < DIM; d++) {
                    nrm[d] = n1[d] * rn;
                    mv1[d] = s.contact_spec == 1 ? p1[d] + f1[d] * dt : p1[d];
                    mv2[d] = s.contact_spec == 1 ? p2[d] + f2[d] * dt : p2[d];
                }
                const double r12 = 1.0 / (m1 + m2);
                const double r1 = 1.0 / m1;
                double chk = 0.0, vn = 0.0, mn = 0.0;
#pragma unroll
                for (int d = 0; d < DIM; d++) {
                    vcm[d] = (mv1[d] + mv2[d]) * r12;
                    chk += (mv1[d] * r1 - vcm[d]) * nrm[d];
                }
                if (!(chk > 0.0)) continue;
#pragma unroll
                for (int d = 0; d < DIM; d++) {
                    vn += vcm[d] * nrm[d];
                    mn += mv1[d] * nrm[d];
                }
                const double fn = m1 / dt * (vn - mn / m1);
                const double am = m1 / dt;
                double s1[DIM], ft2 = 0.0;
#pragma unroll
                for (int d = 0; d < DIM; d++) {
                    s1[d] = (vcm[d] - mv1[d] * r1) * am - nrm[d] * fn;
                    ft2 += s1[d] * s1[d];
                }
                const double ft = sqrt(ft2);
                if (ft != 0.0) {
                    const double rft = 1.0 / ft;
#pragma unroll
                    for (int d = 0; d < DIM; d++) s1[d] *= rft;
                }
                const double fnc = fn < 0.0 ? fn : 0.0;
                const double fa = C.mu * fabs(fn), fb = fabs(ft);
                const double fsc = fb < fa ? fb : fa;
#pragma unroll
                for (int b = 0; b < NBM; b++) {
#pragma unroll
                    for (int d = 0; d < DIM; d++) {
                        const double fc = nrm[d] * fnc + s1[d] * fsc;
                        if (b == C.b1) f[b][d] = f[b][d] + fc;
                        if (b == C.b2) f[b][d] = f[b][d] - fc;
                    }
                }
                if (cr) {
#pragma unroll
                    for (int d = 0; d < DIM; d++) cr[(DIM + d) * S + addr] = nrm[d] * fnc + s1[d] * fsc;
                }
            }
            // the body's boundary, then moveGrid
#pragma unroll
            for (int b = 0; b < NBM; b++) {
                if (b >= nb) continue;
                const BodyDev &B = A.bodies[s.body0 + b];
                if (B.has_boundary == 1 && wall) {
                    // CartesianBox::applyRules (cartesian_box.cpp:73-85): every component on a face node
#pragma unroll
                    for (int d = 0; d < DIM; d++) p[b][d] = f[b][d] = 0.0;
                } else if (B.has_boundary >= 2 && wall) {
                    // CartesianSmoothBox::applyRules (cartesian_smooth_box.cpp:70-84): the component
                    // normal to each face the node lies on. CartesianFrictionalBox::applyRules
                    // (cartesian_frictional_box.cpp:119-143) also takes the closing momentum that
                    // component would carry as the normal impulse.
                    double dm[DIM];
#pragma unroll
                    for (int d = 0; d < DIM; d++) {
                        dm[d] = 0.0;
                        const bool lo = ijk[d] == 0, hi = ijk[d] == s.N[d];
                        if (lo || hi) {
                            const double t = p[b][d] + dt * f0[b][d];
                            if (lo) dm[d] = -(t < 0.0 ? t : 0.0);
                            else dm[d] = -(t > 0.0 ? t : 0.0);
                            p[b][d] = f[b][d] = 0.0;
                        }
                    }
                    if (B.has_boundary == 3) {
                        // Coulomb friction on the remaining motion, at most what stops it (145-150)
                        double n2 = 0.0, l2 = 0.0, mv[DIM];
#pragma unroll
                        for (int d = 0; d < DIM; d++) {
                            n2 += dm[d] * dm[d];
                            mv[d] = p[b][d] + dt * f[b][d];
                            l2 += mv[d] * mv[d];
                        }
                        const double fnorm = sqrt(n2) / dt;
                        const double len = sqrt(l2);
                        if (len > 0.0) {
                            const double a = len / dt, c = fnorm * B.bc_mu;
                            const double mag = a < c ? a : c;
                            const double r = 1.0 / len;
#pragma unroll
                            for (int d = 0; d < DIM; d++) f[b][d] -= (mag * mv[d]) * r;
                        }
                    }
                }
                const long long gi = B.grid_base + addr;
                const bool has = m[b] > 0.0;
                const double rm = has ? 1.0 / m[b] : 0.0;
                if (A.rec) A.rec[gi] = m[b];
#pragma unroll
                for (int d = 0; d < DIM; d++) {
                    const double pn = p[b][d] + f[b][d] * dt;
                    const GO vo = (GO)(has ? pn * rm : 0.0), dvo = (GO)(has ? (f[b][d] * dt) * rm : 0.0);
                    outp[(G::OV + d) * A.nt + gi] = vo;
                    outp[(G::ODV + d) * A.nt + gi] = dvo;
                    if (!isfinite(vo) || !isfinite(dvo)) A.nonfin[par] = 1u;
                    if (A.rec) {
                        A.rec[(1 + d) * A.nt + gi] = pn;
                        A.rec[(1 + DIM + d) * A.nt + gi] = f[b][d];
                    }
                }
            }
        }
        if (lane == 0) A.nflag[gblk] = 0u;
    }
    if (A.stats) {
        // one atomic per wavefront
        unsigned long long tot = active_nodes;
        for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o);
        if (lane == 0 && tot) atomicAdd(&A.stats[0], tot);
    }
}

// ---------------------------------------------------------------------------------------
// Node fields of the last grid update in the reference's node order (CartesianLinear:
// n = i + (N0+1)(j + (N1+1) k)), from the node blocks that update processed (list `lst`,
// `cnt` entries, global block ids); nodes outside them keep what the caller filled in.
// contact < 0: body planes into out, [m, mxt DIM, f DIM] (with_rec) then [xt DIM, dxt DIM]
// from the grid output, nn per plane, body storage at grid_base. contact >= 0: [normal DIM,
// force DIM] of that contact slot's record. One wavefront per list entry.
template <int DIM, class GO>
__global__ __launch_bounds__(256) void k_nodes_out(KArgs A, int si, long long grid_base, const int *lst, int cnt,
                                                    int with_rec, int contact, double *out, long long nn) {
    const GO *outp = nullptr;
    if constexpr (sizeof(GO) == 8) outp = reinterpret_cast<const GO *>(A.out);
    else outp = reinterpret_cast<const GO *>(A.out32);
    using G = GF<DIM>;
    constexpr int NB = Cfg<DIM>::NB, L = Cfg<DIM>::LNB;
    const SimDev &s = A.sims[si];
    const int lane = threadIdx.x & 63;
    const int bi = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (bi >= cnt) return;
    const long long lb = lst[bi] - s.nblk_off;
    if (lb < 0 || lb >= s.nblk) return;  // another sim's block
    int ijk[3] = {0, 0, 0};
    long long t = lb;
    const int bx = (int)(t % s.NBn[0]);
    t /= s.NBn[0];
    ijk[0] = (bx << L) + (lane & (NB - 1)) - NSHIFT;
    if (DIM == 3) {
        ijk[1] = ((int)(t % s.NBn[1]) << L) + ((lane >> L) & (NB - 1)) - NSHIFT;
        ijk[2] = ((int)(t / s.NBn[1]) << L) + (lane >> (2 * L)) - NSHIFT;
    } else {
        ijk[1] = ((int)t << L) + (lane >> L) - NSHIFT;
    }
    bool valid = true;
#pragma unroll
    for (int d = 0; d < DIM; d++) valid = valid && ijk[d] >= 0 && ijk[d] <= s.N[d];
    if (!valid) return;
    long long nid = ijk[DIM - 1];
#pragma unroll
    for (int d = DIM - 2; d >= 0; d--) nid = nid * (s.N[d] + 1) + ijk[d];
    const long long addr = lb * 64 + lane;
    if (contact >= 0) {
        const double *cr = A.crec + A.contacts[contact].rec;
        const long long S = s.nblk * 64;
#pragma unroll
        for (int c = 0; c < 2 * DIM; c++) out[c * nn + nid] = cr[c * S + addr];
        return;
    }
    const long long gi = grid_base + addr;
    int o = 0;
    if (with_rec) {
#pragma unroll
        for (int c = 0; c < 1 + 2 * DIM; c++) out[c * nn + nid] = A.rec[c * A.nt + gi];
        o = 1 + 2 * DIM;
    }
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        out[(o + d) * nn + nid] = (double)outp[(G::OV + d) * A.nt + gi];
        out[(o + DIM + d) * nn + nid] = (double)outp[(G::ODV + d) * A.nt + gi];
    }
}

// ---------------------------------------------------------------------------------------
// Downloads and uploads of one body. k_export: the slots whose original index lies in
// [lo, lo + n) write their state as the C-ABI hands it out at index - lo of planes of n
// doubles, x[DIM] v[DIM] T[6] V, and their activity flag into the ints after those planes.
// k_import: staged points i < n (planes of n doubles: x[DIM] v[DIM] T[NTS] V m extent) into
// slots [0, n) of the particle planes `dst` (np per plane) in the form's representation.
// (The simulation's descriptor comes by value from the host: mpm_gpu_load_sim updates the
// device copy only with the next step.)
template <int DIM, class PT>
__global__ __launch_bounds__(256) void k_export(KArgs A, long long lo, long long n, SimDev s, double *out) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= A.np) return;
    const long long o = (long long)A.pid[p] - lo;
    if (o < 0 || o >= n) return;
    double x[DIM], v[DIM], T[6], V;
    PT::template export_state<DIM>(A, s, p, x, v, T, V);
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        out[d * n + o] = x[d];
        out[(DIM + d) * n + o] = v[d];
    }
#pragma unroll
    for (int k = 0; k < 6; k++) out[(2 * DIM + k) * n + o] = T[k];
    out[(2 * DIM + 6) * n + o] = V;
    reinterpret_cast<int *>(out + (2 * DIM + 7) * n)[o] = A.act[p];
}

template <int DIM, class PT, class E>
__global__ __launch_bounds__(256) void k_import(const double *src, long long n, SimDev s, E *dst, long long np) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    PT::template import_state<DIM>(src, n, i, s, dst, np);
}

// ---------------------------------------------------------------------------------------
// connectivity of the current state (DefaultPoints::generateMap's element lookups), for the
// points whose original index lies in [o_lo, o_lo + o_n); point o goes to entry o - o_lo
template <int DIM, class PT>
__global__ void k_connectivity(KArgs A, int *elem_out, uint8_t *code_out, long long o_lo, long long o_n) {
    constexpr int NC = 1 << DIM;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= A.np) return;
    const int oa = A.pid[p];
    if (oa < 0) return;
    const long long o = (long long)oa - o_lo;
    if (o < 0 || o >= o_n) return;
    int bi = 0;
    {
        int lo = 0, hi = A.nbody - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (A.bodies[mid].p_off <= oa) lo = mid;
            else hi = mid - 1;
        }
        bi = lo;
    }
    const SimDev &s = A.sims[A.bodies[bi].sim];
    for (int c = 0; c < NC; c++) elem_out[(long long)o * NC + c] = -1;
    double x[DIM], ext;
    PT::template position<DIM>(A, s, p, x, ext);
    if (!A.act[p] || !in_domain<DIM>(s, x)) {
        code_out[o] = 2;
        return;
    }
    bool cpdi = s.cpdi_spec == 1;
    if (cpdi) {
        for (int c = 0; c < NC; c++) {
            double xc[DIM];
#pragma unroll
            for (int d = 0; d < DIM; d++) xc[d] = x[d] + 0.5 * ext * (((c >> d) & 1) ? 1.0 : -1.0);
            if (!in_domain<DIM>(s, xc)) {
                cpdi = false;
                break;
            }
        }
    }
    if (cpdi) {
        code_out[o] = 1;
        for (int c = 0; c < NC; c++) {
            double xc[DIM];
#pragma unroll
            for (int d = 0; d < DIM; d++) xc[d] = x[d] + 0.5 * ext * (((c >> d) & 1) ? 1.0 : -1.0);
            elem_out[(long long)o * NC + c] = (int)which_element<DIM>(s, xc);
        }
    } else {
        code_out[o] = 0;
        elem_out[(long long)o * NC] = (int)which_element<DIM>(s, x);
    }
}

}  // namespace mpm
