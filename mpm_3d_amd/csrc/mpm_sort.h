// mpm_sort.h -- counting sort of particle slots by particle block, on the device.
//
// A particle block is a PB^DIM group of dual cells (the cell of x/h - 1/2, so every point
// of a block touches nodes inside one (PB+2)^DIM window). Sorting keeps the particles of
// a block contiguous, which is what lets k_particles accumulate a whole block in one LDS
// patch. Inside a block the key orders points by the octant of their position within the
// dual cell first: the 64 consecutive points a wavefront takes then sit in ~64 different
// dual cells, so their LDS atomics hit different nodes (no same-address serialisation)
// while the loads stay coalesced. With the cell-ordered sort the points of each (block,
// octant) bucket are further ordered by dual cell, in the P2G lane order (x + 4z + 16y in
// 3D): a wavefront over a bucket with one point per dual cell then places every point on
// the lane of its cell, whose LDS accesses fall on distinct banks, without the pass
// reordering its lanes.
// The reference has no counterpart: it scatters straight from point order through COO
// tuples (src/algebra/mpm_sparse.cpp:30-44).
//
//   k_bin_count    key per slot (inactive points go to the last bucket), wavefront-
//                  aggregated atomic histogram, rank inside the bucket (cell-ordered: rank
//                  inside the bucket's dual cell, from a per-cell counter)
//   scan           exclusive prefix sum of the histogram (k_scan_*)
//   k_make_plist   one PBlk per non-empty block, in block order (k_block_flags, a scan);
//                  k_clear then clears the histogram
//   k_cell_offsets (cell-ordered) per non-empty bucket, the prefix of its 64 cell counters,
//                  in place
//   k_bin_scatter  moves every particle plane to its sorted slot
//   k_cell_clear   (cell-ordered) zeroes the listed blocks' cell counters for the next sort
#pragma once

#include "mpm_common.h"
#include "mpm_prec.h"

namespace mpm {

__device__ __forceinline__ int body_of_block(const KArgs &A, long long key) {
    int lo = 0, hi = A.nbody - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (A.bodies[mid].pblk_off <= key) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Keys per slot and ranks inside their bucket. Active keys (a particle block and octant,
// a few dozen points each) count with wavefront-aggregated atomics. The inactive bucket
// can hold most of the slots (unused capacity of a slot pool), so it is ranked without
// atomics: the rank inside the workgroup here, the workgroup's count in inact_count
// (scanned into per-workgroup offsets before the scatter). fcnt (cell-ordered sort, else
// null): 64 counters per (block, octant) key, all zero on entry; an active slot's rank is
// then its rank inside its dual cell of the bucket, and pcell its cell.
template <int DIM, class PT>
__global__ __launch_bounds__(256) void k_bin_count(KArgs A, unsigned *count, int *pkey, int *prank,
                                                   long long inactive_key, unsigned *inact_count, unsigned *fcnt,
                                                   uint8_t *pcell) {
    constexpr int PB = Cfg<DIM>::PB;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0) *const_cast<int *>(A.npblk) = 0;
    const bool live = p < A.np;
    long long key = inactive_key;
    int cell = 0;
    if (live && A.act[p] && A.pid[p] >= 0) {
        const int b = body_of_index(A, A.pid[p]);
        const BodyDev &B = A.bodies[b];
        const SimDev &s = A.sims[B.sim];
        int dual[DIM], oct = 0;
        if (PT::template dual_cell<DIM>(A, s, p, dual, oct)) {
            long long lk = 0, stride = 1;
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                lk += (long long)(dual[d] / PB) * stride;
                stride *= s.PNB[d];
            }
            key = (B.pblk_off + lk) * (1 << DIM) + oct;
            cell = lane_key<DIM, 2>(dual);
        } else {
            A.act[p] = 0;  // leaves the domain: deactivated (default_points.cpp:217-219)
        }
    }
    // inactive slots: rank inside the workgroup by ballot + LDS prefix
    __shared__ unsigned wave_inact[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const bool inact = live && key == inactive_key;
    const unsigned long long im = __ballot(inact);
    if (lane == 0) wave_inact[wv] = (unsigned)__popcll(im);
    __syncthreads();
    unsigned before = 0, total = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
        if (w < wv) before += wave_inact[w];
        total += wave_inact[w];
    }
    if (threadIdx.x == 0) inact_count[blockIdx.x] = total;
    // active keys: wavefront aggregation, one atomic per distinct key in the wavefront
    unsigned long long pending = __ballot(live && !inact);
    int rank = inact ? (int)(before + __popcll(im & ((1ull << lane) - 1ull))) : 0;
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const long long lkey = __shfl(key, leader);
        const unsigned long long same = __ballot(live && !inact && key == lkey) & pending;
        unsigned base = 0;
        if (lane == leader) base = atomicAdd(&count[lkey], (unsigned)__popcll(same));
        base = __shfl(base, leader);
        if ((same >> lane) & 1ull) rank = (int)base + __popcll(same & ((1ull << lane) - 1ull));
        pending &= ~same;
    }
    // cell-ordered: the rank inside the point's dual cell of its bucket instead
    if (fcnt && live && !inact) {
        rank = (int)atomicAdd(&fcnt[key * 64 + cell], 1u);
        pcell[p] = (uint8_t)cell;
    }
    if (live) {
        pkey[p] = (int)key;
        prank[p] = rank;
    }
}

// exclusive scan, 1024 threads x 4 elements per workgroup
__global__ __launch_bounds__(1024) void k_scan_block(const unsigned *in, unsigned *out, unsigned *bsum, long long n) {
    __shared__ unsigned sh[1024];
    const long long base = (long long)blockIdx.x * 4096 + threadIdx.x * 4;
    unsigned v[4], t = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        v[i] = (base + i < n) ? in[base + i] : 0u;
        t += v[i];
    }
    sh[threadIdx.x] = t;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const unsigned add = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0u;
        __syncthreads();
        sh[threadIdx.x] += add;
        __syncthreads();
    }
    unsigned run = sh[threadIdx.x] - t;  // exclusive prefix of this thread
#pragma unroll
    for (int i = 0; i < 4; i++) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
    if (threadIdx.x == 1023 && bsum) bsum[blockIdx.x] = sh[1023];
}

__global__ __launch_bounds__(1024) void k_scan_add(unsigned *out, const unsigned *bscan, long long n) {
    const long long i = (long long)blockIdx.x * 4096 + threadIdx.x * 4;
    const unsigned add = bscan[blockIdx.x];
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (i + k < n) out[i + k] += add;
}

// Block list, in block order: flag[b] = block b holds points (its 2^DIM buckets, keys
// b * 2^DIM + octant, are not all empty); flag[nblk] = 0 so that the scan's last entry is
// the number of listed blocks
__global__ __launch_bounds__(256) void k_block_flags(const unsigned *off, unsigned *flag, long long nblk, int noct) {
    const long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (b > nblk) return;
    flag[b] = b < nblk && off[(b + 1) * noct] > off[b * noct] ? 1u : 0u;
}

// one entry per non-empty particle block at its rank among them (pos: exclusive scan of
// the flags), so the particle passes take blocks in key order; k_clear then clears the
// histogram
__global__ __launch_bounds__(256) void k_make_plist(KArgs A, const unsigned *off, const unsigned *pos, PBlk *plist,
                                                    int *npblk, long long nblk, int noct) {
    const long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (b == 0) *npblk = (int)pos[nblk];
    if (b >= nblk || pos[b + 1] == pos[b]) return;
    const unsigned b0 = off[b * noct], b1 = off[(b + 1) * noct];
    PBlk e;
    e.key = (int)b;
    e.start = (int)b0;
    e.count = (int)(b1 - b0);
    e.body = body_of_block(A, b);
    plist[pos[b]] = e;
}

// Cell-ordered sort: for each (block, octant) bucket of the listed blocks (a workgroup per
// block, grid-stride over the list; one wavefront per octant, lane = dual cell in the lane
// order), the exclusive prefix of the bucket's 64 cell counters, written over them: it
// places the bucket's points cell after cell.
template <int DIM>
__global__ __launch_bounds__(64 << DIM) void k_cell_offsets(const PBlk *plist, const int *npblk, const unsigned *off,
                                                            unsigned *fcnt) {
    const int lane = threadIdx.x & 63;
    const int nb = *npblk;
    for (int e = blockIdx.x; e < nb; e += gridDim.x) {
        const long long k = (long long)plist[e].key * (1 << DIM) + (threadIdx.x >> 6);
        if (off[k + 1] == off[k]) continue;  // empty bucket (wavefront-uniform): its counters are zero
        const long long i = k * 64 + lane;
        const unsigned c = fcnt[i];
        unsigned incl = c;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const unsigned t = __shfl_up(incl, d);
            if (lane >= d) incl += t;
        }
        fcnt[i] = incl - c;
    }
}

// Cell-ordered sort, after the scatter: the listed blocks' cell counters (offsets by now)
// back to zero for the next sort; the other blocks' counters were never touched.
template <int DIM>
__global__ __launch_bounds__(64 << DIM) void k_cell_clear(const PBlk *plist, const int *npblk, unsigned *fcnt) {
    const int nb = *npblk;
    for (int e = blockIdx.x; e < nb; e += gridDim.x) fcnt[(long long)plist[e].key * (64 << DIM) + threadIdx.x] = 0u;
}

__global__ void k_clear_u64(unsigned long long *p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

__global__ __launch_bounds__(256) void k_clear(unsigned *count, long long n) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) count[k] = 0u;
}

template <class E, int NF>
__global__ __launch_bounds__(256) void k_bin_scatter(long long np, const E *P, E *P2, const int8_t *act,
                                                     int8_t *act2, const int *pid, int *pid2, const int *pkey,
                                                     const int *prank, const unsigned *off, int inactive_key,
                                                     const unsigned *inact_off, const unsigned *coff,
                                                     const uint8_t *pcell) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int key = pkey[p];
    long long q = (long long)off[key] + prank[p];
    if (key == inactive_key) q += inact_off[blockIdx.x];  // same 256-slot blocks as k_bin_count
    else if (coff) q += coff[(long long)key * 64 + pcell[p]];  // cell-ordered: the cell's start
    const int o = pid[p];
    if (o < 0) {  // unused capacity: nothing to move
        act2[q] = 0;
        pid2[q] = -1;
        return;
    }
#pragma unroll
    for (int f = 0; f < NF; f++) P2[f * np + q] = P[f * np + p];
    act2[q] = act[p];
    pid2[q] = o;
}

// the slots whose original index lies in [lo, lo + n): selected planes into planes of n
// elements at index - lo (one body's points for a download)
template <class E>
__global__ __launch_bounds__(256) void k_unsort_range(long long np, int nf, const E *src, long long src_stride,
                                                      E *dst, const int *pid, long long lo, long long n) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const long long o = (long long)pid[p] - lo;
    if (o < 0 || o >= n) return;
    for (int f = 0; f < nf; f++) dst[f * n + o] = src[f * src_stride + p];
}

__global__ __launch_bounds__(256) void k_fill_f64(double *a, long long n, double v) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// release the points whose original index lies in [lo, hi): inactive and marked as
// unused capacity (pid -1), so sorts and unsorts no longer move their planes
__global__ __launch_bounds__(256) void k_deactivate(long long np, int8_t *act, int *pid, long long lo,
                                                    long long hi) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int o = pid[p];
    if (o >= lo && o < hi) {
        act[p] = 0;
        pid[p] = -1;
    }
}

// Upload into free slots. k_free_count: unused slots (pid < 0) per 256-slot block; after
// an exclusive scan of those counts, k_place gives the j-th unused slot (in slot order)
// staged point j < n: its planes (staged at slot j of `stage`, same plane stride), its
// activity and original index p_off + j.
__device__ __forceinline__ unsigned long long free_ballot(long long np, const int *pid, bool &fr,
                                                          unsigned *wc) {
    const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
    fr = p < np && pid[p] < 0;
    const unsigned long long bal = __ballot(fr);
    if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = (unsigned)__popcll(bal);
    __syncthreads();
    return bal;
}

__global__ __launch_bounds__(256) void k_free_count(long long np, const int *pid, unsigned *count) {
    __shared__ unsigned wc[4];
    bool fr;
    free_ballot(np, pid, fr, wc);
    if (threadIdx.x == 0) count[blockIdx.x] = wc[0] + wc[1] + wc[2] + wc[3];
}

template <class E>
__global__ __launch_bounds__(256) void k_place(long long np, int nf, const E *stage, E *P, const int8_t *act_stage,
                                               int8_t *act, int *pid, const unsigned *free_off, long long n,
                                               long long p_off) {
    __shared__ unsigned wc[4];
    bool fr;
    const unsigned long long bal = free_ballot(np, pid, fr, wc);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    long long r = free_off[blockIdx.x];
    for (int i = 0; i < w; i++) r += wc[i];
    r += __popcll(bal & ((1ull << lane) - 1ull));
    if (!fr || r >= n) return;
    const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
    for (int f = 0; f < nf; f++) P[f * np + p] = stage[f * np + r];
    act[p] = act_stage[r];
    pid[p] = (int)(p_off + r);
}

__global__ __launch_bounds__(256) void k_count_active(long long np, const int8_t *act, unsigned long long *out) {
    __shared__ unsigned wsum[4];
    unsigned c = 0;
    for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < np; p += (long long)gridDim.x * blockDim.x)
        c += act[p] ? 1u : 0u;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int i = 0; i < (int)(blockDim.x >> 6); i++) t += wsum[i];
        if (t) atomicAdd(out, t);
    }
}

}  // namespace mpm
