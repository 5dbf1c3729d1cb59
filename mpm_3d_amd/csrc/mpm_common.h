// mpm_common.h -- device-side descriptors, layouts and shape-function evaluation shared by
// the MI355X explicit MPM kernels (gfx950, wave64).
//
// Layouts (all structure-of-arrays, one plane per component):
//   particles  P[f * np + slot]      f = x[DIM], v[DIM], T[NTS], V, m, extent (FP64 form)
//              Pw[f * np + slot]     f = cell, q[DIM], v[DIM], T[NTS], lnJ, V0, m, extent
//                                    (precision form: 4-byte planes; x = h (cell + q 2^-32)
//                                    with q a 32-bit fixed-point offset inside the cell,
//                                    V = V0 exp(lnJ), the rest FP32)
//              slots are kept sorted by particle block (see mpm_sort.h) and carry the
//              point's original global index in pid[slot]
//   grid       acc[f * nt + node]    accumulators m, p[DIM], f[DIM], n[DIM] (contact normal)
//              out[f * nt + node]    v[DIM] (velocity after the update), dv[DIM]; FP64,
//                                    or FP32 (out32) in the precision form
//              node = body.grid_base + 64 * node_block + local; a node block is 4x4x4 (3D)
//              or 8x8 (2D) nodes so one wavefront owns one block.
//   stress     NTS = 6 in 3D (xx, yy, zz, xy, xz, yz), 4 in 2D (xx, yy, zz, xy)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpm {

constexpr int MAXB = 4;  // bodies per simulation handled by the grid kernel

// shape-function evaluation modes for one point
constexpr int MODE_LEFT = -1;  // outside [0, L]: deactivated (default_points.cpp:217-219)
constexpr int MODE_NONE = 0;   // active but no element (a corner/point exactly at x = L)
constexpr int MODE_SEP = 1;    // GIMP/CPDI, separable per-dimension weights (<= 3 nodes/dim)
constexpr int MODE_TUP = 2;    // per-tuple evaluation: CPDI off, wide extents, index wrap

struct SimDev {
    int dim;
    int N[3];
    double L[3];
    double h[3];
    double invh[3];
    double NoverL[3];    // N / L, for the fast element search
    double ftol[3];      // distance to an integer below which the exact search is used
    int NBn[3];          // node blocks per dimension
    int PNB[3];          // particle blocks per dimension
    long long nblk_off;  // first global node-block id of this sim
    long long nblk;      // node blocks of this sim
    long long element_count;
    int body0, nbody, contact0, ncontact;
    double g[3];
    double dt;
    int cpdi_spec, contact_spec;
    float hf[3], invhf[3];  // h and 1/h rounded to FP32 (precision form)
};

struct BodyDev {
    int sim;
    int local;           // index within the sim
    long long p_off, n;  // original global index range of the body's points
    long long grid_base; // first node of the body's grid fields
    long long pblk_off;  // first particle-block key of the body
    long long npblk;     // particle blocks of the body
    int has_material;    // 0 none, 1 IsotropicLinearElasticity, 2 Sand_SachithLocal
    int has_boundary;    // 0 none, 1 box (all components), 2 smooth box (normal), 3 frictional box
    int is_b1, sand_dp;  // sand_dp: USE_DENSITY_FOR_PRESSURE
    double twoG, lambda;
    double bc_mu;        // Coulomb coefficient of the frictional box
    double G, K;         // shear and bulk moduli
    double sand[6];      // grain diameter, mu_s, grains_rho, rho_critical, mu_2, I_0
};

struct ContactDev {
    int b1, b2;     // sim-local body indices
    double mu;
    long long rec;  // first element of the contact's record planes in KArgs::crec
};

struct PBlk {  // a run of sorted particles sharing one particle block
    int key, start, count, body;
};

template <int DIM> struct Cfg;
// LDS patch layout: node (x, y, z) of the patch at x + P*y + ZS*z. In 3D the z-plane
// stride is padded from 64 to 68 doubles: the 64 lanes of a wavefront sit in a 4x4x4 group
// of dual cells, and with the padding their 8-byte accesses fall on each of the 32
// double-wide bank slots exactly twice (x + 8y + 4z mod 32) instead of four times.
template <> struct Cfg<3> {
    static constexpr int NB = 4;   // node-block edge (nodes)
    static constexpr int LNB = 2;  // log2(NB)
    static constexpr int PB = 4;   // particle-block edge (dual cells)
    static constexpr int P = 8;    // LDS patch edge (nodes): PB + 2 window + 2 drift
    static constexpr int ZS = 68;  // z-plane stride of the LDS patch
    static constexpr int PATCH = 512;   // nodes in the patch
    static constexpr int PSTR = 544;    // doubles per field plane of the patch (8 * ZS)
    static constexpr int PBLKS = 8;     // node blocks covering the patch (2 x 2 x 2)
    static constexpr int NTS = 6;
};
template <> struct Cfg<2> {
    static constexpr int NB = 8;
    static constexpr int LNB = 3;
    static constexpr int PB = 8;
    static constexpr int P = 12;
    static constexpr int ZS = 0;
    static constexpr int PATCH = 144;
    static constexpr int PSTR = 144;
    static constexpr int PBLKS = 4;     // node blocks covering the patch (2 x 2, the second ones in part)
    static constexpr int NTS = 4;
};

template <int DIM> struct PF {  // particle field planes (FP64 form)
    static constexpr int X = 0, V = DIM, T = 2 * DIM, NTS = Cfg<DIM>::NTS;
    static constexpr int VOL = 2 * DIM + NTS, MASS = VOL + 1, EXT = VOL + 2;
    static constexpr int GP = VOL + 3, GPD = VOL + 4;  // Sand_SachithLocal gammap, gammap_dot
    static constexpr int N = VOL + 5;
};

// Precision-form particle planes: 4-byte words, plane 0 the packed cell index.
template <int DIM> struct PW {
    static constexpr int CELL = 0, Q = 1, V = 1 + DIM, T = 1 + 2 * DIM, NTS = Cfg<DIM>::NTS;
    static constexpr int LNJ = 1 + 2 * DIM + NTS, V0 = LNJ + 1, MASS = LNJ + 2, EXT = LNJ + 3;
    static constexpr int GP = LNJ + 4, GPD = LNJ + 5;  // Sand_SachithLocal gammap, gammap_dot
    static constexpr int N = LNJ + 6;
};

// Packed cell index of the precision form: (i + 1) | (j + 1) << B0 | (k + 1) << (B0 + B1),
// so cells -1..2^B-2 fit (a point may sit one cell outside the grid for the step between
// leaving the domain and being deactivated).
template <int DIM> struct CellBits;
template <> struct CellBits<3> { static constexpr int B[3] = {11, 11, 10}; };
template <> struct CellBits<2> { static constexpr int B[3] = {16, 16, 0}; };

template <int DIM>
__host__ __device__ __forceinline__ uint32_t pack_cell(const int *c) {
    uint32_t w = 0;
    int sh = 0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        w |= (uint32_t)(c[d] + 1) << sh;
        sh += CellBits<DIM>::B[d];
    }
    return w;
}

template <int DIM>
__host__ __device__ __forceinline__ void unpack_cell(uint32_t w, int *c) {
    int sh = 0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        c[d] = (int)((w >> sh) & ((1u << CellBits<DIM>::B[d]) - 1u)) - 1;
        sh += CellBits<DIM>::B[d];
    }
}

template <int DIM> struct GF {  // grid field planes
    static constexpr int M = 0, P = 1, F = 1 + DIM, NRM = 1 + 2 * DIM, NACC = 1 + 3 * DIM;
    static constexpr int OV = 0, ODV = DIM, NOUT = 2 * DIM;
};

struct KArgs {
    const SimDev *sims;
    int nsim;
    const BodyDev *bodies;
    int nbody;
    const ContactDev *contacts;
    double *P;
    long long np;
    int8_t *act;
    int *pid;
    uint32_t *Pw;  // precision-form particle planes
    double *acc;
    double *out;
    float *out32;  // precision-form grid output
    long long nt;
    unsigned *nflag;
    int *nlist;
    int *ncount;
    long long nblk_total;
    const PBlk *plist;
    const int *npblk;
    int par;
    double *Lout;  // 9 planes x np, slot order (optional)
    // node record of the grid update (optional): m, m v after moveGrid, f after contact and
    // boundary, 1 + 2 DIM planes x nt; per contact slot: normal and force, 2 DIM planes over
    // its sim's node storage, at ContactDev::rec
    double *rec;
    double *crec;
    unsigned long long *stats;  // [0] active nodes of the last grid update
    // points the split particle passes leave to their side kernels (window not inside the
    // block's LDS patch, or the per-tuple form): slot list (np) and counters
    // [0] P2G count, [1] G2P count, [2] [3] their side kernels' finished workgroups
    unsigned *rare;
    unsigned *rare_n;
    // block counter of the particle passes ([0] next block, [1] finished workgroups)
    unsigned *work;
    // per list parity: the grid update wrote a non-finite v or dv (set by k_grid, cleared by
    // the next update)
    unsigned *nonfin;
};

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// body of point index o (bodies hold consecutive index ranges starting at p_off)
__device__ __forceinline__ int body_of_index(const KArgs &A, long long o) {
    int lo = 0, hi = A.nbody - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (A.bodies[mid].p_off <= o) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Lane of the P2G lane order for a point of in-block dual cell `dual` (KB = 2 or 3 bits per
// dimension, 64 keys). Order 2 is also the order of the dual cells inside a (block, octant)
// bucket of the cell-ordered sort.
template <int DIM, int ORDER>
__device__ __forceinline__ int lane_key(const int *dual) {
    constexpr int KB = Cfg<DIM>::PB == 4 ? 2 : 3;
    static_assert(KB * DIM == 6, "one key per lane of a wavefront");
    constexpr int m = (1 << KB) - 1;
    const int a0 = dual[0] & m, a1 = dual[1] & m, a2 = DIM == 3 ? dual[DIM - 1] & m : 0;
    if (DIM == 3) return ORDER == 1 ? a0 | a1 << KB | a2 << (2 * KB) : a0 | a2 << KB | a1 << (2 * KB);
    return ORDER == 1 ? a0 | a1 << KB : a1 | a0 << KB;
}

// sim-local node (i,j,k), -1 <= i <= N+1, -> offset inside the sim's node-block storage.
// Storage is shifted by two nodes: the zero-weight slots of a 3-node window that hang
// over the domain edge (nodes -1 and N+1) have a home (those padding nodes stay zero),
// and the LDS patch of particle block pb, nodes pb*PB-2 .. pb*PB+P-3, starts on a
// node-block boundary, so staging and flushing it touch whole 64-node blocks.
constexpr int NSHIFT = 2;
template <int DIM>
__host__ __device__ __forceinline__ long long node_addr(const SimDev &s, const int *ijk) {
    constexpr int NB = Cfg<DIM>::NB, L = Cfg<DIM>::LNB;
    const int i0 = ijk[0] + NSHIFT;
    long long blk = i0 >> L;
    int loc = i0 & (NB - 1);
    if (DIM >= 2) {
        const int i1 = ijk[1] + NSHIFT;
        blk += (long long)(i1 >> L) * s.NBn[0];
        loc += (i1 & (NB - 1)) << L;
    }
    if (DIM == 3) {
        const int i2 = ijk[DIM - 1] + NSHIFT;
        blk += (long long)(i2 >> L) * s.NBn[0] * s.NBn[1];
        loc += (i2 & (NB - 1)) << (2 * L);
    }
    return blk * 64 + loc;
}

// Node i of the LDS patch counted by whole node blocks: the patch of origin org (a node-
// block boundary in storage) is covered by Cfg::PBLKS blocks of 64 nodes; i = 64 * block
// + lane. Sets the node's sim-local ijk, inp (the node lies in the patch) and in (it lies in
// the domain, 0 <= ijk <= N); returns its storage offset (meaningful when in).
template <int DIM>
__device__ __forceinline__ long long patch_node(const SimDev &s, const int *org, int i, int *nijk, bool &inp,
                                                bool &in) {
    constexpr int NB = Cfg<DIM>::NB, L = Cfg<DIM>::LNB, P = Cfg<DIM>::P;
    constexpr int BPD = (P + NB - 1) / NB;  // node blocks per patch edge
    const int b = i >> 6, lane = i & 63;
    int bc[3] = {b % BPD, (b / BPD) % BPD, b / (BPD * BPD)};
    int pl[3] = {bc[0] * NB + (lane & (NB - 1)), bc[1] * NB + ((lane >> L) & (NB - 1)), bc[2] * NB + (lane >> (2 * L))};
    inp = true;
    in = true;
    long long blk = 0, stride = 1;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        nijk[d] = org[d] + pl[d];
        inp = inp && pl[d] < P;
        in = in && nijk[d] >= 0 && nijk[d] <= s.N[d];
        blk += (long long)((org[d] + NSHIFT) / NB + bc[d]) * stride;
        stride *= s.NBn[d];
    }
    in = in && inp;
    return blk * 64 + lane;
}

// CartesianLinear::inDomain (cartesian.cpp:245-253): 0 <= x <= L in every dimension
template <int DIM>
__host__ __device__ __forceinline__ bool in_domain(const SimDev &s, const double *x) {
    bool in = true;
#pragma unroll
    for (int d = 0; d < DIM; d++) in = in && (x[d] <= s.L[d] && x[d] >= 0.0);
    return in;
}

// floor((x / L) * N) exactly as the reference rounds it (cartesian.cpp:226-236), for
// 0 <= x <= L. q = x * (N/L) differs from fl(fl(x/L) * N) by at most 4 ulp(N); when q is
// farther than ftol = N * 2^-48 from an integer both floors agree, otherwise the
// reference's division is evaluated.
__host__ __device__ __forceinline__ int elem_floor(const SimDev &s, int d, double x) {
    const double q = x * s.NoverL[d];
    const double f = floor(q);
    const double r = q - f;
    if (r > s.ftol[d] && r < 1.0 - s.ftol[d]) return (int)f;
    return (int)floor(x / s.L[d] * (double)s.N[d]);
}

// CartesianLinear::cartesianWhichElement (cartesian.cpp:219-240), including the
// double-precision accumulation of the element id. -1 unless 0 <= x < L.
template <int DIM>
__host__ __device__ __forceinline__ long long which_element(const SimDev &s, const double *x) {
#pragma unroll
    for (int d = 0; d < DIM; d++)
        if (!(x[d] < s.L[d] && x[d] >= 0.0)) return -1;
    long long e = elem_floor(s, 0, x[0]);
    if (DIM >= 2) e = (long long)((double)e + (double)elem_floor(s, 1, x[1]) * (double)s.N[0]);
    if (DIM == 3)
        e = (long long)((double)e + (double)elem_floor(s, DIM - 1, x[DIM - 1]) * (double)s.N[1] * (double)s.N[0]);
    return e;
}

// Separable GIMP/CPDI weights of one point (default_points.cpp:264-324 summed per node):
// per dimension the window is the 3 nodes base, base+1, base+2 with
//   sw = hat(x - a) + hat(x + a)      and      dw = hat(x + a) - hat(x - a)
// (a = 0.5*extent); S = prod sw / 2^D and G_d = dw_d prod_{e!=d} sw_e * gscale with
// gscale = 1 / (2^D * 0.5 * extent). Each corner's hat values are those of the element
// that CartesianLinear::whichElement assigns to the corner, so a corner exactly at x = L
// contributes nothing, as in the reference. base is the point's dual cell
// floor(x/h - 1/2), moved into [ep-1, em] so both corner elements lie in the window;
// slots no corner reaches carry exact zeros. The fixed 3^D stencil keeps the node loops
// free of per-lane branches. Points whose corners span more than 2 elements in some
// dimension (extent > h) take the per-tuple path.
template <int DIM, class R = double> struct Win {
    int base[DIM];
    R sw[DIM][3];
    R dw[DIM][3];
    // per dimension, bit a set when slot a is a node of a corner's element (of the point's
    // own element with CPDI off): the slots the reference's tuples cover. The other slot
    // holds exact zeros and must not be multiplied into anything, so that a non-finite node
    // or point value reaches exactly what the reference's tuples reach (0 * NaN is NaN).
    unsigned char rch[DIM];
};

// cpdi_out tells the tuple path whether the point uses its corners (CPDI on) or not.
template <int DIM>
__host__ __device__ __forceinline__ int point_window(const SimDev &s, const double *x, double ext, Win<DIM> &w,
                                            bool &cpdi_out) {
    cpdi_out = false;
    if (!in_domain<DIM>(s, x)) return MODE_LEFT;
    // CPDI off (cpdi_spec 0, or a corner outside the domain: default_points.cpp:220-263) is
    // the linear basis over the point's own element: the same window with both "corners" at
    // the point (sw = 2 (1 - |r|) on the element's two nodes) and the gradient slots set to
    // -/+ extent / h, which the common factors turn into -/+ prod_{e != d} (1 - |r_e|) / h_d.
    bool off = s.cpdi_spec != 1;
    const double a = off ? 0.0 : 0.5 * ext;
    double xm[DIM], xp[DIM];
    bool cin = true;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        xm[d] = x[d] + a * (-1.0);  // x + (0.5*extent)*A(c,d)
        xp[d] = x[d] + a * (1.0);
        cin = cin && (xm[d] <= s.L[d] && xm[d] >= 0.0) && (xp[d] <= s.L[d] && xp[d] >= 0.0);
    }
    if (!cin) {  // a corner leaves the domain: CPDI off
        off = true;
#pragma unroll
        for (int d = 0; d < DIM; d++) xm[d] = xp[d] = x[d];
    }
    if (off && !(ext > 0.0)) return MODE_TUP;
    cpdi_out = !off;
    int mode = MODE_SEP;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        const bool vm = xm[d] < s.L[d], vp = xp[d] < s.L[d];
        if (!vm && !vp) mode = MODE_NONE;
        const int em = vm ? elem_floor(s, d, xm[d]) : 0;
        const int ep = vp ? elem_floor(s, d, xp[d]) : 0;
        if ((vm && em >= s.N[d]) || (vp && ep >= s.N[d])) mode = mode == MODE_NONE ? MODE_NONE : MODE_TUP;
        const int lo = vm ? em : ep;
        const int hi = vp ? ep : em;
        if (hi - lo > 1) mode = mode == MODE_NONE ? MODE_NONE : MODE_TUP;
        int b = (int)floor(x[d] * s.invh[d] - 0.5);
        b = b < hi - 1 ? hi - 1 : (b > lo ? lo : b);
        w.base[d] = b;
        double s0 = 0, s1 = 0, s2 = 0, d0 = 0, d1 = 0, d2 = 0;
        const double hd = s.h[d], ih = s.invh[d];
        if (vm) {
            const double h0 = 1.0 - fabs((xm[d] - hd * em) * ih);
            const double h1 = 1.0 - fabs((xm[d] - hd * (em + 1)) * ih);
            const bool k0 = em == b;  // slot of the corner's element: 0 or 1
            s0 += k0 ? h0 : 0.0;
            s1 += k0 ? h1 : h0;
            s2 += k0 ? 0.0 : h1;
            d0 -= k0 ? h0 : 0.0;
            d1 -= k0 ? h1 : h0;
            d2 -= k0 ? 0.0 : h1;
        }
        if (vp) {
            const double h0 = 1.0 - fabs((xp[d] - hd * ep) * ih);
            const double h1 = 1.0 - fabs((xp[d] - hd * (ep + 1)) * ih);
            const bool k0 = ep == b;
            s0 += k0 ? h0 : 0.0;
            s1 += k0 ? h1 : h0;
            s2 += k0 ? 0.0 : h1;
            d0 += k0 ? h0 : 0.0;
            d1 += k0 ? h1 : h0;
            d2 += k0 ? 0.0 : h1;
        }
        if (off) {
            // on a node line the reference's CPDI-off gradient is 0/0: per-tuple path
            if (vm && x[d] == hd * (double)em) mode = mode == MODE_NONE ? MODE_NONE : MODE_TUP;
            const double g = ext / hd;
            const bool k0 = em == b;
            d0 = k0 ? -g : 0.0;
            d1 = k0 ? g : -g;
            d2 = k0 ? 0.0 : g;
        }
        w.sw[d][0] = s0; w.sw[d][1] = s1; w.sw[d][2] = s2;
        w.dw[d][0] = d0; w.dw[d][1] = d1; w.dw[d][2] = d2;
        w.rch[d] = (unsigned char)((vm ? (em == b ? 3 : 6) : 0) | (vp ? (ep == b ? 3 : 6) : 0));
    }
    return mode;
}

// whether window slot (a0, a1[, a2]) is one the reference's tuples cover (Win::rch)
template <int DIM, class R>
__host__ __device__ __forceinline__ bool slot_reached(const Win<DIM, R> &w, const int *off) {
    unsigned m = 1;
#pragma unroll
    for (int d = 0; d < DIM; d++) m &= (unsigned)w.rch[d] >> off[d];
    return m != 0;
}

// Per-tuple evaluation, exactly the reference's COO construction for one point
// (default_points.cpp:238-324 with cartesian.cpp:280-357): for every corner (CPDI on)
// or the point itself (CPDI off), the 2^D nodes of its element with S and G.
// fn(node_ijk, S, G[DIM]) is called per tuple. Used for CPDI-off points and the rare
// shapes the separable form does not cover.
template <int DIM, class Fn>
__host__ __device__ void for_each_tuple(const SimDev &s, const double *x, double ext, bool cpdi, Fn &&fn) {
    constexpr int NC = 1 << DIM;
    const int ncorner = cpdi ? NC : 1;
    for (int c = 0; c < ncorner; c++) {
        double xc[DIM];
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            const double A = ((c >> d) & 1) ? 1.0 : -1.0;
            xc[d] = cpdi ? x[d] + 0.5 * ext * A : x[d];
        }
        long long e = which_element<DIM>(s, xc);
        if (e < 0 || e >= s.element_count) continue;
        int eijk[DIM];
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            eijk[d] = (int)(e % s.N[d]);
            e /= s.N[d];
        }
        for (int n = 0; n < NC; n++) {
            int nijk[DIM];
            double r[DIM];
            double tmp = 1.0;
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                nijk[d] = eijk[d] + ((n >> d) & 1);
                r[d] = (xc[d] - s.h[d] * (double)nijk[d]) / s.h[d];
                tmp *= (1.0 - fabs(r[d]));
            }
            double S, G[DIM];
            if (cpdi) {
                S = tmp / (double)NC;
#pragma unroll
                for (int d = 0; d < DIM; d++) {
                    const double A = ((c >> d) & 1) ? 1.0 : -1.0;
                    G[d] = tmp / ((double)NC * 0.5 * ext) * A;
                }
            } else {
                S = tmp;
#pragma unroll
                for (int d = 0; d < DIM; d++) {
                    double gi = -tmp;
                    gi = gi / (1.0 - fabs(r[d]));
                    gi = gi * r[d];
                    gi = gi / fabs(r[d]);
                    gi = gi / s.h[d];
                    if (fabs(r[d]) >= 1.0) gi = 0.0;
                    G[d] = gi;
                }
            }
            fn(nijk, S, G);
        }
    }
}

// IsotropicLinearElasticity::calculateStress for one point (isolin.cpp:50-78), in the
// reference's operation order: T += dt * (((2G D + lambda trD I) + W T) - T W).
// T is symmetric (stored xx, yy, zz, xy, xz, yz); L is the full 3x3 velocity gradient.
template <class R>
__device__ __forceinline__ void isolin_update(double twoG, double lambda, double dt, const R L[3][3], R T6[6]) {
    R Tm[3][3] = {{T6[0], T6[3], T6[4]}, {T6[3], T6[1], T6[5]}, {T6[4], T6[5], T6[2]}};
    R D[3][3], W[3][3];
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            D[r][c] = (L[r][c] + L[c][r]) * (R)0.5;
            W[r][c] = (L[r][c] - L[c][r]) * (R)0.5;
        }
    const R trD = D[0][0] + D[1][1] + D[2][2];
    const R tg = (R)twoG;
    const R lt = (R)lambda * trD;
    const int rr[6] = {0, 1, 2, 0, 0, 1};
    const int cc[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const int r = rr[k], c = cc[k];
        const R WT = W[r][0] * Tm[0][c] + W[r][1] * Tm[1][c] + W[r][2] * Tm[2][c];
        const R TW = Tm[r][0] * W[0][c] + Tm[r][1] * W[1][c] + Tm[r][2] * W[2][c];
        const R I = (r == c) ? (R)1.0 : (R)0.0;
        const R tmp = ((D[r][c] * tg + I * lt) + WT) - TW;
        T6[k] = T6[k] + tmp * (R)dt;
    }
}

// Sand_SachithLocal::calculateStress on one active point, SPEC = UPDATE (sand_sachithlocal.cpp:
// 144-232): the Jaumann trial stress of isolin_update, then the local mu(I) return map; m / V
// is the density after this step's volume update. Returns the plastic shear rate nup_tau.
template <class R>
__device__ __forceinline__ R sand_update(const BodyDev &B, double dt, R m, R V, const R L[3][3], R T6[6]) {
    const R d = (R)B.sand[0], MU_S = (R)B.sand[1], GRAINS_RHO = (R)B.sand[2], RHO_CRITICAL = (R)B.sand[3];
    const R MU_2 = (R)B.sand[4], I_0 = (R)B.sand[5];
    const R G = (R)B.G, dtr = (R)dt;
    R s[6] = {T6[0], T6[1], T6[2], T6[3], T6[4], T6[5]};
    isolin_update<R>(B.twoG, B.lambda, dt, L, s);  // trial stress
    R p_tr = -(s[0] + s[1] + s[2]) / (R)3.0;
    R t0[6];
#pragma unroll
    for (int k = 0; k < 6; k++) t0[k] = s[k] + (k < 3 ? p_tr : (R)0);
    // |t0| over the 9 components in row order: xx xy xz yx yy yz zx zy zz
    R nn = t0[0] * t0[0];
    nn += t0[3] * t0[3];
    nn += t0[4] * t0[4];
    nn += t0[3] * t0[3];
    nn += t0[1] * t0[1];
    nn += t0[5] * t0[5];
    nn += t0[4] * t0[4];
    nn += t0[5] * t0[5];
    nn += t0[2] * t0[2];
    const R tau_tr = sqrt(nn) / sqrt((R)2.0);
    if (B.sand_dp) p_tr = (R)B.K * ((m / V - RHO_CRITICAL) / RHO_CRITICAL);
    const bool density_flag = (m / V) < RHO_CRITICAL;
    R nup;
    if (density_flag || p_tr <= (R)0) {
        nup = tau_tr / (G * dtr);
#pragma unroll
        for (int k = 0; k < 6; k++) T6[k] = (R)0;
    } else {
        R tau_tau, scale;
        const R S0 = MU_S * p_tr;
        if (tau_tr <= S0) {
            tau_tau = tau_tr;
            scale = (R)1.0;
        } else {
            const R S2 = MU_2 * p_tr;
            const R alpha = G * I_0 * dtr * sqrt(p_tr / GRAINS_RHO) / d;
            const R b = -(S2 + tau_tr + alpha);
            const R c = S2 * tau_tr + S0 * alpha;
            // negative_root(1, b, c) (sand_sachithlocal.cpp:44-53)
            const R disc = sqrt(b * b - (R)4 * (R)1 * c);
            tau_tau = b > (R)0 ? (-b - disc) / ((R)2 * (R)1) : ((R)2 * c) / (-b + disc);
            scale = tau_tau / tau_tr;
        }
        nup = ((tau_tr - tau_tau) / G) / dtr;
#pragma unroll
        for (int k = 0; k < 6; k++) T6[k] = scale * t0[k] - (k < 3 ? p_tr : (R)0);
    }
    return nup;
}

}  // namespace mpm
