// mpm_kernels.h -- the explicit USL step as HIP kernels for gfx950 (FP64 form).
//
// One step (explicit_usl.cpp:48-87) is three launches over resident HBM state:
//   k_p2g   map (default_points.cpp:196-330) + P2G (explicit_usl.cpp:112-190), and the
//           contact-normal scatter gradS*m (contact_huang.cpp:80-99): one workgroup per
//           particle block accumulates into an LDS node patch with FP64 LDS atomics, then
//           flushes the patch to the grid with FP64 global atomics and marks node blocks.
//   k_grid  per touched node: ContactHuang (contact_huang.cpp:104-197), CartesianBox
//           (cartesian_box.cpp:73-85), moveGrid (explicit_usl.cpp:233-264); one wavefront
//           per 64-node block; consumes and re-zeroes the accumulators.
//   k_g2p   movePoints, calculateStrainRate, updateDensity, gravity, updateStress
//           (explicit_usl.cpp:266-355, isolin.cpp:50-78).
#pragma once

#include "mpm_common.h"

namespace mpm {

__device__ __forceinline__ void mark_block(const KArgs &A, long long gblk) {
    if (atomicOr(&A.nflag[gblk], 1u) == 0u) {
        const int pos = atomicAdd(&A.ncount[A.par], 1);
        A.nlist[(long long)A.par * A.nblk_total + pos] = (int)gblk;
    }
}

// mark the node blocks covering nodes [lo, hi] (inclusive) of a sim; called by one thread
template <int DIM>
__device__ void mark_range(const KArgs &A, const SimDev &s, const int *lo, const int *hi) {
    constexpr int L = Cfg<DIM>::LNB;
    int b0[3] = {0, 0, 0}, b1[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        b0[d] = lo[d] >> L;
        b1[d] = hi[d] >> L;
    }
    for (int k = b0[2]; k <= b1[2]; k++)
        for (int j = b0[1]; j <= b1[1]; j++)
            for (int i = b0[0]; i <= b1[0]; i++) {
                long long lb = i + (long long)s.NBn[0] * (j + (long long)s.NBn[1] * k);
                mark_block(A, s.nblk_off + lb);
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

// the per-point P2G payload: m, m v, m g, V T
template <int DIM> struct P2GVals {
    double m;
    double mv[DIM];
    double mg[DIM];
    double VT[DIM][DIM];
};

template <int DIM>
__device__ __forceinline__ void p2g_node_values(const P2GVals<DIM> &q, double S, const double *G, double *out,
                                                 bool b1) {
    // out: m, p[DIM], f[DIM], n[DIM]
    out[0] = S * q.m;
#pragma unroll
    for (int d = 0; d < DIM; d++) out[1 + d] = S * q.mv[d];
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        double f = S * q.mg[d];
#pragma unroll
        for (int c = 0; c < DIM; c++) f = fma(-q.VT[d][c], G[c], f);
        out[1 + DIM + d] = f;
    }
#pragma unroll
    for (int d = 0; d < DIM; d++) out[1 + 2 * DIM + d] = b1 ? G[d] * q.m : 0.0;
}

template <int DIM>
__global__ __launch_bounds__(256) void k_p2g(KArgs A) {
    constexpr int P = Cfg<DIM>::P, PATCH = Cfg<DIM>::PATCH, NACC = GF<DIM>::NACC;
    using F = PF<DIM>;
    __shared__ double lacc[NACC * PATCH];
    __shared__ int bb[2 * DIM];
    const int tid = threadIdx.x;
    const int nent = *A.npblk;
    for (int e = blockIdx.x; e < nent; e += gridDim.x) {
        const PBlk ent = A.plist[e];
        const int body = uniform(ent.body);
        const int start = uniform(ent.start), count = uniform(ent.count);
        const BodyDev &B = A.bodies[body];
        const SimDev &s = A.sims[B.sim];
        const bool b1 = B.is_b1 != 0;
        int org[DIM];
        patch_origin<DIM>(s, uniform(ent.key) - B.pblk_off, org);
        for (int i = tid; i < NACC * PATCH; i += blockDim.x) lacc[i] = 0.0;
        if (tid < DIM) {
            bb[tid] = 0x7fffffff;
            bb[DIM + tid] = -0x7fffffff;
        }
        __syncthreads();
        const long long gbase = B.grid_base;
        for (int k = tid; k < count; k += blockDim.x) {
            const long long p = start + k;
            if (!A.act[p]) continue;
            double x[DIM];
#pragma unroll
            for (int d = 0; d < DIM; d++) x[d] = A.P[(F::X + d) * A.np + p];
            const double ext = A.P[F::EXT * A.np + p];
            Win<DIM> w;
            bool cpdi;
            const int mode = point_window<DIM>(s, x, ext, w, cpdi);
            if (mode == MODE_LEFT) {
                A.act[p] = 0;
                continue;
            }
            if (mode == MODE_NONE) continue;
            P2GVals<DIM> q;
            q.m = A.P[F::MASS * A.np + p];
            const double V = A.P[F::VOL * A.np + p];
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                q.mv[d] = A.P[(F::V + d) * A.np + p] * q.m;
                q.mg[d] = s.g[d] * q.m;
            }
            {
                double T6[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
                for (int c = 0; c < Cfg<DIM>::NTS; c++) T6[c] = A.P[(F::T + c) * A.np + p];
                const double Tf[3][3] = {{T6[0], T6[3], T6[4]}, {T6[3], T6[1], T6[5]}, {T6[4], T6[5], T6[2]}};
#pragma unroll
                for (int r = 0; r < DIM; r++)
#pragma unroll
                    for (int c = 0; c < DIM; c++) q.VT[r][c] = Tf[r][c] * V;  // T * v (explicit_usl.cpp:157-160)
            }
            if (mode == MODE_SEP) {
                bool fits = true;
#pragma unroll
                for (int d = 0; d < DIM; d++)
                    fits = fits && (w.base[d] >= org[d]) && (w.base[d] + w.cnt[d] <= org[d] + P);
                const double gscale = 1.0 / ((double)(1 << DIM) * 0.5 * ext);
                const double sscale = 1.0 / (double)(1 << DIM);
                if (fits) {
#pragma unroll
                    for (int d = 0; d < DIM; d++) {
                        atomicMin(&bb[d], w.base[d]);
                        atomicMax(&bb[DIM + d], w.base[d] + w.cnt[d] - 1);
                    }
                } else {
                    int lo[DIM], hi[DIM];
#pragma unroll
                    for (int d = 0; d < DIM; d++) {
                        lo[d] = w.base[d];
                        hi[d] = w.base[d] + w.cnt[d] - 1;
                    }
                    mark_range<DIM>(A, s, lo, hi);
                }
#pragma unroll
                for (int c2 = 0; c2 < (DIM == 3 ? 3 : 1); c2++) {
                    if (DIM == 3 && c2 >= w.cnt[DIM - 1]) continue;
#pragma unroll
                    for (int c1 = 0; c1 < 3; c1++) {
                        if (c1 >= w.cnt[1]) continue;
#pragma unroll
                        for (int c0 = 0; c0 < 3; c0++) {
                            if (c0 >= w.cnt[0]) continue;
                            const double sz = DIM == 3 ? w.sw[DIM - 1][c2] : 1.0;
                            const double dz = DIM == 3 ? w.dw[DIM - 1][c2] : 0.0;
                            const double syz = w.sw[1][c1] * sz;
                            const double S = w.sw[0][c0] * syz * sscale;
                            double G[DIM];
                            G[0] = w.dw[0][c0] * syz * gscale;
                            G[1] = w.sw[0][c0] * (w.dw[1][c1] * sz) * gscale;
                            if (DIM == 3) G[DIM - 1] = w.sw[0][c0] * (w.sw[1][c1] * dz) * gscale;
                            double vals[NACC];
                            p2g_node_values<DIM>(q, S, G, vals, b1);
                            int nijk[DIM];
                            nijk[0] = w.base[0] + c0;
                            nijk[1] = w.base[1] + c1;
                            if (DIM == 3) nijk[DIM - 1] = w.base[DIM - 1] + c2;
                            if (fits) {
                                int li = nijk[0] - org[0] + P * (nijk[1] - org[1]);
                                if (DIM == 3) li += P * P * (nijk[DIM - 1] - org[DIM - 1]);
#pragma unroll
                                for (int f = 0; f < NACC; f++) {
                                    if (f >= 1 + 2 * DIM && !b1) continue;
                                    atomicAdd(&lacc[f * PATCH + li], vals[f]);
                                }
                            } else {
                                const long long gi = gbase + node_addr<DIM>(s, nijk);
#pragma unroll
                                for (int f = 0; f < NACC; f++) {
                                    if (f >= 1 + 2 * DIM && !b1) continue;
                                    unsafeAtomicAdd(&A.acc[f * A.nt + gi], vals[f]);
                                }
                            }
                        }
                    }
                }
            } else {  // MODE_TUP
                for_each_tuple<DIM>(s, x, ext, cpdi, [&](const int *nijk, double S, const double *G) {
                    double vals[NACC];
                    p2g_node_values<DIM>(q, S, G, vals, b1);
                    bool in = true;
#pragma unroll
                    for (int d = 0; d < DIM; d++) in = in && nijk[d] >= org[d] && nijk[d] < org[d] + P;
                    if (in) {
                        int li = nijk[0] - org[0] + P * (nijk[1] - org[1]);
                        if (DIM == 3) li += P * P * (nijk[DIM - 1] - org[DIM - 1]);
#pragma unroll
                        for (int d = 0; d < DIM; d++) {
                            atomicMin(&bb[d], nijk[d]);
                            atomicMax(&bb[DIM + d], nijk[d]);
                        }
                        for (int f = 0; f < NACC; f++) {
                            if (f >= 1 + 2 * DIM && !b1) continue;
                            atomicAdd(&lacc[f * PATCH + li], vals[f]);
                        }
                    } else {
                        const long long gi = gbase + node_addr<DIM>(s, nijk);
                        for (int f = 0; f < NACC; f++) {
                            if (f >= 1 + 2 * DIM && !b1) continue;
                            unsafeAtomicAdd(&A.acc[f * A.nt + gi], vals[f]);
                        }
                        mark_range<DIM>(A, s, nijk, nijk);
                    }
                });
            }
        }
        __syncthreads();
        // flush the touched part of the patch
        int lo[DIM], ex[DIM];
        int total = 1;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            lo[d] = bb[d];
            ex[d] = bb[DIM + d] >= bb[d] ? bb[DIM + d] - bb[d] + 1 : 0;  // empty when no LDS-path point
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
            int li = nijk[0] - org[0] + P * (nijk[1] - org[1]);
            if (DIM == 3) li += P * P * (nijk[DIM - 1] - org[DIM - 1]);
            const long long gi = gbase + node_addr<DIM>(s, nijk);
#pragma unroll
            for (int f = 0; f < NACC; f++) {
                if (f >= 1 + 2 * DIM && !b1) continue;
                const double v = lacc[f * PATCH + li];
                if (v != 0.0) unsafeAtomicAdd(&A.acc[f * A.nt + gi], v);
            }
        }
        if (tid == 0 && total > 0) {
            int hi[DIM];
#pragma unroll
            for (int d = 0; d < DIM; d++) hi[d] = lo[d] + ex[d] - 1;
            mark_range<DIM>(A, s, lo, hi);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------
// grid update: contact, boundary, moveGrid. One wavefront per node block.
template <int DIM>
__global__ __launch_bounds__(256) void k_grid(KArgs A) {
    using G = GF<DIM>;
    constexpr int NB = Cfg<DIM>::NB, L = Cfg<DIM>::LNB;
    const int par = A.par;
    if (blockIdx.x == 0 && threadIdx.x == 0) A.ncount[par ^ 1] = 0;
    const int cnt = A.ncount[par];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int wpb = blockDim.x >> 6;
    unsigned long long active_nodes = 0;
    for (int bi = blockIdx.x * wpb + wv; bi < cnt; bi += gridDim.x * wpb) {
        const long long gblk = A.nlist[(long long)par * A.nblk_total + bi];
        int si = 0;
        if (A.nsim > 1) {  // binary search: last sim with nblk_off <= gblk
            int lo = 0, hi = A.nsim - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (A.sims[mid].nblk_off <= gblk) lo = mid;
                else hi = mid - 1;
            }
            si = lo;
        }
        si = uniform(si);
        const SimDev &s = A.sims[si];
        const long long lb = gblk - s.nblk_off;
        int ijk[3] = {0, 0, 0};
        {
            long long t = lb;
            const int bx = (int)(t % s.NBn[0]);
            t /= s.NBn[0];
            ijk[0] = (bx << L) + (lane & (NB - 1));
            if (DIM == 3) {
                const int by = (int)(t % s.NBn[1]);
                const int bz = (int)(t / s.NBn[1]);
                ijk[1] = (by << L) + ((lane >> L) & (NB - 1));
                ijk[2] = (bz << L) + (lane >> (2 * L));
            } else {
                ijk[1] = ((int)t << L) + (lane >> L);
            }
        }
        bool valid = true;
        bool wall = false;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            valid = valid && ijk[d] <= s.N[d];
            wall = wall || ijk[d] == 0 || ijk[d] == s.N[d];
        }
        const long long addr = lb * 64 + lane;
        if (valid) {
            const int nb = s.nbody;
            double m[MAXB], p[MAXB][DIM], f[MAXB][DIM], nr[MAXB][DIM];
#pragma unroll
            for (int b = 0; b < MAXB; b++) {
                m[b] = 0.0;
#pragma unroll
                for (int d = 0; d < DIM; d++) p[b][d] = f[b][d] = nr[b][d] = 0.0;
                if (b < nb) {
                    const BodyDev &B = A.bodies[s.body0 + b];
                    const long long gi = B.grid_base + addr;
                    m[b] = A.acc[G::M * A.nt + gi];
                    A.acc[G::M * A.nt + gi] = 0.0;
#pragma unroll
                    for (int d = 0; d < DIM; d++) {
                        p[b][d] = A.acc[(G::P + d) * A.nt + gi];
                        f[b][d] = A.acc[(G::F + d) * A.nt + gi];
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
            for (int c = 0; c < s.ncontact; c++) {
                const ContactDev C = A.contacts[s.contact0 + c];
                double m1 = 0, m2 = 0, p1[DIM] = {}, p2[DIM] = {}, f1[DIM] = {}, f2[DIM] = {}, n1[DIM] = {};
#pragma unroll
                for (int b = 0; b < MAXB; b++) {
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
                for (int b = 0; b < MAXB; b++) {
#pragma unroll
                    for (int d = 0; d < DIM; d++) {
                        const double fc = nrm[d] * fnc + s1[d] * fsc;
                        if (b == C.b1) f[b][d] = f[b][d] + fc;
                        if (b == C.b2) f[b][d] = f[b][d] - fc;
                    }
                }
            }
            // CartesianBox then moveGrid, per body
#pragma unroll
            for (int b = 0; b < MAXB; b++) {
                if (b >= nb) continue;
                const BodyDev &B = A.bodies[s.body0 + b];
                if (B.has_boundary && wall) {
#pragma unroll
                    for (int d = 0; d < DIM; d++) p[b][d] = f[b][d] = 0.0;
                }
                const long long gi = B.grid_base + addr;
                const bool has = m[b] > 0.0;
                const double rm = has ? 1.0 / m[b] : 0.0;
#pragma unroll
                for (int d = 0; d < DIM; d++) {
                    const double pn = p[b][d] + f[b][d] * dt;
                    A.out[(G::OV + d) * A.nt + gi] = has ? pn * rm : 0.0;
                    A.out[(G::ODV + d) * A.nt + gi] = has ? (f[b][d] * dt) * rm : 0.0;
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
// G2P: movePoints, calculateStrainRate, updateDensity, updateStress
template <int DIM>
__device__ __forceinline__ void g2p_node(const KArgs &A, long long gi, double S, const double *G, double dt,
                                         double *up, double *dvp, double (*Lp)[DIM]) {
    using GG = GF<DIM>;
    double v[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) {
        v[d] = A.out[(GG::OV + d) * A.nt + gi];
        const double dv = A.out[(GG::ODV + d) * A.nt + gi];
        up[d] = fma(S, v[d] * dt, up[d]);
        dvp[d] = fma(S, dv, dvp[d]);
    }
#pragma unroll
    for (int r = 0; r < DIM; r++)
#pragma unroll
        for (int c = 0; c < DIM; c++) Lp[r][c] = fma(v[r], G[c], Lp[r][c]);
}

template <int DIM>
__global__ __launch_bounds__(256) void k_g2p(KArgs A, int want_L) {
    using F = PF<DIM>;
    const int tid = threadIdx.x;
    const int nent = *A.npblk;
    for (int e = blockIdx.x; e < nent; e += gridDim.x) {
        const PBlk ent = A.plist[e];
        const int body = uniform(ent.body);
        const int start = uniform(ent.start), count = uniform(ent.count);
        const BodyDev &B = A.bodies[body];
        const SimDev &s = A.sims[B.sim];
        const long long gbase = B.grid_base;
        const double dt = s.dt;
        for (int k = tid; k < count; k += blockDim.x) {
            const long long p = start + k;
            if (!A.act[p]) continue;
            double x[DIM];
#pragma unroll
            for (int d = 0; d < DIM; d++) x[d] = A.P[(F::X + d) * A.np + p];
            const double ext = A.P[F::EXT * A.np + p];
            Win<DIM> w;
            bool cpdi;
            const int mode = point_window<DIM>(s, x, ext, w, cpdi);
            if (mode == MODE_LEFT) {  // cannot happen after the P2G of this step; kept for safety
                A.act[p] = 0;
                continue;
            }
            double up[DIM], dvp[DIM], Lp[DIM][DIM];
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                up[d] = dvp[d] = 0.0;
#pragma unroll
                for (int c = 0; c < DIM; c++) Lp[d][c] = 0.0;
            }
            if (mode == MODE_SEP) {
                const double gscale = 1.0 / ((double)(1 << DIM) * 0.5 * ext);
                const double sscale = 1.0 / (double)(1 << DIM);
#pragma unroll
                for (int c2 = 0; c2 < (DIM == 3 ? 3 : 1); c2++) {
                    if (DIM == 3 && c2 >= w.cnt[DIM - 1]) continue;
#pragma unroll
                    for (int c1 = 0; c1 < 3; c1++) {
                        if (c1 >= w.cnt[1]) continue;
#pragma unroll
                        for (int c0 = 0; c0 < 3; c0++) {
                            if (c0 >= w.cnt[0]) continue;
                            const double sz = DIM == 3 ? w.sw[DIM - 1][c2] : 1.0;
                            const double dz = DIM == 3 ? w.dw[DIM - 1][c2] : 0.0;
                            const double syz = w.sw[1][c1] * sz;
                            const double S = w.sw[0][c0] * syz * sscale;
                            double G[DIM];
                            G[0] = w.dw[0][c0] * syz * gscale;
                            G[1] = w.sw[0][c0] * (w.dw[1][c1] * sz) * gscale;
                            if (DIM == 3) G[DIM - 1] = w.sw[0][c0] * (w.sw[1][c1] * dz) * gscale;
                            int nijk[DIM];
                            nijk[0] = w.base[0] + c0;
                            nijk[1] = w.base[1] + c1;
                            if (DIM == 3) nijk[DIM - 1] = w.base[DIM - 1] + c2;
                            g2p_node<DIM>(A, gbase + node_addr<DIM>(s, nijk), S, G, dt, up, dvp, Lp);
                        }
                    }
                }
            } else if (mode == MODE_TUP) {
                for_each_tuple<DIM>(s, x, ext, cpdi, [&](const int *nijk, double S, const double *G) {
                    g2p_node<DIM>(A, gbase + node_addr<DIM>(s, nijk), S, G, dt, up, dvp, Lp);
                });
            }
            // movePoints: x += S^T u; x_t += S^T dv (FLIP)
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                A.P[(F::X + d) * A.np + p] = x[d] + up[d];
                A.P[(F::V + d) * A.np + p] = A.P[(F::V + d) * A.np + p] + dvp[d];
            }
            // updateDensity: v *= exp(dt * trace(L))
            double tr = Lp[0][0] + Lp[1][1];
            if (DIM == 3) tr = tr + Lp[DIM - 1][DIM - 1];
            A.P[F::VOL * A.np + p] = A.P[F::VOL * A.np + p] * exp(dt * tr);
            // updateStress
            if (B.has_material) {
                double L3[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
                for (int r = 0; r < DIM; r++)
#pragma unroll
                    for (int c = 0; c < DIM; c++) L3[r][c] = Lp[r][c];
                double T6[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
                for (int c = 0; c < Cfg<DIM>::NTS; c++) T6[c] = A.P[(F::T + c) * A.np + p];
                isolin_update<double>(B.twoG, B.lambda, dt, L3, T6);
#pragma unroll
                for (int c = 0; c < Cfg<DIM>::NTS; c++) A.P[(F::T + c) * A.np + p] = T6[c];
            }
            if (want_L && A.Lout) {
                double L9[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
                for (int r = 0; r < DIM; r++)
#pragma unroll
                    for (int c = 0; c < DIM; c++) L9[3 * r + c] = Lp[r][c];
#pragma unroll
                for (int k = 0; k < 9; k++) A.Lout[k * A.np + p] = L9[k];
            }
        }
    }
}

// ---------------------------------------------------------------------------------------
// connectivity of the current state (DefaultPoints::generateMap's element lookups)
template <int DIM>
__global__ void k_connectivity(KArgs A, int *elem_out, uint8_t *code_out) {
    using F = PF<DIM>;
    constexpr int NC = 1 << DIM;
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= A.np) return;
    const int o = A.pid[p];
    if (o < 0) return;
    int bi = 0;
    {
        int lo = 0, hi = A.nbody - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (A.bodies[mid].p_off <= o) lo = mid;
            else hi = mid - 1;
        }
        bi = lo;
    }
    const SimDev &s = A.sims[A.bodies[bi].sim];
    for (int c = 0; c < NC; c++) elem_out[(long long)o * NC + c] = -1;
    double x[DIM];
#pragma unroll
    for (int d = 0; d < DIM; d++) x[d] = A.P[(F::X + d) * A.np + p];
    const double ext = A.P[F::EXT * A.np + p];
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
