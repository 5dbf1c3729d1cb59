// mpm_prec.h -- the two particle-state forms behind one kernel skeleton.
//
//   Prec64  plain form: FP64 state x, v, T, V, m, extent; the element search and the GIMP
//           corner arithmetic round exactly as the reference's (cartesian.cpp:219-253,
//           default_points.cpp:196-330), so connectivity is bit-identical.
//   Prec32  precision form: a packed cell index plus a 32-bit fixed-point offset inside
//           the cell (x = h * (cell + q 2^-32)), FP32 v, T, m, extent and the volume as
//           V0 exp(lnJ); shape functions, G2P and the IsotropicLinearElasticity update in
//           FP32. Nodal mass,
//           momentum, force and contact normals still accumulate in FP64 (LDS and global
//           atomics), and the grid update (p/m division, ContactHuang) stays FP64; only
//           its output v, dv is stored as FP32 for the gather.
//
// Each form provides: State, load/store, window (separable GIMP weights), tuples (per-tuple
// weights for CPDI-off points and the shapes the separable form does not cover), advance
// (movePoints .. updateStress), fetch (one grid-output value), the sort key inputs and the
// FP64 position used by the connectivity query.
#pragma once

#include "mpm_common.h"

namespace mpm {

struct Prec64 {
    using R = double;   // particle arithmetic
    using GO = double;  // grid output element
    static constexpr bool FP32 = false;
    template <int DIM> struct State {
        double x[DIM], v[DIM], T6[6], V, m, ext;
        double nup;  // Sand_SachithLocal plastic shear rate of the last update
    };

    // a prefetched state: the state itself
    template <int DIM> using Raw = State<DIM>;
    template <int DIM>
    __device__ __forceinline__ static void load_raw(const KArgs &A, long long p, Raw<DIM> &w, bool need_vol) {
        load<DIM>(A, p, w, need_vol);
    }
    template <int DIM>
    __device__ __forceinline__ static void cook(const Raw<DIM> &w, State<DIM> &q, bool) { q = w; }

    template <int DIM>
    __device__ __forceinline__ static void load(const KArgs &A, long long p, State<DIM> &q, bool) {
        using F = PF<DIM>;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            q.x[d] = A.P[(F::X + d) * A.np + p];
            q.v[d] = A.P[(F::V + d) * A.np + p];
        }
#pragma unroll
        for (int c = 0; c < 6; c++) q.T6[c] = 0.0;
#pragma unroll
        for (int c = 0; c < F::NTS; c++) q.T6[c] = A.P[(F::T + c) * A.np + p];
        q.V = A.P[F::VOL * A.np + p];
        q.m = A.P[F::MASS * A.np + p];
        q.ext = A.P[F::EXT * A.np + p];
    }

    template <int DIM>
    __device__ __forceinline__ static void store(const KArgs &A, long long p, const State<DIM> &q) {
        using F = PF<DIM>;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            A.P[(F::X + d) * A.np + p] = q.x[d];
            A.P[(F::V + d) * A.np + p] = q.v[d];
        }
#pragma unroll
        for (int c = 0; c < F::NTS; c++) A.P[(F::T + c) * A.np + p] = q.T6[c];
        A.P[F::VOL * A.np + p] = q.V;
    }

    template <int DIM>
    __device__ __forceinline__ static int window(const SimDev &s, const State<DIM> &q, Win<DIM, double> &w,
                                                 bool &cpdi) {
        return point_window<DIM>(s, q.x, q.ext, w, cpdi);
    }

    template <int DIM, class Fn>
    __device__ __forceinline__ static void tuples(const SimDev &s, const State<DIM> &q, bool cpdi, Fn &&fn) {
        for_each_tuple<DIM>(s, q.x, q.ext, cpdi, fn);
    }

    __device__ __forceinline__ static double fetch(const KArgs &A, long long i) { return A.out[i]; }

    // movePoints / calculateStrainRate / updateDensity / updateStress (explicit_usl.cpp:266-355)
    template <int DIM>
    __device__ __forceinline__ static void advance(const BodyDev &B, const SimDev &s, double dt, State<DIM> &q,
                                                   const double *up, const double *dvp, const double (*Lp)[DIM]) {
        (void)s;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            q.x[d] = q.x[d] + up[d];   // x += S^T u     (explicit_usl.cpp:279)
            q.v[d] = q.v[d] + dvp[d];  // x_t += S^T dv  (explicit_usl.cpp:292, FLIP)
        }
        double tr = Lp[0][0] + Lp[1][1];
        if (DIM == 3) tr = tr + Lp[DIM - 1][DIM - 1];
        // V *= exp(dt tr L) (explicit_usl.cpp:338). For |dt tr L| <= 2^-7 exp is its Taylor
        // polynomial through degree 6 (remainder below 2^-60 relative), by Horner; exp()
        // otherwise. Both are within two ulps of the exact value.
        const double xe = dt * tr;
        double ex;
        if (fabs(xe) <= 0x1p-7)
            ex = fma(xe, fma(xe, fma(xe, fma(xe, fma(xe, fma(xe, 1.0 / 720, 1.0 / 120), 1.0 / 24), 1.0 / 6), 0.5), 1.0),
                     1.0);
        else
            ex = exp(xe);
        q.V = q.V * ex;
        if (B.has_material) {
            double L3[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
            for (int r = 0; r < DIM; r++)
#pragma unroll
                for (int c = 0; c < DIM; c++) L3[r][c] = Lp[r][c];
            if (B.has_material == 2) q.nup = sand_update<double>(B, dt, q.m, q.V, L3, q.T6);
            else isolin_update<double>(B.twoG, B.lambda, dt, L3, q.T6);
        }
    }

    // Sand_SachithLocal: gammap += nup dt, gammap_dot = nup (sand_sachithlocal.cpp:226-229)
    template <int DIM>
    __device__ __forceinline__ static void store_plastic(const KArgs &A, long long p, double nup, double dt) {
        using F = PF<DIM>;
        A.P[F::GP * A.np + p] = A.P[F::GP * A.np + p] + nup * dt;
        A.P[F::GPD * A.np + p] = nup;
    }

    // sort key inputs: dual cell (shifted to >= 0) and octant inside it; false if the
    // point is outside [0, L]
    template <int DIM>
    __device__ __forceinline__ static bool dual_cell(const KArgs &A, const SimDev &s, long long p, int *dual, int &oct) {
        double x[DIM];
#pragma unroll
        for (int d = 0; d < DIM; d++) x[d] = A.P[(PF<DIM>::X + d) * A.np + p];
        if (!in_domain<DIM>(s, x)) return false;
        oct = 0;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            const double t = x[d] * s.invh[d] - 0.5;
            const double ft = floor(t);
            int c = (int)ft + 1;
            dual[d] = c < 0 ? 0 : (c > s.N[d] ? s.N[d] : c);
            oct |= (t - ft >= 0.5 ? 1 : 0) << d;
        }
        return true;
    }

    template <int DIM>
    __device__ __forceinline__ static void position(const KArgs &A, const SimDev &s, long long p, double *x,
                                                    double &ext) {
        (void)s;
#pragma unroll
        for (int d = 0; d < DIM; d++) x[d] = A.P[(PF<DIM>::X + d) * A.np + p];
        ext = A.P[PF<DIM>::EXT * A.np + p];
    }

    // slot p's state as the C-ABI hands it out: x, v (DIM each), T (6: xx yy zz xy xz yz,
    // the components the form does not store as 0), V
    template <int DIM>
    __device__ __forceinline__ static void export_state(const KArgs &A, const SimDev &, long long p, double *x,
                                                        double *v, double *T, double &V) {
        using F = PF<DIM>;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            x[d] = A.P[(F::X + d) * A.np + p];
            v[d] = A.P[(F::V + d) * A.np + p];
        }
#pragma unroll
        for (int c = 0; c < 6; c++) T[c] = c < F::NTS ? A.P[(F::T + c) * A.np + p] : 0.0;
        V = A.P[F::VOL * A.np + p];
    }

    // point i < n of an upload, staged as planes of n doubles (x[DIM] v[DIM] T[NTS] V m ext),
    // into slot i of the planes `dst` (np per plane); gammap and gammap_dot start at 0
    template <int DIM>
    __device__ __forceinline__ static void import_state(const double *src, long long n, long long i,
                                                        const SimDev &, double *dst, long long np) {
        using F = PF<DIM>;
        for (int f = 0; f < F::EXT + 1; f++) dst[f * np + i] = src[f * n + i];
        dst[F::GP * np + i] = 0.0;
        dst[F::GPD * np + i] = 0.0;
    }
};

struct Prec32 {
    using R = float;
    using GO = float;
    static constexpr bool FP32 = true;
    template <int DIM> struct State {
        int c[DIM];       // cell
        uint32_t q[DIM];  // offset inside the cell, 32-bit fixed point (x = h (c + q 2^-32))
        float r[DIM];     // q 2^-32 rounded to FP32, for the shape functions
        float v[DIM], T6[6], lnJ, V, m, ext;  // V = V0 exp(lnJ), loaded for P2G and sand only
        float nup;  // Sand_SachithLocal plastic shear rate of the last update
    };

    // a prefetched state as stored (packed cell, fixed-point offsets, V0): fewer registers
    // while the loads are in flight; cook() derives the rest
    template <int DIM> struct Raw {
        uint32_t cell, q[DIM];
        float v[DIM], T[Cfg<DIM>::NTS], lnJ, V0, m, ext;
    };
    template <int DIM>
    __device__ __forceinline__ static void load_raw(const KArgs &A, long long p, Raw<DIM> &w, bool need_vol) {
        using F = PW<DIM>;
        const float *Pf = reinterpret_cast<const float *>(A.Pw);
        w.cell = A.Pw[F::CELL * A.np + p];
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            w.q[d] = A.Pw[(F::Q + d) * A.np + p];
            w.v[d] = Pf[(F::V + d) * A.np + p];
        }
#pragma unroll
        for (int c = 0; c < F::NTS; c++) w.T[c] = Pf[(F::T + c) * A.np + p];
        w.lnJ = Pf[F::LNJ * A.np + p];
        w.V0 = need_vol ? Pf[F::V0 * A.np + p] : 0.0f;
        w.m = Pf[F::MASS * A.np + p];
        w.ext = Pf[F::EXT * A.np + p];
    }
    template <int DIM>
    __device__ __forceinline__ static void cook(const Raw<DIM> &w, State<DIM> &q, bool need_vol) {
        unpack_cell<DIM>(w.cell, q.c);
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            q.q[d] = w.q[d];
            q.r[d] = offset(w.q[d]);
            q.v[d] = w.v[d];
        }
#pragma unroll
        for (int c = 0; c < 6; c++) q.T6[c] = 0.0f;
#pragma unroll
        for (int c = 0; c < Cfg<DIM>::NTS; c++) q.T6[c] = w.T[c];
        q.lnJ = w.lnJ;
        q.V = need_vol ? w.V0 * expf(w.lnJ) : 0.0f;
        q.m = w.m;
        q.ext = w.ext;
    }

    template <int DIM>
    __device__ __forceinline__ static void load(const KArgs &A, long long p, State<DIM> &q, bool need_vol) {
        using F = PW<DIM>;
        const float *Pf = reinterpret_cast<const float *>(A.Pw);
        unpack_cell<DIM>(A.Pw[F::CELL * A.np + p], q.c);
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            q.q[d] = A.Pw[(F::Q + d) * A.np + p];
            q.r[d] = offset(q.q[d]);
            q.v[d] = Pf[(F::V + d) * A.np + p];
        }
#pragma unroll
        for (int c = 0; c < 6; c++) q.T6[c] = 0.0f;
#pragma unroll
        for (int c = 0; c < F::NTS; c++) q.T6[c] = Pf[(F::T + c) * A.np + p];
        q.lnJ = Pf[F::LNJ * A.np + p];
        q.V = need_vol ? Pf[F::V0 * A.np + p] * expf(q.lnJ) : 0.0f;
        q.m = Pf[F::MASS * A.np + p];
        q.ext = Pf[F::EXT * A.np + p];
    }

    template <int DIM>
    __device__ __forceinline__ static void store(const KArgs &A, long long p, const State<DIM> &q) {
        using F = PW<DIM>;
        float *Pf = reinterpret_cast<float *>(A.Pw);
        A.Pw[F::CELL * A.np + p] = pack_cell<DIM>(q.c);
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            A.Pw[(F::Q + d) * A.np + p] = q.q[d];
            Pf[(F::V + d) * A.np + p] = q.v[d];
        }
#pragma unroll
        for (int c = 0; c < F::NTS; c++) Pf[(F::T + c) * A.np + p] = q.T6[c];
        Pf[F::LNJ * A.np + p] = q.lnJ;
    }

    // q 2^-32 in FP32, rounded toward zero so that r < 1 and the point stays in cell c
    __device__ __forceinline__ static float offset(uint32_t q) { return __uint2float_rz(q) * 0x1p-32f; }

    // x = h (c + r) is inside [0, L] in dimension d: element E = c + floor(r) with
    // 0 <= E < N, or exactly on the last node (E == N, frac 0). NaN is outside.
    __device__ __forceinline__ static bool inside(int c, float r, int N) {
        if (!(fabsf(r) < 8.0f)) return false;  // NaN, or far outside any cell of c
        const float fl = floorf(r);
        const int E = c + (int)fl;
        return (E >= 0 && E < N) || (E == N && r == fl);
    }

    template <int DIM>
    __device__ __forceinline__ static int window(const SimDev &s, const State<DIM> &q, Win<DIM, float> &w,
                                                 bool &cpdi) {
        cpdi = false;
        // CPDI off (cpdi_spec 0, or a corner outside the domain): both "corners" at the point
        // and the gradient slots set to -/+ extent / h, as in point_window; on a node line this
        // is the one-sided derivative of the per-tuple path below
        bool off = s.cpdi_spec != 1;
        float cm[DIM], cp[DIM];
        bool cin = true;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            const float a = off ? 0.0f : 0.5f * q.ext * s.invhf[d];  // half extent in cells
            cm[d] = q.r[d] - a;
            cp[d] = q.r[d] + a;
            cin = cin && inside(q.c[d], cm[d], s.N[d]) && inside(q.c[d], cp[d], s.N[d]);
        }
        // both corners inside put the point, which lies between them, inside too (the domain
        // is a box; inside() is monotone in r); otherwise the point is tested on its own
        if (!cin) {
            bool in = true;
#pragma unroll
            for (int d = 0; d < DIM; d++) in = in && inside(q.c[d], q.r[d], s.N[d]);
            if (!in) return MODE_LEFT;
            off = true;  // a corner leaves the domain: CPDI off
#pragma unroll
            for (int d = 0; d < DIM; d++) cm[d] = cp[d] = q.r[d];
        }
        if (off && !(q.ext > 0.0f)) return MODE_TUP;
        cpdi = !off;
        int mode = MODE_SEP;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            const float flm = floorf(cm[d]), flp = floorf(cp[d]);
            const int em = q.c[d] + (int)flm, ep = q.c[d] + (int)flp;
            const float fm = cm[d] - flm, fp = cp[d] - flp;
            const bool vm = em < s.N[d], vp = ep < s.N[d];  // a corner exactly at L has no element
            if (!vm && !vp) mode = MODE_NONE;
            const int lo = vm ? em : ep;
            const int hi = vp ? ep : em;
            if (hi - lo > 1) mode = mode == MODE_NONE ? MODE_NONE : MODE_TUP;
            int b = q.c[d] + (int)floorf(q.r[d] - 0.5f);
            b = b < hi - 1 ? hi - 1 : (b > lo ? lo : b);
            w.base[d] = b;
            float s0 = 0, s1 = 0, s2 = 0, d0 = 0, d1 = 0, d2 = 0;
            if (vm) {
                const float h0 = 1.0f - fm, h1 = fm;
                const bool k0 = em == b;
                s0 += k0 ? h0 : 0.0f;
                s1 += k0 ? h1 : h0;
                s2 += k0 ? 0.0f : h1;
                d0 -= k0 ? h0 : 0.0f;
                d1 -= k0 ? h1 : h0;
                d2 -= k0 ? 0.0f : h1;
            }
            if (vp) {
                const float h0 = 1.0f - fp, h1 = fp;
                const bool k0 = ep == b;
                s0 += k0 ? h0 : 0.0f;
                s1 += k0 ? h1 : h0;
                s2 += k0 ? 0.0f : h1;
                d0 += k0 ? h0 : 0.0f;
                d1 += k0 ? h1 : h0;
                d2 += k0 ? 0.0f : h1;
            }
            if (off) {
                const float g = q.ext * s.invhf[d];
                const bool k0 = em == b;
                d0 = k0 ? -g : 0.0f;
                d1 = k0 ? g : -g;
                d2 = k0 ? 0.0f : g;
            }
            w.sw[d][0] = s0; w.sw[d][1] = s1; w.sw[d][2] = s2;
            w.dw[d][0] = d0; w.dw[d][1] = d1; w.dw[d][2] = d2;
            w.rch[d] = (unsigned char)((vm ? (em == b ? 3 : 6) : 0) | (vp ? (ep == b ? 3 : 6) : 0));
        }
        return mode;
    }

    // per-tuple weights in cell-relative coordinates (same tuple set as the reference's
    // COO construction: corners, or the point itself with CPDI off)
    template <int DIM, class Fn>
    __device__ __forceinline__ static void tuples(const SimDev &s, const State<DIM> &q, bool cpdi, Fn &&fn) {
        constexpr int NC = 1 << DIM;
        const int ncorner = cpdi ? NC : 1;
        for (int c = 0; c < ncorner; c++) {
            int E[DIM];
            float f[DIM];
            bool valid = true;
#pragma unroll
            for (int d = 0; d < DIM; d++) {
                const float A = ((c >> d) & 1) ? 1.0f : -1.0f;
                const float rr = cpdi ? q.r[d] + 0.5f * q.ext * s.invhf[d] * A : q.r[d];
                const float fl = floorf(rr);
                E[d] = q.c[d] + (int)fl;
                f[d] = rr - fl;
                valid = valid && E[d] >= 0 && E[d] < s.N[d];
            }
            if (!valid) continue;
            for (int n = 0; n < NC; n++) {
                int nijk[DIM];
                float rn[DIM];
                float tmp = 1.0f;
#pragma unroll
                for (int d = 0; d < DIM; d++) {
                    const int bit = (n >> d) & 1;
                    nijk[d] = E[d] + bit;
                    rn[d] = f[d] - (float)bit;  // (x - x_node) / h
                    tmp *= (1.0f - fabsf(rn[d]));
                }
                float S, G[DIM];
                if (cpdi) {
                    S = tmp / (float)NC;
#pragma unroll
                    for (int d = 0; d < DIM; d++) {
                        const float A = ((c >> d) & 1) ? 1.0f : -1.0f;
                        G[d] = tmp / ((float)NC * 0.5f * q.ext) * A;
                    }
                } else {
                    S = tmp;
                    // cartesian.cpp:307-357 computes -N / (1 - |r_d|) * r_d / |r_d| / h_d and
                    // zeroes it at |r_d| >= 1. For the element's own nodes that equals the
                    // product form below except exactly on a node line (r_d = 0: 0/0;
                    // |r_d| = 1: zero instead of the limit). In FP64 a point lands there with
                    // probability ~2^-52 per step; on the 2^-32 fixed-point grid of this form
                    // it happens routinely, so the form takes the one-sided derivative of
                    // the element the point is in -- what FP64 computes a hair off the line.
#pragma unroll
                    for (int d = 0; d < DIM; d++) {
                        float others = 1.0f;
#pragma unroll
                        for (int e = 0; e < DIM; e++)
                            if (e != d) others *= (1.0f - fabsf(rn[e]));
                        G[d] = -copysignf(1.0f, rn[d]) * others / s.hf[d];
                    }
                }
                fn(nijk, S, G);
            }
        }
    }

    __device__ __forceinline__ static float fetch(const KArgs &A, long long i) { return A.out32[i]; }

    // Positions advance in 32-bit fixed point inside the cell, so increments far below an
    // FP32 ulp of the offset (slow points) still accumulate; the volume advances as
    // lnJ += dt tr L, exact in the small increments where V *= exp(dt tr L) in FP32 is not.
    template <int DIM>
    __device__ __forceinline__ static void advance(const BodyDev &B, const SimDev &s, double dt, State<DIM> &q,
                                                   const float *up, const float *dvp, const float (*Lp)[DIM]) {
        const float dtf = (float)dt;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            const float inc = up[d] * s.invhf[d];  // cells
            if (fabsf(inc) < 4.0f) {
                const long long t = (long long)q.q[d] + (long long)rintf(inc * 0x1p32f);
                int c = q.c[d] + (int)(t >> 32);  // floor(t / 2^32)
                uint32_t f = (uint32_t)t;
                // left the domain by more than a cell in one step: keep it outside, within
                // the packed range (cell -1, or half a cell past L)
                if (c < -1) {
                    c = -1;
                    f = 0u;
                } else if (c > s.N[d]) {
                    c = s.N[d];
                    f = 0x80000000u;
                }
                q.c[d] = c;
                q.q[d] = f;
            } else {  // NaN or a jump of 4+ cells: park the point outside the grid
                q.c[d] = -1;
                q.q[d] = 0u;
            }
            q.r[d] = offset(q.q[d]);  // the fused pass maps the moved point
            q.v[d] = q.v[d] + dvp[d];
        }
        float tr = Lp[0][0] + Lp[1][1];
        if (DIM == 3) tr = tr + Lp[DIM - 1][DIM - 1];
        q.lnJ = q.lnJ + dtf * tr;
        // the fused pass scatters with the updated volume and the sand update reads it; the
        // state holds V = 0 where neither needs it (the split G2P of an elastic body)
        if (q.V != 0.0f) q.V = q.V * expf(dtf * tr);
        if (B.has_material) {
            float L3[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
            for (int r = 0; r < DIM; r++)
#pragma unroll
                for (int c = 0; c < DIM; c++) L3[r][c] = Lp[r][c];
            if (B.has_material == 2) q.nup = sand_update<float>(B, dt, q.m, q.V, L3, q.T6);
            else isolin_update<float>(B.twoG, B.lambda, dt, L3, q.T6);
        }
    }

    // Sand_SachithLocal: gammap += nup dt, gammap_dot = nup (sand_sachithlocal.cpp:226-229)
    template <int DIM>
    __device__ __forceinline__ static void store_plastic(const KArgs &A, long long p, float nup, double dt) {
        using F = PW<DIM>;
        float *Pf = reinterpret_cast<float *>(A.Pw);
        Pf[F::GP * A.np + p] = Pf[F::GP * A.np + p] + nup * (float)dt;
        Pf[F::GPD * A.np + p] = nup;
    }

    template <int DIM>
    __device__ __forceinline__ static bool dual_cell(const KArgs &A, const SimDev &s, long long p, int *dual, int &oct) {
        using F = PW<DIM>;
        int c[DIM];
        unpack_cell<DIM>(A.Pw[F::CELL * A.np + p], c);
        uint32_t qq[DIM];
        bool in = true;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            qq[d] = A.Pw[(F::Q + d) * A.np + p];
            in = in && inside(c[d], offset(qq[d]), s.N[d]);
        }
        if (!in) return false;
        oct = 0;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            // dual cell of x/h - 1/2 and the half of it the point is in
            const bool upper = qq[d] >= 0x80000000u;
            int dc = c[d] + (upper ? 1 : 0);
            dual[d] = dc < 0 ? 0 : (dc > s.N[d] ? s.N[d] : dc);
            oct |= (upper ? 0 : 1) << d;
        }
        return true;
    }

    template <int DIM>
    __device__ __forceinline__ static void position(const KArgs &A, const SimDev &s, long long p, double *x,
                                                    double &ext) {
        using F = PW<DIM>;
        const float *Pf = reinterpret_cast<const float *>(A.Pw);
        int c[DIM];
        unpack_cell<DIM>(A.Pw[F::CELL * A.np + p], c);
#pragma unroll
        for (int d = 0; d < DIM; d++)
            x[d] = s.h[d] * ((double)c[d] + (double)A.Pw[(F::Q + d) * A.np + p] * 0x1p-32);
        ext = Pf[F::EXT * A.np + p];
    }

    // slot p's state as the C-ABI hands it out (see Prec64::export_state): x = h (c + q 2^-32)
    // in FP64, V = V0 exp(lnJ) in FP64
    template <int DIM>
    __device__ __forceinline__ static void export_state(const KArgs &A, const SimDev &s, long long p, double *x,
                                                        double *v, double *T, double &V) {
        using F = PW<DIM>;
        const float *Pf = reinterpret_cast<const float *>(A.Pw);
        int c[DIM];
        unpack_cell<DIM>(A.Pw[F::CELL * A.np + p], c);
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            x[d] = s.h[d] * ((double)c[d] + (double)A.Pw[(F::Q + d) * A.np + p] * 0x1p-32);
            v[d] = (double)Pf[(F::V + d) * A.np + p];
        }
#pragma unroll
        for (int k = 0; k < 6; k++) T[k] = k < F::NTS ? (double)Pf[(F::T + k) * A.np + p] : 0.0;
        V = (double)Pf[F::V0 * A.np + p] * exp((double)Pf[F::LNJ * A.np + p]);
    }

    // point i < n of an upload (staged as for Prec64::import_state) into the precision form:
    // cell + 32-bit fixed-point offset (x = h (cell + q 2^-32)), lnJ = 0 with V0 = V, the rest
    // rounded to FP32; a coordinate far outside the grid (or NaN) parks the point in cell -1,
    // where the first sort deactivates it
    template <int DIM>
    __device__ __forceinline__ static void import_state(const double *src, long long n, long long i,
                                                        const SimDev &s, uint32_t *dst, long long np) {
        using F = PW<DIM>;
        float *df = reinterpret_cast<float *>(dst);
        int c[DIM];
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            const double t = src[d * n + i] / s.h[d];
            double cf = floor(t);
            double fr = t - cf;  // in [0, 1)
            if (!(cf >= -1.0)) {
                cf = -1.0;
                fr = 0.0;
            } else if (cf > (double)s.N[d]) {
                cf = (double)s.N[d];
                fr = 0.5;
            }
            double qf = rint(ldexp(fr, 32));
            if (qf >= 4294967296.0) {  // rounds up to the next node
                qf = 0.0;
                cf += 1.0;
            }
            c[d] = (int)cf;
            dst[(F::Q + d) * np + i] = (uint32_t)qf;
            df[(F::V + d) * np + i] = (float)src[(DIM + d) * n + i];
        }
        dst[F::CELL * np + i] = pack_cell<DIM>(c);
#pragma unroll
        for (int k = 0; k < F::NTS; k++) df[(F::T + k) * np + i] = (float)src[(2 * DIM + k) * n + i];
        const int o = 2 * DIM + F::NTS;
        df[F::LNJ * np + i] = 0.0f;
        df[F::V0 * np + i] = (float)src[o * n + i];
        df[F::MASS * np + i] = (float)src[(o + 1) * n + i];
        df[F::EXT * np + i] = (float)src[(o + 2) * n + i];
        df[F::GP * np + i] = 0.0f;
        df[F::GPD * np + i] = 0.0f;
    }
};

}  // namespace mpm
