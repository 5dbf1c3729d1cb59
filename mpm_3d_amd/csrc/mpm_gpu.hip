// mpm_gpu.hip -- libmpm_gpu.so: device state, step orchestration and the C-ABI declared
// in include/mpm_gpu.h. Everything stays resident in HBM between calls; a step call is a
// fixed sequence of kernel launches on the handle's stream with no host synchronisation
// until the call returns.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <limits>
#include <mutex>
#include <vector>

#include <rocprofiler-sdk-roctx/roctx.h>

#include "mpm_gpu.h"
#include "mpm_kernels.h"
#include "mpm_sort.h"

using namespace mpm;

namespace {

// A roctx range over one C-ABI call (rocprofv3 --marker-trace shows uploads, steps,
// downloads and slot reconfiguration beside the kernels they enqueue); no-ops without a tracer.
struct TraceRange {
    explicit TraceRange(const char *name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange &) = delete;
    TraceRange &operator=(const TraceRange &) = delete;
};

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                            \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return fail(e_ == hipErrorOutOfMemory ? MPM_ERR_NOMEM : MPM_ERR_HIP, "%s: %s (%s:%d)", \
                        #expr, hipGetErrorString(e_), __FILE__, __LINE__);                       \
    } while (0)

constexpr int NCLASS = 5;
enum { KC_PASS = 0, KC_GRID = 1, KC_SORT = 2, KC_P2G = 3, KC_G2P = 4 };

long long ceil_div(long long a, long long b) { return (a + b - 1) / b; }

struct Layout {  // sizes derived from the descriptors, shared by create and footprint
    int dim = 0;
    std::vector<SimDev> sims;
    std::vector<BodyDev> bodies;
    std::vector<ContactDev> contacts;
    long long np = 0, nt = 0, nblk_total = 0, nkeys = 0, npblk_total = 0;
    long long crec_total = 0;  // contact record elements (ContactDev::rec), allocated on demand
    bool any_b1 = false;
    bool cell_fit = false;  // the cell-ordered sort's counters fit their budget (CELL_SORT_BUDGET)
};

// The cell-ordered sort keeps 64 counters (4 bytes each, turned into offsets and back to
// zero by every sort) per (block, octant) key over the whole particle-block grid of every
// body; handles whose counters would exceed this many bytes sort by (block, octant) only.
constexpr long long CELL_SORT_BUDGET = 4ll << 30;

// IsotropicLinearElasticity::init (isolin.cpp:34-38); Sand_SachithLocal::init computes the
// same moduli and keeps its six properties (sand_sachithlocal.cpp:58-117)
void set_material(BodyDev &B, const mpm_body_desc &bd) {
    B.has_material = bd.has_material;
    B.twoG = B.lambda = B.G = B.K = 0.0;
    for (int k = 0; k < 6; k++) B.sand[k] = 0.0;
    B.sand_dp = 0;
    if (!bd.has_material) return;
    const double G = bd.E / (2.0 * (1.0 + bd.nu));
    const double K = bd.E / (3.0 * (1.0 - 2 * bd.nu));
    B.G = G;
    B.K = K;
    B.twoG = 2 * G;
    B.lambda = K - 2.0 * G / 3.0;
    if (bd.has_material == 2) {
        for (int k = 0; k < 6; k++) B.sand[k] = bd.sand[k];
        B.sand_dp = bd.sand_density_pressure != 0;
    }
}

int check_material(const mpm_body_desc &bd, int sim, int body) {
    if (bd.has_material < 0 || bd.has_material > 2)
        return fail(MPM_ERR_ARG, "sim %d body %d: has_material must be 0..2, got %d", sim, body, bd.has_material);
    if (bd.has_material == 2 && !(bd.sand[0] > 0 && bd.sand[2] > 0 && bd.sand[3] > 0))
        return fail(MPM_ERR_ARG, "sim %d body %d: sand needs grain diameter, grains_rho, rho_critical > 0", sim, body);
    return MPM_OK;
}

int build_layout(const mpm_sim_desc *sd, int nsim, Layout &Lo) {
    if (nsim <= 0 || !sd) return fail(MPM_ERR_ARG, "need at least one simulation");
    Lo.dim = sd[0].grid.dim;
    if (Lo.dim != 2 && Lo.dim != 3) return fail(MPM_ERR_ARG, "dim must be 2 or 3 (got %d)", Lo.dim);
    long long p_off = 0, grid_base = 0, nblk_off = 0, pblk_off = 0;
    for (int i = 0; i < nsim; i++) {
        const mpm_grid_desc &g = sd[i].grid;
        if (g.dim != Lo.dim) return fail(MPM_ERR_ARG, "all simulations of a handle must share dim");
        if (sd[i].nbody < 1 || sd[i].nbody > MPM_MAX_BODIES_PER_SIM)
            return fail(MPM_ERR_ARG, "sim %d: 1..%d bodies supported (got %d)", i, MPM_MAX_BODIES_PER_SIM,
                        sd[i].nbody);
        if (g.cpdi_spec != 0 && g.cpdi_spec != 1)
            return fail(MPM_ERR_ARG, "Unrecognized Argument for use_cpdi: %d", g.cpdi_spec);
        if (g.contact_spec != 0 && g.contact_spec != 1)
            return fail(MPM_ERR_ARG, "ERROR: Unknown SPEC in contact: %d", g.contact_spec);
        SimDev s{};
        s.dim = g.dim;
        const int NB = g.dim == 3 ? Cfg<3>::NB : Cfg<2>::NB;
        const int PB = g.dim == 3 ? Cfg<3>::PB : Cfg<2>::PB;
        long long nblk = 1, pnb = 1;
        s.element_count = 1;
        for (int d = 0; d < 3; d++) {
            if (d < g.dim) {
                if (!(g.L[d] > 0) || g.N[d] < 1)
                    return fail(MPM_ERR_ARG, "sim %d: need L > 0 and N >= 1 in every dimension", i);
                s.N[d] = g.N[d];
                s.L[d] = g.L[d];
                s.h[d] = g.L[d] / g.N[d];
                s.invh[d] = 1.0 / s.h[d];
                s.NoverL[d] = (double)g.N[d] / g.L[d];
                s.ftol[d] = std::ldexp((double)g.N[d], -48);
                s.hf[d] = (float)s.h[d];
                s.invhf[d] = (float)s.invh[d];
                s.NBn[d] = (int)ceil_div(g.N[d] + 2 + NSHIFT, NB);  // storage of nodes -NSHIFT..N+1
                s.PNB[d] = (int)ceil_div(g.N[d] + 1, PB);
                nblk *= s.NBn[d];
                pnb *= s.PNB[d];
                s.element_count *= g.N[d];
                s.g[d] = g.gravity[d];
            } else {
                s.N[d] = 0;
                s.NBn[d] = 1;
                s.PNB[d] = 1;
            }
        }
        s.nblk_off = nblk_off;
        s.nblk = nblk;
        s.body0 = (int)Lo.bodies.size();
        s.nbody = sd[i].nbody;
        s.contact0 = (int)Lo.contacts.size();
        s.ncontact = sd[i].ncontact;
        s.cpdi_spec = g.cpdi_spec;
        s.contact_spec = g.contact_spec;
        s.dt = 0.0;
        nblk_off += nblk;
        std::vector<int> is_b1(sd[i].nbody, 0);
        for (int c = 0; c < sd[i].ncontact; c++) {
            const mpm_contact_desc &cd = sd[i].contacts[c];
            if (cd.body1 < 0 || cd.body1 >= sd[i].nbody || cd.body2 < 0 || cd.body2 >= sd[i].nbody ||
                cd.body1 == cd.body2)
                return fail(MPM_ERR_ARG, "sim %d contact %d: bodies must be two distinct sim-local indices", i, c);
            is_b1[cd.body1] = 1;
            Lo.any_b1 = true;
            ContactDev C{cd.body1, cd.body2, cd.mu, Lo.crec_total};
            Lo.crec_total += 2LL * Lo.dim * nblk * 64;
            Lo.contacts.push_back(C);
        }
        for (int b = 0; b < sd[i].nbody; b++) {
            const mpm_body_desc &bd = sd[i].bodies[b];
            if (bd.n_points < 0) return fail(MPM_ERR_ARG, "negative point count");
            if (bd.has_boundary < 0 || bd.has_boundary > 3)
                return fail(MPM_ERR_ARG, "sim %d body %d: has_boundary must be 0..3, got %d", i, b, bd.has_boundary);
            if (int rc = check_material(bd, i, b)) return rc;
            BodyDev B{};
            B.sim = i;
            B.local = b;
            B.p_off = p_off;
            B.n = bd.n_points;
            B.grid_base = grid_base;
            B.pblk_off = pblk_off;
            B.npblk = pnb;
            B.has_boundary = bd.has_boundary;
            B.bc_mu = bd.bc_mu;
            B.is_b1 = is_b1[b];
            set_material(B, bd);
            p_off += bd.n_points;
            grid_base += nblk * 64;
            pblk_off += pnb;
            Lo.bodies.push_back(B);
        }
        Lo.sims.push_back(s);
    }
    Lo.np = p_off;
    Lo.nt = grid_base;
    Lo.nblk_total = nblk_off;
    Lo.npblk_total = pblk_off;
    Lo.nkeys = pblk_off * (1ll << Lo.dim) + 1;  // (block, octant) keys + inactive bucket
    if (Lo.np >= (1ll << 31) - 1) return fail(MPM_ERR_ARG, "at most 2^31-2 points per handle");
    if (Lo.nkeys >= (1ll << 31) - 1) return fail(MPM_ERR_ARG, "grid too large for one handle");
    Lo.cell_fit = (Lo.nkeys - 1) * 64 * 4 <= CELL_SORT_BUDGET;
    return MPM_OK;
}

long long footprint_bytes(const Layout &Lo, int precision) {
    const bool f32 = precision == MPM_PREC_FP32;
    const int nf = f32 ? (Lo.dim == 3 ? PW<3>::N : PW<2>::N) : (Lo.dim == 3 ? PF<3>::N : PF<2>::N);
    const int esz = f32 ? 4 : 8;
    const int nacc = Lo.dim == 3 ? GF<3>::NACC : GF<2>::NACC;
    const int nout = Lo.dim == 3 ? GF<3>::NOUT : GF<2>::NOUT;
    long long b = 0;
    b += 2 * Lo.np * (nf * esz + 1 + 4);   // particle planes, double-buffered, + act + pid
    b += Lo.np * (4 + 4);                  // sort key + rank
    b += Lo.nt * (nacc * 8 + nout * esz);  // grid
    b += Lo.nblk_total * (4 + 2 * 4);      // node-block flags + two lists
    b += Lo.nkeys * (4 + 4);               // histogram, offsets
    if (Lo.cell_fit) b += (Lo.nkeys - 1) * 64 * 4 + Lo.np;  // cell counters, a slot's cell
    b += Lo.npblk_total * sizeof(PBlk);    // block list
    b += (Lo.npblk_total + 1) * (4 + 4);   // block flags, ranks
    b += Lo.np * 4 + 16;                   // side list of the particle passes
    b += 8;                                // block counter and non-finite flags
    return b;
}

struct Timer {
    hipEvent_t a, b;
    int cls;
};

}  // namespace

struct mpm_gpu_handle {
    Layout lo;
    int precision = 0;
    int device = 0;  // the device current at create
    hipStream_t stream = nullptr;
    SimDev *d_sims = nullptr;
    BodyDev *d_bodies = nullptr;
    ContactDev *d_contacts = nullptr;
    double *P = nullptr, *P2 = nullptr;     // plain form planes
    uint32_t *Pw = nullptr, *Pw2 = nullptr;  // precision form planes
    int8_t *act = nullptr, *act2 = nullptr;
    int *pid = nullptr, *pid2 = nullptr;
    int *pkey = nullptr, *prank = nullptr;
    double *acc = nullptr, *out = nullptr;
    float *out32 = nullptr;
    unsigned *nflag = nullptr;
    int *nlist = nullptr;
    int *ncount = nullptr;
    unsigned *bcount = nullptr, *boff = nullptr;
    unsigned *fcnt = nullptr;  // cell-ordered sort: per (block, octant) key, 64 counters / offsets
    uint8_t *pcell = nullptr;  // cell-ordered sort: a slot's dual cell in its block
    bool cell_sort = false;    // MPM_OPT_CELL_SORT
    std::vector<unsigned *> scan_tmp;  // per level: block sums, their scan
    unsigned *inact_count = nullptr, *inact_off = nullptr;  // inactive slots per sort workgroup
    std::vector<unsigned *> scan_tmp_i;
    unsigned *bflag = nullptr, *bpos = nullptr;  // non-empty particle blocks, their ranks
    std::vector<unsigned *> scan_tmp_b;
    std::vector<long long> scan_n;
    PBlk *plist = nullptr;
    int *npblk = nullptr;
    double *Lout = nullptr;
    bool keep_L = true;  // MPM_OPT_KEEP_L: the last step of a call writes the velocity gradient
    unsigned long long *stats = nullptr;
    unsigned *rare = nullptr, *rare_n = nullptr;  // side list of the split particle passes
    unsigned *work = nullptr;                     // block counter of the particle passes
    unsigned *nonfin = nullptr;                   // non-finite grid output, per list parity
    int grid_nbm = MAXB;  // bodies per simulation the grid kernel is sized for (1, 2 or MAXB)
    bool record = false;      // MPM_OPT_NODE_RECORD
    double *rec = nullptr;    // node record planes (KArgs::rec)
    double *crec = nullptr;   // contact record planes (KArgs::crec)
    bool recorded = false;    // the last grid update wrote the record
    bool grid_valid = false;  // a step ran since the last upload / load / clear (node downloads)
    double *stage = nullptr;  // download staging
    long long stage_elems = 0;
    int par = 0;
    int buf = 0;  // which of the two particle buffers is current (toggled by every sort)
    int sort_interval = 16;
    int pass_grid = 1024;  // persistent workgroups of the G2P and fused particle passes
    int p2g_grid = 1536;   // persistent workgroups of the P2G-only pass (set from occupancy at create)
    int grid_wgs = 2048;   // workgroups of the grid update
    bool autotune = false, tuned = false;
    bool use_graph = false;
    hipGraphExec_t graph[4] = {nullptr, nullptr, nullptr, nullptr};  // per (parity, buffer)
    int graph_len[4] = {0, 0, 0, 0};
    int graph_captures = 0;
    bool sorted = false;  // slots are in particle-block order (cleared by upload)
    int since_sort = 0;   // steps advanced since the last sort
    int fused = 0;  // 1: one particle pass per step (G2P then P2G); 0: two passes (default: faster on gfx950)
    bool sims_dirty = true;
    std::vector<char> uploaded;
    std::vector<long long> count;  // points in use per body (<= its capacity BodyDev::n)
    std::vector<std::array<int, 3>> grid_cap;  // per sim: N at creation (node storage size)
    std::vector<int> contact_cap;              // per sim: contacts at creation
    std::vector<std::vector<double>> x0;  // per body, dim planes (for displacement)
    bool timing = false;
    std::vector<Timer> timers;
    std::vector<Timer> free_timers;
    long long t_launch[NCLASS] = {0, 0, 0, 0, 0};
    double t_ms[NCLASS] = {0, 0, 0, 0, 0};
    unsigned long long last_active_nodes = 0;
};

namespace {

KArgs kargs(mpm_gpu_handle *h) {
    KArgs A{};
    A.sims = h->d_sims;
    A.nsim = (int)h->lo.sims.size();
    A.bodies = h->d_bodies;
    A.nbody = (int)h->lo.bodies.size();
    A.contacts = h->d_contacts;
    A.P = h->P;
    A.np = h->lo.np;
    A.act = h->act;
    A.pid = h->pid;
    A.Pw = h->Pw;
    A.acc = h->acc;
    A.out = h->out;
    A.out32 = h->out32;
    A.nt = h->lo.nt;
    A.nflag = h->nflag;
    A.nlist = h->nlist;
    A.ncount = h->ncount;
    A.nblk_total = h->lo.nblk_total;
    A.plist = h->plist;
    A.npblk = h->npblk;
    A.par = h->par;
    A.Lout = h->Lout;
    A.stats = h->stats;
    A.rare = h->rare;
    A.rare_n = h->rare_n;
    A.work = h->work;
    A.nonfin = h->nonfin;
    A.rec = nullptr;  // set by launch_grid when recording
    A.crec = nullptr;
    return A;
}

int timer_begin(mpm_gpu_handle *h, int cls, Timer &t) {
    if (!h->timing) return MPM_OK;
    if (h->free_timers.empty()) {
        Timer n;
        HIP_TRY(hipEventCreate(&n.a));
        HIP_TRY(hipEventCreate(&n.b));
        h->free_timers.push_back(n);
    }
    t = h->free_timers.back();
    h->free_timers.pop_back();
    t.cls = cls;
    HIP_TRY(hipEventRecord(t.a, h->stream));
    return MPM_OK;
}

int timer_end(mpm_gpu_handle *h, Timer &t) {
    if (!h->timing) return MPM_OK;
    HIP_TRY(hipEventRecord(t.b, h->stream));
    h->timers.push_back(t);
    return MPM_OK;
}

int timers_collect(mpm_gpu_handle *h) {
    for (Timer &t : h->timers) {
        HIP_TRY(hipEventSynchronize(t.b));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, t.a, t.b));
        h->t_launch[t.cls]++;
        h->t_ms[t.cls] += ms;
        h->free_timers.push_back(t);
    }
    h->timers.clear();
    return MPM_OK;
}

template <class T>
int dalloc(T **p, long long n) {
    if (n <= 0) n = 1;
    HIP_TRY(hipMalloc((void **)p, sizeof(T) * n));
    return MPM_OK;
}

void swap_buffers(mpm_gpu_handle *h);

// exclusive prefix sum of n unsigned values, multi-level (1024 threads x 4 per block);
// tmp holds two buffers per level beyond the first (alloc_scan_tmp)
void scan_exclusive(mpm_gpu_handle *h, const unsigned *in, unsigned *outp, long long n,
                    const std::vector<unsigned *> &tmp) {
    std::vector<std::pair<unsigned *, long long>> stack;
    int lvl = 0;
    while (true) {
        const long long nb = ceil_div(n, 4096);
        unsigned *bsum = nb > 1 ? tmp[2 * lvl] : nullptr;
        hipLaunchKernelGGL(k_scan_block, dim3((unsigned)nb), dim3(1024), 0, h->stream, in, outp, bsum, n);
        if (nb <= 1) break;
        stack.push_back({outp, n});
        in = bsum;
        outp = tmp[2 * lvl + 1];
        n = nb;
        lvl++;
    }
    for (int l = (int)stack.size() - 1; l >= 0; l--) {
        const long long nn = stack[l].second;
        hipLaunchKernelGGL(k_scan_add, dim3((unsigned)ceil_div(nn, 4096)), dim3(1024), 0, h->stream, stack[l].first,
                           tmp[2 * l + 1], nn);
    }
}

// workgroups of the cell-ordered sort's per-block kernels (grid-stride over the block list,
// whose length is known only on the device)
unsigned cell_wgs(const mpm_gpu_handle *h) { return (unsigned)std::min(4096ll, std::max(1ll, h->lo.npblk_total)); }

template <int DIM, class PT>
int launch_sort(mpm_gpu_handle *h) {
    const long long np = h->lo.np, nk = h->lo.nkeys;
    KArgs A = kargs(h);
    Timer t{};
    int rc = timer_begin(h, KC_SORT, t);
    if (rc) return rc;
    const int gp = (int)std::max(1ll, ceil_div(np, 256));
    const bool cells = h->cell_sort;
    hipLaunchKernelGGL((k_bin_count<DIM, PT>), dim3(gp), dim3(256), 0, h->stream, A, h->bcount, h->pkey, h->prank,
                       nk - 1, h->inact_count, cells ? h->fcnt : nullptr, h->pcell);
    scan_exclusive(h, h->bcount, h->boff, nk, h->scan_tmp);             // bucket offsets
    scan_exclusive(h, h->inact_count, h->inact_off, gp, h->scan_tmp_i);  // inactive ranks per workgroup
    const long long nblk = h->lo.npblk_total;
    hipLaunchKernelGGL(k_block_flags, dim3((unsigned)ceil_div(nblk + 1, 256)), dim3(256), 0, h->stream, h->boff,
                       h->bflag, nblk, 1 << DIM);
    scan_exclusive(h, h->bflag, h->bpos, nblk + 1, h->scan_tmp_b);
    hipLaunchKernelGGL(k_make_plist, dim3((unsigned)ceil_div(nblk + 1, 256)), dim3(256), 0, h->stream, A, h->boff,
                       h->bpos, h->plist, h->npblk, nblk, 1 << DIM);
    if (cells)
        hipLaunchKernelGGL(k_cell_offsets<DIM>, dim3(cell_wgs(h)), dim3(64 << DIM), 0, h->stream, h->plist, h->npblk,
                           h->boff, h->fcnt);
    hipLaunchKernelGGL(k_clear, dim3((unsigned)ceil_div(nk, 256)), dim3(256), 0, h->stream, h->bcount, nk);
    if constexpr (PT::FP32) {
        hipLaunchKernelGGL((k_bin_scatter<uint32_t, PW<DIM>::N>), dim3(gp), dim3(256), 0, h->stream, np, h->Pw, h->Pw2,
                           h->act, h->act2, h->pid, h->pid2, h->pkey, h->prank, h->boff, (int)(nk - 1),
                           h->inact_off, cells ? h->fcnt : nullptr, h->pcell);
    } else {
        hipLaunchKernelGGL((k_bin_scatter<double, PF<DIM>::N>), dim3(gp), dim3(256), 0, h->stream, np, h->P, h->P2,
                           h->act, h->act2, h->pid, h->pid2, h->pkey, h->prank, h->boff, (int)(nk - 1),
                           h->inact_off, cells ? h->fcnt : nullptr, h->pcell);
    }
    if (cells)
        hipLaunchKernelGGL(k_cell_clear<DIM>, dim3(cell_wgs(h)), dim3(64 << DIM), 0, h->stream, h->plist, h->npblk,
                           h->fcnt);
    swap_buffers(h);
    HIP_TRY(hipGetLastError());
    return timer_end(h, t);
}

// workgroups of the side kernels (grid-stride over a device-side count)
constexpr int RARE_WGS = 256;

template <int DIM, class PT, bool G2P, bool P2G>
void launch_particles(mpm_gpu_handle *h, int want_L) {
    KArgs A = kargs(h);
    const dim3 grid((!G2P && P2G) ? h->p2g_grid : h->pass_grid);
    if (h->lo.any_b1)
        hipLaunchKernelGGL((k_particles<DIM, G2P, P2G, true, PT>), grid, dim3(256), 0, h->stream, A, want_L);
    else
        hipLaunchKernelGGL((k_particles<DIM, G2P, P2G, false, PT>), grid, dim3(256), 0, h->stream, A, want_L);
    // split passes leave the points outside their LDS path to a side kernel (side_pass)
    if (!side_pass<G2P, P2G, PT::FP32>()) return;
    if (P2G && !G2P) {
        if (h->lo.any_b1)
            hipLaunchKernelGGL((k_p2g_side<DIM, true, PT>), dim3(RARE_WGS), dim3(256), 0, h->stream, A);
        else
            hipLaunchKernelGGL((k_p2g_side<DIM, false, PT>), dim3(RARE_WGS), dim3(256), 0, h->stream, A);
    } else if (G2P && !P2G) {
        hipLaunchKernelGGL((k_g2p_side<DIM, PT>), dim3(RARE_WGS), dim3(256), 0, h->stream, A, want_L);
    }
}

template <int DIM, class PT>
void launch_grid(mpm_gpu_handle *h, bool want_stats) {
    KArgs Ag = kargs(h);
    if (want_stats) hipLaunchKernelGGL(k_clear_u64, dim3(1), dim3(64), 0, h->stream, h->stats, 4);
    else Ag.stats = nullptr;
    if (h->record) {
        Ag.rec = h->rec;
        Ag.crec = h->crec;
    }
    h->recorded = h->record;
    using GO = typename PT::GO;
    if (h->grid_nbm == 1) hipLaunchKernelGGL((k_grid<DIM, GO, 1>), dim3(h->grid_wgs), dim3(256), 0, h->stream, Ag);
    else if (h->grid_nbm == 2) hipLaunchKernelGGL((k_grid<DIM, GO, 2>), dim3(h->grid_wgs), dim3(256), 0, h->stream, Ag);
    else hipLaunchKernelGGL((k_grid<DIM, GO, MAXB>), dim3(h->grid_wgs), dim3(256), 0, h->stream, Ag);
}

// sort when the order is stale: after an upload, when the interval has passed, or at the
// start of every call with interval 0
template <int DIM, class PT>
int sort_if_due(mpm_gpu_handle *h, bool first_of_call) {
    const bool due = !h->sorted || (h->sort_interval == 0 && first_of_call) ||
                     (h->sort_interval > 0 && h->since_sort >= h->sort_interval);
    if (!due) return MPM_OK;
    const int rc = launch_sort<DIM, PT>(h);
    if (rc) return rc;
    h->sorted = true;
    h->since_sort = 0;
    return MPM_OK;
}

// One step with separate passes: [sort] P2G(par) grid(par) G2P. P2G marks node blocks
// into list[par]; k_grid consumes it and clears list[par^1] for the next step.
template <int DIM, class PT>
int split_step(mpm_gpu_handle *h, bool first, bool last) {
    Timer t{};
    int rc = sort_if_due<DIM, PT>(h, first);
    if (rc) return rc;
    if ((rc = timer_begin(h, KC_P2G, t))) return rc;
    launch_particles<DIM, PT, false, true>(h, 0);
    if ((rc = timer_end(h, t))) return rc;
    if ((rc = timer_begin(h, KC_GRID, t))) return rc;
    launch_grid<DIM, PT>(h, last);
    if ((rc = timer_end(h, t))) return rc;
    h->par ^= 1;
    if ((rc = timer_begin(h, KC_G2P, t))) return rc;
    launch_particles<DIM, PT, true, false>(h, last ? 1 : 0);
    if ((rc = timer_end(h, t))) return rc;
    h->since_sort++;
    HIP_TRY(hipGetLastError());
    return MPM_OK;
}

// The fused sequence: sort; P2G; { grid; par^=1; [sort]; G2P+P2G } ...; grid; G2P
template <int DIM, class PT>
int fused_steps(mpm_gpu_handle *h, int nsteps) {
    Timer t{};
    int rc = sort_if_due<DIM, PT>(h, true);
    if (rc) return rc;
    if ((rc = timer_begin(h, KC_P2G, t))) return rc;
    launch_particles<DIM, PT, false, true>(h, 0);
    if ((rc = timer_end(h, t))) return rc;
    for (int k = 0; k < nsteps; k++) {
        const bool last = k == nsteps - 1;
        if ((rc = timer_begin(h, KC_GRID, t))) return rc;
        launch_grid<DIM, PT>(h, last);
        if ((rc = timer_end(h, t))) return rc;
        h->par ^= 1;
        h->since_sort++;
        if (last) {
            if ((rc = timer_begin(h, KC_G2P, t))) return rc;
            launch_particles<DIM, PT, true, false>(h, 1);
            if ((rc = timer_end(h, t))) return rc;
        } else {
            if (h->sort_interval > 0 && h->since_sort >= h->sort_interval) {
                if ((rc = launch_sort<DIM, PT>(h))) return rc;
                h->since_sort = 0;
            }
            if ((rc = timer_begin(h, KC_PASS, t))) return rc;
            launch_particles<DIM, PT, true, true>(h, 0);
            if ((rc = timer_end(h, t))) return rc;
        }
        HIP_TRY(hipGetLastError());
    }
    return MPM_OK;
}

void swap_buffers(mpm_gpu_handle *h) {
    std::swap(h->P, h->P2);
    std::swap(h->Pw, h->Pw2);
    std::swap(h->act, h->act2);
    std::swap(h->pid, h->pid2);
    h->buf ^= 1;
}

// One sort period (sort + sort_interval split steps) as a HIP graph. The captured kernels
// bake the list parity and the particle buffer they start from, so a graph is kept per
// (parity, buffer) and replayed only from that state; replaying advances the host-side
// state exactly as capturing did.
template <int DIM, class PT>
int graph_period(mpm_gpu_handle *h) {
    const int S = h->sort_interval;
    const int key = h->par * 2 + h->buf;
    if (h->graph[key] && h->graph_len[key] == S) {
        HIP_TRY(hipGraphLaunch(h->graph[key], h->stream));
        swap_buffers(h);  // the period's sort
        if (S & 1) h->par ^= 1;
        h->since_sort = S;
        return MPM_OK;
    }
    if (h->graph[key]) {
        HIP_TRY(hipGraphExecDestroy(h->graph[key]));
        h->graph[key] = nullptr;
    }
    HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int rc = MPM_OK;
    for (int j = 0; j < S && rc == MPM_OK; j++) rc = split_step<DIM, PT>(h, false, false);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(MPM_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    hipGraphExec_t ex = nullptr;
    const hipError_t e2 = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e2 != hipSuccess) return fail(MPM_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e2));
    h->graph[key] = ex;
    h->graph_len[key] = S;
    h->graph_captures++;
    HIP_TRY(hipGraphLaunch(ex, h->stream));  // host state was advanced while capturing
    return MPM_OK;
}

// Autotuner of the tile and bin sizes: times whole steps (HIP events) for candidate
// persistent-grid sizes of the G2P (and fused) pass, then of the P2G pass, then of the grid
// kernel, keeping the fastest of each, and then the sort interval (steps per binning of the
// points into particle blocks). The steps it times are real steps of the run (results do not
// depend on launch shape or sort interval beyond atomic summation order). Returns the
// number of steps consumed.
template <int DIM, class PT>
int autotune(mpm_gpu_handle *h, int budget, int *used) {
    *used = 0;
    static const int pass_cand[] = {512, 1024, 2048, 4096, 8192};
    static const int p2g_cand[] = {768, 1024, 1536, 2048, 4096};
    static const int grid_cand[] = {512, 2048, 8192};
    const int reps = 2;
    const int need = reps * (5 + 5 + 3) + 1;
    if (budget < need) return MPM_OK;  // too short a call to tune; tune on a later one
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    const bool timing = h->timing;
    h->timing = false;
    int rc = MPM_OK;
    auto measure = [&](float *ms) -> int {
        if (hipEventRecord(e0, h->stream) != hipSuccess) return MPM_ERR_HIP;
        for (int r = 0; r < reps; r++) {
            const int rr = split_step<DIM, PT>(h, *used == 0, false);
            if (rr) return rr;
            (*used)++;
        }
        if (hipEventRecord(e1, h->stream) != hipSuccess) return MPM_ERR_HIP;
        if (hipEventSynchronize(e1) != hipSuccess) return MPM_ERR_HIP;
        return hipEventElapsedTime(ms, e0, e1) == hipSuccess ? MPM_OK : MPM_ERR_HIP;
    };
    float best = 1e30f, ms = 0;
    int best_pass = h->pass_grid, best_p2g = h->p2g_grid, best_grid = h->grid_wgs;
    for (int c : pass_cand) {
        h->pass_grid = c;
        if ((rc = measure(&ms))) break;
        if (ms < best) { best = ms; best_pass = c; }
    }
    h->pass_grid = best_pass;
    best = 1e30f;
    for (int c : p2g_cand) {
        if (rc) break;
        h->p2g_grid = c;
        if ((rc = measure(&ms))) break;
        if (ms < best) { best = ms; best_p2g = c; }
    }
    h->p2g_grid = best_p2g;
    best = 1e30f;
    for (int c : grid_cand) {
        if (rc) break;
        h->grid_wgs = c;
        if ((rc = measure(&ms))) break;
        if (ms < best) { best = ms; best_grid = c; }
    }
    h->grid_wgs = best_grid;
    // the sort interval (how many steps the particle-block bins serve before points are
    // re-binned): one sort period of each candidate, its time per step, when the call has the
    // steps to spare; the sort at the start of each period is forced. (A 64-step candidate
    // measured no better on the 16M column and picked, on the 2D ensemble, an interval that
    // was 8% slower over the whole run: points move faster later than in the tuned period.)
    static const int sort_cand[] = {8, 16, 32};
    if (!rc && h->sort_interval > 0 && budget - *used >= 8 + 16 + 32 + 1) {
        float best_ps = 1e30f;
        int best_si = h->sort_interval;
        for (int c : sort_cand) {
            h->sort_interval = c;
            h->since_sort = c;
            if (hipEventRecord(e0, h->stream) != hipSuccess) { rc = MPM_ERR_HIP; break; }
            for (int r = 0; r < c && !rc; r++) {
                rc = split_step<DIM, PT>(h, false, false);
                (*used)++;
            }
            if (rc) break;
            if (hipEventRecord(e1, h->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                rc = MPM_ERR_HIP;
                break;
            }
            if (ms / c < best_ps) {
                best_ps = ms / c;
                best_si = c;
            }
        }
        h->sort_interval = best_si;
    }
    h->tuned = true;
    h->timing = timing;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return fail(rc, "autotune: a timed step failed");
    return MPM_OK;
}

template <int DIM, class PT>
int launch_step(mpm_gpu_handle *h, int nsteps) {
    if (h->fused) return fused_steps<DIM, PT>(h, nsteps);
    int k = 0, rc = MPM_OK;
    if (h->autotune && !h->tuned) {
        int used = 0;
        if ((rc = autotune<DIM, PT>(h, nsteps - 1, &used))) return rc;
        k = used;
    }
    // (no graph replay while recording: the record pointers would be baked into the graph)
    const bool graphs = h->use_graph && !h->timing && h->sort_interval > 0 && !h->record;
    while (k < nsteps) {
        const bool last = k == nsteps - 1;
        if (graphs && h->sorted && h->since_sort >= h->sort_interval && nsteps - 1 - k >= h->sort_interval) {
            if ((rc = graph_period<DIM, PT>(h))) return rc;
            k += h->sort_interval;
            continue;
        }
        if ((rc = split_step<DIM, PT>(h, k == 0, last))) return rc;
        k++;
    }
    return MPM_OK;
}

// Workgroups of the P2G pass the device holds at once (its occupancy per CU x CUs): one
// resident round of persistent workgroups. Measured on the 16M column: a grid of 4/3 of
// the resident count costs 15-25% over one or two full rounds. 0 when a query fails.
template <int DIM, class PT>
int p2g_resident(bool nrm) {
    int per_cu = 0, dev = 0, cus = 0;
    const void *k = nrm ? reinterpret_cast<const void *>(&k_particles<DIM, false, true, true, PT>)
                        : reinterpret_cast<const void *>(&k_particles<DIM, false, true, false, PT>);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 256, 0) != hipSuccess) return 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    return per_cu * cus;
}

int create_from_layout(Layout &&lo, int precision, mpm_gpu_handle **out) {
    if (precision != MPM_PREC_FP64 && precision != MPM_PREC_FP32)
        return fail(MPM_ERR_ARG, "unknown precision form %d", precision);
    if (precision == MPM_PREC_FP32) {
        // the packed cell index holds cells -1 .. 2^bits - 2 per dimension
        for (const SimDev &s : lo.sims)
            for (int d = 0; d < lo.dim; d++) {
                const int bits = lo.dim == 3 ? CellBits<3>::B[d] : CellBits<2>::B[d];
                if (s.N[d] + 2 > (1 << bits))
                    return fail(MPM_ERR_ARG, "precision form: N[%d] = %d exceeds %d cells", d, s.N[d], (1 << bits) - 2);
            }
    }
    mpm_gpu_handle *h = new mpm_gpu_handle();
    h->lo = std::move(lo);
    h->precision = precision;
    auto cleanup = [&](int rc) {
        mpm_gpu_destroy(h);
        return rc;
    };
    int rc;
#define TRY(x)                   \
    do {                         \
        rc = (x);                \
        if (rc) return cleanup(rc); \
    } while (0)
    hipError_t e = hipGetDevice(&h->device);
    if (e != hipSuccess) return cleanup(fail(MPM_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(e)));
    e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return cleanup(fail(MPM_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e)));
    const Layout &L = h->lo;
    const int dim = L.dim;
    const bool f32 = precision == MPM_PREC_FP32;
    {
        // no later load_sim gives a simulation more bodies than it was created with
        int mx = 1;
        for (const SimDev &sd : L.sims) mx = std::max(mx, sd.nbody);
        h->grid_nbm = mx <= 1 ? 1 : (mx <= 2 ? 2 : MAXB);
    }
    {
        const int r = dim == 3 ? (f32 ? p2g_resident<3, Prec32>(L.any_b1) : p2g_resident<3, Prec64>(L.any_b1))
                               : (f32 ? p2g_resident<2, Prec32>(L.any_b1) : p2g_resident<2, Prec64>(L.any_b1));
        if (r > 0) h->p2g_grid = r;
        (void)hipGetLastError();  // a failed query leaves the default and no pending error
    }
    const int nf = f32 ? (dim == 3 ? PW<3>::N : PW<2>::N) : (dim == 3 ? PF<3>::N : PF<2>::N);
    const int nacc = dim == 3 ? GF<3>::NACC : GF<2>::NACC;
    const int nout = dim == 3 ? GF<3>::NOUT : GF<2>::NOUT;
    TRY(dalloc(&h->d_sims, (long long)L.sims.size()));
    TRY(dalloc(&h->d_bodies, (long long)L.bodies.size()));
    TRY(dalloc(&h->d_contacts, std::max<long long>(1, (long long)L.contacts.size())));
    if (f32) {
        TRY(dalloc(&h->Pw, L.np * nf));
        TRY(dalloc(&h->Pw2, L.np * nf));
        TRY(dalloc(&h->out32, L.nt * nout));
    } else {
        TRY(dalloc(&h->P, L.np * nf));
        TRY(dalloc(&h->P2, L.np * nf));
        TRY(dalloc(&h->out, L.nt * nout));
    }
    TRY(dalloc(&h->act, L.np));
    TRY(dalloc(&h->act2, L.np));
    TRY(dalloc(&h->pid, L.np));
    TRY(dalloc(&h->pid2, L.np));
    TRY(dalloc(&h->pkey, L.np));
    TRY(dalloc(&h->prank, L.np));
    TRY(dalloc(&h->acc, L.nt * nacc));
    TRY(dalloc(&h->nflag, L.nblk_total));
    TRY(dalloc(&h->nlist, 2 * L.nblk_total));
    TRY(dalloc(&h->ncount, 2));
    TRY(dalloc(&h->bcount, L.nkeys));
    TRY(dalloc(&h->boff, L.nkeys));
    if (L.cell_fit) {
        TRY(dalloc(&h->fcnt, (L.nkeys - 1) * 64));
        TRY(dalloc(&h->pcell, L.np));
        h->cell_sort = true;
    }
    TRY(dalloc(&h->plist, L.npblk_total));
    TRY(dalloc(&h->npblk, 1));
    TRY(dalloc(&h->stats, 4));
    TRY(dalloc(&h->rare, L.np));
    TRY(dalloc(&h->rare_n, 4));
    TRY(dalloc(&h->work, 2));
    TRY(dalloc(&h->nonfin, 2));
    auto alloc_scan_tmp = [&](long long n, std::vector<unsigned *> &tmp) -> int {
        for (; n > 4096;) {
            const long long nb = ceil_div(n, 4096);
            unsigned *a = nullptr, *b = nullptr;
            int r = dalloc(&a, nb);
            if (!r) r = dalloc(&b, nb);
            tmp.push_back(a);
            tmp.push_back(b);
            if (r) return r;
            n = nb;
        }
        return MPM_OK;
    };
    TRY(alloc_scan_tmp(L.nkeys, h->scan_tmp));
    const long long gp_sort = std::max(1ll, ceil_div(L.np, 256));
    TRY(dalloc(&h->inact_count, gp_sort));
    TRY(dalloc(&h->inact_off, gp_sort));
    TRY(alloc_scan_tmp(gp_sort, h->scan_tmp_i));
    TRY(dalloc(&h->bflag, L.npblk_total + 1));
    TRY(dalloc(&h->bpos, L.npblk_total + 1));
    TRY(alloc_scan_tmp(L.npblk_total + 1, h->scan_tmp_b));
    // zero state: accumulators must start at zero, flags clear, counters zero
    e = hipMemsetAsync(h->acc, 0, sizeof(double) * L.nt * nacc, h->stream);
    if (e == hipSuccess)
        e = f32 ? hipMemsetAsync(h->out32, 0, sizeof(float) * L.nt * nout, h->stream)
                : hipMemsetAsync(h->out, 0, sizeof(double) * L.nt * nout, h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(h->nflag, 0, sizeof(unsigned) * std::max(1ll, L.nblk_total), h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(h->ncount, 0, sizeof(int) * 2, h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(h->bcount, 0, sizeof(unsigned) * L.nkeys, h->stream);
    if (e == hipSuccess && h->fcnt)
        e = hipMemsetAsync(h->fcnt, 0, sizeof(unsigned) * std::max(1ll, (L.nkeys - 1) * 64), h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(h->npblk, 0, sizeof(int), h->stream);
    if (e == hipSuccess)
        e = f32 ? hipMemsetAsync(h->Pw, 0, sizeof(uint32_t) * std::max(1ll, L.np * nf), h->stream)
                : hipMemsetAsync(h->P, 0, sizeof(double) * std::max(1ll, L.np * nf), h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(h->act, 0, std::max(1ll, L.np), h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(h->stats, 0, sizeof(unsigned long long) * 4, h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(h->rare_n, 0, sizeof(unsigned) * 4, h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(h->work, 0, sizeof(unsigned) * 2, h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(h->nonfin, 0, sizeof(unsigned) * 2, h->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(h->d_sims, L.sims.data(), sizeof(SimDev) * L.sims.size(), hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(h->d_bodies, L.bodies.data(), sizeof(BodyDev) * L.bodies.size(), hipMemcpyHostToDevice,
                           h->stream);
    if (e == hipSuccess && !L.contacts.empty())
        e = hipMemcpyAsync(h->d_contacts, L.contacts.data(), sizeof(ContactDev) * L.contacts.size(),
                           hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) {
        std::vector<int> ids(std::max(1ll, L.np));
        for (long long i = 0; i < L.np; i++) ids[i] = (int)i;
        e = hipMemcpyAsync(h->pid, ids.data(), sizeof(int) * L.np, hipMemcpyHostToDevice, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    }
    if (e != hipSuccess) return cleanup(fail(MPM_ERR_HIP, "initialising device state: %s", hipGetErrorString(e)));
    h->uploaded.assign(L.bodies.size(), 0);
    for (const SimDev &sd : L.sims) {
        h->grid_cap.push_back({sd.N[0], sd.N[1], sd.N[2]});
        h->contact_cap.push_back(sd.ncontact);
    }
    h->count.resize(L.bodies.size());
    for (size_t b = 0; b < L.bodies.size(); b++) h->count[b] = L.bodies[b].n;
    h->x0.resize(L.bodies.size());
    h->sims_dirty = true;
#undef TRY
    *out = h;
    return MPM_OK;
}

}  // namespace

extern "C" {

const char *mpm_gpu_last_error(void) { return g_err.c_str(); }

int mpm_gpu_create_batch(const mpm_sim_desc *sims, int nsim, int precision, mpm_gpu_handle **out) {
    TraceRange tr_("mpm_gpu_create_batch");
    if (!out) return fail(MPM_ERR_ARG, "out is NULL");
    *out = nullptr;
    Layout lo;
    int rc = build_layout(sims, nsim, lo);
    if (rc) return rc;
    return create_from_layout(std::move(lo), precision, out);
}

int mpm_gpu_create(const mpm_grid_desc *grid, const mpm_body_desc *bodies, int nbody,
                   const mpm_contact_desc *contacts, int ncontact, int precision, mpm_gpu_handle **out) {
    if (!grid || !bodies) return fail(MPM_ERR_ARG, "grid and bodies are required");
    mpm_sim_desc s{};
    s.grid = *grid;
    s.nbody = nbody;
    s.bodies = bodies;
    s.ncontact = ncontact;
    s.contacts = contacts;
    return mpm_gpu_create_batch(&s, 1, precision, out);
}

int64_t mpm_gpu_footprint(const mpm_sim_desc *sims, int nsim, int precision) {
    Layout lo;
    if (build_layout(sims, nsim, lo)) return -1;
    return footprint_bytes(lo, precision);
}

static int ensure_stage(mpm_gpu_handle *h, long long elems);

int mpm_gpu_upload(mpm_gpu_handle *h, int body, const double *x, const double *v, const double *T,
                   const double *V, const double *m, const double *extent, const int32_t *active) {
    TraceRange tr_("mpm_gpu_upload");
    if (!h) return fail(MPM_ERR_ARG, "handle is NULL");
    if (body < 0 || body >= (int)h->lo.bodies.size()) return fail(MPM_ERR_ARG, "body %d out of range", body);
    const BodyDev &B = h->lo.bodies[body];
    const int dim = h->lo.dim;
    const long long n = h->count[body], np = h->lo.np;
    if (n > 0 && (!x || !v || !V || !m || !extent || !active))
        return fail(MPM_ERR_ARG, "x, v, V, m, extent, active are required");
    // The body's old points are released (unused capacity, pid -1) and the new ones are
    // staged at slots [0, n) of the second buffer, then placed in the unused slots of the
    // current layout in slot order; the sort before the next step puts them in block order.
    // Other bodies hold at most their capacity, so at least B.n >= n slots are unused.
    const bool f32 = h->precision == MPM_PREC_FP32;
    const int nf = f32 ? (dim == 3 ? PW<3>::N : PW<2>::N) : (dim == 3 ? PF<3>::N : PF<2>::N);
    const int gp = (int)std::max(1ll, ceil_div(np, 256));
    HIP_TRY(hipStreamSynchronize(h->stream));
    // From here on the body's old points are gone: until this upload completes, the body
    // counts as not uploaded (a step reports MPM_ERR_STATE instead of running without it)
    // and the layout as unsorted.
    h->uploaded[body] = 0;
    h->sorted = false;
    h->grid_valid = false;
    if (B.n > 0) {
        hipLaunchKernelGGL(k_deactivate, dim3(gp), dim3(256), 0, h->stream, np, h->act, h->pid, B.p_off,
                           B.p_off + B.n);
        HIP_TRY(hipGetLastError());
    }
    if (n > 0) {
        {
            std::vector<int8_t> a8(n);
            for (long long i = 0; i < n; i++) a8[i] = active[i] ? 1 : 0;
            HIP_TRY(hipMemcpyAsync(h->act2, a8.data(), n, hipMemcpyHostToDevice, h->stream));
            HIP_TRY(hipStreamSynchronize(h->stream));
        }
        // the body's fields as planes of n doubles in h->stage (x[dim] v[dim] T[NTS] V m ext),
        // then into slots [0, n) of the second buffer in the form's representation
        const int NTS = dim == 3 ? 6 : 4;
        const int nsf = 2 * dim + NTS + 3;
        int rc = ensure_stage(h, (long long)nsf * n);
        if (rc) return rc;
        double *st = h->stage;
        auto put = [&](int plane, int count, const double *src) -> hipError_t {
            return hipMemcpyAsync(st + (long long)plane * n, src, sizeof(double) * n * count, hipMemcpyHostToDevice,
                                  h->stream);
        };
        HIP_TRY(put(0, dim, x));
        HIP_TRY(put(dim, dim, v));
        if (T) {
            HIP_TRY(put(2 * dim, NTS, T));  // 6 planes in (xx, yy, zz, xy, xz, yz); 2D keeps the first four
        } else {
            const long long nz = (long long)NTS * n;
            hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)ceil_div(nz, 256)), dim3(256), 0, h->stream,
                               st + 2ll * dim * n, nz, 0.0);
        }
        HIP_TRY(put(2 * dim + NTS, 1, V));
        HIP_TRY(put(2 * dim + NTS + 1, 1, m));
        HIP_TRY(put(2 * dim + NTS + 2, 1, extent));
        const unsigned gi = (unsigned)ceil_div(n, 256);
        if (dim == 3) {
            if (f32) hipLaunchKernelGGL((k_import<3, Prec32, uint32_t>), dim3(gi), dim3(256), 0, h->stream, st, n, h->lo.sims[B.sim], h->Pw2, np);
            else hipLaunchKernelGGL((k_import<3, Prec64, double>), dim3(gi), dim3(256), 0, h->stream, st, n, h->lo.sims[B.sim], h->P2, np);
        } else {
            if (f32) hipLaunchKernelGGL((k_import<2, Prec32, uint32_t>), dim3(gi), dim3(256), 0, h->stream, st, n, h->lo.sims[B.sim], h->Pw2, np);
            else hipLaunchKernelGGL((k_import<2, Prec64, double>), dim3(gi), dim3(256), 0, h->stream, st, n, h->lo.sims[B.sim], h->P2, np);
        }
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_free_count, dim3(gp), dim3(256), 0, h->stream, np, h->pid, h->inact_count);
        scan_exclusive(h, h->inact_count, h->inact_off, gp, h->scan_tmp_i);
        if (f32)
            hipLaunchKernelGGL(k_place<uint32_t>, dim3(gp), dim3(256), 0, h->stream, np, nf, h->Pw2, h->Pw, h->act2,
                               h->act, h->pid, h->inact_off, n, B.p_off);
        else
            hipLaunchKernelGGL(k_place<double>, dim3(gp), dim3(256), 0, h->stream, np, nf, h->P2, h->P, h->act2,
                               h->act, h->pid, h->inact_off, n, B.p_off);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->x0[body].assign(x, x + (long long)dim * n);
    h->uploaded[body] = 1;
    h->grid_valid = false;
    h->sorted = false;
    return MPM_OK;
}

static int deactivate_body(mpm_gpu_handle *h, int b) {
    const BodyDev &B = h->lo.bodies[b];
    if (B.n == 0) return MPM_OK;
    const long long np = h->lo.np;
    hipLaunchKernelGGL(k_deactivate, dim3((unsigned)ceil_div(np, 256)), dim3(256), 0, h->stream, np, h->act, h->pid,
                       B.p_off, B.p_off + B.n);
    HIP_TRY(hipGetLastError());
    return MPM_OK;
}

int mpm_gpu_load_sim(mpm_gpu_handle *h, int sim, const mpm_sim_desc *d) {
    TraceRange tr_("mpm_gpu_load_sim");
    if (!h || !d) return fail(MPM_ERR_ARG, "handle and descriptor are required");
    if (sim < 0 || sim >= (int)h->lo.sims.size()) return fail(MPM_ERR_ARG, "sim %d out of range", sim);
    SimDev &s = h->lo.sims[sim];
    const mpm_grid_desc &g = d->grid;
    if (g.dim != h->lo.dim) return fail(MPM_ERR_ARG, "dim must match the handle (%d)", h->lo.dim);
    if (d->nbody < 1 || d->nbody > s.nbody) return fail(MPM_ERR_ARG, "sim %d holds 1..%d bodies", sim, s.nbody);
    if (d->ncontact < 0 || d->ncontact > h->contact_cap[sim])
        return fail(MPM_ERR_ARG, "sim %d holds at most %d contacts", sim, h->contact_cap[sim]);
    if (g.cpdi_spec != 0 && g.cpdi_spec != 1) return fail(MPM_ERR_ARG, "Unrecognized Argument for use_cpdi: %d", g.cpdi_spec);
    if (g.contact_spec != 0 && g.contact_spec != 1)
        return fail(MPM_ERR_ARG, "ERROR: Unknown SPEC in contact: %d", g.contact_spec);
    for (int k = 0; k < g.dim; k++)
        if (!(g.L[k] > 0) || g.N[k] < 1 || g.N[k] > h->grid_cap[sim][k])
            return fail(MPM_ERR_ARG, "sim %d: need L > 0 and 1 <= N <= %d in dimension %d", sim, h->grid_cap[sim][k], k);
    for (int b = 0; b < d->nbody; b++) {
        if (d->bodies[b].n_points < 0 || d->bodies[b].n_points > h->lo.bodies[s.body0 + b].n)
            return fail(MPM_ERR_ARG, "sim %d body %d: at most %lld points", sim, b, (long long)h->lo.bodies[s.body0 + b].n);
        if (d->bodies[b].has_boundary < 0 || d->bodies[b].has_boundary > 3)
            return fail(MPM_ERR_ARG, "sim %d body %d: has_boundary must be 0..3, got %d", sim, b, d->bodies[b].has_boundary);
        if (int rc = check_material(d->bodies[b], sim, b)) return rc;
    }
    for (int c = 0; c < d->ncontact; c++) {
        const mpm_contact_desc &cd = d->contacts[c];
        if (cd.body1 < 0 || cd.body1 >= d->nbody || cd.body2 < 0 || cd.body2 >= d->nbody || cd.body1 == cd.body2)
            return fail(MPM_ERR_ARG, "contact %d: bodies must be two distinct sim-local indices", c);
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    // grid geometry; the node and particle-block storage keeps its creation size
    s.element_count = 1;
    for (int k = 0; k < g.dim; k++) {
        s.N[k] = g.N[k];
        s.L[k] = g.L[k];
        s.h[k] = g.L[k] / g.N[k];
        s.invh[k] = 1.0 / s.h[k];
        s.NoverL[k] = (double)g.N[k] / g.L[k];
        s.ftol[k] = std::ldexp((double)g.N[k], -48);
        s.hf[k] = (float)s.h[k];
        s.invhf[k] = (float)s.invh[k];
        s.element_count *= g.N[k];
    }
    for (int k = 0; k < 3; k++) s.g[k] = k < g.dim ? g.gravity[k] : 0.0;
    s.cpdi_spec = g.cpdi_spec;
    s.contact_spec = g.contact_spec;
    s.ncontact = d->ncontact;
    for (int b = 0; b < s.nbody; b++) {
        const int gb = s.body0 + b;
        BodyDev &B = h->lo.bodies[gb];
        int rc = deactivate_body(h, gb);
        if (rc) return rc;
        B.is_b1 = 0;
        if (b < d->nbody) {
            const mpm_body_desc &bd = d->bodies[b];
            B.has_boundary = bd.has_boundary;
            B.bc_mu = bd.bc_mu;
            set_material(B, bd);
            h->count[gb] = bd.n_points;
        } else {
            h->count[gb] = 0;
        }
        h->uploaded[gb] = h->count[gb] == 0;
        h->grid_valid = false;
        h->x0[gb].clear();
    }
    for (int c = 0; c < d->ncontact; c++) {
        ContactDev &C = h->lo.contacts[s.contact0 + c];  // the slot keeps its record planes (C.rec)
        C.b1 = d->contacts[c].body1;
        C.b2 = d->contacts[c].body2;
        C.mu = d->contacts[c].mu;
        h->lo.bodies[s.body0 + d->contacts[c].body1].is_b1 = 1;
    }
    HIP_TRY(hipMemcpyAsync(h->d_bodies, h->lo.bodies.data(), sizeof(BodyDev) * h->lo.bodies.size(),
                           hipMemcpyHostToDevice, h->stream));
    if (!h->lo.contacts.empty())
        HIP_TRY(hipMemcpyAsync(h->d_contacts, h->lo.contacts.data(), sizeof(ContactDev) * h->lo.contacts.size(),
                               hipMemcpyHostToDevice, h->stream));
    // the device copy of the descriptors follows at once, so that every kernel before the next
    // step (connectivity, downloads) sees the new grid
    HIP_TRY(hipMemcpyAsync(h->d_sims, h->lo.sims.data(), sizeof(SimDev) * h->lo.sims.size(), hipMemcpyHostToDevice,
                           h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    h->sims_dirty = true;
    return MPM_OK;
}

int mpm_gpu_clear_sim(mpm_gpu_handle *h, int sim) {
    TraceRange tr_("mpm_gpu_clear_sim");
    if (!h) return fail(MPM_ERR_ARG, "handle is NULL");
    if (sim < 0 || sim >= (int)h->lo.sims.size()) return fail(MPM_ERR_ARG, "sim %d out of range", sim);
    SimDev &s = h->lo.sims[sim];
    HIP_TRY(hipStreamSynchronize(h->stream));
    for (int b = 0; b < s.nbody; b++) {
        int rc = deactivate_body(h, s.body0 + b);
        if (rc) return rc;
        h->count[s.body0 + b] = 0;
        h->uploaded[s.body0 + b] = 1;
    }
    h->grid_valid = false;
    s.ncontact = 0;
    h->sims_dirty = true;
    HIP_TRY(hipStreamSynchronize(h->stream));
    return MPM_OK;
}

int mpm_gpu_set_sort_interval(mpm_gpu_handle *h, int interval) {
    if (!h) return fail(MPM_ERR_ARG, "handle is NULL");
    if (interval < 0) return fail(MPM_ERR_ARG, "interval must be >= 0");
    h->sort_interval = interval;
    return MPM_OK;
}

int mpm_gpu_set_option(mpm_gpu_handle *h, int key, int value) {
    if (!h) return fail(MPM_ERR_ARG, "handle is NULL");
    switch (key) {
        case MPM_OPT_SORT_INTERVAL:
            if (value < 0) return fail(MPM_ERR_ARG, "sort interval must be >= 0");
            h->sort_interval = value;
            return MPM_OK;
        case MPM_OPT_FUSED_PASS:
            h->fused = value != 0;
            return MPM_OK;
        case MPM_OPT_GRAPH:
            h->use_graph = value != 0;
            return MPM_OK;
        case MPM_OPT_AUTOTUNE:
            h->autotune = value != 0;
            h->tuned = false;
            return MPM_OK;
        case MPM_OPT_CELL_SORT:
            if (value != 0 && !h->fcnt)
                return fail(MPM_ERR_ARG, "the cell-ordered sort's counters exceed their budget on this handle");
            if ((value != 0) != h->cell_sort) {
                // captured sort periods bake the sort's launches
                HIP_TRY(hipStreamSynchronize(h->stream));
                for (hipGraphExec_t &g : h->graph)
                    if (g) {
                        HIP_TRY(hipGraphExecDestroy(g));
                        g = nullptr;
                    }
            }
            h->cell_sort = value != 0;
            return MPM_OK;
        case MPM_OPT_NODE_RECORD:
            if (value != 0 && !h->rec) {
                const long long nr = (1LL + 2 * h->lo.dim) * std::max(1ll, h->lo.nt);
                HIP_TRY(hipMalloc(&h->rec, sizeof(double) * nr));
                HIP_TRY(hipMemsetAsync(h->rec, 0, sizeof(double) * nr, h->stream));
                if (h->lo.crec_total > 0) {
                    HIP_TRY(hipMalloc(&h->crec, sizeof(double) * h->lo.crec_total));
                    HIP_TRY(hipMemsetAsync(h->crec, 0, sizeof(double) * h->lo.crec_total, h->stream));
                }
                HIP_TRY(hipStreamSynchronize(h->stream));
            }
            h->record = value != 0;
            return MPM_OK;
        case MPM_OPT_KEEP_L:
            h->keep_L = value != 0;
            if (!h->keep_L && h->Lout) {
                HIP_TRY(hipStreamSynchronize(h->stream));
                HIP_TRY(hipFree(h->Lout));
                h->Lout = nullptr;
            }
            return MPM_OK;
        case MPM_OPT_PASS_GRID:
        case MPM_OPT_P2G_GRID:
        case MPM_OPT_GRID_WGS:
            if (value < 1 || value > (1 << 20)) return fail(MPM_ERR_ARG, "workgroup count out of range");
            if (key == MPM_OPT_PASS_GRID) h->pass_grid = h->p2g_grid = value;
            else if (key == MPM_OPT_P2G_GRID) h->p2g_grid = value;
            else h->grid_wgs = value;
            return MPM_OK;
        default:
            return fail(MPM_ERR_ARG, "unknown option %d", key);
    }
}

int mpm_gpu_get_option(mpm_gpu_handle *h, int key, int *value) {
    if (!h || !value) return fail(MPM_ERR_ARG, "handle and value are required");
    switch (key) {
        case MPM_OPT_SORT_INTERVAL: *value = h->sort_interval; return MPM_OK;
        case MPM_OPT_FUSED_PASS: *value = h->fused; return MPM_OK;
        case MPM_OPT_GRAPH: *value = h->use_graph; return MPM_OK;
        case MPM_OPT_AUTOTUNE: *value = h->autotune ? (h->tuned ? 2 : 1) : 0; return MPM_OK;
        case MPM_OPT_CELL_SORT: *value = h->cell_sort; return MPM_OK;
        case MPM_OPT_PASS_GRID: *value = h->pass_grid; return MPM_OK;
        case MPM_OPT_P2G_GRID: *value = h->p2g_grid; return MPM_OK;
        case MPM_OPT_NODE_RECORD: *value = h->record; return MPM_OK;
        case MPM_OPT_GRID_WGS: *value = h->grid_wgs; return MPM_OK;
        case MPM_OPT_GRAPH_CAPTURES: *value = h->graph_captures; return MPM_OK;
        case MPM_OPT_KEEP_L: *value = h->keep_L; return MPM_OK;
        default: return fail(MPM_ERR_ARG, "unknown option %d", key);
    }
}

static int run_steps(mpm_gpu_handle *h, const double *dt, int nsim, int nsteps);

int mpm_gpu_step(mpm_gpu_handle *h, double dt, int nsteps) {
    if (!h) return fail(MPM_ERR_ARG, "handle is NULL");
    std::vector<double> dts(h->lo.sims.size(), dt);
    return run_steps(h, dts.data(), (int)dts.size(), nsteps);
}

int mpm_gpu_step_batch(mpm_gpu_handle *h, const double *dt, int nsim, int nsteps) {
    if (!h) return fail(MPM_ERR_ARG, "handle is NULL");
    if (!dt) return fail(MPM_ERR_ARG, "dt is NULL");
    return run_steps(h, dt, nsim, nsteps);
}

static int run_steps(mpm_gpu_handle *h, const double *dt, int nsim, int nsteps) {
    TraceRange tr_("mpm_gpu_step");
    if (nsteps < 0) return fail(MPM_ERR_ARG, "nsteps must be >= 0");
    if (nsim != (int)h->lo.sims.size()) return fail(MPM_ERR_ARG, "need %zu time steps, got %d", h->lo.sims.size(), nsim);
    for (int i = 0; i < nsim; i++)
        if (!(dt[i] > 0) || !std::isfinite(dt[i])) return fail(MPM_ERR_ARG, "dt must be positive and finite");
    for (size_t b = 0; b < h->uploaded.size(); b++)
        if (!h->uploaded[b] && h->count[b] > 0) return fail(MPM_ERR_STATE, "body %zu has not been uploaded", b);
    if (nsteps == 0) return MPM_OK;
    bool changed = h->sims_dirty;
    for (int i = 0; i < nsim; i++) {
        SimDev &s = h->lo.sims[i];
        if (s.dt != dt[i]) {
            s.dt = dt[i];
            changed = true;
        }
    }
    if (changed) {
        HIP_TRY(hipMemcpyAsync(h->d_sims, h->lo.sims.data(), sizeof(SimDev) * h->lo.sims.size(), hipMemcpyHostToDevice,
                               h->stream));
        h->sims_dirty = false;
    }
    if (!h->Lout && h->keep_L) {
        HIP_TRY(hipMalloc(&h->Lout, sizeof(double) * 9 * std::max(1ll, h->lo.np)));
        HIP_TRY(hipMemsetAsync(h->Lout, 0, sizeof(double) * 9 * std::max(1ll, h->lo.np), h->stream));
    }
    const bool f32 = h->precision == MPM_PREC_FP32;
    const int rc = h->lo.dim == 3 ? (f32 ? launch_step<3, Prec32>(h, nsteps) : launch_step<3, Prec64>(h, nsteps))
                                  : (f32 ? launch_step<2, Prec32>(h, nsteps) : launch_step<2, Prec64>(h, nsteps));
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (h->timing) {
        const int r2 = timers_collect(h);
        if (r2) return r2;
    }
    HIP_TRY(hipMemcpy(&h->last_active_nodes, h->stats, sizeof(unsigned long long), hipMemcpyDeviceToHost));
    h->grid_valid = true;
    return MPM_OK;
}

static int ensure_stage(mpm_gpu_handle *h, long long elems) {
    if (h->stage_elems >= elems) return MPM_OK;
    if (h->stage) HIP_TRY(hipFree(h->stage));
    h->stage = nullptr;
    HIP_TRY(hipMalloc(&h->stage, sizeof(double) * elems));
    h->stage_elems = elems;
    return MPM_OK;
}

int mpm_gpu_download(mpm_gpu_handle *h, int body, double *x, double *v, double *T, double *V, int32_t *active) {
    TraceRange tr_("mpm_gpu_download");
    if (!h) return fail(MPM_ERR_ARG, "handle is NULL");
    if (body < 0 || body >= (int)h->lo.bodies.size()) return fail(MPM_ERR_ARG, "body %d out of range", body);
    const BodyDev &B = h->lo.bodies[body];
    const int dim = h->lo.dim;
    const long long n = h->count[body], np = h->lo.np;
    if (n == 0) return MPM_OK;
    // one pass over the slots: the body's points write their state in original order into
    // planes of n doubles (x, v, T, V) followed by the activity flags
    const long long nd = (2ll * dim + 7) * n;
    int rc = ensure_stage(h, nd + (n + 1) / 2);
    if (rc) return rc;
    KArgs A = kargs(h);
    const unsigned gp = (unsigned)ceil_div(np, 256);
    const bool f32 = h->precision == MPM_PREC_FP32;
    if (dim == 3) {
        if (f32) hipLaunchKernelGGL((k_export<3, Prec32>), dim3(gp), dim3(256), 0, h->stream, A, B.p_off, n, h->lo.sims[B.sim], h->stage);
        else hipLaunchKernelGGL((k_export<3, Prec64>), dim3(gp), dim3(256), 0, h->stream, A, B.p_off, n, h->lo.sims[B.sim], h->stage);
    } else {
        if (f32) hipLaunchKernelGGL((k_export<2, Prec32>), dim3(gp), dim3(256), 0, h->stream, A, B.p_off, n, h->lo.sims[B.sim], h->stage);
        else hipLaunchKernelGGL((k_export<2, Prec64>), dim3(gp), dim3(256), 0, h->stream, A, B.p_off, n, h->lo.sims[B.sim], h->stage);
    }
    HIP_TRY(hipGetLastError());
    auto get = [&](long long plane, int count, double *dst) -> hipError_t {
        return hipMemcpyAsync(dst, h->stage + plane * n, sizeof(double) * n * count, hipMemcpyDeviceToHost, h->stream);
    };
    if (x) HIP_TRY(get(0, dim, x));
    if (v) HIP_TRY(get(dim, dim, v));
    if (T) HIP_TRY(get(2 * dim, 6, T));
    if (V) HIP_TRY(get(2 * dim + 6, 1, V));
    if (active)
        HIP_TRY(hipMemcpyAsync(active, reinterpret_cast<const int *>(h->stage + nd), sizeof(int) * n,
                               hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return MPM_OK;
}

int mpm_gpu_download_aux(mpm_gpu_handle *h, int body, double *u, double *L) {
    if (!h) return fail(MPM_ERR_ARG, "handle is NULL");
    if (body < 0 || body >= (int)h->lo.bodies.size()) return fail(MPM_ERR_ARG, "body %d out of range", body);
    const BodyDev &B = h->lo.bodies[body];
    const int dim = h->lo.dim;
    const long long n = h->count[body], np = h->lo.np;
    if (n == 0) return MPM_OK;
    if (u) {
        int rc = mpm_gpu_download(h, body, u, nullptr, nullptr, nullptr, nullptr);
        if (rc) return rc;
        const std::vector<double> &x0 = h->x0[body];
        for (long long i = 0; i < (long long)dim * n; i++) u[i] -= x0[i];
    }
    if (L) {
        if (!h->keep_L) return fail(MPM_ERR_STATE, "the velocity gradient is not kept (MPM_OPT_KEEP_L is 0)");
        if (!h->Lout) {  // no step yet
            std::memset(L, 0, sizeof(double) * 9 * n);
            return MPM_OK;
        }
        int rc = ensure_stage(h, n * 9);
        if (rc) return rc;
        const int gp = (int)ceil_div(np, 256);
        hipLaunchKernelGGL(k_unsort_range<double>, dim3(gp), dim3(256), 0, h->stream, np, 9, h->Lout, np, h->stage,
                           h->pid, B.p_off, n);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(L, h->stage, sizeof(double) * 9 * n, hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
    }
    return MPM_OK;
}

int mpm_gpu_download_plastic(mpm_gpu_handle *h, int body, double *gammap, double *gammap_dot) {
    if (!h) return fail(MPM_ERR_ARG, "handle is NULL");
    if (body < 0 || body >= (int)h->lo.bodies.size()) return fail(MPM_ERR_ARG, "body %d out of range", body);
    const BodyDev &B = h->lo.bodies[body];
    const int dim = h->lo.dim;
    const long long n = h->count[body], np = h->lo.np;
    if (n == 0) return MPM_OK;
    const bool f32 = h->precision == MPM_PREC_FP32;
    const int gpf = f32 ? (dim == 3 ? PW<3>::GP : PW<2>::GP) : (dim == 3 ? PF<3>::GP : PF<2>::GP);
    int rc = ensure_stage(h, n * 2);
    if (rc) return rc;
    const int gp = (int)ceil_div(np, 256);
    if (f32) {
        uint32_t *st = reinterpret_cast<uint32_t *>(h->stage);
        hipLaunchKernelGGL(k_unsort_range<uint32_t>, dim3(gp), dim3(256), 0, h->stream, np, 2,
                           h->Pw + (long long)gpf * np, np, st, h->pid, B.p_off, n);
        HIP_TRY(hipGetLastError());
        std::vector<float> g(2 * n);
        HIP_TRY(hipMemcpyAsync(g.data(), st, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        for (long long i = 0; i < n; i++) {
            if (gammap) gammap[i] = g[i];
            if (gammap_dot) gammap_dot[i] = g[n + i];
        }
        return MPM_OK;
    }
    hipLaunchKernelGGL(k_unsort_range<double>, dim3(gp), dim3(256), 0, h->stream, np, 2, h->P + (long long)gpf * np,
                       np, h->stage, h->pid, B.p_off, n);
    HIP_TRY(hipGetLastError());
    if (gammap) HIP_TRY(hipMemcpyAsync(gammap, h->stage, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
    if (gammap_dot)
        HIP_TRY(hipMemcpyAsync(gammap_dot, h->stage + n, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return MPM_OK;
}

}  // extern "C"

// node-block list of the last grid update: list par^1 (split and fused steps flip par
// after the grid update; the list stays intact until the next step's P2G fills the other)
static int last_grid_list(mpm_gpu_handle *h, const int **lst, int *cnt) {
    const int lp = h->par ^ 1;
    HIP_TRY(hipMemcpy(cnt, h->ncount + lp, sizeof(int), hipMemcpyDeviceToHost));
    *lst = h->nlist + (long long)lp * h->lo.nblk_total;
    return MPM_OK;
}

static long long sim_nodes(const SimDev &s, int dim) {
    long long nn = 1;
    for (int d = 0; d < dim; d++) nn *= s.N[d] + 1;
    return nn;
}

// k_connectivity over the points with original index in [o_lo, o_lo + o_n), into elem
// (o_n x 2^dim) and code (o_n)
static hipError_t launch_connectivity(mpm_gpu_handle *h, int *elem, uint8_t *code, long long o_lo, long long o_n) {
    KArgs A = kargs(h);
    const unsigned gp = (unsigned)ceil_div(h->lo.np, 256);
    const bool f32 = h->precision == MPM_PREC_FP32;
    if (h->lo.dim == 3) {
        if (f32) hipLaunchKernelGGL((k_connectivity<3, Prec32>), dim3(gp), dim3(256), 0, h->stream, A, elem, code, o_lo, o_n);
        else hipLaunchKernelGGL((k_connectivity<3, Prec64>), dim3(gp), dim3(256), 0, h->stream, A, elem, code, o_lo, o_n);
    } else {
        if (f32) hipLaunchKernelGGL((k_connectivity<2, Prec32>), dim3(gp), dim3(256), 0, h->stream, A, elem, code, o_lo, o_n);
        else hipLaunchKernelGGL((k_connectivity<2, Prec64>), dim3(gp), dim3(256), 0, h->stream, A, elem, code, o_lo, o_n);
    }
    return hipGetLastError();
}

// Device scratch of mpm_gpu_connectivity_noalloc: corner elements (MPM_NOALLOC_MAX_POINTS x 8
// int32) then codes (MPM_NOALLOC_MAX_POINTS bytes), 16.5 MiB held on every device the library
// runs on until the process ends. In host code `scratch` names a host-side stand-in, not
// device memory: its device address comes from hipGetSymbolAddress, once per device.
constexpr size_t SCRATCH_CODE_OFF = (size_t)MPM_NOALLOC_MAX_POINTS * 8 * sizeof(int);
static __device__ __align__(256) unsigned char scratch[SCRATCH_CODE_OFF + (size_t)MPM_NOALLOC_MAX_POINTS];

constexpr int SCRATCH_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// The scratch's device address on the handle's device (looked up on that device's first call)
static int scratch_of(mpm_gpu_handle *h, unsigned char **out) {
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != h->device)
        return fail(MPM_ERR_STATE, "the handle was created on device %d; device %d is current", h->device, dev);
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return fail(MPM_ERR_ARG, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(MPM_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch_ptr[dev];
    return MPM_OK;
}

template <int DIM, class GO>
static void launch_nodes_out(mpm_gpu_handle *h, int si, long long grid_base, const int *lst, int cnt, int with_rec,
                             int contact, double *out, long long nn) {
    if (cnt <= 0) return;
    KArgs A = kargs(h);
    A.rec = h->rec;  // kargs leaves the record unset (launch_grid sets it per update)
    A.crec = h->crec;
    hipLaunchKernelGGL((k_nodes_out<DIM, GO>), dim3((unsigned)ceil_div(cnt, 4)), dim3(256), 0, h->stream, A, si,
                       grid_base, lst, cnt, with_rec, contact, out, nn);
}

extern "C" {

int mpm_gpu_download_nodes(mpm_gpu_handle *h, int body, double *m, double *mxt, double *f, double *xt, double *dxt) {
    TraceRange tr_("mpm_gpu_download_nodes");
    if (!h) return fail(MPM_ERR_ARG, "handle is NULL");
    if (body < 0 || body >= (int)h->lo.bodies.size()) return fail(MPM_ERR_ARG, "body %d out of range", body);
    if (!h->grid_valid) return fail(MPM_ERR_STATE, "no step since the last upload or load: no node values");
    const bool want_rec = m || mxt || f;
    if (want_rec && (!h->recorded || !h->rec))
        return fail(MPM_ERR_STATE, "m, mxt and f need MPM_OPT_NODE_RECORD on during the last step call");
    const BodyDev &B = h->lo.bodies[body];
    const SimDev &s = h->lo.sims[B.sim];
    const int dim = h->lo.dim;
    const long long nn = sim_nodes(s, dim);
    const int with_rec = want_rec ? 1 : 0;
    const long long planes = (with_rec ? 1 + 2 * dim : 0) + 2 * dim;
    int rc = ensure_stage(h, planes * nn);
    if (rc) return rc;
    HIP_TRY(hipMemsetAsync(h->stage, 0, sizeof(double) * planes * nn, h->stream));
    const int *lst = nullptr;
    int cnt = 0;
    if ((rc = last_grid_list(h, &lst, &cnt))) return rc;
    const bool f32 = h->precision == MPM_PREC_FP32;
    if (dim == 3) {
        if (f32) launch_nodes_out<3, float>(h, B.sim, B.grid_base, lst, cnt, with_rec, -1, h->stage, nn);
        else launch_nodes_out<3, double>(h, B.sim, B.grid_base, lst, cnt, with_rec, -1, h->stage, nn);
    } else {
        if (f32) launch_nodes_out<2, float>(h, B.sim, B.grid_base, lst, cnt, with_rec, -1, h->stage, nn);
        else launch_nodes_out<2, double>(h, B.sim, B.grid_base, lst, cnt, with_rec, -1, h->stage, nn);
    }
    HIP_TRY(hipGetLastError());
    const size_t vb = sizeof(double) * dim * nn;
    const long long o = with_rec ? 1 + 2 * dim : 0;
    if (m) HIP_TRY(hipMemcpyAsync(m, h->stage, sizeof(double) * nn, hipMemcpyDeviceToHost, h->stream));
    if (mxt) HIP_TRY(hipMemcpyAsync(mxt, h->stage + nn, vb, hipMemcpyDeviceToHost, h->stream));
    if (f) HIP_TRY(hipMemcpyAsync(f, h->stage + (1 + dim) * nn, vb, hipMemcpyDeviceToHost, h->stream));
    if (xt) HIP_TRY(hipMemcpyAsync(xt, h->stage + o * nn, vb, hipMemcpyDeviceToHost, h->stream));
    if (dxt) HIP_TRY(hipMemcpyAsync(dxt, h->stage + (o + dim) * nn, vb, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return MPM_OK;
}

int mpm_gpu_download_contact(mpm_gpu_handle *h, int contact, double *normal, double *force) {
    if (!h) return fail(MPM_ERR_ARG, "handle is NULL");
    if (contact < 0 || contact >= (int)h->lo.contacts.size()) return fail(MPM_ERR_ARG, "contact %d out of range", contact);
    int si = 0;
    while (si + 1 < (int)h->lo.sims.size() && h->lo.sims[si + 1].contact0 <= contact) si++;
    const SimDev &s = h->lo.sims[si];
    if (contact - s.contact0 >= s.ncontact) return fail(MPM_ERR_ARG, "contact slot %d is not in use", contact);
    if (!h->grid_valid) return fail(MPM_ERR_STATE, "no step since the last upload or load: no node values");
    if (!h->recorded || !h->crec)
        return fail(MPM_ERR_STATE, "contact values need MPM_OPT_NODE_RECORD on during the last step call");
    const int dim = h->lo.dim;
    const long long nn = sim_nodes(s, dim);
    int rc = ensure_stage(h, 2LL * dim * nn);
    if (rc) return rc;
    // nodes the last update did not reach: normal 0/0 (no mass gradient), force 0
    const double qnan = std::numeric_limits<double>::quiet_NaN();
    hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)ceil_div(dim * nn, 256)), dim3(256), 0, h->stream, h->stage,
                       (long long)dim * nn, qnan);
    HIP_TRY(hipMemsetAsync(h->stage + dim * nn, 0, sizeof(double) * dim * nn, h->stream));
    const int *lst = nullptr;
    int cnt = 0;
    if ((rc = last_grid_list(h, &lst, &cnt))) return rc;
    if (dim == 3) launch_nodes_out<3, double>(h, si, 0, lst, cnt, 0, contact, h->stage, nn);
    else launch_nodes_out<2, double>(h, si, 0, lst, cnt, 0, contact, h->stage, nn);
    HIP_TRY(hipGetLastError());
    const size_t vb = sizeof(double) * dim * nn;
    if (normal) HIP_TRY(hipMemcpyAsync(normal, h->stage, vb, hipMemcpyDeviceToHost, h->stream));
    if (force) HIP_TRY(hipMemcpyAsync(force, h->stage + dim * nn, vb, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return MPM_OK;
}

int mpm_gpu_connectivity(mpm_gpu_handle *h, int body, int32_t *corner_elem, uint8_t *code) {
    if (!h) return fail(MPM_ERR_ARG, "handle is NULL");
    if (body < 0 || body >= (int)h->lo.bodies.size()) return fail(MPM_ERR_ARG, "body %d out of range", body);
    if (!corner_elem || !code) return fail(MPM_ERR_ARG, "corner_elem and code are required");
    const BodyDev &B = h->lo.bodies[body];
    const int dim = h->lo.dim;
    const int nc = 1 << dim;
    const long long n = h->count[body], np = h->lo.np;
    if (n == 0) return MPM_OK;
    int *d_elem = nullptr;
    uint8_t *d_code = nullptr;
    HIP_TRY(hipMalloc(&d_elem, sizeof(int) * np * nc));
    HIP_TRY(hipMalloc(&d_code, np));
    HIP_TRY(launch_connectivity(h, d_elem, d_code, 0, np));
    HIP_TRY(hipMemcpyAsync(corner_elem, d_elem + B.p_off * nc, sizeof(int) * n * nc, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipMemcpyAsync(code, d_code + B.p_off, n, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipFree(d_elem));
    HIP_TRY(hipFree(d_code));
    return MPM_OK;
}

int mpm_gpu_connectivity_noalloc(mpm_gpu_handle *h, int body, int32_t *corner_elem, uint8_t *code) {
    if (!h) return fail(MPM_ERR_ARG, "handle is NULL");
    if (body < 0 || body >= (int)h->lo.bodies.size()) return fail(MPM_ERR_ARG, "body %d out of range", body);
    if (!corner_elem || !code) return fail(MPM_ERR_ARG, "corner_elem and code are required");
    const BodyDev &B = h->lo.bodies[body];
    const int nc = 1 << h->lo.dim;
    const long long n = h->count[body];
    if (n > MPM_NOALLOC_MAX_POINTS)
        return fail(MPM_ERR_ARG, "body %d has %lld points; the allocation-free connectivity takes at most %lld", body,
                    n, (long long)MPM_NOALLOC_MAX_POINTS);
    if (n == 0) return MPM_OK;
    unsigned char *scr = nullptr;
    int rc = scratch_of(h, &scr);
    if (rc) return rc;
    // the body's points only: corner elements of point o at scr[o * nc], codes after them
    int *d_elem = reinterpret_cast<int *>(scr);
    uint8_t *d_code = scr + SCRATCH_CODE_OFF;
    HIP_TRY(launch_connectivity(h, d_elem, d_code, B.p_off, n));
    HIP_TRY(hipMemcpyAsync(corner_elem, d_elem, sizeof(int) * n * nc, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipMemcpyAsync(code, d_code, n, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return MPM_OK;
}

int mpm_gpu_set_timing(mpm_gpu_handle *h, int enable) {
    if (!h) return fail(MPM_ERR_ARG, "handle is NULL");
    h->timing = enable != 0;
    for (int i = 0; i < NCLASS; i++) {
        h->t_launch[i] = 0;
        h->t_ms[i] = 0;
    }
    return MPM_OK;
}

int mpm_gpu_get_timing(mpm_gpu_handle *h, int kernel_class, int64_t *launches, double *ms) {
    if (!h) return fail(MPM_ERR_ARG, "handle is NULL");
    if (kernel_class < 0 || kernel_class >= NCLASS) return fail(MPM_ERR_ARG, "kernel class out of range");
    if (launches) *launches = h->t_launch[kernel_class];
    if (ms) *ms = h->t_ms[kernel_class];
    return MPM_OK;
}

int mpm_gpu_stats(mpm_gpu_handle *h, int64_t *active_nodes, int64_t *active_points) {
    if (!h) return fail(MPM_ERR_ARG, "handle is NULL");
    if (active_nodes) *active_nodes = (int64_t)h->last_active_nodes;
    if (active_points) {
        unsigned long long *d = nullptr;
        HIP_TRY(hipMalloc(&d, sizeof(unsigned long long)));
        HIP_TRY(hipMemsetAsync(d, 0, sizeof(unsigned long long), h->stream));
        hipLaunchKernelGGL(k_count_active, dim3((unsigned)std::max(1ll, ceil_div(h->lo.np, 256))), dim3(256), 0,
                           h->stream, h->lo.np, h->act, d);
        unsigned long long c = 0;
        HIP_TRY(hipMemcpyAsync(&c, d, sizeof(c), hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        HIP_TRY(hipFree(d));
        *active_points = (int64_t)c;
    }
    return MPM_OK;
}

}  // extern "C"

namespace {
__global__ __launch_bounds__(256) void k_copy_probe(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                    long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}
}  // namespace

extern "C" {

int mpm_gpu_copy_probe(int64_t nbytes, int reps, double *gbps) {
    if (nbytes < 16 || reps < 1 || !gbps) return fail(MPM_ERR_ARG, "need nbytes >= 16, reps >= 1, gbps");
    const long long n = nbytes / 16;
    double2 *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipMalloc(&a, n * 16));
    HIP_TRY(hipMalloc(&b, n * 16));
    HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipMemsetAsync(a, 1, n * 16, st));
    const int grid = 256 * 16;
    hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, st, a, b, n);  // warm-up
    HIP_TRY(hipEventRecord(e0, st));
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, st, a, b, n);
    HIP_TRY(hipEventRecord(e1, st));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    *gbps = 2.0 * n * 16 * reps / (ms * 1e-3) / 1e9;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    (void)hipFree(a);
    (void)hipFree(b);
    return MPM_OK;
}

int mpm_gpu_host_weights(const mpm_grid_desc *grid, const double *x, double extent, int32_t *node_ijk, double *S,
                         double *G, int max_tuples) {
    if (!grid || !x || !node_ijk || !S || !G) return fail(MPM_ERR_ARG, "null argument");
    mpm_body_desc bd{};
    bd.n_points = 1;
    mpm_sim_desc sd{};
    sd.grid = *grid;
    sd.nbody = 1;
    sd.bodies = &bd;
    Layout lo;
    int rc = build_layout(&sd, 1, lo);
    if (rc) return rc;
    const SimDev &s = lo.sims[0];
    int count = 0;
    auto emit = [&](const int *nijk, double Sv, const double *Gv, int dim) {
        if (count < max_tuples) {
            for (int d = 0; d < 3; d++) {
                node_ijk[3 * count + d] = d < dim ? nijk[d] : 0;
                G[3 * count + d] = d < dim ? Gv[d] : 0.0;
            }
            S[count] = Sv;
        }
        count++;
    };
    if (s.dim == 3) {
        Win<3> w;
        bool cpdi;
        const int mode = point_window<3>(s, x, extent, w, cpdi);
        if (mode == MODE_LEFT) return -100;
        if (mode == MODE_SEP) {
            for_each_sep<3>(w, extent, [&](const int *off, double Sv, const double *Gv) {
                int n[3] = {w.base[0] + off[0], w.base[1] + off[1], w.base[2] + off[2]};
                emit(n, Sv, Gv, 3);
            });
        } else if (mode == MODE_TUP) {
            for_each_tuple<3>(s, x, extent, cpdi, [&](const int *n, double Sv, const double *Gv) { emit(n, Sv, Gv, 3); });
        }
    } else {
        Win<2> w;
        bool cpdi;
        const int mode = point_window<2>(s, x, extent, w, cpdi);
        if (mode == MODE_LEFT) return -100;
        if (mode == MODE_SEP) {
            for_each_sep<2>(w, extent, [&](const int *off, double Sv, const double *Gv) {
                int n[2] = {w.base[0] + off[0], w.base[1] + off[1]};
                emit(n, Sv, Gv, 2);
            });
        } else if (mode == MODE_TUP) {
            for_each_tuple<2>(s, x, extent, cpdi, [&](const int *n, double Sv, const double *Gv) { emit(n, Sv, Gv, 2); });
        }
    }
    return count;
}

int mpm_gpu_synchronize(mpm_gpu_handle *h) {
    if (!h) return fail(MPM_ERR_ARG, "handle is NULL");
    HIP_TRY(hipStreamSynchronize(h->stream));
    return MPM_OK;
}

int mpm_gpu_destroy(mpm_gpu_handle *h) {
    if (!h) return MPM_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    void *ptrs[] = {h->d_sims, h->d_bodies, h->d_contacts, h->P, h->P2, h->Pw, h->Pw2, h->act, h->act2, h->pid,
                    h->pid2, h->pkey, h->prank, h->acc, h->out, h->out32, h->nflag, h->nlist, h->ncount, h->bcount, h->boff,
                    h->plist, h->npblk, h->Lout, h->stats, h->stage, h->rec, h->crec, h->rare, h->rare_n,
                    h->work, h->nonfin, h->fcnt, h->pcell, h->bflag, h->bpos};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (unsigned *p : h->scan_tmp)
        if (p) (void)hipFree(p);
    for (unsigned *p : h->scan_tmp_i)
        if (p) (void)hipFree(p);
    for (unsigned *p : h->scan_tmp_b)
        if (p) (void)hipFree(p);
    if (h->inact_count) (void)hipFree(h->inact_count);
    if (h->inact_off) (void)hipFree(h->inact_off);
    for (hipGraphExec_t g : h->graph)
        if (g) (void)hipGraphExecDestroy(g);
    for (Timer &t : h->free_timers) {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    for (Timer &t : h->timers) {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return MPM_OK;
}

}  // extern "C"
