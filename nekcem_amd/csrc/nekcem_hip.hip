// nekcem_hip.hip -- C-ABI of libnekcem_hip.so (declared in include/nekcem_hip.h).
//
// Host side of the MI355X Maxwell step.  A mesh of one polynomial order is a
// DevMesh: the compact per-element tables of mesh_tables.hpp on the device
// plus the ping-pong state buffers and the RK register.  A handle
// (nkc_create / nkc_create_box) owns one DevMesh, a stream and the captured
// step graphs; a batch (nkc_batch_create) owns one DevMesh per polynomial
// order holding all its cavities back to back, one stream per order, and a
// graph whose per-order branches run side by side.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <map>
#include <string>
#include <vector>
#include <array>
#include <mutex>

#include "../../include/nekcem_hip.h"
#include "adjoint_kernel.hpp"
#include "diag_kernels.hpp"
#include "geo_kernel.hpp"
#include "geo_launch.hpp"
#include "mesh_tables.hpp"
#include "order_launch.hpp"
#include "restart_io.hpp"
#include "stage_kernel.hpp"

using namespace nkc;

namespace {

thread_local std::string g_err;

int32_t fail(int32_t code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(call)                                                                   \
    do {                                                                               \
        hipError_t _e = (call);                                                        \
        if (_e != hipSuccess)                                                          \
            return fail(NKC_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), \
                        __FILE__, __LINE__);                                           \
    } while (0)

// Low-storage RK45 of Carpenter & Kennedy (1994), the ifrk45 constants of
// rk_storage (src/cem_common.F:86-104).
const double RK4A[5] = {0.0,
                        -567301805773.0 / 1357537059087.0,
                        -2404267990393.0 / 2016746695238.0,
                        -3550918686646.0 / 2091501179385.0,
                        -1275806237668.0 / 842570457699.0};
const double RK4B[5] = {1432997174477.0 / 9575080441755.0,
                        5161836677717.0 / 13612068292357.0,
                        1720146321549.0 / 2090206949498.0,
                        3134564353537.0 / 4481467310338.0,
                        2277821191437.0 / 14882151754819.0};
const double RK4C[5] = {0.0,
                        1432997174477.0 / 9575080441755.0,
                        2526269341429.0 / 6820363962896.0,
                        2006345519317.0 / 3224310063776.0,
                        2802321613138.0 / 2924317926251.0};

bool supported_n1(int n1) { return n1 >= 4 && n1 <= 10; }
// Orders the general-hexahedra kernel takes (nkc_create only).
bool supported_n1_general(int n1) { return n1 >= GEO_MIN_N1 && n1 <= GEO_MAX_N1; }

int32_t dev_alloc(double **p, size_t n) {
    HIPCHK(hipMalloc((void **)p, sizeof(double) * n));
    return NKC_OK;
}

template <class T>
int32_t upload(T **dst, const std::vector<T> &src) {
    HIPCHK(hipMalloc((void **)dst, sizeof(T) * (src.empty() ? 1 : src.size())));
    if (!src.empty()) {
        const hipError_t e = hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            // no half-filled table is left behind
            (void)hipFree(*dst);
            *dst = nullptr;
            return fail(NKC_ERR_HIP, "hipMemcpy: %s", hipGetErrorString(e));
        }
    }
    return NKC_OK;
}

// One polynomial order's worth of elements on the device.
struct DevMesh {
    int n1 = 0, mat = MAT_ELEM;
    int64_t nelt = 0, npts = 0, fs = 0;
    double *u[2] = {nullptr, nullptr};
    double *k = nullptr;
    double *D = nullptr, *geo = nullptr, *fco = nullptr, *ieps = nullptr, *imu = nullptr;
    double *dt_e = nullptr;       // per-element dt (batches), or null
    int32_t *nbr = nullptr;
    int layout = -1;              // stage-kernel layout (PACK | OCC << 1); -1: not tuned yet
    std::vector<int32_t> xcd_lo;  // 9 element bounds of equal-work XCD ranges, or empty
    std::vector<double> jac, w;   // host copies (source weights)
    // general hexahedra (geo_kernel.hpp): per-node metric quotients and
    // bm1, per-face-node normal + lift, partner node; fco per face node
    bool general = false;
    double *gmet = nullptr, *fgeo = nullptr, *bmw = nullptr;
    int32_t *fnb = nullptr;
};

int32_t dev_mesh_upload(const HostMesh &m, DevMesh &dm) {
    dm.n1 = m.n1;
    dm.mat = m.mat ? MAT_NODE : MAT_ELEM;
    dm.nelt = m.nelt;
    dm.npts = m.nelt * m.n1 * m.n1 * m.n1;
    dm.fs = (dm.npts + 31) / 32 * 32;
    dm.jac = m.jac;
    dm.w = m.w;
    const size_t ub = sizeof(double) * 6 * (size_t)dm.fs;
    HIPCHK(hipMalloc((void **)&dm.u[0], ub));
    HIPCHK(hipMalloc((void **)&dm.u[1], ub));
    HIPCHK(hipMalloc((void **)&dm.k, ub));
    HIPCHK(hipMemset(dm.u[0], 0, ub));
    HIPCHK(hipMemset(dm.u[1], 0, ub));
    HIPCHK(hipMemset(dm.k, 0, ub));
    int32_t rc;
    if ((rc = upload(&dm.D, m.D))) return rc;
    if ((rc = upload(&dm.geo, m.geo))) return rc;
    if ((rc = upload(&dm.nbr, m.nbr))) return rc;
    if ((rc = upload(&dm.fco, m.fco))) return rc;
    if ((rc = upload(&dm.ieps, m.ieps))) return rc;
    if ((rc = upload(&dm.imu, m.imu))) return rc;
    dm.general = m.general;
    if (m.general) {
        if ((rc = upload(&dm.gmet, m.gmet))) return rc;
        if ((rc = upload(&dm.fgeo, m.fgeo))) return rc;
        if ((rc = upload(&dm.bmw, m.bmw))) return rc;
        if ((rc = upload(&dm.fnb, m.fnb))) return rc;
        dm.layout = 0;            // one layout: nothing to tune
    }
    return NKC_OK;
}

void dev_mesh_free(DevMesh &dm) {
    double *p[] = {dm.u[0], dm.u[1], dm.k, dm.D, dm.geo, dm.fco, dm.ieps, dm.imu, dm.dt_e, dm.gmet, dm.fgeo, dm.bmw};
    for (double *q : p) (void)hipFree(q);
    (void)hipFree(dm.nbr);
    (void)hipFree(dm.fnb);
    dm = DevMesh();
}

// Stage-kernel layouts: 0..3 one element per workgroup, bit 0 = PACK (lane
// roles packed), bit 1 = OCC (4 waves/SIMD register budget); 4, 5 two
// elements per workgroup, packed, without / with the OCC budget (N <= 5
// only); 6, 7 the six-role kernel without / with the budget of its
// LDS-bound occupancy (N >= 7 only; no PML, incident field or graphene
// sheet).  The residual test entry uses layout 0.
constexpr int NLAYOUT = NKC_NLAYOUT;

// Work of a Drude element's stage relative to a plain one, for the XCD split
// (E lanes of a Drude element spent 33 against 18 us in the level phase at
// N = 9, 24^3, a workgroup's life ~1.6 times the others').
constexpr double DRUDE_WORK = 1.6;

// Elements per workgroup of a layout.
int layout_epw(int layout) { return (layout == 4 || layout == 5) ? 2 : 1; }

// Whether a layout is built for this order.
bool layout_built(int n1, int layout) {
    if (layout < 4) return true;
    if (layout < 6) return n1 <= EPW_MAX_N1;
    return n1 >= S6_MIN_N1;
}

// Whether a built layout carries the physics of these stage arguments (the
// six-role kernel has no PML, incident field or graphene sheet).
bool layout_runs(int n1, int layout, const StageArgs &a) {
    return layout_built(n1, layout) && !(layout >= 6 && (a.pml_off || a.inc_mask || a.gr_off));
}

// Lanes a workgroup of the given layout has.
int stage_block(int n1, int layout) {
    const int n2 = n1 * n1;
    if (layout >= 6) return (6 * n2 + 63) / 64 * 64;
    const int one = (layout & 1) ? (4 * n2 + 63) / 64 * 64 : 4 * ((n2 + 63) / 64 * 64);
    return one * layout_epw(layout);
}

}  // namespace
struct nkc_handle;
namespace {
// Entry points whose physics the general-hexahedra kernel does not carry.
int32_t general_refuses(const nkc_handle *, const char *what) {
    return fail(NKC_ERR_UNSUPPORTED, "%s is not built for general (non-box) hexahedra", what);
}

// One stage on a general mesh: the per-node-geometry kernel.
int32_t launch_geo_stage(const DevMesh &dm, const StageArgs &a, int gmode, hipStream_t s) {
    GeoArgs g{};
    g.uin = a.uin;
    g.uout = a.uout;
    g.kbuf = a.kbuf;
    g.res = a.res;
    g.D = dm.D;
    g.gmet = dm.gmet;
    g.bmw = dm.bmw;
    g.fgeo = dm.fgeo;
    g.fnb = dm.fnb;
    g.fco = dm.fco;
    g.ieps = dm.ieps;
    g.imu = dm.imu;
    g.fs = dm.fs;
    g.npts = dm.npts;
    g.nelt = (int32_t)dm.nelt;
    g.per_xcd = (int)((dm.nelt + 7) / 8);
    g.rka = a.rka;
    g.rkb = a.rkb;
    g.dt = a.dt;
    g.kflags = a.kflags;
    if (!launch_geo(g, dm.n1, gmode, 8 * g.per_xcd, s))
        return fail(NKC_ERR_UNSUPPORTED, "polynomial order N=%d not built for general hexahedra", dm.n1 - 1);
    HIPCHK(hipGetLastError());
    return NKC_OK;
}

int32_t launch_stage(const DevMesh &dm, const StageArgs &a, int mode, hipStream_t s, int layout = -2) {
    if (dm.general) return launch_geo_stage(dm, a, mode == MODE_RK ? GMODE_RK : GMODE_RESIDUAL, s);
    const int pk0 = layout == -2 ? (dm.layout < 0 ? 0 : dm.layout) : layout;
    // the residual entry and orders without the layout run layout 0
    const int pk = (mode == MODE_RK && layout_runs(dm.n1, pk0, a)) ? pk0 : 0;
    const int64_t nwg = (dm.nelt + layout_epw(pk) - 1) / layout_epw(pk);
    const int per_xcd = (int)((nwg + 7) / 8);
    StageArgs b = a;
    b.per_xcd = per_xcd;
    int grid = 8 * per_xcd;
    b.xcd_bal = 0;
    // the equal-work ranges: instances with in-stage physics only (a Drude
    // region sets them), one element per workgroup
    if (dm.xcd_lo.size() == 9 && layout_epw(pk) == 1 && (a.src_off || a.drude_off || a.lor_off || a.pml_off || a.inc_mask || a.gr_off)) {
        int len = 0;
        for (int x = 0; x < 8; ++x) {
            b.xcd_lo[x] = dm.xcd_lo[x];
            len = std::max(len, dm.xcd_lo[x + 1] - dm.xcd_lo[x]);
        }
        b.xcd_lo[8] = dm.xcd_lo[8];
        b.xcd_bal = 1;
        grid = 8 * std::max(len, 1);
    }
    switch (dm.n1) {
    case 4: launch_stage_n4(b, dm.mat, mode, pk, grid, s); break;
    case 5: launch_stage_n5(b, dm.mat, mode, pk, grid, s); break;
    case 6: launch_stage_n6(b, dm.mat, mode, pk, grid, s); break;
    case 7: launch_stage_n7(b, dm.mat, mode, pk, grid, s); break;
    case 8: launch_stage_n8(b, dm.mat, mode, pk, grid, s); break;
    case 9: launch_stage_n9(b, dm.mat, mode, pk, grid, s); break;
    case 10: launch_stage_n10(b, dm.mat, mode, pk, grid, s); break;
    default: return fail(NKC_ERR_UNSUPPORTED, "polynomial order N=%d not built", dm.n1 - 1);
    }
    HIPCHK(hipGetLastError());
    return NKC_OK;
}

StageArgs mesh_args(const DevMesh &dm) {
    StageArgs a{};
    a.D = dm.D;
    a.geo = dm.geo;
    a.nbr = dm.nbr;
    a.fco = dm.fco;
    a.ieps = dm.ieps;
    a.imu = dm.imu;
    a.dt_e = dm.dt_e;
    a.fs = dm.fs;
    a.nelt = (int32_t)dm.nelt;
    a.kflags = 3;
    return a;
}

// RK-register traffic of stage s of a step: no read in the first stage
// (a_1 = 0), no write in the last (the next step's first stage ignores it).
int32_t stage_kflags(int s) {
    return (s == 0 ? 0 : 1) | (s == 4 ? 0 : 2);
}

// Autotuner over the stage kernel's layouts (lane packing x register
// budget): time each on this mesh, alternating, best of three rounds of
// four launches, and keep the fastest.  It writes only a scratch output and
// a scratch RK register, and with the handle's source, Drude, Lorentz, PML,
// incident-field and graphene-sheet set-up (feat, may be null) it times the kernel
// instance the step runs, the auxiliary fields and their registers
// redirected to scratch: the state is untouched.  A packed layout whose
// workgroup equals the unpacked one is not timed.  jsz, lsz, psz, gsz:
// doubles of the Drude / Lorentz / PML / graphene-sheet state blocks.
int32_t tune_stage(DevMesh &dm, const StageArgs *feat, size_t jsz, size_t lsz, size_t psz, size_t gsz,
                   const double *uin, hipStream_t s, double us[NLAYOUT]) {
    for (int l = 0; l < NLAYOUT; ++l) us[l] = NAN;
    if (dm.general) {
        dm.layout = 0;            // the general kernel has one layout
        return NKC_OK;
    }
    // not timed: a layout not built for this order, and a packed one-element
    // layout whose workgroup equals its unpacked twin
    auto skip = [&](int p) {
        return !layout_built(dm.n1, p) || (p < 4 && (p & 1) && stage_block(dm.n1, p) == stage_block(dm.n1, p - 1)) ||
               (p >= 6 && feat && (feat->pml_off || feat->inc_mask || feat->gr_off));
    };
    const size_t blk = 6 * (size_t)dm.fs;
    const size_t jx = (feat && feat->drude_off) ? jsz : 0;
    const size_t lx = (feat && feat->lor_off) ? lsz : 0;
    const size_t px = (feat && feat->pml_off) ? psz : 0;
    const size_t gx = (feat && feat->gr_off) ? gsz : 0;
    double *scr = nullptr;
    int32_t rc = dev_alloc(&scr, 2 * blk + 2 * jx + 2 * lx + 2 * px + 2 * gx);
    if (rc) return rc;
    StageArgs a = mesh_args(dm);
    if (feat) {
        a.tdev = feat->tdev;
        a.src_off = feat->src_off;
        a.src_amp = feat->src_amp;
        for (int c = 0; c < 3; ++c) a.src_dir[c] = feat->src_dir[c];
        a.src_t0 = feat->src_t0;
        a.src_tw = feat->src_tw;
        a.src_kind = feat->src_kind;
        a.src_om = feat->src_om;
        a.src_ph = feat->src_ph;
        a.inc_mask = feat->inc_mask;
        a.xyz = feat->xyz;
        for (int c = 0; c < 3; ++c) {
            a.inc_k[c] = feat->inc_k[c];
            a.inc_h[c] = feat->inc_h[c];
            a.inc_e[c] = feat->inc_e[c];
        }
        a.inc_om = feat->inc_om;
        if (px) {
            a.pml_off = feat->pml_off;
            a.psig = feat->psig;
            a.pP = scr + 2 * blk + 2 * jx + 2 * lx;
            a.pkP = scr + 2 * blk + 2 * jx + 2 * lx + px;
        }
        if (gx) {
            a.gr_off = feat->gr_off;
            a.gpar = feat->gpar;
            a.gA = feat->gA;
            a.gAn = scr + 2 * blk + 2 * jx + 2 * lx + 2 * px;
            a.gK = scr + 2 * blk + 2 * jx + 2 * lx + 2 * px + gx;
        }
        if (jx) {
            a.drude_off = feat->drude_off;
            a.dpar = feat->dpar;
            a.dJ = scr + 2 * blk;
            a.dkJ = scr + 2 * blk + jx;
        }
        if (lx) {
            a.lor_off = feat->lor_off;
            a.lpar = feat->lpar;
            a.lL = scr + 2 * blk + 2 * jx;
            a.lkL = scr + 2 * blk + 2 * jx + lx;
        }
    }
    a.uin = uin;
    a.uout = scr;
    a.kbuf = scr + blk;
    a.rka = 0.0;
    a.rkb = RK4B[0];
    a.dt = 1e-3;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float best[NLAYOUT];
    for (int l = 0; l < NLAYOUT; ++l) best[l] = INFINITY;
    rc = NKC_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = fail(NKC_ERR_HIP, "hipEventCreate");
    for (int round = 0; round < 3 && rc == NKC_OK; ++round)
        for (int p = 0; p < NLAYOUT && rc == NKC_OK; ++p) {
            if (skip(p)) continue;
            if ((rc = launch_stage(dm, a, MODE_RK, s, p))) break;
            if (hipEventRecord(e0, s) != hipSuccess) { rc = fail(NKC_ERR_HIP, "hipEventRecord"); break; }
            for (int r = 0; r < 4 && rc == NKC_OK; ++r) rc = launch_stage(dm, a, MODE_RK, s, p);
            if (rc) break;
            float ms = 0.0f;
            if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                rc = fail(NKC_ERR_HIP, "tuner timing");
                break;
            }
            best[p] = std::min(best[p], ms / 4.0f);
        }
    (void)hipStreamSynchronize(s);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(scr);
    if (rc) return rc;
    int pick = 0;
    for (int l = 0; l < NLAYOUT; ++l) {
        if (std::isfinite(best[l])) us[l] = 1e3 * best[l];
        if (best[l] < best[pick]) pick = l;
    }
    dm.layout = pick;
    return NKC_OK;
}

// time = time + dt at the end of a step (src/cem_drive.F:639), on the device
// so a replayed graph sees the right stage times.
__global__ void advance_time(double *t, double dt) { t[0] = t[0] + dt; }

__global__ void maxabs_kernel(const double *__restrict__ u, int64_t fs, int64_t npts,
                              unsigned long long *out) {
    unsigned long long m = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int f = 0; f < 6; ++f)
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npts; i += stride) {
            const double v = fabs(u[f * fs + i]);
            const unsigned long long b = __double_as_longlong(v);   // monotone for v >= 0, NaN on top
            m = b > m ? b : m;
        }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_down(m, off);
        m = o > m ? o : m;
    }
    __shared__ unsigned long long sm[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sm[w] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long r = 0;
        for (int q = 0; q < (int)(blockDim.x >> 6); ++q) r = sm[q] > r ? sm[q] : r;
        atomicMax(out, r);
    }
}

__global__ void zero_u64(unsigned long long *p) { *p = 0ull; }

}  // namespace

// ===========================================================================
// single-mesh handle
// ===========================================================================
struct nkc_handle {
    DevMesh m;
    int device = 0;
    hipStream_t stream = nullptr;
    double *res = nullptr;
    unsigned long long *red = nullptr;   // check_finite result
    int cur = 0;                          // which u buffer holds the state
    bool use_graph = true;
    bool keep_k = false;                  // last stage writes the RK register (nkc_set_keep_k)
    double graph_dt = NAN;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    // tdev[0] = time at the start of the next step, tdev[1] = stage time for
    // nkc_residual
    double *tdev = nullptr;
    double t_host = 0.0;
    // source
    int32_t *src_off = nullptr;
    double *src_amp = nullptr;
    double src_dir[3] = {0, 0, 0}, src_t0 = 0, src_tw = 1;
    int32_t src_kind = 0;                 // 0 Gaussian pulse, 1 sin(src_om t + src_ph)
    double src_om = 0, src_ph = 0;
    // PML (compact per PML element: sigma 3 n3; (B, D) 6 n3, register, residual)
    int64_t npml = 0;
    std::vector<int64_t> pml_elems;       // PML element -> mesh element
    int32_t *pml_off = nullptr;
    double *psig = nullptr, *pP = nullptr, *pkP = nullptr, *presP = nullptr;
    // incident plane wave: per element face bits, wave parameters
    int64_t ninc = 0;
    int32_t *inc_mask = nullptr;
    double inc_k[3] = {0, 0, 0}, inc_om = 0, inc_h[3] = {0, 0, 0}, inc_e[3] = {0, 0, 0};
    // graphene sheet: blocks of n2 face nodes, one per sheet face side
    // (gr_face[b] = e*6 + kf); listed point l -> (block, node) at gr_list
    // with its 0-based face-point number; state in gA[cur]
    int64_t ngraph = 0;
    std::vector<int64_t> gr_face;
    std::vector<std::array<int64_t, 3>> gr_list;
    int32_t *gr_off = nullptr;
    double *gpar = nullptr, *gA[2] = {nullptr, nullptr}, *gK = nullptr;
    // Drude
    int64_t ndrude = 0, ndelem = 0;
    int32_t *drude_off = nullptr;
    double *dJ = nullptr, *dkJ = nullptr, *dpar = nullptr, *dresJ = nullptr;
    // Lorentz (compact per Lorentz element: (J, P) 6 n3, register, (a,b,c) per node)
    int64_t nlor = 0, nlelem = 0;
    int32_t *lor_off = nullptr;
    double *lL = nullptr, *lkL = nullptr, *lpar = nullptr, *lresL = nullptr;
    std::vector<int64_t> lor_slot;        // listed node -> lelem 6 n3 + vloc
    std::vector<int64_t> drude_slot;      // listed node -> compact (elem*3*n3 + vloc)
    // host copy of the element tables (material updates, adjoint)
    HostMesh hm;
    double C0 = 1.0;
    double *yz = nullptr;                 // (sqrt(eps/mu), sqrt(mu/eps)) per ieps entry, adjoint only
    int64_t adj_store_cap = 0;            // cap on the adjoint's stage-store window (0: from free memory)
    int32_t adj_grads = 3;                // bit 0: eps/mu gradients, bit 1: Drude gradients
    double tune_us[NLAYOUT] = {NAN, NAN, NAN, NAN, NAN, NAN, NAN, NAN};   // tuner's stage times per layout
    // on-device diagnostics: node coordinates (3 x fs; from the box or
    // nkc_set_coords), element Jacobians and GLL weights (bm1 = jac w3),
    // partial norms, the 13 final values
    bool is_box = false;
    int64_t box_nel[3] = {0, 0, 0};
    double box_lo[3] = {0, 0, 0}, box_hi[3] = {0, 0, 0};
    double *xyz = nullptr, *djac = nullptr, *dw = nullptr, *dpart = nullptr, *dout = nullptr;
    std::vector<double> hxyz;             // host copy of xyz
    // adjoint: checkpoints of the state every `every` steps of the last
    // nkc_adjoint_forward, the stage store of one segment, adjoint buffers
    struct Adj {
        int64_t nsteps = 0, every = 0, nseg = 0;
        // A segment (every steps) is reversed in windows of S steps: the
        // window's start states are recomputed from the segment checkpoint
        // into `sub`, then each window is recomputed through the stage store
        // (5 S stages) and reversed.  S = every when the store fits.
        int64_t S = 0, nsub_max = 0;
        int64_t sub_seg = -1;             // segment whose window starts `sub` holds
        // sub_all: `sub` holds the window starts of every segment, kept by
        // the forward (S < every and room for them), so the backward needs
        // no pass to rebuild them
        bool sub_all = false;
        // The store has K window slots; slot_id[k] = seg * nsub_max + q of
        // the window slot k holds (-1: none).  K > 1 only when windows are
        // whole segments: the forward then keeps its last K segments.
        int64_t K = 1;
        std::vector<int64_t> slot_id;
        int cur_slot = 0;                 // slot queue_step_store writes
        size_t bytes = 0;                 // device bytes held by the adjoint
        bool have_grad = false;           // a backward sweep has filled the gradients
        double dt = 0.0, t0 = 0.0;
        double *ckpt = nullptr;           // nseg x 6 fs: the state at the start of each segment
        double *sub = nullptr;            // (sub_all ? nseg : 1) x nsub_max x 6 fs: window start states
        double *subJ = nullptr;           // the same x jsz
        double *st_u = nullptr;           // K x (5 S + 1) x 6 fs: every stage input of a window
        double *st_r = nullptr;           // K x 5 S x 6 fs: every stage residual
        double *lam[2] = {nullptr, nullptr}, *kap[2] = {nullptr, nullptr};
        double *geps = nullptr, *gmu = nullptr;    // fs each
        double *ufin = nullptr;           // final forward state (restored after backward)
        double *kfin = nullptr;           // final RK register
        // Drude region (jsz = ndelem 3 n3 doubles per current block)
        double *ckJ = nullptr;            // nseg x jsz: the current at each segment start
        double *stJ = nullptr;            // K x 5 S x jsz: every stage-input current
        double *lamJ = nullptr, *kapJ = nullptr;   // adjoint current and its RK register
        double *gA = nullptr, *gB = nullptr;       // ndelem n3 each: d/da, d/db per node
        double *Jfin = nullptr, *kJfin = nullptr;  // final current and its RK register
    } adj;
};

namespace {

// doubles in one compact Drude current block (ndelem x 3 x n3)
size_t drude_size(const nkc_handle *h) {
    const size_t n3 = (size_t)h->m.n1 * h->m.n1 * h->m.n1;
    return (size_t)h->ndelem * 3 * n3;
}

// doubles in one compact Lorentz state block (nlelem x 6 x n3)
size_t lorentz_size(const nkc_handle *h) {
    const size_t n3 = (size_t)h->m.n1 * h->m.n1 * h->m.n1;
    return (size_t)h->nlelem * 6 * n3;
}

// doubles in one compact PML state block (npml x 6 x n3)
size_t pml_size(const nkc_handle *h) {
    const size_t n3 = (size_t)h->m.n1 * h->m.n1 * h->m.n1;
    return (size_t)h->npml * 6 * n3;
}

// doubles in one graphene-sheet state block (blocks x 15 x n2)
size_t sheet_size(const nkc_handle *h) {
    const size_t n2 = (size_t)h->m.n1 * h->m.n1;
    return h->gr_face.size() * 15 * n2;
}

StageArgs base_args(nkc_handle *h) {
    StageArgs a = mesh_args(h->m);
    a.tdev = h->tdev;
    a.src_off = h->src_off;
    a.src_amp = h->src_amp;
    for (int c = 0; c < 3; ++c) a.src_dir[c] = h->src_dir[c];
    a.src_t0 = h->src_t0;
    a.src_tw = h->src_tw;
    a.drude_off = h->drude_off;
    a.dJ = h->dJ;
    a.dkJ = h->dkJ;
    a.dpar = h->dpar;
    a.dresJ = h->dresJ;
    a.lor_off = h->lor_off;
    a.lL = h->lL;
    a.lkL = h->lkL;
    a.lpar = h->lpar;
    a.lresL = h->lresL;
    a.src_kind = h->src_kind;
    a.src_om = h->src_om;
    a.src_ph = h->src_ph;
    a.pml_off = h->pml_off;
    a.psig = h->psig;
    a.pP = h->pP;
    a.pkP = h->pkP;
    a.presP = h->presP;
    a.inc_mask = h->inc_mask;
    a.xyz = h->xyz;
    for (int c = 0; c < 3; ++c) {
        a.inc_k[c] = h->inc_k[c];
        a.inc_h[c] = h->inc_h[c];
        a.inc_e[c] = h->inc_e[c];
    }
    a.inc_om = h->inc_om;
    a.gr_off = h->gr_off;
    a.gpar = h->gpar;
    a.gA = h->gA[h->cur];
    a.gAn = h->gA[h->cur ^ 1];
    a.gK = h->gK;
    return a;
}

// RK-register flags of stage s for this handle: with keep_k the last stage
// writes the register too, so khn/ken after a step are rk4_upd's.
int32_t handle_kflags(const nkc_handle *h, int s) {
    return stage_kflags(s) | ((h->keep_k && s == 4) ? 2 : 0);
}

// Queue one RK step starting from buffer `from` (ends in 1-from).
int32_t queue_step(nkc_handle *h, int from, double dt) {
    for (int s = 0; s < 5; ++s) {
        StageArgs a = base_args(h);
        const int src = (from + s) & 1;
        a.uin = h->m.u[src];
        a.uout = h->m.u[src ^ 1];
        a.gA = h->gA[src];
        a.gAn = h->gA[src ^ 1];
        a.kbuf = h->m.k;
        a.rka = RK4A[s];
        a.kflags = handle_kflags(h, s);
        a.rkb = RK4B[s];
        a.rkc = RK4C[s];
        a.dt = dt;
        const int32_t rc = launch_stage(h->m, a, MODE_RK, h->stream);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(advance_time, dim3(1), dim3(1), 0, h->stream, h->tdev, dt);
    HIPCHK(hipGetLastError());
    return NKC_OK;
}

int32_t drop_graphs(nkc_handle *h) {
    for (int i = 0; i < 2; ++i)
        if (h->gexec[i]) {
            HIPCHK(hipGraphExecDestroy(h->gexec[i]));
            h->gexec[i] = nullptr;
        }
    h->graph_dt = NAN;
    return NKC_OK;
}

int32_t capture(nkc_handle *h, double dt) {
    int32_t rc = drop_graphs(h);
    if (rc) return rc;
    for (int from = 0; from < 2; ++from) {
        HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        rc = queue_step(h, from, dt);
        hipGraph_t g = nullptr;
        const hipError_t ee = hipStreamEndCapture(h->stream, &g);
        if (rc || ee != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return rc ? rc : fail(NKC_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ee));
        }
        const hipError_t ei = hipGraphInstantiate(&h->gexec[from], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            h->gexec[from] = nullptr;
            return fail(NKC_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ei));
        }
    }
    h->graph_dt = dt;
    return NKC_OK;
}

void free_adjoint(nkc_handle *h) {
    auto &A = h->adj;
    double *p[] = {A.ckpt, A.st_u, A.st_r, A.lam[0], A.lam[1], A.kap[0], A.kap[1], A.geps, A.gmu, A.ufin,
                   A.kfin, A.ckJ, A.stJ, A.lamJ, A.kapJ, A.gA, A.gB, A.Jfin, A.kJfin, A.sub, A.subJ};
    for (double *q : p) (void)hipFree(q);
    A = nkc_handle::Adj();
}

// Admittance/impedance pairs for the adjoint's material gradient, built
// from the same inverse-material tables the stage kernel reads.
int32_t ensure_yz(nkc_handle *h) {
    if (h->yz) return NKC_OK;
    const size_t n = h->hm.ieps.size();
    std::vector<double> yz(2 * n);
    for (size_t x = 0; x < n; ++x) {
        yz[2 * x] = std::sqrt(h->hm.imu[x] / h->hm.ieps[x]);
        yz[2 * x + 1] = std::sqrt(h->hm.ieps[x] / h->hm.imu[x]);
    }
    return upload(&h->yz, yz);
}

void free_handle(nkc_handle *h) {
    drop_graphs(h);
    free_adjoint(h);
    (void)hipFree(h->yz);
    h->yz = nullptr;
    dev_mesh_free(h->m);
    (void)hipFree(h->res); (void)hipFree(h->red); (void)hipFree(h->tdev);
    (void)hipFree(h->src_off); (void)hipFree(h->src_amp);
    (void)hipFree(h->drude_off); (void)hipFree(h->dJ); (void)hipFree(h->dkJ);
    (void)hipFree(h->dpar); (void)hipFree(h->dresJ);
    (void)hipFree(h->lor_off); (void)hipFree(h->lL); (void)hipFree(h->lkL);
    (void)hipFree(h->lpar); (void)hipFree(h->lresL);
    (void)hipFree(h->pml_off); (void)hipFree(h->psig); (void)hipFree(h->pP); (void)hipFree(h->pkP);
    (void)hipFree(h->presP); (void)hipFree(h->inc_mask);
    (void)hipFree(h->gr_off); (void)hipFree(h->gpar); (void)hipFree(h->gA[0]); (void)hipFree(h->gA[1]);
    (void)hipFree(h->gK);
    (void)hipFree(h->xyz); (void)hipFree(h->djac); (void)hipFree(h->dw);
    (void)hipFree(h->dpart); (void)hipFree(h->dout);
    if (h->stream) (void)hipStreamDestroy(h->stream);
}

constexpr int DIAG_BLOCKS = 1024;

int32_t ensure_diag(nkc_handle *h) {
    int32_t rc;
    if (!h->djac && (rc = upload(&h->djac, h->hm.jac))) return rc;
    if (!h->dw && (rc = upload(&h->dw, h->hm.w))) return rc;
    if (!h->dpart && (rc = dev_alloc(&h->dpart, (size_t)DIAG_BLOCKS * DIAG_NV))) return rc;
    if (!h->dout && (rc = dev_alloc(&h->dout, DIAG_NV))) return rc;
    return NKC_OK;
}

// Node coordinates on the host and the device: a box handle builds them
// from its extents (the GLL lines of genbox, element-major, i fastest);
// others need nkc_set_coords.
int32_t ensure_coords(nkc_handle *h) {
    if (h->xyz) return NKC_OK;
    if (!h->is_box) return fail(NKC_ERR_ARG, "no node coordinates: call nkc_set_coords first");
    const int n1 = h->m.n1, n2 = n1 * n1, n3 = n2 * n1;
    std::vector<double> z, w, D;
    gll_basis(n1, z, w, D);
    const int64_t fs = h->m.fs, npts = h->m.npts;
    std::vector<double> xyz(3 * (size_t)fs, 0.0);
    std::vector<double> line[3];
    for (int a = 0; a < 3; ++a) {
        const int64_t ne = h->box_nel[a];
        line[a].resize(ne * n1);
        const double lo = h->box_lo[a], len = h->box_hi[a] - h->box_lo[a];
        for (int64_t ea = 0; ea < ne; ++ea) {
            const double v0 = lo + len * (double)ea / (double)ne, v1 = lo + len * (double)(ea + 1) / (double)ne;
            for (int i = 0; i < n1; ++i) line[a][ea * n1 + i] = v0 * (0.5 * (1.0 - z[i])) + v1 * (0.5 * (1.0 + z[i]));
        }
    }
    for (int64_t p = 0; p < npts; ++p) {
        const int64_t e = p / n3, r = p % n3;
        const int64_t ex = e % h->box_nel[0], ey = (e / h->box_nel[0]) % h->box_nel[1], ez = e / (h->box_nel[0] * h->box_nel[1]);
        xyz[p] = line[0][ex * n1 + r % n1];
        xyz[fs + p] = line[1][ey * n1 + (r / n1) % n1];
        xyz[2 * fs + p] = line[2][ez * n1 + r / n2];
    }
    h->hxyz = xyz;
    return upload(&h->xyz, xyz);
}

int32_t make_handle(const HostMesh &hm, double C0, nkc_handle **out) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(NKC_ERR_NODEVICE, "no HIP device visible");
    nkc_handle *h = new nkc_handle();
    h->hm = hm;
    h->C0 = C0;
    int32_t rc = NKC_OK;
    do {
        if (hipGetDevice(&h->device) != hipSuccess) { rc = fail(NKC_ERR_HIP, "hipGetDevice"); break; }
        if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
            rc = fail(NKC_ERR_HIP, "hipStreamCreate"); break;
        }
        if ((rc = dev_mesh_upload(hm, h->m))) break;
        // element Jacobians and GLL weights of bm1 for the diagnostics
        if ((rc = upload(&h->djac, hm.jac)) || (rc = upload(&h->dw, hm.w))) break;
        if (hipMalloc((void **)&h->red, 8) != hipSuccess) { rc = fail(NKC_ERR_HIP, "hipMalloc"); break; }
        if (hipMalloc((void **)&h->tdev, 2 * sizeof(double)) != hipSuccess ||
            hipMemset(h->tdev, 0, 2 * sizeof(double)) != hipSuccess) { rc = fail(NKC_ERR_HIP, "hipMalloc time"); break; }
    } while (0);
    if (rc) {
        free_handle(h);
        delete h;
        return rc;
    }
    *out = h;
    return NKC_OK;
}

// copy 6 fields between the caller's arrays (3 + 3 components at stride
// ld, in host or device memory: hipMemcpyDefault tells them apart) and the
// library's device blocks
int32_t copy_fields_h2d(double *dst, int64_t fs, int64_t dst_off, const double *hn, const double *en,
                        int64_t n, int64_t ld) {
    const size_t nb = sizeof(double) * n;
    for (int c = 0; c < 3; ++c) {
        HIPCHK(hipMemcpy(dst + c * fs + dst_off, hn + c * ld, nb, hipMemcpyDefault));
        HIPCHK(hipMemcpy(dst + (3 + c) * fs + dst_off, en + c * ld, nb, hipMemcpyDefault));
    }
    return NKC_OK;
}

int32_t copy_fields_d2h(const double *src, int64_t fs, int64_t src_off, double *hn, double *en,
                        int64_t n, int64_t ld) {
    const size_t nb = sizeof(double) * n;
    for (int c = 0; c < 3; ++c) {
        HIPCHK(hipMemcpy(hn + c * ld, src + c * fs + src_off, nb, hipMemcpyDefault));
        HIPCHK(hipMemcpy(en + c * ld, src + (3 + c) * fs + src_off, nb, hipMemcpyDefault));
    }
    return NKC_OK;
}

// A caller's array (host or device memory) as a host vector, for the entry
// points that build their tables on the host.
template <class T>
int32_t stage_host(std::vector<T> &out, const T *src, int64_t n) {
    out.resize((size_t)n);
    if (n > 0) HIPCHK(hipMemcpy(out.data(), src, sizeof(T) * (size_t)n, hipMemcpyDefault));
    return NKC_OK;
}

int32_t launch_adjoint(const DevMesh &dm, const AdjArgs &a0, hipStream_t s) {
    const int per_xcd = (int)((dm.nelt + 7) / 8);
    const int grid = 8 * per_xcd;
    AdjArgs a = a0;
    a.per_xcd = per_xcd;
    a.D = dm.D;
    a.geo = dm.geo;
    a.nbr = dm.nbr;
    a.ieps = dm.ieps;
    a.imu = dm.imu;
    a.fs = dm.fs;
    a.nelt = (int32_t)dm.nelt;
    switch (dm.n1) {
    case 4: launch_adjoint_n4(a, dm.mat, grid, s); break;
    case 5: launch_adjoint_n5(a, dm.mat, grid, s); break;
    case 6: launch_adjoint_n6(a, dm.mat, grid, s); break;
    case 7: launch_adjoint_n7(a, dm.mat, grid, s); break;
    case 8: launch_adjoint_n8(a, dm.mat, grid, s); break;
    case 9: launch_adjoint_n9(a, dm.mat, grid, s); break;
    case 10: launch_adjoint_n10(a, dm.mat, grid, s); break;
    default: return fail(NKC_ERR_UNSUPPORTED, "polynomial order N=%d not built", dm.n1 - 1);
    }
    HIPCHK(hipGetLastError());
    return NKC_OK;
}

// One forward step of the segment recompute: stage s of step `slot` reads
// its input from stage store slot 5 slot + s and writes the next stage's
// input to the following slot, so every stage input is kept at no extra
// traffic; the stage residual goes to the matching st_r slot (not captured:
// the store offsets change with the slot).
int32_t queue_step_store(nkc_handle *h, double dt, int64_t slot) {
    const int64_t blk = 6 * h->m.fs;
    auto &A = h->adj;
    double *su = A.st_u + (int64_t)A.cur_slot * (5 * A.S + 1) * blk;
    double *sr = A.st_r ? A.st_r + (int64_t)A.cur_slot * 5 * A.S * blk : nullptr;
    double *sj = h->ndrude ? A.stJ + (int64_t)A.cur_slot * 5 * A.S * (int64_t)drude_size(h) : nullptr;
    for (int s = 0; s < 5; ++s) {
        StageArgs a = base_args(h);
        a.uin = su + (slot * 5 + s) * blk;
        a.uout = su + (slot * 5 + s + 1) * blk;
        a.kbuf = h->m.k;
        a.rka = RK4A[s];
        a.kflags = handle_kflags(h, s);
        a.rkb = RK4B[s];
        a.rkc = RK4C[s];
        a.dt = dt;
        a.st_r = sr ? sr + (slot * 5 + s) * blk : nullptr;
        if (h->ndrude) a.st_J = sj + (slot * 5 + s) * (int64_t)drude_size(h);
        const int32_t rc = launch_stage(h->m, a, MODE_RK, h->stream);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(advance_time, dim3(1), dim3(1), 0, h->stream, h->tdev, dt);
    HIPCHK(hipGetLastError());
    return NKC_OK;
}

}  // namespace

// ===========================================================================
// batch of independent box cavities
// ===========================================================================
struct nkc_batch {
    struct Cav {
        int bin = 0;
        int64_t off = 0, nelt = 0;   // element offset and count in the bin
        double dt = NAN, time = 0.0;
    };
    std::vector<DevMesh> bins;
    std::vector<hipStream_t> streams;     // one per bin
    std::vector<hipEvent_t> ev;           // fork/join events
    std::vector<Cav> cav;
    hipStream_t main = nullptr;
    int cur = 0;
    bool graph_ok = false;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
};

namespace {

int32_t batch_drop_graphs(nkc_batch *b) {
    for (int i = 0; i < 2; ++i)
        if (b->gexec[i]) {
            HIPCHK(hipGraphExecDestroy(b->gexec[i]));
            b->gexec[i] = nullptr;
        }
    b->graph_ok = false;
    return NKC_OK;
}

// One RK step of every cavity: each bin's five stages run in order on its own
// stream, the bins side by side; forked from and joined back to b->main.
int32_t batch_queue_step(nkc_batch *b, int from) {
    const int nb = (int)b->bins.size();
    HIPCHK(hipEventRecord(b->ev[0], b->main));
    for (int i = 0; i < nb; ++i) HIPCHK(hipStreamWaitEvent(b->streams[i], b->ev[0], 0));
    for (int i = 0; i < nb; ++i) {
        const DevMesh &dm = b->bins[i];
        for (int s = 0; s < 5; ++s) {
            StageArgs a = mesh_args(dm);
            const int src = (from + s) & 1;
            a.uin = dm.u[src];
            a.uout = dm.u[src ^ 1];
            a.kbuf = dm.k;
            a.rka = RK4A[s];
            a.kflags = stage_kflags(s);
            a.rkb = RK4B[s];
            a.rkc = RK4C[s];
            a.dt = 0.0;              // per element from dt_e
            const int32_t rc = launch_stage(dm, a, MODE_RK, b->streams[i]);
            if (rc) return rc;
        }
        HIPCHK(hipEventRecord(b->ev[1 + i], b->streams[i]));
        HIPCHK(hipStreamWaitEvent(b->main, b->ev[1 + i], 0));
    }
    return NKC_OK;
}

int32_t batch_capture(nkc_batch *b) {
    int32_t rc = batch_drop_graphs(b);
    if (rc) return rc;
    for (int from = 0; from < 2; ++from) {
        HIPCHK(hipStreamBeginCapture(b->main, hipStreamCaptureModeThreadLocal));
        rc = batch_queue_step(b, from);
        hipGraph_t g = nullptr;
        const hipError_t ee = hipStreamEndCapture(b->main, &g);
        if (rc || ee != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return rc ? rc : fail(NKC_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ee));
        }
        const hipError_t ei = hipGraphInstantiate(&b->gexec[from], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            b->gexec[from] = nullptr;
            return fail(NKC_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ei));
        }
    }
    b->graph_ok = true;
    return NKC_OK;
}

void free_batch(nkc_batch *b) {
    batch_drop_graphs(b);
    for (auto &dm : b->bins) dev_mesh_free(dm);
    for (auto s : b->streams) (void)hipStreamDestroy(s);
    for (auto e : b->ev) (void)hipEventDestroy(e);
    if (b->main) (void)hipStreamDestroy(b->main);
}

}  // namespace

extern "C" {

const char *nkc_last_error(void) { return g_err.c_str(); }

int32_t nkc_create(const nkc_desc *d, nkc_handle **out) {
    g_err.clear();
    if (!d || !out) return fail(NKC_ERR_ARG, "null descriptor or output handle");
    *out = nullptr;
    const int n1 = d->N + 1;
    if (!supported_n1_general(n1)) return fail(NKC_ERR_UNSUPPORTED, "polynomial order N=%d outside 3..11", d->N);
    if (d->nelt <= 0 || d->nelt > (int64_t)1 << 30) return fail(NKC_ERR_ARG, "nelt=%lld out of range", (long long)d->nelt);
    const void *req[] = {d->dxm1, d->wxm1, d->rxmn, d->rymn, d->rzmn, d->sxmn, d->symn, d->szmn,
                         d->txmn, d->tymn, d->tzmn, d->jacm, d->bmn, d->unxm, d->unym, d->unzm,
                         d->aream, d->cemface, d->face_partner, d->permittivity, d->permeability,
                         d->Z_0, d->Z_1, d->Y_0, d->Y_1};
    for (const void *p : req)
        if (!p) return fail(NKC_ERR_ARG, "required array pointer is NULL");
    if (d->ncempec < 0 || (d->ncempec > 0 && !d->cempec)) return fail(NKC_ERR_ARG, "bad cempec");
    const int64_t npts = (int64_t)n1 * n1 * n1 * d->nelt;
    if (npts > ((int64_t)1 << 31) - 1) return fail(NKC_ERR_ARG, "npts exceeds int32 indexing");
    HostMesh hm;
    // box meshes above the box kernels' orders run on the general kernel
    const int32_t rc = supported_n1(n1) ? compact_reference(d, hm) : compact_general(d, hm);
    if (rc) return fail(rc, "%s", hm.err.c_str());
    return make_handle(hm, d->ifupwind ? 1.0 : 0.0, out);
}

int32_t nkc_create_box(const nkc_box_desc *d, nkc_handle **out) {
    g_err.clear();
    if (!d || !out) return fail(NKC_ERR_ARG, "null descriptor or output handle");
    *out = nullptr;
    if (!supported_n1(d->N + 1)) return fail(NKC_ERR_UNSUPPORTED, "polynomial order N=%d outside 3..9", d->N);
    const int64_t n1 = d->N + 1;
    const int64_t nelt = (int64_t)d->nel[0] * d->nel[1] * d->nel[2];
    if (nelt <= 0 || nelt * n1 * n1 * n1 > ((int64_t)1 << 31) - 1) return fail(NKC_ERR_ARG, "bad box size");
    HostMesh hm;
    int32_t rc = build_box(d, hm);
    if (rc) return fail(rc, "%s", hm.err.c_str());
    if ((rc = make_handle(hm, d->ifupwind ? 1.0 : 0.0, out))) return rc;
    nkc_handle *h = *out;
    h->is_box = true;
    for (int a = 0; a < 3; ++a) {
        h->box_nel[a] = d->nel[a];
        h->box_lo[a] = d->lo[a];
        h->box_hi[a] = d->hi[a];
    }
    return NKC_OK;
}

int32_t nkc_create_hex(int32_t N, int32_t ifupwind, int64_t nelt, const double *xc, const double *yc,
                       const double *zc, const char *cbc, const double *bc, const double *eps_elem,
                       const double *mu_elem, nkc_handle **out) {
    g_err.clear();
    if (!out || !xc || !yc || !zc || !cbc || !bc) return fail(NKC_ERR_ARG, "null argument");
    *out = nullptr;
    if (!supported_n1(N + 1)) return fail(NKC_ERR_UNSUPPORTED, "polynomial order N=%d outside 3..9", N);
    const int64_t n1 = N + 1;
    if (nelt <= 0 || nelt * n1 * n1 * n1 > ((int64_t)1 << 31) - 1) return fail(NKC_ERR_ARG, "bad element count");
    HostMesh hm;
    const int32_t rc = build_hex(N + 1, ifupwind, nelt, xc, yc, zc, cbc, bc, eps_elem, mu_elem, hm);
    if (rc) return fail(rc, "%s", hm.err.c_str());
    return make_handle(hm, ifupwind ? 1.0 : 0.0, out);
}

int32_t nkc_destroy(nkc_handle *h) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    (void)hipStreamSynchronize(h->stream);
    free_handle(h);
    delete h;
    return NKC_OK;
}

int32_t nkc_set_state(nkc_handle *h, const double *hn, const double *en, const double *khn,
                      const double *ken, int64_t ld) {
    if (!h || !hn || !en) return fail(NKC_ERR_ARG, "null handle or state");
    const DevMesh &m = h->m;
    if (ld < m.npts) return fail(NKC_ERR_ARG, "ld=%lld < npts=%lld", (long long)ld, (long long)m.npts);
    HIPCHK(hipStreamSynchronize(h->stream));
    int32_t rc = copy_fields_h2d(m.u[h->cur], m.fs, 0, hn, en, m.npts, ld);
    if (rc) return rc;
    const size_t nb = sizeof(double) * m.npts;
    for (int c = 0; c < 3; ++c) {
        if (khn) HIPCHK(hipMemcpy(m.k + c * m.fs, khn + c * ld, nb, hipMemcpyDefault));
        else HIPCHK(hipMemset(m.k + c * m.fs, 0, nb));
        if (ken) HIPCHK(hipMemcpy(m.k + (3 + c) * m.fs, ken + c * ld, nb, hipMemcpyDefault));
        else HIPCHK(hipMemset(m.k + (3 + c) * m.fs, 0, nb));
    }
    return NKC_OK;
}

int32_t nkc_get_state(nkc_handle *h, double *hn, double *en, double *khn, double *ken, int64_t ld) {
    if (!h || !hn || !en) return fail(NKC_ERR_ARG, "null handle or state");
    const DevMesh &m = h->m;
    if (ld < m.npts) return fail(NKC_ERR_ARG, "ld=%lld < npts=%lld", (long long)ld, (long long)m.npts);
    HIPCHK(hipStreamSynchronize(h->stream));
    int32_t rc = copy_fields_d2h(m.u[h->cur], m.fs, 0, hn, en, m.npts, ld);
    if (rc) return rc;
    if (khn || ken) {
        std::vector<double> tmp((khn && ken) ? 0 : 3 * ld);
        rc = copy_fields_d2h(m.k, m.fs, 0, khn ? khn : tmp.data(), ken ? ken : tmp.data(), m.npts, ld);
        if (rc) return rc;
    }
    return NKC_OK;
}

int32_t nkc_step(nkc_handle *h, double *time, double dt, int64_t nsteps) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    if (nsteps < 0) return fail(NKC_ERR_ARG, "nsteps < 0");
    if (!std::isfinite(dt)) return fail(NKC_ERR_ARG, "dt not finite");
    if (nsteps == 0) return NKC_OK;
    const double t0 = time ? *time : h->t_host;
    if (!(t0 == h->t_host)) {
        HIPCHK(hipStreamSynchronize(h->stream));
        HIPCHK(hipMemcpy(h->tdev, &t0, sizeof(double), hipMemcpyHostToDevice));
        h->t_host = t0;
    }
    if (h->m.layout < 0) {
        const StageArgs f = base_args(h);
        const int32_t rc = tune_stage(h->m, &f, drude_size(h), lorentz_size(h), pml_size(h), sheet_size(h), h->m.u[h->cur], h->stream,
                       h->tune_us);
        if (rc) return rc;
    }
    if (h->use_graph) {
        if (!(h->graph_dt == dt)) {
            const int32_t rc = capture(h, dt);
            if (rc) return rc;
        }
        for (int64_t s = 0; s < nsteps; ++s) {
            HIPCHK(hipGraphLaunch(h->gexec[h->cur], h->stream));
            h->cur ^= 1;
        }
    } else {
        for (int64_t s = 0; s < nsteps; ++s) {
            const int32_t rc = queue_step(h, h->cur, dt);
            if (rc) return rc;
            h->cur ^= 1;
        }
    }
    for (int64_t s = 0; s < nsteps; ++s) h->t_host = h->t_host + dt;   // same sums as the device
    if (time) *time = h->t_host;
    return NKC_OK;
}

int32_t nkc_residual(nkc_handle *h, double rktime, double *reshn, double *resen, int64_t ld) {
    if (!h || !reshn || !resen) return fail(NKC_ERR_ARG, "null handle or output");
    const DevMesh &m = h->m;
    if (ld < m.npts) return fail(NKC_ERR_ARG, "ld < npts");
    if (!h->res) HIPCHK(hipMalloc((void **)&h->res, sizeof(double) * 6 * m.fs));
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(h->tdev + 1, &rktime, sizeof(double), hipMemcpyHostToDevice));
    StageArgs a = base_args(h);
    a.tdev = h->tdev + 1;
    a.rkc = 0.0;
    a.uin = m.u[h->cur];
    a.res = h->res;
    const int32_t rc = launch_stage(m, a, MODE_RESIDUAL, h->stream);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    return copy_fields_d2h(h->res, m.fs, 0, reshn, resen, m.npts, ld);
}

int32_t nkc_curl(nkc_handle *h, double *reshn, double *resen, int64_t ld) {
    if (!h || !reshn || !resen) return fail(NKC_ERR_ARG, "null handle or output");
    const DevMesh &m = h->m;
    if (ld < m.npts) return fail(NKC_ERR_ARG, "ld < npts");
    int32_t rc;
    if ((rc = ensure_diag(h))) return rc;
    if (!h->res) HIPCHK(hipMalloc((void **)&h->res, sizeof(double) * 6 * m.fs));
    HIPCHK(hipStreamSynchronize(h->stream));
    if (m.general) {
        StageArgs a = mesh_args(m);
        a.uin = m.u[h->cur];
        a.res = h->res;
        if ((rc = launch_geo_stage(m, a, GMODE_CURL, h->stream))) return rc;
        HIPCHK(hipStreamSynchronize(h->stream));
        return copy_fields_d2h(h->res, m.fs, 0, reshn, resen, m.npts, ld);
    }
    const int bs = ((m.n1 * m.n1 * m.n1 + 63) / 64) * 64;
    const dim3 grid((unsigned)m.nelt), block((unsigned)bs);
    const double *u = m.u[h->cur];
    switch (m.n1) {
    case 4: hipLaunchKernelGGL(curl_kernel<4>, grid, block, 0, h->stream, u, h->res, m.fs, m.D, m.geo, h->djac, h->dw); break;
    case 5: hipLaunchKernelGGL(curl_kernel<5>, grid, block, 0, h->stream, u, h->res, m.fs, m.D, m.geo, h->djac, h->dw); break;
    case 6: hipLaunchKernelGGL(curl_kernel<6>, grid, block, 0, h->stream, u, h->res, m.fs, m.D, m.geo, h->djac, h->dw); break;
    case 7: hipLaunchKernelGGL(curl_kernel<7>, grid, block, 0, h->stream, u, h->res, m.fs, m.D, m.geo, h->djac, h->dw); break;
    case 8: hipLaunchKernelGGL(curl_kernel<8>, grid, block, 0, h->stream, u, h->res, m.fs, m.D, m.geo, h->djac, h->dw); break;
    case 9: hipLaunchKernelGGL(curl_kernel<9>, grid, block, 0, h->stream, u, h->res, m.fs, m.D, m.geo, h->djac, h->dw); break;
    case 10: hipLaunchKernelGGL(curl_kernel<10>, grid, block, 0, h->stream, u, h->res, m.fs, m.D, m.geo, h->djac, h->dw); break;
    default: return fail(NKC_ERR_UNSUPPORTED, "polynomial order N=%d not built", m.n1 - 1);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    return copy_fields_d2h(h->res, m.fs, 0, reshn, resen, m.npts, ld);
}

int32_t nkc_set_source(nkc_handle *h, const double *amp, const double dir[3], double t0, double tw) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    if (h->m.general && (amp != nullptr)) return general_refuses(h, "a source");
    HIPCHK(hipStreamSynchronize(h->stream));
    free_adjoint(h);                      // the stored trajectory has the old source
    h->m.layout = -1;                     // the step's kernel instance may change: tune again
    (void)hipFree(h->src_off); h->src_off = nullptr;
    (void)hipFree(h->src_amp); h->src_amp = nullptr;
    int32_t rc = drop_graphs(h);
    if (rc || !amp) return rc;
    if (!dir || !(tw != 0.0) || !std::isfinite(t0) || !std::isfinite(tw)) return fail(NKC_ERR_ARG, "bad source direction/time");
    const DevMesh &m = h->m;
    const int n1 = m.n1, n2 = n1 * n1, n3 = n2 * n1;
    std::vector<int32_t> off(m.nelt, -1);
    std::vector<double> vals;
    for (int64_t e = 0; e < m.nelt; ++e) {
        bool any = false;
        for (int q = 0; q < n3; ++q) any |= amp[e * n3 + q] != 0.0;
        if (!any) continue;
        off[e] = (int32_t)(vals.size() / n3);
        for (int k = 0; k < n1; ++k)
            for (int j = 0; j < n1; ++j)
                for (int i = 0; i < n1; ++i) {
                    // usersrc adds amp (bm1-weighted) before ebm1; the kernel
                    // works per unit J w3
                    const double bm = m.jac[e] * (m.w[i] * m.w[j] * m.w[k]);
                    vals.push_back(amp[e * n3 + i + n1 * j + n2 * k] / bm);
                }
    }
    if (vals.empty()) return NKC_OK;
    // the offset table switches the term on: upload it last, so that a
    // failed upload leaves the source off rather than half set
    if ((rc = upload(&h->src_amp, vals))) return rc;
    if ((rc = upload(&h->src_off, off))) return rc;
    for (int c = 0; c < 3; ++c) h->src_dir[c] = dir[c];
    h->src_t0 = t0;
    h->src_tw = tw;
    h->src_kind = 0;
    return NKC_OK;
}

int32_t nkc_set_source_time(nkc_handle *h, int32_t kind, double omega, double phase) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    if (kind != 0 && kind != 1) return fail(NKC_ERR_ARG, "source time kind %d not 0 or 1", kind);
    if (!std::isfinite(omega) || !std::isfinite(phase)) return fail(NKC_ERR_ARG, "source omega/phase not finite");
    HIPCHK(hipStreamSynchronize(h->stream));
    free_adjoint(h);                      // the stored trajectory has the old source
    h->src_kind = kind;
    h->src_om = omega;
    h->src_ph = phase;
    return drop_graphs(h);                // the captured stages carry the old factor
}

// Element bounds of 8 contiguous ranges of equal work, one per XCD, when the
// elements with off[e] >= 0 cost w times the others (elements in order,
// so each XCD still takes a compact run of the mesh).
std::vector<int32_t> xcd_ranges(const std::vector<int32_t> &off, double w) {
    const int64_t n = (int64_t)off.size();
    double total = 0.0;
    for (int64_t e = 0; e < n; ++e) total += off[e] >= 0 ? w : 1.0;
    std::vector<int32_t> lo(9, (int32_t)n);
    lo[0] = 0;
    double acc = 0.0;
    int x = 1;
    for (int64_t e = 0; e < n && x < 8; ++e) {
        acc += off[e] >= 0 ? w : 1.0;
        while (x < 8 && acc >= total * x / 8.0) lo[x++] = (int32_t)(e + 1);
    }
    return lo;
}

int32_t nkc_set_drude(nkc_handle *h, int64_t ndrude, const int32_t *dindex_in, const double *params_in,
                      const double *J_in, const double *kJ_in) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    // the blocks are built on the host from the caller's arrays (host or device memory)
    std::vector<int32_t> dindex_h;
    std::vector<double> params_h, J_h, kJ_h;
    if (ndrude > 0) {
        int32_t rs;
        if ((dindex_in && (rs = stage_host(dindex_h, dindex_in, ndrude))) ||
            (params_in && (rs = stage_host(params_h, params_in, 2 * ndrude))) ||
            (J_in && (rs = stage_host(J_h, J_in, 3 * ndrude))) || (kJ_in && (rs = stage_host(kJ_h, kJ_in, 3 * ndrude))))
            return rs;
    }
    const int32_t *dindex = dindex_in ? dindex_h.data() : nullptr;
    const double *params = params_in ? params_h.data() : nullptr;
    const double *J = J_in ? J_h.data() : nullptr;
    const double *kJ = kJ_in ? kJ_h.data() : nullptr;
    if (h->m.general && (ndrude > 0)) return general_refuses(h, "a Drude region");
    HIPCHK(hipStreamSynchronize(h->stream));
    (void)hipFree(h->drude_off); h->drude_off = nullptr;
    (void)hipFree(h->dJ); h->dJ = nullptr;
    (void)hipFree(h->dkJ); h->dkJ = nullptr;
    (void)hipFree(h->dpar); h->dpar = nullptr;
    (void)hipFree(h->dresJ); h->dresJ = nullptr;
    free_adjoint(h);                      // the stored trajectory has the old region
    h->m.layout = -1;                     // the step's kernel instance may change: tune again
    h->ndrude = 0;
    h->ndelem = 0;
    h->drude_slot.clear();
    h->m.xcd_lo.clear();
    int32_t rc = drop_graphs(h);
    if (rc || ndrude == 0) return rc;
    if (ndrude < 0 || !dindex || !params) return fail(NKC_ERR_ARG, "bad Drude node list");
    const DevMesh &m = h->m;
    const int n3 = m.n1 * m.n1 * m.n1;
    std::vector<int32_t> off(m.nelt, -1);
    int32_t nde = 0;
    for (int64_t l = 0; l < ndrude; ++l) {
        const int64_t v = (int64_t)dindex[l] - 1;
        if (v < 0 || v >= m.npts) return fail(NKC_ERR_ARG, "dindex[%lld]=%d out of range", (long long)l, dindex[l]);
        const int64_t e = v / n3;
        if (off[e] < 0) off[e] = nde++;
    }
    std::vector<double> par((size_t)nde * n3 * 2, 0.0), Jv((size_t)nde * 3 * n3, 0.0), kv = Jv;
    h->drude_slot.resize(ndrude);
    for (int64_t l = 0; l < ndrude; ++l) {
        const int64_t v = (int64_t)dindex[l] - 1, e = v / n3, vl = v % n3;
        const int64_t b = (int64_t)off[e] * 3 * n3 + vl;
        h->drude_slot[l] = b;
        par[((int64_t)off[e] * n3 + vl) * 2 + 0] = params[2 * l + 0];
        par[((int64_t)off[e] * n3 + vl) * 2 + 1] = params[2 * l + 1];
        for (int c = 0; c < 3; ++c) {
            if (J) Jv[b + c * n3] = J[c * ndrude + l];
            if (kJ) kv[b + c * n3] = kJ[c * ndrude + l];
        }
    }
    if ((rc = upload(&h->dpar, par))) return rc;
    if ((rc = upload(&h->dJ, Jv))) return rc;
    if ((rc = upload(&h->dkJ, kv))) return rc;
    if ((rc = upload(&h->dresJ, kv))) return rc;
    if ((rc = upload(&h->drude_off, off))) return rc;   // last: switches the term on
    h->ndrude = ndrude;
    h->ndelem = nde;
    h->m.xcd_lo = xcd_ranges(off, DRUDE_WORK);
    return NKC_OK;
}

int32_t nkc_get_drude(nkc_handle *h, double *J, double *kJ, double *resJ) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    if (h->ndrude == 0) return fail(NKC_ERR_ARG, "no Drude region set");
    HIPCHK(hipStreamSynchronize(h->stream));
    const int n3 = h->m.n1 * h->m.n1 * h->m.n1;
    const size_t nb = sizeof(double) * h->ndelem * 3 * n3;
    std::vector<double> buf(h->ndelem * 3 * n3);
    double *dst[3] = {J, kJ, resJ};
    double *srcp[3] = {h->dJ, h->dkJ, h->dresJ};
    for (int w = 0; w < 3; ++w) {
        if (!dst[w]) continue;
        HIPCHK(hipMemcpy(buf.data(), srcp[w], nb, hipMemcpyDeviceToHost));
        for (int64_t l = 0; l < h->ndrude; ++l)
            for (int c = 0; c < 3; ++c) dst[w][c * h->ndrude + l] = buf[h->drude_slot[l] + c * n3];
    }
    return NKC_OK;
}

int32_t nkc_set_coords(nkc_handle *h, const double *xm1, const double *ym1, const double *zm1) {
    if (!h || !xm1 || !ym1 || !zm1) return fail(NKC_ERR_ARG, "null handle or coordinate array");
    const int64_t fs = h->m.fs, npts = h->m.npts;
    std::vector<double> xyz(3 * (size_t)fs, 0.0);
    std::copy(xm1, xm1 + npts, xyz.begin());
    std::copy(ym1, ym1 + npts, xyz.begin() + fs);
    std::copy(zm1, zm1 + npts, xyz.begin() + 2 * fs);
    HIPCHK(hipStreamSynchronize(h->stream));
    h->hxyz = xyz;
    if (h->xyz) {
        // in place: captured steps that read the coordinates stay valid
        HIPCHK(hipMemcpy(h->xyz, xyz.data(), sizeof(double) * xyz.size(), hipMemcpyHostToDevice));
        return NKC_OK;
    }
    return upload(&h->xyz, xyz);
}

int32_t nkc_init_solution(nkc_handle *h, int32_t sol, double time) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    if (sol != SOL_BOXPER && sol != SOL_BOXPEC) return fail(NKC_ERR_ARG, "unknown solution %d", sol);
    int32_t rc;
    if ((rc = ensure_coords(h))) return rc;
    const DevMesh &m = h->m;
    const int nb = (int)std::min<int64_t>(DIAG_BLOCKS, (m.npts + DIAG_BS - 1) / DIAG_BS);
    hipLaunchKernelGGL(init_solution, dim3(nb), dim3(DIAG_BS), 0, h->stream, sol, time, m.u[h->cur], m.k, m.fs,
                       m.npts, (const double *)h->xyz);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    return NKC_OK;
}

namespace {

// Workgroups of the cem_error partials kernel for this handle's mesh.
int error_blocks(const nkc_handle *h) {
    return (int)std::min<int64_t>(DIAG_BLOCKS, (h->m.npts + DIAG_BS - 1) / DIAG_BS);
}

// cem_error's two kernels with the partial norms in part (error_blocks(h)
// x DIAG_NV doubles, every entry written by the first kernel before the
// second reads it) and the DIAG_NV totals in dout; out[12] on the host.
int32_t error_launch(nkc_handle *h, int sol, double time, const double *exact, double *part, double *dout,
                     double *out) {
    const DevMesh &m = h->m;
    const int nb = error_blocks(h);
    hipLaunchKernelGGL(error_partials, dim3(nb), dim3(DIAG_BS), 0, h->stream, sol, time, (const double *)m.u[h->cur],
                       exact, m.fs, m.npts, m.n1, (const double *)h->xyz, (const double *)h->djac,
                       (const double *)h->dw, (int)m.general, part);
    hipLaunchKernelGGL(error_final, dim3(1), dim3(64), 0, h->stream, (const double *)part, nb, dout);
    HIPCHK(hipGetLastError());
    double v[DIAG_NV];
    HIPCHK(hipMemcpyAsync(v, dout, sizeof v, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int c = 0; c < 6; ++c) {
        const double r = v[c] / v[12];
        out[c] = r > 0.0 ? std::sqrt(r) : r;
        out[6 + c] = v[6 + c];
    }
    return NKC_OK;
}

int32_t run_error(nkc_handle *h, int sol, double time, const double *exact, double *out) {
    int32_t rc;
    if ((rc = ensure_diag(h))) return rc;
    return error_launch(h, sol, time, exact, h->dpart, h->dout, out);
}

}  // namespace

int32_t nkc_cem_error(nkc_handle *h, int32_t sol, double time, double *out) {
    if (!h || !out) return fail(NKC_ERR_ARG, "null handle or output");
    if (sol != SOL_BOXPER && sol != SOL_BOXPEC) return fail(NKC_ERR_ARG, "unknown solution %d", sol);
    int32_t rc;
    if ((rc = ensure_coords(h))) return rc;
    return run_error(h, sol, time, nullptr, out);
}

int32_t nkc_cem_error_fields(nkc_handle *h, const double *hx, const double *ex, int64_t ld, double *out) {
    if (!h || !hx || !ex || !out) return fail(NKC_ERR_ARG, "null argument");
    const DevMesh &m = h->m;
    if (ld < m.npts) return fail(NKC_ERR_ARG, "ld < npts");
    double *tmp = nullptr;
    int32_t rc = dev_alloc(&tmp, 6 * (size_t)m.fs);
    if (rc) return rc;
    if (!(rc = copy_fields_h2d(tmp, m.fs, 0, hx, ex, m.npts, ld))) rc = run_error(h, SOL_FIELDS, 0.0, tmp, out);
    (void)hipFree(tmp);
    return rc;
}

}  // extern "C"

// ---- scratch of the no-allocation cem_error ------------------------------
// The partial norms (at most DIAG_BLOCKS x DIAG_NV doubles) and the DIAG_NV
// totals of nkc_cem_error_noalloc live in this array, one copy per GPU for
// the life of the process (107 KB).  Host code never passes its name to a
// kernel: the device address comes from hipGetSymbolAddress, once per
// device, and is kept in scratch_dev.
namespace {
constexpr int64_t SCRATCH_PART = (int64_t)DIAG_BLOCKS * DIAG_NV;
constexpr int64_t SCRATCH_DOUBLES = SCRATCH_PART + DIAG_NV;
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex scratch_mu;
double *scratch_dev[SCRATCH_MAX_DEVICES] = {};
}  // namespace

static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

namespace {

// Device address of `scratch` on device dev, looked up with that device
// current the first time and cached.
int32_t scratch_ptr(int dev, double **out) {
    if (dev < 0 || dev >= SCRATCH_MAX_DEVICES) return fail(NKC_ERR_ARG, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        int prev = 0;
        HIPCHK(hipGetDevice(&prev));
        if (prev != dev) HIPCHK(hipSetDevice(dev));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (prev != dev) HIPCHK(hipSetDevice(prev));
        if (e != hipSuccess || !p)
            return fail(NKC_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        scratch_dev[dev] = static_cast<double *>(p);
    }
    *out = scratch_dev[dev];
    return NKC_OK;
}

}  // namespace

extern "C" {

int32_t nkc_cem_error_noalloc(nkc_handle *h, int32_t sol, double time, double *out) {
    if (!h || !out) return fail(NKC_ERR_ARG, "null handle or output");
    if (sol != SOL_BOXPER && sol != SOL_BOXPEC) return fail(NKC_ERR_ARG, "unknown solution %d", sol);
    if (!h->xyz)
        return fail(NKC_ERR_ARG, "no node coordinates on the device: nkc_set_coords (or, on a box handle, one "
                                 "nkc_cem_error or nkc_init_solution call) must come first");
    const int64_t need = (int64_t)error_blocks(h) * DIAG_NV;
    if (need > SCRATCH_PART)
        return fail(NKC_ERR_ARG, "%lld partial norms exceed the scratch of %lld", (long long)need,
                    (long long)SCRATCH_PART);
    double *sp = nullptr;
    const int32_t rc = scratch_ptr(h->device, &sp);
    if (rc) return rc;
    return error_launch(h, sol, time, nullptr, sp, sp + SCRATCH_PART, out);
}

int32_t nkc_set_lorentz(nkc_handle *h, int64_t nlor, const int32_t *lindex, const double *params,
                        const double *L, const double *kL) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    if (h->m.general && (nlor > 0)) return general_refuses(h, "a Lorentz region");
    HIPCHK(hipStreamSynchronize(h->stream));
    free_adjoint(h);
    h->m.layout = -1;                     // the step's kernel instance may change: tune again
    (void)hipFree(h->lor_off); h->lor_off = nullptr;
    (void)hipFree(h->lL); h->lL = nullptr;
    (void)hipFree(h->lkL); h->lkL = nullptr;
    (void)hipFree(h->lpar); h->lpar = nullptr;
    (void)hipFree(h->lresL); h->lresL = nullptr;
    h->nlor = 0;
    h->nlelem = 0;
    h->lor_slot.clear();
    int32_t rc = drop_graphs(h);
    if (rc || nlor == 0) return rc;
    if (nlor < 0 || !lindex || !params) return fail(NKC_ERR_ARG, "bad Lorentz node list");
    const DevMesh &m = h->m;
    const int n3 = m.n1 * m.n1 * m.n1;
    std::vector<int32_t> off(m.nelt, -1);
    int32_t nle = 0;
    for (int64_t l = 0; l < nlor; ++l) {
        const int64_t v = (int64_t)lindex[l] - 1;
        if (v < 0 || v >= m.npts) return fail(NKC_ERR_ARG, "lindex[%lld]=%d out of range", (long long)l, lindex[l]);
        const int64_t e = v / n3;
        if (off[e] < 0) off[e] = nle++;
    }
    std::vector<double> par((size_t)nle * n3 * 3, 0.0), Lv((size_t)nle * 6 * n3, 0.0), kv = Lv;
    h->lor_slot.resize(nlor);
    for (int64_t l = 0; l < nlor; ++l) {
        const int64_t v = (int64_t)lindex[l] - 1, e = v / n3, vl = v % n3;
        const int64_t b = (int64_t)off[e] * 6 * n3 + vl;
        h->lor_slot[l] = b;
        for (int q = 0; q < 3; ++q) par[((int64_t)off[e] * n3 + vl) * 3 + q] = params[3 * l + q];
        for (int c = 0; c < 6; ++c) {
            if (L) Lv[b + c * n3] = L[c * nlor + l];
            if (kL) kv[b + c * n3] = kL[c * nlor + l];
        }
    }
    if ((rc = upload(&h->lpar, par))) return rc;
    if ((rc = upload(&h->lL, Lv))) return rc;
    if ((rc = upload(&h->lkL, kv))) return rc;
    if ((rc = upload(&h->lresL, kv))) return rc;
    if ((rc = upload(&h->lor_off, off))) return rc;     // last: switches the term on
    h->nlor = nlor;
    h->nlelem = nle;
    return NKC_OK;
}

int32_t nkc_get_lorentz(nkc_handle *h, double *L, double *kL, double *resL) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    if (h->nlor == 0) return fail(NKC_ERR_ARG, "no Lorentz region set");
    HIPCHK(hipStreamSynchronize(h->stream));
    const int n3 = h->m.n1 * h->m.n1 * h->m.n1;
    std::vector<double> buf(lorentz_size(h));
    double *dst[3] = {L, kL, resL};
    double *srcp[3] = {h->lL, h->lkL, h->lresL};
    for (int w = 0; w < 3; ++w) {
        if (!dst[w]) continue;
        HIPCHK(hipMemcpy(buf.data(), srcp[w], sizeof(double) * buf.size(), hipMemcpyDeviceToHost));
        for (int64_t l = 0; l < h->nlor; ++l)
            for (int c = 0; c < 6; ++c) dst[w][c * h->nlor + l] = buf[h->lor_slot[l] + c * n3];
    }
    return NKC_OK;
}

namespace {

void free_pml(nkc_handle *h) {
    (void)hipFree(h->pml_off); h->pml_off = nullptr;
    (void)hipFree(h->psig); h->psig = nullptr;
    (void)hipFree(h->pP); h->pP = nullptr;
    (void)hipFree(h->pkP); h->pkP = nullptr;
    (void)hipFree(h->presP); h->presP = nullptr;
    h->npml = 0;
    h->pml_elems.clear();
}

// Install a PML on the mesh elements `elems` (0-based, ascending): sig and
// P are compact (3 n3 and 6 n3 doubles per PML element).
int32_t install_pml(nkc_handle *h, const std::vector<int64_t> &elems, const std::vector<double> &sig,
                    const std::vector<double> &P) {
    std::vector<int32_t> off(h->m.nelt, -1);
    for (size_t l = 0; l < elems.size(); ++l) off[elems[l]] = (int32_t)l;
    std::vector<double> zero(P.size(), 0.0);
    int32_t rc;
    if ((rc = upload(&h->psig, sig))) return rc;
    if ((rc = upload(&h->pP, P))) return rc;
    if ((rc = upload(&h->pkP, zero))) return rc;
    if ((rc = upload(&h->presP, zero))) return rc;
    if ((rc = upload(&h->pml_off, off))) return rc;     // last: switches the term on
    h->npml = (int64_t)elems.size();
    h->pml_elems = elems;
    return NKC_OK;
}

// Common start of the PML set-up entry points: nothing stored survives.
int32_t reset_pml(nkc_handle *h) {
    HIPCHK(hipStreamSynchronize(h->stream));
    free_adjoint(h);
    h->m.layout = -1;                     // the step's kernel instance may change: tune again
    free_pml(h);
    return drop_graphs(h);
}

}  // namespace

int32_t nkc_set_pml(nkc_handle *h, int64_t maxpml, const int32_t *pmlptr, const double *sigma, const double *bn,
                    const double *dn, int64_t ld) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    if (h->m.general && (maxpml > 0)) return general_refuses(h, "a PML");
    int32_t rc = reset_pml(h);
    if (rc || maxpml == 0) return rc;
    const DevMesh &m = h->m;
    if (maxpml < 0 || maxpml > m.nelt || !pmlptr || !sigma) return fail(NKC_ERR_ARG, "bad PML element list");
    if (ld < m.npts) return fail(NKC_ERR_ARG, "ld < npts");
    const int64_t n3 = (int64_t)m.n1 * m.n1 * m.n1;
    std::vector<int64_t> elems(maxpml);
    for (int64_t l = 0; l < maxpml; ++l) {
        elems[l] = (int64_t)pmlptr[l] - 1;
        if (elems[l] < 0 || elems[l] >= m.nelt) return fail(NKC_ERR_ARG, "pmlptr[%lld]=%d out of range", (long long)l, pmlptr[l]);
        if (l > 0 && elems[l] <= elems[l - 1]) return fail(NKC_ERR_ARG, "pmlptr must be ascending");
    }
    std::vector<double> sig((size_t)maxpml * 3 * n3), P((size_t)maxpml * 6 * n3, 0.0);
    for (int64_t l = 0; l < maxpml; ++l)
        for (int64_t v = 0; v < n3; ++v) {
            const int64_t g = elems[l] * n3 + v;
            for (int c = 0; c < 3; ++c) {
                sig[(l * 3 + c) * n3 + v] = sigma[c * ld + g];
                if (bn) P[(l * 6 + c) * n3 + v] = bn[c * ld + g];
                if (dn) P[(l * 6 + 3 + c) * n3 + v] = dn[c * ld + g];
            }
        }
    return install_pml(h, elems, sig, P);
}

int32_t nkc_setup_pml(nkc_handle *h, const int8_t *pml_face, int32_t thick, double order, double referr,
                      double *extents) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    if (h->m.general && (pml_face != nullptr)) return general_refuses(h, "a PML");
    int32_t rc = reset_pml(h);
    if (rc || !pml_face) return rc;
    if ((rc = ensure_coords(h))) return rc;
    // pml_errchk, the layers, their extents and sigma (mesh_tables.hpp)
    std::vector<int64_t> elems;
    std::vector<double> sig;
    double ext[12];
    std::string err;
    if ((rc = pml_tables(h->hm, h->hxyz.data(), h->m.fs, pml_face, thick, order, referr, elems, sig, ext, err)))
        return fail(rc, "%s", err.c_str());
    if (extents)
        for (int f = 0; f < 12; ++f) extents[f] = ext[f];
    if (elems.empty()) return NKC_OK;
    const std::vector<double> P(elems.size() * 6 * (size_t)h->m.n1 * h->m.n1 * h->m.n1, 0.0);
    return install_pml(h, elems, sig, P);
}

int32_t nkc_set_pml_fields(nkc_handle *h, const double *bn, const double *dn, int64_t ld) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    if (h->npml == 0) return fail(NKC_ERR_ARG, "no PML set");
    if (ld < h->m.npts) return fail(NKC_ERR_ARG, "ld < npts");
    HIPCHK(hipStreamSynchronize(h->stream));
    free_adjoint(h);
    const int64_t n3 = (int64_t)h->m.n1 * h->m.n1 * h->m.n1;
    std::vector<double> P(pml_size(h), 0.0);
    for (size_t l = 0; l < h->pml_elems.size(); ++l)
        for (int64_t v = 0; v < n3; ++v) {
            const int64_t g = h->pml_elems[l] * n3 + v;
            for (int c = 0; c < 3; ++c) {
                if (bn) P[(l * 6 + c) * n3 + v] = bn[c * ld + g];
                if (dn) P[(l * 6 + 3 + c) * n3 + v] = dn[c * ld + g];
            }
        }
    HIPCHK(hipMemcpy(h->pP, P.data(), sizeof(double) * P.size(), hipMemcpyHostToDevice));
    std::fill(P.begin(), P.end(), 0.0);
    HIPCHK(hipMemcpy(h->pkP, P.data(), sizeof(double) * P.size(), hipMemcpyHostToDevice));
    return NKC_OK;
}

int32_t nkc_get_pml(nkc_handle *h, int32_t which, double *out, int64_t ld) {
    if (!h || !out) return fail(NKC_ERR_ARG, "null handle or output");
    if (h->npml == 0) return fail(NKC_ERR_ARG, "no PML set");
    if (which < 0 || which > 3) return fail(NKC_ERR_ARG, "which %d not in 0..3", which);
    if (ld < h->m.npts) return fail(NKC_ERR_ARG, "ld < npts");
    HIPCHK(hipStreamSynchronize(h->stream));
    const int64_t n3 = (int64_t)h->m.n1 * h->m.n1 * h->m.n1;
    const int nc = which == 3 ? 3 : 6;
    std::vector<double> buf(h->npml * nc * n3);
    const double *src = which == 0 ? h->pP : which == 1 ? h->pkP : which == 2 ? h->presP : h->psig;
    HIPCHK(hipMemcpy(buf.data(), src, sizeof(double) * buf.size(), hipMemcpyDeviceToHost));
    for (size_t l = 0; l < h->pml_elems.size(); ++l)
        for (int c = 0; c < nc; ++c)
            for (int64_t v = 0; v < n3; ++v) out[c * ld + h->pml_elems[l] * n3 + v] = buf[(l * nc + c) * n3 + v];
    return NKC_OK;
}

namespace {

// sqrt(eps/mu) at in-element node v of element e (yconduc of
// cem_maxwell_materials, src/cem_maxwell.F:276-286).
double node_conduc(const HostMesh &hm, int64_t e, int64_t v) {
    const int64_t n3 = (int64_t)hm.n1 * hm.n1 * hm.n1;
    const int64_t i = hm.mat ? e * n3 + v : e;
    return std::sqrt(hm.imu[i] / hm.ieps[i]);
}

// Yfac = 0.5/Y_0 and yconduc of every node of every sheet block into gpar
// slots 12, 13 (Y_0 as cem_maxwell_materials forms it: the face sum of
// yconduc, doubled on PEC faces, halved).
void sheet_coefs(const nkc_handle *h, std::vector<double> &par) {
    const HostMesh &hm = h->hm;
    const int n1 = hm.n1, n2 = n1 * n1;
    for (size_t b = 0; b < h->gr_face.size(); ++b) {
        const int64_t e = h->gr_face[b] / 6;
        const int kf = (int)(h->gr_face[b] % 6);
        const int32_t nb = hm.nbr[e * 6 + kf];
        for (int q = 0; q < n2; ++q) {
            const double ym = node_conduc(hm, e, face_node(n1, kf, q));
            double sum = ym;
            if (nb >= 0) sum = ym + node_conduc(hm, nb, face_node(n1, kf ^ 1, q));
            else if (nb == HOST_FACE_PEC) sum = 2.0 * ym;
            const double Y0 = 0.5 * sum;
            par[(b * 14 + 12) * n2 + q] = 0.5 / Y0;
            par[(b * 14 + 13) * n2 + q] = ym;
        }
    }
}

void free_sheet(nkc_handle *h) {
    (void)hipFree(h->gr_off); h->gr_off = nullptr;
    (void)hipFree(h->gpar); h->gpar = nullptr;
    (void)hipFree(h->gA[0]); h->gA[0] = nullptr;
    (void)hipFree(h->gA[1]); h->gA[1] = nullptr;
    (void)hipFree(h->gK); h->gK = nullptr;
    h->gr_face.clear();
    h->gr_list.clear();
    h->ngraph = 0;
}

}  // namespace

int32_t nkc_set_graphene(nkc_handle *h, int64_t ngraph, const int32_t *graphindex, const double *graphparams,
                         const double *fjn, int64_t ld) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    if (h->m.general && (ngraph > 0)) return general_refuses(h, "a graphene sheet");
    HIPCHK(hipStreamSynchronize(h->stream));
    free_adjoint(h);
    h->m.layout = -1;                     // the step's kernel instance may change: tune again
    free_sheet(h);
    int32_t rc = drop_graphs(h);
    if (rc || ngraph == 0) return rc;
    const int64_t nelt = h->m.nelt;
    const int n1 = h->m.n1, n2 = n1 * n1;
    const int64_t nxzfl = nelt * 6 * n2;
    if (ngraph < 0 || !graphindex || !graphparams || !fjn) return fail(NKC_ERR_ARG, "bad graphene point list");
    if (ld < nxzfl) return fail(NKC_ERR_ARG, "ld < nxzfl (%lld)", (long long)nxzfl);
    // blocks: every face side with a listed point, and the side across it
    std::vector<int32_t> off(nelt * 6, -1);
    std::vector<int64_t> face;
    auto block = [&](int64_t ef) {
        if (off[ef] < 0) {
            off[ef] = (int32_t)face.size();
            face.push_back(ef);
        }
        return off[ef];
    };
    const int *kf_ed = kf_of_ed();
    std::vector<char> seen(nxzfl, 0);
    std::vector<std::array<int64_t, 3>> list;
    list.reserve(ngraph);
    for (int64_t l = 0; l < ngraph; ++l) {
        const int64_t j = (int64_t)graphindex[l] - 1;
        if (j < 0 || j >= nxzfl) return fail(NKC_ERR_ARG, "graphindex(%lld) = %d out of range", (long long)l + 1, graphindex[l]);
        if (seen[j]) return fail(NKC_ERR_ARG, "graphindex(%lld) = %d listed twice", (long long)l + 1, graphindex[l]);
        seen[j] = 1;
        const int64_t e = j / (6 * n2);
        const int kf = kf_ed[(j / n2) % 6 + 1];
        const int q = (int)(j % n2);
        for (int c = 0; c < 12; ++c)
            if (!std::isfinite(graphparams[c * ld + j])) return fail(NKC_ERR_ARG, "graphparams(%lld, %d) not finite", (long long)j + 1, c + 1);
        for (int c = 0; c < 15; ++c)
            if (!std::isfinite(fjn[j + ld * (3 + c)])) return fail(NKC_ERR_ARG, "fjn(%lld, %d, %d) not finite", (long long)j + 1, c % 3 + 1, c / 3 + 2);
        const int32_t b = block(e * 6 + kf);
        const int32_t nb = h->hm.nbr[e * 6 + kf];
        if (nb >= 0) block((int64_t)nb * 6 + (kf ^ 1));
        list.push_back({j, b, q});
    }
    const size_t nb2 = face.size();
    std::vector<double> par(nb2 * 14 * n2, 0.0), st(nb2 * 15 * n2, 0.0);
    for (const auto &x : list) {
        const int64_t j = x[0], b = x[1], q = x[2];
        for (int c = 0; c < 12; ++c) par[(b * 14 + c) * n2 + q] = graphparams[c * ld + j];
        // fjn(ld, 3, 6): slots 2..6 are the state
        for (int sl = 0; sl < 5; ++sl)
            for (int c = 0; c < 3; ++c) st[(b * 15 + 3 * sl + c) * n2 + q] = fjn[j + ld * (c + 3 * (sl + 1))];
    }
    h->gr_face = face;
    sheet_coefs(h, par);
    const std::vector<double> zero(st.size(), 0.0);
    if ((rc = upload(&h->gpar, par)) || (rc = upload(&h->gA[h->cur], st)) || (rc = upload(&h->gA[h->cur ^ 1], st)) ||
        (rc = upload(&h->gK, zero))) {
        free_sheet(h);
        return rc;
    }
    if ((rc = upload(&h->gr_off, off))) {   // last: a non-null gr_off means the sheet is whole
        free_sheet(h);
        return rc;
    }
    h->gr_list = std::move(list);
    h->ngraph = ngraph;
    return NKC_OK;
}

int32_t nkc_get_graphene(nkc_handle *h, int32_t which, double *fjn, int64_t ld) {
    if (!h || !fjn) return fail(NKC_ERR_ARG, "null handle or output");
    if (!h->ngraph) return fail(NKC_ERR_ARG, "no graphene sheet set");
    if (which < 0 || which > 1) return fail(NKC_ERR_ARG, "which must be 0 (state) or 1 (RK register)");
    const int n2 = h->m.n1 * h->m.n1;
    if (ld < h->m.nelt * 6 * n2) return fail(NKC_ERR_ARG, "ld < nxzfl");
    HIPCHK(hipStreamSynchronize(h->stream));
    std::vector<double> buf(sheet_size(h));
    HIPCHK(hipMemcpy(buf.data(), which == 0 ? h->gA[h->cur] : h->gK, sizeof(double) * buf.size(),
                     hipMemcpyDeviceToHost));
    for (const auto &x : h->gr_list) {
        const int64_t j = x[0], b = x[1], q = x[2];
        for (int sl = 0; sl < 5; ++sl)
            for (int c = 0; c < 3; ++c) fjn[j + ld * (c + 3 * (sl + 1))] = buf[(b * 15 + 3 * sl + c) * n2 + q];
    }
    return NKC_OK;
}

int32_t nkc_set_incident(nkc_handle *h, int64_t nface, const int32_t *faces, const double k[3], double omega,
                         const double hamp[3], const double eamp[3]) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    if (h->m.general && (nface > 0)) return general_refuses(h, "an incident field");
    HIPCHK(hipStreamSynchronize(h->stream));
    free_adjoint(h);
    h->m.layout = -1;                     // the step's kernel instance may change: tune again
    (void)hipFree(h->inc_mask); h->inc_mask = nullptr;
    h->ninc = 0;
    int32_t rc = drop_graphs(h);
    if (rc || nface == 0) return rc;
    if (nface < 0 || !faces || !k || !hamp || !eamp) return fail(NKC_ERR_ARG, "bad incident-field face list");
    for (int c = 0; c < 3; ++c)
        if (!std::isfinite(k[c]) || !std::isfinite(hamp[c]) || !std::isfinite(eamp[c]))
            return fail(NKC_ERR_ARG, "incident-field parameters not finite");
    if ((rc = ensure_coords(h))) return rc;
    const int64_t nelt = h->m.nelt;
    const int *kf_ed = kf_of_ed();
    std::vector<int32_t> mask(nelt, 0);
    for (int64_t l = 0; l < nface; ++l) {
        const int64_t e = (int64_t)faces[2 * l] - 1;
        const int f = faces[2 * l + 1];
        if (e < 0 || e >= nelt || f < 1 || f > 6) return fail(NKC_ERR_ARG, "incident face %lld: (%d, %d) out of range", (long long)l, faces[2 * l], f);
        const int kf = kf_ed[f];
        mask[e] |= 1 << kf;
        const int32_t nb = h->hm.nbr[e * 6 + kf];
        if (nb >= 0) mask[nb] |= 1 << (8 + (kf ^ 1));   // the neighbour sees this trace across its opposite face
    }
    if ((rc = upload(&h->inc_mask, mask))) return rc;
    for (int c = 0; c < 3; ++c) {
        h->inc_k[c] = k[c];
        h->inc_h[c] = hamp[c];
        h->inc_e[c] = eamp[c];
    }
    h->inc_om = omega;
    h->ninc = nface;
    return NKC_OK;
}

int32_t nkc_restart_write(nkc_handle *h, const char *path, int32_t istep, int32_t dumpno, double time, double dt,
                          int32_t single) {
    if (!h || !path) return fail(NKC_ERR_ARG, "null handle or path");
    int32_t rc;
    if ((rc = ensure_coords(h))) return rc;
    const DevMesh &m = h->m;
    std::vector<double> u(6 * (size_t)m.fs);
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(u.data(), m.u[h->cur], sizeof(double) * u.size(), hipMemcpyDeviceToHost));
    RestartHeader hd;
    hd.istep = istep;
    hd.dumpno = dumpno;
    hd.nx1 = m.n1;
    hd.np = 1;
    hd.time = time;
    hd.dt = dt;
    std::string err;
    if ((rc = restart_write(path, hd, m.n1, m.nelt, h->hxyz.data(), m.fs, u.data() + 3 * m.fs, u.data(), m.fs,
                            single == 0, err)))
        return fail(rc, "%s", err.c_str());
    return NKC_OK;
}

int32_t nkc_restart_read(nkc_handle *h, const char *path, int32_t *istep, int32_t *dumpno, double *time, double *dt) {
    if (!h || !path) return fail(NKC_ERR_ARG, "null handle or path");
    const DevMesh &m = h->m;
    std::vector<double> u(6 * (size_t)m.fs, 0.0);
    RestartHeader hd;
    std::string err;
    int32_t rc;
    if ((rc = restart_read(path, m.n1, m.nelt, hd, u.data() + 3 * m.fs, u.data(), m.fs, err)))
        return fail(rc, "%s", err.c_str());
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(m.u[h->cur], u.data(), sizeof(double) * u.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemset(m.k, 0, sizeof(double) * u.size()));
    if (istep) *istep = hd.istep;
    if (dumpno) *dumpno = hd.dumpno;
    if (time) *time = hd.time;
    if (dt) *dt = hd.dt;
    return NKC_OK;
}

int32_t nkc_check_finite(nkc_handle *h, double *maxabs) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    hipLaunchKernelGGL(zero_u64, dim3(1), dim3(1), 0, h->stream, h->red);
    hipLaunchKernelGGL(maxabs_kernel, dim3(1024), dim3(256), 0, h->stream, h->m.u[h->cur], h->m.fs,
                       h->m.npts, h->red);
    HIPCHK(hipGetLastError());
    unsigned long long bits = 0;
    HIPCHK(hipMemcpyAsync(&bits, h->red, 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    double mx;
    std::memcpy(&mx, &bits, 8);
    if (maxabs) *maxabs = mx;
    if (!std::isfinite(mx) || mx > 1e20) return fail(NKC_ERR_BLOWUP, "state blew up: max|u| = %g", mx);
    return NKC_OK;
}

int32_t nkc_sync(nkc_handle *h) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    return NKC_OK;
}

int32_t nkc_stream(nkc_handle *h, void **stream) {
    if (!h || !stream) return fail(NKC_ERR_ARG, "null handle or output");
    *stream = (void *)h->stream;
    return NKC_OK;
}

int32_t nkc_info(nkc_handle *h, int64_t out[8]) {
    if (!h || !out) return fail(NKC_ERR_ARG, "null handle or output");
    const int n1 = h->m.n1;
    out[0] = h->m.npts;
    out[1] = h->m.mat;
    out[2] = 1;
    out[3] = h->m.general ? GEO_BS : stage_block(n1, h->m.layout < 0 ? 0 : h->m.layout);
    out[4] = h->m.general ? (int64_t)sizeof(double) * (6 * n1 * n1 * n1 + 37 * n1 * n1)
                          : (int64_t)sizeof(double) * (6 * n1 * n1 * n1 + 25 * n1 * n1);
    out[5] = h->use_graph ? 1 : 0;
    out[6] = h->m.fs;
    out[7] = h->m.nelt;
    return NKC_OK;
}

int32_t nkc_set_graph(nkc_handle *h, int32_t enable) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    h->use_graph = enable != 0;
    if (!h->use_graph) return drop_graphs(h);
    return NKC_OK;
}

int32_t nkc_set_keep_k(nkc_handle *h, int32_t keep) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    if (h->keep_k == (keep != 0)) return NKC_OK;
    h->keep_k = keep != 0;
    return drop_graphs(h);   // the captured step carries the stage flags
}

int32_t nkc_set_layout(nkc_handle *h, int32_t layout) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    if (h->m.general && (layout > 0)) return general_refuses(h, "a stage layout other than 0");
    if (layout < -1 || layout >= NLAYOUT) return fail(NKC_ERR_ARG, "layout %d not in -1..%d", layout, NLAYOUT - 1);
    if (layout >= 0 && !layout_built(h->m.n1, layout))
        return fail(NKC_ERR_ARG, "layout %d is not built for N=%d", layout, h->m.n1 - 1);
    if (layout >= 0 && !layout_runs(h->m.n1, layout, base_args(h)))
        return fail(NKC_ERR_UNSUPPORTED, "layout %d carries no PML, incident field or graphene sheet", layout);
    HIPCHK(hipStreamSynchronize(h->stream));
    h->m.layout = layout;
    return drop_graphs(h);
}

int32_t nkc_get_layout(nkc_handle *h, double *out) {
    if (!h || !out) return fail(NKC_ERR_ARG, "null handle or output");
    for (int l = 0; l < NLAYOUT; ++l) out[l] = h->tune_us[l];
    out[NLAYOUT] = h->m.layout;
    return NKC_OK;
}

int32_t nkc_tune(nkc_handle *h, double *out) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    const StageArgs f = base_args(h);
    int32_t rc = tune_stage(h->m, &f, drude_size(h), lorentz_size(h), pml_size(h), sheet_size(h), h->m.u[h->cur], h->stream,
                       h->tune_us);
    if (rc) return rc;
    if ((rc = drop_graphs(h))) return rc;
    if (out) {
        for (int l = 0; l < NLAYOUT; ++l) out[l] = h->tune_us[l];
        out[NLAYOUT] = h->m.layout;
    }
    return NKC_OK;
}

int32_t nkc_set_materials(nkc_handle *h, const double *eps_in, const double *mu_in) {
    if (!h || !eps_in || !mu_in) return fail(NKC_ERR_ARG, "null handle or material array");
    if (h->m.general && (true)) return general_refuses(h, "replacing the materials");
    const int64_t npts = h->m.npts;
    // the tables are built on the host from the caller's arrays (host or device memory)
    std::vector<double> eps_h, mu_h;
    int32_t rc0;
    if ((rc0 = stage_host(eps_h, eps_in, npts)) || (rc0 = stage_host(mu_h, mu_in, npts))) return rc0;
    const double *eps = eps_h.data(), *mu = mu_h.data();
    for (int64_t p = 0; p < npts; ++p)
        if (!(eps[p] > 0.0) || !(mu[p] > 0.0)) return fail(NKC_ERR_ARG, "material not positive at node %lld", (long long)p);
    HIPCHK(hipStreamSynchronize(h->stream));
    free_adjoint(h);                      // the stored trajectory has the old materials
    // the new tables replace the old ones (host and device) only once all
    // three are on the device, so a failed upload leaves the handle as it was
    HostMesh hm = h->hm;
    set_node_materials(hm, eps, mu, h->C0);
    double *fco = nullptr, *ieps = nullptr, *imu = nullptr;
    int32_t rc;
    if ((rc = upload(&fco, hm.fco)) || (rc = upload(&ieps, hm.ieps)) || (rc = upload(&imu, hm.imu))) {
        (void)hipFree(fco); (void)hipFree(ieps); (void)hipFree(imu);
        return rc;
    }
    h->hm = std::move(hm);
    (void)hipFree(h->yz); h->yz = nullptr;
    (void)hipFree(h->m.fco); (void)hipFree(h->m.ieps); (void)hipFree(h->m.imu);
    h->m.fco = fco;
    h->m.ieps = ieps;
    h->m.imu = imu;
    h->m.mat = MAT_NODE;
    if (h->ngraph) {
        // the sheet's Yfac and yconduc follow the materials
        const size_t n2 = (size_t)h->m.n1 * h->m.n1;
        std::vector<double> par(h->gr_face.size() * 14 * n2);
        HIPCHK(hipMemcpy(par.data(), h->gpar, sizeof(double) * par.size(), hipMemcpyDeviceToHost));
        sheet_coefs(h, par);
        HIPCHK(hipMemcpy(h->gpar, par.data(), sizeof(double) * par.size(), hipMemcpyHostToDevice));
    }
    return drop_graphs(h);
}

namespace {

int32_t dev_alloc_counted(nkc_handle *h, double **p, size_t n) {
    h->adj.bytes += sizeof(double) * n;
    return dev_alloc(p, n);
}

// Every buffer of the adjoint for the layout fixed by nkc_adjoint_forward:
// segment checkpoints, window starts, the stage store of one window, the
// adjoint state and gradients.  The window S is the largest (at most every)
// whose stage store fits in 70 % of the free memory left after the rest.
int32_t alloc_adjoint(nkc_handle *h) {
    auto &A = h->adj;
    const DevMesh &m = h->m;
    const size_t blk = 6 * (size_t)m.fs, jsz = drude_size(h);
    int32_t rc;
    A.bytes = 0;
    if ((rc = dev_alloc_counted(h, &A.ckpt, blk * A.nseg))) return rc;
    if ((rc = dev_alloc_counted(h, &A.ufin, blk))) return rc;
    if ((rc = dev_alloc_counted(h, &A.kfin, blk))) return rc;
    for (int w = 0; w < 2; ++w) {
        if ((rc = dev_alloc_counted(h, &A.lam[w], blk))) return rc;
        if ((rc = dev_alloc_counted(h, &A.kap[w], blk))) return rc;
    }
    if ((rc = dev_alloc_counted(h, &A.geps, m.fs))) return rc;
    if ((rc = dev_alloc_counted(h, &A.gmu, m.fs))) return rc;
    if (h->ndrude) {
        if ((rc = dev_alloc_counted(h, &A.ckJ, jsz * A.nseg))) return rc;
        if ((rc = dev_alloc_counted(h, &A.Jfin, jsz))) return rc;
        if ((rc = dev_alloc_counted(h, &A.kJfin, jsz))) return rc;
        if ((rc = dev_alloc_counted(h, &A.lamJ, jsz))) return rc;
        if ((rc = dev_alloc_counted(h, &A.kapJ, jsz))) return rc;
        if ((rc = dev_alloc_counted(h, &A.gA, jsz / 3))) return rc;
        if ((rc = dev_alloc_counted(h, &A.gB, jsz / 3))) return rc;
    }
    const bool gm = (h->adj_grads & 1) != 0;   // eps/mu gradients need the stage residuals
    const size_t per_step = sizeof(double) * 5 * ((gm ? 2 : 1) * blk + jsz);
    int64_t S = A.every;
    if (h->adj_store_cap > 0) {
        S = std::min(S, h->adj_store_cap);
    } else {
        size_t fr = 0, tot = 0;
        HIPCHK(hipMemGetInfo(&fr, &tot));
        const size_t state = sizeof(double) * (blk + jsz);
        const size_t room = (size_t)(0.7 * (double)fr);
        // S steps of store plus ceil(every / S) window starts
        while (S > 1 && per_step * S + state * (size_t)((A.every + S - 1) / S + 1) > room) S = (S + 1) / 2;
    }
    A.S = S;
    A.nsub_max = (A.every + S - 1) / S;
    // whole segments fit: keep as many of the last ones as half of the free
    // memory holds, so the backward sweep recomputes only the others (the
    // other half stays free for co-resident allocations)
    int64_t K = 1;
    if (S == A.every && h->adj_store_cap == 0) {
        size_t fr = 0, tot = 0;
        HIPCHK(hipMemGetInfo(&fr, &tot));
        const size_t win = per_step * S + sizeof(double) * blk;
        K = std::max<int64_t>(1, std::min<int64_t>(A.nseg, (int64_t)((0.5 * (double)fr) / (double)win)));
    }
    A.K = K;
    A.slot_id.assign(K, -1);
    A.cur_slot = 0;
    if ((rc = dev_alloc_counted(h, &A.st_u, blk * (5 * S + 1) * K))) return rc;
    if (gm && (rc = dev_alloc_counted(h, &A.st_r, blk * 5 * S * K))) return rc;
    // every segment's window starts, when the windows split segments and
    // they fit in a quarter of the free memory: the backward then recomputes
    // each window once, with no pass to rebuild its segment's window starts
    A.sub_all = false;
    if (A.nsub_max > 1) {
        size_t fr = 0, tot = 0;
        HIPCHK(hipMemGetInfo(&fr, &tot));
        const size_t need = sizeof(double) * (blk + jsz) * (size_t)(A.nseg * A.nsub_max);
        A.sub_all = need <= fr / 4;
    }
    const size_t nsub = (size_t)((A.sub_all ? A.nseg : 1) * A.nsub_max);
    if ((rc = dev_alloc_counted(h, &A.sub, blk * nsub))) return rc;
    if (h->ndrude) {
        if ((rc = dev_alloc_counted(h, &A.stJ, jsz * 5 * S * K))) return rc;
        if ((rc = dev_alloc_counted(h, &A.subJ, jsz * nsub))) return rc;
    }
    return NKC_OK;
}

hipError_t d2d(nkc_handle *h, double *dst, const double *src, size_t n) {
    if (n == 0) return hipSuccess;
    return hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyDeviceToDevice, h->stream);
}

// Time after `step` steps from t0, summed as the forward sums it.
double step_time(const nkc_handle *h, int64_t step) {
    double t = h->adj.t0;
    for (int64_t s = 0; s < step; ++s) t = t + h->adj.dt;
    return t;
}

int32_t set_time(nkc_handle *h, double t) {
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(h->tdev, &t, sizeof(double), hipMemcpyHostToDevice));
    h->t_host = t;
    return NKC_OK;
}

// Steps in segment seg, in its window q, and windows in the segment.
int64_t seg_steps(const nkc_handle *h, int64_t seg) {
    return std::min(h->adj.every, h->adj.nsteps - seg * h->adj.every);
}
int64_t win_steps(const nkc_handle *h, int64_t seg, int64_t q) {
    return std::min(h->adj.S, seg_steps(h, seg) - q * h->adj.S);
}
int64_t nwin(const nkc_handle *h, int64_t seg) { return (seg_steps(h, seg) + h->adj.S - 1) / h->adj.S; }

double *slot_u(nkc_handle *h, int k) {
    return h->adj.st_u + (int64_t)k * (5 * h->adj.S + 1) * 6 * h->m.fs;
}

// Window q of segment seg through the stage store slot cur_slot, from the
// state now in that slot's first stage input (and dJ) at the device time of
// its first step.
int32_t store_window(nkc_handle *h, int64_t seg, int64_t q) {
    int32_t rc;
    const int64_t n = win_steps(h, seg, q);
    for (int64_t s = 0; s < n; ++s) {
        if ((rc = queue_step_store(h, h->adj.dt, s))) return rc;
        h->t_host = h->t_host + h->adj.dt;
    }
    h->adj.slot_id[h->adj.cur_slot] = seg * h->adj.nsub_max + q;
    return NKC_OK;
}

int find_slot(const nkc_handle *h, int64_t seg, int64_t q) {
    const int64_t id = seg * h->adj.nsub_max + q;
    for (int k = 0; k < (int)h->adj.K; ++k)
        if (h->adj.slot_id[k] == id) return k;
    return -1;
}

// Where window q of segment seg keeps its start state (and current).
int64_t sub_index(const nkc_handle *h, int64_t seg, int64_t q) {
    return (h->adj.sub_all ? seg * h->adj.nsub_max : 0) + q;
}

// Segment seg from the state in u[cur] (and dJ) at its start time: keep
// every window's start state, run all windows but the last with the plain
// step, and the last through the stage store.
int32_t run_segment_windows(nkc_handle *h, int64_t seg) {
    auto &A = h->adj;
    const size_t blk = 6 * (size_t)h->m.fs, jsz = drude_size(h);
    const int64_t nq = nwin(h, seg);
    int32_t rc;
    for (int64_t q = 0; q < nq; ++q) {
        HIPCHK(d2d(h, A.sub + sub_index(h, seg, q) * blk, h->m.u[h->cur], blk));
        if (h->ndrude) HIPCHK(d2d(h, A.subJ + sub_index(h, seg, q) * jsz, h->dJ, jsz));
        if (q + 1 < nq) {
            if ((rc = nkc_step(h, nullptr, A.dt, win_steps(h, seg, q)))) return rc;
        } else {
            A.cur_slot = 0;
            HIPCHK(d2d(h, slot_u(h, 0), h->m.u[h->cur], blk));
            if ((rc = store_window(h, seg, q))) return rc;
        }
    }
    A.sub_seg = seg;
    return NKC_OK;
}

// Reverse the stages of window q of segment seg (held by the stage store).
int32_t reverse_window(nkc_handle *h, int64_t seg, int64_t q, int &li) {
    auto &A = h->adj;
    const DevMesh &m = h->m;
    const size_t blk = 6 * (size_t)m.fs, jsz = drude_size(h);
    int32_t rc;
    for (int64_t s = win_steps(h, seg, q) - 1; s >= 0; --s)
        for (int st = 4; st >= 0; --st) {
            AdjArgs a{};
            a.lam_in = A.lam[li];
            a.kap_in = A.kap[li];
            a.lam_out = A.lam[li ^ 1];
            a.kap_out = A.kap[li ^ 1];
            a.us = slot_u(h, A.cur_slot) + (s * 5 + st) * blk;
            a.rs = A.st_r ? A.st_r + ((int64_t)A.cur_slot * 5 * A.S + s * 5 + st) * blk : nullptr;
            a.geps = A.geps;
            a.gmu = A.gmu;
            a.rka = RK4A[st];
            a.rkb = RK4B[st];
            a.dt = A.dt;
            a.C0 = h->C0;
            a.yz = h->yz;
            if (h->ndrude) {
                a.drude_off = h->drude_off;
                a.dpar = h->dpar;
                a.lamJ = A.lamJ;
                a.kapJ = A.kapJ;
                a.stJ = A.stJ + ((int64_t)A.cur_slot * 5 * A.S + s * 5 + st) * jsz;
                a.gA = A.gA;
                a.gB = A.gB;
            }
            if ((rc = launch_adjoint(m, a, h->stream))) return rc;
            li ^= 1;
        }
    return NKC_OK;
}

}  // namespace

int32_t nkc_adjoint_set_store(nkc_handle *h, int64_t max_steps) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    if (max_steps < 0) return fail(NKC_ERR_ARG, "max_steps < 0");
    h->adj_store_cap = max_steps;
    return NKC_OK;
}

int32_t nkc_adjoint_release(nkc_handle *h) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    free_adjoint(h);
    return NKC_OK;
}

int32_t nkc_adjoint_set_grads(nkc_handle *h, int32_t material) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    h->adj_grads = (material ? 1 : 0) | 2;
    return NKC_OK;
}

int32_t nkc_adjoint_info(nkc_handle *h, int64_t *out) {
    if (!h || !out) return fail(NKC_ERR_ARG, "null argument");
    const auto &A = h->adj;
    out[0] = A.every;
    out[1] = A.nseg;
    out[2] = A.S;
    out[3] = A.K;
    out[4] = (int64_t)A.bytes;
    return NKC_OK;
}

int32_t nkc_adjoint_forward(nkc_handle *h, double *time, double dt, int64_t nsteps, int64_t every) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    if (h->m.general && (true)) return general_refuses(h, "the adjoint");
    if (nsteps <= 0 || every <= 0 || !std::isfinite(dt)) return fail(NKC_ERR_ARG, "bad adjoint forward arguments");
    if (h->nlor) return fail(NKC_ERR_UNSUPPORTED, "the adjoint of a Lorentz region is not built");
    if (h->npml) return fail(NKC_ERR_UNSUPPORTED, "the adjoint of a PML is not built");
    if (h->ninc) return fail(NKC_ERR_UNSUPPORTED, "the adjoint with an incident field is not built");
    if (h->ngraph) return fail(NKC_ERR_UNSUPPORTED, "the adjoint with a graphene sheet is not built");
    free_adjoint(h);
    auto &A = h->adj;
    A.nsteps = nsteps;
    A.every = every < nsteps ? every : nsteps;
    A.nseg = (nsteps + A.every - 1) / A.every;
    A.dt = dt;
    A.t0 = time ? *time : h->t_host;
    const size_t blk = 6 * (size_t)h->m.fs, jsz = drude_size(h);
    int32_t rc;
    if ((rc = alloc_adjoint(h))) {
        free_adjoint(h);
        return rc;
    }
    if (!(A.t0 == h->t_host) && (rc = set_time(h, A.t0))) return rc;
    // the last K segments (K slots of whole segments), or the last segment's
    // last window (K = 1), run through the stage store, so that the backward
    // sweep starts without recomputing them
    const int64_t first_kept = A.nseg - A.K;
    const double *fin = nullptr;
    for (int64_t seg = 0; seg < A.nseg; ++seg) {
        if (fin) HIPCHK(d2d(h, h->m.u[h->cur], fin, blk));   // state after the previous kept segment
        HIPCHK(d2d(h, A.ckpt + seg * blk, h->m.u[h->cur], blk));
        if (h->ndrude) HIPCHK(d2d(h, A.ckJ + seg * jsz, h->dJ, jsz));
        if (seg < first_kept && A.sub_all && nwin(h, seg) > 1) {
            // keep the segment's window starts for the backward sweep
            for (int64_t q = 0; q < nwin(h, seg); ++q) {
                HIPCHK(d2d(h, A.sub + sub_index(h, seg, q) * blk, h->m.u[h->cur], blk));
                if (h->ndrude) HIPCHK(d2d(h, A.subJ + sub_index(h, seg, q) * jsz, h->dJ, jsz));
                if ((rc = nkc_step(h, nullptr, dt, win_steps(h, seg, q)))) return rc;
            }
        } else if (seg < first_kept) {
            if ((rc = nkc_step(h, nullptr, dt, A.every))) return rc;
        } else if (A.K > 1) {
            A.cur_slot = (int)(seg - first_kept);
            HIPCHK(d2d(h, slot_u(h, A.cur_slot), h->m.u[h->cur], blk));
            if ((rc = store_window(h, seg, 0))) return rc;
            fin = slot_u(h, A.cur_slot) + 5 * seg_steps(h, seg) * blk;
        } else {
            if ((rc = run_segment_windows(h, seg))) return rc;
            fin = slot_u(h, 0) + 5 * win_steps(h, seg, nwin(h, seg) - 1) * blk;
        }
    }
    HIPCHK(d2d(h, h->m.u[h->cur], fin, blk));
    HIPCHK(d2d(h, A.ufin, fin, blk));
    HIPCHK(d2d(h, A.kfin, h->m.k, blk));
    if (h->ndrude) {
        HIPCHK(d2d(h, A.Jfin, h->dJ, jsz));
        HIPCHK(d2d(h, A.kJfin, h->dkJ, jsz));
    }
    if (time) *time = h->t_host;
    return NKC_OK;
}

int32_t nkc_adjoint_backward(nkc_handle *h, const double *lamH, const double *lamE, int64_t ld,
                             double *geps, double *gmu) {
    if (!h || !lamH || !lamE) return fail(NKC_ERR_ARG, "null handle or adjoint seed");
    auto &A = h->adj;
    if (A.nseg == 0) return fail(NKC_ERR_ARG, "no nkc_adjoint_forward run to differentiate");
    DevMesh &m = h->m;
    if (ld < m.npts) return fail(NKC_ERR_ARG, "ld < npts");
    const size_t blk = 6 * (size_t)m.fs, jsz = drude_size(h);
    const double t_end = h->t_host;
    int32_t rc;
    if ((rc = ensure_yz(h))) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemset(A.lam[0], 0, sizeof(double) * blk));
    HIPCHK(hipMemset(A.kap[0], 0, sizeof(double) * blk));
    HIPCHK(hipMemset(A.geps, 0, sizeof(double) * m.fs));
    HIPCHK(hipMemset(A.gmu, 0, sizeof(double) * m.fs));
    if (h->ndrude) {
        HIPCHK(hipMemset(A.lamJ, 0, sizeof(double) * jsz));
        HIPCHK(hipMemset(A.kapJ, 0, sizeof(double) * jsz));
        HIPCHK(hipMemset(A.gA, 0, sizeof(double) * jsz / 3));
        HIPCHK(hipMemset(A.gB, 0, sizeof(double) * jsz / 3));
    }
    if ((rc = copy_fields_h2d(A.lam[0], m.fs, 0, lamH, lamE, m.npts, ld))) return rc;
    // the RK registers restart from zero at a segment or window start (the
    // first stage multiplies them by a_1 = 0)
    auto start_from = [&](const double *J, int64_t step) -> int32_t {
        if (hipMemsetAsync(m.k, 0, sizeof(double) * blk, h->stream) != hipSuccess ||
            (h->ndrude && (hipMemsetAsync(h->dkJ, 0, sizeof(double) * jsz, h->stream) != hipSuccess ||
                           d2d(h, h->dJ, J, jsz) != hipSuccess)))
            return fail(NKC_ERR_HIP, "adjoint restart copies");
        return set_time(h, step_time(h, step));
    };
    int li = 0;
    for (int64_t seg = A.nseg - 1; seg >= 0; --seg) {
        const int64_t nq = nwin(h, seg);
        for (int64_t q = nq - 1; q >= 0; --q) {
            int k = find_slot(h, seg, q);
            if (k < 0 && nq > 1 && !A.sub_all && A.sub_seg != seg) {
                // window starts of this segment (and its last window's stages)
                HIPCHK(d2d(h, m.u[h->cur], A.ckpt + seg * blk, blk));
                if ((rc = start_from(A.ckJ + seg * jsz, seg * A.every))) return rc;
                if ((rc = run_segment_windows(h, seg))) return rc;
                k = find_slot(h, seg, q);
            }
            if (k < 0) {
                // recompute the window into slot 0 (the kept segments, which
                // come last in time, are reversed before it is reused)
                A.cur_slot = 0;
                const double *u0 = nq > 1 ? A.sub + sub_index(h, seg, q) * blk : A.ckpt + seg * blk;
                const double *J0 = nq > 1 ? A.subJ + sub_index(h, seg, q) * jsz : A.ckJ + seg * jsz;
                HIPCHK(d2d(h, slot_u(h, 0), u0, blk));
                if ((rc = start_from(J0, seg * A.every + q * A.S))) return rc;
                if ((rc = store_window(h, seg, q))) return rc;
                k = 0;
            }
            A.cur_slot = k;
            if ((rc = reverse_window(h, seg, q, li))) return rc;
        }
    }
    // restore the forward's final state, RK registers and time
    HIPCHK(d2d(h, m.u[h->cur], A.ufin, blk));
    HIPCHK(d2d(h, m.k, A.kfin, blk));
    if (h->ndrude) {
        HIPCHK(d2d(h, h->dJ, A.Jfin, jsz));
        HIPCHK(d2d(h, h->dkJ, A.kJfin, jsz));
    }
    if ((rc = set_time(h, t_end))) return rc;
    A.have_grad = true;
    if (geps) HIPCHK(hipMemcpy(geps, A.geps, sizeof(double) * m.npts, hipMemcpyDefault));
    if (gmu) HIPCHK(hipMemcpy(gmu, A.gmu, sizeof(double) * m.npts, hipMemcpyDefault));
    return NKC_OK;
}

int32_t nkc_adjoint_get_drude(nkc_handle *h, double *ga, double *gb) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    if (h->ndrude == 0) return fail(NKC_ERR_ARG, "no Drude region set");
    if (!h->adj.gA || !h->adj.have_grad) return fail(NKC_ERR_ARG, "no nkc_adjoint_backward run with this Drude region");
    HIPCHK(hipStreamSynchronize(h->stream));
    const int64_t n3 = (int64_t)h->m.n1 * h->m.n1 * h->m.n1;
    std::vector<double> buf(drude_size(h) / 3), out((size_t)h->ndrude);
    double *dst[2] = {ga, gb};
    const double *srcp[2] = {h->adj.gA, h->adj.gB};
    for (int w = 0; w < 2; ++w) {
        if (!dst[w]) continue;
        HIPCHK(hipMemcpy(buf.data(), srcp[w], sizeof(double) * buf.size(), hipMemcpyDeviceToHost));
        for (int64_t l = 0; l < h->ndrude; ++l) {
            const int64_t b = h->drude_slot[l];          // dd 3 n3 + vloc
            out[l] = buf[(b / (3 * n3)) * n3 + b % (3 * n3)];
        }
        HIPCHK(hipMemcpy(dst[w], out.data(), sizeof(double) * out.size(), hipMemcpyDefault));   // host or device
    }
    return NKC_OK;
}

int32_t nkc_adjoint_apply(nkc_handle *h, const double *kH, const double *kE, double *aH, double *aE, int64_t ld) {
    if (!h || !kH || !kE || !aH || !aE) return fail(NKC_ERR_ARG, "null argument");
    if (h->m.general && (true)) return general_refuses(h, "the adjoint");
    DevMesh &m = h->m;
    if (ld < m.npts) return fail(NKC_ERR_ARG, "ld < npts");
    const size_t blk = 6 * (size_t)m.fs;
    double *zero = nullptr, *kap = nullptr, *out = nullptr, *kout = nullptr, *g = nullptr;
    int32_t rc = NKC_OK;
    do {
        if ((rc = ensure_yz(h))) break;
        if ((rc = dev_alloc(&zero, blk)) || (rc = dev_alloc(&kap, blk)) || (rc = dev_alloc(&out, blk)) ||
            (rc = dev_alloc(&kout, blk)) || (rc = dev_alloc(&g, 2 * m.fs)))
            break;
        if (hipMemset(zero, 0, sizeof(double) * blk) != hipSuccess || hipMemset(g, 0, sizeof(double) * 2 * m.fs) != hipSuccess) {
            rc = fail(NKC_ERR_HIP, "hipMemset");
            break;
        }
        if ((rc = copy_fields_h2d(kap, m.fs, 0, kH, kE, m.npts, ld))) break;
        if (hipStreamSynchronize(h->stream) != hipSuccess) { rc = fail(NKC_ERR_HIP, "sync"); break; }
        AdjArgs a{};
        a.lam_in = zero;
        a.kap_in = kap;
        a.lam_out = out;
        a.kap_out = kout;
        a.us = m.u[h->cur];
        a.rs = zero;
        a.geps = g;
        a.gmu = g + m.fs;
        a.rka = 0.0;
        a.rkb = 0.0;
        a.dt = 1.0;
        a.C0 = h->C0;
        a.yz = h->yz;
        if ((rc = launch_adjoint(m, a, h->stream))) break;
        if (hipStreamSynchronize(h->stream) != hipSuccess) { rc = fail(NKC_ERR_HIP, "sync"); break; }
        rc = copy_fields_d2h(out, m.fs, 0, aH, aE, m.npts, ld);
    } while (0);
    (void)hipFree(zero); (void)hipFree(kap); (void)hipFree(out); (void)hipFree(kout); (void)hipFree(g);
    return rc;
}

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------
int32_t nkc_batch_create(int32_t ncav, const nkc_box_desc *d, nkc_batch **out) {
    g_err.clear();
    if (ncav <= 0 || !d || !out) return fail(NKC_ERR_ARG, "bad batch arguments");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(NKC_ERR_NODEVICE, "no HIP device visible");
    // build every cavity, then concatenate per polynomial order
    std::map<int, std::vector<int>> by_n1;
    std::vector<HostMesh> hm(ncav);
    for (int c = 0; c < ncav; ++c) {
        if (!supported_n1(d[c].N + 1)) return fail(NKC_ERR_UNSUPPORTED, "cavity %d: N=%d outside 3..9", c, d[c].N);
        const int32_t rc = build_box(&d[c], hm[c]);
        if (rc) return fail(rc, "cavity %d: %s", c, hm[c].err.c_str());
        by_n1[hm[c].n1].push_back(c);
    }
    nkc_batch *b = new nkc_batch();
    b->cav.resize(ncav);
    int32_t rc = NKC_OK;
    for (auto &kv : by_n1) {
        HostMesh cat;
        cat.n1 = kv.first;
        cat.mat = 0;
        const HostMesh &first = hm[kv.second.front()];
        cat.D = first.D;
        cat.w = first.w;
        int64_t off = 0;
        for (int c : kv.second) off += hm[c].nelt;
        if (off * cat.n1 * cat.n1 * cat.n1 > ((int64_t)1 << 31) - 1) {
            rc = fail(NKC_ERR_ARG, "order N=%d: cavities exceed int32 indexing", cat.n1 - 1);
            break;
        }
        off = 0;
        for (int c : kv.second) {
            const HostMesh &m = hm[c];
            b->cav[c].bin = (int)b->bins.size();
            b->cav[c].off = off;
            b->cav[c].nelt = m.nelt;
            cat.geo.insert(cat.geo.end(), m.geo.begin(), m.geo.end());
            for (int32_t nb : m.nbr) cat.nbr.push_back(nb >= 0 ? (int32_t)(nb + off) : nb);
            cat.fco.insert(cat.fco.end(), m.fco.begin(), m.fco.end());
            cat.ieps.insert(cat.ieps.end(), m.ieps.begin(), m.ieps.end());
            cat.imu.insert(cat.imu.end(), m.imu.begin(), m.imu.end());
            cat.jac.insert(cat.jac.end(), m.jac.begin(), m.jac.end());
            off += m.nelt;
        }
        cat.nelt = off;
        b->bins.emplace_back();
        if ((rc = dev_mesh_upload(cat, b->bins.back()))) break;
        if (hipMalloc((void **)&b->bins.back().dt_e, sizeof(double) * cat.nelt) != hipSuccess) {
            rc = fail(NKC_ERR_HIP, "hipMalloc dt table");
            break;
        }
        hipStream_t s;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) { rc = fail(NKC_ERR_HIP, "stream"); break; }
        b->streams.push_back(s);
    }
    if (!rc) {
        b->ev.resize(b->bins.size() + 1);
        for (auto &e : b->ev)
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) { rc = fail(NKC_ERR_HIP, "event"); break; }
        if (!rc && hipStreamCreateWithFlags(&b->main, hipStreamNonBlocking) != hipSuccess) rc = fail(NKC_ERR_HIP, "stream");
    }
    if (rc) {
        free_batch(b);
        delete b;
        return rc;
    }
    *out = b;
    return NKC_OK;
}

int32_t nkc_batch_destroy(nkc_batch *b) {
    if (!b) return fail(NKC_ERR_ARG, "null batch");
    (void)hipStreamSynchronize(b->main);
    for (auto s : b->streams) (void)hipStreamSynchronize(s);
    free_batch(b);
    delete b;
    return NKC_OK;
}

int32_t nkc_batch_set_state(nkc_batch *b, int32_t cav, const double *hn, const double *en, int64_t ld) {
    if (!b || !hn || !en || cav < 0 || cav >= (int32_t)b->cav.size()) return fail(NKC_ERR_ARG, "bad batch state arguments");
    const auto &c = b->cav[cav];
    DevMesh &dm = b->bins[c.bin];
    const int64_t n3 = (int64_t)dm.n1 * dm.n1 * dm.n1, n = c.nelt * n3;
    if (ld < n) return fail(NKC_ERR_ARG, "ld < npts of cavity");
    HIPCHK(hipStreamSynchronize(b->main));
    int32_t rc = copy_fields_h2d(dm.u[b->cur], dm.fs, c.off * n3, hn, en, n, ld);
    if (rc) return rc;
    for (int f = 0; f < 6; ++f) HIPCHK(hipMemset(dm.k + f * dm.fs + c.off * n3, 0, sizeof(double) * n));
    return NKC_OK;
}

int32_t nkc_batch_get_state(nkc_batch *b, int32_t cav, double *hn, double *en, int64_t ld) {
    if (!b || !hn || !en || cav < 0 || cav >= (int32_t)b->cav.size()) return fail(NKC_ERR_ARG, "bad batch state arguments");
    const auto &c = b->cav[cav];
    const DevMesh &dm = b->bins[c.bin];
    const int64_t n3 = (int64_t)dm.n1 * dm.n1 * dm.n1, n = c.nelt * n3;
    if (ld < n) return fail(NKC_ERR_ARG, "ld < npts of cavity");
    HIPCHK(hipStreamSynchronize(b->main));
    return copy_fields_d2h(dm.u[b->cur], dm.fs, c.off * n3, hn, en, n, ld);
}

int32_t nkc_batch_step(nkc_batch *b, const double *dt, double *times, int64_t nsteps) {
    if (!b || !dt || nsteps < 0) return fail(NKC_ERR_ARG, "bad batch step arguments");
    const int ncav = (int)b->cav.size();
    bool changed = false;
    for (int c = 0; c < ncav; ++c) {
        if (!std::isfinite(dt[c])) return fail(NKC_ERR_ARG, "dt[%d] not finite", c);
        changed |= !(b->cav[c].dt == dt[c]);
    }
    if (changed) {
        HIPCHK(hipStreamSynchronize(b->main));
        for (size_t i = 0; i < b->bins.size(); ++i) {
            std::vector<double> dte(b->bins[i].nelt);
            for (int c = 0; c < ncav; ++c)
                if (b->cav[c].bin == (int)i)
                    for (int64_t e = 0; e < b->cav[c].nelt; ++e) dte[b->cav[c].off + e] = dt[c];
            HIPCHK(hipMemcpy(b->bins[i].dt_e, dte.data(), sizeof(double) * dte.size(), hipMemcpyHostToDevice));
        }
        for (int c = 0; c < ncav; ++c) b->cav[c].dt = dt[c];
    }
    for (auto &dm : b->bins)
        if (dm.layout < 0) {
            double us[NLAYOUT];
            const int32_t rc = tune_stage(dm, nullptr, 0, 0, 0, 0, dm.u[b->cur], b->main, us);
            if (rc) return rc;
        }
    if (!b->graph_ok) {
        const int32_t rc = batch_capture(b);
        if (rc) return rc;
    }
    for (int64_t s = 0; s < nsteps; ++s) {
        HIPCHK(hipGraphLaunch(b->gexec[b->cur], b->main));
        b->cur ^= 1;
    }
    for (int c = 0; c < ncav; ++c) {
        for (int64_t s = 0; s < nsteps; ++s) b->cav[c].time = b->cav[c].time + dt[c];
        if (times) times[c] = b->cav[c].time;
    }
    return NKC_OK;
}

int32_t nkc_batch_sync(nkc_batch *b) {
    if (!b) return fail(NKC_ERR_ARG, "null batch");
    HIPCHK(hipStreamSynchronize(b->main));
    return NKC_OK;
}

int32_t nkc_batch_stream(nkc_batch *b, void **stream) {
    if (!b || !stream) return fail(NKC_ERR_ARG, "null batch or output");
    *stream = (void *)b->main;
    return NKC_OK;
}

int32_t nkc_batch_info(nkc_batch *b, int64_t out[8]) {
    if (!b || !out) return fail(NKC_ERR_ARG, "null batch or output");
    int64_t npts = 0, nelt = 0;
    for (auto &dm : b->bins) { npts += dm.npts; nelt += dm.nelt; }
    out[0] = npts;
    out[1] = (int64_t)b->cav.size();
    out[2] = (int64_t)b->bins.size();
    out[3] = nelt;
    out[4] = out[5] = out[6] = out[7] = 0;
    return NKC_OK;
}

int32_t nkc_batch_layout(nkc_batch *b, int32_t *out, int64_t cap) {
    if (!b || !out) return fail(NKC_ERR_ARG, "null batch or output");
    const int64_t nb = (int64_t)b->bins.size();
    if (cap < 2 * nb) return fail(NKC_ERR_ARG, "output holds %lld ints, %lld needed", (long long)cap, (long long)(2 * nb));
    for (int64_t i = 0; i < nb; ++i) {
        out[2 * i] = b->bins[i].n1 - 1;
        out[2 * i + 1] = b->bins[i].layout;
    }
    return NKC_OK;
}

}  // extern "C"
