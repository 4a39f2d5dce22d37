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
#include <map>
#include <string>
#include <vector>

#include "../../include/nekcem_hip.h"
#include "mesh_tables.hpp"
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
int block_threads(int n1) { return ((n1 * n1 + 63) / 64) * 64 * 4; }

template <class T>
int32_t upload(T **dst, const std::vector<T> &src) {
    HIPCHK(hipMalloc((void **)dst, sizeof(T) * (src.empty() ? 1 : src.size())));
    if (!src.empty()) HIPCHK(hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
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
    std::vector<double> jac, w;   // host copies (source weights)
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
    return NKC_OK;
}

void dev_mesh_free(DevMesh &dm) {
    double *p[] = {dm.u[0], dm.u[1], dm.k, dm.D, dm.geo, dm.fco, dm.ieps, dm.imu, dm.dt_e};
    for (double *q : p) (void)hipFree(q);
    (void)hipFree(dm.nbr);
    dm = DevMesh();
}

template <int N1, int MAT, int MODE>
void launch_t(const StageArgs &a, int grid, hipStream_t s) {
    hipLaunchKernelGGL((stage_kernel<N1, MAT, MODE>), dim3(grid), dim3(Shape<N1>::BS), 0, s, a);
}

template <int N1>
void launch_n(const StageArgs &a, int mat, int mode, int grid, hipStream_t s) {
    if (mat == MAT_ELEM) {
        if (mode == MODE_RK) launch_t<N1, MAT_ELEM, MODE_RK>(a, grid, s);
        else launch_t<N1, MAT_ELEM, MODE_RESIDUAL>(a, grid, s);
    } else {
        if (mode == MODE_RK) launch_t<N1, MAT_NODE, MODE_RK>(a, grid, s);
        else launch_t<N1, MAT_NODE, MODE_RESIDUAL>(a, grid, s);
    }
}

int32_t launch_stage(const DevMesh &dm, const StageArgs &a, int mode, hipStream_t s) {
    const int per_xcd = (int)((dm.nelt + 7) / 8);
    const int grid = 8 * per_xcd;
    StageArgs b = a;
    b.per_xcd = per_xcd;
    switch (dm.n1) {
    case 4: launch_n<4>(b, dm.mat, mode, grid, s); break;
    case 5: launch_n<5>(b, dm.mat, mode, grid, s); break;
    case 6: launch_n<6>(b, dm.mat, mode, grid, s); break;
    case 7: launch_n<7>(b, dm.mat, mode, grid, s); break;
    case 8: launch_n<8>(b, dm.mat, mode, grid, s); break;
    case 9: launch_n<9>(b, dm.mat, mode, grid, s); break;
    case 10: launch_n<10>(b, dm.mat, mode, grid, s); break;
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
    return a;
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
    // Drude
    int64_t ndrude = 0, ndelem = 0;
    int32_t *drude_off = nullptr;
    double *dJ = nullptr, *dkJ = nullptr, *dpar = nullptr, *dresJ = nullptr;
    std::vector<int64_t> drude_slot;      // listed node -> compact (elem*3*n3 + vloc)
};

namespace {

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
    return a;
}

// Queue one RK step starting from buffer `from` (ends in 1-from).
int32_t queue_step(nkc_handle *h, int from, double dt) {
    for (int s = 0; s < 5; ++s) {
        StageArgs a = base_args(h);
        const int src = (from + s) & 1;
        a.uin = h->m.u[src];
        a.uout = h->m.u[src ^ 1];
        a.kbuf = h->m.k;
        a.rka = RK4A[s];
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
        if (rc) return rc;
        if (ee != hipSuccess) return fail(NKC_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ee));
        HIPCHK(hipGraphInstantiate(&h->gexec[from], g, nullptr, nullptr, 0));
        HIPCHK(hipGraphDestroy(g));
    }
    h->graph_dt = dt;
    return NKC_OK;
}

void free_handle(nkc_handle *h) {
    drop_graphs(h);
    dev_mesh_free(h->m);
    (void)hipFree(h->res); (void)hipFree(h->red); (void)hipFree(h->tdev);
    (void)hipFree(h->src_off); (void)hipFree(h->src_amp);
    (void)hipFree(h->drude_off); (void)hipFree(h->dJ); (void)hipFree(h->dkJ);
    (void)hipFree(h->dpar); (void)hipFree(h->dresJ);
    if (h->stream) (void)hipStreamDestroy(h->stream);
}

int32_t make_handle(const HostMesh &hm, nkc_handle **out) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(NKC_ERR_NODEVICE, "no HIP device visible");
    nkc_handle *h = new nkc_handle();
    int32_t rc = NKC_OK;
    do {
        if (hipGetDevice(&h->device) != hipSuccess) { rc = fail(NKC_ERR_HIP, "hipGetDevice"); break; }
        if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
            rc = fail(NKC_ERR_HIP, "hipStreamCreate"); break;
        }
        if ((rc = dev_mesh_upload(hm, h->m))) break;
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

// copy 6 fields between host (3 + 3 components at stride ld) and device
int32_t copy_fields_h2d(double *dst, int64_t fs, int64_t dst_off, const double *hn, const double *en,
                        int64_t n, int64_t ld) {
    const size_t nb = sizeof(double) * n;
    for (int c = 0; c < 3; ++c) {
        HIPCHK(hipMemcpy(dst + c * fs + dst_off, hn + c * ld, nb, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dst + (3 + c) * fs + dst_off, en + c * ld, nb, hipMemcpyHostToDevice));
    }
    return NKC_OK;
}

int32_t copy_fields_d2h(const double *src, int64_t fs, int64_t src_off, double *hn, double *en,
                        int64_t n, int64_t ld) {
    const size_t nb = sizeof(double) * n;
    for (int c = 0; c < 3; ++c) {
        HIPCHK(hipMemcpy(hn + c * ld, src + c * fs + src_off, nb, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(en + c * ld, src + (3 + c) * fs + src_off, nb, hipMemcpyDeviceToHost));
    }
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
        if (rc) return rc;
        if (ee != hipSuccess) return fail(NKC_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ee));
        HIPCHK(hipGraphInstantiate(&b->gexec[from], g, nullptr, nullptr, 0));
        HIPCHK(hipGraphDestroy(g));
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
    if (!supported_n1(n1)) return fail(NKC_ERR_UNSUPPORTED, "polynomial order N=%d outside 3..9", d->N);
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
    const int32_t rc = compact_reference(d, hm);
    if (rc) return fail(rc, "%s", hm.err.c_str());
    return make_handle(hm, out);
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
    const int32_t rc = build_box(d, hm);
    if (rc) return fail(rc, "%s", hm.err.c_str());
    return make_handle(hm, out);
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
        if (khn) HIPCHK(hipMemcpy(m.k + c * m.fs, khn + c * ld, nb, hipMemcpyHostToDevice));
        else HIPCHK(hipMemset(m.k + c * m.fs, 0, nb));
        if (ken) HIPCHK(hipMemcpy(m.k + (3 + c) * m.fs, ken + c * ld, nb, hipMemcpyHostToDevice));
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

int32_t nkc_set_source(nkc_handle *h, const double *amp, const double dir[3], double t0, double tw) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
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
    if ((rc = upload(&h->src_off, off))) return rc;
    if ((rc = upload(&h->src_amp, vals))) return rc;
    for (int c = 0; c < 3; ++c) h->src_dir[c] = dir[c];
    h->src_t0 = t0;
    h->src_tw = tw;
    return NKC_OK;
}

int32_t nkc_set_drude(nkc_handle *h, int64_t ndrude, const int32_t *dindex, const double *params,
                      const double *J, const double *kJ) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    (void)hipFree(h->drude_off); h->drude_off = nullptr;
    (void)hipFree(h->dJ); h->dJ = nullptr;
    (void)hipFree(h->dkJ); h->dkJ = nullptr;
    (void)hipFree(h->dpar); h->dpar = nullptr;
    (void)hipFree(h->dresJ); h->dresJ = nullptr;
    h->ndrude = 0;
    h->ndelem = 0;
    h->drude_slot.clear();
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
    if ((rc = upload(&h->drude_off, off))) return rc;
    if ((rc = upload(&h->dpar, par))) return rc;
    if ((rc = upload(&h->dJ, Jv))) return rc;
    if ((rc = upload(&h->dkJ, kv))) return rc;
    if ((rc = upload(&h->dresJ, kv))) return rc;
    h->ndrude = ndrude;
    h->ndelem = nde;
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
    out[3] = block_threads(n1);
    out[4] = (int64_t)sizeof(double) * (6 * n1 * n1 * n1 + 25 * n1 * n1);
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

}  // extern "C"
