// nekcem_hip.hip -- C-ABI of libnekcem_hip.so (declared in include/nekcem_hip.h).
//
// Host side of the MI355X Maxwell step: validates and compacts the
// reference-layout working set handed over by nkc_create into per-element
// tables, owns the device buffers and the stream, launches the fused stage
// kernel (stage_kernel.hpp) five times per RK step and replays the step from
// a captured hipGraph.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/nekcem_hip.h"
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

}  // namespace

struct nkc_handle {
    int N1 = 0;
    int64_t nelt = 0, npts = 0, fs = 0;
    int mat = MAT_ELEM;
    int device = 0;
    hipStream_t stream = nullptr;
    double *u[2] = {nullptr, nullptr};
    double *k = nullptr;
    double *res = nullptr;
    double *D = nullptr, *geo = nullptr, *fco = nullptr, *ieps = nullptr, *imu = nullptr;
    int32_t *nbr = nullptr;
    unsigned long long *red = nullptr;   // check_finite result
    int cur = 0;                          // which u buffer holds the state
    bool use_graph = true;
    double graph_dt = NAN;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    // device time: tdev[0] = time at the start of the next step, tdev[1] =
    // stage time for nkc_residual
    double *tdev = nullptr;
    double t_host = 0.0;
    // per-element geometry kept on the host (source weights)
    std::vector<double> jac_e;
    std::vector<double> wgll;
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

// Ed's face (1..6) of kernel face kf (x-, x+, y-, y+, z-, z+).
const int ED_OF_KF[6] = {4, 2, 1, 3, 5, 6};
const int KF_OF_ED[7] = {-1, 2, 1, 3, 0, 4, 5};

// Volume node (in-element index) of in-face node q = a + n1*b on kernel face kf
// (skpdat order, src/nek5_connect11.F:1483-1528).
inline int64_t face_node(int n1, int kf, int q) {
    int a = q % n1, b = q / n1, m = n1 - 1, n2 = n1 * n1;
    switch (kf) {
    case 0: return 0 + (int64_t)n1 * a + (int64_t)n2 * b;
    case 1: return m + (int64_t)n1 * a + (int64_t)n2 * b;
    case 2: return a + 0 + (int64_t)n2 * b;
    case 3: return a + (int64_t)n1 * m + (int64_t)n2 * b;
    case 4: return a + (int64_t)n1 * b + 0;
    default: return a + (int64_t)n1 * b + (int64_t)n2 * m;
    }
}

bool close(double x, double y, double scale, double tol) { return std::fabs(x - y) <= tol * scale; }

template <int N1, int MAT, int MODE>
void launch_t(const StageArgs &a, int grid, hipStream_t s) {
    hipLaunchKernelGGL((stage_kernel<N1, MAT, MODE>), dim3(grid), dim3(Shape<N1>::BS), 0, s, a);
}

int block_threads(int n1) { return ((n1 * n1 + 63) / 64) * 64 * 4; }

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

bool supported_n1(int n1) { return n1 >= 4 && n1 <= 10; }

int32_t launch_stage(nkc_handle *h, const StageArgs &a, int mode) {
    int per_xcd = (int)((h->nelt + 7) / 8);
    int grid = 8 * per_xcd;
    StageArgs b = a;
    b.per_xcd = per_xcd;
    switch (h->N1) {
    case 4: launch_n<4>(b, h->mat, mode, grid, h->stream); break;
    case 5: launch_n<5>(b, h->mat, mode, grid, h->stream); break;
    case 6: launch_n<6>(b, h->mat, mode, grid, h->stream); break;
    case 7: launch_n<7>(b, h->mat, mode, grid, h->stream); break;
    case 8: launch_n<8>(b, h->mat, mode, grid, h->stream); break;
    case 9: launch_n<9>(b, h->mat, mode, grid, h->stream); break;
    case 10: launch_n<10>(b, h->mat, mode, grid, h->stream); break;
    default: return fail(NKC_ERR_UNSUPPORTED, "polynomial order N=%d not built", h->N1 - 1);
    }
    HIPCHK(hipGetLastError());
    return NKC_OK;
}

StageArgs base_args(nkc_handle *h) {
    StageArgs a{};
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
    a.D = h->D;
    a.geo = h->geo;
    a.nbr = h->nbr;
    a.fco = h->fco;
    a.ieps = h->ieps;
    a.imu = h->imu;
    a.fs = h->fs;
    a.nelt = (int32_t)h->nelt;
    return a;
}

// time = time + dt at the end of a step (src/cem_drive.F:639), on the device
// so a replayed graph sees the right stage times.
__global__ void advance_time(double *t, double dt) { t[0] = t[0] + dt; }

// Queue one RK step starting from buffer `from` (ends in 1-from).
int32_t queue_step(nkc_handle *h, int from, double dt) {
    for (int s = 0; s < 5; ++s) {
        StageArgs a = base_args(h);
        int src = (from + s) & 1;
        a.uin = h->u[src];
        a.uout = h->u[src ^ 1];
        a.kbuf = h->k;
        a.rka = RK4A[s];
        a.rkb = RK4B[s];
        a.rkc = RK4C[s];
        a.dt = dt;
        int32_t rc = launch_stage(h, a, MODE_RK);
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
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        rc = queue_step(h, from, dt);
        hipGraph_t gg = nullptr;
        hipError_t ee = hipStreamEndCapture(h->stream, &gg);
        if (rc) return rc;
        if (ee != hipSuccess) return fail(NKC_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ee));
        g = gg;
        HIPCHK(hipGraphInstantiate(&h->gexec[from], g, nullptr, nullptr, 0));
        HIPCHK(hipGraphDestroy(g));
    }
    h->graph_dt = dt;
    return NKC_OK;
}

__global__ void maxabs_kernel(const double *__restrict__ u, int64_t fs, int64_t npts,
                              unsigned long long *out) {
    unsigned long long m = 0;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int f = 0; f < 6; ++f)
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npts; i += stride) {
            double v = fabs(u[f * fs + i]);
            unsigned long long b = __double_as_longlong(v);   // monotone for v >= 0, NaN on top
            m = b > m ? b : m;
        }
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_down(m, off);
        m = o > m ? o : m;
    }
    __shared__ unsigned long long sm[16];
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sm[w] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long r = 0;
        for (int q = 0; q < (int)(blockDim.x >> 6); ++q) r = sm[q] > r ? sm[q] : r;
        atomicMax(out, r);
    }
}

__global__ void zero_u64(unsigned long long *p) { *p = 0ull; }

void free_all(nkc_handle *h) {
    drop_graphs(h);
    (void)hipFree(h->u[0]); (void)hipFree(h->u[1]); (void)hipFree(h->k); (void)hipFree(h->res);
    (void)hipFree(h->D); (void)hipFree(h->geo); (void)hipFree(h->fco); (void)hipFree(h->ieps); (void)hipFree(h->imu);
    (void)hipFree(h->nbr); (void)hipFree(h->red); (void)hipFree(h->tdev);
    (void)hipFree(h->src_off); (void)hipFree(h->src_amp);
    (void)hipFree(h->drude_off); (void)hipFree(h->dJ); (void)hipFree(h->dkJ);
    (void)hipFree(h->dpar); (void)hipFree(h->dresJ);
    if (h->stream) (void)hipStreamDestroy(h->stream);
}

template <class T>
int32_t upload(T **dst, const std::vector<T> &src) {
    HIPCHK(hipMalloc((void **)dst, sizeof(T) * (src.empty() ? 1 : src.size())));
    if (!src.empty()) HIPCHK(hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
    return NKC_OK;
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
    const int n2 = n1 * n1, n3 = n2 * n1;
    const int64_t nelt = d->nelt, npts = (int64_t)n3 * nelt, nxzfl = (int64_t)6 * n2 * nelt;
    if (npts > ((int64_t)1 << 31) - 1) return fail(NKC_ERR_ARG, "npts exceeds int32 indexing");
    const double C0 = d->ifupwind ? 1.0 : 0.0;
    const double tol = 1e-8;

    std::vector<int8_t> pecslot(nxzfl, 0);
    for (int64_t l = 0; l < d->ncempec; ++l) {
        int64_t s = (int64_t)d->cempec[l] - 1;
        if (s < 0 || s >= nxzfl) return fail(NKC_ERR_ARG, "cempec[%lld]=%d out of range", (long long)l, d->cempec[l]);
        pecslot[s] = 1;
    }

    std::vector<double> geo(nelt * 9);
    std::vector<int32_t> nbr(nelt * 6);
    // material mode: per element iff eps, mu and Y/Z coefficients are constant
    bool elem_mat = true;
    for (int64_t e = 0; e < nelt && elem_mat; ++e) {
        int64_t b = e * n3;
        for (int q = 1; q < n3; ++q)
            if (d->permittivity[b + q] != d->permittivity[b] || d->permeability[b + q] != d->permeability[b]) {
                elem_mat = false;
                break;
            }
        for (int f = 0; f < 6 && elem_mat; ++f) {
            int64_t s0 = e * 6 * n2 + (int64_t)f * n2;
            for (int q = 1; q < n2; ++q)
                if (d->Y_0[s0 + q] != d->Y_0[s0] || d->Y_1[s0 + q] != d->Y_1[s0] ||
                    d->Z_0[s0 + q] != d->Z_0[s0] || d->Z_1[s0 + q] != d->Z_1[s0]) {
                    elem_mat = false;
                    break;
                }
        }
    }
    std::vector<double> fco(elem_mat ? nelt * 6 * 4 : nelt * 6 * n2 * 4);
    std::vector<double> ieps(elem_mat ? nelt : npts), imu(elem_mat ? nelt : npts);

    for (int64_t e = 0; e < nelt; ++e) {
        const int64_t b = e * n3;
        const double J0 = d->jacm[b], rx0 = d->rxmn[b], sy0 = d->symn[b], tz0 = d->tzmn[b];
        const double sc = std::fmax(std::fabs(rx0), std::fmax(std::fabs(sy0), std::fabs(tz0)));
        if (!(J0 > 0.0)) return fail(NKC_ERR_ARG, "element %lld: Jacobian %g not positive", (long long)e, J0);
        for (int q = 0; q < n3; ++q) {
            int64_t p = b + q;
            if (!close(d->jacm[p], J0, std::fabs(J0), tol) || !close(d->rxmn[p], rx0, sc, tol) ||
                !close(d->symn[p], sy0, sc, tol) || !close(d->tzmn[p], tz0, sc, tol) ||
                std::fabs(d->rymn[p]) > tol * sc || std::fabs(d->rzmn[p]) > tol * sc ||
                std::fabs(d->sxmn[p]) > tol * sc || std::fabs(d->szmn[p]) > tol * sc ||
                std::fabs(d->txmn[p]) > tol * sc || std::fabs(d->tymn[p]) > tol * sc)
                return fail(NKC_ERR_UNSUPPORTED,
                            "element %lld is not an axis-aligned affine hexahedron; only box meshes are built",
                            (long long)e);
        }
        geo[e * 9 + 0] = rx0 / J0;
        geo[e * 9 + 1] = sy0 / J0;
        geo[e * 9 + 2] = tz0 / J0;
        for (int kf = 0; kf < 6; ++kf) {
            const int ef = ED_OF_KF[kf];
            const int ax = kf / 2;
            const double sg = (kf & 1) ? 1.0 : -1.0;
            const int64_t s0 = e * 6 * n2 + (int64_t)(ef - 1) * n2;
            double L0 = 0.0;
            for (int q = 0; q < n2; ++q) {
                const int64_t s = s0 + q;
                const int64_t v = (int64_t)d->cemface[s] - 1;
                if (v != b + face_node(n1, kf, q))
                    return fail(NKC_ERR_ARG, "cemface[%lld]=%d does not follow the skpdat face order",
                                (long long)s, d->cemface[s]);
                const double nv[3] = {d->unxm[s], d->unym[s], d->unzm[s]};
                for (int c = 0; c < 3; ++c)
                    if (std::fabs(nv[c] - (c == ax ? sg : 0.0)) > tol)
                        return fail(NKC_ERR_UNSUPPORTED, "element %lld face %d: normal is not an axis direction",
                                    (long long)e, ef);
                const double L = d->aream[s] / d->bmn[v];
                if (q == 0) L0 = L;
                else if (!close(L, L0, std::fabs(L0), tol))
                    return fail(NKC_ERR_UNSUPPORTED, "element %lld face %d: lift factor varies over the face",
                                (long long)e, ef);
            }
            geo[e * 9 + 3 + kf] = L0;
            // neighbour from the gs partner map
            const int64_t p0 = (int64_t)d->face_partner[s0] - 1;
            int32_t nb;
            if (p0 < 0) {
                nb = pecslot[s0] ? FACE_PEC : FACE_NONE;
                for (int q = 0; q < n2; ++q)
                    if (d->face_partner[s0 + q] != 0 || pecslot[s0 + q] != pecslot[s0])
                        return fail(NKC_ERR_UNSUPPORTED, "element %lld face %d: mixed boundary face", (long long)e, ef);
            } else {
                if (p0 >= nxzfl) return fail(NKC_ERR_ARG, "face_partner out of range");
                const int64_t ne = p0 / (6 * n2);
                const int nef = (int)((p0 / n2) % 6) + 1;
                const int nkf = KF_OF_ED[nef];
                if (nkf != (kf ^ 1))
                    return fail(NKC_ERR_UNSUPPORTED, "element %lld face %d pairs with face %d of element %lld; "
                                "only opposite-face box pairing is built", (long long)e, ef, nef, (long long)ne);
                const int64_t ns0 = ne * 6 * n2 + (int64_t)(nef - 1) * n2;
                for (int q = 0; q < n2; ++q)
                    if ((int64_t)d->face_partner[s0 + q] - 1 != ns0 + q || pecslot[s0 + q])
                        return fail(NKC_ERR_UNSUPPORTED, "element %lld face %d: rotated face pairing not built",
                                    (long long)e, ef);
                nb = (int32_t)ne;
            }
            nbr[e * 6 + kf] = nb;
            // flux coefficients (cem_maxwell_flux3d loop 2, src/cem_maxwell.F:970-975)
            const int nq = elem_mat ? 1 : n2;
            for (int q = 0; q < nq; ++q) {
                const int64_t s = s0 + q;
                const double Y0 = d->Y_0[s], Y1 = d->Y_1[s], Z0 = d->Z_0[s], Z1 = d->Z_1[s];
                double *c = elem_mat ? &fco[(e * 6 + kf) * 4] : &fco[((e * 6 + kf) * n2 + q) * 4];
                c[0] = -0.5 / Y0 * Y1;
                c[1] = 0.5 / Y0 * C0;
                c[2] = 0.5 / Z0 * Z1;
                c[3] = 0.5 / Z0 * C0;
            }
        }
        if (elem_mat) {
            ieps[e] = 1.0 / d->permittivity[b];
            imu[e] = 1.0 / d->permeability[b];
        }
    }
    if (!elem_mat)
        for (int64_t p = 0; p < npts; ++p) {
            ieps[p] = 1.0 / d->permittivity[p];
            imu[p] = 1.0 / d->permeability[p];
        }

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(NKC_ERR_NODEVICE, "no HIP device visible");

    nkc_handle *h = new nkc_handle();
    h->jac_e.resize(nelt);
    for (int64_t e = 0; e < nelt; ++e) h->jac_e[e] = d->jacm[e * n3];
    h->wgll.assign(d->wxm1, d->wxm1 + n1);
    h->N1 = n1;
    h->nelt = nelt;
    h->npts = npts;
    h->fs = (npts + 31) / 32 * 32;
    h->mat = elem_mat ? MAT_ELEM : MAT_NODE;
    int32_t rc = NKC_OK;
    auto chk = [&](int32_t r) { if (r && !rc) rc = r; return r; };
    do {
        if (hipGetDevice(&h->device) != hipSuccess) { rc = fail(NKC_ERR_HIP, "hipGetDevice"); break; }
        if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
            rc = fail(NKC_ERR_HIP, "hipStreamCreate"); break;
        }
        const size_t ub = sizeof(double) * 6 * (size_t)h->fs;
        if (hipMalloc((void **)&h->u[0], ub) != hipSuccess || hipMalloc((void **)&h->u[1], ub) != hipSuccess ||
            hipMalloc((void **)&h->k, ub) != hipSuccess || hipMalloc((void **)&h->red, 8) != hipSuccess) {
            rc = fail(NKC_ERR_HIP, "hipMalloc of %zu-byte state buffers failed", ub); break;
        }
        if (hipMemset(h->u[0], 0, ub) != hipSuccess || hipMemset(h->u[1], 0, ub) != hipSuccess ||
            hipMemset(h->k, 0, ub) != hipSuccess) { rc = fail(NKC_ERR_HIP, "hipMemset"); break; }
        if (hipMalloc((void **)&h->tdev, 2 * sizeof(double)) != hipSuccess ||
            hipMemset(h->tdev, 0, 2 * sizeof(double)) != hipSuccess) { rc = fail(NKC_ERR_HIP, "hipMalloc time"); break; }
        std::vector<double> Dv(d->dxm1, d->dxm1 + n2);
        if (chk(upload(&h->D, Dv))) break;
        if (chk(upload(&h->geo, geo))) break;
        if (chk(upload(&h->nbr, nbr))) break;
        if (chk(upload(&h->fco, fco))) break;
        if (chk(upload(&h->ieps, ieps))) break;
        if (chk(upload(&h->imu, imu))) break;
    } while (0);
    if (rc) {
        free_all(h);
        delete h;
        return rc;
    }
    *out = h;
    return NKC_OK;
}

int32_t nkc_destroy(nkc_handle *h) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    hipStreamSynchronize(h->stream);
    free_all(h);
    delete h;
    return NKC_OK;
}

int32_t nkc_set_state(nkc_handle *h, const double *hn, const double *en, const double *khn,
                      const double *ken, int64_t ld) {
    if (!h || !hn || !en) return fail(NKC_ERR_ARG, "null handle or state");
    if (ld < h->npts) return fail(NKC_ERR_ARG, "ld=%lld < npts=%lld", (long long)ld, (long long)h->npts);
    HIPCHK(hipStreamSynchronize(h->stream));
    const size_t nb = sizeof(double) * h->npts;
    for (int c = 0; c < 3; ++c) {
        HIPCHK(hipMemcpy(h->u[h->cur] + c * h->fs, hn + c * ld, nb, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(h->u[h->cur] + (3 + c) * h->fs, en + c * ld, nb, hipMemcpyHostToDevice));
        if (khn) HIPCHK(hipMemcpy(h->k + c * h->fs, khn + c * ld, nb, hipMemcpyHostToDevice));
        else HIPCHK(hipMemset(h->k + c * h->fs, 0, nb));
        if (ken) HIPCHK(hipMemcpy(h->k + (3 + c) * h->fs, ken + c * ld, nb, hipMemcpyHostToDevice));
        else HIPCHK(hipMemset(h->k + (3 + c) * h->fs, 0, nb));
    }
    return NKC_OK;
}

int32_t nkc_get_state(nkc_handle *h, double *hn, double *en, double *khn, double *ken, int64_t ld) {
    if (!h || !hn || !en) return fail(NKC_ERR_ARG, "null handle or state");
    if (ld < h->npts) return fail(NKC_ERR_ARG, "ld=%lld < npts=%lld", (long long)ld, (long long)h->npts);
    HIPCHK(hipStreamSynchronize(h->stream));
    const size_t nb = sizeof(double) * h->npts;
    for (int c = 0; c < 3; ++c) {
        HIPCHK(hipMemcpy(hn + c * ld, h->u[h->cur] + c * h->fs, nb, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(en + c * ld, h->u[h->cur] + (3 + c) * h->fs, nb, hipMemcpyDeviceToHost));
        if (khn) HIPCHK(hipMemcpy(khn + c * ld, h->k + c * h->fs, nb, hipMemcpyDeviceToHost));
        if (ken) HIPCHK(hipMemcpy(ken + c * ld, h->k + (3 + c) * h->fs, nb, hipMemcpyDeviceToHost));
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
            int32_t rc = capture(h, dt);
            if (rc) return rc;
        }
        for (int64_t s = 0; s < nsteps; ++s) {
            HIPCHK(hipGraphLaunch(h->gexec[h->cur], h->stream));
            h->cur ^= 1;
        }
    } else {
        for (int64_t s = 0; s < nsteps; ++s) {
            int32_t rc = queue_step(h, h->cur, dt);
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
    if (ld < h->npts) return fail(NKC_ERR_ARG, "ld < npts");
    if (!h->res) HIPCHK(hipMalloc((void **)&h->res, sizeof(double) * 6 * h->fs));
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(hipMemcpy(h->tdev + 1, &rktime, sizeof(double), hipMemcpyHostToDevice));
    StageArgs a = base_args(h);
    a.tdev = h->tdev + 1;
    a.rkc = 0.0;
    a.uin = h->u[h->cur];
    a.res = h->res;
    int32_t rc = launch_stage(h, a, MODE_RESIDUAL);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(h->stream));
    const size_t nb = sizeof(double) * h->npts;
    for (int c = 0; c < 3; ++c) {
        HIPCHK(hipMemcpy(reshn + c * ld, h->res + c * h->fs, nb, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(resen + c * ld, h->res + (3 + c) * h->fs, nb, hipMemcpyDeviceToHost));
    }
    return NKC_OK;
}

int32_t nkc_set_source(nkc_handle *h, const double *amp, const double dir[3], double t0, double tw) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    (void)hipFree(h->src_off); h->src_off = nullptr;
    (void)hipFree(h->src_amp); h->src_amp = nullptr;
    int32_t rc = drop_graphs(h);
    if (rc || !amp) return rc;
    if (!dir || !(tw != 0.0) || !std::isfinite(t0) || !std::isfinite(tw)) return fail(NKC_ERR_ARG, "bad source direction/time");
    const int n1 = h->N1, n2 = n1 * n1, n3 = n2 * n1;
    std::vector<int32_t> off(h->nelt, -1);
    std::vector<double> vals;
    for (int64_t e = 0; e < h->nelt; ++e) {
        bool any = false;
        for (int q = 0; q < n3; ++q) any |= amp[e * n3 + q] != 0.0;
        if (!any) continue;
        off[e] = (int32_t)(vals.size() / n3);
        for (int k = 0; k < n1; ++k)
            for (int j = 0; j < n1; ++j)
                for (int i = 0; i < n1; ++i) {
                    // usersrc adds amp (bm1-weighted) before ebm1; the kernel
                    // works per unit J w3
                    const double bm = h->jac_e[e] * (h->wgll[i] * h->wgll[j] * h->wgll[k]);
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
    const int n3 = h->N1 * h->N1 * h->N1;
    std::vector<int32_t> off(h->nelt, -1);
    int32_t nde = 0;
    for (int64_t l = 0; l < ndrude; ++l) {
        const int64_t v = (int64_t)dindex[l] - 1;
        if (v < 0 || v >= h->npts) return fail(NKC_ERR_ARG, "dindex[%lld]=%d out of range", (long long)l, dindex[l]);
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
    const int n3 = h->N1 * h->N1 * h->N1;
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
    hipLaunchKernelGGL(maxabs_kernel, dim3(1024), dim3(256), 0, h->stream, h->u[h->cur], h->fs, h->npts, h->red);
    HIPCHK(hipGetLastError());
    unsigned long long bits = 0;
    HIPCHK(hipMemcpyAsync(&bits, h->red, 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    double m;
    std::memcpy(&m, &bits, 8);
    if (maxabs) *maxabs = m;
    if (!std::isfinite(m) || m > 1e20) return fail(NKC_ERR_BLOWUP, "state blew up: max|u| = %g", m);
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
    const int n1 = h->N1;
    out[0] = h->npts;
    out[1] = h->mat;
    out[2] = 1;
    out[3] = block_threads(n1);
    out[4] = (int64_t)sizeof(double) * (6 * n1 * n1 * n1 + 25 * n1 * n1);
    out[5] = h->use_graph ? 1 : 0;
    out[6] = h->fs;
    out[7] = h->nelt;
    return NKC_OK;
}

int32_t nkc_set_graph(nkc_handle *h, int32_t enable) {
    if (!h) return fail(NKC_ERR_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(h->stream));
    h->use_graph = enable != 0;
    if (!h->use_graph) return drop_graphs(h);
    return NKC_OK;
}

}  // extern "C"
