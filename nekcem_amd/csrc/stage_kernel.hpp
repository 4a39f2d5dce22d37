// stage_kernel.hpp -- the fused DG-SEM Maxwell RK stage for gfx950.
//
// One launch = one stage of cem_maxwell_op_rk (src/cem_maxwell.F:327-345):
// volume curl (maxwell_wght_curl :1428-1497 over local_grad3,
// src/nek5_grad.F:2-19), face restriction (:604-652), on-device face exchange
// (replaces gs_op_fields(gsh_face,...) :962), PEC mirror (:1368-1426),
// upwind/central flux (:922-1002), lift (:654-756), inverse mass
// (:1833-1904) and the 2N-storage RK update (rk4_upd,
// src/cem_common.F:18-76), fused so every field value crosses HBM once per
// stage in each direction.
//
// Mapping: one workgroup per element, four roles of n1 x n1 lanes each.
// Every global load of the element (the six fields, this role's RK register,
// the neighbour traces) is issued up front, so an element costs one HBM
// round trip.  The element is staged in LDS; lane (i,j) of a role reads rows
// (r-derivative), columns (s-derivative) and z-columns (t-derivative).  In
// the face phase the four roles take one x/y face each with the lanes mapped
// onto face nodes; the fluxes go through LDS to the owning lanes.  The two z
// faces are evaluated by the roles whose level half holds them.  Geometry is
// per element (axis-aligned affine hexahedra):
// g = (rx, sy, tz)/J and a lift factor per face (area / (J w3) at the face
// node, constant on an affine face).
//
// Field order in HBM: 0 Hx, 1 Hy, 2 Hz, 3 Ex, 4 Ey, 5 Ez, each a contiguous
// array of npts doubles at stride fs, volume index i + n1*j + n1^2*k +
// n1^3*e as in the reference.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Build-time switches (0 turns the non-temporal forms off at every order).
#ifndef NKC_NT_STORE
#define NKC_NT_STORE 1      // non-temporal stores for the streamed outputs
#endif
#ifndef NKC_NT_LOAD_K
#define NKC_NT_LOAD_K 1     // non-temporal loads of the RK register
#endif

namespace nkc {

// Whether the streamed outputs and the RK register go past L2 with
// non-temporal stores and loads, per order.  At N = 7 a role's level is 64
// doubles, whole cache lines from one wave, and streaming them keeps L2 for
// the neighbour traces (32^3: 619 against 639-648 us per stage with plain
// accesses); at the other orders a level's (N+1)^2 doubles straddle lines
// and the plain forms measured faster (N = 4 670 -> 534 us, N = 5 555 ->
// 495, N = 6 538 -> 498-504, N = 8 410 -> 391, N = 9 level).
template <int N1>
constexpr bool nt_streams() { return N1 == 8; }

template <bool NT, class T>
__device__ __forceinline__ void st_stream(T *p, T v) {
    if constexpr (NT && NKC_NT_STORE) __builtin_nontemporal_store(v, p);
    else *p = v;
}

template <bool NT, class T>
__device__ __forceinline__ T ld_stream(const T *p) {
    if constexpr (NT && NKC_NT_LOAD_K) return __builtin_nontemporal_load(p);
    else return *p;
}

// Kernel face order (symmetric): 0 x-, 1 x+, 2 y-, 3 y+, 4 z-, 5 z+.
enum FaceType : int32_t { FACE_PEC = -1, FACE_NONE = -2 };

enum StageMode : int { MODE_RK = 0, MODE_RESIDUAL = 1 };
enum MatMode : int { MAT_ELEM = 0, MAT_NODE = 1 };

struct StageArgs {
    const double *__restrict__ uin;   // 6 fields, stride fs
    double *__restrict__ uout;        // MODE_RK: updated fields
    double *__restrict__ kbuf;        // MODE_RK: RK register, in place
    double *__restrict__ res;         // MODE_RESIDUAL: residual out
    const double *__restrict__ D;     // n1*n1, D(i,l) at D[i + n1*l]
    const double *__restrict__ geo;   // per element: gx gy gz L0..L5 (9)
    const int32_t *__restrict__ nbr;  // per element: 6 neighbour ids / FaceType
    const double *__restrict__ fco;   // ELEM: e*6*4 ; NODE: (e*6+f)*n2*4
    const double *__restrict__ ieps;  // ELEM: per element ; NODE: per node
    const double *__restrict__ imu;
    int64_t fs;                       // field stride (doubles)
    int32_t nelt;
    int32_t per_xcd;                  // elements per XCD slice (grid remap)
    double rka, rkb, dt;
    // MODE_RK: bit 0 reads the RK register, bit 1 writes it.  The first
    // stage's a_1 = 0 makes the read moot and the next step's first stage
    // ignores what the last stage would write, so a step's first stage
    // skips the read and its last stage the write (10 % of a step's bytes).
    int32_t kflags;
    const double *__restrict__ dt_e;  // per-element dt (batched cavities) or null
    // stage time = *tdev + rkc*dt (tdev advanced on the device after stage 5)
    const double *__restrict__ tdev;
    double rkc;
    // usersrc-style source on E: src_amp (un-weighted, per node of the
    // elements with src_off[e] >= 0) * dir * g(t); g(t) = exp(-((t-t0)/tw)^2)
    // (src_kind 0) or sin(src_om t + src_ph) (src_kind 1)
    const int32_t *__restrict__ src_off;
    const double *__restrict__ src_amp;
    double src_dir[3], src_t0, src_tw;
    int32_t src_kind;
    double src_om, src_ph;
    // Drude ADE (cem_maxwell_drude): elements with drude_off[e] >= 0 carry
    // J, kJ ([3][n1^3] each) and (a, b) per node
    const int32_t *__restrict__ drude_off;
    double *__restrict__ dJ;
    double *__restrict__ dkJ;
    const double *__restrict__ dpar;
    double *__restrict__ dresJ;       // MODE_RESIDUAL: resJ out (same layout as J)
    // Lorentz ADE (cem_maxwell_lorentz): elements with lor_off[e] >= 0 carry
    // J then P (6 n1^3 doubles at lor_off[e] 6 n1^3) and their RK register,
    // per-node (a, b, c) at (lor_off[e] n1^3 + vloc) 3
    const int32_t *__restrict__ lor_off;
    double *__restrict__ lL;
    double *__restrict__ lkL;
    const double *__restrict__ lpar;
    double *__restrict__ lresL;       // MODE_RESIDUAL: (resJ, resP) out
    // MODE_RK, adjoint recompute: keep the stage residual (6 fields, stride
    // fs) and the stage-input Drude current (layout of dJ) when non-null
    double *__restrict__ st_r;
    double *__restrict__ st_J;
    // PML (pml_step): elements with pml_off[e] >= 0 carry sigma x, y, z
    // (3 n1^3 doubles at pml_off[e] 3 n1^3) and the auxiliary fields B then
    // D (6 n1^3 at pml_off[e] 6 n1^3) with their RK register
    const int32_t *__restrict__ pml_off;
    const double *__restrict__ psig;
    double *__restrict__ pP;
    double *__restrict__ pkP;
    double *__restrict__ presP;       // MODE_RESIDUAL: (resB, resD) out
    // userinc-style incident plane wave: inc_mask[e] bit f (kernel face
    // order) adds it to the element's own trace on face f, bit 8 + f to the
    // neighbour's trace seen across face f; u = cos(inc_k . x - inc_om t)
    // at the trace's node (coordinates xyz, 3 x fs)
    const int32_t *__restrict__ inc_mask;
    const double *__restrict__ xyz;
    double inc_k[3], inc_om, inc_h[3], inc_e[3];
    // graphene sheet current (userfsrc + cem_3d_graphene_current):
    // gr_off[e*6 + f] = the sheet block of face f (kernel order) or -1.  Per
    // block and face node q (the flux's in-face index, the same on both
    // sides of a face): 14 doubles at gpar[(b*14 + c)*n2 + q] (the twelve
    // sheet parameters, Yfac = 0.5/Y_0, yconduc) and the state (fjn slots
    // 2..6 x 3 components) at [(b*15 + c)*n2 + q], read from gA and written
    // to gAn (both sides of a face read each other's state, so the stages
    // alternate two copies), the RK register gK in place.
    const int32_t *__restrict__ gr_off;
    const double *__restrict__ gpar;
    const double *__restrict__ gA;
    double *__restrict__ gAn;
    double *__restrict__ gK;
    // xcd_bal: XCD x takes elements xcd_lo[x] .. xcd_lo[x + 1] - 1 (ranges
    // of equal work when some elements cost more, e.g. a Drude region)
    // instead of equal runs of per_xcd elements (one element per workgroup)
    int32_t xcd_bal;
    int32_t xcd_lo[9];
};

// Incident plane wave at volume node p (global index), time t, added to a
// face trace (H then E).
__device__ __forceinline__ void add_incident(const StageArgs &a, int64_t p, double t, double tr[6]) {
    const double u = cos(a.inc_k[0] * a.xyz[p] + a.inc_k[1] * a.xyz[a.fs + p] + a.inc_k[2] * a.xyz[2 * a.fs + p] -
                         a.inc_om * t);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        tr[c] = tr[c] + a.inc_h[c] * u;
        tr[3 + c] = tr[3 + c] + a.inc_e[c] * u;
    }
}

// n x v for a general n.
__device__ __forceinline__ void gcross(const double n[3], const double v[3], double o[3]) {
    o[0] = n[1] * v[2] - n[2] * v[1];
    o[1] = n[2] * v[0] - n[0] * v[2];
    o[2] = n[0] * v[1] - n[1] * v[0];
}

// n x v for n = sg * e_AX.
template <int AX>
__device__ __forceinline__ void ncross(double sg, const double v[3], double o[3]) {
    if (AX == 0) { o[0] = 0.0;        o[1] = -sg * v[2]; o[2] = sg * v[1]; }
    if (AX == 1) { o[0] = sg * v[2];  o[1] = 0.0;        o[2] = -sg * v[0]; }
    if (AX == 2) { o[0] = -sg * v[1]; o[1] = sg * v[0];  o[2] = 0.0; }
}

// Sheet current J at node q of sheet block b for the side whose outward
// normal is sg e_AX, from that side's traces tr (H then E, the incident wave
// included) and the stage-input state (cem_3d_graphene_current,
// src/cem_maxwell.F:2860-2886):
//   tmp = Yfac (-n x H + yconduc n x (E x n)),
//   J = (s2 + s3 + s5 - (b_cp1 + b_cp2) tmp) / (1 - (b_cp1 + b_cp2) Yfac).
// With upd the side's state takes this stage's RK update (:2887-2929):
// f = tmp - Yfac J drives the Drude slot 2 and the critical-point pairs
// (3, 4) and (5, 6).
template <int AX, int N2>
__device__ __forceinline__ void sheet_current(const StageArgs &a, int b, int q, double sg, const double tr[6],
                                              bool upd, double dt, double J[3]) {
    const double *p = a.gpar + (int64_t)b * 14 * N2 + q;
    const double *A = a.gA + (int64_t)b * 15 * N2 + q;
    double nH[3];
    ncross<AX>(-sg, tr, nH);
    double nEn[3] = {tr[3], tr[4], tr[5]};
    nEn[AX] = 0.0;   // n x (E x n) = E - n (n . E) for n = +-e_AX
    const double Yfac = p[12 * N2], yc = p[13 * N2];
    const double cpf = p[2 * N2] + p[7 * N2];
    const double jnf = 1.0 - cpf * Yfac;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double A0 = A[c * N2], A1 = A[(3 + c) * N2], A3 = A[(9 + c) * N2];
        const double tmp = Yfac * (nH[c] + yc * nEn[c]);
        J[c] = (A0 + A1 + A3 - cpf * tmp) / jnf;
        if (upd) {
            const double A2 = A[(6 + c) * N2], A4 = A[(12 + c) * N2];
            const double f = tmp - Yfac * J[c];
            double r[5];
            r[0] = -p[0] * A0 + p[N2] * f;
            r[1] = A2 + p[5 * N2] * f;
            r[2] = -p[3 * N2] * A1 - p[4 * N2] * A2 + p[6 * N2] * f;
            r[3] = A4 + p[10 * N2] * f;
            r[4] = -p[8 * N2] * A3 - p[9 * N2] * A4 + p[11 * N2] * f;
            const double Ao[5] = {A0, A1, A2, A3, A4};
            const int64_t o = (int64_t)b * 15 * N2 + q;
#pragma unroll
            for (int sl = 0; sl < 5; ++sl) {
                const int64_t x = o + (3 * sl + c) * N2;
                const double kn = a.rka * a.gK[x] + dt * r[sl];
                a.gK[x] = kn;
                a.gAn[x] = Ao[sl] + a.rkb * kn;
            }
        }
    }
}

// The sheet term of face f's flux at node q: -(J own + J neighbour), the
// sum the face gather forms of what userfsrc subtracts from -n x H on each
// side.  The own side's state advances when upd.
template <int AX, int N2>
__device__ __forceinline__ void sheet_flux(const StageArgs &a, int e, int f, int nbr, int q, double sg,
                                           const double um[6], const double up[6], bool upd, double dt,
                                           double g[3]) {
    g[0] = g[1] = g[2] = 0.0;
    double J[3];
    const int b = a.gr_off[e * 6 + f];
    if (b >= 0) {
        sheet_current<AX, N2>(a, b, q, sg, um, upd, dt, J);
        g[0] -= J[0]; g[1] -= J[1]; g[2] -= J[2];
    }
    if (nbr >= 0) {
        const int bn = a.gr_off[nbr * 6 + (f ^ 1)];
        if (bn >= 0) {
            sheet_current<AX, N2>(a, bn, q, -sg, up, false, dt, J);
            g[0] -= J[0]; g[1] -= J[1]; g[2] -= J[2];
        }
    }
}

// Flux at one face node (cem_maxwell_flux3d loops 1-2 + gs + PEC), scaled by
// the lift factor L.  um/up: own/neighbour fields in HBM order (H then E).
// Returns FH (lift into resH) and FE (lift into resE).  g (or null): the
// sheet term added to the gathered -n x H before the PEC rule.
template <int AX>
__device__ __forceinline__ void face_flux(double sg, int type, const double um[6],
                                          const double up[6], const double cf[4],
                                          double L, double FH[3], double FE[3],
                                          const double *g = nullptr) {
    double sE[3], sH[3];
    if (type >= 0) {
        // n x [E] = (-n x E-) + (-n' x E+), n' = -n
        double dE[3] = {up[3] - um[3], up[4] - um[4], up[5] - um[5]};
        double dH[3] = {up[0] - um[0], up[1] - um[1], up[2] - um[2]};
        ncross<AX>(sg, dE, sE);
        ncross<AX>(sg, dH, sH);
        if (g) { sH[0] += g[0]; sH[1] += g[1]; sH[2] += g[2]; }
    } else {
        // no partner: gs leaves -n x E-, -n x H-; PEC doubles E, zeroes H
        double Em[3] = {-um[3], -um[4], -um[5]};
        double Hm[3] = {-um[0], -um[1], -um[2]};
        ncross<AX>(sg, Em, sE);
        if (type == FACE_PEC) {
            sE[0] *= 2.0; sE[1] *= 2.0; sE[2] *= 2.0;
            sH[0] = 0.0; sH[1] = 0.0; sH[2] = 0.0;
        } else {
            ncross<AX>(sg, Hm, sH);
            if (g) { sH[0] += g[0]; sH[1] += g[1]; sH[2] += g[2]; }
        }
    }
    double fu[3], fw[3];
    ncross<AX>(sg, sH, fu);
    ncross<AX>(sg, sE, fw);
    // cf = (Y02, C02Y, Z02, C02Z)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        FH[c] = L * (cf[0] * sE[c] - cf[1] * fu[c]);
        FE[c] = L * (cf[2] * sH[c] - cf[3] * fw[c]);
    }
}

__device__ __forceinline__ int remap_element(int b, int nelt, int per_xcd) {
    // Workgroups are dealt round-robin to the 8 XCDs; give each XCD a
    // contiguous run of elements so x/y neighbours share its L2.
    int xcd = b & 7;
    int slot = b >> 3;
    return xcd * per_xcd + slot;
}

// The element of workgroup b of a one-element-per-workgroup stage launch
// (nwg when the workgroup has none).  BAL: the instance reads the
// equal-work ranges (those with in-stage physics, whose elements differ in
// cost; the plain instances keep the registers the table would take).
template <bool BAL>
__device__ __forceinline__ int stage_element(const StageArgs &a, int b, int nwg) {
    if (BAL && a.xcd_bal) {
        const int x = b & 7, e = a.xcd_lo[x] + (b >> 3);
        return e < a.xcd_lo[x + 1] ? e : nwg;
    }
    return remap_element(b, nwg, a.per_xcd);
}

// Value of v in lane `lane` of the wave (lane a compile-time constant or
// wave-uniform), for every lane: two v_readlane_b32 into a scalar pair.
__device__ __forceinline__ double bcast_lane(double v, int lane) {
    const long long b = __builtin_bit_cast(long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)b, lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __builtin_bit_cast(double, (long long)(((unsigned long long)hi << 32) | lo));
}

template <int N1, int MAT>
__device__ __forceinline__ void load_coef(const StageArgs &a, int e, int f, int q, double c4[4]) {
    const double *p = (MAT == MAT_ELEM) ? a.fco + ((int64_t)e * 6 + f) * 4
                                        : a.fco + (((int64_t)e * 6 + f) * (N1 * N1) + q) * 4;
    c4[0] = p[0]; c4[1] = p[1]; c4[2] = p[2]; c4[3] = p[3];
}

// Launch shape and LDS layout for polynomial order N1-1.
//   ROLES = 4 waves-groups per element: role & 1 selects the update group
//   (0: H from -curl E, 1: E from curl H), role >> 1 the half of the levels.
//   In phase 1 role r computes the fluxes of x/y face r.
//   LDS element layout: u(i,j,k) of field f at f*N3 + k*N2 + j*N1 + (i ^ j)
//   for power-of-two N1 (XOR swizzle: rows, columns and z-columns are all
//   bank-conflict free), plain i otherwise.
//   PACK = 0: each role owns whole waves (roundup(N2, 64) lanes); PACK = 1:
//   the four roles' N2 lanes are packed back to back and the workgroup is
//   roundup(4 N2, 64) lanes (fewer idle lanes when N2 is not a multiple of
//   64, at the cost of waves that straddle two roles).
template <int N1, int PACK = 0>
struct Shape {
    static constexpr int N2 = N1 * N1, N3 = N2 * N1;
    static constexpr int ROLES = 4;
    static constexpr int LANES = PACK ? N2 : ((N2 + 63) / 64) * 64;
    static constexpr int BS = PACK ? ((ROLES * N2 + 63) / 64) * 64 : LANES * ROLES;
    static constexpr int KH = (N1 + 1) / 2;           // levels in the lower half
    static constexpr bool SWZ = (N1 & (N1 - 1)) == 0;
    static constexpr int LDS_DOUBLES = 6 * N3 + 24 * N2 + N2;
    __device__ static __forceinline__ int idx(int i, int j, int k) {
        return (SWZ ? (i ^ j) : i) + N1 * j + N2 * k;
    }
};


// OCC = 1 asks the register allocator for 4 waves per SIMD (128 VGPRs,
// spilling if it must); OCC = 0 leaves it free, except that the FEAT = 1
// instance of a 4-wave workgroup is held to 3 waves per SIMD (168 VGPRs:
// three elements per CU instead of two; a few spills, 10 % faster at N = 7
// with a source).  The autotuner picks.
// FEAT = 1 compiles the in-stage source, Drude and Lorentz terms, FEAT = 2
// the PML, the incident field and the graphene sheet as well; FEAT = 0
// leaves them all out, so the plain step carries none of their registers or
// instructions.
// EPW = elements per workgroup: EPW independent element blocks of BS
// lanes side by side, each with its own LDS, sharing s_D and the barriers
// (workgroup w takes elements EPW w' .. EPW w' + EPW - 1, w' its slot in the
// XCD-aware order over workgroups).
template <int N1, int MAT, int MODE, int PACK, int OCC, int FEAT, int EPW = 1>
__global__ void __launch_bounds__((Shape<N1, PACK>::BS * EPW)) __attribute__((amdgpu_waves_per_eu(OCC ? 4 : (FEAT == 1 && Shape<N1, PACK>::BS * EPW <= 256 ? 3 : 1))))
stage_kernel(StageArgs a) {
    using S = Shape<N1, PACK>;
    constexpr int N2 = S::N2, N3 = S::N3, KH = S::KH, BS = S::BS, LANES = S::LANES;
    constexpr bool NT = nt_streams<N1>();
    __shared__ double s_u_all[EPW][6 * N3];   // element, swizzled
    // ZF: the z-face fluxes are formed in phase 1 as well (roles 0 and 3,
    // once per face) and read back from LDS at the end levels, so phase 2
    // holds no z-neighbour trace or z-face coefficients (N = 9: 147 -> 125
    // VGPRs, stage 433 -> 387 us plain, 759 -> 673 us with Drude).  Not at
    // N = 7, where the extra 12 n1^2 doubles of LDS would cost the fourth
    // workgroup per CU, nor at N = 6, where it measured 7 % slower plain.
    constexpr bool ZF = N1 >= 9 || N1 <= 6;
    __shared__ double s_f_all[EPW][(ZF ? 36 : 24) * N2];   // [face][comp 0..5][a + n1 b]
    __shared__ double s_D[N2];          // D(i,l) at [i + n1 l]

    const int sub = (EPW > 1) ? (int)threadIdx.x / BS : 0;
    const int e = (EPW == 1 ? stage_element<(FEAT >= 1)>(a, blockIdx.x, a.nelt)
                            : remap_element(blockIdx.x, (a.nelt + EPW - 1) / EPW, a.per_xcd) * EPW) + sub;
    if (e - sub >= a.nelt) return;   // whole workgroup leaves together
    // with EPW > 1 the last workgroup's spare blocks stay for the barriers
    const bool valid = EPW == 1 || e < a.nelt;
    double *const s_u = s_u_all[sub];
    double *const s_f = s_f_all[sub];
    const int t = (EPW > 1) ? (int)threadIdx.x % BS : (int)threadIdx.x;
    // with whole waves per role (PACK = 0) the role, and with it the
    // update group, the level half and the face a lane works on, is the
    // same across a wave: the compiler is told so and keeps them (and what
    // derives from them) in scalar registers
    const int role = PACK ? t / LANES : __builtin_amdgcn_readfirstlane(t / LANES);
    const int q = t % LANES;
    const bool act = valid && q < N2 && role < S::ROLES;
    const int i = q % N1, j = q / N1;
    const int grp = role & 1;
    const int half = (role >> 1) & 1;
    const int k0 = half ? KH : 0;
    const int64_t eo = (int64_t)e * N3;
    const int64_t fs = a.fs;

    int nb[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) nb[f] = valid ? a.nbr[e * 6 + f] : FACE_NONE;

    // ---- phase 0: every global load of the element, issued up front ----
    // At N = 7 the neighbours' traces go first: they then land with the
    // element's own streams (one in-order counter) instead of holding up
    // the face phase (580.7 against 584.7 us per stage at 32^3); at the other
    // orders they follow the own streams (first measured 0.4-0.9 % slower)
    constexpr bool NBR_FIRST = N1 == 8 && FEAT < 2;
    double pre[NBR_FIRST ? 12 : 1];
    if constexpr (NBR_FIRST) {
        int64_t nx = 0;
        if (role == 0) nx = (N1 - 1) + N1 * i + N2 * j;
        if (role == 1) nx = 0 + N1 * i + N2 * j;
        if (role == 2) nx = i + N1 * (N1 - 1) + N2 * j;
        if (role == 3) nx = i + N2 * j;
        const int bxy = nb[role], bz = nb[half ? 5 : 4];
        const int64_t nzn = half ? (int64_t)q : (int64_t)(N1 - 1) * N2 + q;
        const bool zr = !ZF || role == 0 || role == 3;
#pragma unroll
        for (int c = 0; c < 6; ++c)
            pre[c] = (act && bxy >= 0) ? a.uin[c * fs + (int64_t)bxy * N3 + nx] : 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c)
            pre[6 + c] = (act && zr && bz >= 0) ? a.uin[c * fs + (int64_t)bz * N3 + nzn] : 0.0;
    }
    constexpr int NLOAD = (6 * N3 + BS - 1) / BS;
    double ld[NLOAD];
#pragma unroll
    for (int n = 0; n < NLOAD; ++n) {
        const int v = t + BS * n;
        ld[n] = (valid && v < 6 * N3) ? a.uin[(v / N3) * fs + eo + (v % N3)] : 0.0;
    }
    // RK register of this role's fields and levels, up front (KPRE) or, in
    // the 4-waves/SIMD layouts, level by level in phase 2 (3 KH fewer live
    // doubles, which the 128-VGPR budget would otherwise spill)
    constexpr bool KPRE = !OCC;
    // KPF (the 4-waves/SIMD layouts without in-stage physics): the first
    // level's register up front, each next one a level ahead
    constexpr bool KPF = !KPRE && FEAT == 0;
    double kp[3][KPRE ? KH : 1];
    if (MODE == MODE_RK && act) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int kk = 0; kk < (KPRE ? KH : 1); ++kk) kp[c][kk] = 0.0;
    }
    if ((KPRE || KPF) && MODE == MODE_RK && act && (a.kflags & 1)) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int kk = 0; kk < (KPRE ? KH : 1); ++kk) {
                const int k = k0 + kk;
                kp[c][kk] = (k < N1) ? ld_stream<NT>(a.kbuf + (3 * grp + c) * fs + eo + k * N2 + q) : 0.0;
            }
    }
    // source amplitude of this role's levels (E roles), with the HBM
    // round trip of the other streams
    const int so = (FEAT && grp == 1 && a.src_off && valid) ? a.src_off[e] : -1;
    double samp[KH];
#pragma unroll
    for (int kk = 0; kk < KH; ++kk) {
        const int k = k0 + kk;
        samp[kk] = (so >= 0 && act && k < N1) ? a.src_amp[(int64_t)so * N3 + k * N2 + q] : 0.0;
    }
    // neighbour trace for this role's x/y face (node (a,b) = (i,j))
    double nxy[6];
    int64_t nbn_xy = 0;
    if (role == 0) nbn_xy = (N1 - 1) + N1 * i + N2 * j;   // x-: nbr x+ face
    if (role == 1) nbn_xy = 0 + N1 * i + N2 * j;          // x+: nbr x- face
    if (role == 2) nbn_xy = i + N1 * (N1 - 1) + N2 * j;   // y-: nbr y+ face
    if (role == 3) nbn_xy = i + N2 * j;                   // y+: nbr y- face
    const int nbr_xy = nb[role];
#pragma unroll
    for (int c = 0; c < 6; ++c)
        nxy[c] = NBR_FIRST ? pre[NBR_FIRST ? c : 0] : (act && nbr_xy >= 0) ? a.uin[c * fs + (int64_t)nbr_xy * N3 + nbn_xy] : 0.0;
    // neighbour trace of the z face inside this role's level half (with ZF
    // only roles 0 and 3 take the z- / z+ face, in phase 1)
    double nz[6];
    const int zface = half ? 5 : 4;
    const int nbr_z = nb[zface];
    const int64_t nbn_z = half ? (int64_t)q : (int64_t)(N1 - 1) * N2 + q;
    const bool zrole = !ZF || role == 0 || role == 3;
#pragma unroll
    for (int c = 0; c < 6; ++c)
        nz[c] = NBR_FIRST ? pre[NBR_FIRST ? 6 + c : 0] : (act && zrole && nbr_z >= 0) ? a.uin[c * fs + (int64_t)nbr_z * N3 + nbn_z] : 0.0;

    if ((int)threadIdx.x < N2) s_D[threadIdx.x] = a.D[threadIdx.x];
#pragma unroll
    for (int n = 0; n < NLOAD; ++n) {
        const int v = t + BS * n;
        if (v < 6 * N3) {
            const int f = v / N3, r = v % N3;
            s_u[f * N3 + S::idx(r % N1, (r / N1) % N1, r / N2)] = ld[n];
        }
    }
    const double *geo = a.geo + (int64_t)(valid ? e : 0) * 9;
    const int imask = (FEAT == 2 && a.inc_mask && valid) ? a.inc_mask[e] : 0;
    // stage time for userinc (phase 1); usersrc takes it in phase 2
    const double tinc = (FEAT == 2 && imask) ? *a.tdev + a.rkc * (a.dt_e ? a.dt_e[e] : a.dt) : 0.0;   // imask != 0: e valid
    const double dt_stage = (FEAT == 2 && a.dt_e && valid) ? a.dt_e[e] : a.dt;   // the sheet's RK update
    __syncthreads();

    // ---- phase 1: role r computes the flux of x/y face r at node (i,j) ----
    if (act) {
        int vi, vj, vk;
        if (role == 0) { vi = 0;      vj = i;      vk = j; }
        if (role == 1) { vi = N1 - 1; vj = i;      vk = j; }
        if (role == 2) { vi = i;      vj = 0;      vk = j; }
        if (role == 3) { vi = i;      vj = N1 - 1; vk = j; }
        double um[6], cf[4], FH[3], FE[3];
        const int own = S::idx(vi, vj, vk);
#pragma unroll
        for (int c = 0; c < 6; ++c) um[c] = s_u[c * N3 + own];
        load_coef<N1, MAT>(a, e, role, q, cf);
        const double sg = (role & 1) ? 1.0 : -1.0;
        if (FEAT == 2 && (imask >> role) & 1) add_incident(a, eo + own, tinc, um);
        if (FEAT == 2 && (imask >> (8 + role)) & 1) add_incident(a, (int64_t)nbr_xy * N3 + nbn_xy, tinc, nxy);
        double gsh[3];
        const bool sheet = FEAT == 2 && a.gr_off;
        if (sheet) {
            if (role < 2) sheet_flux<0, N2>(a, e, role, nbr_xy, q, sg, um, nxy, MODE == MODE_RK, dt_stage, gsh);
            else          sheet_flux<1, N2>(a, e, role, nbr_xy, q, sg, um, nxy, MODE == MODE_RK, dt_stage, gsh);
        }
        if (role < 2) face_flux<0>(sg, nbr_xy, um, nxy, cf, geo[3 + role], FH, FE, sheet ? gsh : nullptr);
        else          face_flux<1>(sg, nbr_xy, um, nxy, cf, geo[3 + role], FH, FE, sheet ? gsh : nullptr);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            s_f[(role * 6 + c) * N2 + q] = FH[c];
            s_f[(role * 6 + 3 + c) * N2 + q] = FE[c];
        }
        if constexpr (ZF) if (zrole) {
            const int kz = half ? N1 - 1 : 0;
            const int ownz = S::idx(i, j, kz);
            double cz[4];
#pragma unroll
            for (int c = 0; c < 6; ++c) um[c] = s_u[c * N3 + ownz];
            load_coef<N1, MAT>(a, e, zface, q, cz);
            if (FEAT == 2 && (imask >> zface) & 1) add_incident(a, eo + ownz, tinc, um);
            if (FEAT == 2 && (imask >> (8 + zface)) & 1) add_incident(a, (int64_t)nbr_z * N3 + nbn_z, tinc, nz);
            if (sheet) sheet_flux<2, N2>(a, e, zface, nbr_z, q, half ? 1.0 : -1.0, um, nz, MODE == MODE_RK, dt_stage, gsh);
            face_flux<2>(half ? 1.0 : -1.0, nbr_z, um, nz, cz, geo[3 + zface], FH, FE, sheet ? gsh : nullptr);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                s_f[(zface * 6 + c) * N2 + q] = FH[c];
                s_f[(zface * 6 + 3 + c) * N2 + q] = FE[c];
            }
        }
    }
    __syncthreads();
    if (!act) return;

    // ---- phase 2: this role's levels; group 0 updates H from -curl E,
    // group 1 updates E from curl H ----
    const double gx = geo[0], gy = geo[1], gz = geo[2];
    const int src = grp ? 0 : 3;             // fields differentiated
    const double sgn = grp ? 1.0 : -1.0;
    const double *S0 = s_u + (src + 0) * N3;
    const double *S1 = s_u + (src + 1) * N3;
    const double *S2 = s_u + (src + 2) * N3;
    // D rows of this lane's column cached in registers, except where
    // 2 (N+1) doubles would cost occupancy: read from LDS.  At N = 8, 9 with
    // whole waves per role the 4-waves layouts, which read the RK register
    // level by level, hold both rows within their 128 VGPRs (not the PML
    // instance, nor the source / Drude one at N = 9, which would spill), the
    // free-budget ones the r row only: the level loop's LDS reads, which
    // bound it at these orders, drop from 9 (N+1) per level to 6 (N+1)
    constexpr bool DREG = (N1 <= 8 && !OCC) || (N1 >= 9 && !PACK && OCC && (FEAT == 0 || (FEAT == 1 && N1 == 9)));
    constexpr bool DREGR = DREG || (N1 >= 9 && !PACK && FEAT < 2);
    double Dr[DREGR ? N1 : 1], Ds[DREG ? N1 : 1];
    if (DREGR) {
#pragma unroll
        for (int l = 0; l < N1; ++l) Dr[l] = s_D[i + N1 * l];
    }
    if (DREG) {
#pragma unroll
        for (int l = 0; l < N1; ++l) Ds[l] = s_D[j + N1 * l];
    }
    // t derivative with whole waves per role: the level k is the same
    // across the wave, so lane l reads D(k, l) once per level and the wave
    // takes the N+1 entries from lanes 0..N (always active) by readlane, in
    // place of N+1 LDS reads.  At N = 8, 9 the plain and source / Drude
    // instances (not the free-budget source / Drude one at N = 8, whose
    // registers it would push past 128); at N = 5, 6 the plain ones, at
    // N = 7 the plain 4-waves one (the free-budget one would pass 128 and
    // lose its fourth workgroup per CU)
    constexpr bool DTL = (N1 >= 9 && !PACK && (FEAT == 0 || (FEAT == 1 && (OCC || N1 == 10)))) ||
                         (N1 >= 6 && N1 <= 8 && !PACK && FEAT == 0 && (OCC || N1 <= 7));
    const int wl = (int)threadIdx.x & 63;
    double minv = 0.0;
    if (MAT == MAT_ELEM) minv = grp ? a.ieps[e] : a.imu[e];
    const double dt = a.dt_e ? a.dt_e[e] : a.dt;
    double cfz[4];
    if constexpr (!ZF) load_coef<N1, MAT>(a, e, zface, q, cfz);
    // usersrc time factor: one value per stage (and element, with per-element dt)
    double sg = 0.0;
    if (so >= 0) {
        const double tstage = *a.tdev + a.rkc * dt;
        if (a.src_kind == 1) {
            sg = sin(a.src_om * tstage + a.src_ph);
        } else {
            const double tt = (tstage - a.src_t0) / a.src_tw;
            sg = exp(-tt * tt);
        }
    }
    const int dd = (FEAT && grp == 1 && a.drude_off) ? a.drude_off[e] : -1;
    const int lo = (FEAT && grp == 1 && a.lor_off) ? a.lor_off[e] : -1;
    const int po = (FEAT == 2 && a.pml_off) ? a.pml_off[e] : -1;
    // the x face (0 x-, 1 x+) and the y face (2 y-, 3 y+) this lane's nodes
    // lie on, or -1, and the LDS offsets of their fluxes (level 0)
    const int xface = (i == 0) ? 0 : (i == N1 - 1 ? 1 : -1);
    const int yface = (j == 0) ? 2 : (j == N1 - 1 ? 3 : -1);
    const int xbase = ((xface < 0 ? 0 : xface) * 6 + 3 * grp) * N2 + j;
    const int ybase = ((yface < 0 ? 2 : yface) * 6 + 3 * grp) * N2 + i;
#pragma unroll
    for (int kk = 0; kk < KH; ++kk) {
        const int k = k0 + kk;
        if (k >= N1) break;
        const int vloc = k * N2 + q;
        double kold[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) kold[c] = kp[c][KPRE ? kk : 0];
        if (KPF && MODE == MODE_RK && (a.kflags & 1) && kk + 1 < KH && k + 1 < N1) {
            // the next level's RK register, issued before this level's
            // stores: a load issued after stores is waited on together with
            // them (one in-order counter), so per-level loads would wait
            // for every level's write acknowledgements
#pragma unroll
            for (int c = 0; c < 3; ++c) kp[c][0] = ld_stream<NT>(a.kbuf + (3 * grp + c) * fs + eo + vloc + N2);
        } else if (!KPRE && !KPF && MODE == MODE_RK && (a.kflags & 1)) {
#pragma unroll
            for (int c = 0; c < 3; ++c) kold[c] = ld_stream<NT>(a.kbuf + (3 * grp + c) * fs + eo + vloc);
        }
        double S1_r = 0, S2_r = 0, S0_s = 0, S2_s = 0, S0_t = 0, S1_t = 0;
        const double dkrow = DTL ? s_D[k + N1 * (wl < N1 ? wl : 0)] : 0.0;
#pragma unroll
        for (int l = 0; l < N1; ++l) {
            const int pr = S::idx(l, j, k), ps = S::idx(i, l, k), pt = S::idx(i, j, l);
            const double dkl = DTL ? bcast_lane(dkrow, l) : s_D[k + N1 * l];
            const double drl = DREGR ? Dr[DREGR ? l : 0] : s_D[i + N1 * l];
            const double dsl = DREG ? Ds[DREG ? l : 0] : s_D[j + N1 * l];
            S1_r += drl * S1[pr];
            S2_r += drl * S2[pr];
            S0_s += dsl * S0[ps];
            S2_s += dsl * S2[ps];
            S0_t += dkl * S0[pt];
            S1_t += dkl * S1[pt];
        }
        double r[3];
        r[0] = sgn * (gy * S2_s - gz * S1_t);
        r[1] = sgn * (gz * S0_t - gx * S2_r);
        r[2] = sgn * (gx * S1_r - gy * S0_s);
        const int own = S::idx(i, j, k);
        if constexpr (ZF) {
            if (k == (half ? N1 - 1 : 0)) {
#pragma unroll
                for (int c = 0; c < 3; ++c) r[c] += s_f[(zface * 6 + 3 * grp + c) * N2 + q];
            }
        } else if (k == (half ? N1 - 1 : 0)) {
            double um[6], FH[3], FE[3];
#pragma unroll
            for (int c = 0; c < 6; ++c) um[c] = s_u[c * N3 + own];
            if (FEAT == 2 && (imask >> zface) & 1) add_incident(a, eo + own, tinc, um);
            if (FEAT == 2 && (imask >> (8 + zface)) & 1) add_incident(a, (int64_t)nbr_z * N3 + nbn_z, tinc, nz);
            // both groups form this face's flux; group 0 advances the sheet
            double gsh[3];
            const bool sheet = FEAT == 2 && a.gr_off;
            if (sheet) sheet_flux<2, N2>(a, e, zface, nbr_z, q, half ? 1.0 : -1.0, um, nz, MODE == MODE_RK && grp == 0, dt, gsh);
            face_flux<2>(half ? 1.0 : -1.0, nbr_z, um, nz, cfz, geo[3 + zface], FH, FE, sheet ? gsh : nullptr);
#pragma unroll
            for (int c = 0; c < 3; ++c) r[c] += grp ? FE[c] : FH[c];
        }
        const int co = 3 * grp;
        // a node lies on at most one x face and one y face: one lift each,
        // x before y as in the face order
        if (xface >= 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) r[c] += s_f[xbase + c * N2 + N1 * k];
        }
        if (yface >= 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) r[c] += s_f[ybase + c * N2 + N1 * k];
        }
        const int64_t node = eo + k * N2 + q;
        if (FEAT == 2 && po >= 0) {
            // pml_step: with sa, sb, sd = sigma/eps of the component's own,
            // next and last axis, the auxiliary field P (B for H, D for E)
            //   resP = res - sb P,  res += (sa - sb) P - sd mu H  (H)
            //                       res += (sa - sb) P - sigma_d E (E)
            const int64_t sb = (int64_t)po * 3 * N3 + vloc, pb = (int64_t)po * 6 * N3 + 3 * grp * N3 + vloc;
            const double sig[3] = {a.psig[sb], a.psig[sb + N3], a.psig[sb + 2 * N3]};
            const double iep = (MAT == MAT_NODE) ? a.ieps[node] : a.ieps[e];
            const double mu = grp ? 0.0 : 1.0 / ((MAT == MAT_NODE) ? a.imu[node] : a.imu[e]);
            const double sp[3] = {sig[0] * iep, sig[1] * iep, sig[2] * iep};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int cb = (c + 1) % 3, cd = (c + 2) % 3;
                const double P = a.pP[pb + c * N3];
                const double rP = r[c] - sp[cb] * P;
                const double drive = grp ? sig[cd] * s_u[(3 + c) * N3 + own] : sp[cd] * mu * s_u[c * N3 + own];
                r[c] += -sp[cb] * P + sp[c] * P - drive;
                if (MODE == MODE_RESIDUAL) {
                    a.presP[pb + c * N3] = rP;
                } else {
                    const double kq = a.rka * a.pkP[pb + c * N3] + dt * rP;
                    a.pkP[pb + c * N3] = kq;
                    a.pP[pb + c * N3] = P + a.rkb * kq;
                }
            }
        }
        if (so >= 0) {
            // usersrc: resE += amp * bm1 * dir * g(t) before the inverse mass
            const double am = samp[kk] * sg;
#pragma unroll
            for (int c = 0; c < 3; ++c) r[c] += am * a.src_dir[c];
        }
        if (dd >= 0) {
            // cem_maxwell_drude: resE -= J bm1; resJ = -a J + b E; J's RK
            // update uses the stage-start E and J
            const int64_t jb = (int64_t)dd * 3 * N3 + vloc;
            const double pa = a.dpar[((int64_t)dd * N3 + vloc) * 2 + 0];
            const double pb = a.dpar[((int64_t)dd * N3 + vloc) * 2 + 1];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double Jc = a.dJ[jb + c * N3];
                const double rJ = -pa * Jc + pb * s_u[(3 + c) * N3 + own];
                r[c] -= Jc;
                if (MODE == MODE_RESIDUAL) {
                    a.dresJ[jb + c * N3] = rJ;
                } else {
                    if (a.st_J) a.st_J[jb + c * N3] = Jc;
                    const double kj = a.rka * a.dkJ[jb + c * N3] + dt * rJ;
                    a.dkJ[jb + c * N3] = kj;
                    a.dJ[jb + c * N3] = Jc + a.rkb * kj;
                }
            }
        }
        if (lo >= 0) {
            // cem_maxwell_lorentz: resE -= J bm1; resJ = -a J - b P + c E;
            // resP = J; J and P advanced by the same RK stage
            const int64_t lb = (int64_t)lo * 6 * N3 + vloc;
            const double *lp = a.lpar + ((int64_t)lo * N3 + vloc) * 3;
            const double la = lp[0], lbp = lp[1], lc = lp[2];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double Jc = a.lL[lb + c * N3], Pc = a.lL[lb + (3 + c) * N3];
                const double rJ = -la * Jc - lbp * Pc + lc * s_u[(3 + c) * N3 + own];
                r[c] -= Jc;
                if (MODE == MODE_RESIDUAL) {
                    a.lresL[lb + c * N3] = rJ;
                    a.lresL[lb + (3 + c) * N3] = Jc;
                } else {
                    const double kj = a.rka * a.lkL[lb + c * N3] + dt * rJ;
                    const double kp2 = a.rka * a.lkL[lb + (3 + c) * N3] + dt * Jc;
                    a.lkL[lb + c * N3] = kj;
                    a.lkL[lb + (3 + c) * N3] = kp2;
                    a.lL[lb + c * N3] = Jc + a.rkb * kj;
                    a.lL[lb + (3 + c) * N3] = Pc + a.rkb * kp2;
                }
            }
        }
        // inverse mass (the J w3 of hbm1/ebm1 is already divided out)
        if (MAT == MAT_NODE) minv = grp ? a.ieps[node] : a.imu[node];
#pragma unroll
        for (int c = 0; c < 3; ++c) r[c] *= minv;
        if (MODE == MODE_RESIDUAL) {
#pragma unroll
            for (int c = 0; c < 3; ++c) st_stream<NT>(a.res + (co + c) * fs + node, r[c]);
        } else {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int64_t p = (co + c) * fs + node;
                const double kk2 = a.rka * kold[c] + dt * r[c];
                if (a.kflags & 2) st_stream<NT>(a.kbuf + p, kk2);
                st_stream<NT>(a.uout + p, s_u[(co + c) * N3 + own] + a.rkb * kk2);
            }
            if (a.st_r) {
#pragma unroll
                for (int c = 0; c < 3; ++c) st_stream<NT>(a.st_r + (co + c) * fs + node, r[c]);
            }
        }
    }
}

}  // namespace nkc
