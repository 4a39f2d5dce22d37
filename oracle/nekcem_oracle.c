/*
 * nekcem_oracle.c -- CPU restatement of NekCEM's 3-D DG-SEM Maxwell RK step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: it is loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's CPU-baseline leg, and by
 * nothing else.  The product path (nekcem_amd/, libnekcem_hip.so) never links,
 * imports or calls it.
 *
 * It restates, loop for loop and in the reference's data layout (column-major
 * volume index i + n1*j + n1^2*k + n1^3*e, Ed's face order, per-face-node
 * arrays of length nxzfl), the following reference routines:
 *
 *   orc_zwgll          nek5_speclib.F:107-121 (ZWGLL), 208-284 (ZWGLJD) -- GLL
 *                      nodes/weights, restated from the closed form (roots of
 *                      (1-x^2)P'_N, w = 2/(N(N+1)P_N^2)), Newton + deflation.
 *   orc_dgll           nek5_speclib.F:807-838 (DGLL)
 *   orc_geom           nek5_coef.F:877-930 (XYZRST), 555-635 (GLMAPM1),
 *                      255 (W3M1), 755-758 (mass BM1), 1150-1237 (AREA3 +
 *                      UNITVEC)
 *   orc_set_fc_ptr     cem_common.F:214-283 (cem_set_fc_ptr) with skpdat from
 *                      nek5_connect11.F:1483-1528 and EFACE from :1067-1072
 *   orc_get_dxmin      nek5_courant.F:2-79 (get_dxmin)
 *   orc_gs_add         jl/gs.c:2074-2100 (fgs_fields, op=add) on the face
 *                      handle: every entry sharing a gs id gets the sum.
 *   orc_materials      cem_maxwell.F:262-325 (cem_maxwell_materials) and
 *                      :183-186 (ebm1 / hbm1)
 *   orc_pec_init       cem_maxwell.F:1338-1366 (cem_maxwell_pec_init)
 *   orc_maxwell_op     cem_maxwell.F:484-508 (cem_maxwell_op) composed of
 *                      :510-602 (cem_maxwell, 3-D branch + chsign),
 *                      :1428-1497 (maxwell_wght_curl) over nek5_grad.F:2-19
 *                      (local_grad3), :604-652 (restrict_to_face),
 *                      :922-1002 (cem_maxwell_flux3d), :1368-1426
 *                      (cem_maxwell_flux_pec), :654-756 (add_flux_to_res),
 *                      :1833-1904 (cem_maxwell_invqmass); plus the Drude ADE
 *                      :3095-3147, a usersrc-style source term, and (in
 *                      orc_maxwell_op_ade / orc_rk_run_ade) the Lorentz
 *                      ADE :3149-3211.
 *   orc_rk_run         cem_maxwell.F:327-345 (cem_maxwell_op_rk), :1906-1968
 *                      (rk_maxwell_ab), cem_common.F:2-16 (rk_c), :18-76
 *                      (rk4_upd), :78-114 (rk_storage, ifrk45 constants)
 *   orc_cem_error      cem_common.F:1335-1355 (cem_error)
 *   orc_curl           cem_maxwell.F:574-582 + chsign, maxwell_wght_curl
 *                      :1428-1497: the volume term alone
 *   orc_maxwell_op_full, orc_rk_run_full
 *                      the above with pml_step (cem_maxwell_pml.F:508-592)
 *                      and the PML rk4_upd of rk_maxwell_ab (cem_maxwell.F:
 *                      1942-1954), a userinc-style incident wave added to
 *                      the face traces (cem_maxwell.F:497-498) and the
 *                      sinusoidal usersrc time factor
 *   orc_maxwell_op_all, orc_rk_run_all
 *                      the above with userfsrc's graphene sheet current
 *                      (cem_3d_graphene_current, cem_maxwell.F:2827-2932,
 *                      called from cem_maxwell_flux3d :941-942)
 *   orc_maxwell_vjp, orc_drude_vjp
 *                      the transposes of the residual's derivatives (the
 *                      discrete adjoint's checker; no reference counterpart)
 *
 * Each is pinned in tests/test_oracle.py by checks that do not restate it:
 * GLL nodes, weights and D against closed forms and exact polynomial
 * derivatives; the RK constants by their stability polynomial; geometry,
 * face pointers, partners, materials and dxmin against what a box fixes;
 * the operator by the reference's own 3dboxper / 3dboxpec cases at their
 * userchk tolerances, exact curls of polynomials, spectral convergence and
 * the skew-adjointness of the central flux; the Drude and Lorentz ADEs and
 * the sources by closed-form uniform-field solutions; the PML, incident
 * wave and graphene sheet by the reference's 3ddielectric and 3dgraphene
 * exact solutions at their userchk tolerances, with planted mistakes
 * failing; the VJPs by dot-product (transpose) identities and central
 * differences.  No function is left unpinned.
 *
 * Built by oracle/Makefile into oracle/_build/libnekcem_oracle.so (gcc,
 * -fopenmp).  OpenMP parallelises over elements / face nodes only; summation
 * order inside each output value follows the reference.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Basis                                                                     */
/* ------------------------------------------------------------------------ */

/* Legendre P_n(x) by the three-term recurrence (PNLEG, nek5_speclib.F). */
double orc_pnleg(double x, int n)
{
    if (n == 0) return 1.0;
    double p0 = 1.0, p1 = x;
    for (int k = 1; k < n; ++k) {
        double p2 = ((2.0 * k + 1.0) * x * p1 - k * p0) / (k + 1.0);
        p0 = p1;
        p1 = p2;
    }
    return p1;
}

/* P_n and P_n' together. */
static void legendre_pd(double x, int n, double *p, double *dp)
{
    if (n == 0) { *p = 1.0; *dp = 0.0; return; }
    double p0 = 1.0, p1 = x, d0 = 0.0, d1 = 1.0;
    for (int k = 1; k < n; ++k) {
        double p2 = ((2.0 * k + 1.0) * x * p1 - k * p0) / (k + 1.0);
        double d2 = d0 + (2.0 * k + 1.0) * p1;
        p0 = p1; p1 = p2; d0 = d1; d1 = d2;
    }
    *p = p1;
    *dp = d1;
}

/* GLL nodes z[0..n1-1] and weights w (ZWGLL, nek5_speclib.F:107-121).
 * Interior nodes are the roots of P_N'(x); Newton with deflation from the
 * Chebyshev-Gauss-Lobatto guesses.  w_i = 2 / (N (N+1) P_N(z_i)^2). */
void orc_zwgll(int n1, double *z, double *w)
{
    int N = n1 - 1;
    if (n1 == 1) { z[0] = 0.0; w[0] = 2.0; return; }
    z[0] = -1.0;
    z[N] = 1.0;
    for (int i = 1; i < N; ++i) {
        double x = -cos(M_PI * (double)i / (double)N);
        for (int it = 0; it < 100; ++it) {
            /* f = P_N'(x), f' = P_N''(x) from (1-x^2)P'' = 2xP' - N(N+1)P */
            double p, dp;
            legendre_pd(x, N, &p, &dp);
            double d2p = (2.0 * x * dp - N * (N + 1.0) * p) / (1.0 - x * x);
            double defl = 0.0;
            for (int j = 1; j < i; ++j) defl += 1.0 / (x - z[j]);
            double dx = dp / (d2p - dp * defl);
            x -= dx;
            if (fabs(dx) < 1e-17) break;
        }
        z[i] = x;
    }
    /* symmetrise (the GLL set is symmetric about 0) */
    for (int i = 0; i < n1 / 2; ++i) {
        double a = 0.5 * (z[N - i] - z[i]);
        z[i] = -a;
        z[N - i] = a;
    }
    if (n1 % 2) z[N / 2] = 0.0;
    for (int i = 0; i < n1; ++i) {
        double p = orc_pnleg(z[i], N);
        w[i] = 2.0 / ((double)N * (N + 1.0) * p * p);
    }
}

/* Derivative matrix, column-major D(i,j) at D[i + n1*j] (DGLL,
 * nek5_speclib.F:807-838). */
void orc_dgll(int n1, const double *z, double *D)
{
    int N = n1 - 1;
    if (n1 == 1) { D[0] = 0.0; return; }
    double fn = (double)N;
    double d0 = fn * (fn + 1.0) / 4.0;
    for (int i = 0; i < n1; ++i)
        for (int j = 0; j < n1; ++j) {
            double d = 0.0;
            if (i != j) d = orc_pnleg(z[i], N) / (orc_pnleg(z[j], N) * (z[i] - z[j]));
            if (i == j && i == 0) d = -d0;
            if (i == j && i == N) d = d0;
            D[i + n1 * j] = d;
        }
}

/* ------------------------------------------------------------------------ */
/* local_grad3 (nek5_grad.F:2-19)                                            */
/*   ur = D * u        as (n1) x (n1^2)                                      */
/*   us(:,:,k) = u(:,:,k) * D^T                                              */
/*   ut = u * D^T      as (n1^2) x (n1)                                      */
/* ------------------------------------------------------------------------ */
static void local_grad3(int n1, const double *D, const double *u,
                        double *ur, double *us, double *ut)
{
    int n2 = n1 * n1;
    for (int jk = 0; jk < n2; ++jk)
        for (int i = 0; i < n1; ++i) {
            double s = 0.0;
            for (int l = 0; l < n1; ++l) s += D[i + n1 * l] * u[l + n1 * jk];
            ur[i + n1 * jk] = s;
        }
    for (int k = 0; k < n1; ++k)
        for (int j = 0; j < n1; ++j)
            for (int i = 0; i < n1; ++i) {
                double s = 0.0;
                for (int l = 0; l < n1; ++l) s += u[i + n1 * l + n2 * k] * D[j + n1 * l];
                us[i + n1 * j + n2 * k] = s;
            }
    for (int k = 0; k < n1; ++k)
        for (int ij = 0; ij < n2; ++ij) {
            double s = 0.0;
            for (int l = 0; l < n1; ++l) s += u[ij + n2 * l] * D[k + n1 * l];
            ut[ij + n2 * k] = s;
        }
}

/* ------------------------------------------------------------------------ */
/* Geometry                                                                  */
/* ------------------------------------------------------------------------ */

/* XYZRST + GLMAPM1 + W3M1 + BM1 + AREA3/UNITVEC for nelt elements.
 * Face arrays are (n1, n1, 6, nelt) in Ed's face order. */
void orc_geom(int n1, int64_t nelt, const double *D, const double *wx,
              const double *xm1, const double *ym1, const double *zm1,
              double *rx, double *ry, double *rz,
              double *sx, double *sy, double *sz,
              double *tx, double *ty, double *tz,
              double *jac, double *bm1, double *w3,
              double *unx, double *uny, double *unz, double *area)
{
    int n2 = n1 * n1, n3 = n2 * n1;
    for (int k = 0; k < n1; ++k)
        for (int j = 0; j < n1; ++j)
            for (int i = 0; i < n1; ++i) w3[i + n1 * j + n2 * k] = wx[i] * wx[j] * wx[k];

#pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * 12 * n3);
        double *xr = buf, *xs = buf + n3, *xt = buf + 2 * n3;
        double *yr = buf + 3 * n3, *ys = buf + 4 * n3, *yt = buf + 5 * n3;
        double *zr = buf + 6 * n3, *zs = buf + 7 * n3, *zt = buf + 8 * n3;
        double *A = buf + 9 * n3, *B = buf + 10 * n3, *C = buf + 11 * n3;
#pragma omp for schedule(static)
        for (int64_t e = 0; e < nelt; ++e) {
            int64_t o = (int64_t)n3 * e;
            local_grad3(n1, D, xm1 + o, xr, xs, xt);
            local_grad3(n1, D, ym1 + o, yr, ys, yt);
            local_grad3(n1, D, zm1 + o, zr, zs, zt);
            for (int i = 0; i < n3; ++i) {
                double J = 0.0;
                J += xr[i] * ys[i] * zt[i];
                J += xt[i] * yr[i] * zs[i];
                J += xs[i] * yt[i] * zr[i];
                J -= xr[i] * yt[i] * zs[i];
                J -= xs[i] * yr[i] * zt[i];
                J -= xt[i] * ys[i] * zr[i];
                jac[o + i] = J;
                rx[o + i] = ys[i] * zt[i] - yt[i] * zs[i];
                ry[o + i] = xt[i] * zs[i] - xs[i] * zt[i];
                rz[o + i] = xs[i] * yt[i] - xt[i] * ys[i];
                sx[o + i] = yt[i] * zr[i] - yr[i] * zt[i];
                sy[o + i] = xr[i] * zt[i] - xt[i] * zr[i];
                sz[o + i] = xt[i] * yr[i] - xr[i] * yt[i];
                tx[o + i] = yr[i] * zs[i] - ys[i] * zr[i];
                ty[o + i] = xs[i] * zr[i] - xr[i] * zs[i];
                tz[o + i] = xr[i] * ys[i] - xs[i] * yr[i];
                bm1[o + i] = J * w3[i];
            }
            int64_t fo = (int64_t)6 * n2 * e;
#define VI(i, j, k) ((i) + n1 * (j) + n2 * (k))
#define FI(a, b, f) (fo + (a) + n1 * (b) + n2 * ((f)-1))
            /* "R": A,B,C = x_s cross x_t */
            for (int i = 0; i < n3; ++i) {
                A[i] = ys[i] * zt[i] - zs[i] * yt[i];
                B[i] = zs[i] * xt[i] - xs[i] * zt[i];
                C[i] = xs[i] * yt[i] - ys[i] * xt[i];
            }
            for (int iz = 0; iz < n1; ++iz)
                for (int iy = 0; iy < n1; ++iy) {
                    double wgt = wx[iy] * wx[iz];
                    int p2 = VI(n1 - 1, iy, iz), p4 = VI(0, iy, iz);
                    area[FI(iy, iz, 2)] = sqrt(A[p2] * A[p2] + B[p2] * B[p2] + C[p2] * C[p2]) * wgt;
                    area[FI(iy, iz, 4)] = sqrt(A[p4] * A[p4] + B[p4] * B[p4] + C[p4] * C[p4]) * wgt;
                    unx[FI(iy, iz, 4)] = -A[p4]; unx[FI(iy, iz, 2)] = A[p2];
                    uny[FI(iy, iz, 4)] = -B[p4]; uny[FI(iy, iz, 2)] = B[p2];
                    unz[FI(iy, iz, 4)] = -C[p4]; unz[FI(iy, iz, 2)] = C[p2];
                }
            /* "S": x_r cross x_t */
            for (int i = 0; i < n3; ++i) {
                A[i] = yr[i] * zt[i] - zr[i] * yt[i];
                B[i] = zr[i] * xt[i] - xr[i] * zt[i];
                C[i] = xr[i] * yt[i] - yr[i] * xt[i];
            }
            for (int iz = 0; iz < n1; ++iz)
                for (int ix = 0; ix < n1; ++ix) {
                    double wgt = wx[ix] * wx[iz];
                    int p1 = VI(ix, 0, iz), p3 = VI(ix, n1 - 1, iz);
                    area[FI(ix, iz, 1)] = sqrt(A[p1] * A[p1] + B[p1] * B[p1] + C[p1] * C[p1]) * wgt;
                    area[FI(ix, iz, 3)] = sqrt(A[p3] * A[p3] + B[p3] * B[p3] + C[p3] * C[p3]) * wgt;
                    unx[FI(ix, iz, 1)] = A[p1]; unx[FI(ix, iz, 3)] = -A[p3];
                    uny[FI(ix, iz, 1)] = B[p1]; uny[FI(ix, iz, 3)] = -B[p3];
                    unz[FI(ix, iz, 1)] = C[p1]; unz[FI(ix, iz, 3)] = -C[p3];
                }
            /* "T": x_r cross x_s */
            for (int i = 0; i < n3; ++i) {
                A[i] = yr[i] * zs[i] - zr[i] * ys[i];
                B[i] = zr[i] * xs[i] - xr[i] * zs[i];
                C[i] = xr[i] * ys[i] - yr[i] * xs[i];
            }
            for (int ix = 0; ix < n1; ++ix)
                for (int iy = 0; iy < n1; ++iy) {
                    double wgt = wx[ix] * wx[iy];
                    int p5 = VI(ix, iy, 0), p6 = VI(ix, iy, n1 - 1);
                    area[FI(ix, iy, 5)] = sqrt(A[p5] * A[p5] + B[p5] * B[p5] + C[p5] * C[p5]) * wgt;
                    area[FI(ix, iy, 6)] = sqrt(A[p6] * A[p6] + B[p6] * B[p6] + C[p6] * C[p6]) * wgt;
                    unx[FI(ix, iy, 5)] = -A[p5]; unx[FI(ix, iy, 6)] = A[p6];
                    uny[FI(ix, iy, 5)] = -B[p5]; uny[FI(ix, iy, 6)] = B[p6];
                    unz[FI(ix, iy, 5)] = -C[p5]; unz[FI(ix, iy, 6)] = C[p6];
                }
            /* UNITVEC */
            for (int q = 0; q < 6 * n2; ++q) {
                int64_t p = fo + q;
                double s = sqrt(unx[p] * unx[p] + uny[p] * uny[p] + unz[p] * unz[p]);
                if (s > 0.0) { unx[p] /= s; uny[p] /= s; unz[p] /= s; }
            }
#undef VI
#undef FI
        }
        free(buf);
    }
}

/* skpdat (nek5_connect11.F:1483-1528) for the symmetric face f in 1..6
 * (x-, x+, y-, y+, z-, z+); out[0..5] = js1, jf1, jskip1, js2, jf2, jskip2
 * (1-based, as in the reference). */
static void skpdat(int n1, int f, int *out)
{
    int NX = n1, NY = n1, NZ = n1;
    switch (f) {
    case 1: out[0] = 1; out[1] = NX * (NY - 1) + 1; out[2] = NX; out[3] = 1; out[4] = NY * (NZ - 1) + 1; out[5] = NY; break;
    case 2: out[0] = 1 + (NX - 1); out[1] = NX * (NY - 1) + 1 + (NX - 1); out[2] = NX; out[3] = 1; out[4] = NY * (NZ - 1) + 1; out[5] = NY; break;
    case 3: out[0] = 1; out[1] = NX; out[2] = 1; out[3] = 1; out[4] = NY * (NZ - 1) + 1; out[5] = NY; break;
    case 4: out[0] = 1 + NX * (NY - 1); out[1] = NX + NX * (NY - 1); out[2] = 1; out[3] = 1; out[4] = NY * (NZ - 1) + 1; out[5] = NY; break;
    case 5: out[0] = 1; out[1] = NX; out[2] = 1; out[3] = 1; out[4] = NY; out[5] = 1; break;
    default: out[0] = 1 + NX * NY * (NZ - 1); out[1] = NX + NX * NY * (NZ - 1); out[2] = 1; out[3] = 1; out[4] = NY; out[5] = 1; break;
    }
}

static const int EFACE[6] = {4, 2, 1, 3, 5, 6};

/* cemface (0-based) of length 6*n1^2*nelt (cem_common.F:214-283). */
void orc_set_fc_ptr(int n1, int64_t nelt, int32_t *cemface)
{
    int n2 = n1 * n1, n3 = n2 * n1;
    for (int64_t e = 0; e < nelt; ++e)
        for (int f = 1; f <= 6; ++f) {
            int ef = EFACE[f - 1];
            int s[6];
            skpdat(n1, f, s);
            int i = 0;
            for (int j2 = s[3]; j2 <= s[4]; j2 += s[5])
                for (int j1 = s[0]; j1 <= s[1]; j1 += s[2]) {
                    i = i + 1;
                    int64_t k = (i - 1) + (int64_t)n2 * (ef - 1) + (int64_t)n2 * 6 * e;
                    cemface[k] = (int32_t)((j1 - 1) + n1 * (j2 - 1) + (int64_t)n3 * e);
                }
        }
}

/* dxmin (nek5_courant.F:2-79), interior nodes only. */
double orc_get_dxmin(int n1, int64_t nelt, const double *xm1, const double *ym1, const double *zm1)
{
    int n2 = n1 * n1, n3 = n2 * n1;
    double d2m = 1.e20;
    for (int64_t e = 0; e < nelt; ++e)
        for (int k = 1; k < n1 - 1; ++k)
            for (int j = 1; j < n1 - 1; ++j)
                for (int i = 1; i < n1 - 1; ++i) {
                    int64_t o = (int64_t)n3 * e;
#define P(a, b, c) (o + (a) + n1 * (b) + n2 * (c))
                    double dx, dy, dz, d2;
                    dx = xm1[P(i + 1, j, k)] - xm1[P(i - 1, j, k)];
                    dy = ym1[P(i + 1, j, k)] - ym1[P(i - 1, j, k)];
                    dz = zm1[P(i + 1, j, k)] - zm1[P(i - 1, j, k)];
                    d2 = dx * dx + dy * dy + dz * dz;
                    if (d2 < d2m) d2m = d2;
                    dx = xm1[P(i, j + 1, k)] - xm1[P(i, j - 1, k)];
                    dy = ym1[P(i, j + 1, k)] - ym1[P(i, j - 1, k)];
                    dz = zm1[P(i, j + 1, k)] - zm1[P(i, j - 1, k)];
                    d2 = dx * dx + dy * dy + dz * dz;
                    if (d2 < d2m) d2m = d2;
                    dx = xm1[P(i, j, k + 1)] - xm1[P(i, j, k - 1)];
                    dy = ym1[P(i, j, k + 1)] - ym1[P(i, j, k - 1)];
                    dz = zm1[P(i, j, k + 1)] - zm1[P(i, j, k - 1)];
                    d2 = dx * dx + dy * dy + dz * dz;
                    if (d2 < d2m) d2m = d2;
#undef P
                }
    return sqrt(d2m) / 2.0;
}

/* gs_op_fields(gsh_face, s, n, nfld, add): entries sharing a gs id are
 * replaced by their sum.  partner[i] is the other entry with i's id (-1 if
 * none); on a conforming hex mesh every face id is shared by at most two
 * entries, so the sum is s(i) + s(partner(i)). */
void orc_gs_add(const int32_t *partner, double *s, int64_t n, int nfld, int64_t stride)
{
    for (int c = 0; c < nfld; ++c) {
        double *v = s + (int64_t)c * stride;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            int64_t p = partner[i];
            if (p > i) {
                double t = v[i] + v[p];
                v[i] = t;
                v[p] = t;
            }
        }
    }
}

/* cem_maxwell_materials (cem_maxwell.F:262-325) and ebm1/hbm1
 * (cem_maxwell.F:183-186).  pec_face[6*e + (ef-1)] != 0 for CBC = 'PEC'. */
void orc_materials(int n1, int64_t nelt, const int32_t *cemface, const int32_t *partner,
                   const int8_t *pec_face, const double *eps, const double *mu,
                   const double *bm1, double *Z0, double *Z1, double *Y0, double *Y1,
                   double *ebm1, double *hbm1)
{
    int n2 = n1 * n1, n3 = n2 * n1;
    int64_t npts = (int64_t)n3 * nelt, nxzfl = (int64_t)6 * n2 * nelt;
    double *zimp = (double *)malloc(sizeof(double) * nxzfl);
    double *ycon = (double *)malloc(sizeof(double) * nxzfl);
    for (int64_t j = 0; j < nxzfl; ++j) {
        int64_t i = cemface[j];
        zimp[j] = sqrt(mu[i] / eps[i]);
        ycon[j] = sqrt(eps[i] / mu[i]);
    }
    memcpy(Z0, zimp, sizeof(double) * nxzfl);
    memcpy(Y0, ycon, sizeof(double) * nxzfl);
    orc_gs_add(partner, Z0, nxzfl, 1, nxzfl);
    orc_gs_add(partner, Y0, nxzfl, 1, nxzfl);
    for (int64_t j = 0; j < nxzfl; ++j) { Y1[j] = 0.0; Z1[j] = 0.0; }
    for (int64_t e = 0; e < nelt; ++e)
        for (int f = 0; f < 6; ++f)
            if (pec_face[6 * e + f])
                for (int j = 0; j < n2; ++j) {
                    int64_t i = e * n2 * 6 + (int64_t)n2 * f + j;
                    Y0[i] = 2.0 * Y0[i];
                    Y1[i] = 2.0 * Y1[i];
                    Z0[i] = 2.0 * Z0[i];
                    Z1[i] = 2.0 * Z1[i];
                }
    for (int64_t j = 0; j < nxzfl; ++j) {
        Z1[j] = Z0[j] - zimp[j];
        Y1[j] = Y0[j] - ycon[j];
        Z0[j] = 0.5 * Z0[j];
        Y0[j] = 0.5 * Y0[j];
    }
    for (int64_t i = 0; i < npts; ++i) {
        ebm1[i] = 1.0 / (eps[i] * bm1[i]);
        hbm1[i] = 1.0 / (mu[i] * bm1[i]);
    }
    free(zimp);
    free(ycon);
}

/* cem_maxwell_pec_init (cem_maxwell.F:1338-1366): list of face-node indices
 * on PEC faces (0-based).  Returns the count. */
int64_t orc_pec_init(int n1, int64_t nelt, const int8_t *pec_face, int32_t *cempec)
{
    int n2 = n1 * n1;
    int64_t l = 0;
    for (int64_t e = 0; e < nelt; ++e)
        for (int f = 0; f < 6; ++f)
            for (int j = 0; j < n2; ++j)
                if (pec_face[6 * e + f]) cempec[l++] = (int32_t)(e * 6 * n2 + (int64_t)f * n2 + j);
    return l;
}

/* ------------------------------------------------------------------------ */
/* The operator                                                              */
/* ------------------------------------------------------------------------ */

typedef struct orc_mesh {
    int32_t n1;
    int32_t nelt;
    int64_t npts;
    int64_t nxzfl;
    const double *D;    /* n1*n1 */
    const double *w3;   /* n1^3 */
    const double *rx, *ry, *rz, *sx, *sy, *sz, *tx, *ty, *tz; /* npts */
    const double *bm1;  /* npts */
    const double *unx, *uny, *unz, *area;                     /* nxzfl */
    const int32_t *cemface;                                   /* nxzfl */
    const int32_t *partner;                                   /* nxzfl */
    const int32_t *cempec;
    int64_t ncempec;
    const double *Y0, *Y1, *Z0, *Z1;                          /* nxzfl */
    const double *hbm1, *ebm1;                                /* npts */
    double C0;                                                /* 1 upwind, 0 central */
    /* Drude ADE (cem_maxwell.F:3095-3147): J over nodes dindex[0..ndrude) */
    int64_t ndrude;
    const int32_t *dindex;
    const double *drude_a;  /* per listed node: gamma */
    const double *drude_b;  /* per listed node: omega_p^2 */
    /* usersrc-style source added to resE before invqmass:
     * resE_c(i) += src_amp(i)*src_dir_c * g(t), g(t) = exp(-((t-t0)/tw)^2) */
    const double *src_amp;  /* npts or NULL; already includes bm1 weighting */
    double src_dir[3];
    double src_t0, src_tw;
    /* scratch (6*nxzfl) */
    double *srflx;
    /* Lorentz ADE (cem_maxwell_lorentz, cem_maxwell.F:3149-3211) over nodes
     * lindex[0..nlorentz): per listed node (a, b, c) at lor_par[3 l ..] */
    int64_t nlorentz;
    const int32_t *lindex;
    const double *lor_par;
    /* usersrc time factor: kind 0 the Gaussian pulse above, kind 1
     * sin(src_omega t + src_phase) (the dipole of tests/3dboxpml/3dboxpml.usr:30-88) */
    int32_t src_kind;
    double src_omega, src_phase;
    /* PML (src/cem_maxwell_pml.F): elements pmlptr[0..npml) (0-based),
     * pmlsigma(npts,3), permittivity and permeability per node */
    int64_t npml;
    const int32_t *pmlptr;
    const double *pmlsigma;
    const double *eps, *mu;
    /* userinc-style incident plane wave on the face points with inc_on set:
     * fH += inc_h u, fE += inc_e u, u = cos(inc_k . x - inc_omega t) at the
     * face point's volume node (tests/3ddielectric/3ddielectric.usr:6-47) */
    const int8_t *inc_on;                 /* nxzfl or NULL */
    const double *xm1, *ym1, *zm1;        /* npts */
    double inc_k[3], inc_omega, inc_h[3], inc_e[3];
    /* graphene surface current (cem_3d_graphene_current,
     * cem_maxwell.F:2827-2932) at the face points gindex[0..ngraph) (0-based
     * nxzfl numbers); per listed point the twelve sheet parameters of
     * tests/3dgraphene/3dgraphene.usr:390-401 (a_d, b_d, b_cp1, a_211, a_221,
     * b_11, b_21, b_cp2, a_212, a_222, b_12, b_22) at gpar[12 l ..] */
    int64_t ngraph;
    const int32_t *gindex;
    const double *gpar;
} orc_mesh;

/* resE = W curl H; resH = -W curl E (cem_maxwell.F:574-582 + 1428-1497) */
static void wght_curl(const orc_mesh *m, const double *u1, const double *u2, const double *u3,
                      double *w1, double *w2, double *w3o)
{
    int n1 = m->n1, n3 = n1 * n1 * n1;
#pragma omp parallel
    {
        double *b = (double *)malloc(sizeof(double) * 9 * n3);
        double *u1r = b, *u1s = b + n3, *u1t = b + 2 * n3;
        double *u2r = b + 3 * n3, *u2s = b + 4 * n3, *u2t = b + 5 * n3;
        double *u3r = b + 6 * n3, *u3s = b + 7 * n3, *u3t = b + 8 * n3;
#pragma omp for schedule(static)
        for (int64_t e = 0; e < m->nelt; ++e) {
            int64_t j = (int64_t)n3 * e;
            local_grad3(n1, m->D, u1 + j, u1r, u1s, u1t);
            local_grad3(n1, m->D, u2 + j, u2r, u2s, u2t);
            local_grad3(n1, m->D, u3 + j, u3r, u3s, u3t);
            for (int i = 0; i < n3; ++i) {
                int64_t k = i + j;
                double w = m->w3[i];
                double u1rw = u1r[i] * w, u1sw = u1s[i] * w, u1tw = u1t[i] * w;
                double u2rw = u2r[i] * w, u2sw = u2s[i] * w, u2tw = u2t[i] * w;
                double u3rw = u3r[i] * w, u3sw = u3s[i] * w, u3tw = u3t[i] * w;
                double rx = m->rx[k], sx = m->sx[k], tx = m->tx[k];
                double ry = m->ry[k], sy = m->sy[k], ty = m->ty[k];
                double rz = m->rz[k], sz = m->sz[k], tz = m->tz[k];
                w1[k] = u3rw * ry + u3sw * sy + u3tw * ty - u2rw * rz - u2sw * sz - u2tw * tz;
                w2[k] = u1rw * rz + u1sw * sz + u1tw * tz - u3rw * rx - u3sw * sx - u3tw * tx;
                w3o[k] = u2rw * rx + u2sw * sx + u2tw * tx - u1rw * ry - u1sw * sy - u1tw * ty;
            }
        }
        free(b);
    }
}

static void maxwell_op_body(const orc_mesh *m, double rktime, const double *hn, const double *en,
                            double *reshn, double *resen, const double *J, double *resJ,
                            const double *L, double *resL, const double *P, double *resP,
                            const double *G, double *resG);

/* cem_maxwell's 3-D branch alone (cem_maxwell.F:574-582 + chsign):
 * resE = W curl H, resH = -W curl E (maxwell_wght_curl :1428-1497). */
void orc_curl(const orc_mesh *m, const double *hn, const double *en, double *reshn, double *resen)
{
    int64_t npts = m->npts;
    wght_curl(m, hn, hn + npts, hn + 2 * npts, resen, resen + npts, resen + 2 * npts);
    wght_curl(m, en, en + npts, en + 2 * npts, reshn, reshn + npts, reshn + 2 * npts);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < 3 * npts; ++i) reshn[i] = -reshn[i];
}

/* One residual evaluation (cem_maxwell_op, cem_maxwell.F:484-508).
 * hn, en, reshn, resen: 3 components each at stride npts.
 * J, resJ: Drude polarisation current (3*ndrude) or NULL. */
void orc_maxwell_op(const orc_mesh *m, double rktime, const double *hn, const double *en,
                    double *reshn, double *resen, const double *J, double *resJ)
{
    maxwell_op_body(m, rktime, hn, en, reshn, resen, J, resJ, NULL, NULL, NULL, NULL, NULL, NULL);
}

/* The same with the Lorentz state L = (J, P) (6 blocks of nlorentz: J then
 * P, component-major) and its residual resL, or NULL. */
void orc_maxwell_op_ade(const orc_mesh *m, double rktime, const double *hn, const double *en,
                        double *reshn, double *resen, const double *J, double *resJ,
                        const double *L, double *resL)
{
    maxwell_op_body(m, rktime, hn, en, reshn, resen, J, resJ, L, resL, NULL, NULL, NULL, NULL);
}

/* The same with the PML state P = (pmlbn, pmldn): 6 blocks of npts, and
 * its residual resP = (respmlbn, respmldn), or NULL. */
void orc_maxwell_op_full(const orc_mesh *m, double rktime, const double *hn, const double *en,
                         double *reshn, double *resen, const double *J, double *resJ,
                         const double *L, double *resL, const double *P, double *resP)
{
    maxwell_op_body(m, rktime, hn, en, reshn, resen, J, resJ, L, resL, P, resP, NULL, NULL);
}

/* The same with the graphene state G (fjn slots 2..6 of the .usr: 15 blocks
 * of ngraph, block 3 (slot - 2) + component) and its residual resG, or NULL. */
void orc_maxwell_op_all(const orc_mesh *m, double rktime, const double *hn, const double *en,
                        double *reshn, double *resen, const double *J, double *resJ,
                        const double *L, double *resL, const double *P, double *resP,
                        const double *G, double *resG)
{
    maxwell_op_body(m, rktime, hn, en, reshn, resen, J, resJ, L, resL, P, resP, G, resG);
}

/* Face trace of H and E at face point i (restrict_to_face), with the
 * userinc plane wave added where it is on (cem_maxwell.F:497-498: after the
 * restriction, before the flux). */
static void face_trace(const orc_mesh *m, int64_t i, double rktime, const double *hn, const double *en,
                       double fh[3], double fe[3])
{
    int64_t npts = m->npts, v = m->cemface[i];
    for (int c = 0; c < 3; ++c) {
        fh[c] = hn[c * npts + v];
        fe[c] = en[c * npts + v];
    }
    if (m->inc_on && m->inc_on[i]) {
        double u = cos(m->inc_k[0] * m->xm1[v] + m->inc_k[1] * m->ym1[v] + m->inc_k[2] * m->zm1[v] -
                       m->inc_omega * rktime);
        for (int c = 0; c < 3; ++c) {
            fh[c] = fh[c] + m->inc_h[c] * u;
            fe[c] = fe[c] + m->inc_e[c] * u;
        }
    }
}

static void maxwell_op_body(const orc_mesh *m, double rktime, const double *hn, const double *en,
                            double *reshn, double *resen, const double *J, double *resJ,
                            const double *L, double *resL, const double *P, double *resP,
                            const double *G, double *resG)
{
    int64_t npts = m->npts, k = m->nxzfl;
    const double *H1 = hn, *H2 = hn + npts, *H3 = hn + 2 * npts;
    const double *E1 = en, *E2 = en + npts, *E3 = en + 2 * npts;
    double *rH1 = reshn, *rH2 = reshn + npts, *rH3 = reshn + 2 * npts;
    double *rE1 = resen, *rE2 = resen + npts, *rE3 = resen + 2 * npts;

    /* cem_maxwell, 3-D branch */
    wght_curl(m, H1, H2, H3, rE1, rE2, rE3);
    wght_curl(m, E1, E2, E3, rH1, rH2, rH3);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < npts; ++i) { rH1[i] = -rH1[i]; rH2[i] = -rH2[i]; rH3[i] = -rH3[i]; }

    /* restrict_to_face + flux3d loop 1:  srflx = (-n x E, -n x H) */
    double *s = m->srflx;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < k; ++i) {
        double fh[3], fe[3];
        face_trace(m, i, rktime, hn, en, fh, fe);
        double fh1 = fh[0], fh2 = fh[1], fh3 = fh[2];
        double fe1 = fe[0], fe2 = fe[1], fe3 = fe[2];
        double nx = m->unx[i], ny = m->uny[i], nz = m->unz[i];
        s[0 * k + i] = -ny * fe3 + nz * fe2;
        s[1 * k + i] = -nz * fe1 + nx * fe3;
        s[2 * k + i] = -nx * fe2 + ny * fe1;
        s[3 * k + i] = -ny * fh3 + nz * fh2;
        s[4 * k + i] = -nz * fh1 + nx * fh3;
        s[5 * k + i] = -nx * fh2 + ny * fh1;
    }
    /* userfsrc of tests/3dgraphene/3dgraphene.usr:235-262: the sheet
     * current (cem_3d_graphene_current, cem_maxwell.F:2827-2932) from the
     * stage's face traces and the sheet state, subtracted from -n x H on each
     * listed point before the face sum.  The state's RK update is the
     * caller's (same coefficients, same stage). */
    if (m->ngraph > 0 && G && resG) {
        int64_t ng = m->ngraph;
#pragma omp parallel for schedule(static)
        for (int64_t l = 0; l < ng; ++l) {
            int64_t j = m->gindex[l], v = m->cemface[j];
            const double *p = m->gpar + 12 * l;
            double fh[3], fe[3];
            face_trace(m, j, rktime, hn, en, fh, fe);
            double nx = m->unx[j], ny = m->uny[j], nz = m->unz[j];
            double nH[3] = {-ny * fh[2] + nz * fh[1], nx * fh[2] - nz * fh[0], -nx * fh[1] + ny * fh[0]};
            double ndotE = nx * fe[0] + ny * fe[1] + nz * fe[2];
            double nEn[3] = {fe[0] - nx * ndotE, fe[1] - ny * ndotE, fe[2] - nz * ndotE};
            double Yfac = 0.5 / m->Y0[j];
            double yc = sqrt(m->eps[v] / m->mu[v]);
            double cpfac = p[2] + p[7];
            double jnfac = 1.0 - cpfac * Yfac;
            for (int c = 0; c < 3; ++c) {
                double A0 = G[(0 + c) * ng + l], A1 = G[(3 + c) * ng + l], A2 = G[(6 + c) * ng + l];
                double A3 = G[(9 + c) * ng + l], A4 = G[(12 + c) * ng + l];
                double tmp = Yfac * (nH[c] + yc * nEn[c]);
                double J1 = (A0 + A1 + A3 - cpfac * tmp) / jnfac;
                double f = tmp - Yfac * J1;
                resG[(0 + c) * ng + l] = -p[0] * A0 + p[1] * f;
                resG[(3 + c) * ng + l] = A2 + p[5] * f;
                resG[(6 + c) * ng + l] = -p[3] * A1 - p[4] * A2 + p[6] * f;
                resG[(9 + c) * ng + l] = A4 + p[10] * f;
                resG[(12 + c) * ng + l] = -p[8] * A3 - p[9] * A4 + p[11] * f;
                s[(3 + c) * k + j] = s[(3 + c) * k + j] - J1;
            }
        }
    }
    /* gs_op_fields(gsh_face, srflx, nxzfl, 6, add) */
    orc_gs_add(m->partner, s, k, 6, k);
    /* cem_maxwell_flux_pec */
    for (int64_t j = 0; j < m->ncempec; ++j) {
        int64_t i = m->cempec[j];
        s[0 * k + i] = 2.0 * s[0 * k + i];
        s[1 * k + i] = 2.0 * s[1 * k + i];
        s[2 * k + i] = 2.0 * s[2 * k + i];
        s[3 * k + i] = 0;
        s[4 * k + i] = 0;
        s[5 * k + i] = 0;
    }
    /* flux3d loop 2 */
    double C0 = m->C0;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < k; ++i) {
        double Y0 = m->Y0[i], Y1 = m->Y1[i], Z0 = m->Z0[i], Z1 = m->Z1[i];
        double Y02 = -0.5 / Y0 * Y1;
        double Z02 = 0.5 / Z0 * Z1;
        double C02Y = 0.5 / Y0 * C0;
        double C02Z = 0.5 / Z0 * C0;
        double nx = m->unx[i], ny = m->uny[i], nz = m->unz[i];
        double fu1 = ny * s[5 * k + i] - nz * s[4 * k + i];
        double fu2 = nz * s[3 * k + i] - nx * s[5 * k + i];
        double fu3 = nx * s[4 * k + i] - ny * s[3 * k + i];
        double fw1 = ny * s[2 * k + i] - nz * s[1 * k + i];
        double fw2 = nz * s[0 * k + i] - nx * s[2 * k + i];
        double fw3 = nx * s[1 * k + i] - ny * s[0 * k + i];
        s[0 * k + i] = Y02 * s[0 * k + i] - C02Y * fu1;
        s[1 * k + i] = Y02 * s[1 * k + i] - C02Y * fu2;
        s[2 * k + i] = Y02 * s[2 * k + i] - C02Y * fu3;
        s[3 * k + i] = Z02 * s[3 * k + i] - C02Z * fw1;
        s[4 * k + i] = Z02 * s[4 * k + i] - C02Z * fw2;
        s[5 * k + i] = Z02 * s[5 * k + i] - C02Z * fw3;
    }
    /* add_flux_to_res (CPU branch, cem_maxwell.F:725-735).  Face nodes of
     * element e only touch volume nodes of e, so elements are independent. */
    int n2 = m->n1 * m->n1;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < m->nelt; ++e)
        for (int64_t j = e * 6 * n2; j < (e + 1) * 6 * n2; ++j) {
            int64_t i = m->cemface[j];
            double a = m->area[j];
            rH1[i] = rH1[i] + a * s[0 * k + j];
            rH2[i] = rH2[i] + a * s[1 * k + j];
            rH3[i] = rH3[i] + a * s[2 * k + j];
            rE1[i] = rE1[i] + a * s[3 * k + j];
            rE2[i] = rE2[i] + a * s[4 * k + j];
            rE3[i] = rE3[i] + a * s[5 * k + j];
        }
    /* pml_step (src/cem_maxwell_pml.F:508-592), after the flux and before
     * usersrc (cem_maxwell.F:502) */
    if (m->npml > 0 && P && resP) {
        int64_t n3 = (int64_t)m->n1 * m->n1 * m->n1;
        const double *B1 = P, *B2 = P + npts, *B3 = P + 2 * npts;
        const double *D1 = P + 3 * npts, *D2 = P + 4 * npts, *D3 = P + 5 * npts;
        double *rB1 = resP, *rB2 = resP + npts, *rB3 = resP + 2 * npts;
        double *rD1 = resP + 3 * npts, *rD2 = resP + 4 * npts, *rD3 = resP + 5 * npts;
#pragma omp parallel for schedule(static)
        for (int64_t l = 0; l < m->npml; ++l) {
            int64_t e = m->pmlptr[l];
            for (int64_t i = 0; i < n3; ++i) {
                int64_t j = e * n3 + i;
                double bm = m->bm1[j], bminv = 1.0 / bm;
                double sigx = m->pmlsigma[j], sigy = m->pmlsigma[npts + j], sigz = m->pmlsigma[2 * npts + j];
                double sxp = sigx / m->eps[j], syp = sigy / m->eps[j], szp = sigz / m->eps[j];
                double mu = m->mu[j];
                rB1[j] = rH1[j] * bminv - syp * B1[j];
                rB2[j] = rH2[j] * bminv - szp * B2[j];
                rB3[j] = rH3[j] * bminv - sxp * B3[j];
                rD1[j] = rE1[j] * bminv - syp * D1[j];
                rD2[j] = rE2[j] * bminv - szp * D2[j];
                rD3[j] = rE3[j] * bminv - sxp * D3[j];
                double ph1 = -syp * B1[j] + sxp * B1[j] - szp * mu * H1[j];
                double ph2 = -szp * B2[j] + syp * B2[j] - sxp * mu * H2[j];
                double ph3 = -sxp * B3[j] + szp * B3[j] - syp * mu * H3[j];
                double pe1 = -syp * D1[j] + sxp * D1[j] - sigz * E1[j];
                double pe2 = -szp * D2[j] + syp * D2[j] - sigx * E2[j];
                double pe3 = -sxp * D3[j] + szp * D3[j] - sigy * E3[j];
                rH1[j] = rH1[j] + ph1 * bm;
                rH2[j] = rH2[j] + ph2 * bm;
                rH3[j] = rH3[j] + ph3 * bm;
                rE1[j] = rE1[j] + pe1 * bm;
                rE2[j] = rE2[j] + pe2 * bm;
                rE3[j] = rE3[j] + pe3 * bm;
            }
        }
    }
    /* usersrc: source on E (bm1-weighted amplitude) with a Gaussian-pulse
     * or sinusoidal time factor, and the Drude current (cem_maxwell_drude,
     * cem_maxwell.F:3095-3147):
     *   resE -= J * bm1 ;  resJ = -a J + b E */
    if (m->src_amp) {
        double g = m->src_kind == 1
                       ? sin(m->src_omega * rktime + m->src_phase)
                       : exp(-((rktime - m->src_t0) / m->src_tw) * ((rktime - m->src_t0) / m->src_tw));
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < npts; ++i) {
            double a = m->src_amp[i] * g;
            rE1[i] += a * m->src_dir[0];
            rE2[i] += a * m->src_dir[1];
            rE3[i] += a * m->src_dir[2];
        }
    }
    if (m->ndrude > 0 && J && resJ) {
        int64_t nd = m->ndrude;
#pragma omp parallel for schedule(static)
        for (int64_t l = 0; l < nd; ++l) {
            int64_t i = m->dindex[l];
            double bm = m->bm1[i];
            rE1[i] = rE1[i] - J[l] * bm;
            rE2[i] = rE2[i] - J[nd + l] * bm;
            rE3[i] = rE3[i] - J[2 * nd + l] * bm;
            resJ[l] = -m->drude_a[l] * J[l] + m->drude_b[l] * E1[i];
            resJ[nd + l] = -m->drude_a[l] * J[nd + l] + m->drude_b[l] * E2[i];
            resJ[2 * nd + l] = -m->drude_a[l] * J[2 * nd + l] + m->drude_b[l] * E3[i];
        }
    }
    /* Lorentz (cem_maxwell_lorentz, cem_maxwell.F:3149-3211):
     *   resE -= J bm1 ;  resJ = -a J - b P + c E ;  resP = J */
    if (m->nlorentz > 0 && L && resL) {
        int64_t nl = m->nlorentz;
        const double *E[3] = {E1, E2, E3};
        double *rE[3] = {rE1, rE2, rE3};
#pragma omp parallel for schedule(static)
        for (int64_t l = 0; l < nl; ++l) {
            int64_t i = m->lindex[l];
            double a = m->lor_par[3 * l], b = m->lor_par[3 * l + 1], c = m->lor_par[3 * l + 2];
            double bm = m->bm1[i];
            for (int q = 0; q < 3; ++q) {
                double Jq = L[q * nl + l], Pq = L[(3 + q) * nl + l];
                rE[q][i] = rE[q][i] - Jq * bm;
                resL[q * nl + l] = -a * Jq - b * Pq + c * E[q][i];
                resL[(3 + q) * nl + l] = Jq;
            }
        }
    }
    /* cem_maxwell_invqmass */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < npts; ++i) {
        rH1[i] = rH1[i] * m->hbm1[i];
        rH2[i] = rH2[i] * m->hbm1[i];
        rH3[i] = rH3[i] * m->hbm1[i];
        rE1[i] = rE1[i] * m->ebm1[i];
        rE2[i] = rE2[i] * m->ebm1[i];
        rE3[i] = rE3[i] * m->ebm1[i];
    }
}

/* 2N-storage Carpenter-Kennedy RK45 constants (cem_common.F:86-104). */
static const double RK4A[5] = {
    0.0,
    -567301805773.0 / 1357537059087.0,
    -2404267990393.0 / 2016746695238.0,
    -3550918686646.0 / 2091501179385.0,
    -1275806237668.0 / 842570457699.0};
static const double RK4B[5] = {
    1432997174477.0 / 9575080441755.0,
    5161836677717.0 / 13612068292357.0,
    1720146321549.0 / 2090206949498.0,
    3134564353537.0 / 4481467310338.0,
    2277821191437.0 / 14882151754819.0};
static const double RK4C[6] = {
    0.0,
    1432997174477.0 / 9575080441755.0,
    2526269341429.0 / 6820363962896.0,
    2006345519317.0 / 3224310063776.0,
    2802321613138.0 / 2924317926251.0,
    1.0};

void orc_rk_coeffs(double *a, double *b, double *c)
{
    for (int i = 0; i < 5; ++i) { a[i] = RK4A[i]; b[i] = RK4B[i]; c[i] = RK4C[i]; }
    c[5] = RK4C[5];
}

/* rk4_upd (cem_common.F:18-76): kh = ca*kh + dt*resh; h = h + cb*kh */
static void rk4_upd(double *h, double *kh, const double *resh, double cb, double ca, double dt, int64_t n)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        kh[i] = ca * kh[i] + dt * resh[i];
        h[i] = h[i] + cb * kh[i];
    }
}

/* nsteps of cem_maxwell_op_rk with the Lorentz state L/kL/resL (6*nlorentz)
 * and the PML state P/kP/resP (6*npts; rk_maxwell_ab's PML updates,
 * cem_maxwell.F:1942-1954) as well (NULL: none); otherwise as orc_rk_run. */
void orc_rk_run_all(const orc_mesh *m, double *hn, double *en, double *khn, double *ken,
                    double *reshn, double *resen, double *J, double *kJ, double *resJ,
                    double *L, double *kL, double *resL, double *P, double *kP, double *resP,
                    double *G, double *kG, double *resG, double dt, int64_t nsteps, double *time);

void orc_rk_run_full(const orc_mesh *m, double *hn, double *en, double *khn, double *ken,
                     double *reshn, double *resen, double *J, double *kJ, double *resJ,
                     double *L, double *kL, double *resL, double *P, double *kP, double *resP,
                     double dt, int64_t nsteps, double *time)
{
    orc_rk_run_all(m, hn, en, khn, ken, reshn, resen, J, kJ, resJ, L, kL, resL, P, kP, resP, NULL, NULL, NULL,
                   dt, nsteps, time);
}

/* The same with the graphene state G/kG/resG (15*ngraph; the four rk4_upd
 * calls of cem_3d_graphene_current, cem_maxwell.F:2915-2929, at the stage's
 * coefficients) as well. */
void orc_rk_run_all(const orc_mesh *m, double *hn, double *en, double *khn, double *ken,
                    double *reshn, double *resen, double *J, double *kJ, double *resJ,
                    double *L, double *kL, double *resL, double *P, double *kP, double *resP,
                    double *G, double *kG, double *resG, double dt, int64_t nsteps, double *time)
{
    int64_t n3 = 3 * m->npts;
    for (int64_t st = 0; st < nsteps; ++st) {
        for (int r = 0; r < 5; ++r) {
            double rktime = *time + dt * RK4C[r];
            maxwell_op_body(m, rktime, hn, en, reshn, resen, J, resJ, L, resL, P, resP, G, resG);
            rk4_upd(hn, khn, reshn, RK4B[r], RK4A[r], dt, n3);
            rk4_upd(en, ken, resen, RK4B[r], RK4A[r], dt, n3);
            if (m->ndrude > 0 && J) rk4_upd(J, kJ, resJ, RK4B[r], RK4A[r], dt, 3 * m->ndrude);
            if (m->nlorentz > 0 && L) rk4_upd(L, kL, resL, RK4B[r], RK4A[r], dt, 6 * m->nlorentz);
            if (m->npml > 0 && P) rk4_upd(P, kP, resP, RK4B[r], RK4A[r], dt, 6 * m->npts);
            if (m->ngraph > 0 && G) rk4_upd(G, kG, resG, RK4B[r], RK4A[r], dt, 15 * m->ngraph);
        }
        *time = *time + dt;
    }
}

void orc_rk_run_ade(const orc_mesh *m, double *hn, double *en, double *khn, double *ken,
                    double *reshn, double *resen, double *J, double *kJ, double *resJ,
                    double *L, double *kL, double *resL, double dt, int64_t nsteps, double *time)
{
    orc_rk_run_full(m, hn, en, khn, ken, reshn, resen, J, kJ, resJ, L, kL, resL, NULL, NULL, NULL, dt, nsteps,
                    time);
}

/* nsteps of cem_maxwell_op_rk.  u arrays hold 3 components at stride npts.
 * J/kJ/resJ: Drude state (3*ndrude) or NULL.  *time is advanced. */
void orc_rk_run(const orc_mesh *m, double *hn, double *en, double *khn, double *ken,
                double *reshn, double *resen, double *J, double *kJ, double *resJ,
                double dt, int64_t nsteps, double *time)
{
    int64_t n3 = 3 * m->npts;
    for (int64_t st = 0; st < nsteps; ++st) {
        for (int r = 0; r < 5; ++r) {
            double rktime = *time + dt * RK4C[r];
            orc_maxwell_op(m, rktime, hn, en, reshn, resen, J, resJ);
            rk4_upd(hn, khn, reshn, RK4B[r], RK4A[r], dt, n3);
            rk4_upd(en, ken, resen, RK4B[r], RK4A[r], dt, n3);
            if (m->ndrude > 0 && J) rk4_upd(J, kJ, resJ, RK4B[r], RK4A[r], dt, 3 * m->ndrude);
        }
        *time = *time + dt;
    }
}

/* cem_error (cem_common.F:1335-1355) with volvm1 = sum(bm1). */
void orc_cem_error(const double *u, const double *exact, const double *bm1, int64_t n,
                   double *l2, double *linf)
{
    double s = 0.0, v = 0.0, mx = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double err = exact[i] - u[i];
        s += err * bm1[i] * err;
        v += bm1[i];
        if (fabs(err) > mx) mx = fabs(err);
    }
    double r = s / v;
    if (r > 0.0) r = sqrt(r);
    *l2 = r;
    *linf = mx;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------------ */
/* Vector-Jacobian product of one residual evaluation (adjoint checker).     */
/*                                                                           */
/* r(u; eps, mu) = cem_maxwell_op above (no Drude).  Given kt = (kH, kE),    */
/*   aH, aE    += (dr/du)^T kt                                               */
/*   geps, gmu += (dr/deps)^T kt, (dr/dmu)^T kt      (per volume node)       */
/* Each forward step is transposed in reverse order: inverse mass, lift,     */
/* flux loop 2, PEC, gs (symmetric), loop 1 / restriction, and the weighted  */
/* curl through the transpose of local_grad3; the eps/mu dependence runs     */
/* through ebm1/hbm1 and through Z_0/Z_1/Y_0/Y_1 of cem_maxwell_materials.   */
/* ------------------------------------------------------------------------ */

/* u(l,jk) += sum_i D(i,l) ur(i,jk); u(i,l,k) += sum_j D(j,l) us(i,j,k);
 * u(ij,l) += sum_k D(k,l) ut(ij,k) */
static void local_grad3_t(int n1, const double *D, const double *ur, const double *us,
                          const double *ut, double *u)
{
    int n2 = n1 * n1;
    for (int jk = 0; jk < n2; ++jk)
        for (int l = 0; l < n1; ++l) {
            double s = 0.0;
            for (int i = 0; i < n1; ++i) s += D[i + n1 * l] * ur[i + n1 * jk];
            u[l + n1 * jk] += s;
        }
    for (int k = 0; k < n1; ++k)
        for (int l = 0; l < n1; ++l)
            for (int i = 0; i < n1; ++i) {
                double s = 0.0;
                for (int j = 0; j < n1; ++j) s += D[j + n1 * l] * us[i + n1 * j + n2 * k];
                u[i + n1 * l + n2 * k] += s;
            }
    for (int l = 0; l < n1; ++l)
        for (int ij = 0; ij < n2; ++ij) {
            double s = 0.0;
            for (int k = 0; k < n1; ++k) s += D[k + n1 * l] * ut[ij + n2 * k];
            u[ij + n2 * l] += s;
        }
}

/* a1,a2,a3 += W^T (v1,v2,v3) for the weighted curl W of maxwell_wght_curl */
static void wght_curl_t(const orc_mesh *m, const double *v1, const double *v2, const double *v3,
                        double sign, double *a1, double *a2, double *a3)
{
    int n1 = m->n1, n3 = n1 * n1 * n1;
#pragma omp parallel
    {
        double *b = (double *)malloc(sizeof(double) * 9 * n3);
        double *r1 = b, *s1 = b + n3, *t1 = b + 2 * n3;
        double *r2 = b + 3 * n3, *s2 = b + 4 * n3, *t2 = b + 5 * n3;
        double *r3 = b + 6 * n3, *s3 = b + 7 * n3, *t3 = b + 8 * n3;
#pragma omp for schedule(static)
        for (int64_t e = 0; e < m->nelt; ++e) {
            int64_t o = (int64_t)n3 * e;
            for (int i = 0; i < n3; ++i) {
                int64_t k = o + i;
                double w = m->w3[i] * sign;
                double q1 = w * v1[k], q2 = w * v2[k], q3 = w * v3[k];
                r1[i] = m->rz[k] * q2 - m->ry[k] * q3;
                s1[i] = m->sz[k] * q2 - m->sy[k] * q3;
                t1[i] = m->tz[k] * q2 - m->ty[k] * q3;
                r2[i] = m->rx[k] * q3 - m->rz[k] * q1;
                s2[i] = m->sx[k] * q3 - m->sz[k] * q1;
                t2[i] = m->tx[k] * q3 - m->tz[k] * q1;
                r3[i] = m->ry[k] * q1 - m->rx[k] * q2;
                s3[i] = m->sy[k] * q1 - m->sx[k] * q2;
                t3[i] = m->ty[k] * q1 - m->tx[k] * q2;
            }
            local_grad3_t(n1, m->D, r1, s1, t1, a1 + o);
            local_grad3_t(n1, m->D, r2, s2, t2, a2 + o);
            local_grad3_t(n1, m->D, r3, s3, t3, a3 + o);
        }
        free(b);
    }
}

/* n x v */
static inline void crossn(double nx, double ny, double nz, const double *v, double *o)
{
    o[0] = ny * v[2] - nz * v[1];
    o[1] = nz * v[0] - nx * v[2];
    o[2] = nx * v[1] - ny * v[0];
}

void orc_maxwell_vjp(const orc_mesh *m, double rktime, const double *eps, const double *mu,
                     const int8_t *pec_face, const double *hn, const double *en,
                     const double *kH, const double *kE, double *aH, double *aE,
                     double *geps, double *gmu)
{
    int64_t npts = m->npts, k = m->nxzfl;
    int n2 = m->n1 * m->n1;
    /* forward residual at u (for the inverse-mass terms) and the face jumps */
    double *rH = (double *)malloc(sizeof(double) * 3 * npts);
    double *rE = (double *)malloc(sizeof(double) * 3 * npts);
    orc_maxwell_op(m, rktime, hn, en, rH, rE, NULL, NULL);
    /* jumps after gs + PEC: recompute loop 1, gs, PEC (srflx holds loop-2 output) */
    double *s = (double *)malloc(sizeof(double) * 6 * k);
    {
        const double *H1 = hn, *H2 = hn + npts, *H3 = hn + 2 * npts;
        const double *E1 = en, *E2 = en + npts, *E3 = en + 2 * npts;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < k; ++i) {
            int64_t v = m->cemface[i];
            double nx = m->unx[i], ny = m->uny[i], nz = m->unz[i];
            s[0 * k + i] = -ny * E3[v] + nz * E2[v];
            s[1 * k + i] = -nz * E1[v] + nx * E3[v];
            s[2 * k + i] = -nx * E2[v] + ny * E1[v];
            s[3 * k + i] = -ny * H3[v] + nz * H2[v];
            s[4 * k + i] = -nz * H1[v] + nx * H3[v];
            s[5 * k + i] = -nx * H2[v] + ny * H1[v];
        }
        orc_gs_add(m->partner, s, k, 6, k);
        for (int64_t j = 0; j < m->ncempec; ++j) {
            int64_t i = m->cempec[j];
            for (int c = 0; c < 3; ++c) { s[c * k + i] *= 2.0; s[(3 + c) * k + i] = 0.0; }
        }
    }
    /* v = P kt ; inverse-mass gradients: d(ebm1 R)/deps = -r/eps */
    double *vH = (double *)malloc(sizeof(double) * 3 * npts);
    double *vE = (double *)malloc(sizeof(double) * 3 * npts);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < npts; ++i) {
        double dE = 0.0, dH = 0.0;
        for (int c = 0; c < 3; ++c) {
            vH[c * npts + i] = m->hbm1[i] * kH[c * npts + i];
            vE[c * npts + i] = m->ebm1[i] * kE[c * npts + i];
            dE += kE[c * npts + i] * rE[c * npts + i];
            dH += kH[c * npts + i] * rH[c * npts + i];
        }
        geps[i] -= dE / eps[i];
        gmu[i] -= dH / mu[i];
    }
    /* volume: R_E = W(H), R_H = -W(E) */
    wght_curl_t(m, vE, vE + npts, vE + 2 * npts, 1.0, aH, aH + npts, aH + 2 * npts);
    wght_curl_t(m, vH, vH + npts, vH + 2 * npts, -1.0, aE, aE + npts, aE + 2 * npts);
    /* lift^T and flux loop 2^T; coefficient gradients */
    double *sb = (double *)malloc(sizeof(double) * 6 * k);
    double *Yb0 = (double *)malloc(sizeof(double) * k), *Yb1 = (double *)malloc(sizeof(double) * k);
    double *Zb0 = (double *)malloc(sizeof(double) * k), *Zb1 = (double *)malloc(sizeof(double) * k);
    double C0 = m->C0;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < k; ++i) {
        int64_t v = m->cemface[i];
        double a = m->area[i];
        double nx = m->unx[i], ny = m->uny[i], nz = m->unz[i];
        double oE[3], oH[3], sE[3], sH[3], t[3], u2[3];
        for (int c = 0; c < 3; ++c) {
            oE[c] = a * vH[c * npts + v];      /* slots 0..2 lift into resH */
            oH[c] = a * vE[c * npts + v];      /* slots 3..5 lift into resE */
            sE[c] = s[c * k + i];
            sH[c] = s[(3 + c) * k + i];
        }
        double Y0 = m->Y0[i], Y1 = m->Y1[i], Z0 = m->Z0[i], Z1 = m->Z1[i];
        double Y02 = -0.5 / Y0 * Y1, Z02 = 0.5 / Z0 * Z1, C02Y = 0.5 / Y0 * C0, C02Z = 0.5 / Z0 * C0;
        /* s_E-bar = Y02 oE + C02Z n x oH ; s_H-bar = C02Y n x oE + Z02 oH */
        crossn(nx, ny, nz, oH, t);
        crossn(nx, ny, nz, oE, u2);
        for (int c = 0; c < 3; ++c) {
            sb[c * k + i] = Y02 * oE[c] + C02Z * t[c];
            sb[(3 + c) * k + i] = C02Y * u2[c] + Z02 * oH[c];
        }
        /* coefficient gradients */
        double nsH[3], nsE[3];
        crossn(nx, ny, nz, sH, nsH);
        crossn(nx, ny, nz, sE, nsE);
        double gY02 = 0, gC02Y = 0, gZ02 = 0, gC02Z = 0;
        for (int c = 0; c < 3; ++c) {
            gY02 += oE[c] * sE[c];
            gC02Y -= oE[c] * nsH[c];
            gZ02 += oH[c] * sH[c];
            gC02Z -= oH[c] * nsE[c];
        }
        Yb0[i] = gY02 * (0.5 * Y1 / (Y0 * Y0)) + gC02Y * (-0.5 * C0 / (Y0 * Y0));
        Yb1[i] = gY02 * (-0.5 / Y0);
        Zb0[i] = gZ02 * (-0.5 * Z1 / (Z0 * Z0)) + gC02Z * (-0.5 * C0 / (Z0 * Z0));
        Zb1[i] = gZ02 * (0.5 / Z0);
    }
    /* PEC^T: slots 0..2 doubled, 3..5 zeroed */
    for (int64_t j = 0; j < m->ncempec; ++j) {
        int64_t i = m->cempec[j];
        for (int c = 0; c < 3; ++c) { sb[c * k + i] *= 2.0; sb[(3 + c) * k + i] = 0.0; }
    }
    /* gs^T = gs */
    orc_gs_add(m->partner, sb, k, 6, k);
    /* loop 1^T: s_E = -n x E  ->  E-bar += n x s_E-bar (same for H); elements independent */
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < m->nelt; ++e)
        for (int64_t i = e * 6 * n2; i < (e + 1) * 6 * n2; ++i) {
            int64_t v = m->cemface[i];
            double nx = m->unx[i], ny = m->uny[i], nz = m->unz[i];
            double bE[3] = {sb[0 * k + i], sb[1 * k + i], sb[2 * k + i]};
            double bH[3] = {sb[3 * k + i], sb[4 * k + i], sb[5 * k + i]};
            double tE[3], tH[3];
            crossn(nx, ny, nz, bE, tE);
            crossn(nx, ny, nz, bH, tH);
            for (int c = 0; c < 3; ++c) {
                aE[c * npts + v] += tE[c];
                aH[c * npts + v] += tH[c];
            }
        }
    /* materials^T: Z0'' = S G zimp, Z_1 = Z0'' - zimp, Z_0 = 0.5 Z0'' */
    double *zb = (double *)malloc(sizeof(double) * k), *yb = (double *)malloc(sizeof(double) * k);
    for (int64_t i = 0; i < k; ++i) {
        zb[i] = 0.5 * Zb0[i] + Zb1[i];
        yb[i] = 0.5 * Yb0[i] + Yb1[i];
    }
    for (int64_t e = 0; e < m->nelt; ++e)
        for (int f = 0; f < 6; ++f)
            if (pec_face[6 * e + f])
                for (int j = 0; j < n2; ++j) {
                    int64_t i = e * 6 * n2 + (int64_t)f * n2 + j;
                    zb[i] *= 2.0;
                    yb[i] *= 2.0;
                }
    orc_gs_add(m->partner, zb, k, 1, k);
    orc_gs_add(m->partner, yb, k, 1, k);
    for (int64_t i = 0; i < k; ++i) {
        int64_t v = m->cemface[i];
        double zimp = sqrt(mu[v] / eps[v]), ycon = sqrt(eps[v] / mu[v]);
        double zbar = zb[i] - Zb1[i], ybar = yb[i] - Yb1[i];
        geps[v] += zbar * (-0.5 * zimp / eps[v]) + ybar * (0.5 * ycon / eps[v]);
        gmu[v] += zbar * (0.5 * zimp / mu[v]) + ybar * (-0.5 * ycon / mu[v]);
    }
    free(rH); free(rE); free(s); free(vH); free(vE); free(sb);
    free(Yb0); free(Yb1); free(Zb0); free(Zb1); free(zb); free(yb);
}

/* Drude part of the vector-Jacobian product (cem_maxwell_drude,
 * cem_maxwell.F:3095-3147).  At the listed nodes i = dindex[l] the residual
 * holds r_E += -ebm1 bm1 J and r_J = -a J + b E.  Given kt = (kE, kJ):
 *   aE[i] += b kJ ;  aJ[l] += -ebm1 bm1 kE[i] - a kJ
 *   ga[l] += -kJ.J ;  gb[l] += kJ.E[i]
 *   geps[i] += kE.(ebm1 bm1 J)/eps[i]   (the eps-derivative of -J/eps)
 * J, kJ, aJ: 3 blocks of ndrude; en, kE, aE: 3 blocks of npts. */
void orc_drude_vjp(const orc_mesh *m, const double *eps, const double *en, const double *J,
                   const double *kE, const double *kJ, double *aE, double *aJ, double *geps,
                   double *ga, double *gb)
{
    int64_t nd = m->ndrude, npts = m->npts;
    for (int64_t l = 0; l < nd; ++l) {
        int64_t i = m->dindex[l];
        double s = m->ebm1[i] * m->bm1[i];
        double pa = m->drude_a[l], pb = m->drude_b[l];
        double dJ = 0.0, dE = 0.0, dP = 0.0;
        for (int c = 0; c < 3; ++c) {
            double kj = kJ[c * nd + l], ke = kE[c * npts + i];
            aE[c * npts + i] += pb * kj;
            aJ[c * nd + l] += -s * ke - pa * kj;
            dJ += kj * J[c * nd + l];
            dE += kj * en[c * npts + i];
            dP += ke * s * J[c * nd + l];
        }
        ga[l] -= dJ;
        gb[l] += dE;
        geps[i] += dP / eps[i];
    }
}
