// mesh_tables.hpp -- host-side compact tables the stage kernel reads.
//
// Producers of a HostMesh:
//   compact_box()        from the reference-layout arrays of nkc_desc (what
//                        NekCEM's cem_maxwell_init leaves in COMMON), when
//                        the mesh is made of axis-aligned affine hexahedra
//                        paired face to face;
//   build_box()          directly from a box description (nkc_box_desc), the
//                        genbox2 semantics of src/nek5_genbox.F:2-145 without
//                        materialising per-node arrays;
//   build_hex()          from element corners and boundary codes (.rea);
//   compact_general()    from the reference-layout arrays of any conforming
//                        hexahedral mesh (curved, rotated, sheared elements).
// compact_reference() tries compact_box() and falls back to
// compact_general().  The box tables hold, per element: g = (rx, sy, tz)/J,
// the six lift factors area/(J w3), the six neighbours (or PEC / unpaired),
// the upwind/central flux coefficients (Y02, C02Y, Z02, C02Z) of
// cem_maxwell_flux3d (src/cem_maxwell.F:970-975) from the impedance averages
// of cem_maxwell_materials (src/cem_maxwell.F:262-325), and 1/eps, 1/mu.
// The general tables hold the same per node (nine cofactor/J quotients,
// bm1, 1/eps, 1/mu) and per face node (unit normal, area/bm1, the partner's
// volume node, the four flux coefficients).  pml_tables() is the PML set-up
// of src/cem_maxwell_pml.F on the box tables.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>
#include <cstdio>
#include <tuple>
#include <map>

#include "../../include/nekcem_hip.h"

namespace nkc {

// FaceType codes live in stage_kernel.hpp: -1 PEC, -2 no partner.
constexpr int32_t HOST_FACE_PEC = -1;
constexpr int32_t HOST_FACE_NONE = -2;

struct HostMesh {
    int n1 = 0;
    int64_t nelt = 0;
    int mat = 0;                     // 0 per element, 1 per node
    std::vector<double> D;           // n1*n1, D(i,l) at [i + n1 l]
    std::vector<double> w;           // GLL weights
    std::vector<double> geo;         // nelt*9
    std::vector<int32_t> nbr;        // nelt*6 (kernel face order)
    std::vector<double> fco;         // nelt*6*4 or nelt*6*n2*4
    std::vector<double> ieps, imu;   // nelt or npts
    std::vector<double> jac;         // nelt (element Jacobian); per node on a general mesh
    // General (curved, rotated or sheared) hexahedra, read by the
    // per-node-geometry kernel: the nine metric quotients G_ab = (cofactor
    // ab)/J per node (rx, ry, rz, sx, sy, sz, tx, ty, tz over J, component
    // major), per face node in the reference's face order the unit normal
    // and the lift factor area/(J w3) of its volume node (4 doubles), the
    // partner's volume node (0-based; HOST_FACE_PEC / HOST_FACE_NONE), and
    // bm1 = J w3 per node.  fco then holds each face node's flux
    // coefficients in the reference's face order, ieps / imu per node.
    bool general = false;
    std::vector<double> gmet, fgeo, bmw;
    std::vector<int32_t> fnb;
    std::string err;
};

// Ed's face (1..6) of kernel face kf (x-, x+, y-, y+, z-, z+).
inline const int *ed_of_kf() { static const int t[6] = {4, 2, 1, 3, 5, 6}; return t; }
inline const int *kf_of_ed() { static const int t[7] = {-1, 2, 1, 3, 0, 4, 5}; return t; }

// Volume node (in-element index) of in-face node q = a + n1*b on kernel face
// kf, in the skpdat order of src/nek5_connect11.F:1483-1528.
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

inline bool near(double x, double y, double scale, double tol) { return std::fabs(x - y) <= tol * scale; }

inline void flux_coef(double Y0, double Y1, double Z0, double Z1, double C0, double *c) {
    c[0] = -0.5 / Y0 * Y1;
    c[1] = 0.5 / Y0 * C0;
    c[2] = 0.5 / Z0 * Z1;
    c[3] = 0.5 / Z0 * C0;
}

// ---------------------------------------------------------------------------
// from the reference arrays
// ---------------------------------------------------------------------------
inline int32_t compact_box(const nkc_desc *d, HostMesh &m) {
    const int n1 = d->N + 1, n2 = n1 * n1, n3 = n2 * n1;
    const int64_t nelt = d->nelt, npts = (int64_t)n3 * nelt, nxzfl = (int64_t)6 * n2 * nelt;
    const double C0 = d->ifupwind ? 1.0 : 0.0;
    const double tol = 1e-8;
    char buf[320];
    auto bad = [&](int32_t code, const char *msg) { m.err = msg; return code; };

    std::vector<int8_t> pecslot(nxzfl, 0);
    for (int64_t l = 0; l < d->ncempec; ++l) {
        int64_t s = (int64_t)d->cempec[l] - 1;
        if (s < 0 || s >= nxzfl) {
            snprintf(buf, sizeof buf, "cempec[%lld]=%d out of range", (long long)l, d->cempec[l]);
            return bad(NKC_ERR_ARG, buf);
        }
        pecslot[s] = 1;
    }
    bool elem_mat = true;
    for (int64_t e = 0; e < nelt && elem_mat; ++e) {
        const int64_t b = e * n3;
        for (int q = 1; q < n3; ++q)
            if (d->permittivity[b + q] != d->permittivity[b] || d->permeability[b + q] != d->permeability[b]) {
                elem_mat = false;
                break;
            }
        for (int f = 0; f < 6 && elem_mat; ++f) {
            const int64_t s0 = e * 6 * n2 + (int64_t)f * n2;
            for (int q = 1; q < n2; ++q)
                if (d->Y_0[s0 + q] != d->Y_0[s0] || d->Y_1[s0 + q] != d->Y_1[s0] ||
                    d->Z_0[s0 + q] != d->Z_0[s0] || d->Z_1[s0 + q] != d->Z_1[s0]) {
                    elem_mat = false;
                    break;
                }
        }
    }
    m.n1 = n1;
    m.nelt = nelt;
    m.mat = elem_mat ? 0 : 1;
    m.D.assign(d->dxm1, d->dxm1 + n2);
    m.w.assign(d->wxm1, d->wxm1 + n1);
    m.geo.assign(nelt * 9, 0.0);
    m.nbr.assign(nelt * 6, 0);
    m.fco.assign(elem_mat ? nelt * 6 * 4 : nelt * 6 * n2 * 4, 0.0);
    m.ieps.assign(elem_mat ? nelt : npts, 0.0);
    m.imu.assign(elem_mat ? nelt : npts, 0.0);
    m.jac.assign(nelt, 0.0);

    for (int64_t e = 0; e < nelt; ++e) {
        const int64_t b = e * n3;
        const double J0 = d->jacm[b], rx0 = d->rxmn[b], sy0 = d->symn[b], tz0 = d->tzmn[b];
        const double sc = std::fmax(std::fabs(rx0), std::fmax(std::fabs(sy0), std::fabs(tz0)));
        if (!(J0 > 0.0)) {
            snprintf(buf, sizeof buf, "element %lld: Jacobian %g not positive", (long long)e, J0);
            return bad(NKC_ERR_ARG, buf);
        }
        for (int q = 0; q < n3; ++q) {
            const int64_t p = b + q;
            if (!near(d->jacm[p], J0, std::fabs(J0), tol) || !near(d->rxmn[p], rx0, sc, tol) ||
                !near(d->symn[p], sy0, sc, tol) || !near(d->tzmn[p], tz0, sc, tol) ||
                std::fabs(d->rymn[p]) > tol * sc || std::fabs(d->rzmn[p]) > tol * sc ||
                std::fabs(d->sxmn[p]) > tol * sc || std::fabs(d->szmn[p]) > tol * sc ||
                std::fabs(d->txmn[p]) > tol * sc || std::fabs(d->tymn[p]) > tol * sc) {
                snprintf(buf, sizeof buf,
                         "element %lld is not an axis-aligned affine hexahedron; only box meshes are built",
                         (long long)e);
                return bad(NKC_ERR_UNSUPPORTED, buf);
            }
        }
        m.jac[e] = J0;
        m.geo[e * 9 + 0] = rx0 / J0;
        m.geo[e * 9 + 1] = sy0 / J0;
        m.geo[e * 9 + 2] = tz0 / J0;
        for (int kf = 0; kf < 6; ++kf) {
            const int ef = ed_of_kf()[kf];
            const int ax = kf / 2;
            const double sg = (kf & 1) ? 1.0 : -1.0;
            const int64_t s0 = e * 6 * n2 + (int64_t)(ef - 1) * n2;
            double L0 = 0.0;
            for (int q = 0; q < n2; ++q) {
                const int64_t s = s0 + q;
                const int64_t v = (int64_t)d->cemface[s] - 1;
                if (v != b + face_node(n1, kf, q)) {
                    snprintf(buf, sizeof buf, "cemface[%lld]=%d does not follow the skpdat face order",
                             (long long)s, d->cemface[s]);
                    return bad(NKC_ERR_ARG, buf);
                }
                const double nv[3] = {d->unxm[s], d->unym[s], d->unzm[s]};
                for (int c = 0; c < 3; ++c)
                    if (std::fabs(nv[c] - (c == ax ? sg : 0.0)) > tol) {
                        snprintf(buf, sizeof buf, "element %lld face %d: normal is not an axis direction",
                                 (long long)e, ef);
                        return bad(NKC_ERR_UNSUPPORTED, buf);
                    }
                const double L = d->aream[s] / d->bmn[v];
                if (q == 0) L0 = L;
                else if (!near(L, L0, std::fabs(L0), tol)) {
                    snprintf(buf, sizeof buf, "element %lld face %d: lift factor varies over the face",
                             (long long)e, ef);
                    return bad(NKC_ERR_UNSUPPORTED, buf);
                }
            }
            m.geo[e * 9 + 3 + kf] = L0;
            const int64_t p0 = (int64_t)d->face_partner[s0] - 1;
            int32_t nb;
            if (p0 < 0) {
                nb = pecslot[s0] ? HOST_FACE_PEC : HOST_FACE_NONE;
                for (int q = 0; q < n2; ++q)
                    if (d->face_partner[s0 + q] != 0 || pecslot[s0 + q] != pecslot[s0]) {
                        snprintf(buf, sizeof buf, "element %lld face %d: mixed boundary face", (long long)e, ef);
                        return bad(NKC_ERR_UNSUPPORTED, buf);
                    }
            } else {
                if (p0 >= nxzfl) return bad(NKC_ERR_ARG, "face_partner out of range");
                const int64_t ne = p0 / (6 * n2);
                const int nef = (int)((p0 / n2) % 6) + 1;
                if (kf_of_ed()[nef] != (kf ^ 1)) {
                    snprintf(buf, sizeof buf,
                             "element %lld face %d pairs with face %d of element %lld; only opposite-face "
                             "box pairing is built", (long long)e, ef, nef, (long long)ne);
                    return bad(NKC_ERR_UNSUPPORTED, buf);
                }
                const int64_t ns0 = ne * 6 * n2 + (int64_t)(nef - 1) * n2;
                for (int q = 0; q < n2; ++q)
                    if ((int64_t)d->face_partner[s0 + q] - 1 != ns0 + q || pecslot[s0 + q]) {
                        snprintf(buf, sizeof buf, "element %lld face %d: rotated face pairing not built",
                                 (long long)e, ef);
                        return bad(NKC_ERR_UNSUPPORTED, buf);
                    }
                nb = (int32_t)ne;
            }
            m.nbr[e * 6 + kf] = nb;
            const int nq = elem_mat ? 1 : n2;
            for (int q = 0; q < nq; ++q) {
                const int64_t s = s0 + q;
                double *c = elem_mat ? &m.fco[(e * 6 + kf) * 4] : &m.fco[((e * 6 + kf) * n2 + q) * 4];
                flux_coef(d->Y_0[s], d->Y_1[s], d->Z_0[s], d->Z_1[s], C0, c);
            }
        }
        if (elem_mat) {
            m.ieps[e] = 1.0 / d->permittivity[b];
            m.imu[e] = 1.0 / d->permeability[b];
        }
    }
    if (!elem_mat)
        for (int64_t p = 0; p < npts; ++p) {
            m.ieps[p] = 1.0 / d->permittivity[p];
            m.imu[p] = 1.0 / d->permeability[p];
        }
    return NKC_OK;
}

// ---------------------------------------------------------------------------
// General hexahedra from the reference arrays: any element geometry the
// reference's GLMAPM1 / AREA3 produce (curved, rotated, sheared) and any
// conforming face pairing gsh_face describes (face_partner may match a face
// node to any node of the facing face).  Per node the metric quotients and
// bm1, per face node (reference face order) the normal, the lift factor
// area/bm1 of its volume node, the partner's volume node and the flux
// coefficients of cem_maxwell_flux3d from the Y / Z arrays.
// ---------------------------------------------------------------------------
inline int32_t compact_general(const nkc_desc *d, HostMesh &m) {
    const int n1 = d->N + 1, n2 = n1 * n1, n3 = n2 * n1;
    const int64_t nelt = d->nelt, npts = (int64_t)n3 * nelt, nxzfl = (int64_t)6 * n2 * nelt;
    const double C0 = d->ifupwind ? 1.0 : 0.0;
    char buf[320];
    auto bad = [&](int32_t code, const char *msg) { m.err = msg; return code; };
    std::vector<int8_t> pecslot(nxzfl, 0);
    for (int64_t l = 0; l < d->ncempec; ++l) {
        const int64_t s = (int64_t)d->cempec[l] - 1;
        if (s < 0 || s >= nxzfl) {
            snprintf(buf, sizeof buf, "cempec[%lld]=%d out of range", (long long)l, d->cempec[l]);
            return bad(NKC_ERR_ARG, buf);
        }
        pecslot[s] = 1;
    }
    m.n1 = n1;
    m.nelt = nelt;
    m.mat = 1;
    m.general = true;
    m.D.assign(d->dxm1, d->dxm1 + n2);
    m.w.assign(d->wxm1, d->wxm1 + n1);
    m.geo.assign(nelt * 9, 0.0);
    m.nbr.assign(nelt * 6, HOST_FACE_NONE);
    m.gmet.assign((size_t)9 * npts, 0.0);
    m.bmw.assign(npts, 0.0);
    m.jac.assign(npts, 0.0);
    m.ieps.assign(npts, 0.0);
    m.imu.assign(npts, 0.0);
    const double *cof[9] = {d->rxmn, d->rymn, d->rzmn, d->sxmn, d->symn, d->szmn, d->txmn, d->tymn, d->tzmn};
    for (int64_t p = 0; p < npts; ++p) {
        const double J = d->jacm[p];
        if (!(J > 0.0) || !(d->bmn[p] > 0.0)) {
            snprintf(buf, sizeof buf, "node %lld: Jacobian %g or mass %g not positive", (long long)p, J, d->bmn[p]);
            return bad(NKC_ERR_ARG, buf);
        }
        for (int c = 0; c < 9; ++c) m.gmet[(size_t)c * npts + p] = cof[c][p] / J;
        m.bmw[p] = d->bmn[p];
        m.jac[p] = J;
        m.ieps[p] = 1.0 / d->permittivity[p];
        m.imu[p] = 1.0 / d->permeability[p];
    }
    m.fgeo.assign((size_t)4 * nxzfl, 0.0);
    m.fnb.assign(nxzfl, HOST_FACE_NONE);
    m.fco.assign((size_t)4 * nxzfl, 0.0);
    const int *kf_ed = kf_of_ed();
    for (int64_t e = 0; e < nelt; ++e)
        for (int ef = 1; ef <= 6; ++ef) {
            const int kf = kf_ed[ef];
            const int64_t s0 = e * 6 * n2 + (int64_t)(ef - 1) * n2;
            bool any_partner = false;
            for (int q = 0; q < n2; ++q) {
                const int64_t s = s0 + q;
                const int64_t v = (int64_t)d->cemface[s] - 1;
                if (v != e * n3 + face_node(n1, kf, q)) {
                    snprintf(buf, sizeof buf, "cemface[%lld]=%d does not follow the skpdat face order",
                             (long long)s, d->cemface[s]);
                    return bad(NKC_ERR_ARG, buf);
                }
                const double nv[3] = {d->unxm[s], d->unym[s], d->unzm[s]};
                const double nn = std::sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
                if (!(std::fabs(nn - 1.0) < 1e-8)) {
                    snprintf(buf, sizeof buf, "face node %lld: normal is not a unit vector", (long long)s);
                    return bad(NKC_ERR_ARG, buf);
                }
                for (int c = 0; c < 3; ++c) m.fgeo[(size_t)4 * s + c] = nv[c];
                m.fgeo[(size_t)4 * s + 3] = d->aream[s] / d->bmn[v];
                const int64_t pr = (int64_t)d->face_partner[s] - 1;
                int32_t nb;
                if (pr >= 0) {
                    if (pr >= nxzfl || pecslot[s]) {
                        snprintf(buf, sizeof buf, "face node %lld: partner %lld out of range or on a PEC face",
                                 (long long)s, (long long)pr);
                        return bad(NKC_ERR_ARG, buf);
                    }
                    if ((int64_t)d->face_partner[pr] - 1 != s) {
                        snprintf(buf, sizeof buf, "face node %lld: partner %lld does not pair back", (long long)s,
                                 (long long)pr);
                        return bad(NKC_ERR_ARG, buf);
                    }
                    nb = d->cemface[pr] - 1;
                    any_partner = true;
                    // the element across: the first paired node names it
                    if (m.nbr[e * 6 + kf] == HOST_FACE_NONE) m.nbr[e * 6 + kf] = (int32_t)(pr / (6 * n2));
                } else {
                    nb = pecslot[s] ? HOST_FACE_PEC : HOST_FACE_NONE;
                }
                m.fnb[s] = nb;
                flux_coef(d->Y_0[s], d->Y_1[s], d->Z_0[s], d->Z_1[s], C0, &m.fco[(size_t)4 * s]);
            }
            if (!any_partner && pecslot[s0]) m.nbr[e * 6 + kf] = HOST_FACE_PEC;
        }
    return NKC_OK;
}

// The reference arrays of nkc_create: axis-aligned affine boxes paired
// face to face go to the compact per-element tables of the tuned kernels,
// anything else to the per-node tables of the general kernel.
inline int32_t compact_reference(const nkc_desc *d, HostMesh &m) {
    const int32_t rc = compact_box(d, m);
    if (rc != NKC_ERR_UNSUPPORTED) return rc;
    HostMesh g;
    const int32_t rg = compact_general(d, g);
    if (rg) {
        m.err = g.err;
        return rg;
    }
    m = std::move(g);
    return NKC_OK;
}

// ---------------------------------------------------------------------------
// per-node materials on an existing element table (box pairing): the flux
// coefficients of every face node from eps, mu at the node and at its
// partner, in the order cem_maxwell_materials forms them: gs sum, PEC
// doubling, Z_1 = Z_0' - Z-, Z_0 = Z_0'/2.
// ---------------------------------------------------------------------------
inline void set_node_materials(HostMesh &m, const double *eps, const double *mu, double C0) {
    const int n1 = m.n1, n2 = n1 * n1, n3 = n2 * n1;
    const int64_t nelt = m.nelt, npts = nelt * n3;
    m.mat = 1;
    m.ieps.assign(npts, 0.0);
    m.imu.assign(npts, 0.0);
    m.fco.assign(nelt * 6 * n2 * 4, 0.0);
    for (int64_t p = 0; p < npts; ++p) {
        m.ieps[p] = 1.0 / eps[p];
        m.imu[p] = 1.0 / mu[p];
    }
    for (int64_t e = 0; e < nelt; ++e)
        for (int kf = 0; kf < 6; ++kf) {
            const int32_t nb = m.nbr[e * 6 + kf];
            for (int q = 0; q < n2; ++q) {
                const int64_t v = e * n3 + face_node(n1, kf, q);
                const double Zm = std::sqrt(mu[v] / eps[v]), Ym = std::sqrt(eps[v] / mu[v]);
                double Z0 = Zm, Y0 = Ym;
                if (nb >= 0) {
                    const int64_t w = (int64_t)nb * n3 + face_node(n1, kf ^ 1, q);
                    Z0 += std::sqrt(mu[w] / eps[w]);
                    Y0 += std::sqrt(eps[w] / mu[w]);
                } else if (nb == HOST_FACE_PEC) {
                    Z0 *= 2.0;
                    Y0 *= 2.0;
                }
                flux_coef(0.5 * Y0, Y0 - Ym, 0.5 * Z0, Z0 - Zm, C0, &m.fco[((e * 6 + kf) * n2 + q) * 4]);
            }
        }
}

// ---------------------------------------------------------------------------
// GLL basis (ZWGLL / DGLL, src/nek5_speclib.F:107-121, 807-838): interior
// nodes are the roots of P_N' (Newton with deflation), w = 2/(N(N+1)P_N^2),
// D_ij = P_N(z_i)/(P_N(z_j)(z_i - z_j)), D_00 = -N(N+1)/4, D_NN = N(N+1)/4.
// ---------------------------------------------------------------------------
inline void legendre(double x, int n, double &p, double &dp) {
    double p0 = 1.0, p1 = x, d0 = 0.0, d1 = 1.0;
    if (n == 0) { p = 1.0; dp = 0.0; return; }
    for (int k = 1; k < n; ++k) {
        const double p2 = ((2.0 * k + 1.0) * x * p1 - k * p0) / (k + 1.0);
        const double d2 = d0 + (2.0 * k + 1.0) * p1;
        p0 = p1; p1 = p2; d0 = d1; d1 = d2;
    }
    p = p1;
    dp = d1;
}

inline void gll_basis(int n1, std::vector<double> &z, std::vector<double> &w, std::vector<double> &D) {
    const int N = n1 - 1;
    z.assign(n1, 0.0);
    w.assign(n1, 0.0);
    D.assign(n1 * n1, 0.0);
    z[0] = -1.0;
    z[N] = 1.0;
    for (int i = 1; i < N; ++i) {
        double x = -std::cos(M_PI * i / N);
        for (int it = 0; it < 100; ++it) {
            double p, dp;
            legendre(x, N, p, dp);
            const double d2p = (2.0 * x * dp - N * (N + 1.0) * p) / (1.0 - x * x);
            double defl = 0.0;
            for (int j = 1; j < i; ++j) defl += 1.0 / (x - z[j]);
            const double dx = dp / (d2p - dp * defl);
            x -= dx;
            if (std::fabs(dx) < 1e-17) break;
        }
        z[i] = x;
    }
    for (int i = 0; i < n1 / 2; ++i) {
        const double a = 0.5 * (z[N - i] - z[i]);
        z[i] = -a;
        z[N - i] = a;
    }
    if (n1 % 2) z[N / 2] = 0.0;
    std::vector<double> P(n1);
    for (int i = 0; i < n1; ++i) {
        double dp;
        legendre(z[i], N, P[i], dp);
        w[i] = 2.0 / (N * (N + 1.0) * P[i] * P[i]);
    }
    for (int i = 0; i < n1; ++i)
        for (int j = 0; j < n1; ++j)
            if (i != j) D[i + n1 * j] = P[i] / (P[j] * (z[i] - z[j]));
    D[0] = -N * (N + 1.0) / 4.0;
    D[N + n1 * N] = N * (N + 1.0) / 4.0;
}

// ---------------------------------------------------------------------------
// from a box description
// ---------------------------------------------------------------------------
inline int32_t build_box(const nkc_box_desc *d, HostMesh &m) {
    const int n1 = d->N + 1;
    const int64_t nel[3] = {d->nel[0], d->nel[1], d->nel[2]};
    if (nel[0] <= 0 || nel[1] <= 0 || nel[2] <= 0) { m.err = "box element counts must be positive"; return NKC_ERR_ARG; }
    for (int a = 0; a < 3; ++a)
        if (!(d->hi[a] > d->lo[a])) { m.err = "box extent must be positive"; return NKC_ERR_ARG; }
    const int64_t nelt = nel[0] * nel[1] * nel[2];
    const double C0 = d->ifupwind ? 1.0 : 0.0;
    std::vector<double> z;
    gll_basis(n1, z, m.w, m.D);
    m.n1 = n1;
    m.nelt = nelt;
    m.mat = 0;
    m.geo.assign(nelt * 9, 0.0);
    m.nbr.assign(nelt * 6, 0);
    m.fco.assign(nelt * 6 * 4, 0.0);
    m.ieps.assign(nelt, 0.0);
    m.imu.assign(nelt, 0.0);
    m.jac.assign(nelt, 0.0);
    const double h[3] = {(d->hi[0] - d->lo[0]) / nel[0], (d->hi[1] - d->lo[1]) / nel[1],
                         (d->hi[2] - d->lo[2]) / nel[2]};
    // cofactors of the affine map (GLMAPM1): rx = y_s z_t, sy = x_r z_t, tz = x_r y_s
    const double xr = 0.5 * h[0], ys = 0.5 * h[1], zt = 0.5 * h[2];
    const double J = xr * ys * zt;
    const double rx = ys * zt, sy = xr * zt, tz = xr * ys;
    const double w0 = m.w[0];
    // area/(J w3) at a face node: face Jacobian / (J w_end)  (AREA3 + GEODAT1)
    const double Lx = rx / (J * w0), Ly = sy / (J * w0), Lz = tz / (J * w0);
    auto eps_of = [&](int64_t e) { return d->eps_elem ? d->eps_elem[e] : 1.0; };
    auto mu_of = [&](int64_t e) { return d->mu_elem ? d->mu_elem[e] : 1.0; };
    for (int64_t e = 0; e < nelt; ++e) {
        const double ee = eps_of(e), mm = mu_of(e);
        if (!(ee > 0.0) || !(mm > 0.0)) { m.err = "permittivity and permeability must be positive"; return NKC_ERR_ARG; }
    }
    for (int64_t e = 0; e < nelt; ++e) {
        const int64_t ex[3] = {e % nel[0], (e / nel[0]) % nel[1], e / (nel[0] * nel[1])};
        m.jac[e] = J;
        double *g = &m.geo[e * 9];
        g[0] = rx / J;
        g[1] = sy / J;
        g[2] = tz / J;
        g[3] = g[4] = Lx;
        g[5] = g[6] = Ly;
        g[7] = g[8] = Lz;
        const double Zm = std::sqrt(mu_of(e) / eps_of(e)), Ym = std::sqrt(eps_of(e) / mu_of(e));
        m.ieps[e] = 1.0 / eps_of(e);
        m.imu[e] = 1.0 / mu_of(e);
        for (int kf = 0; kf < 6; ++kf) {
            const int ax = kf / 2, side = kf & 1;
            int64_t c[3] = {ex[0], ex[1], ex[2]};
            c[ax] += side ? 1 : -1;
            int32_t nb;
            double Y0, Z0;   // gs sums, as cem_maxwell_materials forms them
            if (c[ax] < 0 || c[ax] >= nel[ax]) {
                if (d->pec[ax]) {
                    nb = HOST_FACE_PEC;
                    Y0 = 2.0 * Ym;
                    Z0 = 2.0 * Zm;
                } else {
                    c[ax] = (c[ax] + nel[ax]) % nel[ax];
                    nb = (int32_t)(c[0] + nel[0] * (c[1] + nel[1] * c[2]));
                    Y0 = Ym + std::sqrt(eps_of(nb) / mu_of(nb));
                    Z0 = Zm + std::sqrt(mu_of(nb) / eps_of(nb));
                }
            } else {
                nb = (int32_t)(c[0] + nel[0] * (c[1] + nel[1] * c[2]));
                Y0 = Ym + std::sqrt(eps_of(nb) / mu_of(nb));
                Z0 = Zm + std::sqrt(mu_of(nb) / eps_of(nb));
            }
            m.nbr[e * 6 + kf] = nb;
            const double Y1 = Y0 - Ym, Z1 = Z0 - Zm;
            flux_coef(0.5 * Y0, Y1, 0.5 * Z0, Z1, C0, &m.fco[(e * 6 + kf) * 4]);
        }
    }
    return NKC_OK;
}

// ---------------------------------------------------------------------------
// A mesh of axis-aligned box elements given the way NekCEM's readers leave
// it (rdmesh / bin_rd1_mesh, rdbdry / bin_rd1_bc, src/nek5_connect2.F):
// corners xc, yc, zc (8 per element, corners 1-4 on t = -1 counter-
// clockwise from (r,s) = (-1,-1), then 5-8 on t = +1), and per face (Nek
// face order 1: s = -1, 2: r = +1, 3: s = +1, 4: r = -1, 5: t = -1,
// 6: t = +1) a 3-character boundary code and 5 parameters.  Codes: 'E'
// interior and 'P' periodic, with the neighbour element and face in
// parameters 1-2 (1-based) or 0 to find it by the face's position (a .re2
// lists only boundary faces); 'PEC' a perfect conductor, and so 'PML' (the
// wall behind an absorbing layer).  Each face must
// pair with the opposite face of its neighbour over the same extent.
// ---------------------------------------------------------------------------
inline int32_t build_hex(int n1, int ifupwind, int64_t nelt, const double *xc, const double *yc,
                         const double *zc, const char *cbc, const double *bc, const double *eps_elem,
                         const double *mu_elem, HostMesh &m) {
    static const int nek_to_kf[6] = {2, 1, 3, 0, 4, 5};   // Nek face f (0-based) -> x-,x+,y-,y+,z-,z+
    char buf[256];
    auto bad = [&](int32_t code, const std::string &msg) { m.err = msg; return code; };
    const double C0 = ifupwind ? 1.0 : 0.0;
    std::vector<double> z;
    gll_basis(n1, z, m.w, m.D);
    m.n1 = n1;
    m.nelt = nelt;
    m.mat = 0;
    m.geo.assign(nelt * 9, 0.0);
    m.nbr.assign(nelt * 6, HOST_FACE_NONE);
    m.fco.assign(nelt * 6 * 4, 0.0);
    m.ieps.assign(nelt, 0.0);
    m.imu.assign(nelt, 0.0);
    m.jac.assign(nelt, 0.0);
    std::vector<double> lo(3 * nelt), hi(3 * nelt);
    for (int64_t e = 0; e < nelt; ++e) {
        const double *c[3] = {xc + 8 * e, yc + 8 * e, zc + 8 * e};
        // corner q has r = -1 for q in {0,3,4,7}, s = -1 for {0,1,4,5}, t = -1 for {0..3}
        static const int side[3][8] = {{0, 1, 1, 0, 0, 1, 1, 0}, {0, 0, 1, 1, 0, 0, 1, 1}, {0, 0, 0, 0, 1, 1, 1, 1}};
        for (int a = 0; a < 3; ++a) {
            double l = 0, h = 0;
            for (int q = 0; q < 8; ++q) (side[a][q] ? h : l) += 0.25 * c[a][q];
            const double tol = 1e-9 * std::max(1.0, std::fabs(h - l));
            for (int q = 0; q < 8; ++q)
                if (std::fabs(c[a][q] - (side[a][q] ? h : l)) > tol) {
                    snprintf(buf, sizeof buf, "element %lld is not an axis-aligned box", (long long)e + 1);
                    return bad(NKC_ERR_UNSUPPORTED, buf);
                }
            if (!(h > l)) {
                snprintf(buf, sizeof buf, "element %lld has non-positive extent", (long long)e + 1);
                return bad(NKC_ERR_ARG, buf);
            }
            lo[3 * e + a] = l;
            hi[3 * e + a] = h;
        }
    }
    auto eps_of = [&](int64_t e) { return eps_elem ? eps_elem[e] : 1.0; };
    auto mu_of = [&](int64_t e) { return mu_elem ? mu_elem[e] : 1.0; };
    for (int64_t e = 0; e < nelt; ++e)
        if (!(eps_of(e) > 0.0) || !(mu_of(e) > 0.0)) return bad(NKC_ERR_ARG, "permittivity and permeability must be positive");
    // faces by position for the links a .re2 leaves out: key = (axis, side, plane, centre)
    auto face_key = [&](int64_t e, int kf) {
        const int ax = kf / 2, sd = kf & 1;
        const double plane = sd ? hi[3 * e + ax] : lo[3 * e + ax];
        double cen[2];
        int t = 0;
        for (int a = 0; a < 3; ++a)
            if (a != ax) cen[t++] = 0.5 * (lo[3 * e + a] + hi[3 * e + a]);
        const double sc = 1e9;
        return std::make_tuple(ax, sd, std::llround(plane * sc), std::llround(cen[0] * sc), std::llround(cen[1] * sc));
    };
    std::map<std::tuple<int, int, long long, long long, long long>, int64_t> faces;
    for (int64_t e = 0; e < nelt; ++e)
        for (int kf = 0; kf < 6; ++kf) faces[face_key(e, kf)] = e;
    for (int64_t e = 0; e < nelt; ++e)
        for (int f = 0; f < 6; ++f) {
            const int kf = nek_to_kf[f], ax = kf / 2;
            std::string code(cbc + (e * 6 + f) * 3, 3);
            while (!code.empty() && code.back() == ' ') code.pop_back();
            const double *p = bc + (e * 6 + f) * 5;
            int32_t nb;
            if (code == "PEC" || code == "PML") {
                // cem_maxwell_pec_init (src/cem_maxwell.F:1340-1366) lists
                // both; the PML layers themselves come from nkc_setup_pml
                nb = HOST_FACE_PEC;
            } else if (code == "E" || code == "P") {
                int64_t ne = (int64_t)std::llround(p[0]) - 1;
                int nf = (int)std::llround(p[1]) - 1;
                if (ne < 0) {
                    if (code == "P") {
                        snprintf(buf, sizeof buf, "periodic face %d of element %lld has no partner", f + 1, (long long)e + 1);
                        return bad(NKC_ERR_ARG, buf);
                    }
                    auto k = face_key(e, kf);
                    std::get<1>(k) ^= 1;   // the neighbour's opposite face at the same place
                    auto it = faces.find(k);
                    if (it == faces.end()) {
                        snprintf(buf, sizeof buf, "no element meets face %d of element %lld", f + 1, (long long)e + 1);
                        return bad(NKC_ERR_ARG, buf);
                    }
                    ne = it->second;
                    nf = -1;
                }
                if (ne < 0 || ne >= nelt) {
                    snprintf(buf, sizeof buf, "face %d of element %lld names element %lld", f + 1, (long long)e + 1, (long long)ne + 1);
                    return bad(NKC_ERR_ARG, buf);
                }
                if (nf >= 0 && (nf > 5 || nek_to_kf[nf] != (kf ^ 1))) {
                    snprintf(buf, sizeof buf, "face %d of element %lld pairs with face %d, not the opposite one", f + 1,
                             (long long)e + 1, nf + 1);
                    return bad(NKC_ERR_UNSUPPORTED, buf);
                }
                for (int a = 0; a < 3; ++a)
                    if (a != ax && (std::fabs(lo[3 * e + a] - lo[3 * ne + a]) > 1e-9 * std::max(1.0, std::fabs(lo[3 * e + a])) ||
                                    std::fabs(hi[3 * e + a] - hi[3 * ne + a]) > 1e-9 * std::max(1.0, std::fabs(hi[3 * e + a])))) {
                        snprintf(buf, sizeof buf, "face %d of element %lld and its partner differ in extent", f + 1, (long long)e + 1);
                        return bad(NKC_ERR_UNSUPPORTED, buf);
                    }
                nb = (int32_t)ne;
            } else {
                snprintf(buf, sizeof buf, "boundary code '%s' (element %lld face %d) is not on this path", code.c_str(),
                         (long long)e + 1, f + 1);
                return bad(NKC_ERR_UNSUPPORTED, buf);
            }
            m.nbr[e * 6 + kf] = nb;
        }
    const double w0 = m.w[0];
    for (int64_t e = 0; e < nelt; ++e) {
        const double xr = 0.5 * (hi[3 * e] - lo[3 * e]), ys = 0.5 * (hi[3 * e + 1] - lo[3 * e + 1]);
        const double zt = 0.5 * (hi[3 * e + 2] - lo[3 * e + 2]);
        const double J = xr * ys * zt, rx = ys * zt, sy = xr * zt, tz = xr * ys;
        m.jac[e] = J;
        double *g = &m.geo[e * 9];
        g[0] = rx / J;
        g[1] = sy / J;
        g[2] = tz / J;
        g[3] = g[4] = rx / (J * w0);
        g[5] = g[6] = sy / (J * w0);
        g[7] = g[8] = tz / (J * w0);
        const double Zm = std::sqrt(mu_of(e) / eps_of(e)), Ym = std::sqrt(eps_of(e) / mu_of(e));
        m.ieps[e] = 1.0 / eps_of(e);
        m.imu[e] = 1.0 / mu_of(e);
        for (int kf = 0; kf < 6; ++kf) {
            const int32_t nb = m.nbr[e * 6 + kf];
            double Y0 = Ym, Z0 = Zm;   // gs sums, as cem_maxwell_materials forms them
            if (nb >= 0) {
                Y0 += std::sqrt(eps_of(nb) / mu_of(nb));
                Z0 += std::sqrt(mu_of(nb) / eps_of(nb));
            } else if (nb == HOST_FACE_PEC) {
                Y0 *= 2.0;
                Z0 *= 2.0;
            }
            flux_coef(0.5 * Y0, Y0 - Ym, 0.5 * Z0, Z0 - Zm, C0, &m.fco[(e * 6 + kf) * 4]);
        }
    }
    return NKC_OK;
}

// ---------------------------------------------------------------------------
// PML layers and sigma as cem_maxwell_init sets them up from the faces whose
// boundary code is 'PML' (pml_face[6 e + f - 1], reference face order):
// pml_errchk (src/cem_maxwell_pml.F:2-29), the layers `thick` elements
// inward from each such wall along its axis (march_faces on a conforming
// axis-aligned mesh), pml_extent_and_tags (:261-431) for the inner / outer
// plane of each of the six layers, and pml_calc_sigma (:433-505):
// sigma_axis = sigmamax ((x - inner) / (outer - inner))^order with
// sigmamax = -(order + 1) log(referr) / (2 eta width), eta = sqrt(mu/eps).
// xyz: node coordinates, 3 components at stride fs.  Out: the PML elements
// (ascending), sigma (3 n3 per PML element, component-major), extents
// (pmlinner[6] then pmlouter[6], order x-, x+, y-, y+, z-, z+).
// ---------------------------------------------------------------------------
inline int32_t pml_tables(const HostMesh &hm, const double *xyz, int64_t fs, const int8_t *pml_face, int thick,
                          double order, double referr, std::vector<int64_t> &elems, std::vector<double> &sig,
                          double extents[12], std::string &err) {
    char buf[200];
    auto bad = [&](const char *msg) { err = msg; return (int32_t)NKC_ERR_ARG; };
    if (thick < 1 || thick > 10) { snprintf(buf, sizeof buf, "PML thickness %d not in 1..10", thick); return bad(buf); }
    if (!(order >= 1.0 && order <= 10.0)) { snprintf(buf, sizeof buf, "PML order %g not in 1..10", order); return bad(buf); }
    if (!(referr > 0.0 && referr <= 1.0)) {
        snprintf(buf, sizeof buf, "PML reflection error %g not in (0, 1]", referr);
        return bad(buf);
    }
    const int64_t nelt = hm.nelt, n3 = (int64_t)hm.n1 * hm.n1 * hm.n1;
    const int *kf_ed = kf_of_ed();
    std::vector<int32_t> tag(nelt, 0);
    for (int64_t e = 0; e < nelt; ++e)
        for (int f = 1; f <= 6; ++f) {
            if (!pml_face[e * 6 + f - 1]) continue;
            const int kf = kf_ed[f];
            if (hm.nbr[e * 6 + kf] >= 0) {
                snprintf(buf, sizeof buf, "PML face %d of element %lld is not a wall", f, (long long)e + 1);
                return bad(buf);
            }
            int64_t cur = e;
            for (int l = 0; l < thick; ++l) {
                tag[cur] |= 1 << kf;
                if ((tag[cur] >> (kf ^ 1)) & 1) {
                    snprintf(buf, sizeof buf, "PML layers collide in element %lld", (long long)cur + 1);
                    return bad(buf);
                }
                if (l + 1 == thick) break;
                const int32_t nb = hm.nbr[cur * 6 + (kf ^ 1)];
                if (nb < 0) {
                    snprintf(buf, sizeof buf, "the PML is thicker than the mesh at element %lld", (long long)cur + 1);
                    return bad(buf);
                }
                cur = nb;
            }
        }
    const double inf = 1e20;
    double inner[6] = {-inf, inf, -inf, inf, -inf, inf}, outer[6] = {inf, -inf, inf, -inf, inf, -inf};
    elems.clear();
    for (int64_t e = 0; e < nelt; ++e) {
        if (!tag[e]) continue;
        elems.push_back(e);
        for (int f = 0; f < 6; ++f) {
            if (!((tag[e] >> f) & 1)) continue;
            const int ax = f / 2;
            double lo = inf, hi = -inf;
            for (int64_t v = 0; v < n3; ++v) {
                const double x = xyz[ax * fs + e * n3 + v];
                lo = std::min(lo, x);
                hi = std::max(hi, x);
            }
            if (f & 1) { inner[f] = std::min(inner[f], lo); outer[f] = std::max(outer[f], hi); }
            else       { inner[f] = std::max(inner[f], hi); outer[f] = std::min(outer[f], lo); }
        }
    }
    sig.assign(elems.size() * 3 * n3, 0.0);
    for (size_t l = 0; l < elems.size(); ++l) {
        const int64_t e = elems[l];
        for (int f = 0; f < 6; ++f) {
            if (!((tag[e] >> f) & 1)) continue;
            const int ax = f / 2;
            const double width = std::fabs(outer[f] - inner[f]);
            for (int64_t v = 0; v < n3; ++v) {
                const int64_t g = e * n3 + v;
                const size_t mi = hm.mat ? (size_t)g : (size_t)e;
                const double eta = std::sqrt((1.0 / hm.imu[mi]) / (1.0 / hm.ieps[mi]));
                const double smax = -(order + 1.0) * std::log(referr) / (2.0 * eta * width);
                const double z01 = (xyz[ax * fs + g] - inner[f]) / (outer[f] - inner[f]);
                sig[(l * 3 + ax) * n3 + v] = smax * std::pow(z01, order);
            }
        }
    }
    for (int f = 0; f < 6; ++f) { extents[f] = inner[f]; extents[6 + f] = outer[f]; }
    return NKC_OK;
}

}  // namespace nkc
