"""User-hook library (the .usr side): initial conditions and exact solutions.

The analytic modes are those of the reference's own test cases:
3dboxper (tests/3dboxper/3dboxper.usr:45-86, periodic (1,1,1) mode, omega =
sqrt(3)) and 3dboxpec (tests/3dboxpec/3dboxpec.usr:91-113, PEC cavity mode,
omega = pi*sqrt(3)).
"""
import numpy as np


def boxper_solution(t, x, y, z):
    om = np.sqrt(3.0)
    th, te = np.sin(om * t) / om, np.cos(om * t)
    H = np.stack([2 * np.cos(x) * np.sin(y) * np.cos(z) * th,
                  -np.sin(x) * np.cos(y) * np.cos(z) * th,
                  np.sin(x) * np.sin(y) * np.sin(z) * th])
    E = np.stack([0.0 * x,
                  np.cos(x) * np.sin(y) * np.sin(z) * te,
                  np.cos(x) * np.cos(y) * np.cos(z) * te])
    return H, E


def boxpec_solution(t, x, y, z):
    ww = np.pi
    th = np.cos(ww * np.sqrt(3.0) * t) / np.sqrt(6.0)
    te = np.sin(ww * np.sqrt(3.0) * t) / np.sqrt(2.0)
    H = np.stack([-np.sin(ww * x) * np.cos(ww * y) * np.cos(ww * z) * th,
                  -np.cos(ww * x) * np.sin(ww * y) * np.cos(ww * z) * th,
                  2 * np.cos(ww * x) * np.cos(ww * y) * np.sin(ww * z) * th])
    E = np.stack([-np.cos(ww * x) * np.sin(ww * y) * np.sin(ww * z) * te,
                  np.sin(ww * x) * np.cos(ww * y) * np.sin(ww * z) * te,
                  0.0 * x])
    return H, E


def plane_waves(x, y, z, seed=2, nwaves=8, kmax=4):
    """Superposition of nwaves plane waves with integer wavevectors
    0 < |k| <= kmax (periodic on [0, 2pi]^3), unit polarisation p orthogonal
    to k, seeded phase: E = sum p cos(k.x + phi), H = sum (khat x p) cos(...)."""
    rng = np.random.default_rng(seed)
    E = np.zeros((3,) + x.shape)
    H = np.zeros((3,) + x.shape)
    n = 0
    while n < nwaves:
        k = rng.integers(-kmax, kmax + 1, size=3).astype(float)
        kn = np.linalg.norm(k)
        if kn == 0 or kn > kmax:
            continue
        khat = k / kn
        v = rng.normal(size=3)
        p = v - np.dot(v, khat) * khat
        p /= np.linalg.norm(p)
        phi = rng.uniform(0, 2 * np.pi)
        ph = np.cos(k[0] * x + k[1] * y + k[2] * z + phi)
        q = np.cross(khat, p)
        for c in range(3):
            E[c] += p[c] * ph
            H[c] += q[c] * ph
        n += 1
    return H, E


def element_eps(nelt, seed=1, lo=1.0, hi=4.0):
    return np.random.default_rng(seed).uniform(lo, hi, nelt)


def cavity_batch(ncav=256, seed=3, nmin=2, nmax=16, Nmin=5, Nmax=9):
    """The batched workload: ncav independent PEC box cavities on [0,1]^3,
    each n^3 elements with n ~ U{nmin..nmax} and polynomial order
    N ~ U{Nmin..Nmax}, per-element eps ~ U[1,4], mu = 1 (one seed)."""
    rng = np.random.default_rng(seed)
    cavs = []
    for _ in range(ncav):
        n = int(rng.integers(nmin, nmax + 1))
        N = int(rng.integers(Nmin, Nmax + 1))
        cavs.append(dict(nel=(n, n, n), N=N, lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0),
                         bc=("PEC", "PEC", "PEC"), eps_elem=rng.uniform(1.0, 4.0, n ** 3)))
    return cavs


def gaussian_ez(x, y, z, center=(0.5, 0.5, 0.5), width=0.2):
    """Initial state with Ez = exp(-|x-c|^2 / (2 w^2)), all other fields 0."""
    r2 = (x - center[0]) ** 2 + (y - center[1]) ** 2 + (z - center[2]) ** 2
    E = np.zeros((3,) + x.shape)
    E[2] = np.exp(-r2 / (2.0 * width * width))
    return np.zeros_like(E), E


def smooth_field(x, y, z, seed, nmodes=4, kmax=2):
    """Smooth field on [0,1]^3: a few low-wavenumber sinusoids with seeded
    integer wavevectors and phases, scaled to max |s| = 1."""
    rng = np.random.default_rng(seed)
    s = np.zeros_like(x)
    for _ in range(nmodes):
        k = rng.integers(-kmax, kmax + 1, size=3)
        if not np.any(k):
            k[0] = 1
        s += rng.normal() * np.cos(np.pi * (k[0] * x + k[1] * y + k[2] * z) + rng.uniform(0, 2 * np.pi))
    return s / np.max(np.abs(s))


def gaussian_source(x, y, z, bm1, center=(0.4, 0.55, 0.45), width=0.1):
    """bm1-weighted spatial amplitude of a Gaussian pulse source on Ez."""
    r2 = (x - center[0]) ** 2 + (y - center[1]) ** 2 + (z - center[2]) ** 2
    return np.exp(-r2 / (2.0 * width * width)) * bm1


# ---------------------------------------------------------------------------
# 3ddielectric (tests/3ddielectric/3ddielectric.usr): a plane wave of
# frequency omega = 2 comes from y > 0 onto the interface y = 0 between
# (eps1, mu1) = (1, 1) above and (eps2, mu2) = (1 or 2, 1) below, with PML
# walls in y.  uservp (lines 171-296) sets the materials from each element's
# probe node ym1(nx1/2, nx1/2, nx1/2, e) and lists the upper elements' faces
# that lie on y = 0 for userinc (6-47); usrdat2 (300-330) maps the box onto
# [-2.5, 2.5] x [-5, 5] x [-2.5, 2.5]; usersol (83-147) is the exact
# reflected / transmitted wave, damped inside the PML; userchk (332-392)
# holds the tolerances.
# ---------------------------------------------------------------------------

DIELECTRIC_L2TOL = (5e-4, 5e-14, 5e-4, 5e-4, 5e-14, 5e-4)
DIELECTRIC_LINFTOL = {False: (5e-3, 5e-12, 5e-3, 5e-3, 5e-12, 5e-3),
                      True: (5e-3, 1e-12, 5e-3, 5e-3, 1e-12, 5e-3)}


def dielectric_params(twomat):
    """The /userparam/ block of uservp: omega, eps1, eps2, mu1, mu2 and the
    TE / TM reflection and transmission coefficients."""
    omega, eps1, eps2, mu1, mu2 = 2.0, 1.0, (2.0 if twomat else 1.0), 1.0, 1.0
    z1, z2 = np.sqrt(mu1 / eps1), np.sqrt(mu2 / eps2)
    return dict(omega=omega, eps1=eps1, eps2=eps2, mu1=mu1, mu2=mu2, reflte=(z1 - z2) / (z1 + z2),
                trante=2 * z1 / (z1 + z2), refltm=(z2 - z1) / (z1 + z2), trantm=2 * z2 / (z1 + z2))


def rescale_usrdat2(sizes):
    """usrdat2 of the box cases: each axis mapped onto [-s/2, s/2]."""
    def f(xc, yc, zc):
        out = []
        for c, sz in zip((xc, yc, zc), sizes):
            a, b = c.min(), c.max()
            out.append(sz * (c - a) / (b - a) - sz / 2.0)
        return out
    return f


def probe_upper(lo_y, hi_y, n1):
    """uservp's region test per element: ym1 at the probe node (1-based
    nx1/2 along each direction) > 0."""
    from .mesh import gll
    z, _ = gll(n1)
    zp = z[n1 // 2 - 1]
    return lo_y + (hi_y - lo_y) * 0.5 * (1.0 + zp) > 0.0


def dielectric_materials(upper, n1, twomat):
    """Per-node permittivity and permeability (npts each) of uservp."""
    p = dielectric_params(twomat)
    per = np.where(upper, p["eps1"], p["eps2"])
    eps = np.repeat(per, n1 ** 3)
    mu = np.repeat(np.where(upper, p["mu1"], p["mu2"]), n1 ** 3)
    return eps, mu


def dielectric_incident(upper, lo_y, hi_y, twomat):
    """uservp's incident faces and userinc's wave: (faces, k, omega, hamp,
    eamp) for MaxwellSolver.set_incident.  Faces are the upper elements'
    faces whose every point has |y| <= 1e-8: on an axis-aligned box only a
    y face can, face 1 (s = -1, y = lo) or 3 (s = +1, y = hi)."""
    p = dielectric_params(twomat)
    faces = []
    for e in np.nonzero(upper)[0]:
        if abs(lo_y[e]) <= 1e-8:
            faces.append((e, 1))
        if abs(hi_y[e]) <= 1e-8:
            faces.append((e, 3))
    eta = np.sqrt(p["mu1"] / p["eps1"])
    ky = p["omega"] * np.sqrt(p["mu1"] * p["eps1"])
    # uinc = cos(-ky y - omega t); hz += u, ex += eta u, ez += u, hx -= u / eta
    return faces, (0.0, -ky, 0.0), p["omega"], (-1.0 / eta, 0.0, 1.0), (eta, 0.0, 1.0)


def dielectric_solution(t, y, eps, mu, upper_nodes, pml_nodes, inner, outer, order, referr, twomat):
    """usersol: (H, E), (3, npts) each; y, eps, mu per node; upper_nodes /
    pml_nodes: the node's element is in the upper region / a PML element;
    inner, outer: pmlinner / pmlouter (x-, x+, y-, y+, z-, z+)."""
    p = dielectric_params(twomat)
    om = p["omega"]
    eta = np.sqrt(mu / eps)
    ky = om * np.sqrt(eps * mu)
    H = np.zeros((3, len(y)))
    E = np.zeros((3, len(y)))
    # upper region: the reflected wave, damped in the +y PML
    d = outer[3] - inner[3]
    smax = -(order + 1) * np.log(referr) / (2 * eta * d)
    fac = np.where(pml_nodes, smax * d / (order + 1) * np.abs((y - inner[3]) / d) ** (order + 1), 0.0)
    uu = np.exp(-eta * fac) * np.cos(ky * y - om * t)
    up = upper_nodes
    H[2] = np.where(up, p["reflte"] * uu, H[2])
    E[0] = np.where(up, -eta * p["reflte"] * uu, E[0])
    E[2] = np.where(up, p["refltm"] * uu, E[2])
    H[0] = np.where(up, p["refltm"] * uu / eta, H[0])
    # lower region: the transmitted wave, damped in the -y PML
    d = inner[2] - outer[2]
    smax = -(order + 1) * np.log(referr) / (2 * eta * d)
    fac = np.where(pml_nodes, smax * d / (order + 1) * np.abs((inner[2] - y) / d) ** (order + 1), 0.0)
    uu = np.exp(-eta * fac) * np.cos(-ky * y - om * t)
    lo = ~upper_nodes
    H[2] = np.where(lo, p["trante"] * uu, H[2])
    E[0] = np.where(lo, eta * p["trante"] * uu, E[0])
    E[2] = np.where(lo, p["trantm"] * uu, E[2])
    H[0] = np.where(lo, -p["trantm"] * uu / eta, H[0])
    return H, E


def dipole_amplitude(x, y, z, bm1, width=0.1):
    """usersrc of 3dboxpml (tests/3dboxpml/3dboxpml.usr:30-88): the Gaussian
    Hertzian dipole at the origin on Ez, srcez -= i0 norm exp(-r^2 / (2 w^2))
    sin(-omega t) bm1 with i0 norm = 1: the bm1-weighted amplitude for
    set_source(..., direction=(0, 0, 1), omega=-2.0)."""
    return -np.exp(-0.5 * (x / width) ** 2 - 0.5 * (y / width) ** 2 - 0.5 * (z / width) ** 2) * bm1


# ---------------------------------------------------------------------------
# 3dgraphene (tests/3dgraphene/3dgraphene.usr): two superimposed normally
# incident plane waves striking a graphene sheet on y = 0, whose surface
# current is a Drude term plus two critical-point terms.
# ---------------------------------------------------------------------------
# userchk tolerances (3dgraphene.usr:483-497), order Hx Hy Hz Ex Ey Ez
GRAPHENE_L2TOL = (5e-4, 1e-14, 5e-4, 5e-4, 1e-14, 5e-4)
GRAPHENE_LINFTOL = (5e-3, 5e-12, 5e-3, 5e-3, 5e-12, 5e-3)
# sheet parameters of uservp (3dgraphene.usr:390-401): a_d, b_d, b_cp1,
# a_211, a_221, b_11, b_21, b_cp2, a_212, a_222, b_12, b_22
GRAPHENE_SHEET = (0.0, 1.499, -2.599e-03, 4.632e+05, 1.090e+03, -1.391, -3.125e+02, -1.049e-03, 4.271e+05,
                  2.742e+02, -7.769e-02, 4.268e+02)


def graphene_params(sheet=GRAPHENE_SHEET):
    """The /userparam/ block of uservp (3dgraphene.usr:385-418): omega,
    eps1, eps2, mu1, mu2, the complex sheet conductivity at omega and the TE /
    TM reflection and transmission coefficients (sheet: the twelve
    parameters)."""
    omega, eps1, eps2, mu1, mu2 = 5.0, 1.0, 1.0, 1.0, 1.0
    a_d, b_d, b_cp1, a_211, a_221, b_11, b_21, b_cp2, a_212, a_222, b_12, b_22 = sheet
    w = omega
    s_d = b_d / (a_d - 1j * w)
    s_1 = (1j / w) * ((a_211 * b_11 + 1j * w * b_21) / (w ** 2 - a_211 + 1j * w * a_221) + b_11) - b_cp1
    s_2 = (1j / w) * ((a_212 * b_12 + 1j * w * b_22) / (w ** 2 - a_212 + 1j * w * a_222) + b_12) - b_cp2
    sg = s_d + s_1 + s_2
    z1, z2 = np.sqrt(mu1 / eps1), np.sqrt(mu2 / eps2)
    den = z1 + z2 + sg * z1 * z2
    return dict(omega=omega, eps1=eps1, eps2=eps2, mu1=mu1, mu2=mu2, sigma=sg,
                reflte=(z1 - z2 + sg * z1 * z2) / den, trante=2 * z1 / den,
                refltm=(z2 - z1 - z1 * z2 * sg) / den, trantm=2 * z2 / den)


def sheet_faces(lo_y, hi_y, elements=None):
    """Faces (element, Ed's face) whose every point has |y| <= 1e-8, the
    tests of uservp (3dgraphene.usr:420-470): on an axis-aligned box only a
    y face can, face 1 (y = lo) or 3 (y = hi).  elements: the elements to
    look at (default all)."""
    faces = []
    for e in (range(len(lo_y)) if elements is None else elements):
        if abs(lo_y[e]) <= 1e-8:
            faces.append((int(e), 1))
        if abs(hi_y[e]) <= 1e-8:
            faces.append((int(e), 3))
    return faces


def graphene_incident(upper, lo_y, hi_y):
    """userinc (3dgraphene.usr:7-47) on the upper elements' sheet faces:
    (faces, k, omega, hamp, eamp) for MaxwellSolver.set_incident, with
    uinc = cos(-ky y - omega t): hz += u, ex += eta u, ez += u, hx -= u / eta
    (the upper region's eta and ky; eps = mu = 1 on both sides)."""
    p = graphene_params()
    faces = sheet_faces(lo_y, hi_y, np.nonzero(upper)[0])
    eta = np.sqrt(p["mu1"] / p["eps1"])
    ky = p["omega"] * np.sqrt(p["mu1"] * p["eps1"])
    return faces, (0.0, -ky, 0.0), p["omega"], (-1.0 / eta, 0.0, 1.0), (eta, 0.0, 1.0)


def graphene_initial_state(sheet=GRAPHENE_SHEET):
    """userini's sheet state (3dgraphene.usr:97-147): fjn slots 2..6 of
    every sheet point, (5, 3), from half the tangential complex E field at
    the interface at t = 0."""
    p = graphene_params(sheet)
    a_d, b_d, b_cp1, a_211, a_221, b_11, b_21, b_cp2, a_212, a_222, b_12, b_22 = sheet
    w = p["omega"]
    z1 = np.sqrt(p["mu1"] / p["eps1"])
    enpar = np.array([0.5 * z1 * (1.0 - p["reflte"]), 0.0, 0.5 * (1.0 + p["refltm"])])
    out = np.zeros((5, 3))
    out[0] = (b_d / (a_d - 1j * w) * enpar).real
    f4 = (a_211 * b_11 + 1j * w * b_21) / (w ** 2 - a_211 + 1j * w * a_221)
    out[2] = (f4 * enpar).real
    out[1] = ((1j / w) * (f4 + b_11) * enpar).real
    f6 = (a_212 * b_12 + 1j * w * b_22) / (w ** 2 - a_212 + 1j * w * a_222)
    out[4] = (f6 * enpar).real
    out[3] = ((1j / w) * (f6 + b_12) * enpar).real
    return out


def graphene_solution(t, y, eps, mu, upper_nodes, pml_nodes, inner, outer, order, referr, sheet=GRAPHENE_SHEET):
    """usersol (3dgraphene.usr:150-221): (H, E), (3, npts) each, the complex
    reflected / transmitted waves' real parts, damped inside the PML; y,
    eps, mu per node, upper_nodes / pml_nodes per node, inner / outer the
    PML planes (x-, x+, y-, y+, z-, z+)."""
    p = graphene_params(sheet)
    om = p["omega"]
    eta = np.sqrt(mu / eps)
    ky = om * np.sqrt(eps * mu)
    H = np.zeros((3, len(y)))
    E = np.zeros((3, len(y)))
    d = outer[3] - inner[3]
    smax = -(order + 1) * np.log(referr) / (2 * eta * d)
    fac = np.where(pml_nodes, smax * d / (order + 1) * np.abs((y - inner[3]) / d) ** (order + 1), 0.0)
    uu = np.exp(1j * (ky * y - om * t) - eta * fac)
    up = upper_nodes
    H[2] = np.where(up, (p["reflte"] * uu).real, H[2])
    E[0] = np.where(up, -(p["reflte"] * eta * uu).real, E[0])
    E[2] = np.where(up, (p["refltm"] * uu).real, E[2])
    H[0] = np.where(up, (p["refltm"] * uu / eta).real, H[0])
    d = inner[2] - outer[2]
    smax = -(order + 1) * np.log(referr) / (2 * eta * d)
    fac = np.where(pml_nodes, smax * d / (order + 1) * np.abs((inner[2] - y) / d) ** (order + 1), 0.0)
    uu = np.exp(1j * (-ky * y - om * t) - eta * fac)
    lo = ~upper_nodes
    H[2] = np.where(lo, (p["trante"] * uu).real, H[2])
    E[0] = np.where(lo, (p["trante"] * eta * uu).real, E[0])
    E[2] = np.where(lo, (p["trantm"] * uu).real, E[2])
    H[0] = np.where(lo, -(p["trantm"] * uu / eta).real, H[0])
    return H, E


# ---------------------------------------------------------------------------
# tests/cylwave: the TM_01 mode of a circular PEC waveguide, periodic in z
# (tests/cylwave/cylwave.usr)
# ---------------------------------------------------------------------------
CYLWAVE_L2TOL = (5e-9,) * 6
CYLWAVE_LINFTOL = (5e-8,) * 6


def xy_radius(xc, yc):
    """geom_xyradius (src/cem_common.F:817-843): the largest corner radius
    in the x-y plane."""
    return float(np.sqrt(np.max(np.asarray(xc) ** 2 + np.asarray(yc) ** 2)))


def cylwave_usrdat(xc, yc, zc):
    """cylwave's usrdat: corners within 1e-4 (in r^2) of the radius, rounded
    down to a tenth, are moved onto that circle."""
    xc, yc = np.array(xc, dtype=np.float64), np.array(yc, dtype=np.float64)
    radius = int(10 * xy_radius(xc, yc) + 0.1) / 10.0
    rr = xc * xc + yc * yc
    on = rr > radius * radius - 1e-4
    rn = radius / np.sqrt(rr[on])
    xc[on] *= rn
    yc[on] *= rn
    return xc, yc, np.array(zc, dtype=np.float64)


def cylwave_usrdat2(x, y, z):
    """cylwave's usrdat2: z scaled onto [0, 2 pi xmax]."""
    zmin, zmax = float(np.min(z)), float(np.max(z))
    sz = 2 * np.pi * float(np.max(x)) / (zmax - zmin)
    return x, y, sz * (z - zmin)


def cylwave_solution(t, x, y, z, radius):
    """cylwave's usersol with m = 0, the first root of J_0, one wavelength
    along z: E_z = Re(a J_0(K rho)), E_rho = Re(-i a kz/K J_0'(K rho)),
    H_phi = Re(-i a omega/K J_0'(K rho)), a = exp(-i kz z) exp(i omega t),
    K = j_01 / radius, omega = sqrt(kz^2 + K^2)."""
    from scipy.special import jn_zeros, jv, jvp
    zsize = float(np.max(z) - np.min(z))
    kz = 2 * np.pi / zsize if abs(zsize) > 1e-20 else 0.0
    K = float(jn_zeros(0, 1)[0]) / radius
    om = np.sqrt(kz * kz + K * K)
    rho = np.sqrt(x * x + y * y)
    phi = np.arctan2(y, x)
    a = np.exp(-1j * kz * z) * np.exp(1j * om * t)
    ezz = np.real(a * jv(0, K * rho))
    erho = np.real(a * (-1j) * kz / K * jvp(0, K * rho))
    hphi = np.real(a * (-1j) * om / K * jvp(0, K * rho))
    zero = np.zeros_like(x)
    H = np.stack([-np.sin(phi) * hphi, np.cos(phi) * hphi, zero])
    E = np.stack([np.cos(phi) * erho, np.sin(phi) * erho, ezz])
    return H, E
