"""Seeded synthetic GRMHD snapshots on an FMKS grid (input data for the
GRMHD imaging path; no dump files are available here).

* Grid: n1 x n2 x n3 cells, x1 = ln r uniform over [ln r_in, ln r_out],
  x2 uniform over [0, 1] with the FMKS theta(x1, x2) map (hslope, mks_smooth,
  poly_alpha, poly_xt; poly_norm as the iharm reader derives it,
  src/grmhd/iharm_loader.cpp:322-326), phi uniform over [0, 2 pi).
* Fluid: the Fishbone & Moncrief (1976) torus with inner edge r_fm_in and
  pressure maximum r_fm_max around a Kerr hole of spin a, polytrope
  u = kappa rho^gam / (gam - 1), density normalised to 1 at its maximum,
  internal energy multiplied by u_heat (the equilibrium torus is far colder
  than an accreted one: without it the R_high = 40 electrons radiate nothing
  at 230 GHz),
  in a power-law atmosphere rho = 1e-5 r^-1.5, u = 1e-7 r^-2.5 at rest with
  respect to the normal observer.  Torus velocities are purely azimuthal in
  Boyer-Lindquist, converted to the native primitive U^i = u^i - u^t g^{0i} / g^{00}.
* Perturbation: rho and u multiplied by exp(sigma G - sigma^2 / 2), G a
  unit-variance Gaussian field with Gaussian correlation length `corr` cells
  (seeded; periodic smoothing by FFT).
* Field: poloidal loops from A_phi = max(rho / rho_max - 0.2, 0),
  B^1 = -d_2 A_phi / sqrt(-g), B^2 = d_1 A_phi / sqrt(-g) (centred
  differences), scaled so that min beta over the torus equals beta_min.
* Primitives are rounded to float32 (the iharm reader's precision); the
  derived scalars are left for the device (or the oracle) to compute.
"""
from __future__ import annotations

import math

import numpy as np

from kpolaris_amd import abi

CELL_DTYPE = np.dtype([("prim", np.float32, (8,)), ("derived", np.float64, (5,))])
assert CELL_DTYPE.itemsize == 72


def poly_norm(poly_alpha: float, poly_xt: float) -> float:
    return 0.5 * math.pi / (1.0 + 1.0 / (poly_alpha + 1.0) / poly_xt ** poly_alpha)


def fmks_theta(x1, x2, startx1, hslope, mks_smooth, poly_alpha, poly_xt, pnorm):
    thG = np.pi * x2 + ((1.0 - hslope) / 2.0) * np.sin(2.0 * np.pi * x2)
    y = 2.0 * x2 - 1.0
    thJ = pnorm * y * (1.0 + np.abs(y / poly_xt) ** poly_alpha * np.sign(y) ** int(round(poly_alpha)) /
                       (poly_alpha + 1.0)) + 0.5 * np.pi
    return thG + np.exp(mks_smooth * (startx1 - x1)) * (thJ - thG)


def fmks_dtheta(x1, x2, startx1, hslope, mks_smooth, poly_alpha, poly_xt, pnorm):
    """d theta / d x1, d theta / d x2."""
    y = 2.0 * x2 - 1.0
    expfac = np.exp(mks_smooth * (startx1 - x1))
    thG = np.pi * x2 + ((1.0 - hslope) / 2.0) * np.sin(2.0 * np.pi * x2)
    za = (y / poly_xt) ** int(round(poly_alpha))
    zam1 = (y / poly_xt) ** int(round(poly_alpha - 1))
    thJ = pnorm * y * (1.0 + za / (poly_alpha + 1.0)) + 0.5 * np.pi
    dthG = np.pi + (1.0 - hslope) * np.pi * np.cos(2.0 * np.pi * x2)
    dthJ = 2.0 * pnorm * (1.0 + za / (poly_alpha + 1.0)) + \
        (2.0 * poly_alpha * pnorm * y * zam1) / ((poly_alpha + 1.0) * poly_xt)
    return -mks_smooth * expfac * (thJ - thG), dthG + expfac * (dthJ - dthG)


def fm_l(r, a):
    """Specific angular momentum of the Fishbone-Moncrief torus with pressure
    maximum at r (FM 1976, eq. 3.8 in the form used by HARM-family codes)."""
    sr = math.sqrt(r)
    num = (a * a - 2.0 * a * sr + r * r) * (
        (-2.0 * a * r * (a * a - 2.0 * a * sr + r * r)) / math.sqrt(2.0 * a * sr + (-3.0 + r) * r)
        + ((a + (-2.0 + r) * sr) * (r ** 3 + a * a * (2.0 + r))) / math.sqrt(1.0 + (2.0 * a) / r ** 1.5 - 3.0 / r))
    den = r ** 3 * math.sqrt(2.0 * a * sr + (-3.0 + r) * r) * (a * a + (-2.0 + r) * r)
    return num / den


def _lnh(r, th, a, l):
    sth, cth = np.sin(th), np.cos(th)
    DD = r * r - 2.0 * r + a * a
    AA = (r * r + a * a) ** 2 - DD * a * a * sth * sth
    SS = r * r + a * a * cth * cth
    q = np.sqrt(1.0 + 4.0 * (l * l * SS * SS) * DD / (AA * AA * sth * sth))
    return 0.5 * np.log((1.0 + q) / (SS * DD / AA)) - 0.5 * q - 2.0 * a * r * l / AA


def gaussian_field(shape, corr: float, rng) -> np.ndarray:
    """Unit-variance Gaussian random field with Gaussian correlation length
    `corr` cells (white noise smoothed in Fourier space)."""
    w = rng.standard_normal(shape)
    f = np.fft.rfftn(w)
    kx = np.fft.fftfreq(shape[0])[:, None, None]
    ky = np.fft.fftfreq(shape[1])[None, :, None]
    kz = np.fft.rfftfreq(shape[2])[None, None, :]
    k2 = kx * kx + ky * ky + kz * kz
    f *= np.exp(-2.0 * (np.pi ** 2) * corr * corr * k2)
    g = np.fft.irfftn(f, s=shape, axes=(0, 1, 2))
    g -= g.mean()
    return g / max(g.std(), 1e-300)


def make_snapshot(n1=288, n2=128, n3=128, *, spin=0.9375, r_in=1.2, r_out=1000.0, r_fm_in=6.0,
                  r_fm_max=12.0, gam=4.0 / 3.0, kappa=1e-3, sigma=0.3, corr=4.0, beta_min=100.0,
                  seed=20261202, hslope=0.3, mks_smooth=0.5, poly_alpha=14.0, poly_xt=0.82,
                  freq_cgs=230e9, mbh_solar=6.5e9, M_unit=1e28, r_high=40.0, r_low=1.0, u_heat=25.0,
                  background_field=0.0):
    """Returns (cells: structured array [n1*n2*n3] of CELL_DTYPE with derived
    scalars zero, model: abi.GRMHDModel without the device pointer).
    background_field > 0 adds a uniform toroidal B^3 of that fraction of the
    largest |B^1| of the loops, so no cell is field-free (the loops vanish
    outside the torus, where the field direction, and with it the emission
    angle, is undefined)."""
    rng = np.random.default_rng(seed)
    np.seterr(invalid="ignore", divide="ignore")  # inside the horizon the torus formulas are masked out
    startx1 = math.log(r_in)
    dx1 = (math.log(r_out) - startx1) / n1
    dx2 = 1.0 / n2
    dx3 = 2.0 * math.pi / n3
    pn = poly_norm(poly_alpha, poly_xt)
    geo = (startx1, hslope, mks_smooth, poly_alpha, poly_xt, pn)
    x1 = startx1 + (np.arange(n1) + 0.5) * dx1
    x2 = (np.arange(n2) + 0.5) * dx2
    X1, X2 = np.meshgrid(x1, x2, indexing="ij")
    r = np.exp(X1)
    th = fmks_theta(X1, X2, *geo)
    dth1, dth2 = fmks_dtheta(X1, X2, *geo)
    a = spin
    l = fm_l(r_fm_max, a)
    lnh = _lnh(r, th, a, l) - _lnh(np.array(r_fm_in), np.array(np.pi / 2), a, l)
    inside = (r >= r_fm_in) & (lnh > 0.0)
    hm1 = np.where(inside, np.expm1(np.where(inside, lnh, 0.0)), 0.0)
    rho_t = np.where(inside, (hm1 * (gam - 1.0) / (kappa * gam)) ** (1.0 / (gam - 1.0)), 0.0)
    rho_max = rho_t.max()
    rho_t /= rho_max
    u_t = u_heat * kappa * (rho_t * rho_max) ** gam / (gam - 1.0) / rho_max
    # Boyer-Lindquist azimuthal velocity of the torus, then u^t from u.u = -1
    sth, cth = np.sin(th), np.cos(th)
    DD = r * r - 2.0 * r + a * a
    AA = (r * r + a * a) ** 2 - DD * a * a * sth * sth
    SS = r * r + a * a * cth * cth
    expm2chi = SS * SS * DD / (AA * AA * sth * sth)
    up1 = np.sqrt((-1.0 + np.sqrt(1.0 + 4.0 * l * l * expm2chi)) / 2.0)
    uphi = 2.0 * a * r * np.sqrt(1.0 + up1 * up1) / np.sqrt(AA * SS * DD) + np.sqrt(SS / AA) * up1 / sth
    f = 2.0 * r / SS
    s2 = sth * sth
    gtt, gtp, gpp = -1.0 + f, -a * f * s2, s2 * (SS + a * a * s2 * (1.0 + f))
    A, B, C = gtt, 2.0 * gtp * uphi, gpp * uphi * uphi + 1.0
    ut = (-B - np.sqrt(np.maximum(B * B - 4.0 * A * C, 0.0))) / (2.0 * A)
    # native g^{0i}/g^{00}: g^{00} = -(1+f), g^{01} = f / r, g^{02} = -f dth1 / (r dth2)
    g00 = -(1.0 + f)
    U1_t = -ut * (f / r) / g00
    U2_t = -ut * (-f * dth1 / (r * dth2)) / g00
    U3_t = uphi
    # atmosphere at rest with respect to the normal observer
    rho_a = 1e-5 * r ** -1.5
    u_a = 1e-7 * r ** -2.5
    shape = (n1, n2, n3)
    pert = np.exp(sigma * gaussian_field(shape, corr, rng) - 0.5 * sigma * sigma)
    ins3 = inside[:, :, None]
    rho = np.where(ins3, rho_t[:, :, None] * pert, rho_a[:, :, None])
    uu = np.where(ins3, u_t[:, :, None] * pert, u_a[:, :, None])
    rho = np.maximum(rho, rho_a[:, :, None])
    uu = np.maximum(uu, u_a[:, :, None])
    U1 = np.where(ins3, U1_t[:, :, None], 0.0) * np.ones(shape)
    U2 = np.where(ins3, U2_t[:, :, None], 0.0) * np.ones(shape)
    U3 = np.where(ins3, U3_t[:, :, None], 0.0) * np.ones(shape)
    # poloidal loops from A_phi (axisymmetric part of the unperturbed torus density)
    Aphi = np.maximum(rho_t - 0.2, 0.0)
    gdet = (SS * sth) * r * dth2
    dA1 = np.gradient(Aphi, dx1, axis=0)
    dA2 = np.gradient(Aphi, dx2, axis=1)
    B1 = (-dA2 / gdet)[:, :, None] * np.ones(shape)
    B2 = (dA1 / gdet)[:, :, None] * np.ones(shape)
    B3 = np.zeros(shape)
    model = grmhd_model(n1, n2, n3, spin=spin, gam=gam, startx1=startx1, dx1=dx1, dx2=dx2, dx3=dx3, r_in=r_in,
                        r_out=r_out, hslope=hslope, mks_smooth=mks_smooth, poly_alpha=poly_alpha, poly_xt=poly_xt,
                        poly_norm_=pn, freq_cgs=freq_cgs, mbh_solar=mbh_solar, M_unit=M_unit, r_high=r_high,
                        r_low=r_low)
    # beta normalisation over torus cells with the native metric at cell centres
    bsq = _bsq_native(r, th, dth1, dth2, a, U1, U2, U3, B1, B2, B3)
    ok = ins3 & (rho > 0.1 * rho.max()) & (bsq > 0.0)
    beta = (gam - 1.0) * uu / (0.5 * np.where(bsq > 0.0, bsq, np.inf))
    bmin = float(beta[ok].min()) if ok.any() else beta_min
    scale = math.sqrt(bmin / beta_min)
    B1 *= scale
    B2 *= scale
    if background_field > 0.0:
        B3 = B3 + background_field * float(np.abs(B1).max())
    cells = np.zeros(n1 * n2 * n3, dtype=CELL_DTYPE)
    prim = np.stack([rho, uu, U1, U2, U3, B1, B2, B3], axis=-1).reshape(-1, 8)
    cells["prim"] = prim.astype(np.float32)
    return cells, model


def advect(cells: np.ndarray, model, t: float) -> np.ndarray:
    """The snapshot at time t of a flow rotating rigidly on each radial shell
    with the prograde Keplerian angular velocity Omega(r) = 1 / (r^1.5 + a):
    every primitive at azimuth phi is the t = 0 value at phi - Omega t
    (periodic linear interpolation in the azimuthal index).  Derived scalars
    are left zero (recompute them on the device or with the oracle)."""
    n1, n2, n3 = model.n1, model.n2, model.n3
    prim = cells["prim"].reshape(n1, n2, n3, 8)
    r = np.exp(model.startx1 + (np.arange(n1) + 0.5) * model.dx1)
    omega = 1.0 / (r ** 1.5 + model.spin)
    shift = omega * t / model.dx3           # cells moved in +phi, per radial index
    k = np.arange(n3)
    out = np.empty_like(prim)
    for i in range(n1):
        s0 = math.floor(shift[i])
        w = shift[i] - s0
        a = prim[i][:, (k - s0) % n3]
        b = prim[i][:, (k - s0 - 1) % n3]
        out[i] = ((1.0 - w) * a.astype(np.float64) + w * b.astype(np.float64)).astype(np.float32)
    res = np.zeros(cells.shape, dtype=CELL_DTYPE)
    res["prim"] = out.reshape(-1, 8)
    return res


def make_sequence(nsnap=32, dt=5.0, t0=0.0, n1=192, n2=96, n3=96, **kw):
    """A slow-light snapshot sequence: the seeded torus of make_snapshot at
    t0, advected by Keplerian rotation to t0 + k dt, k < nsnap.  Returns
    (list of cell arrays, times, model)."""
    base, model = make_snapshot(n1, n2, n3, **kw)
    times = t0 + dt * np.arange(nsnap)
    cells = [base if k == 0 else advect(base, model, dt * k) for k in range(nsnap)]
    return cells, times, model


def _bsq_native(r, th, dth1, dth2, a, U1, U2, U3, B1, B2, B3):
    """b^2 at cell centres (the derived-scalar formulas, vectorised)."""
    cth, sth = np.cos(th), np.sin(th)
    s2 = sth * sth
    rho2 = r * r + a * a * cth * cth
    f = 2.0 * r / rho2
    g = {}
    g[0, 0] = -1.0 + f
    g[0, 1] = f * r
    g[0, 3] = -a * f * s2
    g[1, 1] = (1.0 + f) * r * r + rho2 * dth1 * dth1
    g[1, 2] = rho2 * dth1 * dth2
    g[1, 3] = -a * s2 * (1.0 + f) * r
    g[2, 2] = rho2 * dth2 * dth2
    g[3, 3] = s2 * (rho2 + a * a * s2 * (1.0 + f))
    for (i, j) in list(g):
        g[j, i] = g[i, j]

    def G(i, j):
        v = g.get((i, j))
        return 0.0 if v is None else v[:, :, None]

    gc0 = [-(1.0 + f), f / r, -f * dth1 / (r * dth2), 0.0 * f]
    gc0 = [c[:, :, None] for c in gc0]
    v = [None, U1, U2, U3]
    spatial = sum(G(i, j) * v[i] * v[j] for i in range(1, 4) for j in range(1, 4))
    vfac = np.sqrt(np.maximum(-1.0 / gc0[0] * (1.0 + np.abs(spatial)), 0.0))
    ucon = [-vfac * gc0[0]] + [v[i] - vfac * gc0[i] for i in range(1, 4)]
    ucov = [sum(G(m, n) * ucon[n] for n in range(4)) for m in range(4)]
    Bv = [None, B1, B2, B3]
    udotB = sum(ucov[i] * Bv[i] for i in range(1, 4))
    bcon = [udotB] + [(Bv[i] + ucon[i] * udotB) / np.maximum(ucon[0], 1e-300) for i in range(1, 4)]
    bcov = [sum(G(m, n) * bcon[n] for n in range(4)) for m in range(4)]
    return np.abs(sum(bcon[m] * bcov[m] for m in range(4)))


def grmhd_model(n1, n2, n3, *, spin, gam, startx1, dx1, dx2, dx3, r_in, r_out, hslope, mks_smooth, poly_alpha,
                poly_xt, poly_norm_, freq_cgs, mbh_solar, M_unit, r_high, r_low) -> "abi.GRMHDModel":
    m = abi.grmhd_defaults()
    m.n1, m.n2, m.n3 = n1, n2, n3
    m.spin, m.gam = spin, gam
    m.startx1, m.startx2, m.startx3 = startx1, 0.0, 0.0
    m.dx1, m.dx2, m.dx3 = dx1, dx2, dx3
    m.r_in, m.r_out = r_in, r_out
    m.hslope, m.mks_smooth, m.poly_alpha, m.poly_xt, m.poly_norm = hslope, mks_smooth, poly_alpha, poly_xt, poly_norm_
    m.freq_cgs, m.mbh_solar, m.M_unit = freq_cgs, mbh_solar, M_unit
    m.trat_large, m.trat_small = r_high, r_low
    return m


def pass_a_params(model, *, nx=512, ny=512, inclination_deg=163.0, xspan=10.0, camera_radius=1000.0,
                  adaptive=1, tolerance=1e-6, step=0.1, max_steps=200000, outer_radius=None):
    """Camera and integrator parameters of a GRMHD image on the model's FMKS
    coordinates; inner/outer radii per the reference tool
    (tools/kpolaris_model_image.cpp:468-475, 5120-5125)."""
    p = abi.pass_a_defaults()
    p.camera.nx, p.camera.ny = nx, ny
    p.camera.radius = camera_radius
    p.camera.inclination = math.radians(inclination_deg)
    # pinhole extents as the tool resolves them (tools/kpolaris_model_image_options.cpp:160-180):
    # the half-width xspan (in M) becomes the slope range fov = 2 xspan / R
    p.camera.fov = 2.0 * xspan / camera_radius
    p.camera.fov_y = 2.0 * xspan * ny / nx / camera_radius
    p.camera.xspan = p.camera.yspan = -1.0
    p.camera.model = abi.CAMERA_PINHOLE
    p.coordinate_system = abi.COORD_FMKS
    p.mass, p.spin = 1.0, model.spin
    horizon = 1.0 + math.sqrt(max(0.0, 1.0 - model.spin ** 2))
    p.inner_radius = max(model.r_in, 1.05 * horizon)
    p.outer_radius = min(model.r_out, 0.9 * camera_radius) if outer_radius is None else outer_radius
    p.step, p.max_steps, p.adaptive = step, max_steps, adaptive
    p.adaptive_tolerance, p.min_step, p.max_step = tolerance, 1e-10, 5.0
    p.max_radiation_step, p.max_radiation_depth = 1.0, 1.0
    p.max_absorption_depth, p.max_faraday_depth = 1.0, 4.0
    p.fmks_startx1, p.fmks_hslope, p.fmks_mks_smooth = model.startx1, model.hslope, model.mks_smooth
    p.fmks_poly_alpha, p.fmks_poly_xt, p.fmks_poly_norm = model.poly_alpha, model.poly_xt, model.poly_norm
    return p
