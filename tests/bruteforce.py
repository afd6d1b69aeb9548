"""All-pairs float64 evaluation of the per-particle SPH sums, written with numpy broadcasting
and no cell lists. It pins the oracle's neighbour machinery (which cells are visited,
boundary cells, self terms, the r^2 cut-offs) on small scenes: every pair within the
kernel radius is summed regardless of cells, which agrees with the 27-cell walk because
cell size = large kernel length (SPH-3D-2/sph_system.cpp:24).

Formulas follow the reference kernels: density SPH-3D-2/sph_system.cu:338-677, force
:679-1137, energy :1169-1546 (2D: SPH-2D/sph_system.cu:268-460, :462-790, :827-1065)."""
import math

import numpy as np


def _lv(level):
    return np.where(level == 1, 0, 1)


def _params(p, variant):
    g = lambda k: float(p[k][0])
    if variant == 0:
        return dict(h=np.array([g("large_kernel"), g("small_kernel")]),
                    m=np.array([g("large_mass"), g("small_mass")]),
                    poly6=np.array([g("large_poly6"), g("small_poly6")]),
                    spiky=np.array([g("large_spiky"), g("small_spiky")]),
                    visco=np.array([g("large_visco"), g("small_visco")]),
                    grad=np.array([g("large_grad_poly6"), g("small_grad_poly6")]),
                    lapc=np.array([g("large_lapc_poly6"), g("small_lapc_poly6")]),
                    k2=np.array([g("large_kernel2"), g("small_kernel2")]),
                    k6=np.array([g("large_kernel6"), g("small_kernel6")]),
                    rest=g("rest_density"), gas=g("gas_constant"), visc=g("viscosity"),
                    st=g("surface_normal"), coef=float(np.float32(0.04)),
                    eps=float(np.float32(1e-12)), dim=2)
    return dict(h=np.array([g("large_kernel"), g("small_kernel")]),
                m=np.array([g("large_mass"), g("small_mass")]),
                poly6=np.array([g("large_poly6"), g("small_poly6")]),
                spiky=np.array([g("large_spiky"), g("small_spiky")]),
                visco=np.array([g("large_visco"), g("small_visco")]),
                grad=np.array([g("large_grad_poly6"), g("small_grad_poly6")]),
                lapc=np.array([g("large_lapc_poly6"), g("small_lapc_poly6")]),
                k2=np.array([g("large_kernel_2"), g("small_kernel_2")]),
                k6=np.array([g("large_kernel_6"), g("small_kernel_6")]),
                rest=g("rest_density"), gas=g("gas_constant"), visc=g("viscosity"),
                st=g("surface_tension"),
                coef=float(np.float32(0.02)) if variant == 2 else float(np.float32(0.1)),
                eps=float(np.float32(1e-4)), dim=3)


def density(p, variant, pos, level):
    """Density and pressure of every particle (float64)."""
    q = _params(p, variant)
    L = _lv(level)
    x = pos.astype(np.float64)
    d = x[None, :, :] - x[:, None, :]
    r2 = (d * d).sum(-1)
    k2i = q["k2"][L][:, None]
    mj = q["m"][L][None, :]
    w = np.where((r2 < k2i) & (r2 >= q["eps"]), (k2i - r2) ** 3, 0.0)
    np.fill_diagonal(w, 0.0)
    dens = q["poly6"][L] * ((mj * w).sum(1) + q["m"][L] * q["k6"][L])
    press = ((dens / q["rest"]) ** 3 - 1.0) * q["gas"]
    return dens, press


def force(p, variant, pos, ev, dens, press, level):
    """Acceleration (before the 2D mouse force) and |grad c| of every particle (float64)."""
    q = _params(p, variant)
    dim = q["dim"]
    L = _lv(level)
    x = pos.astype(np.float64)[:, :dim]
    e = ev.astype(np.float64)[:, :dim]
    rel = x[:, None, :] - x[None, :, :]          # p_i - p_j
    r2 = (rel * rel).sum(-1)
    hi, hj = q["h"][L][:, None], q["h"][L][None, :]
    k2i, k2j = q["k2"][L][:, None], q["k2"][L][None, :]
    mi, mj = q["m"][L][:, None], q["m"][L][None, :]
    di, dj = dens[:, None], dens[None, :]
    pi_, pj = press[:, None], press[None, :]
    inter = (r2 < np.maximum(k2i, k2j)) & (r2 >= float(np.float32(1e-4)))
    np.fill_diagonal(inter, False)
    r = np.sqrt(np.where(inter, r2, 1.0))
    ikr, jkr = hi - r, hj - r
    ion, jon = inter & (ikr > 0), inter & (jkr > 0)
    ipk = np.where(ion, q["spiky"][L][:, None] * ikr * ikr / q["k6"][L][:, None], 0.0)
    ivk = np.where(ion, q["visco"][L][:, None] / q["k6"][L][:, None] * ikr, 0.0)
    jpk = np.where(jon, q["spiky"][L][None, :] * jkr * jkr / q["k6"][L][None, :], 0.0)
    jvk = np.where(jon, q["visco"][L][None, :] / q["k6"][L][None, :] * jkr, 0.0)
    tp = np.where(inter, mi * mj * (pi_ / di ** 2 + pj / dj ** 2) * (ipk + jpk) / 2, 0.0)
    tv = np.where(inter, (mi / di) * (mj / dj) * q["visc"] * (ivk + jvk) / 2, 0.0)
    f = -(rel * (tp / r)[..., None]).sum(1) + ((e[None, :, :] - e[:, None, :]) * tv[..., None]).sum(1)
    dd = k2i - r2
    if dim == 3:
        tg = np.where(ion, -q["grad"][L][:, None] * mj / dj * dd ** 2, 0.0)
        lap = np.where(ion, q["lapc"][L][:, None] * mj / dj * dd * r2, 0.0).sum(1)
    else:
        tg = np.where(ion, -q["grad"][L][:, None] * (mj / dj) * dd ** 2, 0.0)
        lap = np.where(ion, q["lapc"][L][:, None] * (mj / dj) * dd * r2, 0.0).sum(1)
    g = (tg[..., None] * rel).sum(1)
    nrm = np.sqrt((g * g).sum(1))
    on = nrm > q["st"]
    f = f + np.where(on[:, None], q["coef"] * lap[:, None] * g / np.where(on, nrm, 1.0)[:, None], 0.0)
    return f, nrm


def wavelet(x):
    A = 2.0 / (float(np.float32(3.141592)) ** 0.25 * math.sqrt(6.0))
    return A * (x * x / 4.0 - 1.0) * np.power(float(np.float32(2.718281)), -x * x / 8.0)


def energy(p, variant, pos, vel, level):
    """Wavelet 'energy' of every particle (float64); NaN for a particle without neighbours."""
    q = _params(p, variant)
    dim = q["dim"]
    L = _lv(level)
    x = pos.astype(np.float64)[:, :dim]
    v = vel.astype(np.float64)[:, :dim]
    rel = x[:, None, :] - x[None, :, :]
    r2 = (rel * rel).sum(-1)
    if dim == 3:
        ker = np.full(len(x), q["h"][0])
        k2 = np.full(len(x), q["k2"][0])
        inter = (r2 < k2[:, None]) & (r2 >= float(np.float32(1e-4)))
    else:
        ker, k2 = q["h"][L], q["k2"][L]
        inter = r2 < k2[:, None]
    np.fill_diagonal(inter, False)
    t = np.where(inter[..., None], wavelet(rel / ker[:, None, None]), 0.0)
    tt = t.sum(1)
    tv = (v[None, :, :] * t).sum(1) / np.sqrt(ker)[:, None]
    with np.errstate(invalid="ignore", divide="ignore"):
        u = tv / tt
    e2 = u[:, 0] ** 2 + u[:, 1] ** 2
    m = q["m"][L] if variant == 2 else 1.0
    return 0.5 * m * e2
