"""Pure-PyTorch FP64 restatement of one RK2 step of the hydro flux path on a single
periodic block (test infrastructure: the FP64 reference gradient comes from torch
autograd through this code).  Same formulas and operand order as oracle/hydro_oracle.c;
branches are torch.where selections, so autograd follows the active branch exactly like
the hand-written adjoint.

State layout: a tensor (NV, nz, ny, nx) of interior cells (no ghosts), periodic.
"""
import math

import numpy as np
import torch

C_LIGHT = 2.99792458e10
C2L = C_LIGHT * C_LIGHT
SMALL = 10.0 * 2.2250738585072014e-308


class _CSqrt(torch.autograd.Function):
    """Correctly rounded sqrt (torch's CPU float64 sqrt is not, on this build)."""

    @staticmethod
    def forward(ctx, x):
        y = torch.from_numpy(np.sqrt(x.detach().cpu().numpy())).to(x.device)
        ctx.save_for_backward(y)
        return y

    @staticmethod
    def backward(ctx, g):
        (y,) = ctx.saved_tensors
        return g / (2.0 * y)


def csqrt(x):
    return _CSqrt.apply(x)


def SIGN(x):
    return torch.where(x < 0.0, -1.0, 1.0).to(x.dtype)


def mn(a, b):
    return torch.where(b < a, b, a)


def mx(a, b):
    return torch.where(a < b, b, a)


def cons_to_prim(u, gm1e, dfloor, pfloor):
    u_d = torch.where(u[0] > dfloor, u[0], torch.full_like(u[0], dfloor))
    di = 1.0 / u_d
    w = [u_d, u[1] * di, u[2] * di, u[3] * di]
    e_k = 0.5 * di * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3])
    sie = (u[4] - e_k) / u_d
    w_p = mx(torch.full_like(u_d, SMALL), gm1e * u_d * sie)
    # floors are assumed inactive in the reference-gradient tests
    bulk = mx(torch.full_like(u_d, SMALL), gm1e * (gm1e + 1.0) * u_d * sie)
    gc = bulk / w_p
    ge = w_p / (u[4] - e_k) + 1
    w.append(w_p)
    for n in range(5, u.shape[0]):
        w.append(u[n] * di)
    return torch.stack(w), torch.stack([gc, ge])


def plm(qm, q0, qp):
    dql = q0 - qm
    dqr = qp - q0
    dq2 = dql * dqr
    dqm = torch.where(dq2 > 0.0, dq2 / torch.where(dq2 > 0.0, dql + dqr, torch.ones_like(dq2)),
                      torch.zeros_like(dq2))
    return q0 + dqm, q0 - dqm


def ppm(q0, q1, q2, q3, q4):
    C2 = 1.25
    qa = q2 - q1
    qb = q3 - q2
    dd_im1 = 0.5 * qa + 0.5 * (q1 - q0)
    dd = 0.5 * qb + 0.5 * qa
    dd_ip1 = 0.5 * (q4 - q3) + 0.5 * qb
    dph = 0.5 * (q1 + q2) + (dd_im1 - dd) / 6.0
    dph_ip1 = 0.5 * (q2 + q3) + (dd - dd_ip1) / 6.0
    d2qc_im1 = q0 + q2 - 2.0 * q1
    d2qc = q1 + q3 - 2.0 * q2
    d2qc_ip1 = q2 + q4 - 2.0 * q3
    qa_tmp = dph - q1
    qb_tmp = q2 - dph
    A = 3.0 * (q1 + q2 - 2.0 * dph)
    cond = (SIGN(A) == SIGN(d2qc_im1)) & (SIGN(A) == SIGN(d2qc))
    qd = torch.where(cond, SIGN(A) * mn(C2 * torch.abs(d2qc_im1),
                                        mn(C2 * torch.abs(d2qc), torch.abs(A))),
                     torch.zeros_like(A))
    dph_tmp = 0.5 * (q1 + q2) - qd / 6.0
    dph = torch.where(qa_tmp * qb_tmp < 0.0, dph_tmp, dph)
    qa_tmp = dph_ip1 - q2
    qb_tmp = q3 - dph_ip1
    A = 3.0 * (q2 + q3 - 2.0 * dph_ip1)
    cond = (SIGN(A) == SIGN(d2qc)) & (SIGN(A) == SIGN(d2qc_ip1))
    qd = torch.where(cond, SIGN(A) * mn(C2 * torch.abs(d2qc),
                                        mn(C2 * torch.abs(d2qc_ip1), torch.abs(A))),
                     torch.zeros_like(A))
    dphip1_tmp = 0.5 * (q2 + q3) - qd / 6.0
    dph_ip1 = torch.where(qa_tmp * qb_tmp < 0.0, dphip1_tmp, dph_ip1)
    d2qf = 6.0 * (dph + dph_ip1 - 2.0 * q2)
    qr = dph
    ql = dph_ip1
    dqf_minus = q2 - qr
    dqf_plus = ql - q2
    qa_tmp = dqf_minus * dqf_plus
    qb_tmp = (q3 - q2) * (q2 - q1)
    cond = ((SIGN(d2qc_im1) == SIGN(d2qc)) & (SIGN(d2qc_im1) == SIGN(d2qc_ip1)) &
            (SIGN(d2qc_im1) == SIGN(d2qf)))
    qe = torch.where(cond, SIGN(d2qf) * mn(mn(C2 * torch.abs(d2qc_im1), C2 * torch.abs(d2qc)),
                                           mn(C2 * torch.abs(d2qc_ip1), torch.abs(d2qf))),
                     torch.zeros_like(d2qf))
    amax = mx(torch.abs(q1), torch.abs(q0))
    bmax = mx(mx(torch.abs(q2), torch.abs(q3)), torch.abs(q4))
    on = torch.abs(d2qf) > 1.0e-12 * mx(amax, bmax)
    rho = torch.where(on, qe / torch.where(on, d2qf, torch.ones_like(d2qf)),
                      torch.zeros_like(d2qf))
    tmp_m = q2 - rho * dqf_minus
    tmp_p = q2 + rho * dqf_plus
    tmp2_m = q2 - 2.0 * dqf_plus
    tmp2_p = q2 + 2.0 * dqf_minus
    ext = (qa_tmp <= 0.0) | (qb_tmp <= 0.0)
    lim = ext & (rho <= (1.0 - 1.0e-12))
    qr_e = torch.where(lim, tmp_m, qr)
    ql_e = torch.where(lim, tmp_p, ql)
    o1 = torch.abs(dqf_minus) >= 2.0 * torch.abs(dqf_plus)
    o2 = torch.abs(dqf_plus) >= 2.0 * torch.abs(dqf_minus)
    qr_o = torch.where(o1, tmp2_m, qr)
    ql_o = torch.where(o2, tmp2_p, ql)
    return torch.where(ext, ql_e, ql_o), torch.where(ext, qr_e, qr_o)


def sound(g, ig, rho, p):
    return C_LIGHT * csqrt(g / (1 + (ig + rho * C2L / p)))


def _vmag(w):
    """|v| with a zero derivative at v = 0 (as the hand-written LHLLC adjoint takes it)."""
    s = w[1] * w[1] + w[2] * w[2] + w[3] * w[3]
    pos = s > 0.0
    return torch.where(pos, csqrt(torch.where(pos, s, torch.ones_like(s))), torch.zeros_like(s))


def hllc(wl, wr, gl_, gr_, lowmach=False):
    gammal, gammar = gl_[0], gr_[0]
    igl = 1.0 / (gl_[1] - 1)
    igr = 1.0 / (gr_[1] - 1)
    cl = sound(gammal, igl, wl[0], wl[4])
    cr = sound(gammar, igr, wr[0], wr[4])
    el = wl[4] * igl + 0.5 * wl[0] * (wl[1] * wl[1] + wl[2] * wl[2] + wl[3] * wl[3])
    er = wr[4] * igr + 0.5 * wr[0] * (wr[1] * wr[1] + wr[2] * wr[2] + wr[3] * wr[3])
    rhoa = .5 * (wl[0] + wr[0])
    ca = .5 * (cl + cr)
    pvrs = .5 * (wr[1] - wl[1]) * rhoa * ca
    if lowmach:  # hydro_lhllc.hpp:70-79: phi = chi (2 - chi), chi = max(|v_L|, |v_R|)
        chi = mx(_vmag(wl), _vmag(wr))
        pvrs = pvrs * (chi * (2.0 - chi))
    pstar = .5 * (wl[4] + wr[4]) - pvrs
    K = (gammal + 1) / (2 * gammar)
    one = torch.ones_like(pstar)
    ql = torch.where(pstar <= wl[4], one,
                     csqrt(torch.where(pstar <= wl[4], one, 1.0 + K * (pstar / wl[4] - 1.0))))
    qr = torch.where(pstar <= wr[4], one,
                     csqrt(torch.where(pstar <= wr[4], one, 1.0 + K * (pstar / wr[4] - 1.0))))
    sl = wl[1] - cl * ql
    sr = wr[1] + cr * qr
    bp = torch.where(sr > 0.0, sr, torch.full_like(sr, 1.0e-20))
    bm = torch.where(sl < 0.0, sl, torch.full_like(sl, -1.0e-20))
    vxl = wl[1] - sl
    vxr = wr[1] - sr
    tl = wl[4] + vxl * wl[0] * wl[1]
    tr = wr[4] + vxr * wr[0] * wr[1]
    ml = wl[0] * vxl
    mr = -(wr[0] * vxr)
    sstar = (tl - tr) / (ml + mr)
    cp = (ml * tr + mr * tl) / (ml + mr)
    cp = torch.where(cp > 0.0, cp, torch.zeros_like(cp))
    vxl = wl[1] - bm
    vxr = wr[1] - bp
    fl = [wl[0] * vxl, wl[0] * wl[1] * vxl + wl[4], wl[0] * wl[2] * vxl, wl[0] * wl[3] * vxl,
          el * vxl + wl[4] * wl[1]]
    fr = [wr[0] * vxr, wr[0] * wr[1] * vxr + wr[4], wr[0] * wr[2] * vxr, wr[0] * wr[3] * vxr,
          er * vxr + wr[4] * wr[1]]
    left = sstar >= 0.0
    zero = torch.zeros_like(sstar)
    ul = torch.where(left, sstar / (sstar - bm), zero)
    ur = torch.where(left, zero, -sstar / (bp - sstar))
    ust = torch.where(left, -bm / (sstar - bm), bp / (bp - sstar))
    return [ul * fl[0] + ur * fr[0], ul * fl[1] + ur * fr[1] + ust * cp, ul * fl[2] + ur * fr[2],
            ul * fl[3] + ur * fr[3], ul * fl[4] + ur * fr[4] + ust * cp * sstar]


def hlle(wl, wr, gl_, gr_):
    gammal, gammar = gl_[0], gr_[0]
    igl = 1.0 / (gl_[1] - 1)
    igr = 1.0 / (gr_[1] - 1)
    cl = sound(gammal, igl, wl[0], wl[4])
    cr = sound(gammar, igr, wr[0], wr[4])
    el = wl[4] * igl + 0.5 * wl[0] * (wl[1] * wl[1] + wl[2] * wl[2] + wl[3] * wl[3])
    er = wr[4] * igr + 0.5 * wr[0] * (wr[1] * wr[1] + wr[2] * wr[2] + wr[3] * wr[3])
    rhoa = .5 * (wl[0] + wr[0])
    ca = .5 * (cl + cr)
    pmid = .5 * (wl[4] + wr[4] + (wl[1] - wr[1]) * rhoa * ca)
    one = torch.ones_like(pmid)
    ql = torch.where(pmid <= wl[4], one,
                     1.0 + (gammal + 1) / csqrt(2 * gammal) * (pmid / wl[4] - 1.0))
    qr = torch.where(pmid <= wr[4], one,
                     1.0 + (gammar + 1) / csqrt(2 * gammar) * (pmid / wr[4] - 1.0))
    al = wl[1] - cl * ql
    ar = wr[1] + cr * qr
    bp = torch.where(ar > 0.0, ar, torch.zeros_like(ar))
    bm = torch.where(al < 0.0, al, torch.zeros_like(al))
    vxl = wl[1] - bm
    vxr = wr[1] - bp
    fl = [wl[0] * vxl, wl[0] * wl[1] * vxl, wl[0] * wl[2] * vxl, wl[0] * wl[3] * vxl]
    fr = [wr[0] * vxr, wr[0] * wr[1] * vxr, wr[0] * wr[2] * vxr, wr[0] * wr[3] * vxr]
    fl[1] = fl[1] + wl[4]
    fr[1] = fr[1] + wr[4]
    fl.append(el * vxl + wl[4] * wl[1])
    fr.append(er * vxr + wr[4] * wr[1])
    split = bp != bm
    tmp = torch.where(split, 0.5 * (bp + bm) / torch.where(split, bp - bm, one),
                      torch.zeros_like(bp))
    return [0.5 * (fl[n] + fr[n]) + (fl[n] - fr[n]) * tmp for n in range(5)]


def riemann_fn(riemann):
    if riemann == "hllc":
        return hllc
    if riemann == "lhllc":
        return lambda wl, wr, gl_, gr_: hllc(wl, wr, gl_, gr_, lowmach=True)
    return hlle


def recon_fn(recon, tk):
    """(ql, qr) of the cells at offset 0 from the stencil accessor tk(k)."""
    if recon == "ppm":
        return ppm(tk(-2), tk(-1), tk(0), tk(1), tk(2))
    if recon == "plm":
        return plm(tk(-1), tk(0), tk(1))
    q = tk(0)
    return q, q


# the upstream low-storage stage tables (gam0, gam1, beta) and vl2's first-order predictor
STAGES = {"rk1": [(0.0, 1.0, 1.0)],
          "rk2": [(0.0, 1.0, 1.0), (0.5, 0.5, 0.5)],
          "rk3": [(0.0, 1.0, 1.0), (0.25, 0.75, 0.25), (2.0 / 3.0, 1.0 / 3.0, 2.0 / 3.0)],
          "vl2": [(0.0, 1.0, 0.5), (0.0, 1.0, 1.0)]}


def stage(u_in, u_n, gam0, gam1, beta_dt, recon, riemann, eos, dx, ndim):
    """One fused stage on a periodic block: u (NV, nz, ny, nx) tensors."""
    gm1e, dfloor, pfloor = eos
    nv = u_in.shape[0]
    w, gm = cons_to_prim(u_in, gm1e, dfloor, pfloor)
    q = torch.cat([w, gm])  # (NV+2, ...)
    area = (dx[1] * dx[2], dx[0] * dx[2], dx[0] * dx[1])
    vol = dx[0] * dx[1] * dx[2]
    du = None
    for d in range(ndim):
        ax = 3 - d  # tensor axis of spatial axis d
        sh = lambda k: torch.roll(q, shifts=-k, dims=ax)  # value at cell i+k
        ql, qr = recon_fn(recon, sh)
        # face i+1/2 of cell i: left state ql[i], right state qr[i+1]
        WL = ql
        WR = torch.roll(qr, shifts=-1, dims=ax)
        ivx, ivy, ivz = 1 + d, 1 + (1 + d) % 3, 1 + (2 + d) % 3
        wl = [WL[0], WL[ivx], WL[ivy], WL[ivz], WL[4]]
        wr = [WR[0], WR[ivx], WR[ivy], WR[ivz], WR[4]]
        f = riemann_fn(riemann)(wl, wr, WL[nv:], WR[nv:])
        F = [None] * nv
        F[0], F[ivx], F[ivy], F[ivz], F[4] = f
        for n in range(5, nv):
            F[n] = torch.where(F[0] >= 0.0, F[0] * WL[n], F[0] * WR[n])
        F = torch.stack(F)  # flux at face i+1/2, stored at cell i
        Fm = torch.roll(F, shifts=1, dims=ax)  # face i-1/2
        dd = area[d] * F - area[d] * Fm
        du = dd if du is None else du + dd
    div = -du / vol
    # floors inactive: the floored u equals u
    return gam0 * u_in + gam1 * u_n + beta_dt * div


def rk2_step(u, dt, recon, riemann, eos, dx, ndim):
    u1 = stage(u, u, 0.0, 1.0, 1.0 * dt, recon, riemann, eos, dx, ndim)
    return stage(u1, u, 0.5, 0.5, 0.5 * dt, recon, riemann, eos, dx, ndim)


def step(u, dt, integrator, recon, riemann, eos, dx, ndim):
    """One step of an integrator: stage s maps x -> gam0 x + gam1 u + beta dt L(x)."""
    x = u
    for s, (g0, g1, b) in enumerate(STAGES[integrator]):
        r = "dc" if (integrator == "vl2" and s == 0) else recon
        x = stage(x, u, g0, g1, b * dt, r, riemann, eos, dx, ndim)
    return x


def _pad(u, ax, G, kind, d):
    """Pad conserved state u along tensor axis ax by G cells per side with the boundary
    rule of spatial axis d: periodic, outflow (copy edge) or reflect (mirror, negate the
    normal momentum)."""
    n = u.shape[ax]
    if kind == "periodic":
        lo, hi = u.narrow(ax, n - G, G), u.narrow(ax, 0, G)
    elif kind == "outflow":
        lo = u.narrow(ax, 0, 1).expand(*[G if a == ax else -1 for a in range(u.dim())])
        hi = u.narrow(ax, n - 1, 1).expand(*[G if a == ax else -1 for a in range(u.dim())])
    else:
        sign = torch.ones(u.shape[0], 1, 1, 1, dtype=u.dtype, device=u.device)
        sign[1 + d] = -1.0
        lo = torch.flip(u.narrow(ax, 0, G), dims=[ax]) * sign
        hi = torch.flip(u.narrow(ax, n - G, G), dims=[ax]) * sign
    return torch.cat([lo, u, hi], dim=ax)


def stage_bc(u_in, u_n, gam0, gam1, beta_dt, recon, riemann, eos, dx, ndim, bc, g=None):
    """One fused stage on a single mesh with per-axis boundary rules bc[d] in
    {"periodic","outflow","reflect"} (applied to the conserved state before ConsToPrim,
    like the fused kernel) and an optional uniform gravity g (stage-start primitives)."""
    gm1e, dfloor, pfloor = eos
    nv = u_in.shape[0]
    G = 3
    area = (dx[1] * dx[2], dx[0] * dx[2], dx[0] * dx[1])
    vol = dx[0] * dx[1] * dx[2]
    du = None
    for d in range(ndim):
        ax = 3 - d
        n = u_in.shape[ax]
        U = _pad(u_in, ax, G, bc[d], d)
        w, gm = cons_to_prim(U, gm1e, dfloor, pfloor)
        q = torch.cat([w, gm])
        tk = lambda k: q.narrow(ax, G - 1 + k, n + 2)  # stencil offset k of cells -1..n
        ql, qr = recon_fn(recon, tk)
        WL = ql.narrow(ax, 0, n + 1)   # ql output of cell f-1, faces f = 0..n
        WR = qr.narrow(ax, 1, n + 1)   # qr output of cell f
        ivx, ivy, ivz = 1 + d, 1 + (1 + d) % 3, 1 + (2 + d) % 3
        wl = [WL[0], WL[ivx], WL[ivy], WL[ivz], WL[4]]
        wr = [WR[0], WR[ivx], WR[ivy], WR[ivz], WR[4]]
        f = riemann_fn(riemann)(wl, wr, WL[nv:], WR[nv:])
        F = [None] * nv
        F[0], F[ivx], F[ivy], F[ivz], F[4] = f
        for m in range(5, nv):
            F[m] = torch.where(F[0] >= 0.0, F[0] * WL[m], F[0] * WR[m])
        F = torch.stack(F)
        dd = area[d] * F.narrow(ax, 1, n) - area[d] * F.narrow(ax, 0, n)
        du = dd if du is None else du + dd
    div = -du / vol
    out = gam0 * u_in + gam1 * u_n + beta_dt * div
    if g is not None and any(v != 0.0 for v in g):
        w, _ = cons_to_prim(u_in, gm1e, dfloor, pfloor)
        parts = [out[v] for v in range(nv)]
        parts[1] = parts[1] + g[0] * w[0] * beta_dt
        de = g[0] * w[1]
        if ndim >= 2:
            parts[2] = parts[2] + g[1] * w[0] * beta_dt
            de = de + g[1] * w[2]
        if ndim >= 3:
            parts[3] = parts[3] + g[2] * w[0] * beta_dt
            de = de + g[2] * w[3]
        parts[4] = parts[4] + de * w[0] * beta_dt
        out = torch.stack(parts)
    return out


def rk2_step_bc(u, dt, recon, riemann, eos, dx, ndim, bc, g=None):
    u1 = stage_bc(u, u, 0.0, 1.0, 1.0 * dt, recon, riemann, eos, dx, ndim, bc, g)
    return stage_bc(u1, u, 0.5, 0.5, 0.5 * dt, recon, riemann, eos, dx, ndim, bc, g)


def lattice_to_global(bs, field):
    """Interior of a single lattice mesh (pack layout) as one (NV, nz, ny, nx) array."""
    import numpy as np

    nv = bs.nvar
    bx = bs.blocks["nx"][0]
    dx = bs.blocks["dx"][0]
    x0 = bs.blocks["xmin"].min(axis=0)
    pos = np.rint((bs.blocks["xmin"] - x0) / (dx * bx)).astype(int)
    shape = [(pos[:, d].max() + 1) * bx[d] for d in (2, 1, 0)]
    out = np.zeros((nv, *shape))
    for b in range(bs.nblocks):
        lx, ly, lz = pos[b]
        out[:, lz * bx[2]:(lz + 1) * bx[2], ly * bx[1]:(ly + 1) * bx[1],
            lx * bx[0]:(lx + 1) * bx[0]] = bs.interior(field, b, nv)
    return out
