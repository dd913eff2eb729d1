"""An independent NumPy evaluation of the DG right-hand side in STRONG form,
used to pin the CPU oracle (which follows the reference's WEAK form).

For LGL collocation the summation-by-parts identity D^T W + W D = B
(B = diag(-1, 0, ..., 0, 1)) turns the weak form of the reference
(euler3d_set2c.jl:144-348)

    rhs_i = sum_d sum_n D[n, i] (MJ (grad xi_d . f))_n  -  sum_faces sMJ F*

into

    rhs_i = -w_i sum_d (D_d G_d)_i + sum_faces sMJ (n . f(q_M) - F*),
    G_d = J grad(xi_d) . f,

which this module evaluates with dense tensor contractions over all
elements at once.  The two agree to rounding only if the oracle's fluxes,
signs, index maps and face normals are right.
"""
import numpy as np


def _fluxes(eqn, dim, q, phys, y=None):
    """Physical flux f[c][s] (c = x, y[, z]) at every node; q (Np, ns, nelem);
    y the node heights (the total-energy pressure needs them)."""
    if eqn == 2:  # total energy (set3c): (U, V, rho, E)
        U, V, rho, E = (q[:, k, :] for k in range(4))
        P = (phys.R_gas / phys.c_v) * (E - 0.5 * (U * U + V * V) / rho - rho * phys.gravity * y)
        out = []
        for c, m in enumerate((U, V)):
            u = m / rho
            mom = [U * u + (P if c == 0 else 0.0), V * u + (P if c == 1 else 0.0)]
            out.append(mom + [m, (E + P) * u])
        return out, P
    if eqn == 0:
        if dim == 3:
            U, V, W, rho, E = (q[:, k, :] for k in range(5))
            vel = (U, V, W)
        else:
            U, V, rho, E = (q[:, k, :] for k in range(4))
            vel = (U, V)
        P = phys.p0 * (phys.R_gas * E / phys.p0) ** (phys.c_p / phys.c_v)
        out = []
        for c in range(dim):
            u = vel[c] / rho
            mom = [vel[k] * u + (P if k == c else 0.0) for k in range(dim)]
            out.append(mom + [vel[c], E * u])  # (mom..., rho, E) in state order
        return out, P
    # shallow water: (U, V, h, b)
    U, V, h, b = (q[:, k, :] for k in range(4))
    H = h + b
    g, dnl = phys.swe_gravity, phys.delta_nl
    u, v = U / H, V / H
    fx = [(H * u * u + 0.5 * g * h * h) * dnl + g * h * b, H * v * u * dnl, U, 0 * h]
    fy = [H * u * v * dnl, (H * v * v + 0.5 * g * h * h) * dnl + g * h * b, V, 0 * h]
    return [fx, fy], None


def _nsr(eqn, dim):
    return 3 if eqn == 1 else dim + 2


def _deriv(D, F, d, dim, Nq):
    """(D_d F): derivative along reference direction d of F (Np, nelem)."""
    nel = F.shape[-1]
    A = F.reshape((Nq,) * dim + (nel,), order="F")
    if d == 0:
        R = np.einsum("in,n...->i...", D, A)
    elif d == 1:
        R = np.einsum("jn,in...->ij...", D, A)
    else:
        R = np.einsum("kn,ijn...->ijk...", D, A)
    return R.reshape(F.shape, order="F")


def rhs_strong(s, Q, include_source=True):
    """Strong-form RHS (no rhs accumulator) for a problems.Setup."""
    eqn, dim, N = s.eqn, s.dim, s.N
    Nq, Np = N + 1, (N + 1) ** dim
    Nfp, nface = Nq ** (dim - 1), 2 * dim
    nsr = _nsr(eqn, dim)
    vg = s.vgeo
    iMJ = 9 if dim == 3 else 4
    MJ = vg[:, iMJ, :]
    w = s.w
    if dim == 3:
        Wt = np.einsum("k,j,i->ijk", w, w, w).reshape(-1, order="F")
    else:
        Wt = np.einsum("j,i->ij", w, w).reshape(-1, order="F")
    J = MJ / Wt[:, None]
    ynode = vg[:, 7, :] if eqn == 2 else None
    f, P = _fluxes(eqn, dim, Q, s.phys, ynode)
    rhs = np.zeros((Np, Q.shape[1], Q.shape[2]), order="F")
    for st in range(nsr):
        div = np.zeros_like(J)
        for d in range(dim):
            G = J * sum(vg[:, d + dim * c, :] * f[c][st] for c in range(dim))
            div += _deriv(s.D, G, d, dim, Nq)
        rhs[:, st, :] = -Wt[:, None] * div
    if include_source and eqn in (0, 2):
        iW = dim - 1
        irho = dim
        rhs[:, iW, :] -= MJ * Q[:, irho, :] * s.phys.gravity
    # faces
    nel = Q.shape[2]
    vM = s.vmapM.reshape(-1, order="F") - 1
    vP = s.vmapP.reshape(-1, order="F") - 1
    sg = s.sgeo.reshape(s.sgeo.shape[0], -1, order="F")
    nrm = sg[:dim]
    sMJ = sg[dim]
    bnd = np.repeat(s.elemtobndy.reshape(-1, order="F"), Nfp)
    ns = Q.shape[1]
    # node-major view: Qflat[global_node, state]
    Qflat = np.stack([Q[:, k, :].reshape(-1, order="F") for k in range(ns)], axis=1)
    qM = Qflat[vM]
    qP = Qflat[vP].copy()
    if eqn == 1:  # the reference reads the P-side b from the M element
        eM = np.repeat(np.arange(nel), nface * Nfp)
        qP[:, 3] = Qflat[eM * Np + vP % Np, 3]
    wall = bnd == 1
    if np.any(wall):
        mom = slice(0, dim)
        un = np.sum(nrm.T[wall] * qM[wall][:, mom], axis=1)
        qP[wall] = qM[wall]
        qP[wall, mom] = qM[wall][:, mom] - 2 * un[:, None] * nrm.T[wall]

    yflat = vg[:, 7, :].reshape(-1, order="F") if eqn == 2 else None

    def fl(qq, yy=None):
        ff, PP = _fluxes(eqn, dim, qq[:, :, None], s.phys, None if yy is None else yy[:, None])
        flat = [[np.asarray(x).reshape(-1) for x in fc] for fc in ff]
        return flat, (None if PP is None else np.asarray(PP).reshape(-1))

    fM, PM = fl(qM, None if yflat is None else yflat[vM])
    # a wall's P state is the mirrored M state at the M node's height
    yPn = None if yflat is None else np.where(wall, yflat[vM], yflat[vP])
    fPP, PPP = fl(qP, yPn)
    if eqn in (0, 2):
        irho = dim
        unM = np.sum(nrm.T * qM[:, :dim], axis=1) / qM[:, irho]
        unP = np.sum(nrm.T * qP[:, :dim], axis=1) / qP[:, irho]
        lam = np.maximum(np.abs(unM) + np.sqrt(s.phys.gamma * PM / qM[:, irho]),
                         np.abs(unP) + np.sqrt(s.phys.gamma * PPP / qP[:, irho]))
    else:
        g, dnl = s.phys.swe_gravity, s.phys.delta_nl
        HM, HP = qM[:, 2] + qM[:, 3], qP[:, 2] + qP[:, 3]
        unM = np.sum(nrm.T * qM[:, :2], axis=1) / HM
        unP = np.sum(nrm.T * qP[:, :2], axis=1) / HP
        lam = np.maximum((np.abs(unM) + np.sqrt(g * HM)) * dnl + np.sqrt(g * qM[:, 3]) * (1 - dnl),
                         (np.abs(unP) + np.sqrt(g * HP)) * dnl + np.sqrt(g * qP[:, 3]) * (1 - dnl))
    for st in range(nsr):
        nfM = sum(nrm[c] * fM[c][st] for c in range(dim))
        nfP = sum(nrm[c] * fPP[c][st] for c in range(dim))
        Fstar = (nfM + nfP - lam * (qP[:, st] - qM[:, st])) / 2
        contrib = sMJ * (nfM - Fstar)
        acc = np.zeros(Np * nel)
        np.add.at(acc, vM, contrib)
        rhs[:, st, :] += acc.reshape(Np, nel, order="F")
    return rhs
