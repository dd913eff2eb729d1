"""Problem setup: mesh -> maps -> geometry -> initial state, as the reference
drivers do it (euler3d_set2c.jl:1049-1110, euler2d_set2c.jl:924-984,
shallow_water.jl:751-800), with every physical constant a parameter.

Everything here is host-side NumPy in the reference's column-major layout.
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from . import mesh as M
from ._lib import check, lib

NSTATE = {(_lib.EULER, 3): 5, (_lib.EULER, 2): 4, (_lib.SWE, 2): 4, (_lib.EULER_TE, 2): 4}
NVGEO = {3: 14, 2: 8}
NSGEO = {3: 5, 2: 4}
IMJ = {3: 9, 2: 4}
IMJI = {3: 10, 2: 5}


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def linrange(a, b, n):
    """Julia range(a, stop=b, length=n) for the endpoints used here:
    element i is the correctly rounded (a*(n-1-i) + b*i)/(n-1) whenever the
    numerator is exact (integer endpoints)."""
    return np.array([(a * (n - 1 - i) + b * i) / (n - 1) for i in range(n)], dtype=np.float64)


@dataclass
class Setup:
    eqn: int
    dim: int
    N: int
    nelem: int
    D: np.ndarray
    r: np.ndarray
    w: np.ndarray
    vgeo: np.ndarray
    sgeo: np.ndarray
    vmapM: np.ndarray
    vmapP: np.ndarray
    elemtobndy: np.ndarray
    mesh: dict = field(repr=False)
    phys: object = None
    Q: np.ndarray = None
    label: str = ""
    Q0: np.ndarray = None  # background at rest for the precision form (None: not defined)

    @property
    def Nq(self):
        return self.N + 1

    @property
    def Np(self):
        return (self.N + 1) ** self.dim

    @property
    def nstate(self):
        return NSTATE[(self.eqn, self.dim)]

    @property
    def nvgeo(self):
        return NVGEO[self.dim]

    def zeros_state(self):
        return np.zeros((self.Np, self.nstate, self.nelem), order="F")

    def coords(self):
        """Node coordinates (Np, nelem) per dimension."""
        ix = 11 if self.dim == 3 else 6
        return tuple(self.vgeo[:, ix + c, :] for c in range(self.dim))

    def ndof(self):
        return self.nelem * self.Np * self.nstate


def build(eqn, dim, N, x, periodic, boundary=None, phys=None, label="", geometry=True):
    """brickmesh -> partition -> connectmesh -> mappings -> lglpoints /
    spectralderivative -> computegeometry (one rank).  geometry=False stops
    before computegeometry (vgeo = sgeo = None): the device path builds them
    on the GPU (to_device)."""
    etv, etc, etb, fcs = M.brickmesh(x, periodic, boundary=boundary)
    etv, etc, etb, fcs = M.partition(etv, etc, etb, fcs)
    mesh = M.connectmesh(etv, etc, etb, fcs)
    vmapM, vmapP = M.mappings(N, mesh["elemtoelem"], mesh["elemtoface"], mesh["elemtoordr"])
    r, w = M.lglpoints(N)
    D = M.spectralderivative(r)
    vgeo, sgeo = M.computegeometry(dim, N, mesh["elemtocoord"], vmapM) if geometry else (None, None)
    return Setup(eqn=eqn, dim=dim, N=N, nelem=etc.shape[2], D=D, r=r, w=w, vgeo=vgeo, sgeo=sgeo,
                 vmapM=vmapM, vmapP=vmapP, elemtobndy=mesh["elemtobndy"], mesh=mesh,
                 phys=phys if phys is not None else _lib.default_phys(), label=label)


# ------------------------------------------------------------- initial data
def bubble_state(s, theta0=300.0, dtheta=0.5, rc=250.0, center=(500.0, 500.0, 350.0),
                 spherical=False, jitter=0.0, seed=0):
    b = _lib.Bubble(theta0=theta0, dtheta=dtheta, rc=rc, xc=center[0], yc=center[1],
                    zc=center[2] if s.dim == 3 else center[-1], spherical=1 if spherical else 0,
                    jitter=jitter, seed=seed)
    Q = s.zeros_state()
    check(lib().canary_ic_bubble(s.dim, s.N, s.nelem, _p(s.vgeo), C.byref(b), C.byref(s.phys),
                                 _p(Q)), "ic_bubble")
    return Q


def vortex_params(beta=5.0, x0=0.0, y0=0.0, u0=1.0, v0=1.0, rho_inf=1.0, p_inf=1.0, gamma=1.4,
                  Lx=10.0, Ly=10.0):
    return _lib.Vortex(beta=beta, x0=x0, y0=y0, u0=u0, v0=v0, rho_inf=rho_inf, p_inf=p_inf,
                       gamma=gamma, Lx=Lx, Ly=Ly)


def vortex_state(s, t=0.0, params=None):
    params = params or vortex_params()
    Q = s.zeros_state()
    check(lib().canary_ic_vortex2d(s.N, s.nelem, _p(s.vgeo), C.byref(params), float(t), _p(Q)),
          "ic_vortex2d")
    return Q


def uniform_rest_state(s, rho, p):
    """Euler state at rest with density rho and pressure p everywhere:
    rho*theta from P = p0 (R rho theta / p0)^(c_p / c_v)."""
    ph = s.phys
    rt = (ph.p0 / ph.R_gas) * (p / ph.p0) ** (ph.c_v / ph.c_p)
    Q = s.zeros_state()
    Q[:, s.dim, :] = rho
    Q[:, s.dim + 1, :] = rt
    return Q


def swe_state(s, bumps, depth=10.0, L=(1.0, 1.0)):
    bumps = np.ascontiguousarray(np.asarray(bumps, dtype=np.float64).reshape(-1, 4))
    Q = s.zeros_state()
    check(lib().canary_ic_swe_bumps(s.N, s.nelem, _p(s.vgeo), float(depth), len(bumps),
                                    _p(bumps), float(L[0]), float(L[1]), _p(Q)), "ic_swe_bumps")
    return Q


def cfl(s, Q):
    """Reference cfl() value (before the `/ N^sqrt(2)`)."""
    Qf = np.asfortranarray(Q)
    return lib().canary_cfl(s.eqn, s.dim, s.N, _p(s.vgeo), _p(Qf), s.nelem, C.byref(s.phys))


def swe_courant(s, Q, advection=False):
    """Time step of the shallow-water example (swe2d.jl:100-124):
    min over nodes of 0.5 dx / wave_speed / N with
    dx = sqrt((1/(2 xi_x))^2 + (1/(2 eta_y))^2) and
    wave_speed = |u| + delta_wave sqrt(g H) delta_nl + g b (1 - delta_nl)."""
    Q = np.asarray(Q)
    g, dnl = s.phys.swe_gravity, s.phys.delta_nl
    h, b, U, V = Q[:, 2, :], Q[:, 3, :], Q[:, 0, :], Q[:, 1, :]
    xix, etay = s.vgeo[:, 0, :], s.vgeo[:, 3, :]
    H = h + b
    u, v = U / H, V / H
    dx = np.sqrt((1.0 / (2 * xix)) ** 2 + (1.0 / (2 * etay)) ** 2)
    vel = np.sqrt(u ** 2 + v ** 2)
    dwave = 0.0 if advection else 1.0
    wave = vel + dwave * np.sqrt(g * H) * dnl + g * b * (1 - dnl)
    return float(np.min(0.5 * dx / wave / s.N))


def euler_courant(s, Q, courant=0.5):
    """Courant-number time step for the Euler sets, in the form of the
    shallow-water example's courantnumber (swe2d.jl:100-124):
    min over nodes of courant * dx / wave_speed / N with
    dx = sqrt(sum_d (1/(2 xi^d_d))^2) and wave_speed = |u| + c,
    P = p0 (R rho theta / p0)^(c_p/c_v), c = sqrt(gamma P / rho).  (The
    Euler drivers' cfl() leaves the sound speed unscaled by the metric, which
    is only stable on kilometre-scale boxes.)"""
    Q = np.asarray(Q)
    d, ph = s.dim, s.phys
    rho, rt = Q[:, d, :], Q[:, d + 1, :]
    vel = np.sqrt(sum((Q[:, c, :] / rho) ** 2 for c in range(d)))
    P = ph.p0 * (ph.R_gas * rt / ph.p0) ** (ph.c_p / ph.c_v)
    c = np.sqrt(ph.gamma * P / rho)
    diag = (0, 4, 8) if d == 3 else (0, 3)   # xi_x, eta_y(, zeta_z) in vgeo
    dx = np.sqrt(sum((1.0 / (2 * s.vgeo[:, k, :])) ** 2 for k in diag))
    return float(np.min(courant * dx / (vel + c) / s.N))


def l2energysquared(s, Q):
    Qf = np.asfortranarray(Q)
    return lib().canary_l2energysquared(s.Np, s.nstate, s.nvgeo, IMJ[s.dim], _p(Qf), _p(s.vgeo),
                                        s.nelem)


# ------------------------------------------------- device-side preparation
def _torch():
    import torch
    return torch


def to_device(s, device="cuda"):
    """Upload the maps and build vgeo / sgeo on the GPU (device
    computegeometry, bit-identical to the host one).  Sets s.dev = {vgeo,
    sgeo, D, vmapM, vmapP, bndy} (flat CUDA tensors, reference layouts)."""
    torch = _torch()
    from . import kernels as K
    dv = torch.device(device)

    def up(a):
        return torch.from_numpy(np.ascontiguousarray(np.asarray(a).reshape(-1, order="F"))).to(dv)

    vM = up(s.vmapM)
    vgeo, sgeo = K.computegeometry(s.dim, s.N, up(s.mesh["elemtocoord"]), vM)
    s.dev = dict(vgeo=vgeo, sgeo=sgeo, D=up(s.D), vmapM=vM, vmapP=up(s.vmapP),
                 bndy=up(s.elemtobndy))
    return s


def device_arrays(s, device="cuda"):
    """The drop-in arrays of a setup on the device: s.dev when the setup was
    prepared there, else uploads of the host arrays."""
    if getattr(s, "dev", None) is not None:
        return s.dev
    torch = _torch()
    dv = torch.device(device)

    def up(a):
        return torch.from_numpy(np.ascontiguousarray(np.asarray(a).reshape(-1, order="F"))).to(dv)

    return dict(vgeo=up(s.vgeo), sgeo=up(s.sgeo), D=up(s.D), vmapM=up(s.vmapM),
                vmapP=up(s.vmapP), bndy=up(s.elemtobndy))


def device_state(s, which="Q", device="cuda"):
    """Q (or Q0) of a setup as a flat CUDA float64 tensor."""
    d = getattr(s, which + "d", None)
    if d is not None:
        return d
    torch = _torch()
    a = getattr(s, which)
    return torch.from_numpy(np.ascontiguousarray(a.reshape(-1, order="F"))).to(torch.device(device))


def _stream():
    return C.c_void_p(_torch().cuda.current_stream().cuda_stream)


def bubble_state_device(s, theta0=300.0, dtheta=0.5, rc=250.0, center=(500.0, 500.0, 350.0),
                        spherical=False, jitter=0.0, seed=0):
    torch = _torch()
    b = _lib.Bubble(theta0=theta0, dtheta=dtheta, rc=rc, xc=center[0], yc=center[1],
                    zc=center[2] if s.dim == 3 else center[-1], spherical=1 if spherical else 0,
                    jitter=jitter, seed=seed)
    vgeo = s.dev["vgeo"]
    Q = torch.empty(s.Np * s.nstate * s.nelem, dtype=torch.float64, device=vgeo.device)
    check(lib().canary_ic_bubble_device(s.dim, s.N, s.nelem, C.c_void_p(vgeo.data_ptr()),
                                        C.byref(b), C.byref(s.phys), C.c_void_p(Q.data_ptr()),
                                        _stream()), "ic_bubble_device")
    return Q


def vortex_state_device(s, t=0.0, params=None):
    torch = _torch()
    params = params or vortex_params()
    vgeo = s.dev["vgeo"]
    Q = torch.empty(s.Np * 4 * s.nelem, dtype=torch.float64, device=vgeo.device)
    check(lib().canary_ic_vortex2d_device(s.N, s.nelem, C.c_void_p(vgeo.data_ptr()),
                                          C.byref(params), float(t), C.c_void_p(Q.data_ptr()),
                                          _stream()), "ic_vortex2d_device")
    return Q


def _reduce_device(fn, *args):
    torch = _torch()
    work = torch.empty(_lib.REDUCE_WORK, dtype=torch.float64, device="cuda")
    out = torch.empty((), dtype=torch.float64, device="cuda")
    check(fn(*args, C.c_void_p(work.data_ptr()), C.c_void_p(out.data_ptr()), _stream()),
          "device reduction")
    return float(out.item())


def cfl_device(s, Q):
    """cfl() (before the / N^sqrt(2)) of a device state, on the device."""
    return _reduce_device(lib().canary_cfl_device, s.eqn, s.dim, s.N,
                          C.c_void_p(s.dev["vgeo"].data_ptr()), C.c_void_p(Q.data_ptr()),
                          s.nelem, C.byref(s.phys))


def euler_courant_device(s, Q, courant=0.5):
    """euler_courant of a device state, on the device."""
    return _reduce_device(lib().canary_courant_device, s.dim, s.N,
                          C.c_void_p(s.dev["vgeo"].data_ptr()), C.c_void_p(Q.data_ptr()),
                          s.nelem, C.byref(s.phys), float(courant))


def bubble3d_mesh(Ne=32, N=4, L=1000.0, gravity=9.81, walls=True):
    """Host part of bubble3d without geometry (no GPU work)."""
    phys = _lib.default_phys(gravity=gravity)
    x = linrange(0.0, L, Ne + 1)
    per = (False, False, False) if walls else (True, True, False)
    return build(_lib.EULER, 3, N, (x, x, x), per, phys=phys,
                 label="euler3d bubble %d^3 N=%d" % (Ne, N), geometry=False)


def bubble3d_on_device(s, dtheta=2.0, jitter=1e-3, seed=1, spherical=True):
    """Device part of bubble3d: geometry, state and hydrostatic background on
    the GPU (s.Qd, s.Q0d); s.Q = s.Q0 = None."""
    to_device(s)
    s.Qd = bubble_state_device(s, dtheta=dtheta, spherical=spherical, jitter=jitter, seed=seed)
    s.Q0d = bubble_state_device(s, dtheta=0.0, spherical=spherical)
    return s


def vortex2d_mesh(Ne=10, N=4, half=5.0, params=None):
    """Host part of vortex2d without geometry (no GPU work)."""
    x = linrange(-half, half, Ne + 1)
    params = params or vortex_params(Lx=2 * half, Ly=2 * half)
    s = build(_lib.EULER, 2, N, (x, x), (True, True), phys=vortex_phys(params.gamma),
              label="euler2d vortex %d^2 N=%d" % (Ne, N), geometry=False)
    s.vortex = params
    return s


def vortex2d_on_device(s):
    """Device part of vortex2d: geometry, exact state at t=0 and the uniform
    background at rest on the GPU."""
    torch = _torch()
    to_device(s)
    p = s.vortex
    s.Qd = vortex_state_device(s, 0.0, p)
    ph = s.phys
    rt = (ph.p0 / ph.R_gas) * (p.p_inf / ph.p0) ** (ph.c_v / ph.c_p)
    Q0 = torch.zeros((s.nelem, s.nstate, s.Np), dtype=torch.float64, device=s.Qd.device)
    Q0[:, s.dim, :] = p.rho_inf
    Q0[:, s.dim + 1, :] = rt
    s.Q0d = Q0.reshape(-1)
    return s


# -------------------------------------------------------- named workloads
def bubble3d(Ne=32, N=4, L=1000.0, gravity=9.81, dtheta=2.0, jitter=1e-3, seed=1,
             walls=True, spherical=True):
    """3D rising thermal bubble: [0,L]^3, Ne^3 hexes, no-flux walls, g,
    theta0 = 300 K, +dtheta K cosine bubble of radius 250 m at (500,500,350),
    seeded theta jitter."""
    phys = _lib.default_phys(gravity=gravity)
    x = linrange(0.0, L, Ne + 1)
    per = (False, False, False) if walls else (True, True, False)
    s = build(_lib.EULER, 3, N, (x, x, x), per, phys=phys,
              label="euler3d bubble %d^3 N=%d" % (Ne, N))
    s.Q = bubble_state(s, dtheta=dtheta, spherical=spherical, jitter=jitter, seed=seed)
    s.Q0 = bubble_state(s, dtheta=0.0, spherical=spherical)  # hydrostatic background
    return s


def bubble2d(Ne=10, N=4, L=1000.0, gravity=10.0, dtheta=0.5, jitter=0.0, seed=1):
    """2D bubble of the reference main() (euler2d_set2c.jl:1046-1118)."""
    phys = _lib.default_phys(gravity=gravity)
    x = linrange(0.0, L, Ne + 1)
    s = build(_lib.EULER, 2, N, (x, x), (True, False), phys=phys,
              label="euler2d bubble %d^2 N=%d" % (Ne, N))
    s.Q = bubble_state(s, dtheta=dtheta, center=(500.0, 350.0), jitter=jitter, seed=seed)
    s.Q0 = bubble_state(s, dtheta=0.0, center=(500.0, 350.0))
    return s


def bubble2d_te(Ne=10, N=4, L=1000.0, gravity=10.0, theta_c=0.5, rc=250.0, center=(500.0, 350.0),
                theta_ref=300.0):
    """The 2D rising bubble of the total-energy example (euler2d_set3c.jl
    main :1474-1556: [0, L]^2, periodic in x, walls in y): ic :1489-1515 in
    (rho, u, v, theta), then convert_set2nc_to_set3c :1306-1329 to
    (rho u, rho v, rho, E) with E = rho (c_v T + (u^2 + v^2)/2 + g y).  Q0 is
    the same state without the bubble."""
    phys = _lib.default_phys(gravity=gravity)
    x = linrange(0.0, L, Ne + 1)
    s = build(_lib.EULER_TE, 2, N, (x, x), (True, False), phys=phys,
              label="euler2d total-energy bubble %d^2 N=%d" % (Ne, N))
    s.Q = te_bubble_state(s, theta_c=theta_c, rc=rc, center=center, theta_ref=theta_ref)
    s.Q0 = te_bubble_state(s, theta_c=0.0, rc=rc, center=center, theta_ref=theta_ref)
    return s


def te_bubble_state(s, theta_c=0.5, rc=250.0, center=(500.0, 350.0), theta_ref=300.0):
    """ic + convert_set2nc_to_set3c of euler2d_set3c.jl on a setup's nodes."""
    ph = s.phys
    x, y = s.vgeo[:, 6, :], s.vgeo[:, 7, :]
    r = np.sqrt((x - center[0]) ** 2 + (y - center[1]) ** 2)
    dth = np.where(r <= rc, 0.5 * theta_c * (1.0 + np.cos(np.pi * r / rc)), 0.0)
    th = theta_ref + dth
    pi_k = 1.0 - ph.gravity / (ph.c_p * th) * y
    rho = ph.p0 / (ph.R_gas * th) * pi_k ** (ph.c_v / ph.R_gas)
    u = np.zeros_like(rho)
    v = np.zeros_like(rho)
    P = ph.p0 * (rho * ph.R_gas * th / ph.p0) ** (ph.c_p / ph.c_v)
    T = P / (rho * ph.R_gas)
    E = ph.c_v * T + 0.5 * (u ** 2 + v ** 2) + ph.gravity * y
    Q = s.zeros_state()
    Q[:, 0, :], Q[:, 1, :], Q[:, 2, :], Q[:, 3, :] = rho * u, rho * v, rho, rho * E
    return Q


def te_dt(s, Q):
    """courantnumber's time step of the total-energy example
    (euler2d_set3c.jl:121-148): min over nodes of dx / (|u| + c) / N."""
    ph = s.phys
    U, V, rho, E = (Q[:, k, :] for k in range(4))
    xix, etay, y = s.vgeo[:, 0, :], s.vgeo[:, 3, :], s.vgeo[:, 7, :]
    P = (ph.R_gas / ph.c_v) * (E - (U * U + V * V) / (2 * rho) - rho * ph.gravity * y)
    u, v = U / rho, V / rho
    dx = np.sqrt((1.0 / (2 * xix)) ** 2 + (1.0 / (2 * etay)) ** 2)
    ws = np.sqrt(u * u + v * v) + np.sqrt(ph.gamma * P / rho)
    return float(np.min(1.0 * dx / ws / s.N))


def te_to_set2nc(s, Q):
    """convert_set3c_to_set2nc (euler2d_set3c.jl:1331-1356): (u, v, rho,
    theta) from the total-energy state, for the energy norm and plots."""
    ph = s.phys
    U, V, rho, E = (np.array(Q[:, k, :]) for k in range(4))
    y = s.vgeo[:, 7, :]
    u, v, e = U / rho, V / rho, E / rho
    P = (ph.R_gas / ph.c_v) * rho * (e - 0.5 * (u * u + v * v) - ph.gravity * y)
    th = ph.p0 / (rho * ph.R_gas) * (P / ph.p0) ** (ph.c_v / ph.c_p)
    out = np.empty_like(np.asarray(Q), order="F")
    out[:, 0, :], out[:, 1, :], out[:, 2, :], out[:, 3, :] = u, v, rho, th
    return out


def vortex_phys(gamma=1.4):
    """set2c constants for which P = (rho theta)^gamma: p0 = R = 1, c_p/c_v = gamma, g = 0."""
    return _lib.default_phys(gamma=gamma, p0=1.0, R_gas=1.0, c_p=gamma, c_v=1.0, gravity=0.0)


def vortex2d(Ne=10, N=4, half=5.0, params=None):
    """Isentropic vortex on the periodic box [-half, half]^2."""
    x = linrange(-half, half, Ne + 1)
    params = params or vortex_params(Lx=2 * half, Ly=2 * half)
    s = build(_lib.EULER, 2, N, (x, x), (True, True), phys=vortex_phys(params.gamma),
              label="euler2d vortex %d^2 N=%d" % (Ne, N))
    s.Q = vortex_state(s, 0.0, params)
    s.vortex = params
    s.Q0 = uniform_rest_state(s, params.rho_inf, params.p_inf)
    return s


def swe2d(Ne=10, N=4, depth=10.0, bumps=((0.5, 0.5, 0.05, 0.1),), periodic=True,
          gravity=10.0):
    """Shallow water at rest on [0,1]^2 with Gaussian bumps over depth b."""
    phys = _lib.default_phys(swe_gravity=gravity, delta_nl=1.0)
    x = linrange(0.0, 1.0, Ne + 1)
    per = (True, True) if periodic else (False, False)
    s = build(_lib.SWE, 2, N, (x, x), per, phys=phys, label="swe2d %d^2 N=%d" % (Ne, N))
    s.Q = swe_state(s, bumps, depth=depth)
    return s


def swe_ensemble(n=64, seed=42, ne=(8, 64), orders=(3, 4, 5, 7), depth=10.0,
                 height=(0.01, 0.1), width=(0.05, 0.2), nbumps=2, gravity=10.0):
    """n independent periodic shallow-water solves on [0,1]^2: Ne^2 quads with
    Ne uniform in ne, N drawn from orders, still depth `depth` with nbumps
    Gaussian bumps of seeded height and width (domain widths).  Returns
    [(setup, dt)] with the shallow-water example's Courant step."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        Ne = int(rng.integers(ne[0], ne[1] + 1))
        N = int(rng.choice(orders))
        bumps = [(rng.uniform(0, 1), rng.uniform(0, 1), rng.uniform(*height), rng.uniform(*width))
                 for _ in range(nbumps)]
        s = swe2d(Ne=Ne, N=N, depth=depth, bumps=bumps, periodic=True, gravity=gravity)
        out.append((s, swe_courant(s, s.Q)))
    return out


@dataclass
class RankShare:
    """One rank's share of a mesh split into contiguous element chunks of the
    (Hilbert) element order, the reference's partition (src/mesh.jl
    partition): the real elements first, then the ghost elements their faces
    reach, as the reference appends them after realelems (src/mesh.jl
    :944-955).  Maps are renumbered locally; vgeo/sgeo hold the real
    elements only (ghosts are read through vmapP, never computed)."""
    nreal: int
    nelem_total: int
    real: np.ndarray     # global element index of each real element
    ghost: np.ndarray    # global element index of each ghost element, ascending
    vgeo: np.ndarray
    sgeo: np.ndarray
    vmapM: np.ndarray
    vmapP: np.ndarray
    elemtobndy: np.ndarray
    sendelems: np.ndarray  # local indices of real elements some other rank reads (1-based)
    send_nodes: np.ndarray  # local node ids (1-based) other ranks read through their vmapP
    recv_nodes: np.ndarray  # local node ids (1-based) of the ghost nodes this rank's vmapP reads

    def local_state(self, Q):
        """(Np, nstate, nreal + nghost) state of this share from a global one."""
        return np.asfortranarray(np.concatenate([Q[:, :, self.real], Q[:, :, self.ghost]], axis=2))


def rank_share(s, nranks, rank):
    """The share of rank `rank` when setup `s` is split into `nranks`
    contiguous chunks (sizes differing by at most one, the reference's
    linearpartition)."""
    ne, Np = s.nelem, s.Np
    lo = rank * ne // nranks
    hi = (rank + 1) * ne // nranks
    real = np.arange(lo, hi)
    nreal = hi - lo
    vmapP = np.asarray(s.vmapP)[..., lo:hi]
    eP = (vmapP - 1) // Np
    nodeP = (vmapP - 1) % Np
    outside = (eP < lo) | (eP >= hi)
    ghost = np.unique(eP[outside])
    local = np.full(ne, -1, dtype=np.int64)
    local[real] = np.arange(nreal)
    local[ghost] = nreal + np.arange(ghost.size)
    vmapP_l = np.asfortranarray(local[eP] * Np + nodeP + 1)
    vmapM = np.asarray(s.vmapM)[..., lo:hi]
    vmapM_l = np.asfortranarray(vmapM - lo * Np)
    # the real elements some other rank's ghosts are (what a send buffer packs)
    sends, snodes = [], []
    for r in range(nranks):
        if r == rank:
            continue
        rlo, rhi = r * ne // nranks, (r + 1) * ne // nranks
        vPr = np.asarray(s.vmapP)[..., rlo:rhi]
        ePr = (vPr - 1) // Np
        mine = (ePr >= lo) & (ePr < hi)
        sends.append(ePr[mine])
        snodes.append(vPr[mine] - lo * Np)
    sendelems = (np.unique(np.concatenate(sends)) - lo + 1) if sends else np.zeros(0, np.int64)
    send_nodes = np.unique(np.concatenate(snodes)) if snodes else np.zeros(0, np.int64)
    recv_nodes = np.unique(vmapP_l[vmapP_l > nreal * Np])
    return RankShare(nreal=nreal, nelem_total=nreal + ghost.size, real=real, ghost=ghost,
                     vgeo=np.asfortranarray(np.asarray(s.vgeo)[..., lo:hi]),
                     sgeo=np.asfortranarray(np.asarray(s.sgeo)[..., lo:hi]),
                     vmapM=vmapM_l, vmapP=vmapP_l,
                     elemtobndy=np.asfortranarray(np.asarray(s.elemtobndy)[..., lo:hi]),
                     sendelems=sendelems.astype(np.int64), send_nodes=send_nodes.astype(np.int64),
                     recv_nodes=recv_nodes.astype(np.int64))


def rk_coefs():
    """LSRK54 (Carpenter & Kennedy 1994) as DFloat(num)/DFloat(den):
    euler3d_set2c.jl:772-788."""
    A = (0.0, -567301805773 / 1357537059087, -2404267990393 / 2016746695238,
         -3550918686646 / 2091501179385, -1275806237668 / 842570457699)
    B = (1432997174477 / 9575080441755, 5161836677717 / 13612068292357,
         1720146321549 / 2090206949498, 3134564353537 / 4481467310338,
         2277821191437 / 14882151754819)
    Cc = (0.0, 1432997174477 / 9575080441755, 2526269341429 / 6820363962896,
          2006345519317 / 3224310063776, 2802321613138 / 2924317926251)
    return A, B, Cc
