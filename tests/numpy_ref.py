"""Independent numpy restatements of the reference formulas on dense (y, x)
arrays, used to pin the oracle.  Ghost cells are numpy.pad(mode="edge"),
i.e. BlockLabOpen's zero-gradient rule (Definitions.h:125-157); the block
structure plays no role here, so agreement with the block/lab oracle pins
the block and ghost indexing.  numpy's elementwise float64 arithmetic is
IEEE without contraction, so stencil values compare bit for bit.
"""
import numpy as np

BS = 32
EPS = np.finfo(np.float64).eps


def _smax(a, b):  # std::max
    return b if a < b else a


def _smin(a, b):  # std::min
    return b if b < a else a


def pad(f):
    return np.pad(f, 1, mode="edge")


def advdiff(u, v, dt, h, nu, uinf):
    """advDiff::operator() (advDiff.cpp:36-118). Returns (u, v, corr)."""
    NY, NX = u.shape
    U, V = pad(u), pad(v)
    nor = _smax(_smax(abs(uinf[0]), abs(uinf[1])), EPS)
    t = [_smax(uinf[0], 0.0) / nor, _smax(uinf[1], 0.0) / nor,
         _smin(uinf[0], 0.0) / nor, _smin(uinf[1], 0.0) / nor]
    fW, fS, fE, fN = [1 - x * x / 100 for x in t]
    # advDiff.cpp:68-71: the ghost layer of a domain-face block is faded
    U[1:-1, 0] *= fW; V[1:-1, 0] *= fW
    U[0, 1:-1] *= fS; V[0, 1:-1] *= fS
    U[1:-1, -1] *= fE; V[1:-1, -1] *= fE
    U[-1, 1:-1] *= fN; V[-1, 1:-1] *= fN
    afac, dfac = -0.5 * dt / h, (nu / h) * (dt / h)
    uc, vc = U[1:-1, 1:-1], V[1:-1, 1:-1]
    upx, upy, ulx, uly = U[1:-1, 2:], U[2:, 1:-1], U[1:-1, :-2], U[:-2, 1:-1]
    vpx, vpy, vlx, vly = V[1:-1, 2:], V[2:, 1:-1], V[1:-1, :-2], V[:-2, 1:-1]
    dUadv = (uc + uinf[0]) * (upx - ulx) + (vc + uinf[1]) * (upy - uly)
    dUdif = upx + upy + ulx + uly - 4 * uc
    dVadv = (uc + uinf[0]) * (vpx - vlx) + (vc + uinf[1]) * (vpy - vly)
    dVdif = vpx + vpy + vlx + vly - 4 * vc
    u1 = uc + (afac * dUadv + dfac * dUdif)
    v1 = vc + (afac * dVadv + dfac * dVdif)
    IF = u1[:, -1].sum() - u1[:, 0].sum() + v1[-1, :].sum() - v1[0, :].sum()
    corr = IF / (2 * NY + 2 * NX)
    u1[:, 0] += corr
    u1[:, -1] -= corr
    v1[0, :] += corr
    v1[-1, :] -= corr
    return u1, v1, corr


def advdiff_stencil_only(u, v, dt, h, nu, uinf):
    """advDiff stencil without the boundary flux correction."""
    NY, NX = u.shape
    u1, v1, corr = advdiff(u, v, dt, h, nu, uinf)
    u1[:, 0] -= corr
    u1[:, -1] += corr
    v1[0, :] -= corr
    v1[-1, :] += corr
    return u1, v1


def pressure_rhs(u, v, dt, h):
    """PressureSingle.cpp:110-128 (no shapes)."""
    U, V = pad(u), pad(v)
    facDiv = 0.5 * h / dt
    return facDiv * ((U[1:-1, 2:] - U[1:-1, :-2]) + (V[2:, 1:-1] - V[:-2, 1:-1]))


def pressure_correction(u, v, p, dt, h):
    """PressureSingle.cpp:216-241."""
    P = pad(p)
    pFac = -0.5 * dt / h
    return u + pFac * (P[1:-1, 2:] - P[1:-1, :-2]), v + pFac * (P[2:, 1:-1] - P[:-2, 1:-1])


def cell_positions(NX, NY, hb, h):
    """Cubism BlockInfo::pos: origin (block index * hb) + h*(i+0.5)."""
    gx, gy = np.arange(NX), np.arange(NY)
    px = (gx // BS).astype(np.float64) * hb + h * ((gx % BS) + 0.5)
    py = (gy // BS).astype(np.float64) * hb + h * ((gy % BS) + 0.5)
    return px, py


def fade_arg(NX, NY, hb, h, extent, fadeLen):
    """PoissonSolver.cpp:87-92 on every cell."""
    px, py = cell_positions(NX, NY, hb, h)
    ext0, ext1 = (NX // BS) * h * BS, (NY // BS) * h * BS
    fl = fadeLen * extent
    inv = 1 / max(fl, EPS)
    X, Y = np.meshgrid(px, py)
    yt = inv * np.maximum(0.0, fl - ext1 + Y)
    yb = inv * np.maximum(0.0, fl - Y)
    xt = inv * np.maximum(0.0, fl - ext0 + X)
    xb = inv * np.maximum(0.0, fl - X)
    return np.minimum(np.maximum(np.maximum(np.maximum(yt, yb), xt), xb), 1.0)


def cub2rhs(r, hb, h, extent, fadeLen):
    """PoissonSolver::cub2rhs fade (PoissonSolver.cpp:61-122)."""
    NY, NX = r.shape
    fa = fade_arg(NX, NY, hb, h, extent, fadeLen)
    f = r * (fa * fa)
    sumRHS = f.sum()
    sumPOS = max(f[f > 0].sum(), EPS)
    sumNEG = max(-f[f < 0].sum(), EPS)
    corr = sumRHS / sumPOS if sumRHS > 0 else sumRHS / sumNEG
    fc = corr * (np.maximum(f, 0.0) if corr > 0 else -np.minimum(f, 0.0))
    return r * (1 - fa * fa) + fc


def poisson_periodic(r):
    """FFTW_periodic::_solve_spectral (FFTW_periodic.h:77-92) with numpy.fft."""
    NY, NX = r.shape
    R = np.fft.rfft2(r)
    ky = np.arange(NY)
    ky = np.where(ky <= NY // 2, ky, NY - ky)[:, None] * (2 * np.pi / NY)
    kx = np.arange(NX // 2 + 1)[None, :] * (2 * np.pi / NX)
    with np.errstate(divide="ignore"):
        S = -1 / (kx * kx + ky * ky)
    S[0, 0] = 0
    return np.fft.irfft2(R * S, s=(NY, NX))


def green(NX, NY, h):
    """Free-space kernel G(dx, dy) on |dx|<NX, |dy|<NY (FFTW_freespace.cpp:108-119)."""
    dx = np.arange(-(NX - 1), NX)[None, :]
    dy = np.arange(-(NY - 1), NY)[:, None]
    with np.errstate(divide="ignore"):
        G = np.log(h * np.sqrt(dx * dx + dy * dy)) / (2 * np.pi)
    G[NY - 1, NX - 1] = 1 / (2 * np.pi) / 2 * (2 * np.log(h / np.sqrt(np.pi)) - 1)
    return G


def poisson_freespace_direct(r, h):
    """Brute-force linear convolution p_i = sum_j G(x_i - x_j) r_j."""
    NY, NX = r.shape
    G = green(NX, NY, h)
    p = np.zeros_like(r)
    # p[j,i] = sum_{b,a} G[b - j + NY-1, a - i + NX-1] r[b,a]
    for j in range(NY):
        Gj = G[NY - 1 - j:2 * NY - 1 - j, :]
        for i in range(NX):
            p[j, i] = np.sum(Gj[:, NX - 1 - i:2 * NX - 1 - i] * r)
    return p
