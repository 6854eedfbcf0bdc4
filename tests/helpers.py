"""Shared test helpers: state generators and the mapping between the product's
reference-valued enums and the oracle's internal scheme ids."""
import numpy as np

from cfd_inwsl_amd import BC, FluxMethod, InterMethod, Recon, cases
from oracle import oracle as O

SCHEME_OF = {
    InterMethod.WCNS5: O.S_JS,
    InterMethod.WCNSZ5: O.S_Z,
    InterMethod.TCNS5: O.S_TENO5,
    InterMethod.TCNS5CT4: O.S_TENO5_CT4,
    InterMethod.TCNS5CT7: O.S_TENO5_CT7,
    InterMethod.WCNS5CONGZ: O.S_TENOS,
    InterMethod.WCNS5CONGZCT4: O.S_TENOS_CT4,
    InterMethod.WCNS5CONGZCT7: O.S_TENOS_CT7,
    InterMethod.CONGTCNS5CT5: O.S_TENO5,
    InterMethod.CONGTCNS5CT10: O.S_TENO5_CT10,
    InterMethod.NEW_WHFTCNSA: O.S_TENOA,
    InterMethod.WHFZYCTCNSNMYASF002_1: O.S_TENOSA_F002,
    InterMethod.NEW_WHFTCNSASF203_NOS: O.S_TENOSA_F203,
    InterMethod.NEW_WHFTCNSASF202_NOS: O.S_TENOSA_F202,
    InterMethod.WHFTCNSN: O.S_WHF_N,
    InterMethod.WHFTCNSNA: O.S_WHF_A,
    InterMethod.WHFTCNSNAS: O.S_WHF_AS,
    InterMethod.WHFTCNSNS: O.S_WHF_S,
    InterMethod.WHFTCNSNLAD: O.S_WHF_LAD,
    InterMethod.WHFAITCNSNS: O.S_AI_S,
    InterMethod.WHFAITCNSNA: O.S_AI_A,
    InterMethod.WHFAITCNSNAS_1: O.S_AI_AS,
    InterMethod.WHFAITCNSNAS_2: O.S_AI_AS,
    InterMethod.WHFAITCNSNLADS: O.S_AI_LADS,
    InterMethod.WHFAITCNSNAZS: O.S_AI_AZS,
    InterMethod.WHFAITCNSNMYASF002_1: O.S_AI_F002,
    InterMethod.WHFAITCNSNMYASF002_2: O.S_AI_F002,
    InterMethod.WHFAITCNSNMYASF002_AI1: O.S_AI_F002,
    InterMethod.WHFCOMPARE: O.S_WHF_COMPARE,
    InterMethod.MUCSLIN5: O.S_MUSCL5,
}


def scheme_of(inter, dim):
    """The oracle scheme the reference wires to an InterMethod in dim
    dimensions (its 1D Euler path takes Teno5_CongAA for WCNS5CONGZCT7,
    solverType.cpp:248-250)."""
    if dim == 1 and inter == InterMethod.WCNS5CONGZCT7:
        return O.S_CONGAA
    return SCHEME_OF[inter]


def oracle_grid(dim, n, h, inter, flux, recon, bc, bc_prim=None, x0=0.0, t=0.0, gravity=(0.0, 0.0, 0.0)):
    return O.make_grid(dim, n, h, scheme_of(inter, dim), int(flux), 1 if recon == Recon.CHAR else 0,
                       [int(b) for b in bc], bc_prim, x0=x0, t=t, gravity=gravity)


def random_state(dim, n, seed, jumps=True, amp=0.2):
    """Smooth random field plus (optionally) a few discontinuities, positive
    density and pressure; AoS [cell][nv] with x fastest."""
    rng = np.random.default_rng(seed)
    shape = tuple(reversed(list(n)[:dim]))  # (nz, ny, nx)
    grids = np.meshgrid(*[np.linspace(0, 2 * np.pi, s, endpoint=False) for s in shape], indexing="ij")

    def field():
        f = np.zeros(shape)
        for _ in range(3):
            k = rng.integers(1, 4, size=dim)
            ph = rng.uniform(0, 2 * np.pi)
            f += np.sin(sum(k[i] * grids[i] for i in range(dim)) + ph)
        return f / 3

    prim = np.empty(shape + (dim + 2,))
    prim[..., 0] = 1.0 + amp * field()
    for v in range(dim):
        prim[..., 1 + v] = 0.5 * field()
    prim[..., -1] = 1.0 + amp * field()
    if jumps:
        for _ in range(2):
            ax = rng.integers(0, dim)
            cut = rng.integers(2, shape[ax] - 2)
            sl = [slice(None)] * dim
            sl[ax] = slice(cut, None)
            prim[tuple(sl) + (0,)] *= rng.uniform(0.2, 0.5)
            prim[tuple(sl) + (dim + 1,)] *= rng.uniform(0.1, 0.5)
    return cases.prim_to_cons(prim).reshape(-1, dim + 2)


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    scale = np.abs(b).max(axis=0)
    scale = np.where(scale > 0, scale, 1.0)
    return float((np.abs(a - b).max(axis=0) / scale).max())
