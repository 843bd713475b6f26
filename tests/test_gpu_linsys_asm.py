"""GPU parity of initialize_linsys / assemble_linsys with the panel wake
(dust_initialize_linsys, dust_assemble_linsys through the C ABI) against the
oracle on identical inputs: a surface-panel wing with implicit and explicit
wake panels and lifting-line panels, a vortex-lattice plate behind it (mixed
row kinds), static/moving splits, body velocities, uvort and the pressure
system.  FP64: relative max-norm <= 1e-12 on A, B, L_static, A_pres, b and
b_pres; LU factors of the static blocks <= 1e-11 with the same pivots."""
import numpy as np
import pytest

from dust_amd import linsys as LS

import linsys_cases as LC

pytestmark = pytest.mark.gpu


def relmax(a, b):
    s = np.abs(b).max()
    return np.abs(a - b).max() / (s if s > 0 else 1.0)


@pytest.mark.parametrize("split", ["static", "mixed", "moving"])
def test_assembly_vs_oracle(orc, split):
    c = LC.mixed_case()
    n = len(c["n_ver"])
    nst = {"static": c["n_wing"], "mixed": 3 * c["strip"], "moving": 0}[split]
    nss = min(nst, c["n_wing"])
    ub = None
    if split != "static":
        ub = np.random.default_rng(7).normal(scale=0.5, size=(n, 3))
        ub[:nst] = 0.0  # static elements do not move
    S = LC.mixed_system(c, nst, nss, 1 if nst else 0, ub=ub)
    print(f"assembly split={split} rank={n} nstatic={nst} wake={S.n_wake} expl={S.n_expl}",
          flush=True)
    g = LS.initialize_linsys(S)
    o = orc.initialize_linsys(S)
    assert g["info"] == o["info"] == 0
    if nst:
        assert np.array_equal(g["ipiv"][:nst], o["ipiv"][:nst])
        assert relmax(g["A"][:nst, :nst], o["A"][:nst, :nst]) <= 1e-11
        assert relmax(g["B"][:nst, :nst], o["Bst"][:nst, :nst]) <= 1e-12
        assert relmax(g["L_static"][:nst, :1], o["L_static"][:nst, :1]) <= 1e-12
    if nss:
        assert np.array_equal(g["ipiv_pres"][:nss], o["ipiv_pres"][:nss])
        assert relmax(g["A_pres"][:nss, :nss], o["A_pres"][:nss, :nss]) <= 1e-11
    for step in range(2):
        # a new uvort and new explicit strengths every step, the wake panels
        # move back
        S.uvort = S.uvort * (1.0 + 0.1 * step)
        S.wake_mag = S.wake_mag * (1.0 - 0.2 * step)
        S.wake_ver = S.wake_ver + np.array([0.05 * step, 0.0, 0.0])
        bg, bpg = LS.assemble_linsys(S, g, pressure=True)
        bo, bpo = orc.assemble_linsys(S, o)
        eb, ebp = relmax(bg, bo), relmax(bpg, bpo)
        ea = relmax(g["A"][:, nst:], o["A"][:, nst:])
        ea2 = relmax(g["A"][nst:, :nst], o["A"][nst:, :nst]) if nst < n and nst else 0.0
        print(f"step {step}: b {eb:.2e} b_pres {ebp:.2e} A {ea:.2e} {ea2:.2e}")
        assert eb <= 1e-12 and ebp <= 1e-12 and ea <= 1e-12 and ea2 <= 1e-12
        nsp = c["n_wing"]
        if nss and nss < nsp:
            assert relmax(g["A_pres"][:nss, nss:nsp], o["A_pres"][:nss, nss:nsp]) <= 1e-12
        if nss < nsp:
            assert relmax(g["A_pres"][nss:nsp, :nsp], o["A_pres"][nss:nsp, :nsp]) <= 1e-12


def test_wing_with_wake_solves_like_oracle(orc):
    """The surface-panel wing at rest with its implicit wake row: device
    assembly and factorisation, then dgetrs, against the oracle's solution."""
    ver, nv, dy, wake, up, lo = LC.morino_wing(nc=20, ns=20)
    S = LC.system(ver, nv, LS.ROW_SURFPAN, len(nv), wake=wake, u_inf=(10.0, 0.0, 0.0))
    g = LS.initialize_linsys(S)
    b = LS.assemble_linsys(S, g)
    x = LS.dgetrs(g["A"], g["ipiv"], np.asfortranarray(b.copy()))
    o = orc.initialize_linsys(S)
    bo, _ = orc.assemble_linsys(S, o)
    xo = bo.copy()
    orc.lib().ao_dgetrs(S.rank, 1, orc._p(o["A"]), S.rank, orc._p(o["ipiv"]), orc._p(xo), S.rank)
    print(f"wing with wake: x {relmax(x, xo):.2e}")
    assert relmax(x, xo) <= 1e-10
