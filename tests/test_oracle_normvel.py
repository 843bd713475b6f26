"""press_normvel_der (src/linsys/mod_pressure_equation.f90:396-462) in the
oracle, pinned by hand-worked cases on a flat square grid with central
difference stencils: for a linear surface velocity U the interior panels
give GradS_Un = grad(U.n) and DivS_U = div U exactly, so the source is
n.(ub - U_old)/dt - grad(U_z).ub + (dUx/dx + dUy/dy) ub_z; a boundary edge
takes U_e - U = -2 U, an edge whose neighbour is not a surface panel adds
nothing, and R_g multiplies the stencils."""
import numpy as np

import normvel_cases as NC

A, B, Cc, D = 0.7, -0.3, 1.1, 0.4


def setup(dt=0.01):
    g = NC.flat_grid()
    U = NC.linear_field(g["cen"], A, B, Cc, D)
    rng = np.random.default_rng(2)
    Uold = U - rng.uniform(-0.1, 0.1, U.shape)
    ub = np.tile([0.3, -0.2, 0.9], (g["n"], 1))
    return g, U, Uold, ub, dt


def test_interior_panels_are_exact_for_linear_fields(orc):
    g, U, Uold, ub, dt = setup()
    out = orc.press_normvel_der(g["nor"], ub, U, Uold, g["st"], g["n_ver"], g["neigh"], g["R"], dt)
    for j in range(1, g["ny"] - 1):
        for i in range(1, g["nx"] - 1):
            k = g["idx"](i, j)
            ref = (ub[k, 2] - Uold[k, 2]) / dt - (Cc * ub[k, 0] + D * ub[k, 1]) + (A + B) * ub[k, 2]
            assert abs(out[k] - ref) <= 1e-12 * max(1.0, abs(ref))


def test_boundary_and_foreign_neighbours(orc):
    g, U, Uold, ub, dt = setup()
    h, k = g["h"], g["idx"](0, 2)  # the -x boundary, interior in y
    base = (ub[k, 2] - Uold[k, 2]) / dt
    # edge 3 (-x) has no neighbour: (U_e - U) = -2U
    gx = Cc / 2 + U[k, 2] / h
    gy = D
    div = A / 2 + U[k, 0] / h + B
    out = orc.press_normvel_der(g["nor"], ub, U, Uold, g["st"], g["n_ver"], g["neigh"], g["R"], dt)
    ref = base - (gx * ub[k, 0] + gy * ub[k, 1]) + div * ub[k, 2]
    assert abs(out[k] - ref) <= 1e-12 * abs(ref)
    # the same edge facing a panel of another kind: no term at all
    ng = g["neigh"].copy()
    ng[k, 3] = -2
    out2 = orc.press_normvel_der(g["nor"], ub, U, Uold, g["st"], g["n_ver"], ng, g["R"], dt)
    ref2 = base - (Cc / 2 * ub[k, 0] + D * ub[k, 1]) + (A / 2 + B) * ub[k, 2]
    assert abs(out2[k] - ref2) <= 1e-12 * abs(ref2)


def test_rotation_scales_the_surface_terms(orc):
    g, U, Uold, ub, dt = setup()
    o1 = orc.press_normvel_der(g["nor"], ub, U, Uold, g["st"], g["n_ver"], g["neigh"], g["R"], dt)
    o2 = orc.press_normvel_der(g["nor"], ub, U, Uold, g["st"], g["n_ver"], g["neigh"],
                               2.0 * g["R"], dt)
    base = np.einsum("ij,ij->i", g["nor"], ub - Uold) / dt
    assert np.allclose(o2 - base, 2.0 * (o1 - base), rtol=1e-12, atol=1e-12)
    # a body at rest: only the time derivative of the normal velocity is left
    z = orc.press_normvel_der(g["nor"], 0 * ub, U, Uold, g["st"], g["n_ver"], g["neigh"], g["R"],
                              dt)
    assert np.allclose(z, -Uold[:, 2] / dt, rtol=1e-14, atol=0)
