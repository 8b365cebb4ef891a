"""Sutton-Chen EAM set-up shared by the CPU and GPU EAM tests: the reference's
test parameters (test/setup_params.jl:79-87: n = 8.482, m = 4.692,
eps = 0.0370 eV, C = 27.561, a = 0.25 nm) and start configurations."""
import numpy as np

NM_TO_BOHR = 18.8973
EV_TO_HARTREE = 0.0367493
N_EXP, M_EXP, EPS, C_SC, A_LAT = 8.482, 4.692, EV_TO_HARTREE * 0.0370, 27.561, 0.25 * NM_TO_BOHR


def eam_constants():
    """(n, m, ean, eCam) as EmbeddedAtomPotential(n, m, eps, C, a) stores them
    (src/EnergyEvaluation.jl:1078-1082: ean = eps a^n / 2, eCam = eps C a^(m/2))."""
    return N_EXP, M_EXP, EPS * A_LAT**N_EXP / 2, EPS * C_SC * A_LAT ** (M_EXP / 2)


def cuboctahedron():
    """13-atom fcc fragment (centre + 12 nearest neighbours at a / sqrt 2) and
    a container radius of 2.5 nearest-neighbour distances."""
    d = A_LAT / np.sqrt(2)
    v = [(0, 0, 0)] + [(sx, sy, 0) for sx in (1, -1) for sy in (1, -1)] + \
        [(sx, 0, sz) for sx in (1, -1) for sz in (1, -1)] + [(0, sy, sz) for sy in (1, -1) for sz in (1, -1)]
    return np.array(v, float) * d / np.sqrt(2), 2.5 * d


def fcc_box(n_cells):
    """fcc crystal of 4 n^3 atoms at the lattice constant a in a cubic box."""
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    pts = [(np.array([i, j, k]) + b) * A_LAT for i in range(n_cells) for j in range(n_cells)
           for k in range(n_cells) for b in base]
    L = n_cells * A_LAT
    return np.array(pts) - L / 2 + A_LAT / 4, L


def brute_force_energy(x, box=None):
    """Independent numpy evaluation of E = sum_i ean phi_i - eCam sqrt(rho_i)
    (src/EnergyEvaluation.jl:1059-1073) with an explicit minimum image."""
    n, m, ean, ecam = eam_constants()
    d = x[:, None, :] - x[None, :, :]
    if box is not None:
        d = d - box * np.round(d / box)
    r = np.sqrt((d**2).sum(-1))
    np.fill_diagonal(r, np.inf)
    phi = (r**-n).sum(1)
    rho = (r**-m).sum(1)
    return float(np.sum(ean * phi - ecam * np.sqrt(rho))), phi, rho
