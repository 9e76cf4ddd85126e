"""Synthetic copper-like funcfl EAM tables (Python statement, used to write fixtures).

The product writes the same tables from C (``mdb_write_funcfl_synthetic``); a CPU test checks
that both agree value by value.  Functions (r in Angstrom, energies in eV):

    taper(r)  = x^4 / (1 + x^4),  x = (r - rc) / h,  for r < rc; 0 beyond
    rho(r)    = exp(-3 (r / re - 1)) * taper(r)
    Z(r)      = Z0 exp(-beta (r - re)) * taper(r)          (Morse-like effective charge)
    F(rho)    = -E0 sqrt(rho)

with re = 2.556 (fcc Cu nearest-neighbour distance at a = 3.615), rc = 4.95, h = 0.5,
Z0 = 0.1776, beta = 1.5, E0 = 1.0 (Z0 puts dE/da ~ 0 at a = 3.615).  Grids: nr = nrho = 2000,
dr = 0.0025 (r up to 4.9975), drho = 0.02.  The file layout is the DYNAMO funcfl one that
MD-Bench's readEamFile parses (src/common/eam_utils.c:42-93): comment line, "ielem mass a lat",
"nrho drho nr dr cut", then F, Z and rho, five values per line.
"""
import math

RE, RC, H = 2.556, 4.95, 0.5
Z0, BETA, E0 = 0.1776, 1.5, 1.0
NR = NRHO = 2000
DR, DRHO = 0.0025, 0.02
MASS, ALAT = 63.55, 3.615


def taper(r):
    if r >= RC:
        return 0.0
    x = (r - RC) / H
    x4 = x * x * x * x
    return x4 / (1.0 + x4)


def rho_of_r(r):
    return math.exp(-3.0 * (r / RE - 1.0)) * taper(r)


def z_of_r(r):
    return Z0 * math.exp(-BETA * (r - RE)) * taper(r)


def f_of_rho(rho):
    return -E0 * math.sqrt(rho)


def tables():
    frho = [f_of_rho(i * DRHO) for i in range(NRHO)]
    zr = [z_of_r(i * DR) for i in range(NR)]
    rhor = [rho_of_r(i * DR) for i in range(NR)]
    return frho, zr, rhor


def write_funcfl_synthetic(path):
    frho, zr, rhor = tables()
    with open(path, "w") as fh:
        fh.write("synthetic Cu-like funcfl: rho=exp(-3(r/2.556-1))*taper, Z Morse-like, "
                 "F=-sqrt(rho)\n")
        fh.write("%d %.6f %.6f FCC\n" % (29, MASS, ALAT))
        fh.write("%d %.16e %d %.16e %.16e\n" % (NRHO, DRHO, NR, DR, RC))
        for arr in (frho, zr, rhor):
            for k in range(0, len(arr), 5):
                fh.write(" ".join("%.16e" % v for v in arr[k:k + 5]) + "\n")
