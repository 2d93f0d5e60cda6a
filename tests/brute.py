"""Brute-force global assembly in NumPy/SciPy, independent of the oracle's
node-walking code: every element's local matrix is scattered into one sparse
global matrix through explicit element -> node connectivity.

Conventions (chfem/src/femhmg/femhmg_3D.c:4-75, chfem/src/femhmg/utils.h:81-82):
element and periodic node index i = row + col*ny + layer*nx*ny; local node k of
element (r, c, l) sits at (r + DR[k], c + DC[k], l + DL[k]) wrapped periodically.
"""
import numpy as np
import scipy.sparse as sp

DR = np.array([1, 1, 0, 0, 1, 1, 0, 0])
DC = np.array([0, 1, 1, 0, 0, 1, 1, 0])
DL = np.array([0, 0, 0, 0, 1, 1, 1, 1])

# load case -> (seam axis, seam local nodes, jumping component)
# X: right face of the last column; Y: top of row 0; Z: near face of layer 0;
# YZ/XZ/XY jump one component across the Y / Z / Y seam.
LOADS = [(0, (1, 2, 5, 6), 0), (1, (2, 3, 6, 7), 1), (2, (0, 1, 2, 3), 2),
         (1, (2, 3, 6, 7), 2), (2, (0, 1, 2, 3), 0), (1, (2, 3, 6, 7), 0)]


def connectivity(nx, ny, nz):
    """(N, 8) periodic node index of each element's local nodes."""
    e = np.arange(nx * ny * nz)
    r, c, l = e % ny, (e // ny) % nx, e // (nx * ny)
    rr = (r[:, None] + DR[None]) % ny
    cc = (c[:, None] + DC[None]) % nx
    ll = (l[:, None] + DL[None]) % nz
    return rr + cc * ny + ll * nx * ny


def packed_keys(elem, nx, ny, nz):
    """uint16 key per node: bits [2k, 2k+1] = material of the element whose
    local node k is this node."""
    conn = connectivity(nx, ny, nz)
    keys = np.zeros(nx * ny * nz, dtype=np.uint32)
    for k in range(8):
        keys[conn[:, k]] |= elem.astype(np.uint32) << (2 * k)
    return keys.astype(np.uint16)


def assemble(elem, nx, ny, nz, Kloc, nd):
    """Global periodic stiffness: sum_e P_e^T K_{mat(e)} P_e, local DOF nd*k+a,
    global DOF nd*node+a.  Kloc: (nmat, 8nd, 8nd)."""
    conn = connectivity(nx, ny, nz)
    ldof = (nd * conn[:, :, None] + np.arange(nd)[None, None, :]).reshape(len(conn), 8 * nd)
    rows = np.repeat(ldof, 8 * nd, axis=1).ravel()
    cols = np.tile(ldof, (1, 8 * nd)).ravel()
    vals = Kloc[elem].reshape(len(conn), -1).ravel()
    n = nd * nx * ny * nz
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def seam_jump(elem, nx, ny, nz, nd, load):
    """Element-wise jump vectors j_e (N, 8nd): L at the seam nodes' jumping
    component on seam elements, zero elsewhere; L = the grid length across
    the seam (element units)."""
    axis, S, comp = LOADS[load]
    N = nx * ny * nz
    e = np.arange(N)
    r, c, l = e % ny, (e // ny) % nx, e // (nx * ny)
    on = [c == nx - 1, r == 0, l == 0][axis]
    L = [nx, ny, nz][axis]
    j = np.zeros((N, 8 * nd))
    for k in S:
        j[on, nd * k + (comp if nd == 3 else 0)] = L
    return j


def rhs(elem, nx, ny, nz, Kloc, nd, load):
    """b = -sum_e P_e^T K_e j_e."""
    conn = connectivity(nx, ny, nz)
    j = seam_jump(elem, nx, ny, nz, nd, load)
    f = -np.einsum("eij,ej->ei", Kloc[elem], j)
    b = np.zeros(nd * nx * ny * nz)
    ldof = (nd * conn[:, :, None] + np.arange(nd)[None, None, :]).reshape(len(conn), 8 * nd)
    np.add.at(b, ldof.ravel(), f.ravel())
    return b


def effective_column(elem, nx, ny, nz, CBloc, nd, load, x):
    """C[:, load] = (1/N) sum_e CB_e (P_e x + j_e)."""
    conn = connectivity(nx, ny, nz)
    ldof = (nd * conn[:, :, None] + np.arange(nd)[None, None, :]).reshape(len(conn), 8 * nd)
    ue = x[ldof] + seam_jump(elem, nx, ny, nz, nd, load)
    return np.einsum("eij,ej->i", CBloc[elem], ue) / (nx * ny * nz)


def fluid_maps(elem, nx, ny, nz):
    """Fluid keys (bit k: the element whose local node k is this node is pore,
    material 0), DOF ids (velocity nodes -- key 255 -- numbered first in node
    order, then border nodes 0 < key < 255) and the inverse map."""
    conn = connectivity(nx, ny, nz)
    N = nx * ny * nz
    fk = np.zeros(N, dtype=np.uint32)
    for k in range(8):
        fk[conn[:, k]] |= (elem == 0).astype(np.uint32) << k
    vel = np.flatnonzero(fk == 255)
    bord = np.flatnonzero((fk > 0) & (fk < 255))
    ids = np.zeros(N, dtype=np.uint32)
    ids[vel] = np.arange(len(vel))
    ids[bord] = len(vel) + np.arange(len(bord))
    d2n = np.concatenate([vel, bord]).astype(np.uint32)
    return fk.astype(np.uint8), ids, d2n, len(vel), len(bord)


def assemble_fluid(elem, nx, ny, nz, K32):
    """Stokes system over pore elements with the vector layout
    [u,v,w of velocity nodes | p of velocity and border nodes]; velocity is
    zero (no DOF) on border nodes; pressure-pressure block enters as -S."""
    conn = connectivity(nx, ny, nz)
    fk, ids, d2n, nvel, nb = fluid_maps(elem, nx, ny, nz)
    Kt = K32.reshape(32, 32).copy()
    Kt[24:, 24:] *= -1.0
    rows, cols, vals = [], [], []
    for e in np.flatnonzero(elem == 0):
        gd = np.full(32, -1, dtype=np.int64)
        for k in range(8):
            node = conn[e, k]
            if fk[node] == 255:
                gd[3 * k:3 * k + 3] = 3 * ids[node] + np.arange(3)
            gd[24 + k] = 3 * nvel + ids[node]
        ok = gd >= 0
        g = gd[ok]
        sub = Kt[np.ix_(ok, ok)]
        rows.append(np.repeat(g, len(g)))
        cols.append(np.tile(g, len(g)))
        vals.append(sub.ravel())
    n = 4 * nvel + nb
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))


def isotropic_C(E, nu):
    """Voigt stiffness (xx, yy, zz, yz, xz, xy; engineering shears)."""
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    C = np.zeros((6, 6))
    C[:3, :3] = lam
    C[range(3), range(3)] = lam + 2 * mu
    C[range(3, 6), range(3, 6)] = mu
    return C


def laminate_C(Cs, fractions, normal):
    """Exact effective stiffness of a laminate of homogeneous layers normal to
    axis `normal` (0 = x, 1 = y, 2 = z): the tangential strains and the normal
    tractions are uniform across layers."""
    Nn = {0: [0, 5, 4], 1: [1, 5, 3], 2: [2, 4, 3]}[normal]
    T = [i for i in range(6) if i not in Nn]
    f = np.asarray(fractions, dtype=float)
    inv_nn = sum(fi * np.linalg.inv(C[np.ix_(Nn, Nn)]) for fi, C in zip(f, Cs))
    a = sum(fi * np.linalg.inv(C[np.ix_(Nn, Nn)]) @ C[np.ix_(Nn, T)] for fi, C in zip(f, Cs))
    b = sum(fi * C[np.ix_(T, Nn)] @ np.linalg.inv(C[np.ix_(Nn, Nn)]) for fi, C in zip(f, Cs))
    c = sum(fi * (C[np.ix_(T, T)] - C[np.ix_(T, Nn)] @ np.linalg.inv(C[np.ix_(Nn, Nn)]) @ C[np.ix_(Nn, T)])
            for fi, C in zip(f, Cs))
    P = np.linalg.inv(inv_nn)
    Ceff = np.zeros((6, 6))
    Ceff[np.ix_(Nn, Nn)] = P
    Ceff[np.ix_(Nn, T)] = P @ a
    Ceff[np.ix_(T, Nn)] = b @ P
    Ceff[np.ix_(T, T)] = c + b @ P @ a
    return Ceff
