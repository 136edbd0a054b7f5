"""The cell partition of a general-path step launch (fdfk_general.hip:
gstep_job over gbox_cell, gring_box and gring_job), restated on the CPU.
The launch starts with the ring cells the step updates, in the ring box's
order rounded up to whole 256-lane workgroups; the band around the tiled
interior follows and leaves its ring cells to that first range.  Every cell
of the band must be taken exactly once, and the ring box minus its hole must
be exactly the cells of the reference's ring arrays (src/FDFK2D.f90:330-355,
plus the top ring with an absorbing top layer)."""
import pytest


def gbox_cell(t, x0, x1, z0, z1, hx0, hx1, hz0, hz1):
    W, H = x1 - x0 + 1, z1 - z0 + 1
    if hx0 > hx1 or hz0 > hz1:
        if t >= W * H: return None
        return (z0 + t % H, x0 + t // H)
    na, nb, nm = hz0 - z0, z1 - hz1, hz1 - hz0 + 1
    if t < na * W: return (z0 + t % na, x0 + t // na)
    t -= na * W
    if t < nb * W: return (hz1 + 1 + t % nb, x0 + t // nb)
    t -= nb * W
    nl = (hx0 - x0) * nm
    if t < nl: return (hz0 + t % nm, x0 + t // nm)
    t -= nl
    if t < (x1 - hx1) * nm: return (hz0 + t % nm, hx1 + 1 + t // nm)
    return None


def gbox_count(x0, x1, z0, z1, hx0, hx1, hz0, hz1):
    n = (x1 - x0 + 1) * (z1 - z0 + 1)
    if hx0 <= hx1 and hz0 <= hz1: n -= (hx1 - hx0 + 1) * (hz1 - hz0 + 1)
    return n


def check(nx, nz, npml, norder, top, tiled, top_ring):
    nbd = 2 * norder - 1
    zt = 1 - npml if top else 1
    NZ = nz + npml - zt + 1
    box = (1 - npml, nx + npml, zt, zt + NZ - 1)
    hole = (1, nx, 2 * norder, nz) if tiled else (1, 0, 1, 0)
    rb = [1 - nbd, nx + nbd, 1 - nbd if top_ring else 1, nz + nbd, nbd + 1, nx - nbd, nbd + 1 if top_ring else 1, nz - nbd]
    nq = gbox_count(*rb)
    iz_first = 1 if zt == 1 else zt + 1
    def edge(iz, ix):
        return ix < 2 - npml or ix > nx + npml - 1 or iz < iz_first or iz > nz + npml - 1
    def in_hole(iz, ix):
        return hole[0] <= ix <= hole[1] and hole[2] <= iz <= hole[3]
    def ring(iz, ix):   # membership of the ring arrays (gring >= 0)
        if 1 <= iz <= nz - nbd and (1 <= ix <= nbd or nx - nbd + 1 <= ix <= nx): return True
        if nz - nbd + 1 <= iz <= nz and 1 <= ix <= nx: return True
        if 1 <= iz <= nz and (1 - nbd <= ix <= 0 or nx + 1 <= ix <= nx + nbd): return True
        if nz + 1 <= iz <= nz + nbd and 1 - nbd <= ix <= nx + nbd: return True
        if top_ring and ((1 <= iz <= nbd and nbd + 1 <= ix <= nx - nbd) or (1 - nbd <= iz <= 0 and 1 - nbd <= ix <= nx + nbd)): return True
        return False
    nc = gbox_count(*box, *hole)
    nqb = (nq + 255) // 256 * 256
    seen = {}
    for t in range(nqb + nc + 255):
        if t < nqb:
            if t >= nq: continue
            c = gbox_cell(t, *rb)
            assert c is not None and ring(*c)
            if edge(*c) or in_hole(*c): continue
            seen[c] = seen.get(c, 0) + 1
        else:
            c = gbox_cell(t - nqb, box[0], box[1], box[2], box[3], *hole)
            if c is None: continue
            if not edge(*c) and ring(*c): continue
            seen[c] = seen.get(c, 0) + 1
    want = set((iz, ix) for ix in range(box[0], box[1] + 1) for iz in range(box[2], box[3] + 1) if not in_hole(iz, ix))
    assert set(seen) == want, (len(seen), len(want))
    assert all(v == 1 for v in seen.values())
    # the ring box minus its hole is exactly the ring arrays
    rset = set(gbox_cell(t, *rb) for t in range(nq))
    allr = set((iz, ix) for ix in range(box[0], box[1] + 1) for iz in range(box[2], box[3] + 1) if ring(iz, ix))
    assert rset == allr


@pytest.mark.parametrize("nx,nz,npml,norder,top,tiled,top_ring", [
    (121, 60, 15, 3, True, True, True), (121, 60, 15, 3, True, True, False), (121, 60, 15, 3, False, True, False),
    (90, 40, 12, 1, True, False, True), (110, 50, 16, 5, False, False, False), (30, 20, 10, 3, True, True, True),
    (12, 8, 6, 3, True, True, True), (100, 45, 14, 6, True, True, True)])
def test_step_launch_takes_every_cell_once(nx, nz, npml, norder, top, tiled, top_ring):
    check(nx, nz, npml, norder, top, tiled, top_ring)
