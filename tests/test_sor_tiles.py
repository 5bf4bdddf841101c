"""CPU proof of the tiled GPU SOR schedules (core_collapse_amd/csrc/poisson2d.hip, sorsched.cpp).

The tiled kernels claim to reproduce lexicographic Gauss-Seidel SOR (M2:207-218) with clamped boundaries; the
periodic / mixed-boundary kernels (time skew c = max(Nx, Ny), one thread per row) claim the same with wrap-around
neighbours.  Here each kernel's schedule is emulated cell by cell in Python (same tile geometry, same level order,
same per-cell formula) and compared bit for bit with the plain lexicographic loop; further checks show that no two
tiles of one level touch a cell that the other writes, that the dataflow launch's waits (taken from the library's
own host code) order every conflicting pair of tiles, and, for the sheared tiles, that the kernel's staging slots
cover the LDS box exactly once and that every value it carries in registers or reads a step early is the value
the step would read from the box.  (The wave-streamed SOR has its own model in test_sor_schedule.py.)
"""
import numpy as np
import pytest

from core_collapse_amd import mhd2d as M


def cell_update(phi, rhs, i, j, Nx, Ny, idx2, idy2, inv_denom, omega):
    pE = phi[j, min(i + 1, Nx - 1)]
    pW = phi[j, max(i - 1, 0)]
    pN = phi[min(j + 1, Ny - 1), i]
    pS = phi[max(j - 1, 0), i]
    pc = phi[j, i]
    pnew = ((pE + pW) * idx2 + (pN + pS) * idy2 - rhs[j, i]) * inv_denom
    phi[j, i] = pc + omega * (pnew - pc)


def lexicographic(rhs, nsw, idx2, idy2, omega):
    Ny, Nx = rhs.shape
    phi = np.zeros_like(rhs)
    inv = 1.0 / (2.0 * (idx2 + idy2))
    for _ in range(nsw):
        for j in range(Ny):
            for i in range(Nx):
                cell_update(phi, rhs, i, j, Nx, Ny, idx2, idy2, inv, omega)
    return phi


def tile_levels(Nx, Ny, nsw, W, WJ, S):
    """The host-side schedule of Poisson2D::init: non-empty tiles grouped by TS+TJ+TI."""
    nTS = (nsw + S - 1) // S
    nTI = (Nx + nsw - 1 + W - 1) // W
    nTJ = (Ny + nsw - 1 + WJ - 1) // WJ
    levels = {}
    for TS in range(nTS):
        s0, s1 = S * TS, min(S * TS + S, nsw)
        for TJ in range(nTJ):
            for TI in range(nTI):
                ok = any(W * TI - s + W - 1 >= 0 and W * TI - s <= Nx - 1 and WJ * TJ - s + WJ - 1 >= 0 and
                         WJ * TJ - s <= Ny - 1 for s in range(s0, s1))
                if ok:
                    levels.setdefault(TS + TJ + TI, []).append((TS, TJ, TI))
    return [levels[k] for k in sorted(levels)]


def tile_cells(tile, Nx, Ny, nsw, W, WJ, S):
    """(step, sweep, i, j) of every update of a tile, in the kernel's step order."""
    TS, TJ, TI = tile
    s0, s1 = S * TS, min(S * TS + S, nsw)
    out = []
    for d in range(W + WJ - 1):
        for s in range(s0, s1):
            for lane in range(W):
                lj = d - lane
                if not (0 <= lj < WJ):
                    continue
                i, j = W * TI - s + lane, WJ * TJ - s + lj
                if 0 <= i < Nx and 0 <= j < Ny:
                    out.append((d, s, i, j))
    return out


def tile_updates(i, j, s0, s1, TI, TJ, W, WJ):
    """The kernel's write-back predicate (tile_updates in poisson2d.hip)."""
    lo = max(s0, W * TI - i, WJ * TJ - j)
    hi = min(s1 - 1, W * TI - i + W - 1, WJ * TJ - j + WJ - 1)
    return lo <= hi


def run_tile(phi, rhs, tile, nsw, idx2, idy2, omega, W, WJ, S):
    """One tile of the kernel: copy the tile's box, update the copy in step order, write back the
    cells the predicate selects."""
    Ny, Nx = rhs.shape
    inv = 1.0 / (2.0 * (idx2 + idy2))
    TS, TJ, TI = tile
    s0, s1 = S * TS, min(S * TS + S, nsw)
    box = phi.copy()
    for (_, _, i, j) in tile_cells(tile, Nx, Ny, nsw, W, WJ, S):
        cell_update(box, rhs, i, j, Nx, Ny, idx2, idy2, inv, omega)
    ib0, ib1 = max(0, W * TI - (s1 - 1) - 1), min(Nx - 1, W * TI - s0 + W)
    jb0, jb1 = max(0, WJ * TJ - (s1 - 1) - 1), min(Ny - 1, WJ * TJ - s0 + WJ)
    changed = box != phi
    for j in range(jb0, jb1 + 1):
        for i in range(ib0, ib1 + 1):
            if tile_updates(i, j, s0, s1, TI, TJ, W, WJ):
                phi[j, i] = box[j, i]
    # every cell the tile changed lies in its box and passes the predicate
    cj, ci = np.nonzero(changed)
    assert all(ib0 <= i <= ib1 and jb0 <= j <= jb1 and tile_updates(i, j, s0, s1, TI, TJ, W, WJ)
               for i, j in zip(ci, cj))


def tiled(rhs, nsw, idx2, idy2, omega, W, WJ, S, reverse_levels=False):
    """The kernel's tiles in level order (each level's tiles in either order)."""
    Ny, Nx = rhs.shape
    phi = np.zeros_like(rhs)
    for lev in tile_levels(Nx, Ny, nsw, W, WJ, S):
        for tile in (reversed(lev) if reverse_levels else lev):
            run_tile(phi, rhs, tile, nsw, idx2, idy2, omega, W, WJ, S)
    return phi


@pytest.mark.parametrize("Nx,Ny,nsw,W,WJ,S", [(9, 7, 5, 4, 3, 2), (11, 13, 9, 4, 4, 3), (6, 10, 7, 3, 5, 4),
                                               (70, 66, 18, 64, 64, 16)])
def test_tiled_schedule_equals_lexicographic_bitwise(Nx, Ny, nsw, W, WJ, S):
    rng = np.random.default_rng(Nx * 100 + Ny)
    rhs = rng.standard_normal((Ny, Nx))
    idx2, idy2 = float(Nx * Nx), float(Ny * Ny)
    ref = lexicographic(rhs, nsw, idx2, idy2, 1.8)
    for rev in (False, True):
        got = tiled(rhs, nsw, idx2, idy2, 1.8, W, WJ, S, reverse_levels=rev)
        assert np.array_equal(got, ref)


@pytest.mark.parametrize("Nx,Ny,nsw,W,WJ,S", [(9, 7, 9, 4, 3, 2), (13, 11, 12, 4, 4, 3), (150, 140, 40, 64, 64, 16),
                                               (150, 300, 70, 64, 64, 32), (150, 300, 70, 64, 128, 32)])
def test_no_conflicts_within_a_level(Nx, Ny, nsw, W, WJ, S):
    for lev in tile_levels(Nx, Ny, nsw, W, WJ, S):
        writes, reads = [], []
        for tile in lev:
            cells = tile_cells(tile, Nx, Ny, nsw, W, WJ, S)
            w = {(i, j) for (_, _, i, j) in cells}
            r = set(w)
            for (_, _, i, j) in cells:
                r |= {(min(i + 1, Nx - 1), j), (max(i - 1, 0), j), (i, min(j + 1, Ny - 1)), (i, max(j - 1, 0))}
            writes.append(w)
            reads.append(r)
        for a in range(len(lev)):
            for b in range(len(lev)):
                if a != b:
                    assert not (writes[a] & reads[b]), (lev[a], lev[b])


def test_every_update_happens_once():
    Nx, Ny, nsw, W, WJ, S = 37, 29, 23, 8, 6, 4
    count = {}
    for lev in tile_levels(Nx, Ny, nsw, W, WJ, S):
        for tile in lev:
            for (_, s, i, j) in tile_cells(tile, Nx, Ny, nsw, W, WJ, S):
                count[(s, i, j)] = count.get((s, i, j), 0) + 1
    assert len(count) == nsw * Nx * Ny and set(count.values()) == {1}


def tiled_ghost(rhs, nsw, idx2, idy2, omega, W, WJ, S, reverse_levels=False):
    """The kernel as written since the ghost-margin form: the tile's box is copied with a one-cell
    margin; margin cells on a domain edge mirror the boundary cell (clamped neighbour) and are
    rewritten together with it, margin cells inside the domain are never read (NaN here, so a read
    would show); every stencil read is an unclamped box offset.  Each lane walks its column over the
    steps [lane + ljlo, lane + ljhi) that the kernel computes once per lane."""
    Ny, Nx = rhs.shape
    phi = np.zeros_like(rhs)
    inv = 1.0 / (2.0 * (idx2 + idy2))
    for lev in tile_levels(Nx, Ny, nsw, W, WJ, S):
        for tile in (reversed(lev) if reverse_levels else lev):
            TS, TJ, TI = tile
            s0, s1 = S * TS, min(S * TS + S, nsw)
            ib0, ib1 = max(0, W * TI - (s1 - 1) - 1), min(Nx - 1, W * TI - s0 + W)
            jb0, jb1 = max(0, WJ * TJ - (s1 - 1) - 1), min(Ny - 1, WJ * TJ - s0 + WJ)
            ni, nj = ib1 - ib0 + 1, jb1 - jb0 + 1
            assert ni + 2 <= W + S + 3 and nj + 2 <= WJ + S + 3
            box = np.full((nj + 2, ni + 2), np.nan)
            box[1:-1, 1:-1] = phi[jb0:jb1 + 1, ib0:ib1 + 1]
            if ib0 == 0:
                box[1:-1, 0] = box[1:-1, 1]
            if ib1 == Nx - 1:
                box[1:-1, -1] = box[1:-1, -2]
            if jb0 == 0:
                box[0, 1:-1] = box[1, 1:-1]
            if jb1 == Ny - 1:
                box[-1, 1:-1] = box[-2, 1:-1]
            for d in range(W + WJ - 1):
                for s in range(s0, s1):
                    for lane in range(W):
                        ci = W * TI - s + lane
                        if not (0 <= ci < Nx):
                            continue
                        ljlo, ljhi = max(0, s - WJ * TJ), min(WJ, Ny - WJ * TJ + s)
                        if not (lane + ljlo <= d < lane + ljhi):
                            continue
                        cj = WJ * TJ - s + d - lane
                        assert 0 <= cj < Ny
                        x, y = ci - ib0 + 1, cj - jb0 + 1
                        assert 1 <= x <= ni and 1 <= y <= nj  # LDS reads stay inside the box + margin
                        pc = box[y, x]
                        pnew = ((box[y, x + 1] + box[y, x - 1]) * idx2 + (box[y + 1, x] + box[y - 1, x]) * idy2
                                - rhs[cj, ci]) * inv
                        v = pc + omega * (pnew - pc)
                        assert np.isfinite(v)
                        box[y, x] = v
                        if ci == 0:
                            box[y, x - 1] = v
                        if ci == Nx - 1:
                            box[y, x + 1] = v
                        if cj == 0:
                            box[y - 1, x] = v
                        if cj == Ny - 1:
                            box[y + 1, x] = v
            for j in range(jb0, jb1 + 1):
                for i in range(ib0, ib1 + 1):
                    if tile_updates(i, j, s0, s1, TI, TJ, W, WJ):
                        phi[j, i] = box[j - jb0 + 1, i - ib0 + 1]
    return phi


@pytest.mark.parametrize("Nx,Ny,nsw,W,WJ,S", [(9, 7, 5, 4, 3, 2), (11, 13, 9, 4, 4, 3), (10, 9, 12, 4, 8, 4),
                                               (21, 19, 13, 8, 16, 8), (70, 40, 9, 64, 64, 4)])
def test_ghost_margin_kernel_equals_lexicographic_bitwise(Nx, Ny, nsw, W, WJ, S):
    rng = np.random.default_rng(Nx * 100 + Ny + 7)
    rhs = rng.standard_normal((Ny, Nx))
    idx2, idy2 = float(Nx * Nx), float(Ny * Ny)
    ref = lexicographic(rhs, nsw, idx2, idy2, 1.8)
    for rev in (False, True):
        assert np.array_equal(tiled_ghost(rhs, nsw, idx2, idy2, 1.8, W, WJ, S, reverse_levels=rev), ref)


# ------------------------------------------------------------------ periodic / mixed boundaries
# k_sor_wavefront_1wg and k_sor_skew_rows: cell (i, j) of sweep s at step i + j + c s, c = max(Nx, Ny).

def lexicographic_bc(rhs, nsw, idx2, idy2, omega, px, py):
    Ny, Nx = rhs.shape
    phi = np.zeros_like(rhs)
    inv = 1.0 / (2.0 * (idx2 + idy2))
    bm = lambda i, n, p: (i % n) if p else min(max(i, 0), n - 1)  # noqa: E731
    for _ in range(nsw):
        for j in range(Ny):
            for i in range(Nx):
                pE, pW = phi[j, bm(i + 1, Nx, px)], phi[j, bm(i - 1, Nx, px)]
                pN, pS = phi[bm(j + 1, Ny, py), i], phi[bm(j - 1, Ny, py), i]
                pc = phi[j, i]
                pnew = ((pE + pW) * idx2 + (pN + pS) * idy2 - rhs[j, i]) * inv
                phi[j, i] = pc + omega * (pnew - pc)
    return phi


def skewed_bc(rhs, nsw, idx2, idy2, omega, px, py, order):
    """The time-skewed schedule; within a step the cells are updated in `order` (0: rows up,
    1: rows down, 2: all reads before all writes) -- equal results show they are independent."""
    Ny, Nx = rhs.shape
    phi = np.zeros_like(rhs)
    inv = 1.0 / (2.0 * (idx2 + idy2))
    bm = lambda i, n, p: (i % n) if p else min(max(i, 0), n - 1)  # noqa: E731
    c = max(Nx, Ny)
    for T in range(c * (nsw - 1) + Nx + Ny - 1):
        pending = []
        for j in (range(Ny) if order != 1 else range(Ny - 1, -1, -1)):
            s, i = divmod(T - j, c)
            if T - j < 0 or s >= nsw or i >= Nx:
                continue
            pE, pW = phi[j, bm(i + 1, Nx, px)], phi[j, bm(i - 1, Nx, px)]
            pN, pS = phi[bm(j + 1, Ny, py), i], phi[bm(j - 1, Ny, py), i]
            pc = phi[j, i]
            v = pc + omega * (((pE + pW) * idx2 + (pN + pS) * idy2 - rhs[j, i]) * inv - pc)
            if order == 2:
                pending.append((j, i, v))
            else:
                phi[j, i] = v
        for j, i, v in pending:
            phi[j, i] = v
    return phi


def skew_rows(rhs, nsw, idx2, idy2, omega, px, py, PD, loads_first=False):
    """k_sor_skew_rows step by step: one thread per row, west from the thread's previous result,
    south through the double-buffered LDS exchange, periodic seams from registers and the rings of
    rows 0 and Ny-1, the rest loaded PD steps ahead (a load sees memory as it is when issued).
    loads_first issues a step's loads before the step's stores: the kernel orders them only within
    a thread, so equal results show no load reads a cell stored in the same step."""
    Ny, Nx = rhs.shape
    N, RING = Nx * Ny, 2048
    phi = np.zeros(N + 64)
    r = rhs.reshape(-1)
    inv = 1.0 / (2.0 * (idx2 + idy2))
    c = max(Nx, Ny)
    steps = c * (nsw - 1) + Nx + Ny - 1
    xs = np.zeros((2, 64)); ring0 = np.zeros(RING); ringN = np.zeros(RING)
    out = np.zeros(64); first = np.zeros(64); wlast = np.zeros(64)
    pos = [[-j, 0] for j in range(64)]

    def issue(j, ii, ss):
        act = j < Ny and ss < nsw and 0 <= ii < Nx
        k = j * Nx + (ii if act else 0)
        dump = N + j
        return (phi[k + 1 if act and ii + 1 < Nx else dump], phi[k + Nx if act and j + 1 < Ny else dump],
                phi[k if act else dump], r[k if act else 0])

    q = [[None] * 64 for _ in range(PD)]
    pre = [[-j, 0] for j in range(64)]
    for u in range(PD):
        for j in range(64):
            q[u][j] = issue(j, *pre[j])
            pre[j][0] += 1
            if pre[j][0] == c:
                pre[j] = [0, pre[j][1] + 1]
    for T in range(steps):
        u, par = T % PD, T & 1
        stores = []
        for j in range(64):
            i, s = pos[j]
            E0, N0, pc, rr = q[u][j]
            act = j < Ny and s < nsw and 0 <= i < Nx
            Sr = ((0.0 if s == 0 else ringN[(T - (c - Ny + 1)) % RING]) if py else pc)
            S = xs[par ^ 1][j - 1] if j > 0 else Sr
            Nb = N0 if j < Ny - 1 else (ring0[(T - (Ny - 1)) % RING] if py else pc)
            E = E0 if i < Nx - 1 else (first[j] if px else pc)
            W = out[j] if i > 0 else ((0.0 if s == 0 else wlast[j]) if px else pc)
            with np.errstate(all="ignore"):
                v = pc + omega * (((E + W) * idx2 + (Nb + S) * idy2 - rr) * inv - pc)
            stores.append((j * Nx + i if act else N + j, v))
            xs[par][j] = v
            if j == 0:
                ring0[T % RING] = v
            if j == Ny - 1:
                ringN[T % RING] = v
            first[j] = v if i == 0 else first[j]
            wlast[j] = v if i == Nx - 1 else wlast[j]
            out[j] = v
        if not loads_first:
            for k, v in stores:
                phi[k] = v
        for j in range(64):
            q[u][j] = issue(j, *pre[j])
            pre[j][0] += 1
            if pre[j][0] == c:
                pre[j] = [0, pre[j][1] + 1]
            pos[j][0] += 1
            if pos[j][0] == c:
                pos[j] = [0, pos[j][1] + 1]
        if loads_first:
            for k, v in stores:
                phi[k] = v
    return phi[:N].reshape(Ny, Nx)


@pytest.mark.parametrize("Nx,Ny,px,py,nsw", [(12, 12, 1, 1, 5), (10, 17, 1, 1, 4), (19, 8, 1, 1, 4), (15, 9, 1, 0, 4),
                                             (9, 14, 0, 1, 4), (13, 13, 0, 0, 3), (2, 5, 1, 1, 3)])
def test_time_skewed_periodic_schedule_equals_lexicographic_bitwise(Nx, Ny, px, py, nsw):
    rng = np.random.default_rng(Nx * 31 + Ny)
    rhs = 4.0 * np.pi * (rng.uniform(0, 1, (Ny, Nx)) - 0.5)
    want = lexicographic_bc(rhs, nsw, float(Nx * Nx), float(Ny * Ny), 1.8, px, py)
    for order in (0, 1, 2):
        assert np.array_equal(skewed_bc(rhs, nsw, float(Nx * Nx), float(Ny * Ny), 1.8, px, py, order), want)


@pytest.mark.parametrize("Nx,Ny,px,py,nsw,PD", [(18, 18, 1, 1, 4, 8), (20, 30, 1, 0, 3, 8), (40, 7, 1, 1, 3, 8),
                                                (12, 12, 1, 1, 5, 2), (9, 14, 0, 1, 4, 4), (25, 19, 0, 0, 3, 8)])
def test_row_thread_kernel_equals_lexicographic_bitwise(Nx, Ny, px, py, nsw, PD):
    if max(Nx, Ny) < 2 * PD + 2:
        pytest.skip("the kernel needs c >= 2 PD + 2")
    rng = np.random.default_rng(Nx * 7 + Ny)
    rhs = 4.0 * np.pi * (rng.uniform(0, 1, (Ny, Nx)) - 0.5)
    want = lexicographic_bc(rhs, nsw, float(Nx * Nx), float(Ny * Ny), 1.8, px, py)
    for loads_first in (False, True):
        got = skew_rows(rhs, nsw, float(Nx * Nx), float(Ny * Ny), 1.8, px, py, PD, loads_first)
        assert np.array_equal(got, want), loads_first


# ------------------------------------------------------------------ dataflow schedule
# The default tiled solve is one persistent launch: workgroups take tiles in launch (level) order from
# an atomic counter and start a tile once every tile it waits for has set its done flag.  The waits
# come from the library itself (cc_sor_flow_schedule, the host code that builds the device tables).


def _footprint(cells, Nx, Ny):
    w = {(i, j) for (_, _, i, j) in cells}
    r = set(w)
    for (i, j) in w:
        r |= {(min(i + 1, Nx - 1), j), (max(i - 1, 0), j), (i, min(j + 1, Ny - 1)), (i, max(j - 1, 0))}
    return w, r


FLOW_CASES = [(9, 7, 9, 4, 3, 2), (13, 11, 12, 4, 4, 3), (11, 13, 9, 4, 4, 3), (20, 6, 30, 5, 2, 7),
              (6, 20, 17, 3, 6, 5), (150, 140, 40, 64, 64, 16), (100, 200, 70, 64, 128, 32)]


def _geometry(sheared):
    """(levels, cells of a tile, the kernel's tile body) of rectangular or sheared tiles."""
    if sheared:
        return sheared_levels, sheared_cells, sheared_run_tile
    return tile_levels, tile_cells, run_tile


@pytest.mark.parametrize("sheared", [False, True])
@pytest.mark.parametrize("Nx,Ny,nsw,W,WJ,S", FLOW_CASES)
def test_dataflow_waits_order_every_conflicting_pair(Nx, Ny, nsw, W, WJ, S, sheared):
    """Two tiles of which one writes a cell the other reads are never in flight together: the later
    one (in level order) waits for the earlier one, directly or through a chain of waits.  Pairs
    without such a cell commute, so every execution the waits allow gives the level order's result."""
    levels, cells_of, _ = _geometry(sheared)
    tiles, deps = M.sor_flow_schedule(Nx, Ny, nsw, W, WJ, S, sheared=sheared)
    assert tiles == [t for lev in levels(Nx, Ny, nsw, W, WJ, S) for t in lev]
    n = len(tiles)
    anc = []
    for k in range(n):
        a = set()
        for d in deps[k]:
            # a wait points to a tile below in every coordinate, hence earlier in launch order: the
            # tiles a workgroup waits for were handed out before its own, so the waits cannot deadlock
            assert d < k and tiles[d] != tiles[k] and all(tiles[d][c] <= tiles[k][c] for c in range(3))
            a |= {d} | anc[d]
        anc.append(a)
    fp = [_footprint(cells_of(t, Nx, Ny, nsw, W, WJ, S), Nx, Ny) for t in tiles]
    for b in range(n):
        for a in range(b):
            if fp[a][0] & fp[b][1] or fp[b][0] & fp[a][1]:
                assert a in anc[b], (tiles[a], tiles[b])
    # the waits are minimal: no wait is implied by another one
    for k in range(n):
        for d in deps[k]:
            assert not any(d in anc[e] for e in deps[k] if e != d)


@pytest.mark.parametrize("sheared", [False, True])
@pytest.mark.parametrize("Nx,Ny,nsw,W,WJ,S", FLOW_CASES[:5])
def test_dataflow_random_orders_equal_lexicographic_bitwise(Nx, Ny, nsw, W, WJ, S, sheared):
    _, _, body = _geometry(sheared)
    tiles, deps = M.sor_flow_schedule(Nx, Ny, nsw, W, WJ, S, sheared=sheared)
    rng = np.random.default_rng(Nx * 31 + Ny)
    rhs = rng.standard_normal((Ny, Nx))
    idx2, idy2 = float(Nx * Nx), float(Ny * Ny)
    ref = lexicographic(rhs, nsw, idx2, idy2, 1.8)
    for _ in range(3):
        phi = np.zeros_like(rhs)
        done = [False] * len(tiles)
        while not all(done):
            ready = [k for k in range(len(tiles)) if not done[k] and all(done[d] for d in deps[k])]
            k = ready[rng.integers(len(ready))]
            body(phi, rhs, tiles[k], nsw, idx2, idy2, 1.8, W, WJ, S)
            done[k] = True
        assert np.array_equal(phi, ref)


def test_dataflow_schedule_rejects_bad_arguments():
    from core_collapse_amd._lib import CCError
    with pytest.raises(CCError):
        M.sor_flow_schedule(0, 8, 10)
    with pytest.raises(CCError):
        M.sor_flow_schedule(8, 8, 10, W=0)


# ------------------------------------------------------------------ sheared tiles
# A cell (i, j) of sweep s sits at k = i + j + 2s on the skewed time axis; every value it reads was
# written at k - 1 or k - 2.  A sheared tile is a box of S sweeps x W values of i' = i + s x K values
# of k: at step d (k = K TK + d) lane l of sweep s updates i = W TI + l - s, j = k - W TI - l - s, so
# every lane works at every step of an interior tile.  Its LDS box is kept in sheared coordinates
# (x, y) = (i - ib + 1, i + j - ub + 1): E/W neighbours are (x +- 1, y +- 1), N/S are (x, y +- 1).


def sheared_levels(Nx, Ny, nsw, W, K, S):
    """Non-empty sheared tiles (TS, TK, TI) grouped by TS + TK + TI (the host schedule)."""
    nTS = (nsw + S - 1) // S
    nTI = (Nx + nsw - 1 + W - 1) // W
    nTK = (Nx + Ny + 2 * nsw - 3 + K - 1) // K
    levels = {}
    for TS in range(nTS):
        s0, s1 = S * TS, min(S * TS + S, nsw)
        for TK in range(nTK):
            for TI in range(nTI):
                ok = False
                for s in range(s0, s1):
                    imin, imax = max(0, W * TI - s), min(Nx - 1, W * TI + W - 1 - s)
                    if imin > imax:
                        continue
                    jmin, jmax = max(0, K * TK - imax - 2 * s), min(Ny - 1, K * TK + K - 1 - imin - 2 * s)
                    if jmin <= jmax:
                        ok = True
                        break
                if ok:
                    levels.setdefault(TS + TK + TI, []).append((TS, TK, TI))
    return [levels[k] for k in sorted(levels)]


def sheared_cells(tile, Nx, Ny, nsw, W, K, S):
    TS, TK, TI = tile
    s0, s1 = S * TS, min(S * TS + S, nsw)
    out = []
    for d in range(K):
        for s in range(s0, s1):
            for lane in range(W):
                i, j = W * TI + lane - s, K * TK + d - W * TI - lane - s
                if 0 <= i < Nx and 0 <= j < Ny:
                    out.append((d, s, i, j))
    return out


def sheared_updates(i, j, s0, s1, TI, TK, W, K):
    """The kernel's write-back predicate: some sweep s in [s0, s1) puts (i, j) inside the tile."""
    u = i + j
    lo = max(s0, W * TI - i, (K * TK - u + 1) >> 1)
    hi = min(s1 - 1, W * TI + W - 1 - i, (K * TK + K - 1 - u) >> 1)
    return lo <= hi


def sheared_slots(BX, BY):
    """The kernel's staging slots (sh_slot): slot c -> (y, x, r), r the grid row of the box.  Grid row r holds
    x in [max(0, BX-1-r), min(BX-1, BX+BY-2-r)]; when BY >= BX - 1 rows v < BX - 1 and v + BY share a slot
    row, else a slot row is one grid row and slots outside the box are skipped (y out of range)."""
    paired = BY >= BX - 1
    n = (BY if paired else BX + BY - 1) * BX
    for c in range(n):
        x, v = c % BX, c // BX
        r = v + BY if (paired and v < BX - 1 and x < BX - 1 - v) else v
        y = x + r - (BX - 1)
        if 0 <= y < BY:
            yield y, x, r


def sheared_run_tile(phi, rhs, tile, nsw, idx2, idy2, omega, W, K, S):
    """One sheared tile as the kernel runs it: stage the in-domain cells of the box through the staging slots,
    K steps (all reads of a step before its writes; no cell written at a step is read by another lane at that
    step), write back the cells the predicate selects.  In an interior tile the step uses the kernel's
    registers: the old value is the N neighbour read at the previous step, the S neighbour the lane's previous
    result, the source term read a step ahead; each is checked against the box value it stands for."""
    Ny, Nx = rhs.shape
    inv = 1.0 / (2.0 * (idx2 + idy2))
    TS, TK, TI = tile
    s0, s1 = S * TS, min(S * TS + S, nsw)
    ib, ub = W * TI - (s0 + S - 1), K * TK - 2 * (s0 + S - 1)
    BX, BY = W + S + 1, K + 2 * S
    box = np.full((BY, BX), np.nan)
    i0, u0 = ib - 1, ub - 1
    jlo = u0 - i0 - (BX - 1)
    staged = {}
    for y, x, r in sheared_slots(BX, BY):
        i, j = i0 + x, jlo + r
        assert i + j - u0 == y
        staged[(y, x)] = staged.get((y, x), 0) + 1
        if 0 <= i < Nx and 0 <= j < Ny:
            box[y, x] = phi[j, i]
    assert len(staged) == BX * BY and set(staged.values()) == {1}
    interior = ib >= 1 and ib + W + S - 2 <= Nx - 2 and \
        K * TK - (W * TI + W - 1) - (s0 + S - 1) >= 1 and K * TK + K - 1 - W * TI - s0 <= Ny - 2
    regs = {}  # interior tiles: (s, lane) -> [prev, nold, rn] after the lane's last step
    for d in range(K):
        new, reads, writes = [], set(), set()
        old = {k: list(v) for k, v in regs.items()}  # the registers as the step starts
        for s in range(s0, s1):
            for lane in range(W):
                i, j = W * TI + lane - s, K * TK + d - W * TI - lane - s
                if not (0 <= i < Nx and 0 <= j < Ny):
                    assert not interior
                    continue
                x, y = i - ib + 1, i + j - ub + 1
                assert 1 <= x <= BX - 2 and 1 <= y <= BY - 2
                e = (y + 1, x + 1) if i < Nx - 1 else (y, x)
                w = (y - 1, x - 1) if i > 0 else (y, x)
                n = (y + 1, x) if j < Ny - 1 else (y, x)
                so = (y - 1, x) if j > 0 else (y, x)
                if interior:
                    assert i < Nx - 1 and i > 0 and j < Ny - 1 and j > 0
                pc, pE, pW, pN, pS, r = box[y, x], box[e], box[w], box[n], box[so], rhs[j, i]
                if interior and d > 0:
                    prev, nold, rn = old[(s, lane)]
                    assert (nold, prev, rn) == (pc, pS, r), (tile, d, s, lane)
                pnew = ((pE + pW) * idx2 + (pN + pS) * idy2 - r) * inv
                v = pc + omega * (pnew - pc)
                assert np.isfinite(v)
                new.append(((y, x), v))
                if interior:
                    regs.setdefault((s, lane), [0.0] * 3)
                    regs[(s, lane)][0] = v
                    regs[(s, lane)][1] = pN
                reads |= {e, w, n, so} - {(y, x)}  # a clamped neighbour is the cell itself
                writes.add((y, x))
        assert len(writes) == len(new) and not (reads & writes), (tile, d)
        for c, v in new:
            box[c] = v
        if interior and d + 1 < K:  # the source term of the next step, read a step ahead
            for s in range(s0, s1):
                for lane in range(W):
                    i, j = W * TI + lane - s, K * TK + d + 1 - W * TI - lane - s
                    regs[(s, lane)][2] = rhs[j, i]
    for y in range(BY):
        for x in range(BX):
            i, u = ib - 1 + x, ub - 1 + y
            j = u - i
            if 0 <= i < Nx and 0 <= j < Ny:
                upd = sheared_updates(i, j, s0, s1, TI, TK, W, K)
                if upd:
                    phi[j, i] = box[y, x]
                else:
                    assert box[y, x] == phi[j, i] or np.isnan(box[y, x])


SHEARED_CASES = [(9, 7, 5, 4, 3, 2), (11, 13, 9, 4, 4, 3), (6, 10, 7, 3, 5, 4), (13, 11, 12, 4, 6, 2),
                 (20, 6, 30, 5, 2, 7), (70, 66, 18, 16, 8, 4), (60, 50, 9, 8, 12, 3), (90, 70, 10, 16, 40, 2)]


@pytest.mark.parametrize("Nx,Ny,nsw,W,K,S", SHEARED_CASES)
def test_sheared_tiles_equal_lexicographic_bitwise(Nx, Ny, nsw, W, K, S):
    rng = np.random.default_rng(Nx * 100 + Ny + 7)
    rhs = rng.standard_normal((Ny, Nx))
    idx2, idy2 = float(Nx * Nx), float(Ny * Ny)
    ref = lexicographic(rhs, nsw, idx2, idy2, 1.8)
    count = {}
    for lev in sheared_levels(Nx, Ny, nsw, W, K, S):
        for tile in lev:
            for (_, s, i, j) in sheared_cells(tile, Nx, Ny, nsw, W, K, S):
                count[(s, i, j)] = count.get((s, i, j), 0) + 1
    assert len(count) == nsw * Nx * Ny and set(count.values()) == {1}
    for rev in (False, True):
        phi = np.zeros_like(rhs)
        for lev in sheared_levels(Nx, Ny, nsw, W, K, S):
            for tile in (reversed(lev) if rev else lev):
                sheared_run_tile(phi, rhs, tile, nsw, idx2, idy2, 1.8, W, K, S)
        assert np.array_equal(phi, ref)
