"""The wave-streamed SOR schedule of core_collapse_amd/csrc/sorws.hip, modelled on the CPU.

The model follows the kernel step for step: the same register rings and slots, the same loads (issued when
the kernel issues them, kWsPD steps ahead), stores, records, waits and publications, with each tile a
generator that a random scheduler advances one action at a time whenever the wait before that action is
satisfied.  Memory is checked for races the way the hardware could expose them: a store stays private to
its tile until the tile publishes; a load of a location that another tile has stored to and not yet
published returns NaN (a race whose value the sweep would then carry into the result), and a store to a
location another tile has loaded and not yet published (its load may still be in flight) fails the test.

Pinned against a direct lexicographic sweep written from the reference's loop (clamped neighbours,
M2:196-223) -- FP64: bit for bit; mixed: against the same sweep written as per-block iterative refinement
(phi = phi_b + d, FP64 residual of phi_b, FP32 sweeps of d), bit for bit, and within FP32 rounding of the
FP64 sweep.  Grids cover interior and edge tiles, a partial last sweep block, widths below the tile
height, one-row and one-column grids.
"""
import numpy as np
import pytest

W = 64
PD, PUB = 6, 12   # kWsPD, steps per publication
RS, RP = 12, 9    # kWsRS, kWsRP: record ring slots, steps a record is loaded ahead


def lex_sor(src, nsw, kx, ky, om, dtype=np.float64, x0=None):
    """Lexicographic SOR from x0 (zero), clamped neighbours, the kernel's operation order."""
    Ny, Nx = src.shape
    x = np.zeros((Ny, Nx), dtype) if x0 is None else x0.astype(dtype).copy()
    kx, ky, om = dtype(kx), dtype(ky), dtype(om)
    for _ in range(nsw):
        for r in range(Ny):
            for c in range(Nx):
                s = x[r, c]
                Wv = x[r, c - 1] if c > 0 else s
                Ev = x[r, c + 1] if c < Nx - 1 else s
                Sv = x[r - 1, c] if r > 0 else s
                Nv = x[r + 1, c] if r < Ny - 1 else s
                x[r, c] = s + om * (((src[r, c] + kx * (Ev + Wv)) + ky * (Nv + Sv)) - s)
    return x


def block_refined_sor(src, nsw, SW, kx, ky, om):
    """The mixed form: per block of SW sweeps, the FP64 residual of the block's base phi_b (clamped), SOR
    sweeps of an FP32 correction d from zero with that residual as source, then phi = phi_b + d in FP64."""
    Ny, Nx = src.shape
    phi = np.zeros((Ny, Nx))
    for s0 in range(0, nsw, SW):
        b = phi
        E = np.concatenate([b[:, 1:], b[:, -1:]], 1)
        Wb = np.concatenate([b[:, :1], b[:, :-1]], 1)
        N = np.concatenate([b[1:], b[-1:]], 0)
        S = np.concatenate([b[:1], b[:-1]], 0)
        res = (((src + kx * (E + Wb)) + ky * (N + S)) - b).astype(np.float32)
        d = lex_sor(res, min(SW, nsw - s0), kx, ky, om, np.float32)
        phi = b + d.astype(np.float64)
    return phi


class Mem:
    """A flat array with private (unpublished) stores and in-flight loads per tile."""

    def __init__(self, n, ntiles, dtype):
        self.vis = np.zeros(n, dtype)
        self.cur = np.zeros(n, dtype)
        self.writer = np.full(n, -1)
        self.inflight = np.zeros((ntiles, n), bool)

    def load(self, t, idx):
        idx = np.asarray(idx)
        assert idx.min() >= 0 and idx.max() < self.vis.size, "load out of bounds"
        w = self.writer[idx]
        out = np.where(w == t, self.cur[idx], self.vis[idx])
        out = np.where((w >= 0) & (w != t), np.nan, out)
        self.inflight[t, idx] = True
        return out

    def store(self, t, idx, val):
        idx = np.asarray(idx)
        if idx.size == 0:
            return
        assert idx.min() >= 0 and idx.max() < self.vis.size, "store out of bounds"
        others = np.delete(self.inflight[:, idx], t, axis=0)
        assert not others.any(), "store to a location with another tile's load in flight"
        assert not ((self.writer[idx] >= 0) & (self.writer[idx] != t)).any(), "two unpublished stores"
        self.cur[idx] = val
        self.writer[idx] = t

    def publish(self, t):
        m = self.writer == t
        self.vis[m] = self.cur[m]
        self.writer[m] = -1
        self.inflight[t] = False


class Model:
    def __init__(self, src, nsw, SW, kx, ky, om, mixed):
        Ny, Nx = src.shape
        self.Nx, self.Ny, self.nsw, self.SW = Nx, Ny, nsw, SW
        self.kx, self.ky, self.om, self.mixed = kx, ky, om, mixed
        self.R = np.float32 if mixed else np.float64
        self.NB = (Ny + SW - 1 + 63) // 64
        self.pitch = 64 * self.NB + 2
        self.K = Nx + 64 * self.NB + SW + 16
        self.ntiles = self.NB * ((nsw + SW - 1) // SW)
        self.RECN = 2 * SW + (4 if mixed else 0)  # record: SW values, SW source terms, (mixed) FP64 base, pad
        self.rec_stride = (Nx + SW + 160) * self.RECN
        n = self.K * self.pitch
        self.D = Mem(n, self.ntiles, np.float64)
        self.C = np.zeros(n)
        for r in range(Ny):
            for c in range(Nx):
                self.C[(c + r) * self.pitch + r] = src[r, c]
        self.rec = Mem(self.ntiles * self.rec_stride, self.ntiles, self.R)
        self.progress = np.zeros(self.ntiles, np.int64)

    def publish(self, t, g):
        self.D.publish(t)
        self.rec.publish(t)
        self.progress[t] = g

    # -------------------------------------------------------------- one tile, as the kernel runs it
    def tile(self, tile):
        SW, R, P = self.SW, self.R, self.pitch
        q, b = divmod(tile, self.NB)
        nm = min(SW, self.nsw - q * SW)
        gbase = 64 * b + 2 * SW * q
        pB = tile - 1 if b > 0 else -1
        pL = tile - self.NB if q > 0 else -1
        pD = tile - self.NB + 1 if (q > 0 and b + 1 < self.NB) else -1
        Tend = (self.Nx + 63 + nm + 11) // 12 * 12
        lane = np.arange(W)
        r0 = 64 * b + lane
        X = np.zeros((3, SW, W), R)
        Sr = np.zeros((3, SW, W), R)
        RX = np.zeros((RS, SW, W), R)   # record ring (broadcast form; the kernel keeps it spread over lanes)
        RSr = np.zeros((RS, SW, W), R)
        E, N, Cc, Bw = (np.zeros((PD, W)) for _ in range(4))
        RB = np.zeros((RS, W))
        RECN = self.RECN

        def prefetch(T):
            e = T % PD
            k = T + 64 * b
            E[e] = self.D.load(tile, (k + 1) * P + r0)
            N[e] = self.D.load(tile, (k + 1) * P + r0 + 1)
            Cc[e] = self.C[k * P + r0]
            if self.mixed:
                kw, rw = k - 2 * (nm - 1), r0 - (nm - 1)
                Bw[e] = self.D.load(tile, np.where((rw >= 0) & (kw >= 0), kw * P + rw, k * P + r0))

        def load_record(T, slot):  # the producer's local step T (uniform address: every lane the same)
            assert (T + 1) * RECN <= self.rec_stride, "record outside the producer's"
            base = pB * self.rec_stride + T * RECN
            v = self.rec.load(tile, np.arange(base, base + RECN))
            RX[slot] = v[:SW, None]
            RSr[slot] = v[SW:2 * SW, None]
            if self.mixed:
                RB[slot] = v[2 * SW:2 * SW + 2].view(np.float64)[0]

        def shr(v, old):
            o = np.empty(W, v.dtype)
            o[1:] = v[:-1]
            o[0] = old[0]
            return o

        state = {"Eprev": None, "Eprev2": np.zeros(W)}

        def step(T):
            cur, p1, p2, e = T % 3, (T + 2) % 3, (T + 1) % 3, T % PD
            r1, r2 = (T + 63) % RS, (T + 62) % RS
            kx, ky, om = R(self.kx), R(self.ky), R(self.om)
            for m in range(SW):
                c, r = T - lane - m, r0 - m
                if m == 0:
                    if self.mixed:
                        sb, eb, nb, wb = state["Eprev"], E[e].copy(), N[e].copy(), state["Eprev2"].copy()
                        s_b = shr(state["Eprev2"], RB[r1])  # lane l-1's base of the previous step
                        eb = np.where(c == self.Nx - 1, sb, eb)
                        wb = np.where(c == 0, sb, wb)
                        s_b = np.where(r == 0, sb, s_b)
                        nb = np.where(r == self.Ny - 1, sb, nb)
                        with np.errstate(invalid="ignore"):
                            src = (((Cc[e] + self.kx * (eb + wb)) + self.ky * (nb + s_b)) - sb).astype(R)
                        selfv = Ev = Nv = np.zeros(W, R)
                    else:
                        selfv, Ev, Nv, src = state["Eprev"].copy(), E[e].copy(), N[e].copy(), Cc[e].copy()
                else:
                    selfv = shr(X[p2][m - 1], RX[r2][m - 1])
                    Ev = shr(X[p1][m - 1], RX[r1][m - 1])
                    Nv = X[p1][m - 1].copy()
                    src = shr(Sr[p2][m - 1], RSr[r2][m - 1])
                Sv = shr(X[p1][m], RX[r1][m])
                Wv = X[p1][m].copy()
                if m == 0:
                    Ev = np.where(c == self.Nx - 1, selfv, Ev)
                Sv = np.where(r == 0, selfv, Sv)
                Nv = np.where(r == self.Ny - 1, selfv, Nv)
                with np.errstate(invalid="ignore"):
                    xn = selfv + om * (((src + kx * (Ev + Wv)) + ky * (Nv + Sv)) - selfv)
                active = (c >= 0) & (c < self.Nx) & (r >= 0) & (r < self.Ny) & (m < nm)
                X[cur][m] = np.where(active, xn, np.where(c < 0, Ev, Wv))
                Sr[cur][m] = src
                st = active & (m == nm - 1)
                if st.any():
                    val = Bw[e] + xn.astype(np.float64) if self.mixed else xn
                    self.D.store(tile, ((c + r) * P + r)[st], val[st])
            assert (T + 1) * RECN <= self.rec_stride, "record outside the tile's"
            base = tile * self.rec_stride + T * RECN
            v = np.concatenate([X[cur][:, 63], Sr[cur][:, 63]])
            if self.mixed:
                v = np.concatenate([v, np.array([state["Eprev"][63], 0.0]).view(np.float32)])
            self.rec.store(tile, np.arange(base, base + RECN), v)
            state["Eprev2"] = state["Eprev"]
            state["Eprev"] = E[e].copy()

        # ---- the kernel's sequence of waits and actions
        yield [(pL, gbase + PD), (pD, gbase + PD)]
        for T in range(PD):
            prefetch(T)
        state["Eprev"] = self.D.load(tile, (64 * b) * P + r0)
        if not self.mixed:  # W of sweep 0 at column 0: the value held at step -1, the E input of that step
            X[2][0] = state["Eprev"]
        if pB >= 0:
            yield [(pB, gbase + RP)]
            for Tp in range(62, 64 + RP):
                load_record(Tp, Tp % RS)
        for T in range(Tend):
            if pB >= 0:
                yield [(pB, gbase + T + RP + 1)]
                load_record(T + 64 + RP, (T + 64 + RP) % RS)
            step(T)
            yield [(pL, gbase + T + PD + 1), (pD, gbase + T + PD + 1)]
            prefetch(T + PD)
            if (T + 1) % PUB == 0:
                self.publish(tile, gbase + T + 1)
        self.publish(tile, 1 << 60)

    def run(self, rng):
        gens = {t: self.tile(t) for t in range(self.ntiles)}
        waits = {t: next(g) for t, g in gens.items()}
        while gens:
            ready = [t for t, w in waits.items() if all(p < 0 or self.progress[p] >= need for p, need in w)]
            assert ready, "deadlock"
            t = ready[rng.integers(len(ready))]
            try:
                waits[t] = next(gens[t])
            except StopIteration:
                del gens[t], waits[t]
        x = np.empty((self.Ny, self.Nx))
        for r in range(self.Ny):
            for c in range(self.Nx):
                x[r, c] = self.D.vis[(c + r) * self.pitch + r]
        return x


CASES = [(20, 9, 7, 4), (70, 130, 11, 4), (33, 65, 17, 8), (130, 70, 9, 8), (5, 3, 5, 4), (1, 70, 6, 4),
         (70, 1, 9, 8), (2, 2, 3, 8), (64, 64, 16, 4)]


def _problem(Nx, Ny, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((Ny, Nx)), 0.3, 0.2, 1.8


@pytest.mark.parametrize("Nx,Ny,nsw,SW", CASES)
def test_fp64_schedule_is_the_lexicographic_sweep(Nx, Ny, nsw, SW):
    src, kx, ky, om = _problem(Nx, Ny, Nx * 7 + Ny)
    want = lex_sor(src, nsw, kx, ky, om)
    for seed in range(2):
        got = Model(src, nsw, SW, kx, ky, om, mixed=False).run(np.random.default_rng(seed))
        assert np.array_equal(got, want), (seed, np.nanmax(np.abs(got - want)))


@pytest.mark.parametrize("Nx,Ny,nsw,SW", CASES)
def test_mixed_schedule_is_block_refined_sor(Nx, Ny, nsw, SW):
    src, kx, ky, om = _problem(Nx, Ny, Nx * 5 + Ny)
    want = block_refined_sor(src, nsw, SW, kx, ky, om)
    got = Model(src, nsw, SW, kx, ky, om, mixed=True).run(np.random.default_rng(3))
    assert np.array_equal(got, want), np.nanmax(np.abs(got - want))
    # and it is the FP64 sweep to FP32 rounding of the corrections
    ref = lex_sor(src, nsw, kx, ky, om)
    assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()


def test_the_model_catches_a_missing_wait():
    """Dropping the wait on the (b+1, q-1) producer lets a tile read cells before they are published."""
    src, kx, ky, om = _problem(70, 130, 1)
    m = Model(src, 11, 4, kx, ky, om, mixed=False)
    tile_gen = m.tile

    def no_diag_wait(t):
        for w in tile_gen(t):
            yield [pw for pw in w if pw[0] != t - m.NB + 1 or t % m.NB == m.NB - 1]
    m.tile = no_diag_wait
    got = None
    try:
        got = m.run(np.random.default_rng(0))
    except AssertionError:
        return
    assert not np.array_equal(got, lex_sor(src, 11, kx, ky, om))
