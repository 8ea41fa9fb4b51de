"""Schedule model of the band -> tridiagonal bulge chasing (hip/chase.hip, k_chase), on the CPU.

The kernel keeps each sweep's current 32-column block in work-group memory and makes three kinds of
global-memory events per step: the load of the new diagonal block, the mid-step write-back of the
block the chase leaves (all but its first column), and the end-of-step write-back of the first column
of the current block (the whole block on the last step) followed by the progress flag.  Step j of
sweep s may load once sweep s-1 has published step j+1.  This model replays the same band-storage
arithmetic with those local copies and events, in random interleavings allowed by that rule, and
checks the result against strictly sequential sweeps bit for bit; without the end-of-step column
write-back the same rule must fail (the model detects the race).
"""
import numpy as np

def larfg(x):
    alpha = x[0]
    xn2 = float(np.vdot(x[1:], x[1:]).real)
    if xn2 == 0 and alpha.imag == 0:
        return alpha.real, 0j, None
    nrm = np.sqrt(abs(alpha) ** 2 + xn2)
    beta = -nrm if alpha.real >= 0 else nrm
    tau = (beta - alpha.real) / beta - 1j * alpha.imag / beta
    scl = 1.0 / (alpha - beta)
    return beta, tau, scl

class Sweep:
    def __init__(self, s, n, B):
        self.s, self.j = s, 0
        self.ns = (n - 1 - s + B - 1) // B
        self.phase = 0
        self.prev = None
        self.nxt = None

def run(AB0, n, B, order, seed, dep, eager):
    LDB = 2 * B
    G = AB0.copy()
    rng = np.random.default_rng(seed)
    sw = [Sweep(s, n, B) for s in range(n - 1)]
    done = [0] * (n - 1)
    V = {}

    def reflector(col, Lx):
        x = col[:Lx].copy()
        beta, tau, scl = larfg(x)
        v = np.zeros(B, complex); v[0] = 1
        if scl is not None:
            v[1:Lx] = scl * x[1:Lx]
        col[:Lx] = 0; col[0] = beta
        return v, tau

    def ev_load(w):
        s, j = w.s, w.j
        c = s if j == 0 else s + 1 + (j - 1) * B
        r0 = s + 1 + j * B
        L = min(B, n - r0)
        w.cur = G[r0:r0 + L].copy()
        if j == 0:
            w.prev = np.zeros((B, LDB), complex)
            w.prev[0] = G[s].copy()

    def ev_mid(w):
        s, j, n_ = w.s, w.j, n
        c = s if j == 0 else s + 1 + (j - 1) * B
        r0 = s + 1 + j * B
        L = min(B, n - r0); r1 = r0 + L
        L2 = min(B, n - r1) if r1 < n else 0
        off0 = r0 - c
        prev, cur = w.prev, w.cur
        if w.nxt is None:
            v, tau = reflector(prev[0, off0:], L)
        else:
            v, tau = w.nxt
        V[(s, j)] = (v.copy(), tau)
        # (i) left, columns c+t, t = 1 .. B-1 (entries prev[t][B + i - t])
        if j > 0:
            for t in range(1, B):
                a = np.array([prev[t, B + i - t] if i < L else 0 for i in range(B)])
                wv = np.sum(np.conj(v) * a)
                ct = np.conj(tau) * wv
                for i in range(L):
                    prev[t, B + i - t] = a[i] - v[i] * ct
        # (iii) right on the block below: C(i, k) = cur[k][B + i - k]
        for i in range(L2):
            cc = np.array([cur[k, B + i - k] if k < L else 0 for k in range(B)])
            u = np.sum(cc * v)
            tu = tau * u
            for k in range(L):
                cur[k, B + i - k] = cc[k] - tu * np.conj(v[k])
        # (ii) y = D v
        y = np.zeros(B, complex)
        for i in range(L):
            for k in range(L):
                d = cur[k, i - k] if i >= k else np.conj(cur[i, k - i])
                y[i] += d * v[k]
        # early reflector of the next step
        if j + 1 < w.ns and eager:
            w.nxt = reflector(cur[0, B:], L2)
        else:
            w.nxt = None
        # write-back of prev
        if j == 0:
            G[s] = prev[0]
        else:
            for t in range(1 if eager else 0, B):
                G[c + t] = prev[t]
        # two-sided
        al = float(np.sum(np.conj(v[:L]) * y[:L]).real)
        t2 = 0.5 * abs(tau) ** 2 * al
        wv = tau * y - t2 * v
        for i in range(L):
            for k in range(i + 1):
                val = cur[k, i - k] - wv[i] * np.conj(v[k]) - v[i] * np.conj(wv[k])
                if i == k:
                    val = val.real + 0j
                cur[k, i - k] = val

    def ev_final(w):
        s, j = w.s, w.j
        r0 = s + 1 + j * B
        L = min(B, n - r0)
        if j == w.ns - 1:
            G[r0:r0 + L] = w.cur
        elif eager:
            G[r0] = w.cur[0]
        w.prev = w.cur
        w.j += 1
        done[s] = w.j

    active = list(range(n - 1))
    while True:
        cand = []
        for s in range(n - 1):
            w = sw[s]
            if w.j >= w.ns:
                continue
            if w.phase == 0:
                need = min(w.j + dep, sw[s - 1].ns) if s > 0 else 0
                if s > 0 and done[s - 1] < need:
                    continue
            cand.append(s)
        if not cand:
            break
        s = cand[rng.integers(len(cand))] if order == "random" else (cand[0] if order == "seq" else cand[-1])
        w = sw[s]
        if w.phase == 0: ev_load(w); w.phase = 1
        elif w.phase == 1: ev_mid(w); w.phase = 2
        else: ev_final(w); w.phase = 0
    assert all(w.j == w.ns for w in sw)
    return G, V



def _band(n, B, rng):
    A = np.tril(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    AB = np.zeros((n, 2 * B), complex)
    for c in range(n):
        for d in range(B + 1):
            if c + d < n:
                AB[c, d] = A[c + d, c]
        AB[c, 0] = AB[c, 0].real
    return AB


def test_chase_schedule_dependency_rule():
    rng = np.random.default_rng(7)
    for n, B in [(23, 4), (29, 8)]:
        AB = _band(n, B, rng)
        ref, _ = run(AB, n, B, "seq", 0, 10**9, True)
        for seed in range(4):
            G, _ = run(AB, n, B, "random", seed, 2, True)
            assert np.array_equal(G, ref), (n, B, seed)
        G, _ = run(AB, n, B, "greedy", 0, 2, True)
        assert np.array_equal(G, ref)
        # the same dependency without the end-of-step column write-back is a race
        bad = [np.array_equal(run(AB, n, B, "random", sd, 2, False)[0], ref) for sd in range(3)]
        assert not all(bad)
