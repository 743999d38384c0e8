"""The tiling of the multi-plane RT sweep (rt_sweep_multi_kernel in
miso_amd/csrc/rt.hip), emulated on the CPU: workgroups that own 32 columns
of a ray over the short transverse axis and advance P planes per launch,
recomputing the upwind cone in a local plane buffer, store every swept node
exactly once, touch no index outside the ray's field, and give the same
values as the plane-by-plane sweep (rt_host.hpp:456-508 order) -- with a
stand-in node function, so the check is of the dependency argument: the
cone's stale edge never reaches an owned column.  The kernel itself is
checked bit for bit against the plane kernel on the GPU (test_rt.py)."""
import itertools

import numpy as np
import pytest

W, PMAX, ROWS, NT = 32, 16, 33, 1024


def _node(I00, I10, I01, I11, S00, S10, S01, S11, K00, K10, K01, K11, kc, sc):
    return (0.3 * I00 + 0.25 * I10 + 0.2 * I01 + 0.1 * I11
            + 0.01 * (S00 + 2 * S10 + 3 * S01 + 4 * S11)
            + 0.001 * (K00 - K10 + K01 * K11) + 0.05 * kc * sc)


def _ray(dims, margin, dom, t, tr):
    """The per-ray constants as misohip_rt_solve's set-up forms them."""
    It, Jt, Kt = dims
    sizes = [d - 2 * margin for d in dims]
    b0 = [margin - 1] * 3
    b1 = [b0[a] + sizes[a] for a in range(3)]
    strides = [Jt, 1, It * Jt] if tr else [1, It, It * Jt]
    beg, end = [0] * 3, [0] * 3
    for a in range(3):
        beg[a], end[a] = (b0[a] + 1, b1[a]) if t[a] > 0 else (b1[a] - 1, b0[a])
    b = 1 if dom == 0 else 0
    c = 1 if dom == 2 else 2
    idx = [0] * 3
    idx[dom], idx[b], idx[c] = beg[dom], min(beg[b], end[b]), min(beg[c], end[c])
    return dict(start=sum(idx[a] * strides[a] for a in range(3)), da=t[dom] * strides[dom],
                db=strides[b], dc=strides[c], ua=t[dom] * strides[dom], ub=t[b] * strides[b],
                uc=t[c] * strides[c], nb=abs(end[b] - beg[b]) + 1, nc=abs(end[c] - beg[c]) + 1,
                nplanes=abs(end[dom] - beg[dom]) + 1, sb=t[b], sc=t[c])


def _plane_sweep(R, I, K, S):
    for p in range(R["nplanes"]):
        for ic in range(R["nc"]):
            for ib in range(R["nb"]):
                c = R["start"] + p * R["da"] + ib * R["db"] + ic * R["dc"]
                c00 = c - R["ua"]
                c10, c01, c11 = c00 - R["ub"], c00 - R["uc"], c00 - R["ub"] - R["uc"]
                I[c] = _node(I[c00], I[c10], I[c01], I[c11], S[c00], S[c10], S[c01], S[c11],
                             K[c00], K[c10], K[c01], K[c11], K[c], S[c])


def _cone_sweep(R, I, K, S, P, n):
    X = W + PMAX
    CPT = (X * ROWS + NT - 1) // NT
    stores = {}

    def at(c):
        assert 0 <= c < n, c
        return c

    for p0 in range(0, R["nplanes"], P):
        start = I.copy()   # within a launch only owned nodes are written
        writes = {}
        for U0 in range(0, R["nb"], W):
            ext, pend = W + P, min(p0 + P, R["nplanes"])
            ncell = ext * (R["nc"] + 1)
            Il, Kl, Sl = (np.full((ROWS, X), np.nan) for _ in range(3))
            cells = []
            for tid in range(NT):
                for q in range(CPT):
                    idx = tid + q * NT
                    e, r = idx % ext, idx // ext
                    u, v = U0 - P + e, r - 1
                    role, off = 0, 0
                    if not (idx >= ncell or u < -1 or u >= R["nb"]):
                        ib = u if R["sb"] > 0 else R["nb"] - 1 - u
                        ic = v if R["sc"] > 0 else R["nc"] - 1 - v
                        off = R["start"] + ib * R["db"] + ic * R["dc"]
                        if u == -1 or v == -1:
                            role = 1
                        elif e == 0:
                            role = 4
                        else:
                            role = 3 if U0 <= u < U0 + W else 2
                    if role:
                        cells.append((role, off, e, r))
            for role, off, e, r in cells:
                c = at(off + (p0 - 1) * R["da"])
                Il[r, e], Kl[r, e], Sl[r, e] = start[c], K[c], S[c]
            for p in range(p0, pend):
                new = []
                for role, off, e, r in cells:
                    if role == 4:
                        new.append(None)
                        continue
                    c = at(off + p * R["da"])
                    if role == 1:
                        new.append((start[c], K[c], S[c]))
                        continue
                    v = _node(Il[r, e], Il[r, e - 1], Il[r - 1, e], Il[r - 1, e - 1],
                              Sl[r, e], Sl[r, e - 1], Sl[r - 1, e], Sl[r - 1, e - 1],
                              Kl[r, e], Kl[r, e - 1], Kl[r - 1, e], Kl[r - 1, e - 1], K[c], S[c])
                    new.append((v, K[c], S[c]))
                    if role == 3:
                        assert c not in writes
                        writes[c] = v
                        stores[c] = stores.get(c, 0) + 1
                for (role, off, e, r), val in zip(cells, new):
                    if val is not None:
                        Il[r, e], Kl[r, e], Sl[r, e] = val
        for c, v in writes.items():
            I[c] = v
    return stores


@pytest.mark.parametrize("dims", [(12, 40, 9), (70, 38, 7)])
@pytest.mark.parametrize("P", [5, 16])
def test_cone_tiling_matches_the_plane_sweep(dims, P):
    rng = np.random.default_rng(0)
    n = int(np.prod(dims))
    cases = 0
    for dom, tr in ((0, 1), (1, 0)):
        for t in itertools.product((1, -1), repeat=3):
            R = _ray(dims, 2, dom, t, tr)
            if R["nc"] + 1 > ROWS:
                continue
            K, S, I0 = rng.random(n), rng.random(n), rng.random(n)
            Ia = I0.copy()
            _plane_sweep(R, Ia, K, S)
            Ib = I0.copy()
            st = _cone_sweep(R, Ib, K, S, P, n)
            assert len(st) == R["nb"] * R["nc"] * R["nplanes"]
            assert all(v == 1 for v in st.values())
            assert np.array_equal(Ia, Ib)
            cases += 1
    assert cases > 0
