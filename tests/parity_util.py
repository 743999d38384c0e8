"""Shared helpers of the parity tests: run the same seeded input through the
oracle (MISO order, host) and through the HIP backend (C ABI), compare."""
import numpy as np

FIELDS = ("ro", "vx", "vy", "vz", "bx", "by", "bz", "ei", "ph")


def interior_slices(grid):
    return (slice(None),) + grid.interior()


def rel_linf(a, b, floor=1e-300):
    """max over cells of |a-b| / max|b| per field (relative L-infinity)."""
    out = {}
    for m, n in enumerate(FIELDS):
        scale = max(np.max(np.abs(b[m])), floor)
        out[n] = float(np.max(np.abs(a[m] - b[m])) / scale)
    return out


def assert_bitwise(a, b, what=""):
    """Equal as numbers (signed zeros compare equal), NaN-free."""
    assert np.isfinite(a).all() and np.isfinite(b).all(), what + ": non-finite values"
    bad = np.argwhere(a != b)
    if len(bad):
        m = bad[0]
        raise AssertionError(
            f"{what}: {len(bad)} cells differ; first at {tuple(m)}: {a[tuple(m)]!r} vs "
            f"{b[tuple(m)]!r}; rel linf {rel_linf(a, b)}")


def random_state(shape, seed, dtype=np.float64, amp=0.3):
    """A smooth-ish random MHD state: positive ro and ei, O(1) v and B."""
    rng = np.random.default_rng(seed)
    q = np.empty((9,) + tuple(shape))
    q[0] = 1.0 + amp * rng.random(shape)
    for m in (1, 2, 3):
        q[m] = amp * rng.standard_normal(shape)
    for m in (4, 5, 6):
        q[m] = amp * rng.standard_normal(shape)
    q[7] = 1.5 + amp * rng.random(shape)
    q[8] = 0.01 * rng.standard_normal(shape)
    return q.astype(dtype)
