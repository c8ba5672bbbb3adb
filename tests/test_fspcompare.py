"""The vectorised sparse ℓ1 distance of the GPU tests against a per-state dictionary walk."""
import numpy as np

from fspcompare import l1_distance


def _dict_l1(sa, va, sb, vb):
    da = {tuple(x): v for x, v in zip(sa.tolist(), va)}
    db = {tuple(x): v for x, v in zip(sb.tolist(), vb)}
    return sum(abs(da.get(x, 0.0) - db.get(x, 0.0)) for x in set(da) | set(db))


def test_l1_distance_matches_dictionary_walk():
    rng = np.random.default_rng(3)
    for ns, hi in [(1, 50), (2, 12), (3, 9), (4, 300)]:
        pool = np.unique(rng.integers(0, hi, size=(400, ns)), axis=0)
        for _ in range(5):
            ia = rng.choice(len(pool), size=rng.integers(0, len(pool)), replace=False)
            ib = rng.choice(len(pool), size=rng.integers(0, len(pool)), replace=False)
            sa, sb = pool[ia], pool[ib]
            va, vb = rng.random(len(ia)), rng.random(len(ib))
            assert np.isclose(l1_distance(sa, va, sb, vb), _dict_l1(sa, va, sb, vb), rtol=1e-13, atol=0)


def test_l1_distance_special_cases():
    s = np.array([[0, 1], [2, 0], [1, 1]], dtype=np.int32)
    v = np.array([0.5, 0.25, 0.25])
    assert l1_distance(s, v, s[::-1], v[::-1]) == 0.0  # same vector, other order
    assert l1_distance(s, v, s[:0], v[:0]) == 1.0  # against the empty vector
    assert l1_distance(s[:0], v[:0], s[:0], v[:0]) == 0.0
    assert l1_distance(s[:1], v[:1], s[1:], v[1:]) == 1.0  # disjoint supports
    # a coordinate that only the larger base separates: (0, 10) vs (1, 0)
    assert l1_distance([[0, 10]], [1.0], [[1, 0]], [1.0]) == 2.0
