"""Property tests of the oracle over random small inputs (hypothesis): each
oracle routine against an independent restatement written here from the
definition, so a slip in either shows up on some generated case.

* face accumulation (jfa.py:703-762) against the set definition
  (``faces_reference_sets``);
* the majority filter (jfa.py:640-699) against a direct 7-point tally;
* the hash-grid CSR (grid.py:189-309) against numpy's stable argsort and
  cumulative sum of the cells.
"""
import numpy as np
from hypothesis import given, settings
from hypothesis import strategies as st

from test_oracle_jfa import faces_reference_sets

F = np.float32
SETTINGS = settings(max_examples=60, deadline=None)


def random_labels(seed, res, n, holes):
    """Labels in [-1, n + 1] (some at or above n, which the path ignores),
    drawn from a few blobs so faces repeat, plus a fraction of -1 holes."""
    rng = np.random.default_rng(seed)
    k = int(rng.integers(1, max(2, n + 2)))
    lab = rng.integers(0, n + 2, k)[rng.integers(0, k, res ** 3)].astype(np.int32)
    lab[rng.random(res ** 3) < holes] = -1
    return lab


@SETTINGS
@given(seed=st.integers(0, 2 ** 31 - 1), res=st.integers(2, 7), n=st.integers(1, 12),
       holes=st.sampled_from([0.0, 0.1, 0.5]))
def test_faces_match_set_definition(orc, seed, res, n, holes):
    lab = random_labels(seed, res, n, holes)
    ids, cts = orc.accumulate_faces(lab, res, n)
    eids, ects = faces_reference_sets(lab, res, n)
    assert (ids == eids).all() and (cts == ects).all()


def majority_reference(g, res):
    """jfa.py:640-699 read directly: the centre and its six face neighbours
    (x-1, x+1, y-1, y+1, z-1, z+1, periodic) are tallied in that order with
    -1 skipped; the first label whose count is strictly the largest wins; an
    unassigned centre stays -1."""
    L = g.reshape(res, res, res)
    out = np.empty_like(L)
    for x in range(res):
        for y in range(res):
            for z in range(res):
                c = L[x, y, z]
                if c < 0:
                    out[x, y, z] = -1
                    continue
                seq = [c, L[(x - 1) % res, y, z], L[(x + 1) % res, y, z],
                       L[x, (y - 1) % res, z], L[x, (y + 1) % res, z],
                       L[x, y, (z - 1) % res], L[x, y, (z + 1) % res]]
                best, bc = c, 0
                for lab in seq:
                    if lab < 0:
                        continue
                    cnt = sum(1 for t in seq if t == lab)
                    if cnt > bc:
                        best, bc = lab, cnt
                out[x, y, z] = best
    return out.ravel()


@SETTINGS
@given(seed=st.integers(0, 2 ** 31 - 1), res=st.integers(2, 6), n=st.integers(1, 9),
       holes=st.sampled_from([0.0, 0.2]))
def test_majority_matches_tally(orc, seed, res, n, holes):
    lab = random_labels(seed, res, n, holes)
    assert (orc.majority(lab, res) == majority_reference(lab, res)).all()


@SETTINGS
@given(seed=st.integers(0, 2 ** 31 - 1), n=st.integers(1, 400), G=st.integers(3, 9),
       spread=st.sampled_from([0.5, 0.75, 2.0]))
def test_grid_csr_is_stable_counting_sort(orc, seed, n, G, spread):
    """Positions partly outside the box (spread > 0.5) must be wrapped first;
    the CSR is then the stable counting sort of the wrapped positions' cells."""
    rng = np.random.default_rng(seed)
    pos = ((rng.random((n, 3)) - 0.5) * 2 * spread).astype(F)
    p, cc, cs, ci = orc.grid_build(pos, n, G, 1.0)
    assert (np.abs(p) <= F(0.5)).all()
    # the cells of the wrapped positions, as grid.py:202-219 computes them
    q = ((p + F(0.5)) + F(1e-7)) * F(1.0)
    cell3 = np.floor(q * F(G)).astype(np.int64) % G
    cells = (cell3[:, 0] * G + cell3[:, 1]) * G + cell3[:, 2]
    assert (cc == np.bincount(cells, minlength=G ** 3)).all()
    assert (cs == np.concatenate([[0], np.cumsum(cc)[:-1]])).all()
    assert (ci == np.argsort(cells, kind="stable")).all()
