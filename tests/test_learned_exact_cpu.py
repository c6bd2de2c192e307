"""The exact correction of the learned PTT proposal, pinned on the CPU
oracle by what the mathematics fixes.

The learned sampler draws a face from the network's estimate g(f) and
accepts it with probability F(f) / (M g(f)), F the reference's face value
(sum of the 3 exact vertex supports, 0 if one is below 0.4).  Whenever
F <= M g on every face, an accepted candidate is distributed as the
reference's: face with probability F(f) / sum F, uniform in the triangle
(ptt.cu:270-344), then the support test.

A network with zero weights and output bias c predicts the same support c
for every vertex, so g is the same on every face and the proposal is
uniform over faces:
- c = 9 (the largest support of an ODF normalised to 1 per voxel): F <= 27
  <= M g, the envelope holds, no draw is counted as a miss, and the winners'
  faces follow F(f) / sum F (chi-square), not the uniform proposal (the
  test's power), with points uniform in their triangle;
- c = 0.1: F exceeds M g on the supported faces, and those draws are
  counted.
On the test ODF every candidate's support exceeds 0.4, so the winner of a
round is its first accepted candidate and the face law is exactly F / sum F.
"""

import json
import os

import numpy as np
from scipy import stats

from oracle.oracle import Oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _disc():
    with open(os.path.join(GOLD, "reference_tables.json")) as f:
        t = json.load(f)
    return (np.asarray(t["disc2_vert"], np.float64).reshape(24, 2),
            np.asarray(t["disc2_face"], np.int64).reshape(31, 3))


def _constant_net(c):
    from gpustreamlines_amd import learned as LP

    p = np.zeros(LP.NPARAM, np.float32)
    p[LP.B3:LP.B3 + LP.N_OUT] = c
    return LP.to_bf16(p)


def _lobed(w):
    """The same ODF in every voxel, 0.06 + 0.9 x^2 over the sphere (x the
    vertex's first coordinate), normalised to a maximum of 1: every probe tap
    is >= 0.06, so every candidate's support is >= 0.54, and curves that bend
    towards x gather more, so face values differ several-fold."""
    v = w.vertices.astype(np.float32)
    odf = 0.06 + 0.9 * v[:, 0] ** 2
    odf = odf / odf.max()
    return np.broadcast_to(odf[None, None, None, :], w.dataf.shape).astype(np.float32).copy()


def _setup(w, c):
    stop = np.full(w.stop_map.shape, 0.8, np.float32)
    o = Oracle(w.params, _lobed(w), stop, w.vertices, w.edges, ptt_lut=w.ptt_lut)
    o.set_learned(_constant_net(c))
    return o


def _faces_of(ks, mc):
    verts, faces = _disc()
    tri = verts[faces] * mc
    a, b, c = tri[:, 0], tri[:, 1], tri[:, 2]
    minv = np.linalg.inv(np.stack([b - a, c - a], axis=2))
    rs = np.einsum("fij,nfj->nfi", minv, ks[:, None, :] - a[None])
    inside = (rs[..., 0] >= -1e-5) & (rs[..., 1] >= -1e-5) & (rs.sum(-1) <= 1 + 1e-5)
    assert inside.any(axis=1).all(), "a winner lies outside every disc face"
    face = np.argmax(inside, axis=1)
    return face, rs[np.arange(len(face)), face]


def test_learned_winners_follow_the_reference_face_law(small_workloads):
    w = small_workloads["ptt"]
    shape = w.dataf.shape
    o = _setup(w, 9.0)
    mc = float(w.params["max_curvature"])
    _, faces = _disc()
    pos = np.array([shape[0] / 2, shape[1] / 2, shape[2] / 2], np.float32)
    frame = np.array([0, 0, 1, 1, 0, 0, 0, 1, 0], np.float32)
    vp = o.vertex_supports(pos, frame).astype(np.float64)
    vp = np.where(vp < 0.4, 0.0, vp)
    fv = vp[faces].sum(axis=1) * (vp[faces] != 0).all(axis=1)
    pf = fv / fv.sum()
    assert pf.max() > 4 * pf.min()
    ks, misses = [], 0
    for call in range(4000):
        st, k, _, cnt = o.ptt_learned_substep(pos, frame, sub0=32 * call + 1)
        assert st == 1, "a learned round should win on this ODF"
        ks.append(k.astype(np.float64))
        misses += cnt[3]
    assert misses == 0
    face, r = _faces_of(np.array(ks), mc)
    counts = np.bincount(face, minlength=31)
    chi = stats.chisquare(counts, pf * counts.sum())
    assert chi.pvalue > 1e-4, (chi, counts, np.round(pf * counts.sum()))
    # power: the uncorrected (uniform) proposal is rejected
    assert stats.chisquare(counts, np.full(31, counts.sum() / 31)).pvalue < 1e-8
    for j in range(2):
        assert stats.kstest(np.clip(r[:, j], 0, 1), lambda x: 1 - (1 - x) ** 2).pvalue > 1e-4


def test_learned_envelope_misses_are_counted(small_workloads):
    w = small_workloads["ptt"]
    shape = w.dataf.shape
    o = _setup(w, 0.1)
    pos = np.array([shape[0] / 2, shape[1] / 2, shape[2] / 2], np.float32)
    frame = np.array([0, 0, 1, 1, 0, 0, 0, 1, 0], np.float32)
    tot = 0
    for call in range(50):
        st, _, _, cnt = o.ptt_learned_substep(pos, frame, sub0=32 * call + 1)
        assert st in (1, 2)
        tot += cnt[3]
    assert tot > 0
