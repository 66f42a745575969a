"""BoxPairs.carry pinned by its closed-form rules, on each backend, through the ABI.

core/.../box/BoxPairs.java:184-267 (core/ = vexelray-sim-rigid-core/src/main/java/dev/vexelray/sim/rigid/). After the
collide, each slot's pair finds the slot its first body held the same partner in the step before (any of its 32,
the last that matches); each new point takes the seven impulse words of the old point whose anchor on the first
body is nearest, if it lies within SAME_POINT = 0.05 of that body's smallest half-extent (strictly nearer than any
before it); one point then and one now always carry; a pair with no kept slot, or a point with no old point in
reach, starts from zeros.

Each case runs the collide list alone (vr_box_collide_list, BoxStepper.java:62-95), so no substep changes what the
carry wrote: the kept buffers are written by name (pairs, manifold, impulses become pairsWas, manifoldWas,
impulsesWas at the keep), and the impulses are read straight after.
"""

import numpy as np
import pytest

import scene as sc  # noqa: F401  (sets up the package)
from boxscene import BoxContact, BoxScene

BACKENDS = ["oracle", pytest.param("gpu", marks=pytest.mark.gpu)]
WORDS = 28  # BoxPairs.IMPULSE_WORDS: 4 points x (normal, tangent 3, rolling 3)
M = BoxContact.MANIFOLD_WORDS


def anchor_at(k):
    """Where point k's anchor on the first body sits in a manifold."""
    return BoxContact.POINTS + k * BoxContact.POINT_WORDS + BoxContact.ANCHOR_FIRST


def pair(backend, ball_on_top=False, half=0.1):
    """A cube of half `half` at rest on the floor with a cube (or a ball of radius `half`) pressed 1 cm into its top:
    one pair, body 0 first, in body 0's slot 0."""
    s = BoxScene(2).uniform(half, half, half, 1.0)
    if ball_on_top:
        s.sphere(1, half)
    s.x[:] = 0.5
    s.z[:] = 0.5
    s.y[0] = half
    s.y[1] = 3 * half - 0.01
    s.start(backend)
    return s


def collide_list(s):
    return s.sim.collideList()


def marked(seed=1):
    """28 distinct impulse words, none zero."""
    return (np.arange(1, WORDS + 1, dtype=np.float32) * np.float32(0.25) + np.float32(seed)).astype(np.float32)


def slot_words(s, slot=0):
    return s.floats("impulses")[slot * WORDS:(slot + 1) * WORDS].copy()


def write_slot(s, name, slot, values, width):
    buf = s.floats(name) if name != "pairs" else s.ints(name)
    buf = buf.copy()
    buf[slot * width:(slot + 1) * width] = values
    s.sim.write(name, buf)


@pytest.mark.parametrize("backend", BACKENDS)
def test_points_that_stay_put_carry_their_words(backend):
    with pair(backend) as s:
        rep = collide_list(s)
        assert (rep.touching, rep.points) == (1, 4)
        assert not slot_words(s).any(), "carried from nothing: zeros"
        w = marked()
        write_slot(s, "impulses", 0, w, WORDS)
        collide_list(s)
        assert np.array_equal(slot_words(s).view(np.int32), w.view(np.int32))


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("share, carried", [(0.04, True), (0.06, False)])
def test_a_point_carries_only_within_a_twentieth_of_the_smallest_half(backend, share, carried):
    """Point 2's old anchor moved along x by `share` of the first box's smallest half-extent (0.1 m): within 0.05 it
    still takes its words, beyond it starts from zeros; the other points carry as before."""
    with pair(backend) as s:
        collide_list(s)
        m = s.floats("manifold")[:M].copy()
        m[anchor_at(2)] = np.float32(m[anchor_at(2)] + np.float32(share * 0.1))
        write_slot(s, "manifold", 0, m, M)
        w = marked()
        write_slot(s, "impulses", 0, w, WORDS)
        collide_list(s)
        got = slot_words(s).reshape(4, 7)
        want = w.reshape(4, 7).copy()
        if not carried:
            want[2] = 0
        assert np.array_equal(got.view(np.int32), want.view(np.int32))


@pytest.mark.parametrize("backend", BACKENDS)
def test_the_nearest_old_point_gives_its_words(backend):
    """Two old points within reach of new point 0: the nearer one's words are taken, whichever comes first."""
    with pair(backend) as s:
        collide_list(s)
        m = s.floats("manifold")[:M].copy()
        new0 = m[anchor_at(0):anchor_at(0) + 3].copy()
        m[anchor_at(0):anchor_at(0) + 3] = new0 + np.array([0.003, 0, 0], np.float32)  # within reach (0.005), farther
        m[anchor_at(3):anchor_at(3) + 3] = new0 + np.array([0.001, 0, 0], np.float32)  # within reach, nearer
        write_slot(s, "manifold", 0, m, M)
        w = marked()
        write_slot(s, "impulses", 0, w, WORDS)
        collide_list(s)
        got = slot_words(s).reshape(4, 7)
        assert np.array_equal(got[0].view(np.int32), w.reshape(4, 7)[3].view(np.int32))
        # new point 3 has lost its old anchor to point 0's neighbourhood: none within reach, zeros
        assert not got[3].any()


@pytest.mark.parametrize("backend", BACKENDS)
def test_of_two_old_points_at_the_same_distance_the_first_gives_its_words(backend):
    """Old points 0 and 3 both at one anchor, within reach of new point 0: only an old point strictly nearer than
    those before it is taken, so old point 0's words carry, not point 3's."""
    with pair(backend) as s:
        collide_list(s)
        m = s.floats("manifold")[:M].copy()
        moved = m[anchor_at(0):anchor_at(0) + 3] + np.array([0.002, 0, 0], np.float32)
        m[anchor_at(0):anchor_at(0) + 3] = moved
        m[anchor_at(3):anchor_at(3) + 3] = moved
        write_slot(s, "manifold", 0, m, M)
        w = marked()
        write_slot(s, "impulses", 0, w, WORDS)
        collide_list(s)
        got = slot_words(s).reshape(4, 7)
        assert np.array_equal(got[0].view(np.int32), w.reshape(4, 7)[0].view(np.int32))
        assert not got[3].any()


@pytest.mark.parametrize("backend", BACKENDS)
def test_one_point_then_and_now_always_carries(backend):
    """A ball on a box: one point. Its old anchor moved far off (a rolling ball's would) still carries; with two old
    points, the nearest rule applies instead."""
    with pair(backend, ball_on_top=True) as s:
        rep = collide_list(s)
        assert (rep.touching, rep.points) == (1, 1)
        m = s.floats("manifold")[:M].copy()
        far = m.copy()
        far[anchor_at(0)] = np.float32(far[anchor_at(0)] + np.float32(0.05))  # ten times the reach
        write_slot(s, "manifold", 0, far, M)
        w = marked()
        write_slot(s, "impulses", 0, w, WORDS)
        collide_list(s)
        got = slot_words(s).reshape(4, 7)
        assert np.array_equal(got[0].view(np.int32), w.reshape(4, 7)[0].view(np.int32))
        assert not got[1:].any()

        two = far.copy()
        two[BoxContact.COUNT] = np.float32(2)
        two[anchor_at(1):anchor_at(1) + 3] = m[anchor_at(0):anchor_at(0) + 3]  # old point 1 where the new one is
        write_slot(s, "manifold", 0, two, M)
        write_slot(s, "impulses", 0, w, WORDS)
        collide_list(s)
        got = slot_words(s).reshape(4, 7)
        assert np.array_equal(got[0].view(np.int32), w.reshape(4, 7)[1].view(np.int32))

        far[anchor_at(0)] = np.float32(m[anchor_at(0)] + np.float32(0.05))
        two[anchor_at(1)] = far[anchor_at(0)]  # both old points out of reach: nothing carries
        write_slot(s, "manifold", 0, two, M)
        write_slot(s, "impulses", 0, w, WORDS)
        collide_list(s)
        assert not slot_words(s).any()


@pytest.mark.parametrize("backend", BACKENDS)
def test_a_pair_with_no_kept_slot_starts_from_zeros(backend):
    with pair(backend) as s:
        collide_list(s)
        w = marked()
        write_slot(s, "impulses", 0, w, WORDS)
        write_slot(s, "pairs", 0, np.array([0, -1], dtype=np.int32), 2)  # the partner not kept in any slot
        collide_list(s)
        assert not slot_words(s).any()


@pytest.mark.parametrize("backend", BACKENDS)
def test_the_kept_slot_is_found_among_all_32(backend):
    """The step before held the partner in slot 5 of the first body: its manifold and words there carry into slot 0."""
    with pair(backend) as s:
        collide_list(s)
        m = s.floats("manifold")[:M].copy()
        w = marked()
        write_slot(s, "pairs", 0, np.array([0, -1], dtype=np.int32), 2)
        write_slot(s, "pairs", 5, np.array([0, 1], dtype=np.int32), 2)
        write_slot(s, "manifold", 5, m, M)
        write_slot(s, "impulses", 5, w, WORDS)
        write_slot(s, "impulses", 0, marked(7), WORDS)
        collide_list(s)
        assert np.array_equal(slot_words(s, 0).view(np.int32), w.view(np.int32))
