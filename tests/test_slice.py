"""SphereSimulation.slice and StepReport.submissions (gui/.../SphereSimulation.java:136-148, 201-248).

The reference submits each list its stepper runs (SphereStepper.step: one runAndRead a wait, then the closing run;
SphereStepper.java:66-103) cut into pieces of at most `slice` dispatches, and counts the pieces. What the reference
fixes whatever its lists' lengths is pinned here:

- no slice, or a slice at least as long as every list, is one submission a list: waits + 1 a step, and 1 for a step
  run as one list;
- a smaller slice never makes fewer submissions, and one dispatch a submission makes the most;
- the state is the same under any slice (PhysicsThreadTest.java:60 checks this and the count growing);
- a negative slice is refused with the reference's message, and a box world has none.

The count under a slice shorter than a list also takes the list's length, which includes CountingSort.scan's passes:
CountingSort is SupirVast's, absent from the reference, and its scan is taken as three passes. That count is parity
unpinned: it is only checked to agree between the oracle and the HIP library, in both of the library's launch models.
"""

import numpy as np
import pytest

import scene as sc
from boxscene import BoxScene

pkg = sc.pkg
STATE = ("x", "y", "z", "u", "v", "w", "ax", "ay", "az")
LONGEST = 1 << 20  # longer than any list a step runs


def pile_reports(backend, dispatches, steps=3, solve=None, friction=True):
    s = sc.pile()
    if solve is not None:
        s.solve = solve
    if friction:
        s.muStatic = s.muKinetic = 0.5
    s.start(backend)
    s.sim.slice(dispatches)
    reports = [s.sim.stepAndWait() for _ in range(steps)]
    state = {k: s.floats(k).copy() for k in STATE}
    s.close()
    return reports, state


@pytest.mark.parametrize("friction", [True, False])
def test_whole_lists_are_one_submission_a_wait_and_the_close(friction):
    for dispatches in (0, LONGEST):
        reports, _ = pile_reports("oracle", dispatches, friction=friction)
        for r in reports:
            assert r.solve.waits > 0
            assert r.submissions == r.solve.waits + 1, (dispatches, r)


def test_one_list_steps_are_one_submission():
    reports, _ = pile_reports("oracle", 0, solve=pkg.Solve.JACOBI)
    assert [r.submissions for r in reports] == [1, 1, 1]
    assert [r.solve.waits for r in reports] == [0, 0, 0]


def test_smaller_slices_make_more_submissions_and_the_same_state():
    counts = []
    base = None
    for dispatches in (0, 64, 16, 4, 1):
        reports, state = pile_reports("oracle", dispatches)
        counts.append(sum(r.submissions for r in reports))
        if base is None:
            base = state
        for k in STATE:
            assert np.array_equal(state[k].view(np.int32), base[k].view(np.int32)), (dispatches, k)
    assert counts == sorted(counts)
    assert counts[-1] > counts[0]


def refuses_a_negative_slice(backend):
    s = sc.pile().start(backend)
    with pytest.raises(pkg.IllegalArgumentException, match="a slice is some dispatches, or zero for no bound; got -1"):
        s.sim.slice(-1)
    s.close()


def test_a_negative_slice_is_refused():
    refuses_a_negative_slice("oracle")


@pytest.mark.gpu
def test_hip_refuses_a_negative_slice():
    refuses_a_negative_slice("gpu")


def test_a_box_world_has_no_slice_and_counts_its_two_lists():
    s = BoxScene(8)
    for k in range(8):
        s.x[k], s.y[k], s.z[k] = 0.1 + 0.1 * k, 0.5, 0.5
        s.half[:, k] = 0.04
    s.start("oracle")
    with pytest.raises(pkg.IllegalStateException):
        s.sim._world.slice(1)
    s.sim.stepAndWait()
    assert s.sim.submissions == 2  # BoxStepper.step: the collide list read back, then the step's list
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("host_waits", ["0", "1"])
def test_hip_submissions_match_the_oracle(host_waits, monkeypatch):
    monkeypatch.setenv("VR_HOST_WAITS", host_waits)
    for dispatches in (0, 1, 7, 40, LONGEST):
        want, want_state = pile_reports("oracle", dispatches)
        got, got_state = pile_reports("gpu", dispatches)
        assert [(r.submissions, r.solve.waits) for r in got] == [(r.submissions, r.solve.waits) for r in want], \
            dispatches
        for k in STATE:
            assert np.array_equal(got_state[k].view(np.int32), want_state[k].view(np.int32)), (dispatches, k)


def test_a_gridless_jacobi_list_is_cut_as_the_reference_cuts_it():
    """Without a grid a Jacobi step is one list the reference fixes whole (SphereStep.java:146-208): per substep
    predict, then solve and apply each iteration, then velocity, bounce and bounced; then show. So its submissions
    under any slice are pinned: ceil(length / slice), one with no slice."""
    s = sc.crowd()
    assert s.solve == pkg.Solve.JACOBI and not s.grid
    length = s.substeps * (1 + 2 * s.iterations + 3) + 1
    for dispatches, want in ((0, 1), (1, length), (7, -(-length // 7)), (length, 1), (length + 1, 1)):
        s.start("oracle")
        s.sim.slice(dispatches)
        assert s.sim.stepAndWait().submissions == want, dispatches
        s.close()
