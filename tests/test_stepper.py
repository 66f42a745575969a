"""The run-until-done stepper's schedule on the oracle: the rules SphereStepperTest checks against a runner that says
how many rounds each pass needs (ctest/sphere/SphereStepperTest.java:60-115), here on real scenes whose passes need
nothing (spheres apart) or many rounds (spheres crowded together).

SphereStepper.step (SphereStepper.java:66-103) opens each pass with its opening batch, then runs batches of BATCH = 8
until the device reports nothing open, one wait per batch. SphereStepper.adapt (:118-126) grows a pass's opening
to the rounds it ran (at most CYCLE = 48), and shrinks it by a batch after SHRINK_AFTER = 120 passes in a row that ran
no more than it. The HIP library's device-decided step settles the same counts on the device; its step reports are
checked against the oracle's over 40 steps in test_gpu_parity.py.
"""

import ctypes

import numpy as np

import scene as sc
from scene import Scene

pkg = sc.pkg
BATCH, CYCLE, SHRINK_AFTER = 8, 48, 120


def opening(s, iteration):
    """The oracle stepper's opening batch for a pass (SphereStepper.opening, :115-117)."""
    fn = sc.oracle_abi().lib.vro_world_opening
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int32]
    fn.restype = ctypes.c_int
    return fn(s.sim._world.handle, iteration)


def spheres(substeps, iterations, crowded):
    """64 spheres of 3 cm with no gravity in a 1 m box, as the reference test's step has 64 (SphereStepperTest.java:
    56-59): crowded into 0.2 m a side, or on a 4 x 4 x 4 lattice 20 cm apart, which touch nothing."""
    s = Scene(64).uniform(0.03, 0.1)
    s.gy = 0
    s.solve = pkg.Solve.GAUSS_SEIDEL_BY_CONTACT
    s.substeps, s.iterations = substeps, iterations
    if crowded:
        rnd = pkg.JavaRandom(17)
        for k in range(s.n):
            s.x[k], s.y[k], s.z[k] = (np.float32(0.4 + 0.2 * rnd.nextDouble()) for _ in range(3))
    else:
        s.x[:], s.y[:], s.z[:] = lattice()
    return s


def lattice():
    g = (0.2 + 0.2 * np.arange(4)).astype(np.float32)
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    return x.ravel(), y.ravel(), z.ravel()


def test_a_pass_that_needs_no_more_than_its_opening_is_one_wait():
    """SphereStepperTest.aPassThatNeedsNoMoreThanItsOpeningIsOneWait (:63-73): three substeps whose passes need no
    round past the opening are three waits and 3 x BATCH rounds, then the closing: four lists."""
    with spheres(3, 1, crowded=False).start("oracle") as s:
        for _ in range(3):
            r = s.sim.stepAndWait()
            assert (r.solve.waits, r.solve.rounds, r.solve.mostRounds, r.solve.unlisted) == (3, 3 * BATCH, BATCH, 0)
            assert r.submissions == 4


def test_a_pass_goes_on_a_batch_at_a_time_until_nothing_is_open():
    """SphereStepperTest.aPassGoesOnABatchAtATimeUntilNothingIsOpen (:75-82): one pass that needs more than its
    opening runs it, then batches of 8, one wait each: waits = 1 + (rounds - BATCH) / BATCH."""
    with spheres(1, 1, crowded=True).start("oracle") as s:
        r = s.sim.stepAndWait().solve
        assert r.rounds > 2 * BATCH, "the crowd should need several batches"
        assert r.rounds % BATCH == 0 and r.mostRounds == r.rounds
        assert r.waits == 1 + (r.rounds - BATCH) // BATCH


def listed_pairs(s):
    """Every pair the list takes at the given positions, by brute force in f32: nearer than the radii's sum times
    (1 + LIST_MARGIN) (Spheres.java:795 and the list pass)."""
    m = np.float32(1) + np.float32(pkg.Spheres.LIST_MARGIN)
    count = 0
    for a in range(s.n):
        d = [(c[a] - c[a + 1:]).astype(np.float32) for c in (s.x, s.y, s.z)]
        d2 = d[0] * d[0] + (d[1] * d[1] + d[2] * d[2])
        near = (s.r[a] + s.r[a + 1:]) * m
        count += int(np.count_nonzero(d2 < near * near))
    return count


def test_contacts_the_list_had_no_room_for_are_reported():
    """SphereStepperTest.contactsTheListHadNoRoomForAreReported (:154-160): each pass adds what its list found past
    the list's capacity, so a substep of two passes reports it twice. The crowd is still and has no gravity, so the
    first substep lists the pairs of its starting positions, counted here by brute force."""
    s = spheres(1, 2, crowded=True)
    s.contactsPerSphere = 1
    listed = listed_pairs(s)
    assert listed > s.n, "the list should overflow"
    with s.start("oracle"):
        r = s.sim.stepAndWait().solve
        assert int(s.ints("contactCount")[pkg.Spheres.LISTED]) == listed
        assert r.unlisted == 2 * (listed - s.n)


def test_the_opening_grows_to_what_a_pass_needed_and_shrinks_slowly():
    """SphereStepperTest.theOpeningGrowsToWhatAPassNeededAndShrinksSlowly (:98-113): one substep of two passes. A
    crowded first step grows each pass's opening to the rounds it ran (at most CYCLE). Moved apart, the spheres need
    nothing: the opening still runs whole, one wait a pass, and holds for SHRINK_AFTER - 1 more steps; the next one
    shrinks it by a batch."""
    with spheres(1, 2, crowded=True).start("oracle") as s:
        assert (opening(s, 0), opening(s, 1)) == (BATCH, BATCH)
        r = s.sim.stepAndWait().solve
        ran = sorted((r.mostRounds, r.rounds - r.mostRounds))
        assert ran[0] > BATCH, "both passes should run past their opening"
        grown = (opening(s, 0), opening(s, 1))
        assert sorted(grown) == [min(k, CYCLE) for k in ran]

        x, y, z = lattice()
        for name, values in (("x", x), ("y", y), ("z", z)):
            s.sim.write(name, values)
        for name in ("u", "v", "w", "ax", "ay", "az"):
            s.sim.write(name, np.zeros(s.n, dtype=np.float32))
        for _ in range(SHRINK_AFTER - 1):
            r = s.sim.stepAndWait().solve
            assert (r.waits, r.rounds) == (2, sum(grown))
        assert (opening(s, 0), opening(s, 1)) == grown, "not before a long run of passes that needed less"
        s.sim.stepAndWait()
        assert (opening(s, 0), opening(s, 1)) == tuple(k - BATCH for k in grown), "then a batch smaller"
