"""The HIP path against the oracle on the same seeded scenes, through the C ABI.

Integer results (contact lists, counts, the rounds each pass took) must match exactly. The f32 state is compared
bit for bit as well: both sides do the same operations in the same order without fused multiply-adds, so they
agree to the bit; the stated contract, 1e-4 relative, is checked too, and is what a failure reports against.
"""

import numpy as np
import pytest

import scene as sc
from scene import Scene

pkg = sc.pkg
GS = pkg.Solve.GAUSS_SEIDEL_BY_CONTACT
pytestmark = pytest.mark.gpu

STATE = ("x", "y", "z", "u", "v", "w", "ax", "ay", "az", "qx", "qy", "qz", "qw", "dx", "dy", "dz")


def run_both(make, steps, extra=()):
    out = {}
    for backend in ("oracle", "gpu"):
        s = make()
        s.start(backend)
        reports = []
        for _ in range(steps):
            rep = s.sim.stepAndWait()
            reports.append((rep.solve.waits, rep.solve.rounds, rep.solve.mostRounds, rep.solve.unlisted))
        o = {k: s.floats(k).copy() for k in STATE}
        o["reports"] = reports
        for name in extra:
            info = s.sim._world.info(name)
            o[name] = s.sim._world.read(name, np.float32 if info[1] == 1 else np.int32).copy()
        s.close()
        out[backend] = o
    return out["oracle"], out["gpu"]


def assert_close(a, b, what):
    for k in STATE:
        x, y = a[k], b[k]
        scale = np.maximum(np.abs(x), 1e-3)
        rel = float(np.max(np.abs(x - y) / scale)) if x.size else 0.0
        assert rel <= 1e-4, f"{what}: {k} differs by {rel:.3g} relative"


def assert_bits(a, b, what):
    for k in STATE:
        diff = int(np.count_nonzero(a[k].view(np.int32) != b[k].view(np.int32)))
        assert diff == 0, f"{what}: {k} differs in {diff} of {a[k].size} words"


def friction_pile():
    s = sc.pile()
    s.muStatic = s.muKinetic = 0.5
    s.muRolling = 0.05
    s.restitution = 0.3
    return s


def test_pile_with_friction_matches_oracle_bitwise():
    a, b = run_both(friction_pile, 6, extra=("contactA", "contactB", "contactCount", "contactImpulse"))
    assert a["reports"] == b["reports"]
    listed = int(a["contactCount"][0])
    assert listed == int(b["contactCount"][0]) and listed > 0
    assert np.array_equal(a["contactA"][:listed], b["contactA"][:listed])
    assert np.array_equal(a["contactB"][:listed], b["contactB"][:listed])
    assert np.array_equal(a["contactCount"], b["contactCount"])
    assert_close(a, b, "friction pile")
    assert_bits(a, b, "friction pile")
    imp_a = a["contactImpulse"][:7 * listed]
    imp_b = b["contactImpulse"][:7 * listed]
    assert np.array_equal(imp_a.view(np.int32), imp_b.view(np.int32))


@pytest.mark.parametrize("plan", ["wide", "local"])
def test_both_pass_plans_match_the_oracle_bitwise(plan, monkeypatch):
    """A Gauss-Seidel pass runs as chip-wide round launches finished by one workgroup ("wide"), or as wide rounds
    while more than 1,024 contacts are open and then one workgroup with the claim arrays in LDS ("local");
    VR_SPHERE_PLAN fixes the plan. Both run the reference's rounds: each matches the oracle to the bit, reports
    included."""
    monkeypatch.setenv("VR_SPHERE_PLAN", plan)
    a, b = run_both(friction_pile, 4, extra=("contactCount", "contactImpulse"))
    assert a["reports"] == b["reports"]
    assert np.array_equal(a["contactCount"], b["contactCount"])
    assert_bits(a, b, f"friction pile, {plan} plan")
    listed = int(a["contactCount"][0])
    assert np.array_equal(a["contactImpulse"][:7 * listed].view(np.int32),
                          b["contactImpulse"][:7 * listed].view(np.int32))


def crowded_pile():
    """900 spheres in the pile's box: more listed contacts a pass than the local plan's 1,024 threads take at once."""
    s = sc.pile(n=900)
    s.muStatic = s.muKinetic = 0.5
    s.muRolling = 0.05
    return s


def test_the_local_plan_over_rounds_longer_than_its_workgroup_matches_the_oracle(monkeypatch):
    monkeypatch.setenv("VR_SPHERE_PLAN", "local")
    a, b = run_both(crowded_pile, 3, extra=("contactCount",))
    assert int(a["contactCount"][2]) > 1024, "the rounds must take the workgroup more than once"  # LONGEST
    assert a["reports"] == b["reports"]
    assert_bits(a, b, "crowded pile, local plan")


def test_the_local_plans_wide_rounds_changing_between_steps_match_the_oracle(monkeypatch):
    """Unfixed, a world of at most 6,144 spheres takes the local plan, whose wide rounds a pass are those the step
    before needed to bring the pass to 1,024 open contacts: on the crowded pile (up to 5,261 contacts listed a
    substep in its first step) they change from step to step, and the state stays bit-exact."""
    monkeypatch.delenv("VR_SPHERE_PLAN", raising=False)
    a, b = run_both(crowded_pile, 3, extra=("contactCount",))
    assert a["reports"] == b["reports"]
    assert_bits(a, b, "crowded pile, chosen plan")


def test_frictionless_pile_fixed_rounds_matches_oracle():
    def make():
        s = sc.pile()
        s.fixedRounds = True
        s.rounds = 20
        return s
    a, b = run_both(make, 5, extra=("contactCount",))
    assert np.array_equal(a["contactCount"], b["contactCount"])
    assert_bits(a, b, "fixed rounds")


def test_jacobi_grid_crowd_matches_oracle():
    def make():
        s = sc.crowd()
        s.grid = True
        s.restitution = 0.5
        return s
    a, b = run_both(make, 3)
    assert_bits(a, b, "jacobi grid")


def test_jacobi_every_pair_matches_oracle():
    def make():
        s = sc.crowd()
        s.averaged = False
        s.omega = 0.5
        return s
    a, b = run_both(make, 2)
    assert_bits(a, b, "jacobi pairs")


def test_wide_priorities_match_oracle():
    def make():
        s = friction_pile()
        s.widePairs = True
        return s
    a, b = run_both(make, 4, extra=("contactCount",))
    assert a["reports"] == b["reports"]
    assert_bits(a, b, "wide priorities")


def test_list_overflow_is_counted_the_same():
    def make():
        s = sc.pile()
        s.contactsPerSphere = 1
        return s
    a, b = run_both(make, 3, extra=("contactCount",))
    assert a["reports"] == b["reports"]
    assert any(r[3] > 0 for r in a["reports"]), "the list should overflow"
    assert np.array_equal(a["contactCount"], b["contactCount"])
    assert_bits(a, b, "overflow")


def test_reports_and_state_over_many_steps_match_oracle():
    """40 steps of the friction pile: the per-step reports (waits, rounds, most rounds) follow the reference's stepper,
    its opening batch growing and shrinking (SphereStepper.java:118-126), and the state stays bit-identical."""
    a, b = run_both(friction_pile, 40, extra=("contactCount",))
    assert a["reports"] == b["reports"]
    assert_bits(a, b, "40 steps")
