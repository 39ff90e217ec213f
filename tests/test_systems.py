"""The System mirror's schedule (genesis_amd/systems.py) against the
reference's build_schedule order (src/universe/mod.rs:355-388)."""
from genesis_amd import systems as S
from genesis_amd.params import Physics


def test_schedule_order_follows_build_schedule():
    assert S.build_schedule(Physics()).names() == [
        "velocity_half_kick", "position_drift", "boundary", "forces", "velocity_half_kick"]
    ph = Physics(enable_boundaries=False, enable_thermostat=True, enable_bond_observation=True)
    assert S.build_schedule(ph).names() == [
        "velocity_half_kick", "position_drift", "forces", "velocity_half_kick", "thermostat", "bond_observation"]


class _Probe:
    """Stands in for the resident universe: records the calls in order."""

    def __init__(self):
        self.calls = []

    def half_kick(self):
        self.calls.append("kick")

    def drift(self):
        self.calls.append("drift")

    def wrap(self):
        self.calls.append("wrap")

    def forces(self):
        self.calls.append("forces")
        return -1.5, 0.25

    def thermostat(self):
        self.calls.append("thermostat")
        return 0.99

    def bond_observe(self, tick):
        self.calls.append(f"observe@{tick}")


def test_tick_advances_the_clock_before_the_schedule():
    ph = Physics(enable_thermostat=True, enable_bond_observation=True)
    probe = _Probe()
    ctx = S.SystemContext(probe)
    sched = S.build_schedule(ph)
    S.tick(sched, ctx)
    S.tick(sched, ctx)
    one = ["kick", "drift", "wrap", "forces", "kick", "thermostat"]
    assert probe.calls == one + ["observe@1"] + one + ["observe@2"]
    assert ctx.resources == {"PotentialEnergy": -1.5, "BondEnergy": 0.25, "lambda": 0.99}
