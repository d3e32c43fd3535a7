"""Batched stepping: every member of a batch ends exactly where the same
number of steps on it alone leaves it (positions, counts, omega, accepts),
whatever the lane shape, with members of mixed size, architecture, cadence
and gridding path (atomic and tiled with reordering between graph replays)."""
import numpy as np
import pytest

from new_arch_amd import scmf
from new_arch_amd import workloads as W
from new_arch_amd.abi import make_opts

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu-batch] {msg}", flush=True)


def mix():
    systems = W.batched_mix(n_systems=7, seed=3, sizes=(16, 24, 32), lengths=(16, 24, 32))
    # one member large enough for the tiled gridding, reordered every 3 sweeps
    big = W.small_diblock_melt()
    systems.append((big, 77, 99))
    return systems


def make(systems, lanes=0, sort_every=3):
    out = []
    for k, (spec, seed, sid) in enumerate(systems):
        out.append(scmf.System(spec, make_opts(seed=seed, system_id=sid,
                                               omega_every_step=(k % 2 == 0),
                                               density_bits=16 if k % 3 else 32,
                                               sort_every=sort_every, lanes_per_chain=lanes)))
    for g in out:
        g.grid_density()
        g.update_omega()
    return out


def same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint32),
                          np.ascontiguousarray(b).view(np.uint32))


@pytest.mark.parametrize("lanes", [0, 1, 4, 16])
def test_batch_equals_members_alone(lanes):
    banner(f"batch vs alone lanes={lanes}")
    systems = mix()
    alone = make(systems, lanes)
    mem = make(systems, lanes)
    steps = 12
    for g in alone:
        for k in range(1, steps + 1):
            g.step(k)
    b = scmf.Batch(mem)
    if lanes:
        b.set_lanes(lanes)
    b.step(steps)
    b.sync()
    for i, (a, m) in enumerate(zip(alone, mem)):
        assert same(a.positions(), m.positions()), f"member {i} positions"
        np.testing.assert_array_equal(a.density(), m.density(), err_msg=f"member {i} density")
        assert same(a.omega(), m.omega()), f"member {i} omega"
        sa, sm = a.stats(), m.stats()
        assert sa.accepts == sm.accepts and sa.sweeps == sm.sweeps == steps
    b.close()
    # members stay usable on their own after the batch
    mem[0].step(steps + 1)
    alone[0].step(steps + 1)
    assert same(alone[0].positions(), mem[0].positions())
    for g in alone + mem:
        g.close()


def test_batch_rejects_duplicates():
    banner("batch duplicates")
    spec, seed, sid = W.batched_mix(n_systems=1, seed=1)[0]
    g = scmf.System(spec, make_opts(seed=seed))
    with pytest.raises(scmf.ScmfError):
        scmf.Batch([g, g])
    g.close()


def test_stops_and_membership_changes_while_steps_are_queued():
    """Members stop on the device at their own sweep counts (the batch keeps
    stepping past them), and membership changes take effect for the steps
    queued after them without waiting for those queued before: every member
    ends where its own number of scmf_step calls alone leaves it."""
    banner("stops and membership changes")
    systems = W.batched_mix(n_systems=6, seed=9, sizes=(16, 24), lengths=(16, 24, 32))
    mem = make(systems)
    stops = [5, 9, 14, 3, 11, 20]
    b = scmf.Batch(mem[:4])
    b.update(mem[:4], stops[:4])
    b.step(7)                                 # members 0 and 3 stop inside these steps
    b.update([mem[1], mem[2], mem[4]], [stops[1], stops[2], stops[4]])   # 0, 3 leave; 4 joins
    b.step(6)
    b.update([mem[2], mem[4], mem[5]], [stops[2], stops[4], stops[5]])   # 5 joins
    b.step(25)
    b.sync()
    alone = make(systems)
    for i, g in enumerate(alone):
        for k in range(1, stops[i] + 1):
            g.step(k)
    for i, (a, m) in enumerate(zip(alone, mem)):
        assert m.stats().sweeps == stops[i], (i, m.stats().sweeps)
        assert same(a.positions(), m.positions()), f"member {i} positions"
        np.testing.assert_array_equal(a.density(), m.density(), err_msg=f"member {i} density")
        assert same(a.omega(), m.omega()), f"member {i} omega"
        assert a.stats().accepts == m.stats().accepts
    b.close()
    for g in alone + mem:
        g.close()


def test_membership_changes_reuse_the_captured_graph():
    """A membership that fits the captured graph's room is installed by a
    table upload alone: members join and leave, the graph is captured only
    when the room runs out, vacant slots do no work, and every member ends
    where its own steps alone leave it."""
    banner("static graph across membership changes")
    systems = W.batched_mix(n_systems=8, seed=21, sizes=(16, 24), lengths=(16, 24, 32))
    mem = make(systems, sort_every=0)          # atomic gridding: no per-member launches
    stops = [6, 10, 4, 12, 8, 9, 5, 7]
    b = scmf.Batch(mem[:2])
    b.set_lanes(8)
    st0 = b.stats()
    assert st0.captures >= 1 and st0.members == 2
    b.update(mem[:3], stops[:3])              # grows past the exact first graph: captured with room
    st1 = b.stats()
    assert st1.captures == st0.captures + 1
    assert st1.member_capacity >= 4 and st1.blocks_used <= st1.blocks_launched
    done = [0] * len(mem)

    def advance(members, n):
        for i in members:
            done[i] = min(stops[i], done[i] + n)

    b.step(3)
    advance([0, 1, 2], 3)
    plan = [[0, 1, 2, 3], [1, 2, 3], [1, 3, 4], [1, 3, 4, 5], [4, 5]]
    for members in plan:
        b.update([mem[i] for i in members], [stops[i] for i in members])
        b.step(2)
        advance(members, 2)
    st2 = b.stats()
    assert st2.captures == st1.captures, "a membership inside the room captured a graph"
    assert st2.table_uploads >= st1.table_uploads + len(plan)
    b.step(12)
    advance(plan[-1], 12)
    b.sync()
    alone = make(systems, sort_every=0)
    for i, g in enumerate(alone):
        for k in range(1, done[i] + 1):
            g.step(k)
    for i, (a, m) in enumerate(zip(alone, mem)):
        assert m.stats().sweeps == done[i], (i, m.stats().sweeps, done[i])
        assert same(a.positions(), m.positions()), f"member {i} positions"
        np.testing.assert_array_equal(a.density(), m.density(), err_msg=f"member {i} density")
        assert same(a.omega(), m.omega()), f"member {i} omega"
        assert a.stats().accepts == m.stats().accepts
    assert done[6] == done[7] == 0 and done[0] < stops[0]   # 6, 7 never ran; 0 left early
    b.close()
    for g in alone + mem:
        g.close()
