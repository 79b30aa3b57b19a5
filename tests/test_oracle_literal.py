"""The LITERAL replay (oracle, CPU only): the shipped driver's scratch volumes
with their stale contents (Segmenter.cpp:537-554, 977, 985-1019;
FilterDuplicates.cu:90-100), against the clean REF semantics that every parity
claim of the engine rests on.

The replay is shown to reproduce each artefact where it must fire (small
volumes built for it), and to leave the lists identical to REF on instances of
the benchmark configurations; the full configurations are run by the same
function (oracle.literal_diagnostic) and recorded in the design notes.
"""
import numpy as np
import pytest

import oracle as O
from aworkefficientgpualgorithmforlevelsetsegmentation_amd import synth


def _dense_case():
    n = 32
    z, y, x = np.mgrid[0:n, 0:n, 0:n]
    vol = ((x + y + z) % 7 * 30).astype(np.uint8)
    phi0 = np.clip((x - 15.5) / 3.0, -4, 4).astype(np.float32)
    return vol, phi0


def test_stale_initial_flags_fire_when_the_lists_reach_them():
    """Seed next to the origin of a 32^3 volume: the initial list's tau values
    are small, so 7 align256(n) passes them within a few iterations; the stale
    flags are compacted (artefact 1), some outside the scanned positions, and
    duplicates appear in the list (artefact 3)."""
    n = 32
    vol = np.full((n, n, n), 60, np.uint8)
    vol[2:30, 2:30, 2:30] = 180
    phi0 = synth.sphere_sdf(n, n, n, (3.0, 3.0, 3.0), 2.0)
    r = O.literal_diagnostic(vol, phi0, 1.0, 180, 30, 0.1, 0.5, 200)
    assert r["events"][0] > 0 and r["first_event"] is not None
    assert r["first_divergence"] is not None and r["first_divergence"] >= r["first_event"]
    assert r["events"][2] > 0


def test_stale_tags_after_a_switch_from_the_dense_path():
    """A wide band starts on the dense path and narrows onto the few path: the
    temporal flags the dense iterations left in the tag volume drop candidates
    of lists 1..6 (artefact 2), first at the switch."""
    vol, phi0 = _dense_case()
    st = O.State(vol)
    st.set_phi(phi0, 3.5)
    nh, _, _ = st.iterate(60.0, 20.0, 0.2, 0.5, 80)
    switch = next(k for k, v in enumerate(nh) if O.path_is_few(int(v), st.N))
    r = O.literal_diagnostic(vol, phi0, 3.5, 60.0, 20.0, 0.2, 0.5, 80, initial_flags=False)
    assert r["events"][0] == 0 and r["events"][1] > 0
    assert r["first_event"] == switch and r["first_divergence"] == switch


def test_a_first_dense_iteration_reactivates_the_initial_list():
    """With the initial flags left in the valid volume, a first iteration on the
    dense path compacts all of them: the next list is REF's plus the initial
    list, in ascending tau."""
    vol, phi0 = _dense_case()
    ref = O.State(vol)
    ref.set_phi(phi0, 3.5)
    V0 = ref.active()
    assert not O.path_is_few(len(V0), ref.N)
    ref.iterate(100.0, 60.0, 0.2, 0.5, 1)
    lit = O.State(vol)
    lit.set_phi(phi0, 3.5)
    lit.literal_begin()
    _, _, m, ev = lit.iterate_literal(100.0, 60.0, 0.2, 0.5, 1)
    want = np.union1d(ref.active(), V0).astype(np.uint32)
    assert int(ev[0, 0]) == len(V0)
    assert np.array_equal(lit.active(), want) and int(m[0]) == len(want)
    assert len(want) > len(ref.active())


def test_clean_start_without_switches_equals_ref():
    """Without stale flags and without a dense-to-few switch the replay is REF."""
    w = synth.ellipsoids(edge=48, iterations=80)
    r = O.literal_diagnostic(w.volume, w.phi0, w.band, w.T, w.eps, w.alpha, w.dt, 80, initial_flags=False)
    assert r == {"first_divergence": None, "events": [0, 0, 0], "first_event": None, "iterations": 80,
                 "many_iterations": 0}


@pytest.mark.parametrize("case", ["ellipsoids-128", "adjoint-0", "adjoint-15", "batch-0", "batch-63"])
def test_literal_equals_ref_on_the_benchmark_configurations(case):
    """The artefacts never fire on these instances of the benchmark workloads
    (initial flags included): the lists stay identical to REF every iteration."""
    if case == "ellipsoids-128":
        w = synth.ellipsoids()
    elif case.startswith("adjoint"):
        w = synth.adjoint()[int(case.split("-")[1])]
    else:
        w = synth.batch()[int(case.split("-")[1])]
    r = O.literal_diagnostic(w.volume, w.phi0, w.band, w.T, w.eps, w.alpha, w.dt, w.iterations, threads=4)
    assert r["first_divergence"] is None and r["events"] == [0, 0, 0], r
