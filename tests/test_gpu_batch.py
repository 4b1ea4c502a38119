"""Batching form (ragged molecules, one submission) on the GPU through the
C ABI: qmckl_get_mo_basis_mo_vgl_batch_device over many contexts must give,
for every molecule, exactly what that molecule's own inplace getter gives,
and the FP64 results must match the oracle."""
import numpy as np
import pytest

from qmckl_sycl_amd import systems as S
from qmckl_sycl_amd.context import BatchGraph, Context, DeviceArray, QmcklError, batch_mo_vgl

from helpers import RTOL_VGL, assert_vgl_close, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def molecules():
    # the 48-molecule ragged batch, walker counts scaled down for test time
    return S.ragged_batch(48, scale=0.04)


def setup(mols, forms=None):
    ctxs, outs = [], []
    for i, m in enumerate(mols):
        c = Context(0)
        c.set_system(m)
        if forms is not None:
            c.call("qmckl_set_mo_basis_precision_device", forms[i % len(forms)])
        ctxs.append(c)
        outs.append(DeviceArray(c, (m.point_num, 5, m.mo_num)))
    return ctxs, outs


def single(c, m):
    out = DeviceArray(c, (m.point_num, 5, m.mo_num))
    c.call("qmckl_get_mo_basis_mo_vgl_inplace_device", out.ptr, out.size)
    return out.numpy()


def test_batch_equals_single_fp64(molecules, oracle_lib):
    say("points per molecule:", [m.point_num for m in molecules])
    say("AOs per molecule:", [m.ao_num for m in molecules])
    ctxs, outs = setup(molecules)
    try:
        batch_mo_vgl(ctxs, outs)
        got = [o.numpy() for o in outs]
        for i, (c, m) in enumerate(zip(ctxs, molecules)):
            ref = single(c, m)
            assert np.array_equal(got[i], ref), m.name
        # FP64 parity with the oracle on the smallest, the largest and two
        # others (all their points)
        order = np.argsort([m.point_num * m.ao_num for m in molecules])
        for i in (order[0], order[len(order) // 3], order[2 * len(order) // 3], order[-1]):
            m = molecules[i]
            ref, mag = oracle_lib.mo_vgl(m, with_mag=True)
            assert_vgl_close(got[i], ref, mag, what=f"batch {m.name} P={m.point_num} A={m.ao_num}")
    finally:
        for c in ctxs:
            c.destroy()


@pytest.mark.parametrize("form", [0, 2])
def test_batch_full_size_against_oracle(oracle_lib, form):
    """The batch at its full size (48 molecules, 16-512 walkers, 1.52 M
    points) in the FP64 form and in FP16X2: the two largest molecules by
    P A M and six spread over the batch checked against the oracle on 64
    sampled points each (FP64: RTOL_VGL relative to the magnitude sums;
    FP16X2: 1e-6 of each (point, component) row's maximum), and the largest
    also bitwise against its own inplace getter."""
    mols = S.ragged_batch(48)
    say("full batch points:", sum(m.point_num for m in mols))
    ctxs, outs = setup(mols, forms=[form])
    try:
        batch_mo_vgl(ctxs, outs)
        by_work = sorted(range(len(mols)), key=lambda i: mols[i].point_num * mols[i].ao_num * mols[i].mo_num)
        picks = sorted(set(by_work[-2:]) | set(np.linspace(0, len(mols) - 1, 6).astype(int).tolist()))
        for i in picks:
            m = mols[i]
            got = outs[i].numpy()
            if i == by_work[-1]:
                assert np.array_equal(got, single(ctxs[i], m)), m.name
            idx = np.sort(np.random.default_rng(i).choice(m.point_num, size=min(64, m.point_num), replace=False))
            coord = np.ascontiguousarray(m.coord_T()[:, idx])
            what = f"full batch form {form} {m.name} P={m.point_num} A={m.ao_num} M={m.mo_num}"
            if form == 0:
                ref, mag = oracle_lib.mo_vgl(m, coord_T=coord, with_mag=True)
                assert_vgl_close(got[idx], ref, mag, what=what)
            else:
                ref = oracle_lib.mo_vgl(m, coord_T=coord)
                rm = np.abs(ref).max(axis=2, keepdims=True)
                e = float((np.abs(got[idx] - ref) / np.where(rm > 0, rm, 1.0)).max())
                say(f"{what}: max row-relative error {e:.2e}")
                assert e <= 1e-6, (what, e)
    finally:
        for c in ctxs:
            c.destroy()


def test_batch_mixed_forms_and_moving_points(molecules):
    # the split forms' contractions run as persistent launches over the
    # batch; every molecule must still equal its own single-context result
    mols = molecules[:25]
    ctxs, outs = setup(mols, forms=[0, 2, 3, 1, 4, 5])
    try:
        rng = np.random.default_rng(5)
        for step in range(2):
            batch_mo_vgl(ctxs, outs)
            for c, m, o in zip(ctxs, mols, outs):
                assert np.array_equal(o.numpy(), single(c, m)), (step, m.name)
            # move every electron and run again: the batch re-stamps each context
            for c, m in zip(ctxs, mols):
                c.set_electron_coord(m.elec_coord + rng.normal(0, 0.05, m.elec_coord.shape))
    finally:
        for c in ctxs:
            c.destroy()


def test_batch_rejects_small_output(molecules):
    mols = molecules[:3]
    ctxs, outs = setup(mols)
    try:
        short = DeviceArray(ctxs[1], (1,))
        with pytest.raises(QmcklError) as e:
            batch_mo_vgl(ctxs, [outs[0], short, outs[2]])
        assert e.value.code == 4
    finally:
        for c in ctxs:
            c.destroy()


def _move(ctxs, mols, rng):
    """New electron positions for every molecule (same counts)."""
    for c, m in zip(ctxs, mols):
        e = np.asarray(m.elec_coord)
        c.set_electron_coord(e + rng.normal(0.0, 0.1, size=e.shape))


def test_batch_graph_replays_match_eager(molecules):
    """The batch recorded into a HIP graph: every replay, with the points
    moved and eager batches run in between, equals the eager batch on the
    same points bitwise (FP64, FP16X2 and INT8X5 contexts mixed)."""
    mols = molecules[:12]
    ctxs, outs = setup(mols, forms=[0, 2, 5])
    ref_outs = [DeviceArray(c, (m.point_num, 5, m.mo_num)) for c, m in zip(ctxs, mols)]
    g = None
    try:
        g = BatchGraph(ctxs, outs)
        rng = np.random.default_rng(11)
        for rep in range(3):
            _move(ctxs, mols, rng)
            batch_mo_vgl(ctxs, ref_outs)          # eager work between replays
            ref = [o.numpy() for o in ref_outs]
            assert g.launch() == 0
            got = [o.numpy() for o in outs]
            for i, m in enumerate(mols):
                assert np.array_equal(got[i], ref[i]), (rep, m.name)
            say(f"replay {rep}: {len(mols)} molecules equal to the eager batch")
    finally:
        if g is not None:
            g.destroy()
        for c in ctxs:
            c.destroy()


def test_batch_graph_refuses_stale_launch(molecules):
    """After a change that alters the recorded kernels' arguments (here the
    precision form) a launch fails instead of running stale work; unknown
    handles are rejected."""
    mols = molecules[:3]
    ctxs, outs = setup(mols)
    g = None
    try:
        g = BatchGraph(ctxs, outs)
        assert g.launch() == 0
        ctxs[1].call("qmckl_set_mo_basis_precision_device", 2)
        assert g.launch() == 101          # QMCKL_FAILURE_DEVICE
        ctxs[1].call("qmckl_set_mo_basis_precision_device", 0)
        assert g.launch() == 0
        # a finalize rebuilds the device tables (they may land at the old
        # addresses with other sizes): refused, whatever the addresses
        ctxs[2].call("qmckl_finalize_ao_basis_device")
        assert g.launch() == 101
        assert lib_rc_launch(987654321) == 1   # unknown handle: ARG_1
    finally:
        if g is not None:
            g.destroy()
        for c in ctxs:
            c.destroy()


def lib_rc_launch(handle):
    from qmckl_sycl_amd.context import lib
    return lib().qmckl_mo_basis_mo_vgl_batch_graph_launch_device(handle)



def test_batch_persistent_and_per_context_contractions_mixed():
    """One batch mixing contexts whose split-form contraction goes into the
    persistent launches (up to 192 MOs) with one that keeps its own launch
    (300 MOs, tile width 5), a three-term form and an FP64 context: every
    molecule equals its single-context result bitwise, eagerly and on graph
    replays after the points move."""
    mols = [S.water_cluster(n_side=2, walk_num=2, mo_num=300, name="water8_300mo"),
            S.water_cluster(n_side=2, walk_num=2, mo_num=100, seed=7, name="water8_100mo"),
            S.n2_like(walk_num=3, seed=11),
            S.water_cluster(n_side=2, walk_num=1, mo_num=40, seed=9, name="water8_40mo")]
    forms = [2, 4, 2, 0]
    ctxs, outs = [], []
    for m, f in zip(mols, forms):
        c = Context(0)
        c.set_system(m)
        c.call("qmckl_set_mo_basis_precision_device", f)
        ctxs.append(c)
        outs.append(DeviceArray(c, (m.point_num, 5, m.mo_num)))
    g = None
    try:
        batch_mo_vgl(ctxs, outs)
        for c, m, o in zip(ctxs, mols, outs):
            assert np.array_equal(o.numpy(), single(c, m)), m.name
        g = BatchGraph(ctxs, outs)
        rng = np.random.default_rng(3)
        for rep in range(2):
            _move(ctxs, mols, rng)
            assert g.launch() == 0
            for c, m, o in zip(ctxs, mols, outs):
                assert np.array_equal(o.numpy(), single(c, m)), (rep, m.name)
        say("persistent + per-context batch: eager and 2 graph replays equal to single calls")
    finally:
        if g is not None:
            g.destroy()
        for c in ctxs:
            c.destroy()
