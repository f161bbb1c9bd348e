"""Restart from the reference's checkpoint format.  A run of k steps writes
a checkpoint; the file holds exactly the downloaded state; a fresh context
that reads it (and gets the checkpointed run's T as its Newton guess, which
the format does not carry) continues like the oracle restarted from the same
state, bit for bit.  As in the reference, a restart starts the GLF running
maxima afresh (Read_Ubak restores U only, src/XFLUIDS.cpp:689-724)."""
import numpy as np
import pytest

import xfluids_amd as xf
from parity import announce, make, oracle_for

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["vortex", "shock_bubble_z_janaf", "reactive_channel"])
def test_restart_from_checkpoint_matches_oracle_restart(case, tmp_path):
    announce("checkpoint restart " + case)
    s, U, T, nsteps = make(case)
    k = max(1, nsteps // 2)
    b = s.context()
    b.upload(U, T)
    b.initialize()
    b.steps(k, graph=True)
    path = tmp_path / "ck"
    b.save_checkpoint(path, k, 1.5e-6, 2.5)
    Uk, Tk = b.download(0)
    b.close()
    Uf, step, tm, tc = xf.checkpoint_read(path, s.desc)
    assert (step, tm, tc) == (k, 1.5e-6, 2.5)
    assert np.array_equal(Uf.view(np.uint64), Uk.view(np.uint64))

    c = s.context()
    c.upload(U, Tk)
    assert c.load_checkpoint(path)[0] == k
    c.initialize()
    c.steps(nsteps - k, graph=True)
    Uc, Tc = c.download(0)
    c.close()

    o = oracle_for(s, Uk, Tk)
    assert o.initialize() == 0
    st, _ = o.steps(nsteps - k)
    assert st == 0
    assert np.array_equal(o.field("U").view(np.uint64), Uc.view(np.uint64))
    assert np.array_equal(o.field("T").view(np.uint64), Tc.view(np.uint64))
