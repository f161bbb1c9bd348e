"""Write tests/golden/shock_tube_1024_170us.npz: the reactive shock tube
(H2/O2/Ar 2:1:7, 9 species, 19 reactions, Strang; cases/reactive-shock-tube.json)
on 1024 cells, advanced by the CPU oracle from its seeded initial field to
t = 170 us, the start of the reference's output window
(settings/1d-reactive-st.json): the shock has reflected from the left wall and
heated the gas there to about 1200 K.

Stored: U (AoS with ghost cells), T, the time and step count, and the number
of steps.  A GPU test advances both from this state and compares them.

    python tests/golden/make_shock_tube_fixture.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import xfluids_amd as xf  # noqa: E402
from oracle_lib import Oracle, config_from_setup  # noqa: E402

NX, T_END, SEED = 1024, 170e-6, 1
OUT = os.path.join(HERE, "shock_tube_1024_170us.npz")


def main():
    s = xf.Setup(xf.case_path("reactive-shock-tube"), nx=NX)
    U, T = s.initial("reactive-shock-tube", seed=SEED)
    o = Oracle(config_from_setup(s))
    o.set_state(U, T)
    assert o.initialize() == 0
    o.set_target_time(T_END)
    t0 = time.time()
    while True:
        st, _ = o.steps(500)
        assert st == 0, st
        t, steps, reached = o.time()
        print("t = %.3f us, %d steps, %.0f s" % (t * 1e6, steps, time.time() - t0), flush=True)
        if reached:
            break
    np.savez_compressed(OUT, U=o.field("U").copy(), T=o.field("T").copy(), time=t, steps=steps, nx=NX, seed=SEED)
    print("wrote", OUT, "T max %.1f K" % o.field("T").max())


if __name__ == "__main__":
    main()
