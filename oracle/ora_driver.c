/*
 * Time-step ordering of the oracle (test infrastructure only).
 *
 * ora_initialize: src/main.cpp:36-47 (BoundaryCondition(0), UpdateStates(0)).
 * ora_steps:      XFLUIDS::Evolution src/XFLUIDS.cpp:172-225 without output
 *                 clamps: dt = ComputeTimeStep; Strang: Reaction; RK3 stages
 *                 (RungeKuttaSP3rd src/XFLUIDS.cpp:441-525); Reaction.
 * ora_reaction:   XFLUIDS::Reaction src/XFLUIDS.cpp:625-642.
 */
#include "ora_internal.h"

int ora_chemistry_status(void);

int ora_initialize(ora_fluid *f)
{
    ora_boundary(f, 0);
    return ora_update_states(f, 0);
}

int ora_reaction(ora_fluid *f, double dt)
{
    int err = ora_update_states(f, 0);
    if (err) return err;
    double tem_dt = dt;
    if (f->cfg.strang) tem_dt *= 0.5;
    ora_chemistry(f, tem_dt);
    if (ora_chemistry_status()) return 5;
    return ora_estimate_nan(f, 4);
}

static int rk_stage(ora_fluid *f, int flag, double dt)
{
    int which = (flag == 1) ? 0 : 1;
    ora_boundary(f, which);
    int err = ora_update_states(f, which);
    if (err) return err;
    ora_compute_lu(f, which);
    err = ora_estimate_nan(f, flag);
    if (err) return err;
    ora_rk3(f, flag, dt);
    return 0;
}

int ora_steps(ora_fluid *f, int nsteps, double *dt_out)
{
    for (int s = 0; s < nsteps; s++) {
        double dt = ora_get_dt(f);
        if (dt_out) dt_out[s] = dt;
        int err = 0;
        if (f->cfg.react && f->cfg.strang)
            if ((err = ora_reaction(f, dt))) return err;
        for (int flag = 1; flag <= 3; flag++)
            if ((err = rk_stage(f, flag, dt))) return err;
        if (f->cfg.react)
            if ((err = ora_reaction(f, dt))) return err;
    }
    return 0;
}
