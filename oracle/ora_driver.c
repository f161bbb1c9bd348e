/*
 * Time-step ordering of the oracle (test infrastructure only).
 *
 * ora_initialize: src/main.cpp:36-47 (BoundaryCondition(0), UpdateStates(0)).
 * ora_steps:      XFLUIDS::Evolution src/XFLUIDS.cpp:170-225 without output:
 *                 while physicalTime < target_t: Iteration++; dt =
 *                 ComputeTimeStep, clamped so that physicalTime lands on
 *                 target_t (:196-198); Strang: Reaction; RK3 stages
 *                 (RungeKuttaSP3rd src/XFLUIDS.cpp:441-525); Reaction.  The
 *                 first error stops the loop (:285-287) and is kept with its
 *                 step and stage.
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

static int failed(ora_fluid *f, int err, int where)
{
    f->err_step = f->step;
    f->err_where = where;
    return err;
}

int ora_steps(ora_fluid *f, int nsteps, double *dt_out)
{
    for (int s = 0; s < nsteps; s++) {
        if (!(f->time < f->target)) {
            f->done = 1;
            return 0;
        }
        f->step++;
        double dt = ora_get_dt(f);
        if (f->time + dt > f->target) dt = f->target - f->time;
        f->time += dt;
        if (dt_out) dt_out[s] = dt;
        int err = 0;
        if (f->cfg.react && f->cfg.strang)
            if ((err = ora_reaction(f, dt))) return failed(f, err, 4);
        for (int flag = 1; flag <= 3; flag++)
            if ((err = rk_stage(f, flag, dt))) return failed(f, err, flag);
        if (f->cfg.react)
            if ((err = ora_reaction(f, dt))) return failed(f, err, 4);
    }
    return 0;
}

void ora_set_time(ora_fluid *f, double time, int step)
{
    f->time = time;
    f->step = step;
}

void ora_set_target_time(ora_fluid *f, double target)
{
    f->target = target;
    f->done = 0;
}

void ora_get_time(const ora_fluid *f, double *time, int *step, int *reached)
{
    if (time) *time = f->time;
    if (step) *step = f->step;
    if (reached) *reached = f->done;
}

int ora_last_error(const ora_fluid *f, int *i, int *j, int *k, int *step, int *where)
{
    const ora_config *c = &f->cfg;
    const size_t id = f->err_id;
    if (i) *i = f->err_code ? (int)(id % (size_t)c->Xmax) : -1;
    if (j) *j = f->err_code ? (int)((id / (size_t)c->Xmax) % (size_t)c->Ymax) : -1;
    if (k) *k = f->err_code ? (int)(id / ((size_t)c->Xmax * c->Ymax)) : -1;
    if (step) *step = f->err_code ? f->err_step : 0;
    if (where) *where = f->err_code ? f->err_where : 0;
    return f->err_code;
}
