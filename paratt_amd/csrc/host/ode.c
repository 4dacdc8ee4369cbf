/*
 * ode.c -- step-truncation explicit Euler (paratt_ode.c:13-37) over the
 * GPU-resident primitives.  The vector field is the caller's host C code.
 */
#include "paratt.h"
#include "paratt_ode.h"
#include "ptt_host.h"

/* paratt_ode.c:26-37: f_new = T2(f_old + dt*T1(F(f_old))), caps = current ranks. */
void pattac_exp_euler_st(paratt *f_old, paratt *f_new, paratt *work,
                         void (*vecField)(paratt *, paratt *, int, ...), double dt, double eps_1,
                         double eps_2) {
    vecField(f_old, work, 0);
    paratt_truncate(work, eps_1, work->x->ttRank);
    attac_tt_axpby(1.0, f_old->x, dt, work->x, f_new->x);
    paratt_truncate(f_new, eps_2, f_new->x->ttRank);
}

/* paratt_ode.c:13-24, with the working GPU rounding in place of the
 * reference's serial truncation (whose SVD calls are commented out). */
void attac_exp_euler_st(attac_tt *f_old, attac_tt *f_new, attac_tt *work,
                        void (*vecField)(attac_tt *, attac_tt *, int, ...), double dt,
                        double eps_1, double eps_2) {
    vecField(f_old, work, 0);
    ptt_host_round(work, eps_1, work->ttRank);
    attac_tt_axpby(1.0, f_old, dt, work, f_new);
    ptt_host_round(f_new, eps_2, f_new->ttRank);
}
