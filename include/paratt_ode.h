/*
 * paratt_ode.h -- step-truncation explicit ODE steps on the GPU-resident TT
 * (drop-in for source/paratt_ode.h:28-120; own text).
 */
#ifndef PARATT_ODE_SOLVE_H
#define PARATT_ODE_SOLVE_H

#include <stdlib.h>
#include <stdarg.h>
#include "paratt.h"

#ifdef __cplusplus
extern "C" {
#endif

/* f_new = T2( f_old + dt*T1( vecField(f_old) ) ), T = paratt_truncate with
 * tolerances eps_1 / eps_2 and the current ranks as caps.
 * Replaces paratt_ode.c:26-37.  vecField is host C code; it is called once,
 * synchronously, with nargs = 0, and is expected to compose device-resident
 * primitives (stencils, axpby, Hadamard, truncate) on its TT arguments. */
void pattac_exp_euler_st(paratt *f_old_ptt, paratt *f_new_ptt,
                         paratt *work_ptt,
                         void (*vecField)(paratt *, paratt *, int, ...),
                         double dt, double eps_1, double eps_2);

/* Serial-struct variant (paratt_ode.c:13-24).  The reference routes it
 * through its serial truncation, whose SVD calls are commented out; here it
 * uses the same GPU rounding as pattac_exp_euler_st. */
void attac_exp_euler_st(attac_tt *f_old_tt, attac_tt *f_new_tt,
                        attac_tt *work_tt,
                        void (*vecField)(attac_tt *, attac_tt *, int, ...),
                        double dt, double eps_1, double eps_2);

#ifdef __cplusplus
}
#endif
#endif
