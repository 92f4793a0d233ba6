/* Source compatibility for applications written against ODE 0.12:
 * `#include <ode/ode.h>` resolves to the MI355X library's declarations. */
#ifndef ODEMI_ODE_ODE_H
#define ODEMI_ODE_ODE_H
#include "../odemi.h"
#endif
