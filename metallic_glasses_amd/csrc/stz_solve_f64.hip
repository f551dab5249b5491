// stz_solve_f64.hip -- the stress-solve launch sequence with double spectra.
#define STZ_SOLVE_REAL double
#include "stz_solve_launch.h"
