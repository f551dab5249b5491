// stz_solve_f32.hip -- the stress-solve launch sequence with float spectra.
#define STZ_SOLVE_REAL float
#include "stz_solve_launch.h"
