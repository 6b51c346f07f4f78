// Integrator kernels for scenes whose grids hold packed (reduced-precision)
// records: the kernels of r2_render_kernels.hpp compiled with the packed
// record path fixed at compile time.
#define R2_PREC_MODE 1
#define R2K_NS r2k_packed
#include "r2_render_kernels.hpp"
