// Stage and adjoint kernel instances of polynomial order N = 3.
#define NKC_N1 4
#include "order_inst.hpp"
