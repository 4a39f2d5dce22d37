// Stage and adjoint kernel instances of polynomial order N = 5.
#define NKC_N1 6
#include "order_inst.hpp"
