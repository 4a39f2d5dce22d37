// Stage and adjoint kernel instances of polynomial order N = 6.
#define NKC_N1 7
#include "order_inst.hpp"
