// Stage and adjoint kernel instances of polynomial order N = 4.
#define NKC_N1 5
#include "order_inst.hpp"
