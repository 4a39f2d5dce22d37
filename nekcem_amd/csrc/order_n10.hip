// Stage and adjoint kernel instances of polynomial order N = 9.
#define NKC_N1 10
#include "order_inst.hpp"
