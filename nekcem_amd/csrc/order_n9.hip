// Stage and adjoint kernel instances of polynomial order N = 8.
#define NKC_N1 9
#include "order_inst.hpp"
