// Stage and adjoint kernel instances of polynomial order N = 7.
#define NKC_N1 8
#include "order_inst.hpp"
