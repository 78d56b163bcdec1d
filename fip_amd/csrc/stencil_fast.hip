// Fast stencil: hardware rcp/sqrt/log2/exp2 and FMA contraction (-ffp-contract=fast).
#define FIP_FAST_TU 1
#include "stencil_impl.inc"
